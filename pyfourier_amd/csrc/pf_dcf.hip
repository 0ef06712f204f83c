// Voronoi cell areas of 2-D k-space points (density compensation), gfx950.
//
// The reference tessellates with scipy.spatial.Voronoi (Qhull) on the host
// and sums the area of every bounded cell (src/pyfourier/_dcomp/voronoi.py:
// 220-258, 277-300); open cells (points on the convex hull) get 0.  Here each
// cell is built directly on the GPU by half-plane clipping, one thread per
// point, in double precision: the cell of p starts as a box far larger than
// the point set and is cut by the bisector of p and every candidate q,
// {x : (x - p).(q - p) <= |q - p|^2 / 2}.  Candidates come from a uniform
// bucket grid (points sorted by bucket, bucket start offsets), visited in
// square rings around p's bucket; after ring r every unvisited point is at
// least r*h from p.  A point can only cut the cell if it lies in one of the
// disks centred on the cell's vertices that pass through p: empty buckets and
// buckets that miss the disks' bounding box or the disks are skipped, and the
// scan stops once that box lies inside the visited square (cells next to the
// hull, with long thin tails, then only test the buckets of a strip along the
// boundary).  Points the caller knows to lie on the convex hull (their cells
// are the unbounded ones, whose disks never close) are given 0 unscanned.
// A cell that still has a side of the far box after all rings is open (area
// 0).  Each edge remembers the line it lies on, and the area is taken from
// the vertices recomputed as exact intersections of two bisectors, so the
// far box's rounding does not reach it.  Points that coincide to rounding
// (closer than 1e-13 of the extent: e.g. the same spoke reached twice by the
// rotations of a full-spoke disk) are one site to each other, each getting
// the whole cell, as Qhull treats them in the reference.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "pf_internal.h"
#include "pfhip.h"

namespace {

constexpr int VOR_NV = 64;         // vertices a cell may hold in the first pass
constexpr double VOR_FAR = 1.0e6;     // box half-size, in units of the point set's extent
constexpr double VOR_MERGE = 1.0e-13;  // points closer than this (same units) are one site
constexpr double VOR_PARALLEL = 1.0e-9;  // |sin| below which two bisectors count as parallel
// device-memory bytes per vertex of a second-pass cell: v, tv (16 each), e, te (4 each)
constexpr int VOR_WORK_BYTES = 2 * 16 + 2 * 4;

// The cell is a convex polygon with vertices relative to p (approximate
// positions, for the inside tests) and, per edge k (vertex k -> k + 1), the
// line it lies on: a candidate's index (the bisector of p and that point) or
// -1 for a side of the initial box.  Clip by the bisector of candidate `id`
// (relative position q): {x : x.q <= |q|^2 / 2}.  False on overflow.
__device__ bool vor_clip(double2* v, int* e, int& nv, double2* tv, int* te, double2 q, int id, int nvmax) {
    const double c = 0.5 * (q.x * q.x + q.y * q.y);
    int n = 0;
    for (int k = 0; k < nv; ++k) {
        const int k1 = k + 1 < nv ? k + 1 : 0;
        const double2 a = v[k], b = v[k1];
        const double sa = a.x * q.x + a.y * q.y - c, sb = b.x * q.x + b.y * q.y - c;
        if (sa <= 0.0) {
            if (n >= nvmax) return false;
            tv[n] = a;
            te[n++] = (sa == 0.0 && sb > 0.0) ? id : e[k];  // on the line and leaving: the edge is the new line
        }
        if ((sa < 0.0 && sb > 0.0) || (sa > 0.0 && sb < 0.0)) {
            if (n >= nvmax) return false;
            const double f = sa / (sa - sb);
            tv[n] = make_double2(a.x + f * (b.x - a.x), a.y + f * (b.y - a.y));
            te[n++] = sa < 0.0 ? id : e[k];  // leaving: along the new line; entering: rest of the old edge
        }
    }
    for (int k = 0; k < n; ++k) {
        v[k] = tv[k];
        e[k] = te[k];
    }
    nv = n;
    return true;
}

// The cell's security region: a point q can cut the cell only if it is
// closer to some vertex than p is, i.e. lies in a disk centred on a vertex
// and passing through p.  Clipping only shrinks that union of disks (a new
// vertex lies between two old ones, and its disk inside theirs), so a bucket
// that misses it never needs a visit.  Kept as the disks' bounding box
// (absolute coordinates) and the largest squared vertex distance.
struct VorFlower {
    double x0, x1, y0, y1, r2max;
};

__device__ VorFlower vor_flower(const double2* v, int nv, double2 p) {
    VorFlower f{1e300, -1e300, 1e300, -1e300, 0.0};
    for (int k = 0; k < nv; ++k) {
        const double r2 = v[k].x * v[k].x + v[k].y * v[k].y, r = sqrt(r2);
        f.x0 = fmin(f.x0, p.x + v[k].x - r);
        f.x1 = fmax(f.x1, p.x + v[k].x + r);
        f.y0 = fmin(f.y0, p.y + v[k].y - r);
        f.y1 = fmax(f.y1, p.y + v[k].y + r);
        f.r2max = fmax(f.r2max, r2);
    }
    return f;
}

// does the rectangle [rx0, rx1] x [ry0, ry1] meet the disk of some vertex?
__device__ bool vor_touches(const double2* v, int nv, double2 p, double rx0, double rx1, double ry0, double ry1) {
    for (int k = 0; k < nv; ++k) {
        const double cx = p.x + v[k].x, cy = p.y + v[k].y;
        const double dx = fmax(fmax(rx0 - cx, cx - rx1), 0.0), dy = fmax(fmax(ry0 - cy, cy - ry1), 0.0);
        if (dx * dx + dy * dy <= v[k].x * v[k].x + v[k].y * v[k].y) return true;
    }
    return false;
}

// The cell of point i (sorted index) in the caller's buffers of nvmax
// vertices; its area, 0 for an open cell, or -1 when nvmax is too small.
__device__ double vor_cell(const double2* __restrict__ pts, const int32_t* __restrict__ start, int nx, int ny,
                           double x0, double y0, double h, double extent, int64_t i, int nvmax, double2* v, int* e,
                           double2* tv, int* te) {
    const double2 p = pts[i];
    const double far = VOR_FAR * extent;
    const double merge2 = VOR_MERGE * extent * VOR_MERGE * extent;
    const double eps = 1e-9 * h;  // bucket rectangles widened past the rounding of the bucket assignment
    int nv = 4;
    v[0] = make_double2(-far, -far);
    v[1] = make_double2(far, -far);
    v[2] = make_double2(far, far);
    v[3] = make_double2(-far, far);
    e[0] = e[1] = e[2] = e[3] = -1;
    VorFlower f = vor_flower(v, nv, p);
    const int cx = min(max((int)floor((p.x - x0) / h), 0), nx - 1);
    const int cy = min(max((int)floor((p.y - y0) / h), 0), ny - 1);
    const int rmax = max(nx, ny);
    bool ok = true;
    // one bucket: skipped unless it holds points and meets the security
    // region, else clip by each of its points
    auto visit = [&](int bx, int by) {
        if (bx < 0 || bx >= nx || by < 0 || by >= ny) return;
        const int b0 = by * nx + bx;
        if (start[b0] == start[b0 + 1]) return;
        const double rx0 = x0 + bx * h - eps, rx1 = x0 + (bx + 1) * h + eps;
        const double ry0 = y0 + by * h - eps, ry1 = y0 + (by + 1) * h + eps;
        if (rx1 < f.x0 || rx0 > f.x1 || ry1 < f.y0 || ry0 > f.y1) return;
        if (!vor_touches(v, nv, p, rx0, rx1, ry0, ry1)) return;
        const int b = by * nx + bx;
        for (int k = start[b]; k < start[b + 1]; ++k) {
            if (k == i) continue;
            const double2 q = make_double2(pts[k].x - p.x, pts[k].y - p.y);
            const double d2 = q.x * q.x + q.y * q.y;
            // bisector beyond every vertex, or a point that coincides with p to
            // rounding (one site, as Qhull merges it: both get the whole cell)
            if (d2 >= 4.0 * f.r2max || d2 <= merge2) continue;
            if (!vor_clip(v, e, nv, tv, te, q, k, nvmax)) {
                ok = false;
                return;
            }
            f = vor_flower(v, nv, p);
        }
    };
    // square rings around p's bucket (close to distance order, so a cell
    // takes its final shape from its nearest points first) until the
    // security region lies inside the visited square
    for (int r = 0; r <= rmax && ok; ++r) {
        if (r == 0) {
            visit(cx, cy);
        } else {
            for (int dx = -r; dx <= r && ok; ++dx) {
                visit(cx + dx, cy - r);
                visit(cx + dx, cy + r);
            }
            for (int dy = -r + 1; dy <= r - 1 && ok; ++dy) {
                visit(cx - r, cy + dy);
                visit(cx + r, cy + dy);
            }
        }
        const int ix0 = (int)fmin(fmax(floor((f.x0 - x0) / h), 0.0), (double)(nx - 1));
        const int ix1 = (int)fmin(fmax(floor((f.x1 - x0) / h), 0.0), (double)(nx - 1));
        const int iy0 = (int)fmin(fmax(floor((f.y0 - y0) / h), 0.0), (double)(ny - 1));
        const int iy1 = (int)fmin(fmax(floor((f.y1 - y0) / h), 0.0), (double)(ny - 1));
        if (ix0 >= cx - r && ix1 <= cx + r && iy0 >= cy - r && iy1 <= cy + r) break;
    }
    if (!ok) return -1.0;
    bool open = nv < 3;
    for (int k = 0; k < nv; ++k) open = open || e[k] < 0;  // a side of the box survived: unbounded cell
    if (open) return 0.0;
    // exact vertices: vertex k joins edges k - 1 and k, the bisectors of p
    // with points a and b: x.qa = |qa|^2/2, x.qb = |qb|^2/2
    auto vertex = [&](int k) {
        const int ia = e[k > 0 ? k - 1 : nv - 1], ib = e[k];
        const double2 qa = make_double2(pts[ia].x - p.x, pts[ia].y - p.y);
        const double2 qb = make_double2(pts[ib].x - p.x, pts[ib].y - p.y);
        const double ca = 0.5 * (qa.x * qa.x + qa.y * qa.y), cb = 0.5 * (qb.x * qb.x + qb.y * qb.y);
        const double det = qa.x * qb.y - qa.y * qb.x;
        // (nearly) parallel bisectors, e.g. of two points that coincide to
        // rounding, meet in a zero-length edge: keep the clipped position
        if (fabs(det) <= VOR_PARALLEL * sqrt((2.0 * ca) * (2.0 * cb))) return v[k];
        return make_double2((ca * qb.y - cb * qa.y) / det, (qa.x * cb - qb.x * ca) / det);
    };
    double s = 0.0;
    const double2 a0 = vertex(0);
    double2 a = a0;
    for (int k = 0; k < nv; ++k) {
        const double2 b = k + 1 < nv ? vertex(k + 1) : a0;
        s += a.x * b.y - a.y * b.x;
        a = b;
    }
    return 0.5 * fabs(s);
}

// every point, VOR_NV-vertex buffers in private memory
__global__ void __launch_bounds__(64) k_voronoi2d(const double2* __restrict__ pts, const int32_t* __restrict__ start,
                                                  int nx, int ny, double x0, double y0, double h, double extent,
                                                  int64_t M, const uint8_t* __restrict__ open,
                                                  double* __restrict__ area) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    if (open && open[i]) {  // known to lie on the convex hull: an unbounded cell
        area[i] = 0.0;
        return;
    }
    double2 v[VOR_NV], tv[VOR_NV];
    int e[VOR_NV], te[VOR_NV];
    area[i] = vor_cell(pts, start, nx, ny, x0, y0, h, extent, i, VOR_NV, v, e, tv, te);
}

// the listed points again, nvmax-vertex buffers in device memory (cells
// with more edges than VOR_NV: e.g. the k-space centre of a radial
// trajectory, one edge per spoke direction)
__global__ void __launch_bounds__(64) k_voronoi2d_list(const double2* __restrict__ pts,
                                                       const int32_t* __restrict__ start, int nx, int ny, double x0,
                                                       double y0, double h, double extent,
                                                       const int32_t* __restrict__ list, int64_t nlist, int nvmax,
                                                       char* __restrict__ work, double* __restrict__ area,
                                                       int32_t* __restrict__ status) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nlist) return;
    char* w = work + j * (int64_t)nvmax * VOR_WORK_BYTES;
    double2* v = reinterpret_cast<double2*>(w);
    double2* tv = v + nvmax;
    int* e = reinterpret_cast<int*>(tv + nvmax);
    int* te = e + nvmax;
    const int64_t i = list[j];
    const double a = vor_cell(pts, start, nx, ny, x0, y0, h, extent, i, nvmax, v, e, tv, te);
    if (a < 0.0) atomicOr(status, 1);
    area[i] = a < 0.0 ? 0.0 : a;
}

}  // namespace

extern "C" int pf_voronoi_area_2d(const double* pts, const int* bucket_start, int nx, int ny, double x0, double y0,
                                  double h, long long npoints, const unsigned char* open, const int* list,
                                  long long nlist, void* work, int work_vertices, double* area, int* status,
                                  void* stream) {
    if (npoints < 0 || nlist < 0 || nx < 1 || ny < 1 || !(h > 0.0)) return pf_fail(PF_ERR_ARG, "bad Voronoi bucket grid");
    if (npoints == 0 || (list && nlist == 0)) return PF_OK;
    if (!pts || !bucket_start || !area) return pf_fail(PF_ERR_ARG, "null array");
    if (list && (!work || !status || work_vertices < 4 || work_vertices % 2))
        return pf_fail(PF_ERR_ARG, "a point list needs work buffers of an even vertex count");
    static_assert(VOR_WORK_BYTES == 40, "work layout documented in pfhip.h");
    if ((int64_t)nx * ny > 0x7ffffffe || npoints > 0x7fffffff) return pf_fail(PF_ERR_UNSUPPORTED, "too many points");
    const double extent = h * (double)(nx > ny ? nx : ny);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const double2* p2 = reinterpret_cast<const double2*>(pts);
    if (!list) {
        hipLaunchKernelGGL(k_voronoi2d, dim3((unsigned)((npoints + 63) / 64)), dim3(64), 0, s, p2, bucket_start, nx, ny,
                           x0, y0, h, extent, npoints, open, area);
    } else {
        hipLaunchKernelGGL(k_voronoi2d_list, dim3((unsigned)((nlist + 63) / 64)), dim3(64), 0, s, p2, bucket_start, nx,
                           ny, x0, y0, h, extent, list, nlist, work_vertices, static_cast<char*>(work), area, status);
    }
    PF_HIP(hipGetLastError());
    return PF_OK;
}
