#!/usr/bin/env python
"""Headline benchmark: 3D multicoil NUFFT forward+adjoint on one MI355X.

Workload: 128^3 complex64 image, 32 coils batched in one call (smooth
Gaussian sensitivity maps applied before the timed region), golden-means 3D
radial trajectory 8192 spokes x 256 samples, oversampling 2.0 (256^3 grid),
ramp density weights (|k|/k_max)^2.  One step = one nufft + one nufft_adj on
a precomputed plan (plan creation excluded), inputs resident in HBM.

The headline uses Kaiser-Bessel width 6, the smallest width whose error
against the FP64 NUDFT is below 1e-4 at oversampling 2 (width 4: ~1.2e-3,
width 5: ~1.7e-4); the configured width 4 is timed beside it.

Prints one JSON line (rank 0) with the throughput, the roofline of the
dominant hand-written kernel (HIP events around every stage on the stream the
kernels run on, over the timed region), per-stage times, the error against an
FP64 NUDFT on random subsets, and the CPU baseline: the reference algorithm
restated in C/OpenMP + torch CPU ops (oracle/port.py) timed on this host's
cores on a bounded sample.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
METRIC = "NUFFT fwd+adj/s, 3D 128³×32 coils @rel err<1e-4; batched+grad vs plain, %HBM"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--width", type=int, default=6, help="KB width of the headline (6: rel err < 1e-4)")
    ap.add_argument("--compare-width", type=int, default=4, help="second width timed beside it (0: none)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the NUDFT error check")
    ap.add_argument("--cpu-coils", type=int, default=32, help="coils in the CPU baseline sample")
    ap.add_argument("--cpu-runs", type=int, default=5, help="timed CPU runs (median) after one warm-up")
    ap.add_argument("--small", action="store_true", help="64^3 x 8 coils quick run")
    ap.add_argument("--no-extras", action="store_true", help="skip the batched / autograd comparisons")
    ap.add_argument("--only", default=None, help="run one comparison of the extras by name and print its JSON")
    ap.add_argument("--summarize-pmc", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"), default=None,
                    help="write the HBM traffic file from two rocprofv3 --pmc runs of this script "
                         "(FETCH_SIZE, WRITE_SIZE; csv output named run) and exit")
    return ap.parse_args()


def make_inputs(N, C, nspokes, nsamples):
    from pyfourier_amd import demo

    rng = np.random.default_rng(1)
    x = demo.complex_gaussian(rng, (N, N, N))
    maps = demo.gaussian_coil_maps(np.random.default_rng(2), C, (N, N, N))
    xc = (maps * x[None]).astype(np.complex64)
    coord = demo.golden_means_radial_3d(nspokes, nsamples)
    k = np.sqrt((coord.astype(np.float64) ** 2).sum(-1))
    w = ((k / k.max()) ** 2).astype(np.float32)
    return x, xc, coord, w


def stage_bytes(name, C, N3, G3, M, d):
    b = {
        "pad_fft": C * N3 * 8 + C * G3 * 8,
        "ifft_crop": C * G3 * 8 + C * N3 * 8,
        "deapod_pad": C * N3 * 8 + C * G3 * 8,
        "fft_fwd": 2 * C * G3 * 8,
        "interp": C * G3 * 8 + C * M * 8 + M * 4 * (d + 1),
        "unsort": 2 * C * M * 8 + M * 8,
        "sort": 2 * C * M * 8 + M * 8,
        "spread": C * M * 8 + M * 4 * (d + 1) + C * G3 * 8,
        "fft_adj": 2 * C * G3 * 8,
        "crop_deapod": 2 * C * N3 * 8,
    }
    return b[name]


def cpu_baseline(coord, xc, w, N, width, ncoils_sample, C_full, runs=5):
    """Time the reference algorithm on the host: C/OpenMP restatement of the
    Numba loops + torch CPU apod/pad/FFT.  One warm-up fwd+adj, then the
    median of `runs` fwd+adj on the coil sample.  Also returns the last run's
    forward output and the adjoint of that output, for the parity check."""
    import torch

    from oracle import port as Pt
    from oracle import restate as R

    ncores = len(os.sched_getaffinity(0))
    Pt.set_threads(ncores)
    P = R.Plan(coord, N, width=width, oversamp=2.0)
    pp = Pt.PortPlan(P)
    wf = w.reshape(1, 1, -1)
    xs = xc[:ncoils_sample, None]
    times = []
    for r in range(runs + 1):
        t0 = time.perf_counter()
        y = Pt.nufft(xs, pp) * wf
        xa = Pt.nufft_adj(y * wf, pp)
        dt = time.perf_counter() - t0
        log(f"cpu baseline run {r} ({'warm-up' if r == 0 else 'timed'}): {dt:.2f} s")
        if r > 0:
            times.append(dt)
    med = float(np.median(times))
    per_full = med * C_full / ncoils_sample
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    info = {
        "value": 1.0 / per_full,
        "unit": "fwd+adj/s",
        "cores": ncores,
        "kind": "port",
        "sample": f"{ncoils_sample} of {C_full} coils, full trajectory; median of {runs} fwd+adj after a "
                  f"full-size warm-up ({', '.join(f'{t:.2f}' for t in times)} s)"
                  + (f", scaled x{C_full / ncoils_sample:g}" if ncoils_sample != C_full else "")
                  + f"; C/OpenMP restatement of the Numba loops + torch CPU FFT, no gc.collect overheads; "
                  f"CPU: {model}",
        "torch_threads": torch.get_num_threads(),
    }
    return info, y, xa


def parity_vs_port(pf, plan, y_port, x_port, xc_dev, w_dev, dev):
    """Relative L2 of the GPU forward (on the same coil images) and of the GPU
    adjoint (on the port's own forward output, so both adjoints see the same
    input) against the CPU port's full-size outputs."""
    import torch

    C = y_port.shape[0]
    y = pf.nufft(xc_dev[:C], nufft_plan=plan, weight=w_dev, device=dev)
    yp = torch.from_numpy(np.ascontiguousarray(y_port.reshape(y.shape)).astype(np.complex64)).to(dev)
    e_fwd = float(torch.linalg.vector_norm(y - yp) / torch.linalg.vector_norm(yp))
    xa = pf.nufft_adj(yp, nufft_plan=plan, weight=w_dev, device=dev)
    xp = torch.from_numpy(np.ascontiguousarray(x_port.reshape(xa.shape)).astype(np.complex64)).to(dev)
    e_adj = float(torch.linalg.vector_norm(xa - xp) / torch.linalg.vector_norm(xp))
    del y, yp, xa, xp
    return {"fwd_rel_l2": float(f"{e_fwd:.3e}"), "adj_rel_l2": float(f"{e_adj:.3e}"), "coils": int(C),
            "checker": "oracle/port.py (C/OpenMP + torch CPU FFT, complex64), full size"}


def adjoint_accuracy(pf, plan, y_dev, w_dev, coord_grid, N, dev, nlines=8, chunk=1 << 16):
    """Relative L2 error of the adjoint (coil 0, weighted input) against the
    FP64 adjoint NUDFT x_r = sum_j w_j y_j exp(+2 pi i k_j . r / N) scaled by
    1/Ghat (norm=None), on `nlines` random x-lines of the image."""
    import torch

    rng = np.random.default_rng(12)
    zs = rng.integers(0, N, nlines)
    ys = rng.integers(0, N, nlines)
    xa = pf.nufft_adj(y_dev[:1], nufft_plan=plan, weight=w_dev, device=dev)[0]
    got = xa[torch.from_numpy(zs).to(dev), torch.from_numpy(ys).to(dev)].cpu().numpy()  # (nlines, N)
    yw = (y_dev[0].reshape(-1) * w_dev.reshape(-1)).cpu().numpy().astype(np.complex128)
    c = coord_grid.astype(np.float64)
    r = np.arange(N) - N // 2
    ref = np.zeros((nlines, N), dtype=np.complex128)
    for s in range(0, c.shape[0], chunk):
        cs = c[s:s + chunk]
        ex = np.exp(2j * math.pi * np.outer(cs[:, 0], r) / N)  # (m, N) along x
        v = yw[s:s + chunk][None, :] * np.exp(2j * math.pi * (np.outer(zs - N // 2, cs[:, 2])
                                                                + np.outer(ys - N // 2, cs[:, 1])) / N)
        ref += v @ ex
    ref /= float(np.prod(plan.os_shape))
    return float(np.linalg.norm(got - ref) / np.linalg.norm(ref))


def accuracy(pf, plan, x, xc_dev, y_dev, coord_grid, N, dev, nsub=1024):
    """Relative L2 error vs FP64 NUDFT on random subsets (coil 0)."""
    import torch

    from oracle import nudft as U

    rng = np.random.default_rng(11)
    M = coord_grid.shape[0]
    sel = np.sort(rng.choice(M, size=nsub, replace=False))
    y0 = y_dev[0].reshape(-1)[torch.from_numpy(sel).to(dev)].cpu().numpy()
    x0 = xc_dev[0].cpu().numpy()
    yt = U.nudft(x0[None], coord_grid[sel], [N, N, N], chunk=256)[0]
    # the weight multiplies y: remove it for the comparison
    return float(np.linalg.norm(y0 - yt) / np.linalg.norm(yt))


def time_width(pf, stage_timer, coord, N, width, xc_dev, w_dev, dev, args):
    """Plan at one kernel width, warm up, time args.steps fwd+adj steps (wall
    clock) and the per-stage HIP-event times over the same region."""
    import torch

    t0 = time.perf_counter()
    plan = pf.plan_nufft(torch.from_numpy(coord), N, width=width, oversamp=2.0)
    plan.to(dev)
    torch.cuda.synchronize()
    plan_s = time.perf_counter() - t0
    log(f"width {width}: plan built in {plan_s:.2f} s")

    def step():
        y = pf.nufft(xc_dev, nufft_plan=plan, weight=w_dev, device=dev)
        return y, pf.nufft_adj(y, nufft_plan=plan, weight=w_dev, device=dev)

    for _ in range(args.warmup):
        step()
    # the timed steps run the fused operators (pf_nufft / pf_nufft_adj), as a caller does
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1000.0 / args.steps
    log(f"width {width}: timed {args.steps} steps: {ms:.3f} ms/step")
    # then the same steps stage by stage (the stage entry points with HIP
    # events around each) for the per-stage times
    with stage_timer() as timer:
        for _ in range(max(1, min(args.steps, 5))):
            step()
        torch.cuda.synchronize()
    stages = {k: float(np.mean(v)) for k, v in timer.summary().items()}
    return plan, ms, stages, plan_s


TRAFFIC_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                            "nufft3d_128cube_32coil_w6_traffic.json")
# kernels of a stage's fused-operator call at the headline workload (the
# stage-by-stage spread also runs k_smax, which the fused adjoint folds into
# the sort)
STAGE_KERNELS = {"spread": ("k_tile_max", "k_spread_mfma", "k_spread_reduce"), "interp": ("k_gather_col",)}


def measured_traffic(stage, args):
    """HBM bytes per call of a stage at the headline workload, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE passes (calibrated against a
    device copy of known size); None for other workloads."""
    if args.width != 6 or stage not in STAGE_KERNELS or not os.path.exists(TRAFFIC_FILE):
        return None
    with open(TRAFFIC_FILE) as f:
        k = json.load(f)["kernels"]
    total = 0
    for prefix in STAGE_KERNELS[stage]:
        hits = [v for n, v in k.items() if n.split("<")[0] == prefix]
        if len(hits) != 1:
            return None
        total += hits[0]["read_bytes"] + hits[0]["write_bytes"]
    return int(total)


def _pmc_rows(d, counter):
    """(kernel name, counter value) per dispatch of one rocprofv3 csv pass."""
    import csv

    path = os.path.join(d, "run_counter_collection.csv")
    with open(path) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] == counter:
                name = r["Kernel_Name"].replace("(anonymous namespace)::", "").replace("void ", "").split("(")[0]
                yield name.strip(), float(r["Counter_Value"])


def summarize_pmc(fetch_dir, write_dir, copy_bytes=2 << 30, N=128, C=32, M=8192 * 256, G=256 ** 3):
    """Per-dispatch HBM bytes of every kernel of the headline run, from a
    FETCH_SIZE and a WRITE_SIZE pass (KiB per dispatch), each scaled by what
    the same counter reads for the 2 GiB device copy of stream_copy
    (`__amd_rocclr_copyBuffer`: copy_bytes read and copy_bytes written)."""
    import collections

    per = {}
    for key, d, counter in (("read", fetch_dir, "FETCH_SIZE"), ("write", write_dir, "WRITE_SIZE")):
        vals = collections.defaultdict(list)
        for name, v in _pmc_rows(d, counter):
            vals[name].append(v)
        copy = [v for n, vs in vals.items() if "copyBuffer" in n for v in vs]
        if not copy:
            raise SystemExit(f"{d}: no device-copy dispatch to calibrate {counter} against")
        kib = max(copy)  # the 2 GiB copy is the largest copyBuffer dispatch
        per[key] = (copy_bytes / (kib * 1024.0), kib, vals)
    alg = C * G * 8 + C * M * 8 + M * 16  # grid + samples + coordinates and weight
    kernels = {}
    for name in sorted(set(per["read"][2]) | set(per["write"][2])):
        if "copyBuffer" in name:
            continue
        rv, wv = per["read"][2].get(name, []), per["write"][2].get(name, [])
        kernels[name] = {
            "read_bytes": int(round(np.mean(rv) * 1024 * per["read"][0])) if rv else None,
            "write_bytes": int(round(np.mean(wv) * 1024 * per["write"][0])) if wv else None,
            "dispatches": max(len(rv), len(wv)),
            "algorithmic_bytes": alg if name.split("<")[0] in ("k_gather_col", "k_spread_mfma") else None,
        }
    out = {"calibration": {"fetch_factor": round(per["read"][0], 4), "write_factor": round(per["write"][0], 4),
                           "copy_fetch_kib": per["read"][1], "copy_write_kib": per["write"][1]},
           "kernels": kernels}
    with open(TRAFFIC_FILE, "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")
    return out


def _time(fn, dev, steps, warmup=2):
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize(dev)
    return (time.perf_counter() - t0) * 1000.0 / steps


def stream_copy(dev, nbytes=2 << 30, reps=10):
    """Achievable HBM bandwidth of a plain device copy (read + write bytes / s)."""
    import torch

    src = torch.empty(nbytes // 4, dtype=torch.float32, device=dev).fill_(1.0)
    dst = torch.empty_like(src)
    for _ in range(2):
        dst.copy_(src)
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize(dev)
    gbs = 2 * nbytes * reps / (time.perf_counter() - t0) / 1e9
    del src, dst
    torch.cuda.empty_cache()
    return round(gbs, 1)


def bench_uneven_batch(pf, dev, steps):
    """64 2-D slices, matrix sizes from {128, 192, 256}^2, 8 coils, golden-angle
    radial 16..256 spokes x 256 samples, oversampling 2, width 4: fwd+adj of
    every slice as one graph replay vs the per-slice drop-in calls."""
    import torch

    from pyfourier_amd import demo

    rng = np.random.default_rng(5)
    sizes = [int(v) for v in rng.choice([128, 192, 256], 64)]
    spokes = [int(v) for v in rng.integers(16, 257, 64)]
    coords = [demo.golden_angle_radial_2d(n, 256) for n in spokes]
    C = 8
    batch = pf.plan_nufft_batch(coords, sizes, nbatch=C, width=4, oversamp=2.0, device=dev, autotune=True)
    for buf in batch.images:
        buf.copy_(torch.randn(buf.shape, dtype=torch.complex64, device=dev))
    plans = batch.plans
    imgs = [b[:, 0] for b in batch.images]

    def plain():
        for p, x in zip(plans, imgs):
            pf.nufft_adj(pf.nufft(x, nufft_plan=p, device=dev), nufft_plan=p, device=dev)

    def batched():
        batch.forward()
        batch.adjoint()

    t_plain = _time(plain, dev, steps)
    t_batch = _time(batched, dev, steps)
    return {"workload": "64 uneven 2D slices {128,192,256}^2 x 8 coils, radial 16-256 spokes x 256, W=4",
            "samples": int(sum(spokes) * 256), "groups": len(batch.groups), "plain_ms": round(t_plain, 3),
            "batched_ms": round(t_batch, 3), "speedup": round(t_plain / t_batch, 2),
            "autotuned": [{"tile": list(g.tile) if g.tile else None, "chunk": g.chunk} for g in batch.groups]}


def bench_subspace(pf, dev, steps):
    """500-frame variable-density spiral (2048 samples per frame), 256^2 x 8
    coils, 4-coefficient T2 basis: the fused subspace fwd+adj (one call)
    vs the plain path that expands the 500 frame images and runs the
    multi-frame NUFFT and its adjoint, then projects back onto the basis."""
    import torch

    from pyfourier_amd import demo

    rng = np.random.default_rng(6)
    F, K, N, C = 500, 4, 256, 8
    base = demo.variable_density_spiral_2d(F, 2048, N)
    coord = base.reshape(F, 1, 2048, 2)
    basis = torch.from_numpy(demo.t2_subspace_basis(rng, F, K)).to(dev)
    plan = pf.plan_nufft(torch.from_numpy(coord), N, width=4, oversamp=2.0).to(dev)
    x = torch.randn((C, K, N, N), dtype=torch.complex64, device=dev)

    def fused():
        y = pf.nufft(x, nufft_plan=plan, basis=basis, device=dev)
        pf.nufft_adj(y, nufft_plan=plan, basis=basis, device=dev)

    def plain():
        xf = torch.einsum("fk,ckyx->cfyx", basis, x)
        y = pf.nufft(xf, nufft_plan=plan, device=dev)
        xa = pf.nufft_adj(y, nufft_plan=plan, device=dev)
        torch.einsum("fk,cfyx->ckyx", basis.conj(), xa)

    t_fused = _time(fused, dev, steps)
    t_plain = _time(plain, dev, steps)
    return {"workload": "500-frame spiral x 2048, 256^2 x 8 coils, K=4 subspace, W=4",
            "plain_ms": round(t_plain, 3), "fused_ms": round(t_fused, 3), "speedup": round(t_plain / t_fused, 2)}


def bench_autograd_3d(pf, dev, steps):
    """3-D radial 128^3 x 8 coils (golden-means 8192 x 256), W=4, sigma 2: the
    same L2 data-consistency loss with image and coordinate gradients as
    bench_autograd, through the autograd function vs the plain calls (adjoint
    for the image, one forward NUFFT of the ramp-weighted image per axis,
    here three)."""
    import torch

    from pyfourier_amd import demo

    N, C = 128, 8
    coord = demo.golden_means_radial_3d(8192, 256)
    plan = pf.plan_nufft(torch.from_numpy(coord), N, width=4, oversamp=2.0).to(dev)
    x0 = torch.randn((C, N, N, N), dtype=torch.complex64, device=dev)
    target = pf.nufft(x0, nufft_plan=plan, device=dev)
    x = (x0 * 0.9).requires_grad_()
    ct = torch.from_numpy(coord).to(dev).requires_grad_()

    def autograd():
        x.grad = None
        ct.grad = None
        loss = (pf.nufft(x, ct, nufft_plan=plan, device=dev) - target).abs().pow(2).sum()
        loss.backward()

    r = torch.arange(N, device=dev, dtype=torch.float32) - N // 2
    ramps = [(-2j * math.pi / N) * r.reshape(1, 1, 1, N), (-2j * math.pi / N) * r.reshape(1, 1, N, 1),
             (-2j * math.pi / N) * r.reshape(1, N, 1, 1)]
    unit = N / (2 * math.pi)

    def plain():
        xd = x.detach()
        res = pf.nufft(xd, nufft_plan=plan, device=dev) - target
        gx = 2 * pf.nufft_adj(res, nufft_plan=plan, device=dev)
        gk = [(2 * res.conj() * pf.nufft(ramp * xd, nufft_plan=plan, device=dev)).real.sum(0) * unit
              for ramp in ramps]
        return gx, gk

    t_auto = _time(autograd, dev, steps)
    t_plain = _time(plain, dev, steps)
    return {"workload": "3-D radial 128^3 x 8 coils, golden-means 8192x256, image + coordinate gradients, W=4",
            "plain_ms": round(t_plain, 3), "autograd_ms": round(t_auto, 3), "speedup": round(t_plain / t_auto, 2)}


def bench_cfg1_latency(pf, dev, steps):
    """2-D plain NUFFT, the reference's CPU-runnable case: 128^2 complex64
    image (seed 0), golden-angle radial 64 spokes x 256 samples in radians,
    one coil, oversampling 2, width 4.  Latency of one fwd+adj (drop-in
    calls, wall clock), the same pair replayed from a captured HIP graph,
    and parity of both directions against the float64 restatement plus the
    forward's error against the FP64 NUDFT, at this size."""
    import torch

    from oracle import nudft as U
    from oracle import restate as R
    from pyfourier_amd import demo

    rng = np.random.default_rng(0)
    N = 128
    coord = demo.golden_angle_radial_2d(64, 256)
    x = demo.complex_gaussian(rng, (N, N))
    plan = pf.plan_nufft(torch.from_numpy(coord), N, width=4, oversamp=2.0).to(dev)
    xd = torch.from_numpy(x).to(dev)
    yk = torch.from_numpy(demo.complex_gaussian(rng, coord.shape[:-1])).to(dev)

    def pair():
        pf.nufft(xd, nufft_plan=plan, device=dev)
        pf.nufft_adj(yk, nufft_plan=plan, device=dev)

    n = max(50, steps * 20)
    t_call = _time(pair, dev, n, warmup=5)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        y_g = pf.nufft(xd, nufft_plan=plan, device=dev)
        x_g = pf.nufft_adj(yk, nufft_plan=plan, device=dev)
    t_graph = _time(g.replay, dev, n, warmup=5)
    g.replay()
    torch.cuda.synchronize(dev)
    P = R.Plan(coord, N, width=4, oversamp=2.0)
    yr = R.nufft(x[None, None], P)[0, 0]
    xr = R.nufft_adj(yk.cpu().numpy().reshape(1, 1, -1), P)[0, 0]
    e_f = float(np.linalg.norm(y_g.cpu().numpy().ravel() - yr) / np.linalg.norm(yr))
    e_a = float(np.linalg.norm(x_g.cpu().numpy() - xr) / np.linalg.norm(xr))
    yt = U.nudft(x[None], P.coord.reshape(-1, 2), [N, N])[0]
    e_n = float(np.linalg.norm(y_g.cpu().numpy().ravel() - yt) / np.linalg.norm(yt))
    return {"workload": "2-D 128^2, golden-angle radial 64x256, 1 coil, sigma 2, W=4",
            "fwd_adj_us": round(t_call * 1000.0, 1), "fwd_adj_graph_us": round(t_graph * 1000.0, 1),
            "parity_vs_restatement": {"fwd_rel_l2": float(f"{e_f:.2e}"), "adj_rel_l2": float(f"{e_a:.2e}")},
            "fwd_rel_err_vs_nudft": float(f"{e_n:.2e}")}


def _gram_exact_rows(x0, coord_grid, w, N, rows, chunk=2048):
    """FP64 A^H W A x on image rows (2-D, one coil): A the exact NUDFT."""
    from oracle import nudft as U

    y = U.nudft(x0[None], coord_grid, [N, N], chunk=chunk)[0] * w.reshape(-1)
    c = coord_grid.astype(np.float64)
    r = np.arange(N) - N // 2
    out = np.zeros((len(rows), N), dtype=np.complex128)
    for s in range(0, c.shape[0], 1 << 15):
        cs = c[s:s + (1 << 15)]
        ex = np.exp(2j * math.pi * np.outer(cs[:, 0], r) / N)
        v = y[s:s + (1 << 15)][None, :] * np.exp(2j * math.pi * np.outer(np.asarray(rows) - N // 2, cs[:, 1]) / N)
        out += v @ ex
    return out


def bench_toeplitz(pf, dev, steps):
    """2-D Toeplitz Gram: 48 variable-density spiral interleaves x 4096
    samples, 256^2 x 16 coils, |k| density weights, kernel on the 512^2
    grid.  sparse_gram (pad -> FFT -> per-voxel kernel multiply -> IFFT ->
    crop, one native call) vs the explicit normal operator
    nufft_adj(w * nufft(x)) it stands for, at the nufft defaults (W=4,
    oversampling 1.25), with their relative difference (the two kernel
    approximations) and the multiply's own byte count
    2*C*(2N)^2*8 + (2N)^2*8."""
    import torch

    from pyfourier_amd import demo

    N, C = 256, 16
    coord = demo.variable_density_spiral_2d(48, 4096, N)
    kr = np.sqrt((coord.astype(np.float64) ** 2).sum(-1))
    w = torch.from_numpy((kr / kr.max()).astype(np.float32)).to(dev)
    ct = torch.from_numpy(coord)
    gm = pf.plan_spgram(ct, N, weight=w, oversamp=1.25, device=dev)
    plan = pf.plan_nufft(ct, N, width=4, oversamp=1.25).to(dev)
    x = torch.randn((C, N, N), dtype=torch.complex64, device=dev)

    def gram():
        return pf.sparse_gram(x, gm, device=dev)

    def explicit():
        return pf.nufft_adj(pf.nufft(x, nufft_plan=plan, device=dev) * w, nufft_plan=plan, device=dev)

    a, b = gram(), explicit()
    diff = float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))
    t_g = _time(gram, dev, steps)
    t_e = _time(explicit, dev, steps)
    # both against the FP64 normal operator (1/Ghat) A^H W A x (exact NUDFT,
    # Ghat = ceil(1.25 N)^2 as both are planned) on 8 random rows of coil 0
    rows = sorted(np.random.default_rng(13).choice(N, 8, replace=False).tolist())
    cg = plan.interpolator.coord.reshape(-1, 2).numpy()
    ref = _gram_exact_rows(x[0].cpu().numpy(), cg, w.cpu().numpy(), N, rows) / float(np.prod(plan.os_shape))
    ridx = torch.tensor(rows, device=dev)
    e_g = float(np.linalg.norm(a[0, ridx].cpu().numpy() - ref) / np.linalg.norm(ref))
    e_e = float(np.linalg.norm(b[0, ridx].cpu().numpy() - ref) / np.linalg.norm(ref))
    return {"workload": "2-D Toeplitz Gram, 256^2 x 16 coils, 48 spirals x 4096, kernel on 512^2",
            "gram_ms": round(t_g, 3), "explicit_ms": round(t_e, 3), "speedup": round(t_e / t_g, 2),
            "rel_diff_vs_explicit": round(diff, 5),
            "gram_rel_err_vs_fp64_AhWA": float(f"{e_g:.2e}"),
            "explicit_rel_err_vs_fp64_AhWA": float(f"{e_e:.2e}"),
            "fp64_check": "coil 0, 8 image rows, exact NUDFT normal operator",
            "multiply_bytes": 2 * C * (2 * N) ** 2 * 8 + (2 * N) ** 2 * 8}


def bench_cartesian(pf, dev, steps):
    """Cartesian operators: masked centred FFT fwd+adj (fftn + ifftn) of
    256^2 x 32 coils with an R=4 mask, a 100-contrast K=4 subspace version
    (256^2 x 16 coils), a 3-D 256^3 x 8 coils masked FFT, and the dense
    subspace Gram; each against the same operator composed from torch.fft
    (rocFFT) with explicit shifts, basis contraction and mask multiply, as
    the reference computes it (src/pyfourier/_fft/_fft.py:101-151).
    eff_GB_s counts the operator's own input + output bytes once per call."""
    import torch

    from pyfourier_amd import demo

    rng = np.random.default_rng(8)
    out = {}

    def torch_fwd(x, mask, basis, d):
        ax = tuple(range(-d, 0))
        k = torch.fft.fftshift(torch.fft.fftn(torch.fft.ifftshift(x, dim=ax), dim=ax), dim=ax)
        if basis is not None:
            k = torch.einsum("tk,ck...->ct...", basis.conj(), k)
        return k * mask

    def torch_adj(y, mask, basis, d):
        ax = tuple(range(-d, 0))
        y = y * mask
        if basis is not None:
            y = torch.einsum("tk,ct...->ck...", basis, y)
        return torch.fft.fftshift(torch.fft.ifftn(torch.fft.ifftshift(y, dim=ax), dim=ax), dim=ax)

    def case(name, shape, C, T, K, mframes):
        d = len(shape)
        mask = torch.from_numpy((rng.random((mframes,) + shape) < 0.25).astype(np.float32)).to(dev)
        basis = torch.from_numpy(demo.t2_subspace_basis(rng, T, K)).to(dev) if K else None
        nimg = K if K else T
        x = torch.randn((C, nimg) + shape, dtype=torch.complex64, device=dev)
        if nimg == 1:
            x = x[:, 0]
        plan = pf.plan_fft(mask, shape)

        def native():
            y = pf.fftn(d, x, plan=plan, basis=basis, device=dev)
            pf.ifftn(d, y, plan=plan, basis=basis, device=dev)

        xs = x if nimg > 1 else x[:, None]
        mk = mask if mframes > 1 else mask[0]

        def plain():
            y = torch_fwd(xs, mk, basis, d)
            torch_adj(y, mk, basis, d)

        t_n = _time(native, dev, steps)
        t_p = _time(plain, dev, steps)
        y = pf.fftn(d, x, plan=plan, basis=basis, device=dev)
        err = float((y.reshape(C, T, *shape) - torch_fwd(xs, mk, basis, d)).norm() / y.norm())
        nbytes = 2 * 8 * C * (nimg + T) * int(np.prod(shape))
        out[name] = {"native_ms": round(t_n, 3), "torch_fft_ms": round(t_p, 3), "speedup": round(t_p / t_n, 2),
                     "eff_GB_s": round(nbytes / (t_n * 1e-3) / 1e9, 1), "rel_err_vs_torch": float(f"{err:.2e}")}
        del x, y
        torch.cuda.empty_cache()

    case("masked_2d_256_32coils", (256, 256), 32, 1, 0, 1)
    case("subspace_2d_256_16coils_T100_K4", (256, 256), 16, 100, 4, 100)
    case("masked_3d_256_8coils", (256, 256, 256), 8, 1, 0, 1)

    # dense Gram with the subspace (per-voxel 4x4) vs ifftn(fftn(.)) in torch
    shape, C, T, K = (256, 256), 16, 100, 4
    mask = torch.from_numpy((rng.random((T,) + shape) < 0.25).astype(np.float32)).to(dev)
    basis = torch.from_numpy(demo.t2_subspace_basis(rng, T, K)).to(dev)
    gm = pf.plan_gram(mask, shape, basis=basis, device=dev)
    x = torch.randn((C, K) + shape, dtype=torch.complex64, device=dev)
    t_n = _time(lambda: pf.gram(x, gm, device=dev), dev, steps)
    t_p = _time(lambda: torch_adj(torch_fwd(x, mask, basis, 2), mask, basis, 2), dev, steps)
    out["gram_subspace_2d_256_16coils_T100_K4"] = {"native_ms": round(t_n, 3), "torch_fft_ms": round(t_p, 3),
                                                  "speedup": round(t_p / t_n, 2)}
    return out


def bench_voronoi(pf, dev, steps):
    """Voronoi density compensation on the GPU (pyfourier_amd.voronoi): the
    reference's own use, one 4096-sample variable-density spiral interleaf
    rotated into a 48-shot full disk (196,608 points, 48 of them the k-space
    centre), and 196,608 uniform random points (cell areas only)."""
    import torch

    from pyfourier_amd import demo

    k = torch.from_numpy(demo.variable_density_spiral_2d(1, 4096, 256)).to(dev)
    pts = torch.rand((196608, 2), dtype=torch.float64, device=dev, generator=torch.Generator(device=dev).manual_seed(5))
    n = max(1, min(steps, 5))
    t_disk = _time(lambda: pf.voronoi.voronoi(k, nshots=48, fullspoke=True), dev, n, warmup=1)
    t_uni = _time(lambda: pf.voronoi.voronoi_areas(pts, dev), dev, n, warmup=1)
    return {"workload": "2-D Voronoi DCF: spiral 48 x 4096 full disk; 196,608 uniform points (areas)",
            "full_disk_ms": round(t_disk, 1), "uniform_areas_ms": round(t_uni, 1)}


def bench_grog_sparse(pf, dev, steps):
    """GROG of a 16-slice 2-D radial acquisition (402 spokes x 512 samples,
    32 coils, per-slice operators from 24^2 calibration patches, 11 shift
    steps) onto a 256^2 grid, then the sparse FFT of the gridded points
    (16 x 32 images, fwd+adj).  grog_ms: the whole call (host operator
    training and fractional powers included); apply_ms: the native
    coil-operator kernel alone on the same data (HIP events), priced against
    HBM (data in + out once) and FP32 (8 flops per complex MAC, ndim nc^2
    MACs per sample).  The sparse FFT is compared with torch.fft plus an
    index gather / index_add zero-fill."""
    import ctypes

    import torch

    from pyfourier_amd import _grog, _lib, demo

    rng = np.random.default_rng(9)
    S, nc, nv, ns, N = 16, 32, 402, 512, 256
    coord = (demo.golden_angle_radial_2d(nv, ns)[None] * (N / (2 * math.pi))).astype(np.float32)  # grid units
    calib = torch.randn((nc, S, 24, 24), dtype=torch.complex64)
    x = torch.randn((S, nc, 1, nv, ns), dtype=torch.complex64, device=dev)
    out = {}
    t_grog = _time(lambda: pf.grog_interp(x, calib, coord, N), dev, max(2, steps // 2), warmup=1)
    y, idx, w = pf.grog_interp(x, calib, coord, N)

    # the apply kernel alone
    npts = nv * ns
    cal = calib.numpy().astype(np.complex128)
    cal = cal / np.linalg.norm(cal)
    ops = _grog.train_operators(cal, 2, 0.01)
    deltas = (np.arange(11) - 5) / 10
    tab = torch.from_numpy(np.stack([_grog.operator_powers(g, deltas) for g in ops]).astype(np.complex64)).to(dev)
    L = _lib.lib()
    cd = torch.from_numpy(coord.reshape(-1, 2)).to(dev)
    lut = torch.empty(npts, dtype=torch.int32, device=dev)
    ind = torch.empty((npts, 2), dtype=torch.int32, device=dev)
    flat = torch.empty(npts, dtype=torch.int64, device=dev)
    sh = (ctypes.c_int * 2)(N, N)
    st = _lib.stream_of(dev)
    _lib.check(L.pf_grog_index(_lib.ptr(cd), 2, sh, npts, 11, _lib.ptr(lut), _lib.ptr(ind), _lib.ptr(flat), st))
    xin = x.reshape(S, nc, npts).contiguous()
    yo = torch.empty_like(xin)

    def apply():
        _lib.check(L.pf_grog_apply(_lib.ptr(xin), _lib.ptr(yo), S, nc, npts, 2, 11, _lib.ptr(tab), S,
                                   _lib.ptr(lut), st))

    for _ in range(2):
        apply()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(torch.cuda.current_stream(dev))
    for _ in range(steps):
        apply()
    e1.record(torch.cuda.current_stream(dev))
    torch.cuda.synchronize(dev)
    t_apply = e0.elapsed_time(e1) / steps
    err = float((yo.reshape(y.shape) - y).norm() / y.norm())
    nbytes = 2 * S * nc * npts * 8
    flops = 8.0 * 2 * nc * nc * S * npts
    out["grog_2d_16slices_32coils_402x512"] = {
        "grog_ms": round(t_grog, 3), "apply_ms": round(t_apply, 3),
        "apply_GB_s": round(nbytes / (t_apply * 1e-3) / 1e9, 1),
        "apply_TFLOP_s": round(flops / (t_apply * 1e-3) / 1e12, 2),
        "apply_frac_fp32_peak": round(flops / (t_apply * 1e-3) / 157.3e12, 3),
        "apply_vs_call_rel_err": float(f"{err:.1e}")}

    # sparse FFT of the gridded samples: (16 slices x 32 coils) images of 256^2
    plan = pf.plan_spfft(idx.cpu(), N)
    img = torch.randn((S * nc, N, N), dtype=torch.complex64, device=dev)
    flat_t = (idx[..., 1].reshape(-1).to(dev) * N + idx[..., 0].reshape(-1).to(dev))

    def native():
        k = pf.sparse_fftn(img, plan=plan, device=dev)
        pf.sparse_ifftn(k, plan=plan, device=dev)

    def plain():
        ax = (-2, -1)
        k = torch.fft.fftshift(torch.fft.fftn(torch.fft.ifftshift(img, dim=ax), dim=ax), dim=ax)
        ks = k.reshape(S * nc, -1)[:, flat_t]
        g = torch.zeros((S * nc, N * N), dtype=torch.complex64, device=dev).index_add_(1, flat_t, ks)
        g = g.reshape(S * nc, N, N)
        torch.fft.fftshift(torch.fft.ifftn(torch.fft.ifftshift(g, dim=ax), dim=ax), dim=ax)

    t_n = _time(native, dev, steps)
    t_p = _time(plain, dev, steps)
    out["sparse_fft_2d_256_512images_206k_points"] = {"native_ms": round(t_n, 3), "torch_ms": round(t_p, 3),
                                                      "speedup": round(t_p / t_n, 2)}
    return out


def bench_autograd(pf, dev, steps):
    """256^2 x 8 coils, 16 spiral interleaves x 4096 samples, 10 ms readout,
    8 B0 time segments (smooth +-100 Hz field map): one L2 data-consistency
    loss with gradients for the image and the k-space coordinates, through
    the autograd function (KB-derivative gather) vs the plain calls that give
    the same gradients (adjoint for the image, one forward NUFFT of the
    ramp-weighted image per axis for the coordinates)."""
    import torch

    from pyfourier_amd import demo

    rng = np.random.default_rng(7)
    N, C, L = 256, 8, 8
    coord = demo.variable_density_spiral_2d(16, 4096, N)
    zmap = torch.from_numpy(demo.smooth_field_map(rng, (N, N), 100.0))
    t = torch.linspace(0, 10e-3, 4096)
    plan = pf.plan_nufft(torch.from_numpy(coord), N, width=4, oversamp=2.0, zmap=zmap, T=t, L=L).to(dev)
    x0 = torch.randn((C, N, N), dtype=torch.complex64, device=dev)
    target = pf.nufft(x0, nufft_plan=plan, device=dev)
    x = (x0 * 0.9).requires_grad_()
    ct = torch.from_numpy(coord).to(dev).requires_grad_()  # trajectory parameters live on the GPU

    def autograd():
        x.grad = None
        ct.grad = None
        loss = (pf.nufft(x, ct, nufft_plan=plan, device=dev) - target).abs().pow(2).sum()
        loss.backward()

    r_ax = [torch.arange(N, device=dev, dtype=torch.float32) - N // 2] * 2
    ramps = [(-2j * math.pi / N) * r_ax[1].reshape(1, 1, N), (-2j * math.pi / N) * r_ax[0].reshape(1, N, 1)]
    unit = N / (2 * math.pi)

    def plain():
        xd = x.detach()
        r = pf.nufft(xd, nufft_plan=plan, device=dev) - target
        gx = 2 * pf.nufft_adj(r, nufft_plan=plan, device=dev)
        gk = [(2 * r.conj() * pf.nufft(ramp * xd, nufft_plan=plan, device=dev)).real.sum(0) * unit
              for ramp in ramps]
        return gx, gk

    t_auto = _time(autograd, dev, steps)
    t_plain = _time(plain, dev, steps)
    return {"workload": "256^2 x 8 coils, 16 spirals x 4096, 8 B0 segments, image + coordinate gradients, W=4",
            "plain_ms": round(t_plain, 3), "autograd_ms": round(t_auto, 3), "speedup": round(t_plain / t_auto, 2)}


def bench_default_params(pf, dev, steps):
    """The headline workload (3-D 128^3 x 32 coils, golden-means 8192 x 256,
    |k|^2 weights) at the reference's default kernel parameters, width 4 and
    oversampling 1.25 (a 160^3 grid, register 16 x 10-point FFT passes):
    fwd+adj (fused drop-in calls, wall clock), the stage-by-stage times (HIP
    events) and the forward's error against the FP64 NUDFT on 1024 samples
    of coil 0.  Then the 2-D B0 workload's forward and adjoint at the
    field-map defaults (6 segments, width 4, oversampling 1.25)."""
    import torch

    from pyfourier_amd._nufft import stage_timer

    N, C = 128, 32
    x, xc, coord, w = make_inputs(N, C, 8192, 256)
    xc_dev = torch.from_numpy(xc).to(dev)
    w_dev = torch.from_numpy(w).to(dev)
    plan = pf.plan_nufft(torch.from_numpy(coord), N, width=4, oversamp=1.25).to(dev)

    def step():
        y = pf.nufft(xc_dev, nufft_plan=plan, weight=w_dev, device=dev)
        return pf.nufft_adj(y, nufft_plan=plan, weight=w_dev, device=dev)

    ms = _time(step, dev, max(5, steps), warmup=3)
    with stage_timer() as timer:
        for _ in range(5):
            step()
        torch.cuda.synchronize(dev)
    stages = {k: round(float(np.mean(v)), 4) for k, v in timer.summary().items()}
    y0 = pf.nufft(xc_dev[:1], nufft_plan=plan, device=dev)
    err = accuracy(pf, plan, x, xc_dev, y0, plan.interpolator.coord.reshape(-1, 3).numpy(), N, dev)
    out = {"nufft3d_128cube_32coil_w4_os1.25": {
        "grid": list(plan.interpolator.os_shape), "ms_per_fwd_adj": round(ms, 3),
        "fwd_adj_per_s": round(1000.0 / ms, 1), "stages_ms": stages, "fwd_rel_err_vs_nudft": float(f"{err:.2e}")}}
    del plan, xc_dev, w_dev
    torch.cuda.empty_cache()

    # the B0 workload's operators (256^2 x 8 coils, 16 spirals x 4096, 10 ms,
    # +-100 Hz map) at the reference's field-map defaults: 6 segments, width
    # 4, oversampling 1.25 (320-point rows with segments)
    from pyfourier_amd import demo

    rng = np.random.default_rng(7)
    N2, C2 = 256, 8
    coord2 = demo.variable_density_spiral_2d(16, 4096, N2)
    zmap = torch.from_numpy(demo.smooth_field_map(rng, (N2, N2), 100.0))
    t = torch.linspace(0, 10e-3, 4096)
    x2 = torch.from_numpy(demo.complex_gaussian(rng, (C2, N2, N2))).to(dev)
    plan2 = pf.plan_nufft(torch.from_numpy(coord2), N2, zmap=zmap, T=t).to(dev)
    y2 = pf.nufft(x2, nufft_plan=plan2, device=dev)
    n = max(20, steps * 4)
    t_f = _time(lambda: pf.nufft(x2, nufft_plan=plan2, device=dev), dev, n, warmup=3)
    t_a = _time(lambda: pf.nufft_adj(y2, nufft_plan=plan2, device=dev), dev, n, warmup=3)
    out["b0_2d_256_8coil_spiral16_defaults"] = {
        "segments": int(plan2.zmap_s_kernel.shape[0]), "grid": list(plan2.interpolator.os_shape),
        "fwd_ms": round(t_f, 4), "adj_ms": round(t_a, 4)}
    return out


def bench_fft_sizes(pf, dev, steps):
    """Fused fwd+adj (drop-in calls, wall clock) and the pad+FFT / IFFT+crop
    stages (HIP events) on grids whose FFT lines run the register four-step
    passes of 16 x Q or 8 x Q points (160, 200, 224, 320, 400, 480, 640) and
    the 256-point ones: random coordinates (2^18 per 2-D/3-D plan), complex
    Gaussian images, width 4."""
    import torch

    from pyfourier_amd._nufft import stage_timer

    cases = [((256, 256), 32, 1.25), ((320, 320), 32, 1.25), ((384, 384), 32, 1.25), ((512, 512), 32, 1.25),
             ((160, 160, 160), 8, 1.25), ((112, 112, 112), 16, 2.0), ((160, 160, 160), 8, 2.0),
             ((128, 128, 128), 8, 2.0)]
    out = {}
    for shape, C, os_ in cases:
        d = len(shape)
        rng = np.random.default_rng(1)
        coord = torch.from_numpy(rng.uniform(-np.pi, np.pi, (1 << 18, d)).astype(np.float32))
        plan = pf.plan_nufft(coord, shape, width=4, oversamp=os_).to(dev)
        g = torch.Generator(device=dev)
        g.manual_seed(1)
        x = torch.randn((C,) + shape, dtype=torch.complex64, device=dev, generator=g)

        def step():
            return pf.nufft_adj(pf.nufft(x, nufft_plan=plan, device=dev), nufft_plan=plan, device=dev)

        ms = _time(step, dev, max(5, steps), warmup=2)
        with stage_timer() as t:
            for _ in range(3):
                step()
            torch.cuda.synchronize(dev)
        st = t.summary()
        key = "x".join(str(n) for n in shape) + f"_{C}coil_os{os_}"
        out[key] = {"grid": list(plan.interpolator.os_shape), "fwd_adj_ms": round(ms, 4),
                    "pad_fft_ms": round(float(np.mean(st["pad_fft"])), 4),
                    "ifft_crop_ms": round(float(np.mean(st["ifft_crop"])), 4)}
        del plan, x
        torch.cuda.empty_cache()
    return out


EXTRAS = (("cfg1_latency", bench_cfg1_latency), ("batched_uneven", bench_uneven_batch), ("batched_subspace", bench_subspace),
          ("autograd_b0", bench_autograd), ("autograd_3d", bench_autograd_3d), ("toeplitz", bench_toeplitz),
          ("cartesian", bench_cartesian),
          ("grog_sparse", bench_grog_sparse), ("voronoi", bench_voronoi), ("default_params", bench_default_params),
          ("fft_sizes", bench_fft_sizes))


def main():
    args = parse()
    if args.summarize_pmc:
        out = summarize_pmc(*args.summarize_pmc)
        log(f"wrote {TRAFFIC_FILE}: {len(out['kernels'])} kernels")
        return
    if args.gpus != 1:
        print(
            f"bench.py measures one GPU only (this build has no multi-GPU path); got --gpus {args.gpus}",
            file=sys.stderr,
        )
        sys.exit(2)
    import torch

    import pyfourier_amd as pf
    from pyfourier_amd._nufft import stage_timer

    dev = torch.device("cuda", 0)
    if args.only is not None:
        fns = dict(EXTRAS)
        if args.only not in fns:
            print(f"--only takes one of {sorted(fns)}", file=sys.stderr)
            sys.exit(2)
        print(json.dumps({args.only: fns[args.only](pf, dev, args.steps)}), flush=True)
        return
    N, C, NSP, NS = (64, 8, 2048, 128) if args.small else (128, 32, 8192, 256)
    x, xc, coord, w = make_inputs(N, C, NSP, NS)
    log(f"inputs ready: {N}^3 x {C} coils, {NSP}x{NS} samples")
    xc_dev = torch.from_numpy(xc).to(dev)
    w_dev = torch.from_numpy(w).to(dev)
    plan, ms, stages, plan_s = time_width(pf, stage_timer, coord, N, args.width, xc_dev, w_dev, dev, args)
    cmp = None
    if args.compare_width:
        plan_c, ms_c, stages_c, _ = time_width(pf, stage_timer, coord, N, args.compare_width, xc_dev, w_dev, dev,
                                              args)
        cmp = {"width": args.compare_width, "ms_per_step": round(ms_c, 3), "value": round(1000.0 / ms_c, 3),
               "stages_ms": {k: round(v, 4) for k, v in stages_c.items()}}

    G3 = int(np.prod(plan.os_shape))
    N3 = N ** 3
    M = NSP * NS
    stage_info = {}
    for k, v in stages.items():
        by = stage_bytes(k, C, N3, G3, M, 3)
        stage_info[k] = {"ms": round(v, 4), "GB_s": round(by / (v * 1e-3) / 1e9, 1),
                         "frac_hbm": round(by / (v * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    # the dominant kernel: pad+FFT and IFFT+crop are three FFT pass kernels
    # each (none longer than ~1 ms), so the longest single kernel is the
    # gather or the spread
    own = {k: v for k, v in stages.items() if k not in ("pad_fft", "ifft_crop")}
    dom = max(own, key=own.get)
    dom_bytes = stage_bytes(dom, C, N3, G3, M, 3)
    achieved = dom_bytes / (stages[dom] * 1e-3) / 1e9
    traffic = measured_traffic(dom, args) if not args.small else None

    acc = None
    if not args.no_accuracy:
        coord_grid = plan.interpolator.coord.reshape(-1, 3).numpy()
        y_nw = pf.nufft(xc_dev[:1], nufft_plan=plan, device=dev)
        acc = {"fwd_rel_err_vs_nudft": accuracy(pf, plan, x, xc_dev, y_nw, coord_grid, N, dev),
               "width": args.width, "samples_checked": 1024}
        if cmp is not None:
            y_nw = pf.nufft(xc_dev[:1], nufft_plan=plan_c, device=dev)
            cmp["fwd_rel_err_vs_nudft"] = accuracy(pf, plan_c, x, xc_dev, y_nw, coord_grid, N, dev)

    if acc is not None:
        log(f"accuracy vs NUDFT: {acc}")
    if acc is not None:
        acc["adj_rel_err_vs_nudft"] = adjoint_accuracy(pf, plan, y_nw, w_dev, coord_grid, N, dev)
        acc["adj_lines_checked"] = 8
        log(f"adjoint accuracy vs NUDFT: {acc['adj_rel_err_vs_nudft']:.3e}")
    cpu = None
    parity = None
    if not args.no_cpu:
        cpu, y_port, x_port = cpu_baseline(coord, xc, w, N, args.width, min(args.cpu_coils, C), C,
                                           runs=args.cpu_runs)
        log(f"cpu baseline: {cpu['value']:.4f} fwd+adj/s on {cpu['cores']} cores")
        parity = parity_vs_port(pf, plan, y_port, x_port, xc_dev, w_dev, dev)
        del y_port, x_port
        log(f"parity vs port: {parity}")

    copy_gbs = stream_copy(dev)
    log(f"device copy: {copy_gbs} GB/s")
    extras = None
    if not args.no_extras and not args.small:
        extras = {}
        for name, fn in EXTRAS:
            extras[name] = fn(pf, dev, max(3, args.steps // 2))
            log(f"{name}: {extras[name]}")

    out = {
        "metric": METRIC,
        "value": round(1000.0 / ms, 3),
        "unit": "fwd+adj/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic",
        "config": {
            "workload": f"3D radial NUFFT fwd+adj, {N}^3 x {C} coils batched, golden-means {NSP}x{NS}, "
                        f"oversamp 2.0, KB width {args.width}, |k|^2 weights",
            "image": N, "coils": C, "spokes": NSP, "samples": NS, "oversamp": 2.0, "width": args.width,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom + (" (" + " + ".join(STAGE_KERNELS[dom]) + ")" if dom in STAGE_KERNELS else ""),
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "copy_GBs": copy_gbs,
        },
        "stages": stage_info,
        "plan_s": round(plan_s, 3),
        "accuracy": acc,
        "parity_vs_port": parity,
        "compare_width": cmp,
        "vs_plain": extras,
        "cpu_baseline": cpu,
    }
    if cpu is not None:
        out["gpu_vs_cpu"] = round(out["value"] / cpu["value"], 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
