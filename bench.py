#!/usr/bin/env python
"""Headline benchmark: 3D multicoil NUFFT forward+adjoint on one MI355X.

Workload: 128^3 complex64 image, 32 coils batched in one call (smooth
Gaussian sensitivity maps applied before the timed region), golden-means 3D
radial trajectory 8192 spokes x 256 samples, oversampling 2.0 (256^3 grid),
Kaiser-Bessel width 4, ramp density weights (|k|/k_max)^2.  One step = one
nufft + one nufft_adj on a precomputed plan (plan creation excluded), inputs
resident in HBM.

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
    ap.add_argument("--width", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the NUDFT error check")
    ap.add_argument("--cpu-coils", type=int, default=8, help="coils in the CPU baseline sample")
    ap.add_argument("--small", action="store_true", help="64^3 x 8 coils quick run")
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
        "interp": C * G3 * 8 + C * M * 8 + M * 4 * d,
        "unsort": 2 * C * M * 8 + M * 8,
        "sort": 2 * C * M * 8 + M * 8,
        "spread": C * M * 8 + M * 4 * d + C * G3 * 8,
        "fft_adj": 2 * C * G3 * 8,
        "crop_deapod": 2 * C * N3 * 8,
    }
    return b[name]


def cpu_baseline(coord, xc, w, N, width, ncoils_sample, C_full):
    """Time the reference algorithm on the host: C/OpenMP restatement of the
    Numba loops + torch CPU apod/pad/FFT, one fwd+adj on a coil sample."""
    import torch

    from oracle import port as Pt
    from oracle import restate as R

    ncores = len(os.sched_getaffinity(0))
    Pt.set_threads(ncores)
    P = R.Plan(coord, N, width=width, oversamp=2.0)
    pp = Pt.PortPlan(P)
    wf = w.reshape(1, 1, -1)
    # warm-up on one coil (thread pools, page-in)
    y = Pt.nufft(xc[:1, None], pp) * wf
    Pt.nufft_adj(y * wf, pp)
    xs = xc[:ncoils_sample, None]
    t0 = time.perf_counter()
    y = Pt.nufft(xs, pp) * wf
    Pt.nufft_adj(y * wf, pp)
    dt = time.perf_counter() - t0
    per_full = dt * C_full / ncoils_sample
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {
        "value": 1.0 / per_full,
        "unit": "fwd+adj/s",
        "cores": ncores,
        "kind": "port",
        "sample": f"{ncoils_sample} of {C_full} coils, full trajectory, one fwd+adj after a 1-coil warm-up "
                  f"({dt:.1f} s), scaled x{C_full / ncoils_sample:g}; C/OpenMP restatement of the Numba "
                  f"loops + torch CPU FFT, no gc.collect overheads; CPU: {model}",
        "torch_threads": torch.get_num_threads(),
    }


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


def main():
    args = parse()
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
    N, C, NSP, NS = (64, 8, 2048, 128) if args.small else (128, 32, 8192, 256)
    x, xc, coord, w = make_inputs(N, C, NSP, NS)
    log(f"inputs ready: {N}^3 x {C} coils, {NSP}x{NS} samples")
    t0 = time.perf_counter()
    plan = pf.plan_nufft(torch.from_numpy(coord), N, width=args.width, oversamp=2.0)
    plan.to(dev)
    torch.cuda.synchronize()
    plan_s = time.perf_counter() - t0
    xc_dev = torch.from_numpy(xc).to(dev)
    w_dev = torch.from_numpy(w).to(dev)

    def step():
        y = pf.nufft(xc_dev, nufft_plan=plan, weight=w_dev, device=dev)
        return y, pf.nufft_adj(y, nufft_plan=plan, weight=w_dev, device=dev)

    log(f"plan built in {plan_s:.2f} s")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    log("warm-up done")
    with stage_timer() as timer:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            y, xa = step()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
    ms = (t1 - t0) * 1000.0 / args.steps
    log(f"timed {args.steps} steps: {ms:.3f} ms/step")
    stages = {k: float(np.mean(v)) for k, v in timer.summary().items()}

    G3 = int(np.prod(plan.os_shape))
    N3 = N ** 3
    M = NSP * NS
    stage_info = {}
    for k, v in stages.items():
        by = stage_bytes(k, C, N3, G3, M, 3)
        stage_info[k] = {"ms": round(v, 4), "GB_s": round(by / (v * 1e-3) / 1e9, 1),
                         "frac_hbm": round(by / (v * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    own = {k: v for k, v in stages.items() if not k.startswith("fft")}
    dom = max(own, key=own.get)
    dom_bytes = stage_bytes(dom, C, N3, G3, M, 3)
    achieved = dom_bytes / (stages[dom] * 1e-3) / 1e9

    acc = None
    if not args.no_accuracy:
        coord_grid = plan.interpolator.coord.reshape(-1, 3).numpy()
        y_nw = pf.nufft(xc_dev[:1], nufft_plan=plan, device=dev)
        acc = {"fwd_rel_err_vs_nudft": accuracy(pf, plan, x, xc_dev, y_nw, coord_grid, N, dev),
               "width": args.width, "samples_checked": 1024}

    if acc is not None:
        log(f"accuracy vs NUDFT: {acc}")
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(coord, xc, w, N, args.width, min(args.cpu_coils, C), C)
        log(f"cpu baseline: {cpu['value']:.4f} fwd+adj/s on {cpu['cores']} cores")

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
            "kernel": dom,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
        },
        "stages": stage_info,
        "plan_s": round(plan_s, 3),
        "accuracy": acc,
        "cpu_baseline": cpu,
    }
    if cpu is not None:
        out["gpu_vs_cpu"] = round(out["value"] / cpu["value"], 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
