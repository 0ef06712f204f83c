"""bench.py's contract on the host: one GPU only, and the byte model its
roofline figure is priced on."""

import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""


def test_headline_algorithmic_bytes():
    # 3-D 128^3 x 32 coils, sigma 2 (256^3 grid), 8192 x 256 samples: grid
    # + samples + 4 (d + 1) bytes per sample for the gather and the spread
    C, N, G, M, d = 32, 128, 256, 8192 * 256, 3
    grid, samples, coords = C * G**3 * 8, C * M * 8, M * 4 * (d + 1)
    assert bench.stage_bytes("interp", C, N**3, G**3, M, d) == grid + samples + coords == 4_865_392_640
    assert bench.stage_bytes("spread", C, N**3, G**3, M, d) == 4_865_392_640
    assert bench.stage_bytes("unsort", C, N**3, G**3, M, d) == 2 * samples + M * 8


def test_headline_traffic_found_for_the_dominant_kernels():
    # the committed rocprof counter summary names the kernels bench.py reports
    import argparse

    args = argparse.Namespace(width=6, small=False)
    for stage in ("interp", "spread"):
        t = bench.measured_traffic(stage, args)
        assert t is not None, stage
        # at least the compulsory bytes (grid once, samples once) moved
        assert t >= 32 * 256**3 * 8 + 32 * 8192 * 256 * 8


def test_fp64_gram_rows_equal_full_normal_operator():
    """bench.py's row-wise FP64 A^H W A (the Toeplitz workload's reference)
    equals the full exact normal operator on the same rows."""
    import numpy as np

    from oracle import nudft as U
    from pyfourier_amd import demo

    rng = np.random.default_rng(3)
    N = 16
    c = demo.variable_density_spiral_2d(3, 64, N).reshape(-1, 2).astype(np.float64) * N / (2 * np.pi)
    w = rng.uniform(0.5, 1.5, c.shape[0])
    x = demo.complex_gaussian(rng, (N, N))
    full = U.nudft_adj((w * U.nudft(x[None], c, [N, N])[0])[None], c, [N, N])[0]
    rows = [1, 6, 11]
    got = bench._gram_exact_rows(x, c, w, N, rows)
    assert np.linalg.norm(got - full[rows]) / np.linalg.norm(full[rows]) < 1e-12
