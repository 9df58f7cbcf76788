"""bench.py's single-GPU contract on the CPU: N > 1 exits with a message, before touching a GPU."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_rejects_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_batch96_specs_are_seeded_and_in_range():
    sys.path.insert(0, ROOT)
    import bench
    import flashfftstencil_amd as ffs

    a = bench.batch96_specs(ffs, 0x5EED0004)
    b = bench.batch96_specs(ffs, 0x5EED0004)
    assert len(a) == 96
    for s, t in zip(a, b):
        assert s["sizes"] == t["sizes"] and s["steps"] == t["steps"] and s["shape"] == t["shape"]
        assert all(512 <= n <= 4096 and n % 128 == 0 for n in s["sizes"])
        assert 16 <= s["steps"] <= 256 and s["steps"] % 16 == 0
    assert len({s["shape"] for s in a}) == 4


def test_conv_step_matches_oracle():
    # the library-convolution baseline computes the same periodic step as the oracle (ref convention)
    sys.path.insert(0, ROOT)
    import numpy as np
    import torch

    import bench
    from tests import _oracle

    rng = np.random.default_rng(5)
    for shape, r in (((37,), 1), ((19, 23), 2), ((7, 9, 11), 1)):
        w = rng.random((2 * r + 1) ** len(shape))
        x = rng.random(shape)
        step = bench.conv_step(torch, torch.from_numpy(x), w, r)
        got = step(step(torch.from_numpy(x))).numpy()
        ref = _oracle.sweep(x, r, w, 2)
        assert _oracle.errors(got, ref)[1] < 1e-13, shape


def test_roofline_follows_algorithmic_counts():
    # Box2D25P 8192^2, 32-point tiles at t = 4 (V = 16): 131072 tile pairs of 11 output blocks x 4 real
    # GEMMs x 3 FP16 products (2,162,688 executed MFMA flops per pair); each real GEMM counted once that is
    # 1/3 of it.  A 0.3657 ms block is 0.103 of the FP16 matrix peak (0.150 by the dense 8 d P^(d+1) count)
    # and 0.367 of 8 TB/s, so HBM bounds it.
    sys.path.insert(0, ROOT)
    import bench

    info = dict(dims=2, sizes=[8192, 8192, 1], tile=[32, 32, 1], tiles=512 * 512, fuse=4,
                flops_per_block=2162688.0 * 131072)
    rl = bench.roofline(info, 0.3657e-3, "f16", 2, 25)
    assert rl["dense_flops_per_launch"] == 8 * 2 * 32 ** 3 * 512 * 512
    assert rl["flops_per_launch"] == 2162688.0 * 131072 / 3
    assert rl["bound"] == "hbm" and abs(rl["frac"] - 0.367) < 0.001 and rl["unit"] == "GB/s"
    assert abs(rl["matrix_frac"] - 0.1033) < 0.001 and abs(rl["matrix_frac_dense"] - 0.1503) < 0.001
    # FP64 form (Gauss 3 GEMMs, 11 output blocks: 540,672 flops per pair) at 1.074 ms: 0.84 of the FP64 peak
    info64 = dict(info, flops_per_block=540672.0 * 131072)
    rl64 = bench.roofline(info64, 1.074e-3, "fp64", 2, 25)
    assert rl64["bound"] == "mfma" and abs(rl64["frac"] - 0.8395) < 0.001 and rl64["matrix_frac_dense"] > 1.5
    # 1D prime-factor window (dense count) and the direct kernel (2 |S| t flops per point)
    info1 = dict(dims=1, sizes=[1 << 20, 1, 1], tile=[240, 1, 1], tiles=9363, fuse=64, flops_per_block=1.0)
    assert bench.roofline(info1, 20e-6, "fp64", 1, 3)["dense_flops_per_launch"] == 8 * 240 * 31 * 9363
    rld = bench.roofline(info, 0.36e-3, "direct", 2, 25)
    assert rld["flops_per_launch"] == 2 * 25 * 8192 * 8192 * 4 and rld["executed_mfma_tflops"] is None
