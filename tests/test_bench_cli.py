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
