"""The drop-in 1d/2d/3d programs: argv handling and exit codes on the CPU, the CSV line on the GPU.

Reference behaviour kept (src/2D/2d_main.cu:43-47, src/1D/1d_main.cu:60-64): wrong argument count prints
the help text and exits 1; an unknown shape prints a message and exits 1.  Implemented here where the
reference only documents it (README.md:56-58): --help (exit 0) and --custom.
"""
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "flashfftstencil_amd", "bin")


def run(prog, *args, stdin=None, timeout=120):
    exe = os.path.join(BIN, prog)
    if not os.path.exists(exe):
        pytest.fail(f"{exe} not built")
    return subprocess.run([exe, *args], capture_output=True, text=True, input=stdin, timeout=timeout)


@pytest.mark.parametrize("prog", ["1d", "2d", "3d"])
def test_usage_errors(prog):
    r = run(prog)
    assert r.returncode == 1 and "Usage" in r.stdout
    r = run(prog, "Heat-1D", "10")
    assert r.returncode == 1 and "Usage" in r.stdout
    r = run(prog, "--help")
    assert r.returncode == 0 and "Stencil-shape" in r.stdout and "--custom" in r.stdout
    r = run(prog, "--bogus-flag", "a", "b", "c")
    assert r.returncode == 1


@pytest.mark.parametrize("flag,value", [("--algorithm", "fourier"), ("--precision", "fp8"), ("--convention", "left"),
                                        ("--mapping", "spiral")])
def test_bad_option_values_fail_before_the_gpu(flag, value):
    # rejected by ffs_plan_create's validation, which runs before any HIP call
    r = run("2d", "Box2D9P", "64", "4", flag, value)
    assert r.returncode == 1 and "error" in r.stderr


def test_help_lists_algorithms_and_forms():
    r = run("2d", "--help")
    assert "--algorithm" in r.stdout and "direct" in r.stdout and "f16" in r.stdout


@pytest.mark.parametrize("prog,shape", [("1d", "Box2D9P"), ("2d", "Heat-1D"), ("3d", "NoSuchShape"), ("2d", "Box3D27P")])
def test_wrong_shape_for_program(prog, shape):
    r = run(prog, shape, "64", "4")
    assert r.returncode == 1
    assert "shape" in r.stdout


def test_custom_weight_count_checked(tmp_path):
    f = tmp_path / "w.txt"
    f.write_text("0.1 0.2 0.3 0.4\n")
    r = run("2d", "Box2D9P", "64", "4", "--custom", str(f))
    assert r.returncode == 1 and "needs 9 weights" in r.stderr
    r = run("1d", "Heat-1D", "64", "4", "--custom", "-", stdin="1 2\n")
    assert r.returncode == 1 and "needs 3 weights" in r.stderr
    r = run("2d", "Box2D9P", "64", "4", "--custom", str(tmp_path / "missing.txt"))
    assert r.returncode == 1


def test_bad_numbers():
    assert run("2d", "Box2D9P", "abc", "4").returncode == 1
    assert run("2d", "Box2D9P", "64", "4x").returncode == 1


CSV = re.compile(r"^FlashFFTStencil, (\S+), (\d+), (\d+), (\d+), (\d+), ([0-9.e+-]+), ([0-9.e+-]+)$")


@pytest.mark.gpu
@pytest.mark.parametrize("args,tag,dims", [
    (("1d", "Heat-1D", "100000", "64"), "1d1r", (1, 1, 100000)),
    (("2d", "Box2D9P", "1000", "16"), "box_2d1r", (1, 1000, 1000)),
    (("2d", "Box2D25P", "512", "300", "32"), "box_2d2r", (1, 512, 300)),
    (("2d", "Star2D9P", "257", "8", "--precision", "bf16"), "star_2d2r", (1, 257, 257)),
    (("2d", "Box2D25P", "700", "24", "--precision", "f16"), "box_2d2r", (1, 700, 700)),
    (("2d", "Box2D25P", "700", "24", "--algorithm", "direct"), "box_2d2r", (1, 700, 700)),
    (("3d", "Heat-3D", "40", "6", "--algorithm", "auto"), "star_3d1r", (40, 40, 40)),
])
def test_csv_line_on_gpu(args, tag, dims):
    r = run(*args)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().splitlines()
    m = CSV.match(lines[-1])
    assert m, lines
    assert m.group(1) == tag
    assert tuple(int(m.group(i)) for i in (2, 3, 4)) == dims
    T = int(next(a for a in reversed(args[:5]) if a.isdigit()))
    ms, gst = float(m.group(6)), float(m.group(7))
    assert ms > 0
    pts = dims[0] * dims[1] * dims[2]
    assert abs(gst - pts * T / ms / 1e6) <= 1e-3 * gst


@pytest.mark.gpu
def test_custom_weights_on_gpu(tmp_path):
    # w[1][1] = 1 in the ref convention (out[p] = sum_a w[a] in[p - a]) moves the grid by (1, 1) per step:
    # after 7 steps the output is the initial grid rolled by (7, 7)
    f = tmp_path / "w.txt"
    f.write_text("0 0 0\n0 1 0\n0 0 0\n")
    out = tmp_path / "out.bin"
    r = run("2d", "Box2D9P", "96", "7", "--custom", str(f), "--output", str(out))
    assert r.returncode == 0, r.stderr
    assert r.stdout.startswith("FlashFFTStencil, box_2d1r, 1, 96, 96, 7, ")
    from tests import _oracle
    x = _oracle.uniform(0x5EED0000 + 1, (96, 96))
    got = np.fromfile(out, dtype=np.float64).reshape(96, 96)
    assert np.abs(got - np.roll(x, (7, 7), axis=(0, 1))).max() < 1e-13


@pytest.mark.gpu
@pytest.mark.parametrize("dims,shape,size,steps,extra", [
    ("2d", "Box2D25P", (70, 50), 9, ()), ("2d", "Box2D25P", (70, 50), 9, ("--precision", "f16")),
    ("3d", "Box3D27P", (20, 18, 22), 4, ()), ("1d", "1D7P", (3000,), 12, ("--convention", "centred")),
])
def test_custom_weights_match_oracle(tmp_path, dims, shape, size, steps, extra):
    # random (unnormalised, signed) --custom weights: the driver's grid equals the oracle's on the
    # driver's own seeded input
    from tests import _oracle
    d = len(size)
    radius = {"Box2D25P": 2, "Box3D27P": 1, "1D7P": 3}[shape]
    w = np.random.default_rng(len(size) * 7 + steps).standard_normal((2 * radius + 1) ** d) * 0.2
    f = tmp_path / "w.txt"
    f.write_text(" ".join(repr(float(v)) for v in w))
    out = tmp_path / "out.bin"
    r = run(dims, shape, *map(str, size), str(steps), "--custom", str(f), "--output", str(out), *extra)
    assert r.returncode == 0, r.stderr
    x = _oracle.uniform(0x5EED0000 + 1, size)
    conv = 1 if "centred" in extra else 0
    ref = _oracle.sweep(x, radius, w, steps, convention=conv)
    got = np.fromfile(out, dtype=np.float64).reshape(size)
    tol = 1e-6 if "f16" in extra else 1e-12
    assert _oracle.errors(got, ref)[1] <= tol
