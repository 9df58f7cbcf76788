"""Pinning the oracle (oracle/oracle.c, the FP64 direct periodic sweep) -- CPU only.

Each check is chosen so that a plausible mistake in the oracle (a flipped tap, a wrong wrap, an
off-by-one shift, a dropped axis, the wrong convention) fails it:
  * against the reference's own CPU check code compiled from its sources (oracle/_ref):
    stencil1D (src/1D/check_correct.hpp:15-38), the 1D driver composition (src/1D/1d_main.cu:220-223)
    and the 2D driver composition (src/2D/2d_main.cu:190-192 over src/2D/check_correct.hpp:17-100),
    step by step over several steps;
  * against the committed golden vectors in tests/golden (made by tests/golden/make_golden.py from the
    reference check code and from exact rational arithmetic);
  * against an independent algorithm: numpy FFT circular convolution in 1D/2D/3D;
  * closed forms: Heat-1D T-step kernel = C(2T, j) / 4^T (exact in binary for T <= 20), delta kernel =
    pure roll, mass conservation sum(out) = sum(in) (sum w)^T, Fourier modes are eigenvectors with
    eigenvalue K(theta)^T.
"""
import os
from math import comb

import numpy as np
import pytest

from tests import _oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))


def fft_conv(x, w, radius, steps, convention=0):
    """Independent periodic convolution by FFT (numpy)."""
    d = x.ndim
    K = 2 * radius + 1
    w = np.asarray(w, dtype=np.float64).reshape((K,) * d)
    ker = np.zeros(x.shape)
    for idx in np.ndindex(*w.shape):
        off = tuple(((i if convention == 0 else radius - i) % n) for i, n in zip(idx, x.shape))
        ker[off] += w[idx]
    return np.real(np.fft.ifftn(np.fft.fftn(x) * np.fft.fftn(ker) ** steps))


@pytest.mark.parametrize("shape", [(7,), (64,), (33,), (5, 9), (16, 17), (3, 4, 5), (9, 8, 7)])
@pytest.mark.parametrize("radius", [0, 1, 2])
@pytest.mark.parametrize("convention", [0, 1])
def test_oracle_vs_fft(shape, radius, convention):
    rng = np.random.default_rng(hash((shape, radius, convention)) % 2**32)
    x = rng.random(shape)
    w = rng.random((2 * radius + 1) ** len(shape))
    w /= w.sum()
    for steps in (1, 3):
        got = O.sweep(x, radius, w, steps, convention=convention)
        ref = fft_conv(x, w, radius, steps, convention)
        assert np.max(np.abs(got - ref)) < 1e-13, (shape, radius, convention, steps)


def test_ref_lib_built_when_reference_present():
    if os.path.isdir("/root/reference"):
        assert O.ref_lib(1) is not None and O.ref_lib(2) is not None, "make -C oracle ref did not run"


@pytest.mark.parametrize("n,k", [(37, 3), (56, 29), (100, 5), (12, 7)])
def test_oracle_vs_reference_1d(n, k):
    L = O.ref_lib(1)
    if L is None:
        pytest.skip("reference check code not built (no reference checkout on this machine)")
    rng = np.random.default_rng(n * 31 + k)
    x = rng.random(n)
    w = rng.random(k)
    r = k // 2
    cur_ref, cur_c = x.copy(), x.copy()
    for step in range(4):
        nxt = np.empty(n)
        L.ref_step1d_driver(O._p(cur_ref), n, O._p(w), k, O._p(nxt))
        cur_ref = nxt
        nxt2 = np.empty(n)
        L.ref_stencil1d(O._p(cur_c), n, O._p(w), k, O._p(nxt2))
        cur_c = nxt2
        assert O.errors(O.sweep(x, r, w, step + 1, convention=0), cur_ref)[1] < 1e-14
        assert O.errors(O.sweep(x, r, w, step + 1, convention=1), cur_c)[1] < 1e-14


@pytest.mark.parametrize("width,k", [(19, 3), (24, 3), (31, 5), (8, 3)])
def test_oracle_vs_reference_2d(width, k):
    L = O.ref_lib(2)
    if L is None:
        pytest.skip("reference check code not built (no reference checkout on this machine)")
    rng = np.random.default_rng(width * 7 + k)
    x = rng.random((width, width))
    w = rng.random(k * k)
    cur = x.copy()
    for step in range(3):
        nxt = np.empty_like(cur)
        L.ref_step2d_driver(O._p(np.ascontiguousarray(cur)), width, O._p(w), k, O._p(nxt))
        cur = nxt
        assert O.errors(O.sweep(x, k // 2, w, step + 1, convention=0), cur)[1] < 1e-14


def test_oracle_golden_vectors():
    path = os.path.join(HERE, "golden", "golden.npz")
    g = np.load(path, allow_pickle=False)
    names = sorted({k.split("__")[0] for k in g.files})
    assert len(names) >= 6
    for name in names:
        x = g[name + "__in"]
        w = g[name + "__w"]
        meta = g[name + "__meta"]  # radius, steps, convention
        radius, steps, conv = (int(v) for v in meta)
        want = g[name + "__out"]
        got = O.sweep(x, radius, w, steps, convention=conv)
        assert np.max(np.abs(got - want)) <= 1e-13 * max(1.0, np.max(np.abs(want))), name


@pytest.mark.parametrize("steps", [1, 2, 7, 14, 20])
def test_heat1d_binomial_closed_form(steps):
    n = 2 * steps + 9
    x = np.zeros(n)
    x[0] = 1.0
    got = O.sweep(x, 1, [0.25, 0.5, 0.25], steps, convention=0)
    want = np.zeros(n)
    for j in range(2 * steps + 1):
        want[j] = comb(2 * steps, j) / 4.0 ** steps
    assert np.array_equal(got, want)  # all values are dyadic rationals: exact in FP64


@pytest.mark.parametrize("dims", [1, 2, 3])
def test_delta_kernel_is_roll(dims):
    rng = np.random.default_rng(dims)
    shape = (11, 6, 5)[:dims]
    x = rng.random(shape)
    r = 1
    w = np.zeros((3,) * dims)
    tap = (2, 0, 1)[:dims]
    w[tap] = 1.0
    got = O.sweep(x, r, w, 3, convention=0)
    assert np.array_equal(got, np.roll(x, shift=tuple(3 * t for t in tap), axis=tuple(range(dims))))
    got_c = O.sweep(x, r, w, 3, convention=1)  # centred: out[p] = x[p + (tap - r)]
    assert np.array_equal(got_c, np.roll(x, shift=tuple(-3 * (t - r) for t in tap), axis=tuple(range(dims))))


def test_mass_conservation_and_constant_field():
    rng = np.random.default_rng(4)
    x = rng.random((23, 19))
    w = rng.random(25)
    got = O.sweep(x, 2, w, 5)
    assert abs(got.sum() - x.sum() * w.sum() ** 5) < 1e-10 * abs(x.sum() * w.sum() ** 5)
    c = np.full((10, 12), 0.75)
    wn = w / w.sum()
    assert np.max(np.abs(O.sweep(c, 2, wn, 9) - 0.75)) < 1e-14


def test_fourier_mode_eigenvalue():
    n0, n1 = 24, 40
    k0, k1 = 3, 5
    i0, i1 = np.meshgrid(np.arange(n0), np.arange(n1), indexing="ij")
    th0, th1 = 2 * np.pi * k0 / n0, 2 * np.pi * k1 / n1
    x = np.cos(th0 * i0 + th1 * i1)
    rng = np.random.default_rng(12)
    w = rng.random(9)
    w /= w.sum()
    W = w.reshape(3, 3)
    Kh = sum(W[a, b] * np.exp(-1j * (th0 * a + th1 * b)) for a in range(3) for b in range(3))
    T = 6
    want = np.real(Kh ** T * np.exp(1j * (th0 * i0 + th1 * i1)))
    got = O.sweep(x, 1, w, T)
    assert np.max(np.abs(got - want)) < 1e-13


def test_splitmix_uniform_known_values():
    # SplitMix64 (Steele, Lea, Flood 2014) from state 0: first outputs are published test values
    u = O.uniform(0, (3,))
    first = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    assert list(u) == [(v >> 11) * 2.0 ** -53 for v in first]


@pytest.mark.parametrize("dims,shape,radius", [(1, (301,), 3), (2, (67, 53), 2), (3, (17, 19, 13), 1)])
def test_t_fold_operator_matches_sweep(oracle, dims, shape, radius):
    # the full-size GPU tests evaluate T steps at sample points as one convolution with the T-fold kernel
    # (the oracle's response to a delta); on small grids that must equal the oracle's own sweep there,
    # including the wrap-around rows and columns
    from tests.test_gpu_fullsize import apply_at, sample_points, t_fold_kernel
    rng = np.random.default_rng(dims)
    w = rng.random((2 * radius + 1) ** dims)
    w /= w.sum()
    x = oracle.uniform(5 + dims, shape)
    for steps in (1, 3, 5):
        pts = sample_points(shape, 300, steps)
        got = apply_at(x, t_fold_kernel(oracle, dims, radius, w, steps), pts)
        ref = oracle.sweep(x, radius, w, steps)[tuple(pts)]
        assert np.abs(got - ref).max() <= 1e-14 * np.abs(ref).max()


@pytest.mark.parametrize("dims,shapes", [
    (1, [(1,), (3,), (7,), (31,), (33,), (4099,), (70001,)]),
    (2, [(1, 1), (2, 5), (5, 3), (17, 40), (64, 97), (33, 1100)]),
    (3, [(1, 2, 3), (4, 5, 6), (9, 8, 37), (6, 7, 70)]),
])
def test_fast_sweep_matches_strict_sweep(dims, shapes):
    """oracle/sweep_fast.c (the timed CPU baseline: register blocks, overlapping tail block, cheap wraps,
    column strips) against the strict sweep, over rows shorter and longer than its 32-point blocks and
    512-point strips, grids smaller than the stencil, radii 1..3, zero taps and both conventions."""
    rng = np.random.default_rng(dims)
    for shape in shapes:
        for radius in (1, 2, 3):
            for conv in (0, 1):
                x = rng.random(shape) - 0.25
                w = rng.random((2 * radius + 1) ** dims) - 0.3
                w[rng.random(w.size) < 0.2] = 0.0
                a = O.sweep(x, radius, w, 6, conv)
                b = O.sweep_fast(x, radius, w, 6, conv)
                err = np.abs(a - b).max() / np.abs(a).max()
                assert err <= 1e-13, (shape, radius, conv, err)
    # zero steps copy the grid; a repeated call gives the same result
    x = rng.random(shapes[-1])
    w = rng.random(3 ** dims)
    ref = O.sweep(x, 1, w, 3)
    assert np.abs(O.sweep_fast(x, 1, w, 0) - x).max() == 0.0
    assert np.abs(O.sweep_fast(x, 1, w, 3) - ref).max() <= 1e-13 * np.abs(ref).max()


def test_fast_sweep_wide_stencil():
    """A radius much larger than the grid (every read wraps, (2r+1)^2 source rows per output row)."""
    rng = np.random.default_rng(7)
    x = rng.random((23, 41))
    w = rng.random(121 * 121) - 0.5
    a = O.sweep(x, 60, w, 2)
    b = O.sweep_fast(x, 60, w, 2)
    assert np.abs(a - b).max() <= 1e-13 * np.abs(a).max()
