"""Unit checks of the plan tables the GPU generates (read back with ffs_plan_read_table), against closed
forms, so that a wrong DFT entry, sign, index map or normalisation is found in the table itself and not
only through an end-to-end result:
  * DFT operands: F[j][k] = exp(-2 pi i j k / L) and F F^H = L I, in every family's per-lane layout
    (reference: src/2D/create_fft_pfa_plan.cu:25-39, src/1D/create_fft_pfa_plan.cu:43-76);
  * split operands of the mixed forms: the terms sum back to F within the split's precision, and the
    FP32 operands with their low-order terms give F to ~2^-48;
  * spectra: H = K^t / prod(P) against a numpy FFT of the zero-padded kernel, K(0)^t = (sum w)^t, and
    the Heat-1D closed form K^t = ((1 + cos th)/2)^t e^{-i th t} on the 240-point prime-factor map
    (reference: src/2D/create_fft_pfa_plan.cu:76-107, src/1D/create_fft_pfa_plan.cu:143-178).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def dft(L):
    """F[j][k] = exp(-2 pi i j k / L) with the angle reduced exactly, (j k mod L) / L, before the
    exponential (numpy's own rounding of 2 pi j k / L for large j k would exceed the tables' error)."""
    j = np.arange(L)
    return np.exp(-2j * np.pi * (np.outer(j, j) % L) / L)


def table(plan, which, dtype):
    return np.frombuffer(plan.read_table(which), dtype=dtype)


def check_dft(F, L):
    assert np.abs(F - dft(L)).max() < 4e-16 * L
    assert np.abs(F @ F.conj().T - L * np.eye(L)).max() < 1e-13


def ops16(plan, axis):
    """16-point operands of one axis: [arr: Fr, -(Fr+Fi), Fi-Fr][chunk c][lane], entry (4c + lane/16, lane%16)."""
    t = table(plan, 0, np.float64)[axis * 768:(axis + 1) * 768].reshape(3, 4, 64)
    R, S, D = (np.zeros((16, 16)) for _ in range(3))
    for c in range(4):
        for lane in range(64):
            row, col = 4 * c + lane // 16, lane % 16
            R[row, col], S[row, col], D[row, col] = t[0, c, lane], t[1, c, lane], t[2, c, lane]
    return R, S, D


def gauss_arrays(arrays, F):
    """The three arrays of the Gauss form, each generated in FP64 from the exact entry: R = Re F,
    S = -(Re F + Im F), D = Im F - Re F, each checked against the closed form F on its own (recovering
    Im F from D + R rounds, so the arrays are not compared with each other bitwise)."""
    R, S, D = arrays
    tol = 4e-16 * F.shape[0]
    assert np.abs(R - F.real).max() < tol
    assert np.abs(S + F.real + F.imag).max() < tol
    assert np.abs(D - (F.imag - F.real)).max() < tol
    return R + 1j * (D + R)


def test_dft_operands_16(ffs, torch_cuda):
    _, r, w, _ = ffs.named_stencil("Box2D9P", seed=1)
    plan = ffs.Plan(2, (64, 64), w, r, 3, precision="fp64", tile=16)
    for axis in (0, 1):
        check_dft(gauss_arrays(ops16(plan, axis), dft(16)), 16)
    plan.close()


def test_dft_operands_1d_prime_factor(ffs, torch_cuda):
    # the 240-point window runs as a 15 x 16 Good-Thomas matrix: axis 0 holds F_15 zero-padded to 16
    plan = ffs.Plan(1, (5000,), np.array([0.25, 0.5, 0.25]), 1, 8, precision="fp64")
    want = np.zeros((16, 16), complex)
    want[:15, :15] = dft(15)
    F15 = gauss_arrays(ops16(plan, 0), want)
    check_dft(F15[:15, :15], 15)
    assert np.all(F15[15, :] == 0) and np.all(F15[:, 15] == 0)
    check_dft(gauss_arrays(ops16(plan, 1), dft(16)), 16)
    plan.close()


def test_dft_operands_32_fp64(ffs, torch_cuda):
    # [R][chunk v][column block Kb][arr][lane], lane (n + 16k) = entry (16 R + 4 v + k, 16 Kb + n)
    _, r, w, _ = ffs.named_stencil("Box2D25P", seed=2)
    plan = ffs.Plan(2, (256, 256), w, r, 8, precision="fp64", tile=32)
    assert plan.info["tile"][0] == 32
    t = table(plan, 0, np.float64)[:3072].reshape(2, 4, 2, 3, 64)
    R, S, D = (np.zeros((32, 32)) for _ in range(3))
    for Rb in range(2):
        for v in range(4):
            for Kb in range(2):
                for lane in range(64):
                    row, col = 16 * Rb + 4 * v + lane // 16, 16 * Kb + lane % 16
                    R[row, col], S[row, col], D[row, col] = t[Rb, v, Kb, :, lane]
    check_dft(gauss_arrays((R, S, D), dft(32)), 32)
    plan.close()


def _split_terms(words, f16):
    """[arr][Kb][term][word][lane] words of two 16-bit values -> per-term 32 x 32 matrices of arr."""
    w = words.reshape(3, 2, 3, 4, 64)
    half = np.float16 if f16 else None
    out = np.zeros((3, 3, 32, 32))
    for arr in range(3):
        for Kb in range(2):
            for term in range(3):
                for wd in range(4):
                    for lane in range(64):
                        word = int(w[arr, Kb, term, wd, lane])
                        for h in range(2):
                            bits = (word >> (16 * h)) & 0xFFFF
                            if f16:
                                val = float(np.array([bits], dtype=np.uint16).view(half)[0])
                            else:
                                val = float(np.array([bits << 16], dtype=np.uint32).view(np.float32)[0])
                            e = 2 * wd + h
                            row, col = 16 * (e >> 2) + 4 * (lane // 16) + (e & 3), 16 * Kb + lane % 16
                            out[arr, term, row, col] = val
    return out


@pytest.mark.parametrize("form,bound", [("bf16", 2.0 ** -24), ("f16", 2.0 ** -22)])
def test_split_operands_32(ffs, torch_cuda, form, bound):
    # BF16: F = h + m + l (3 terms, ~24 bits); FP16: F = h + l (2 terms, ~22 bits); arrays Fr, Fi, -Fi
    _, r, w, _ = ffs.named_stencil("Box2D25P", seed=3)
    plan = ffs.Plan(2, (256, 256), w, r, 8, precision=form)
    terms = _split_terms(table(plan, 1, np.uint32)[:4608], form == "f16")
    F = dft(32)
    for arr, want in enumerate((F.real, F.imag, -F.imag)):
        got = terms[arr].sum(axis=0)
        assert np.abs(got - want).max() <= bound, (form, arr)
        if form == "f16":
            assert np.all(terms[arr, 2] == 0)  # no third FP16 term
    plan.close()


def test_split_operands_16(ffs, torch_cuda):
    """16-point mixed operands (a 3D plan: every axis 16 points).  BF16 words [axis][arr][pairing][word]
    [lane], element e of lane (j + 16q) = row 4q + (e & 3), column j; pairings 0, 1, 2 carry the terms
    h, m, l in their first four elements, which sum back to F within the 3-term split's precision.  FP32
    chunks [axis][arr][c][lane] = F[4q + c][j] rounded to FP32, then their low-order terms, with which
    they give F to ~2^-48 (the 3D FP32 form's compensated matrix)."""
    with ffs.Plan(3, (32, 32, 32), np.full(27, 1 / 27), 1, 2, precision="f32") as plan:
        words = table(plan, 1, np.uint32)
    F = dft(16)
    wants = (F.real, F.imag, -F.imag)
    bf = words[:3 * 3072].reshape(3, 3, 4, 4, 64)
    f32 = words[3 * 3072:3 * 3072 + 2 * 3 * 768].view(np.float32).astype(np.float64)
    hi, lo = f32[:2304].reshape(3, 3, 4, 64), f32[2304:].reshape(3, 3, 4, 64)
    lanes = np.arange(64)
    rows_c = 4 * (lanes // 16)
    cols = lanes % 16
    for axis in range(3):
        for arr, want in enumerate(wants):
            split = np.zeros((16, 16))
            for g in range(3):
                for wd in range(2):  # elements 0..3: words 0 and 1
                    for h in range(2):
                        bits = ((bf[axis, arr, g, wd] >> (16 * h)) & 0xFFFF).astype(np.uint32) << 16
                        split[rows_c + 2 * wd + h, cols] += bits.view(np.float32).astype(np.float64)
            assert np.abs(split - want).max() <= 2.0 ** -24, (axis, arr)
            H, L = np.zeros((16, 16)), np.zeros((16, 16))
            for c in range(4):
                H[rows_c + c, cols] = hi[axis, arr, c]
                L[rows_c + c, cols] = lo[axis, arr, c]
            assert np.abs(H - want).max() <= 2.0 ** -25, (axis, arr)  # FP32 rounding of entries <= 1
            assert np.abs(H + L - want).max() <= 2.0 ** -46, (axis, arr)
            assert np.abs(H - want).max() > 2.0 ** -30  # the low-order terms carry something


def spectrum(plan, length):
    t = table(plan, 2, np.float64)
    main = t[:length] + 1j * t[length:2 * length]
    tail = t[2 * length:3 * length] + 1j * t[3 * length:4 * length]
    return main, tail


def kernel_dft(w, radius, dims, P, convention):
    K = 2 * radius + 1
    ker = np.zeros((P,) * dims)
    ker[(slice(0, K),) * dims] = np.asarray(w).reshape((K,) * dims)
    if convention == "centred":  # out[p] = sum_o w[o + r] in[p + o]: taps at offsets r - a
        ker = np.zeros((P,) * dims)
        for a in np.ndindex(*(K,) * dims):
            ker[tuple((radius - ai) % P for ai in a)] += np.asarray(w).reshape((K,) * dims)[a]
    return np.fft.fftn(ker)


@pytest.mark.parametrize("convention", ["ref", "centred"])
@pytest.mark.parametrize("form", ["fp64", "f16"])
def test_spectrum_2d32(ffs, torch_cuda, convention, form):
    # entry e = (kb * 4 + v) * 64 + lane of block kb = 2 Rb + C: row 16 Rb + (q + 4 v | 4 q + v), column
    # 16 C + j (FP64 | FP32 accumulator order); unnormalised weights (sum 1.3), main depth 4 and tail 1
    rng = np.random.default_rng(5)
    w = rng.random(25)
    w *= 1.3 / w.sum()
    plan = ffs.Plan(2, (200, 200), w, 2, 9, precision=form, convention=convention, fuse=4, tile=32)
    info = plan.info
    assert (info["fuse"], info["fuse_tail"]) == (4, 1)
    main, tail = spectrum(plan, 1024)
    Kf = kernel_dft(w, 2, 2, 32, convention)
    for got, t in ((main, 4), (tail, 1)):
        want = np.zeros(1024, complex)
        for e in range(1024):
            kb, v, lane = e // 256, (e // 64) % 4, e % 64
            q, j = lane // 16, lane % 16
            row = 16 * (kb >> 1) + (q + 4 * v if form == "fp64" else 4 * q + v)
            want[e] = Kf[row, 16 * (kb & 1) + j] ** t / 1024
        assert np.abs(got - want).max() < 1e-15 * np.abs(want).max() * 8, t
        assert abs(got[0] - 1.3 ** t / 1024) < 1e-16  # K(0)^t = (sum w)^t
    plan.close()


def test_spectrum_heat1d_closed_form(ffs, torch_cuda):
    # 240-point prime-factor map: entry (v, lane) -> row frequency k0 = q + 4 v (< 15), column frequency
    # k1 = lane % 16, spectrum index k = (15 k1 + 16 k0) mod 240; Heat-1D: K^t = ((1 + cos th)/2)^t e^{-i th t}
    T, t = 40, 32
    plan = ffs.Plan(1, (5000,), np.array([0.25, 0.5, 0.25]), 1, T, precision="fp64", fuse=t)
    info = plan.info
    assert (info["fuse"], info["fuse_tail"]) == (t, T - t)
    main, tail = spectrum(plan, 256)
    for got, tt in ((main, t), (tail, T - t)):
        for e in range(256):
            v, lane = e // 64, e % 64
            k0, k1 = lane // 16 + 4 * v, lane % 16
            if k0 >= 15:
                assert got[e] == 0
                continue
            m = (15 * k1 + 16 * k0) % 240
            th = 2 * np.pi * m / 240
            want = ((1 + np.cos(th)) / 2) ** tt * np.exp(-2j * np.pi * ((m * tt) % 240) / 240) / 240
            assert abs(got[e] - want) < 1e-16, (e, tt)
    plan.close()


def test_spectrum_3d(ffs, torch_cuda):
    # entry (k2 block, reg v, lane): k0 = lane % 16, k1 = q + 4 v (FP64 order), k2 = block
    _, r, w, _ = ffs.named_stencil("Box3D27P", seed=7)
    plan = ffs.Plan(3, (40, 40, 40), w, r, 2, precision="fp64", fuse=2)
    main, _ = spectrum(plan, 4096)
    Kf = kernel_dft(w, 1, 3, 16, "ref")
    want = np.zeros(4096, complex)
    for e in range(4096):
        kb, v, lane = e // 256, (e // 64) % 4, e % 64
        want[e] = Kf[lane % 16, lane // 16 + 4 * v, kb] ** 2 / 4096
    assert np.abs(main - want).max() < 1e-15 * np.abs(want).max() * 8
    plan.close()


def test_read_table_errors(ffs, torch_cuda):
    _, r, w, _ = ffs.named_stencil("Box2D9P", seed=1)
    plan = ffs.Plan(2, (64, 64), w, r, 3, precision="fp64", tile=16)
    with pytest.raises(ffs.FFSError, match="UNSUPPORTED"):
        plan.read_table(1)  # no mixed operands in an FP64 plan
    with pytest.raises(ffs.FFSError, match="INVALID_ARG"):
        plan.read_table(9)
    plan.close()
    direct = ffs.Plan(2, (64, 64), w, r, 3, precision="fp64", algorithm="direct")
    with pytest.raises(ffs.FFSError, match="UNSUPPORTED"):
        direct.read_table(0)
    direct.close()
