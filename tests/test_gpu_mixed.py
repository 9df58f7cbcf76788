"""GPU parity of the mixed-precision forms against the FP64 direct-sweep oracle, through the C ABI.

Bar (stated for the mixed forms): max|gpu - oracle| / max|oracle| <= 1e-6, on the reference's shapes
and boundary convention, including long horizons (T = 256) where per-block rounding accumulates.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL_MIXED = 1e-6


def _run(ffs, torch, x, radius, w, steps, **kw):
    plan = ffs.Plan(x.ndim, x.shape, w, radius, steps, **kw)
    d_in = torch.from_numpy(np.ascontiguousarray(x)).cuda()
    d_out = torch.empty_like(d_in)
    plan.run(d_in, d_out)
    torch.cuda.synchronize()
    info = plan.info
    plan.close()
    return d_out.cpu().numpy(), info


@pytest.mark.parametrize("form", ["bf16", "f32", "f16"])
@pytest.mark.parametrize("shape", ["Heat-2D", "Box2D9P", "Star2D9P", "Box2D25P"])
@pytest.mark.parametrize("size", [(37, 53), (16, 16), (64, 40), (3, 5)])
@pytest.mark.parametrize("steps", [1, 3, 8])
def test_2d_mixed_named(ffs, oracle, torch_cuda, form, shape, size, steps):
    _, r, w, _ = ffs.named_stencil(shape, seed=21)
    x = oracle.uniform(4321 + size[0] * 7 + size[1], size)
    got, info = _run(ffs, torch_cuda, x, r, w, steps, precision=form)
    ref = oracle.sweep(x, r, w, steps)
    _, rel, _ = oracle.errors(got, ref)
    assert rel <= TOL_MIXED, (form, shape, size, steps, info, rel)


@pytest.mark.parametrize("form", ["bf16", "f32", "f16"])
def test_2d_mixed_long_horizon(ffs, oracle, torch_cuda, form):
    _, r, w, _ = ffs.named_stencil("Box2D25P", seed=0x5EED0002)
    x = oracle.uniform(0x5EED0003, (256, 256))
    got, info = _run(ffs, torch_cuda, x, r, w, 256, precision=form)
    ref = oracle.sweep(x, r, w, 256)
    _, rel, l2 = oracle.errors(got, ref)
    print(form, info["fuse"], rel, l2)
    assert rel <= TOL_MIXED, (form, info, rel, l2)


@pytest.mark.parametrize("form", ["bf16", "f32", "f16"])
@pytest.mark.parametrize("convention", ["ref", "centred"])
@pytest.mark.parametrize("fuse", [1, 2, 3])
def test_2d_mixed_fuse_convention(ffs, oracle, torch_cuda, form, convention, fuse):
    rng = np.random.default_rng(40 + fuse)
    w = rng.random(9)
    w /= w.sum()
    x = oracle.uniform(77, (50, 70))
    got, info = _run(ffs, torch_cuda, x, 1, w, 2 * fuse + 1, precision=form, convention=convention, fuse=fuse)
    ref = oracle.sweep(x, 1, w, 2 * fuse + 1, convention=0 if convention == "ref" else 1)
    assert oracle.errors(got, ref)[1] <= TOL_MIXED


# Valid widths on both sides of 16 in the 32-point family: V = 32 - 2 r t <= 16 takes the reduced inverse
# stages (shifted spectrum, one output block), V > 16 the full ones.  Steps 2 t + 1 add a tail block of
# depth 1 whose shift differs from the main block's.
@pytest.mark.parametrize("form", ["bf16", "f16"])
@pytest.mark.parametrize("convention", ["ref", "centred"])
@pytest.mark.parametrize("radius,fuse", [(1, 7), (1, 8), (1, 9), (2, 3), (2, 4), (2, 5), (3, 3), (3, 5)])
def test_2d_mixed_valid_16_boundary(ffs, oracle, torch_cuda, form, convention, radius, fuse):
    rng = np.random.default_rng(100 * radius + fuse)
    w = rng.random((2 * radius + 1) ** 2)
    w /= w.sum()
    x = oracle.uniform(91 + fuse, (53, 71))
    steps = 2 * fuse + 1
    got, info = _run(ffs, torch_cuda, x, radius, w, steps, precision=form, convention=convention, fuse=fuse)
    assert info["tile"][:2] == [32, 32] and info["valid"][0] == 32 - 2 * radius * fuse
    ref = oracle.sweep(x, radius, w, steps, convention=0 if convention == "ref" else 1)
    assert oracle.errors(got, ref)[1] <= TOL_MIXED, (form, convention, radius, fuse, info)


@pytest.mark.parametrize("form", ["bf16", "f32", "f16"])
def test_2d_mixed_gain_not_one(ffs, oracle, torch_cuda, form):
    # weights summing to 1.3: the constant mode is carried in FP64 with gain (sum w)^t
    rng = np.random.default_rng(9)
    w = rng.random(25)
    w *= 1.3 / w.sum()
    x = oracle.uniform(8, (60, 45)) + 5.0
    got, _ = _run(ffs, torch_cuda, x, 2, w, 4, precision=form)
    ref = oracle.sweep(x, 2, w, 4)
    assert oracle.errors(got, ref)[1] <= TOL_MIXED


@pytest.mark.parametrize("form", ["bf16", "f32"])
@pytest.mark.parametrize("shape", ["Heat-1D", "1D5P", "1D7P"])
@pytest.mark.parametrize("n", [5, 240, 1001, 9999])
def test_1d_mixed_named(ffs, oracle, torch_cuda, form, shape, n):
    _, r, w, _ = ffs.named_stencil(shape, seed=5)
    x = oracle.uniform(1000 + n, (n,))
    for steps in (1, 64):
        got, info = _run(ffs, torch_cuda, x, r, w, steps, precision=form)
        ref = oracle.sweep(x, r, w, steps)
        _, rel, _ = oracle.errors(got, ref)
        assert rel <= TOL_MIXED, (form, shape, n, steps, info, rel)


def test_mixed_deterministic(ffs, oracle, torch_cuda):
    torch = torch_cuda
    _, r, w, _ = ffs.named_stencil("Box2D9P", seed=2)
    x = oracle.uniform(6, (300, 200))
    d_in = torch.from_numpy(x).cuda()
    plan = ffs.Plan(2, x.shape, w, r, 20, precision="bf16")
    a = torch.empty_like(d_in)
    b = torch.empty_like(d_in)
    plan.run(d_in, a)
    plan.run(d_in, b)
    torch.cuda.synchronize()
    plan.close()
    assert torch.equal(a, b)


@pytest.mark.parametrize("tile", [16, 32])
@pytest.mark.parametrize("size", [(37, 53), (100, 64), (7, 300)])
def test_2d_bf16_tiles(ffs, oracle, torch_cuda, tile, size):
    form = "bf16"
    _, r, w, _ = ffs.named_stencil("Box2D25P", seed=31)
    x = oracle.uniform(55 + size[0], size)
    for steps in (1, 4, 11):
        got, info = _run(ffs, torch_cuda, x, r, w, steps, precision=form, tile=tile)
        assert info["tile"][:2] == [tile, tile]
        ref = oracle.sweep(x, r, w, steps)
        _, rel, _ = oracle.errors(got, ref)
        assert rel <= TOL_MIXED, (form, tile, size, steps, info, rel)


@pytest.mark.parametrize("form", ["bf16", "f32"])
@pytest.mark.parametrize("shape", ["Heat-3D", "Box3D27P"])
@pytest.mark.parametrize("size", [(16, 16, 16), (11, 13, 17), (40, 24, 30)])
def test_3d_mixed_named(ffs, oracle, torch_cuda, form, shape, size):
    _, r, w, _ = ffs.named_stencil(shape, seed=17)
    x = oracle.uniform(5000 + sum(size), size)
    for steps in (1, 6):
        got, info = _run(ffs, torch_cuda, x, r, w, steps, precision=form)
        ref = oracle.sweep(x, r, w, steps)
        _, rel, _ = oracle.errors(got, ref)
        assert rel <= TOL_MIXED, (form, shape, size, steps, info, rel)


def test_3d_f32_form_is_compensated(ffs, torch_cuda):
    # the 3D FP32 form runs two products per real GEMM (FP32 matrix entries and their low-order terms):
    # twice the matrix flops of the FP32 form in 2D, per 16-point stage
    w3 = np.full(27, 1 / 27)
    with ffs.Plan(3, (32, 32, 32), w3, 1, 2, precision="f32", fuse=2) as p32, \
            ffs.Plan(3, (32, 32, 32), w3, 1, 2, precision="fp64", fuse=2) as p64:
        assert p32.info["tile"][0] == 16
        # per pair: (5 x 16 + V = 12) slice stages (the last inverse stage runs only the valid slices)
        # x 4 real GEMMs x 2 products x 2*16^3 flops
        assert p32.info["valid"][0] == 12
        stages = 5 * 16 + 12
        pairs = p32.info["flops_per_block"] / (stages * 4 * 2 * 2 * 16 ** 3)
        assert pairs == int(pairs) and pairs > 0
        # FP64: Gauss form, 3 real GEMMs per complex product, same pairs
        assert p64.info["flops_per_block"] == pairs * stages * 3 * 2 * 16 ** 3


@pytest.mark.parametrize("form", ["bf16", "f32"])
def test_3d_mixed_sinusoid_field(ffs, oracle, torch_cuda, form):
    # The 3D workload's field scaled down with the same spectral content: random-phase sinusoids with
    # integer wavevectors up to n/32 per axis (16 on 512^3, 2 on 64^3) plus 1e-3 noise, T = 32.
    n = 64
    rng = np.random.default_rng(3)
    idx = np.indices((n, n, n)).astype(np.float64)
    x = np.zeros((n, n, n))
    for _ in range(8):
        k = rng.integers(1, 3, size=3)
        x += rng.random() / 8 * np.cos(2 * np.pi * (k[0] * idx[0] + k[1] * idx[1] + k[2] * idx[2]) / n + 2 * np.pi * rng.random())
    x += 1e-3 * rng.random((n, n, n))
    _, r, w, _ = ffs.named_stencil("Box3D27P", seed=0x5EED0003)
    got, info = _run(ffs, torch_cuda, x, r, w, 32, precision=form)
    ref = oracle.sweep(x, r, w, 32)
    _, rel, l2 = oracle.errors(got, ref)
    print(form, rel, l2)
    assert rel <= TOL_MIXED, (info, rel, l2)


@pytest.mark.parametrize("scale", [1e-30, 1e-6, 1.0, 1e6, 1e30])
def test_2d_f16_scale_invariance(ffs, oracle, torch_cuda, scale):
    # the FP16 form scales every tile pair to its own range: grids of any magnitude meet the same bar
    _, r, w, _ = ffs.named_stencil("Box2D25P", seed=4)
    x = oracle.uniform(12, (96, 80)) * scale
    got, info = _run(ffs, torch_cuda, x, r, w, 7, precision="f16")
    ref = oracle.sweep(x, r, w, 7)
    rel = oracle.errors(got, ref)[1]
    assert rel <= TOL_MIXED, (scale, info, rel)


def test_2d_f16_unnormalised_gain(ffs, oracle, torch_cuda):
    # |K| > 1 at some frequencies (signed weights): the host lowers the scale exponent by log2 |K|^t
    rng = np.random.default_rng(3)
    w = rng.standard_normal(25) * 0.3
    x = oracle.uniform(5, (70, 90))
    got, info = _run(ffs, torch_cuda, x, 2, w, 6, precision="f16")
    ref = oracle.sweep(x, 2, w, 6)
    assert oracle.errors(got, ref)[1] <= TOL_MIXED, info


def test_f16_form_is_2d_only(ffs, torch_cuda):
    with pytest.raises(ffs.FFSError, match="UNSUPPORTED"):
        ffs.Plan(1, (100,), np.array([0.25, 0.5, 0.25]), 1, 4, precision="f16")
    with pytest.raises(ffs.FFSError, match="UNSUPPORTED"):
        ffs.Plan(2, (64, 64), np.full(9, 1 / 9), 1, 4, precision="f16", tile=16)


@pytest.mark.parametrize("form", ["f16", "bf16", "fp64"])
@pytest.mark.parametrize("shape,depth", [("Heat-2D", 8), ("Box2D9P", 8), ("Box2D25P", 4)])
def test_planner_depth_8192(ffs, torch_cuda, form, shape, depth):
    # The cost model's choice on the 8192^2 workloads: the deepest fusion whose valid region is one
    # 16 x 16 block of the 32-point window (V = 32 - 2 r t = 16), not a deeper one with fewer valid points.
    _, r, w, _ = ffs.named_stencil(shape, seed=1)
    plan = ffs.Plan(2, (8192, 8192), w, r, 256, precision=form)
    info = plan.info
    plan.close()
    assert info["tile"][:2] == [32, 32] and info["fuse"] == depth and info["valid"][:2] == [16, 16], info
    # AUTO mapping: XCD bands for the FP64 32-point kernel, hardware order for the mixed forms
    assert info["mapping"] == (1 if form == "fp64" else 2), info


def test_bf16x2_form_withdrawn(ffs, torch_cuda):
    # the 2-term BF16 split misses 1e-6 before the field has smoothed: plan creation refuses it
    with pytest.raises(ffs.FFSError, match="UNSUPPORTED"):
        ffs.Plan(2, (64, 64), np.full(9, 1 / 9), 1, 4, precision="bf16x2")


@pytest.mark.parametrize("form", ["f16", "bf16", "fp64"])
@pytest.mark.parametrize("convention", ["ref", "centred"])
@pytest.mark.parametrize("radius,fuse,size", [(2, 4, (8192, 96)), (1, 8, (100, 70)), (2, 3, (66, 130)),
                                              (3, 2, (40, 77)), (2, 4, (45, 70))])
def test_out16_bitwise_equals_full_stages(ffs, oracle, torch_cuda, form, convention, radius, fuse, size):
    # 32-point tiles with V <= 16: the reduced inverse stages (only the kept output block, from the
    # shifted inverse operands) must give bitwise the values of the full stages at those positions.
    # Steps 2 t + 1: a main block and a tail block of another shift.  (3, 2): V = 20 > 16, full both ways.
    torch = torch_cuda
    rng = np.random.default_rng(700 + 10 * radius + fuse)
    w = rng.random((2 * radius + 1) ** 2)
    w /= w.sum()
    x = oracle.uniform(800 + fuse, size)
    d_in = torch.from_numpy(x).cuda()
    outs = []
    for full in (False, True):
        plan = ffs.Plan(2, size, w, radius, 2 * fuse + 1, precision=form, convention=convention, fuse=fuse, tile=32,
                        full_stages=full)
        d = torch.empty_like(d_in)
        plan.run(d_in, d)
        torch.cuda.synchronize()
        info = plan.info
        plan.close()
        outs.append(d)
    assert info["valid"][0] == 32 - 2 * radius * fuse
    assert torch.equal(outs[0], outs[1]), (form, convention, radius, fuse, float((outs[0] - outs[1]).abs().max()))
