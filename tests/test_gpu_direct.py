"""GPU parity of the direct algorithm (FP64 stencil sum with temporal blocking) against the oracle.

Bar: max|gpu - oracle| / max|oracle| <= 1e-12, the FP64 drop-in tolerance.  Cases: every named shape
(box and star supports, radius 0..3), both conventions, every instantiated fused depth, ragged sizes,
grids smaller than one window (multiple periodic wraps inside a window), several segments along axis 0,
T not a multiple of the depth (tail launch), T = 0, and bitwise equality of graph and plain launches.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _run(ffs, torch, x, radius, w, steps, **kw):
    plan = ffs.Plan(x.ndim, x.shape, w, radius, steps, algorithm=kw.pop("algorithm", "direct"), **kw)
    d_in = torch.from_numpy(np.ascontiguousarray(x)).cuda()
    d_out = torch.empty_like(d_in)
    plan.run(d_in, d_out)
    torch.cuda.synchronize()
    info = plan.info
    plan.close()
    return d_out.cpu().numpy(), info


@pytest.mark.parametrize("shape", ["Heat-2D", "Box2D9P", "Star2D9P", "Box2D25P"])
@pytest.mark.parametrize("size", [(37, 53), (5, 7), (1, 19), (64, 40), (300, 700), (19, 1030)])
@pytest.mark.parametrize("steps", [1, 3, 13])
def test_direct_2d_named(ffs, oracle, torch_cuda, shape, size, steps):
    _, r, w, _ = ffs.named_stencil(shape, seed=21)
    x = oracle.uniform(4321 + size[0] + 7 * size[1], size)
    got, info = _run(ffs, torch_cuda, x, r, w, steps)
    assert info["algorithm"] == 1
    ref = oracle.sweep(x, r, w, steps)
    rel = oracle.errors(got, ref)[1]
    assert rel <= TOL, (shape, size, steps, info, rel)


@pytest.mark.parametrize("radius,star,depths", [(0, False, range(1, 9)), (1, False, range(1, 9)), (1, True, range(1, 9)),
                                                (2, False, range(1, 5)), (2, True, range(1, 7)), (3, False, range(1, 3))])
@pytest.mark.parametrize("convention", [0, 1])
def test_direct_2d_every_depth(ffs, oracle, torch_cuda, radius, star, depths, convention):
    rng = np.random.default_rng(radius * 10 + star)
    K = 2 * radius + 1
    w = rng.random((K, K))
    if star:
        mask = np.zeros((K, K), bool)
        mask[radius, :] = mask[:, radius] = True
        w = np.where(mask, w, 0.0)
    w = (w / w.sum()).ravel()
    x = oracle.uniform(55 + radius, (90, 530))
    for t in depths:
        steps = 2 * t + 1  # two full blocks and a tail of 1
        got, info = _run(ffs, torch_cuda, x, radius, w, steps, fuse=t, convention=convention)
        ref = oracle.sweep(x, radius, w, steps, convention=convention)
        rel = oracle.errors(got, ref)[1]
        assert rel <= TOL, (radius, star, t, convention, info, rel)


@pytest.mark.parametrize("shape", ["Heat-1D", "1D5P", "1D7P"])
@pytest.mark.parametrize("n", [1, 7, 241, 2048, 4099, 70001])
@pytest.mark.parametrize("steps", [1, 13, 64, 150])
def test_direct_1d_named(ffs, oracle, torch_cuda, shape, n, steps):
    _, r, w, _ = ffs.named_stencil(shape, seed=23)
    x = oracle.uniform(99 + n, (n,))
    got, info = _run(ffs, torch_cuda, x, r, w, steps)
    ref = oracle.sweep(x, r, w, steps)
    rel = oracle.errors(got, ref)[1]
    assert rel <= TOL, (shape, n, steps, info, rel)


@pytest.mark.parametrize("convention", [0, 1])
def test_direct_1d_radius0_and_centred(ffs, oracle, torch_cuda, convention):
    x = oracle.uniform(5, (3001,))
    for r, w in ((0, np.array([0.999])), (2, np.array([0.1, 0.2, 0.4, 0.2, 0.1]))):
        got, info = _run(ffs, torch_cuda, x, r, w, 77, convention=convention, fuse=7)
        ref = oracle.sweep(x, r, w, 77, convention=convention)
        assert oracle.errors(got, ref)[1] <= TOL, (r, info)


@pytest.mark.parametrize("shape", ["Heat-3D", "Box3D27P"])
@pytest.mark.parametrize("size", [(16, 16, 16), (11, 13, 17), (5, 5, 5), (40, 70, 36), (33, 9, 65)])
@pytest.mark.parametrize("steps", [1, 2, 5])
def test_direct_3d_named(ffs, oracle, torch_cuda, shape, size, steps):
    _, r, w, _ = ffs.named_stencil(shape, seed=29)
    x = oracle.uniform(700 + sum(size), size)
    got, info = _run(ffs, torch_cuda, x, r, w, steps)
    ref = oracle.sweep(x, r, w, steps)
    rel = oracle.errors(got, ref)[1]
    assert rel <= TOL, (shape, size, steps, info, rel)


@pytest.mark.parametrize("convention", [0, 1])
@pytest.mark.parametrize("fuse", [1, 2])
def test_direct_3d_fuse_convention(ffs, oracle, torch_cuda, convention, fuse):
    rng = np.random.default_rng(fuse)
    w = rng.random(27)
    w /= w.sum()
    x = oracle.uniform(8, (24, 70, 40))
    got, info = _run(ffs, torch_cuda, x, 1, w, 2 * fuse + 1, convention=convention, fuse=fuse)
    ref = oracle.sweep(x, 1, w, 2 * fuse + 1, convention=convention)
    assert oracle.errors(got, ref)[1] <= TOL, info


def test_direct_zero_steps_and_graph_bitwise(ffs, oracle, torch_cuda):
    torch = torch_cuda
    _, r, w, _ = ffs.named_stencil("Box2D25P", seed=3)
    x = oracle.uniform(2, (200, 150))
    got, _ = _run(ffs, torch, x, r, w, 0)
    assert np.array_equal(got, x)
    d_in = torch.from_numpy(x).cuda()
    outs = []
    for use_graph in (True, False):  # 41 steps at depth 4: an 11-launch chain, replayed from a graph
        plan = ffs.Plan(2, x.shape, w, r, 41, use_graph=use_graph, algorithm="direct")
        d_out = torch.empty_like(d_in)
        plan.run(d_in, d_out)
        plan.run(d_in, d_out)
        torch.cuda.synchronize()
        outs.append(d_out.cpu().numpy())
        plan.close()
    assert np.array_equal(outs[0], outs[1])


def test_direct_agrees_with_fft_fp64_at_scale(ffs, oracle, torch_cuda):
    # a 2048^2 grid over 24 steps: the two FP64 algorithms agree to the FP64 bar
    torch = torch_cuda
    _, r, w, _ = ffs.named_stencil("Box2D25P", seed=0x5EED0002)
    x = ffs.fill_uniform(0x5EED0003, 2048 * 2048).reshape(2048, 2048)
    a, _ = _run(ffs, torch, x, r, w, 24, algorithm="direct")
    b, _ = _run(ffs, torch, x, r, w, 24, algorithm="fft")
    assert np.abs(a - b).max() / np.abs(b).max() <= TOL


def test_auto_and_precision_rules(ffs, oracle, torch_cuda):
    _, r, w, _ = ffs.named_stencil("Box2D9P", seed=1)
    x = oracle.uniform(3, (50, 60))
    got, info = _run(ffs, torch_cuda, x, r, w, 4, algorithm="auto", precision="bf16")
    assert info["algorithm"] == 1 and info["precision"] == 0  # direct, computed in FP64
    assert oracle.errors(got, oracle.sweep(x, r, w, 4))[1] <= TOL
    with pytest.raises(ffs.FFSError, match="FP64"):
        ffs.Plan(2, (50, 60), w, r, 4, precision="bf16", algorithm="direct")
    with pytest.raises(ffs.FFSError, match="UNSUPPORTED"):
        ffs.Plan(3, (8, 8, 8), np.full(125, 1 / 125), 2, 4, algorithm="direct")
    # auto falls back to the FFT path where the direct kernel has no instance (3D radius 2)
    plan = ffs.Plan(3, (20, 20, 20), np.full(125, 1 / 125), 2, 2, algorithm="auto")
    assert plan.info["algorithm"] == 0
    plan.close()
    # and takes the direct path in 1D as well
    plan = ffs.Plan(1, (5000,), np.array([0.25, 0.5, 0.25]), 1, 8, algorithm="auto")
    assert plan.info["algorithm"] == 1
    plan.close()


def test_direct_run_block_is_one_depth(ffs, oracle, torch_cuda):
    torch = torch_cuda
    _, r, w, _ = ffs.named_stencil("Heat-2D", seed=1)
    x = oracle.uniform(4, (130, 700))
    plan = ffs.Plan(2, x.shape, w, r, 10, fuse=4, algorithm="direct")
    d_in = torch.from_numpy(x).cuda()
    d_out = torch.empty_like(d_in)
    plan.run_block(d_in, d_out)
    torch.cuda.synchronize()
    assert oracle.errors(d_out.cpu().numpy(), oracle.sweep(x, r, w, 4))[1] <= TOL
    plan.run_block(d_in, d_out, tail=True)
    torch.cuda.synchronize()
    assert oracle.errors(d_out.cpu().numpy(), oracle.sweep(x, r, w, 2))[1] <= TOL
    plan.close()
