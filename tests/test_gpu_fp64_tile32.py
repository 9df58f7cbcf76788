"""GPU parity of the FP64 form on 32-point tiles (k_tile32_f64) against the FP64 direct-sweep oracle.

Bar: max|gpu - oracle| / max|oracle| <= 1e-12, as for the 16-point FP64 form.  The depths cover the
reduced inverse stages (valid width V = 32 - 2 r t <= 16, at the window offset of either index
convention) and the full ones, with tail blocks of another depth, and grids smaller than a tile.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-12


def _run(ffs, torch, x, radius, w, steps, **kw):
    plan = ffs.Plan(x.ndim, x.shape, w, radius, steps, **kw)
    d_in = torch.from_numpy(np.ascontiguousarray(x)).cuda()
    d_out = torch.empty_like(d_in)
    plan.run(d_in, d_out)
    torch.cuda.synchronize()
    info = plan.info
    plan.close()
    return d_out.cpu().numpy(), info


@pytest.mark.parametrize("shape", ["Heat-2D", "Box2D9P", "Star2D9P", "Box2D25P"])
@pytest.mark.parametrize("size", [(37, 53), (16, 16), (5, 7), (64, 40), (1, 19)])
@pytest.mark.parametrize("steps", [1, 4, 9, 17])
def test_2d_fp64_tile32_named(ffs, oracle, torch_cuda, shape, size, steps):
    _, r, w, _ = ffs.named_stencil(shape, seed=11)
    x = oracle.uniform(777 + size[0] * 31 + size[1], size)
    got, info = _run(ffs, torch_cuda, x, r, w, steps, tile=32)
    assert info["tile"][:2] == [32, 32]
    ref = oracle.sweep(x, r, w, steps)
    rel = oracle.errors(got, ref)[1]
    assert rel <= TOL_FP64, (shape, size, steps, info, rel)


@pytest.mark.parametrize("convention", ["ref", "centred"])
@pytest.mark.parametrize("radius,fuse", [(1, 7), (1, 8), (2, 3), (2, 4), (3, 2), (3, 3), (4, 2)])
def test_2d_fp64_tile32_depths(ffs, oracle, torch_cuda, convention, radius, fuse):
    rng = np.random.default_rng(10 * radius + fuse)
    w = rng.random((2 * radius + 1) ** 2)
    w /= w.sum()
    x = oracle.uniform(5 + fuse, (70, 45))
    steps = 2 * fuse + 1
    got, info = _run(ffs, torch_cuda, x, radius, w, steps, tile=32, fuse=fuse, convention=convention)
    assert info["valid"][0] == 32 - 2 * radius * fuse
    ref = oracle.sweep(x, radius, w, steps, convention=0 if convention == "ref" else 1)
    assert oracle.errors(got, ref)[1] <= TOL_FP64, (convention, radius, fuse, info)


def test_2d_fp64_tile32_matches_tile16_at_scale(ffs, torch_cuda):
    # 2048^2 x 16 steps: the two FP64 tile families agree to FP64 rounding (size-independent check)
    torch = torch_cuda
    _, r, w, _ = ffs.named_stencil("Box2D25P", seed=0x5EED0002)
    n = 2048
    d_in = torch.empty(n * n, dtype=torch.float64, device="cuda").reshape(n, n)
    ffs.fill_uniform_device(99, d_in)
    outs = []
    for tile in (16, 32):
        plan = ffs.Plan(2, (n, n), w, r, 16, tile=tile)
        d_out = torch.empty_like(d_in)
        plan.run(d_in, d_out)
        torch.cuda.synchronize()
        plan.close()
        outs.append(d_out)
    rel = float((outs[0] - outs[1]).abs().max() / outs[0].abs().max())
    assert rel <= TOL_FP64, rel


@pytest.mark.parametrize("n1", [1 << 22, (1 << 22) + 96])
@pytest.mark.parametrize("precision", ["fp64", "f16", "bf16"])
def test_2d_tile32_long_rows(ffs, oracle, torch_cuda, n1, precision):
    # Rows of 2^22 columns are the widest whose interior windows are read through the 32-bit buffer
    # offsets; wider rows take the general gather for every window.  Both agree with the oracle.
    _, r, w, _ = ffs.named_stencil("Heat-2D", seed=3)
    x = oracle.uniform(4242 + n1, (34, n1))
    got, info = _run(ffs, torch_cuda, x, r, w, 3, tile=32, precision=precision)
    assert info["tile"][:2] == [32, 32]
    ref = oracle.sweep(x, r, w, 3)
    tol = TOL_FP64 if precision == "fp64" else 1e-6
    assert oracle.errors(got, ref)[1] <= tol, (precision, n1, info)
