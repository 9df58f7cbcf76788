"""GPU parity of the FP64 form against the FP64 direct-sweep oracle, through the C ABI.

Bar: max|gpu - oracle| / max|oracle| <= 1e-12 for the FP64 form (the stated drop-in tolerance).
Cases follow the reference's shapes and its boundary convention (periodic, ref index convention),
plus the cases the reference cannot run: sizes not divisible by the valid width, grids smaller than a
tile, odd tile counts, T not a multiple of the fused depth, T = 0.
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


SHAPES_2D = ["Heat-2D", "Box2D9P", "Star2D9P", "Box2D25P"]


@pytest.mark.parametrize("shape", SHAPES_2D)
@pytest.mark.parametrize("size", [(37, 53), (16, 16), (5, 7), (64, 40), (1, 19)])
@pytest.mark.parametrize("steps", [1, 2, 5])
def test_2d_fp64_named(ffs, oracle, torch_cuda, shape, size, steps):
    dims, r, w, _ = ffs.named_stencil(shape, seed=7)
    x = oracle.uniform(1234 + size[0] * 31 + size[1], size)
    got, info = _run(ffs, torch_cuda, x, r, w, steps)
    ref = oracle.sweep(x, r, w, steps)
    _, rel, _ = oracle.errors(got, ref)
    print(shape, size, steps, info["fuse"], info["fuse_tail"], rel)
    assert rel <= TOL_FP64, (shape, size, steps, info, rel)


@pytest.mark.parametrize("convention", ["ref", "centred"])
@pytest.mark.parametrize("fuse", [1, 2, 3])
def test_2d_fp64_fuse_and_convention(ffs, oracle, torch_cuda, convention, fuse):
    rng = np.random.default_rng(fuse)
    w = rng.random(9)
    w /= w.sum()
    x = oracle.uniform(99, (45, 61))
    conv = 0 if convention == "ref" else 1
    for steps in (fuse, 2 * fuse + 1, 7):
        got, info = _run(ffs, torch_cuda, x, 1, w, steps, convention=convention, fuse=fuse)
        ref = oracle.sweep(x, 1, w, steps, convention=conv)
        _, rel, _ = oracle.errors(got, ref)
        assert rel <= TOL_FP64, (convention, fuse, steps, info, rel)


def test_2d_unnormalised_weights(ffs, oracle, torch_cuda):
    # weights that neither sum to 1 nor are positive: the spectrum power must still be exact
    rng = np.random.default_rng(3)
    w = rng.standard_normal(25) * 0.3
    x = oracle.uniform(5, (50, 50))
    got, _ = _run(ffs, torch_cuda, x, 2, w, 3)
    ref = oracle.sweep(x, 2, w, 3)
    _, rel, _ = oracle.errors(got, ref)
    assert rel <= TOL_FP64


@pytest.mark.parametrize("shape", ["Heat-1D", "1D5P", "1D7P"])
@pytest.mark.parametrize("n", [1, 7, 240, 241, 1000, 4099])
@pytest.mark.parametrize("steps", [1, 13, 64])
def test_1d_fp64_named(ffs, oracle, torch_cuda, shape, n, steps):
    dims, r, w, _ = ffs.named_stencil(shape, seed=11)
    x = oracle.uniform(777 + n, (n,))
    got, info = _run(ffs, torch_cuda, x, r, w, steps)
    ref = oracle.sweep(x, r, w, steps)
    _, rel, _ = oracle.errors(got, ref)
    assert rel <= TOL_FP64, (shape, n, steps, info, rel)


def test_1d_centred_and_forced_fuse(ffs, oracle, torch_cuda):
    rng = np.random.default_rng(8)
    w = rng.random(5)
    x = oracle.uniform(17, (3001,))
    for fuse in (1, 7, 50):
        for conv in (0, 1):
            got, info = _run(ffs, torch_cuda, x, 2, w / w.sum(), 101, convention=conv, fuse=fuse)
            ref = oracle.sweep(x, 2, w / w.sum(), 101, convention=conv)
            _, rel, _ = oracle.errors(got, ref)
            assert rel <= TOL_FP64, (fuse, conv, info, rel)


def test_zero_steps_is_copy(ffs, oracle, torch_cuda):
    _, r, w, _ = ffs.named_stencil("Box2D9P")
    x = oracle.uniform(1, (33, 17))
    got, _ = _run(ffs, torch_cuda, x, r, w, 0)
    assert np.array_equal(got, x)


def test_deterministic_and_graph_equals_plain(ffs, oracle, torch_cuda):
    torch = torch_cuda
    _, r, w, _ = ffs.named_stencil("Box2D25P", seed=3)
    x = oracle.uniform(2, (200, 150))
    d_in = torch.from_numpy(x).cuda()
    outs = []
    for use_graph in (True, True, False):
        plan = ffs.Plan(2, x.shape, w, r, 9, use_graph=use_graph)
        d_out = torch.empty_like(d_in)
        plan.run(d_in, d_out)
        plan.run(d_in, d_out)  # replay
        torch.cuda.synchronize()
        outs.append(d_out.cpu().numpy())
        plan.close()
    assert np.array_equal(outs[0], outs[1])
    assert np.array_equal(outs[0], outs[2])


def test_run_host_matches_device(ffs, oracle, torch_cuda):
    _, r, w, _ = ffs.named_stencil("Heat-2D", seed=1)
    x = oracle.uniform(3, (70, 30))
    plan = ffs.Plan(2, x.shape, w, r, 6)
    got = plan.run_host(x)
    plan.close()
    ref = oracle.sweep(x, r, w, 6)
    assert oracle.errors(got, ref)[1] <= TOL_FP64


@pytest.mark.parametrize("shape", ["Heat-3D", "Box3D27P"])
@pytest.mark.parametrize("size", [(16, 16, 16), (11, 13, 17), (20, 9, 33), (5, 5, 5), (40, 24, 30)])
@pytest.mark.parametrize("steps", [1, 2, 5])
def test_3d_fp64_named(ffs, oracle, torch_cuda, shape, size, steps):
    _, r, w, _ = ffs.named_stencil(shape, seed=13)
    x = oracle.uniform(3000 + sum(size), size)
    got, info = _run(ffs, torch_cuda, x, r, w, steps)
    ref = oracle.sweep(x, r, w, steps)
    _, rel, _ = oracle.errors(got, ref)
    assert rel <= TOL_FP64, (shape, size, steps, info, rel)


@pytest.mark.parametrize("convention", ["ref", "centred"])
@pytest.mark.parametrize("fuse", [1, 3])
def test_3d_fp64_fuse_convention(ffs, oracle, torch_cuda, convention, fuse):
    rng = np.random.default_rng(70 + fuse)
    w = rng.random(27)
    w /= w.sum()
    x = oracle.uniform(71, (18, 22, 26))
    got, info = _run(ffs, torch_cuda, x, 1, w, 2 * fuse + 1, convention=convention, fuse=fuse)
    ref = oracle.sweep(x, 1, w, 2 * fuse + 1, convention=0 if convention == "ref" else 1)
    assert oracle.errors(got, ref)[1] <= TOL_FP64, info


@pytest.mark.parametrize("precision", ["fp64", "f16", "bf16"])
def test_linear_mapping_is_bitwise_equal(ffs, oracle, torch_cuda, precision):
    # the wave-to-tile mapping only reorders independent tile pairs: results are bitwise identical
    torch = torch_cuda
    _, r, w, _ = ffs.named_stencil("Box2D25P", seed=8)
    x = oracle.uniform(31, (333, 290))
    d_in = torch.from_numpy(x).cuda()
    outs = []
    for mapping in ("xcd", "linear"):
        plan = ffs.Plan(2, x.shape, w, r, 11, precision=precision, mapping=mapping)
        assert plan.info["mapping"] == (1 if mapping == "xcd" else 2)
        d_out = torch.empty_like(d_in)
        plan.run(d_in, d_out)
        torch.cuda.synchronize()
        outs.append(d_out.cpu().numpy())
        plan.close()
    assert np.array_equal(outs[0], outs[1])


@pytest.mark.parametrize("algorithm,precision", [("fft", "fp64"), ("fft", "f16"), ("direct", "fp64")])
def test_run_captured_in_caller_graph(ffs, oracle, torch_cuda, algorithm, precision):
    # with use_graph = 0 a run only enqueues kernels on the caller's stream, so it can be recorded into the
    # caller's own graph; every replay (with eager work in between) must equal the eager result
    torch = torch_cuda
    _, r, w, _ = ffs.named_stencil("Box2D25P", seed=12)
    x = oracle.uniform(41, (200, 180))
    d_in = torch.from_numpy(x).cuda()
    plan = ffs.Plan(2, x.shape, w, r, 9, precision=precision, algorithm=algorithm, use_graph=False)
    eager = torch.empty_like(d_in)
    plan.run(d_in, eager)
    torch.cuda.synchronize()
    out = torch.empty_like(d_in)
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        plan.run(d_in, out)  # warm-up on the side stream before capture
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    with torch.cuda.graph(g):
        plan.run(d_in, out)
    for k in range(3):
        out.zero_()
        scratch = torch.ones(1000, device="cuda") * k  # eager work between replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(out, eager), (k, float(scratch.sum()))
    plan.close()


def test_fill_uniform_device_matches_host(ffs, oracle, torch_cuda):
    torch = torch_cuda
    d = torch.empty(100003, dtype=torch.float64, device="cuda")
    ffs.fill_uniform_device(0x5EED0002, d)
    torch.cuda.synchronize()
    assert np.array_equal(d.cpu().numpy(), ffs.fill_uniform(0x5EED0002, d.numel()))


def test_planner_heat1d_single_launch(ffs, torch_cuda):
    # 2^20 points x 64 steps: the run is launch-bound, and the planner's whole-run model takes the one
    # block of depth 64 (V = 112 of the 240-point window) over two blocks of depth 32.
    _, r, w, _ = ffs.named_stencil("Heat-1D", seed=1)
    plan = ffs.Plan(1, (1 << 20,), w, r, 64, precision="fp64")
    info = plan.info
    plan.close()
    assert info["fuse"] == 64 and info["blocks"] == 1 and info["valid"][0] == 112, info


@pytest.mark.parametrize("precision", ["fp64", "f16"])
def test_graph_cache_cycles_buffer_pairs(ffs, oracle, torch_cuda, precision):
    # one plan driven with five (d_in, d_out) pairs in turn, twice round: the plan keeps the captured chains
    # of four pairs, so the fifth evicts the least recently used one and every pair is re-captured on its
    # next turn; every run must equal the first run of its input bitwise, and the FP64 ones the oracle
    torch = torch_cuda
    _, r, w, _ = ffs.named_stencil("Box2D9P", seed=13)
    plan = ffs.Plan(2, (96, 80), w, r, 48, precision=precision, fuse=4)
    assert plan.info["blocks"] > 4  # graph replay is used
    xs = [oracle.uniform(60 + i, (96, 80)) for i in range(5)]
    ins = [torch.from_numpy(x).cuda() for x in xs]
    outs = [torch.empty_like(d) for d in ins]
    first = []
    for rnd in range(2):
        for i in range(5):
            outs[i].fill_(float("nan"))
            plan.run(ins[i], outs[i])
            torch.cuda.synchronize()
            if rnd == 0:
                first.append(outs[i].clone())
            else:
                assert torch.equal(outs[i], first[i]), i
    tol = 1e-12 if precision == "fp64" else 1e-6
    for i in range(5):
        assert oracle.errors(first[i].cpu().numpy(), oracle.sweep(xs[i], r, w, 48))[1] <= tol
    plan.close()


@pytest.mark.parametrize("radius", [2, 3])
@pytest.mark.parametrize("precision,algorithm", [("fp64", "fft"), ("bf16", "fft"), ("f32", "fft"), ("fp64", "auto")])
@pytest.mark.parametrize("convention", ["ref", "centred"])
def test_3d_wider_stencils(ffs, oracle, torch_cuda, radius, precision, algorithm, convention):
    """Custom 3D stencils wider than the named ones (5^3 and 7^3 random weights): the 16^3 tiles then
    keep V = 16 - 2 r t (down to 4), with tails; ragged grids; against the oracle at the form's bar.
    AUTO takes the FFT path here (the direct kernels stop at radius 1 in 3D)."""
    rng = np.random.default_rng(40 + radius)
    w = rng.random((2 * radius + 1) ** 3)
    w /= w.sum()
    tol = 1e-12 if precision == "fp64" else 1e-6
    for sizes, steps in [((23, 19, 29), 1), ((17, 26, 21), 3), ((30, 14, 33), 5)]:
        x = oracle.uniform(7 * radius + steps, sizes)
        try:
            plan = ffs.Plan(3, sizes, w, radius, steps, precision=precision, algorithm=algorithm,
                            convention=convention)
        except ffs.FFSError as e:
            pytest.fail(f"plan rejected: radius {radius} {precision} {algorithm}: {e}")
        d_in = torch_cuda.from_numpy(x).cuda()
        d_out = torch_cuda.empty_like(d_in)
        plan.run(d_in, d_out)
        torch_cuda.cuda.synchronize()
        info = plan.info
        plan.close()
        ref = oracle.sweep(x, radius, w, steps, convention=0 if convention == "ref" else 1)
        rel = oracle.errors(d_out.cpu().numpy(), ref)[1]
        assert rel <= tol, (radius, precision, algorithm, convention, sizes, steps, info, rel)


@pytest.mark.parametrize("radius", [2, 3])
def test_3d_direct_refuses_wider_stencils(ffs, torch_cuda, radius):
    """The direct kernels cover radius <= 1 in 3D; wider stencils are a clear FFS_ERR_UNSUPPORTED."""
    w = np.full((2 * radius + 1) ** 3, 1.0 / (2 * radius + 1) ** 3)
    with pytest.raises(ffs.FFSError, match="not offered in 3D"):
        ffs.Plan(3, (20, 20, 20), w, radius, 2, precision="fp64", algorithm="direct")


@pytest.mark.parametrize("dims,radius", [(2, 5), (2, 7), (1, 5), (1, 12)])
@pytest.mark.parametrize("weights", ["signed", "growing"])
def test_wide_stencils_unusual_weights(ffs, oracle, torch_cuda, dims, radius, weights):
    """Wide custom stencils in 1D and 2D, every FFT form the dimension offers, with weights that are not
    a positive normalised set: signed taps summing to 1 (an oscillating kernel) or positive taps summing
    to 1.3 (a growing field, so the DC gain (sum w)^t and the mixed forms' tile scales are exercised)."""
    rng = np.random.default_rng(10 * dims + radius)
    n = (2 * radius + 1) ** dims
    if weights == "signed":
        w = rng.random(n) - 0.45
        w /= w.sum()
    else:
        w = rng.random(n)
        w *= 1.3 / w.sum()
    sizes = (3001,) if dims == 1 else (71, 90)
    steps = 6
    x = oracle.uniform(31 * radius, sizes)
    ref = oracle.sweep(x, radius, w, steps)
    forms = ["fp64", "bf16", "f32"] + (["f16"] if dims == 2 else [])
    for form in forms:
        plan = ffs.Plan(dims, sizes, w, radius, steps, precision=form)
        d_in = torch_cuda.from_numpy(x).cuda()
        d_out = torch_cuda.empty_like(d_in)
        plan.run(d_in, d_out)
        torch_cuda.cuda.synchronize()
        info = plan.info
        plan.close()
        rel = oracle.errors(d_out.cpu().numpy(), ref)[1]
        assert rel <= (1e-12 if form == "fp64" else 1e-6), (dims, radius, weights, form, info, rel)


@pytest.mark.parametrize("precision", ["fp64", "bf16"])
def test_3d_mappings_bitwise_equal(ffs, oracle, torch_cuda, precision):
    """The 3D kernel's wave-to-tile orders (XCD bands: the mixed form's default; hardware order: the FP64
    form's) only reorder independent tile pairs: bitwise equal results, both against the oracle."""
    _, r, w, _ = ffs.named_stencil("Box3D27P", seed=5)
    x = oracle.uniform(77, (40, 24, 30))
    outs = []
    for mapping in ("xcd", "linear", "auto"):
        plan = ffs.Plan(3, x.shape, w, r, 5, precision=precision, mapping=mapping)
        info = plan.info
        d_in = torch_cuda.from_numpy(x).cuda()
        d_out = torch_cuda.empty_like(d_in)
        plan.run(d_in, d_out)
        torch_cuda.cuda.synchronize()
        plan.close()
        outs.append(d_out.cpu().numpy())
        if mapping == "auto":
            assert info["mapping"] == (2 if precision == "fp64" else 1), info
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])
    assert oracle.errors(outs[0], oracle.sweep(x, r, w, 5))[1] <= (1e-12 if precision == "fp64" else 1e-6)
