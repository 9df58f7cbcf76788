"""ffs_run_static, the entry point that allocates no device memory, against ffs_plan_create + ffs_run.

Several calls in a row on one GPU, each input smaller than the one before (the static array keeps the
previous call's tables and grid, which every call must overwrite before reading), across the families
that use the array differently: 32-point FP16 with reduced inverse stages and a tail block, 16-point
FP64, 3D BF16, the 1D prime-factor window, the direct kernel (ping-pong grid only) and a single-block run
(no ping-pong grid).  Each result must be bitwise the planned run's, and within the form's bar of the
oracle (FP64 1e-12, mixed 1e-6).  Sizes beyond the array are refused before any device work.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [  # (shape, sizes, steps, keyword arguments), largest first
    ("Box2D25P", (1536, 1200), 9, dict(precision="f16")),
    ("Box2D9P", (700, 513), 7, dict(precision="fp64", tile=16)),
    ("Box3D27P", (40, 37, 44), 5, dict(precision="bf16")),
    ("1D7P", (5003,), 40, dict(precision="fp64")),
    ("Box2D25P", (200, 130), 6, dict(precision="fp64", algorithm="direct")),
    ("Box2D9P", (96, 80), 3, dict(precision="bf16", fuse=3)),
]


def test_run_static_matches_planned_runs(ffs, oracle, torch_cuda):
    torch = torch_cuda
    for shape, sizes, steps, kw in CASES:
        dims, r, w, _ = ffs.named_stencil(shape, seed=0x5EED0010 + len(sizes))
        x = oracle.uniform(0x5EED0020 + sum(sizes), sizes)
        d_in = torch.from_numpy(x).cuda()
        planned, static = torch.empty_like(d_in), torch.empty_like(d_in)
        static.fill_(float("nan"))
        with ffs.Plan(dims, sizes, w, r, steps, **kw) as plan:
            plan.run(d_in, planned)
        ffs.run_static(dims, sizes, w, r, steps, d_in, static, **kw)
        torch.cuda.synchronize()
        assert torch.equal(planned, static), (shape, sizes, kw, float((planned - static).abs().max()))
        tol = 1e-12 if kw["precision"] == "fp64" else 1e-6
        assert oracle.errors(static.cpu().numpy(), oracle.sweep(x, r, w, steps))[1] <= tol, (shape, sizes, kw)


def test_run_static_twice_on_one_stream(ffs, oracle, torch_cuda):
    # two calls queued back to back without a synchronisation between them, the second smaller: stream
    # order alone must keep the second call's table writes behind the first call's kernels
    torch = torch_cuda
    _, r, w, _ = ffs.named_stencil("Box2D25P", seed=3)
    big = oracle.uniform(1, (1024, 1100))
    small = oracle.uniform(2, (300, 260))
    outs = []
    for x, prec in ((big, "f16"), (small, "fp64")):
        d_in = torch.from_numpy(x).cuda()
        d_out = torch.empty_like(d_in)
        ffs.run_static(2, x.shape, w, r, 12, d_in, d_out, precision=prec)
        outs.append((x, d_in, d_out, prec))
    torch.cuda.synchronize()
    for x, d_in, d_out, prec in outs:
        with ffs.Plan(2, x.shape, w, r, 12, precision=prec) as plan:
            ref = torch.empty_like(d_in)
            plan.run(d_in, ref)
        torch.cuda.synchronize()
        assert torch.equal(ref, d_out), prec


def test_run_static_rejects_grids_beyond_the_array(ffs, torch_cuda):
    # 2900^2 points x 8 B > 63 MiB of ping-pong space: refused on the host, nothing enqueued, the output
    # untouched
    torch = torch_cuda
    w = np.full(9, 1 / 9)
    d_in = torch.zeros((2900, 2900), dtype=torch.float64, device="cuda")
    d_out = torch.full_like(d_in, 7.0)
    with pytest.raises(ffs.FFSError, match="UNSUPPORTED"):
        ffs.run_static(2, (2900, 2900), w, 1, 20, d_in, d_out, precision="fp64")
    torch.cuda.synchronize()
    assert bool((d_out == 7.0).all())
    small = d_in[:64, :64].contiguous()
    with pytest.raises(ffs.FFSError, match="INVALID_ARG"):
        ffs.run_static(2, (64, 64), w, 1, 20, small, small, precision="fp64")  # in == out
