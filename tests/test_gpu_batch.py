"""GPU parity of the multi-grid batch (one persistent launch per fused block over all grids still
running): every grid against the FP64 oracle (<= 1e-6 max-rel) and bitwise against the same grid run
alone through the single-grid plan (same per-tile arithmetic, so identical results)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _specs(ffs, seed):
    shapes = ["Heat-2D", "Star2D9P", "Box2D9P", "Box2D25P"]
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(7):
        shape = shapes[i % 4]
        _, r, w, _ = ffs.named_stencil(shape, seed=100 + i)
        n0, n1 = (int(v) for v in rng.integers(5, 90, size=2))
        steps = int(rng.choice([0, 1, 3, 7, 16, 25]))
        specs.append(dict(shape=shape, radius=r, weights=w, sizes=(n0, n1), steps=steps))
    specs.append(dict(shape="Box2D25P", radius=2, weights=ffs.named_stencil("Box2D25P", 9)[2], sizes=(130, 97), steps=40))
    return specs


@pytest.mark.parametrize("form", ["bf16", "f16", "fp64"])
def test_batch_matches_oracle_and_single_plans(ffs, oracle, torch_cuda, form):
    torch = torch_cuda
    specs = _specs(ffs, 5)
    xs = [oracle.uniform(200 + i, s["sizes"]) for i, s in enumerate(specs)]
    d_in = [torch.from_numpy(x).cuda() for x in xs]
    d_out = [torch.empty_like(d) for d in d_in]
    batch = ffs.Batch(specs, precision=form)
    batch.run(d_in, d_out)
    batch.run(d_in, d_out)  # graph replay
    torch.cuda.synchronize()
    for i, s in enumerate(specs):
        got = d_out[i].cpu().numpy()
        ref = oracle.sweep(xs[i], s["radius"], s["weights"], s["steps"])
        tol = 1e-12 if form == "fp64" else 1e-6
        assert oracle.errors(got, ref)[1] <= tol, (i, s["shape"], s["sizes"], s["steps"], batch.info(i))
        plan = ffs.Plan(2, s["sizes"], s["weights"], s["radius"], s["steps"], precision=form, tile=32)
        single = torch.empty_like(d_in[i])
        plan.run(d_in[i], single)
        torch.cuda.synchronize()
        plan.close()
        assert torch.equal(single, d_out[i]), (i, s["shape"])
    batch.close()


@pytest.mark.parametrize("form", ["bf16", "f16", "fp64"])
def test_batch_mixes_valid_widths(ffs, oracle, torch_cuda, form):
    # one launch holds tile pairs with V <= 16 (reduced inverse stages at window offsets h = 8, 12, 16 and
    # 18) and V > 16 (full ones), in both index conventions
    torch = torch_cuda
    specs = []
    for i, (r, fuse, sizes, steps, conv) in enumerate([(2, 4, (70, 45), 9, "ref"), (2, 3, (33, 61), 7, "ref"),
                                                       (1, 8, (50, 50), 17, "ref"), (1, 5, (40, 29), 11, "ref"),
                                                       (3, 3, (64, 64), 6, "ref"), (2, 4, (45, 70), 9, "centred"),
                                                       (3, 4, (37, 52), 8, "centred")]):
        rng = np.random.default_rng(300 + i)
        w = rng.random((2 * r + 1) ** 2)
        w /= w.sum()
        specs.append(dict(radius=r, weights=w, sizes=sizes, steps=steps, fuse=fuse, convention=conv))
    xs = [oracle.uniform(500 + i, s["sizes"]) for i, s in enumerate(specs)]
    d_in = [torch.from_numpy(x).cuda() for x in xs]
    d_out = [torch.empty_like(d) for d in d_in]
    batch = ffs.Batch(specs, precision=form)
    batch.run(d_in, d_out)
    torch.cuda.synchronize()
    for i, s in enumerate(specs):
        ref = oracle.sweep(xs[i], s["radius"], s["weights"], s["steps"], convention=0 if s["convention"] == "ref" else 1)
        assert oracle.errors(d_out[i].cpu().numpy(), ref)[1] <= (1e-12 if form == "fp64" else 1e-6), (i, batch.info(i))
        plan = ffs.Plan(2, s["sizes"], s["weights"], s["radius"], s["steps"], precision=form, fuse=s["fuse"], tile=32,
                        convention=s["convention"])
        single = torch.empty_like(d_in[i])
        plan.run(d_in[i], single)
        torch.cuda.synchronize()
        plan.close()
        assert torch.equal(single, d_out[i]), i
    batch.close()


def test_batch_rejects_mixed_precisions(ffs):
    _, r, w, _ = ffs.named_stencil("Box2D9P")
    specs = [dict(radius=r, weights=w, sizes=(32, 32), steps=2, precision="bf16"),
             dict(radius=r, weights=w, sizes=(32, 32), steps=2, precision="f16")]
    with pytest.raises(ffs.FFSError):
        ffs.Batch(specs)


@pytest.mark.parametrize("form", ["f16", "fp64"])
def test_batch_autotune(ffs, oracle, torch_cuda, form, tmp_path):
    # the batch tuner times whole batch runs; the depths it returns must be admissible (each batch
    # creates), the tuned batch must meet the form's bar against the oracle and equal single plans of
    # the same depths bitwise, and a second call must come from the cache file (same answer, one line)
    torch = torch_cuda
    specs = _specs(ffs, 11)
    cache = str(tmp_path / "batch_tune.jsonl")
    tuned, ms = ffs.Batch.autotune(specs, precision=form, cache=cache)
    assert ms > 0 and len(tuned) == len(specs)
    again, ms2 = ffs.Batch.autotune(specs, precision=form, cache=cache)
    assert [t["fuse"] for t in again] == [t["fuse"] for t in tuned] and ms2 == pytest.approx(ms, rel=1e-5)
    assert len(open(cache).readlines()) == 1
    xs = [oracle.uniform(900 + i, s["sizes"]) for i, s in enumerate(tuned)]
    d_in = [torch.from_numpy(x).cuda() for x in xs]
    d_out = [torch.empty_like(d) for d in d_in]
    batch = ffs.Batch(tuned, precision=form)
    batch.run(d_in, d_out)
    torch.cuda.synchronize()
    for i, s in enumerate(tuned):
        assert 1 <= s["fuse"] <= max(1, min(15 // s["radius"], s["steps"]))
        ref = oracle.sweep(xs[i], s["radius"], s["weights"], s["steps"])
        assert oracle.errors(d_out[i].cpu().numpy(), ref)[1] <= (1e-12 if form == "fp64" else 1e-6), (i, s["fuse"])
        plan = ffs.Plan(2, s["sizes"], s["weights"], s["radius"], s["steps"], precision=form, fuse=s["fuse"], tile=32,
                        mapping=s["mapping"])
        single = torch.empty_like(d_in[i])
        plan.run(d_in[i], single)
        torch.cuda.synchronize()
        plan.close()
        assert torch.equal(single, d_out[i]), (i, s["fuse"])
    batch.close()
