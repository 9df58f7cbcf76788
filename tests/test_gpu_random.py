"""Seeded random configurations against the FP64 direct-sweep oracle.

Each case draws the dimension, grid sizes (small, so the oracle runs in milliseconds), radius, random
weights normalised to sum 1, step count, index convention, precision form and (2D FP64) tile size, and
lets the planner choose the depth, so planner choices, tails, wrap-around windows and reduced inverse
stages are exercised together.  Bars: FP64 <= 1e-12, mixed forms <= 1e-6 (max|gpu - oracle| / max|oracle|).
"""
import random

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(rng):
    dims = rng.choice([1, 2, 2, 2, 3])
    r = rng.choice([1, 2, 3]) if dims < 3 else 1
    if dims == 1:
        sizes, precs = (rng.randint(20, 3000),), ["fp64", "bf16", "f32"]
    elif dims == 2:
        sizes, precs = (rng.randint(3, 150), rng.randint(3, 150)), ["fp64", "f16", "bf16", "f32"]
    else:
        sizes, precs = (rng.randint(3, 40), rng.randint(3, 40), rng.randint(3, 40)), ["fp64", "bf16", "f32"]
    kw = dict(precision=rng.choice(precs), convention=rng.choice(["ref", "centred"]))
    if dims == 2 and kw["precision"] == "fp64" and rng.random() < 0.5:
        kw["tile"] = rng.choice([16, 32])
    steps = rng.randint(1, 40)
    w = np.array([rng.random() for _ in range((2 * r + 1) ** dims)])
    return dims, sizes, r, w / w.sum(), steps, kw, rng.randint(1, 10 ** 6)


@pytest.mark.parametrize("seed", range(8))
def test_random_configurations(ffs, oracle, torch_cuda, seed):
    rng = random.Random(1000 + seed)
    for _ in range(40):
        dims, sizes, r, w, steps, kw, xseed = _case(rng)
        x = oracle.uniform(xseed, sizes)
        try:
            plan = ffs.Plan(dims, sizes, w, r, steps, **kw)
        except ffs.FFSError:
            continue  # a combination the planner rejects (e.g. radius too large for the window)
        d_in = torch_cuda.from_numpy(np.ascontiguousarray(x)).cuda()
        d_out = torch_cuda.empty_like(d_in)
        plan.run(d_in, d_out)
        torch_cuda.cuda.synchronize()
        info = plan.info
        plan.close()
        ref = oracle.sweep(x, r, w, steps, convention=0 if kw["convention"] == "ref" else 1)
        tol = 1e-12 if kw["precision"] == "fp64" else 1e-6
        rel = oracle.errors(d_out.cpu().numpy(), ref)[1]
        assert rel <= tol, (dims, sizes, r, steps, kw, info, rel)
