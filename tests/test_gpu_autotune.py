"""Autotuner on the GPU: it measures the admissible (tile, depth) plans, returns one of them, caches the
winner (a second call reads the cache), and the tuned plan is as exact as any other."""
import json
import os

import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape,size,steps,form", [("Box2D9P", (300, 260), 12, "fp64"),
                                                   ("Box2D25P", (256, 256), 9, "bf16"),
                                                   ("Heat-1D", (5000,), 40, "fp64")])
def test_autotune_picks_admissible_and_caches(ffs, oracle, torch_cuda, tmp_path, shape, size, steps, form):
    dims, r, w, _ = ffs.named_stencil(shape, seed=4)
    cache = str(tmp_path / "tune.jsonl")
    tile, fuse, ms = ffs.autotune(dims, size, w, r, steps, precision=form, cache=cache)
    assert ms > 0 and fuse >= 1
    # one of the form's candidate tiles (2D FP64 and BF16 have both 16 and 32); which one wins is a timing
    assert tile in ((240,) if dims == 1 else (16, 32))
    lines = [json.loads(l) for l in open(cache)]
    assert len(lines) == 1 and lines[0]["tile"] == tile and lines[0]["fuse"] == fuse
    assert lines[0]["mapping"] in (1, 2)
    assert ffs.autotune(dims, size, w, r, steps, precision=form, cache=cache)[:2] == (tile, fuse)
    assert len(open(cache).readlines()) == 1  # served from the cache
    x = oracle.uniform(9, size)
    plan = ffs.Plan(dims, size, w, r, steps, precision=form, tile=tile, fuse=fuse)
    got = plan.run_host(x)
    plan.close()
    tol = 1e-12 if form == "fp64" else 1e-6
    assert oracle.errors(got, oracle.sweep(x, r, w, steps))[1] <= tol


@pytest.mark.parametrize("shape,size,steps,form", [("Box2D25P", (300, 280), 12, "f16"), ("Heat-3D", (40, 36, 50), 6, "fp64")])
def test_autotune_across_algorithms(ffs, oracle, torch_cuda, tmp_path, shape, size, steps, form):
    # algorithm "auto": FFT depths and direct depths are all timed; the winner's plan meets the bar of
    # its own arithmetic (FP64 for a direct winner)
    dims, r, w, _ = ffs.named_stencil(shape, seed=6)
    cache = str(tmp_path / "tune.jsonl")
    algo, tile, fuse, mapping, ms = ffs.autotune_full(dims, size, w, r, steps, precision=form, cache=cache,
                                                      algorithm="auto")
    assert algo in ("fft", "direct") and fuse >= 1 and ms > 0 and mapping in ("auto", "xcd", "linear")
    assert ffs.autotune_full(dims, size, w, r, steps, precision=form, cache=cache, algorithm="auto")[:4] == \
        (algo, tile, fuse, mapping)
    prec = "fp64" if algo == "direct" else form
    x = oracle.uniform(10, size)
    plan = ffs.Plan(dims, size, w, r, steps, precision=prec, tile=tile, fuse=fuse, algorithm=algo, mapping=mapping)
    got = plan.run_host(x)
    plan.close()
    tol = 1e-12 if prec == "fp64" else 1e-6
    assert oracle.errors(got, oracle.sweep(x, r, w, steps))[1] <= tol


def test_autotune_cache_keyed_by_weights(ffs, oracle, torch_cuda, tmp_path):
    # Star2D9P and Box2D25P share dims and radius but not their admissible direct depths (star support
    # allows deeper fusion): tuning both into one cache file must give each its own entry, and a plan
    # from each entry must be creatable.
    cache = str(tmp_path / "tune.jsonl")
    got = {}
    for shape in ("Star2D9P", "Box2D25P"):
        dims, r, w, _ = ffs.named_stencil(shape, seed=8)
        algo, tile, fuse, mapping, ms = ffs.autotune_full(dims, (200, 180), w, r, 10, precision="fp64", cache=cache,
                                                          algorithm="direct")
        plan = ffs.Plan(dims, (200, 180), w, r, 10, precision="fp64", tile=tile, fuse=fuse, algorithm=algo)
        x = oracle.uniform(3, (200, 180))
        assert oracle.errors(plan.run_host(x), oracle.sweep(x, r, w, 10))[1] <= 1e-12
        plan.close()
        got[shape] = fuse
    assert len(open(cache).readlines()) == 2
    assert got["Box2D25P"] <= 4
