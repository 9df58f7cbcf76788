"""Writes tests/golden/golden.npz: input/output vectors of one or more periodic stencil steps.

Sources (no values come from this repository's own oracle):
  * the reference's own CPU check functions, compiled from its sources into oracle/_ref by
    oracle/Makefile (`make -C oracle ref`): stencil1D (src/1D/check_correct.hpp:15-38), the 1D driver
    check composition (src/1D/1d_main.cu:220-223) and the 2D one (src/2D/2d_main.cu:190-192 over
    src/2D/check_correct.hpp:17-100), applied step by step;
  * exact rational arithmetic (fractions.Fraction) of out[p] = sum_a w[a] in[(p - a) mod n] for small
    1D/2D/3D cases with dyadic data, rounded once to FP64 at the end.
Each case stores <name>__in, <name>__w, <name>__meta = (radius, steps, convention), <name>__out.
Run from the repository root: python tests/golden/make_golden.py
"""
import itertools
import os
import sys
from fractions import Fraction

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from tests import _oracle as O  # noqa: E402  (only for the _ref loader and pointer helper)


def ref_cases():
    out = {}
    L1, L2 = O.ref_lib(1), O.ref_lib(2)
    if L1 is None or L2 is None:
        raise SystemExit("build the reference check code first: make -C oracle ref")
    rng = np.random.default_rng(20250301)
    for n, k, steps in ((37, 5, 3), (56, 29, 1), (40, 3, 4)):
        x = rng.random(n)
        w = rng.random(k)
        cur = x.copy()
        for _ in range(steps):
            nxt = np.empty(n)
            L1.ref_step1d_driver(O._p(cur), n, O._p(w), k, O._p(nxt))
            cur = nxt
        out[f"ref1d_driver_n{n}_k{k}_T{steps}"] = (x, w, (k // 2, steps, 0), cur)
    for n, k, steps in ((50, 3, 2), (17, 7, 3)):
        x = rng.random(n)
        w = rng.random(k)
        cur = x.copy()
        for _ in range(steps):
            nxt = np.empty(n)
            L1.ref_stencil1d(O._p(cur), n, O._p(w), k, O._p(nxt))
            cur = nxt
        out[f"ref1d_centred_n{n}_k{k}_T{steps}"] = (x, w, (k // 2, steps, 1), cur)
    for width, k, steps in ((19, 3, 2), (12, 5, 1), (24, 3, 3)):
        x = rng.random((width, width))
        w = rng.random(k * k)
        cur = x.copy()
        for _ in range(steps):
            nxt = np.empty_like(cur)
            L2.ref_step2d_driver(O._p(np.ascontiguousarray(cur)), width, O._p(w), k, O._p(nxt))
            cur = nxt
        out[f"ref2d_driver_w{width}_k{k}_T{steps}"] = (x, w, (k // 2, steps, 0), cur)
    return out


def exact_step(x, w, radius, convention):
    shape = x.shape
    K = 2 * radius + 1
    res = np.empty(shape, dtype=object)
    for p in itertools.product(*(range(n) for n in shape)):
        acc = Fraction(0)
        for a in itertools.product(range(K), repeat=len(shape)):
            off = [(ai if convention == 0 else radius - ai) for ai in a]
            src = tuple((pi - oi) % n for pi, oi, n in zip(p, off, shape))
            acc += w[a] * x[src]
        res[p] = acc
    return res


def exact_cases():
    out = {}
    rng = np.random.default_rng(7)
    for shape, radius, steps, conv in (((13,), 1, 5, 0), ((6, 7), 1, 2, 1), ((4, 5, 3), 1, 2, 0), ((9, 8), 2, 2, 0)):
        xi = rng.integers(0, 1024, size=shape)
        wi = rng.integers(1, 16, size=(2 * radius + 1,) * len(shape))
        x = np.vectorize(lambda v: Fraction(int(v), 1024), otypes=[object])(xi)
        wsum = int(wi.sum())
        w = np.vectorize(lambda v: Fraction(int(v), wsum), otypes=[object])(wi)
        cur = x
        for _ in range(steps):
            cur = exact_step(cur, w, radius, conv)
        name = f"exact_{len(shape)}d_{'x'.join(map(str, shape))}_r{radius}_T{steps}_c{conv}"
        out[name] = (xi / 1024.0, (wi / wsum).ravel(), (radius, steps, conv), np.vectorize(float)(cur).astype(np.float64))
    return out


def main():
    cases = {}
    cases.update(ref_cases())
    cases.update(exact_cases())
    arrays = {}
    for name, (x, w, meta, y) in cases.items():
        arrays[name + "__in"] = np.asarray(x, dtype=np.float64)
        arrays[name + "__w"] = np.asarray(w, dtype=np.float64).ravel()
        arrays[name + "__meta"] = np.asarray(meta, dtype=np.int64)
        arrays[name + "__out"] = np.asarray(y, dtype=np.float64)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden.npz")
    np.savez_compressed(path, **arrays)
    print(f"wrote {len(cases)} cases to {path}")


if __name__ == "__main__":
    main()
