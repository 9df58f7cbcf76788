"""The rounding of the two MFMA instructions the mixed forms accumulate with, on small known cases
(tests/gpu_probe/mfma_round.cpp: one wave, every output entry = C + sum_k a_k b_k).  The kernels' stage
order rests on these facts (the BF16 stages keep their small and dominant products in separate
accumulators, the FP32 stages use two accumulator chains), so a change in them shows here first.

  v_mfma_f32_16x16x4f32: every product rounded to FP32, then added in sequence, round to nearest even.
  v_mfma_f32_16x16x32_bf16: the 32 products of an output in four groups of 8 (the K values one 16-lane
    group of A holds); within a group every product truncated toward zero 24 bits below the group's
    largest exponent sum e(a) + e(b); C truncated toward zero 24 bits below the largest exponent sum of
    all 32 products; the sum rounded once, to nearest even.
"""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

PROBE = os.path.join(os.path.dirname(__file__), "gpu_probe", "mfma_round")
U = 2.0 ** -23
E24 = 2.0 ** -24

# case: (FP32 result, BF16 result, what it shows).  FP32 sees only K 0..3 (one per 16-lane group).
EXPECT = {
    0: (1 + U, 1.0, "1 + 0.75 ulp: FP32 rounds to nearest; BF16 truncates the small product to 2^-24, "
                    "and 1 + 2^-24 ties to even"),
    1: (1 + U, 1 + U, "C = 1 + 0.75 ulp: C above the products keeps its bits, the sum rounds up"),
    2: (1.0, 1.0, "1 + 3 x 0.375 ulp: FP32 adds in sequence (exact sum 1.125 ulp would round up)"),
    3: (0.0, 0.0, "(1 + u)^2 - (1 + 2u): FP32 rounds the product before the sum (exact: 2^-46)"),
    4: (2.0 ** -20, 0.0, "2^10 - 2^10 + 2^-20: the BF16 group drops bits 24 below its largest product"),
    5: (1 - U, 1 - E24, "1 - 0.75 ulp: BF16 truncates toward zero (-1.5 x 2^-24 -> -2^-24)"),
    6: (1.5 * E24, 4.5 * E24, "C = -1, 1 + 16 x 2^-25 at K 1..16: the small products in K 8..15 and 16 "
                              "form groups of their own and survive; those beside the 1 do not"),
    7: (1.0, 0.0, "C = 0.75 x 2^-24, products 1 and -1: C truncated 24 bits below the products' largest"),
    8: (1.5 * E24, 1.5 * E24, "C = -0.5, 1 x 0.5 + 1.5 x 2^-24: window from the exponent sum (-1)"),
    9: (E24, E24, "C = -2.25, 1.5 x 1.5 + 2^-24: window from the exponent sum 0, not from 2.25"),
    10: (0.0, E24, "C = -1, 1 at K 0 and 1.5 x 2^-24 at K 4: one group, truncated"),
    11: (0.0, 1.5 * E24, "C = -1, 1 at K 0 and 1.5 x 2^-24 at K 8: two groups, exact"),
}


def test_mfma_rounding_model():
    if not os.path.exists(PROBE):
        pytest.fail(f"{PROBE} is not built (run __graft_entry__.build())")
    res = subprocess.run([PROBE], capture_output=True, text=True, timeout=60)
    assert res.returncode == 0, res.stderr
    got = {}
    for line in res.stdout.split("\n"):
        parts = line.split()
        if not parts:
            continue
        assert parts[1] != "DIFFER", line  # every output entry of an instruction equal
        got[int(parts[0])] = (float.fromhex(parts[1]), float.fromhex(parts[2]))
    assert sorted(got) == sorted(EXPECT)
    for case, (f32, bf16, what) in EXPECT.items():
        assert got[case] == (f32, bf16), (case, what, got[case])
