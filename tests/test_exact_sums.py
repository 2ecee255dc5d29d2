"""The running-sum helper the device sweeps use to check their boundaries
(kd_add_repeated: L sequential W = W + x in O(log) steps) against the plain
loop of L additions, bit for bit, on the cases its binade stepping has to get
right: remainders below, above and exactly at half an ulp (ties to even),
binade crossings, increments that vanish, and long runs."""
import math

import numpy as np
import pytest


def _cases():
    rng = np.random.default_rng(0)
    cases = []
    for _ in range(300):
        x = math.exp(rng.uniform(-20, 20))
        W = x * math.exp(rng.uniform(-3, 12)) if rng.uniform() < 0.8 else x
        L = int(np.exp(rng.uniform(0, math.log(200000))))
        cases.append((W, x, L))
    # Exact-half remainders: x with its lowest set bit at half the ulp of W.
    for k in range(40):
        x = 1.0 + 2.0 ** -52 * (2 * k + 1)       # odd last bit
        for W in (2.0 ** 1, 2.0 ** 1 + 2.0 ** -51, 3.0, 2.0 ** 2 + 2.0 ** -50 * 3):
            cases.append((W, x, 5000 + 17 * k))
    cases += [(1.0, 1e-17, 10 ** 6), (1e300, 1.0, 1000), (0.0, 0.1, 100000), (5.0, 5.0, 3),
              (1.5, 0.75, 99999), (0.1, 0.1, 123457), (2.0 ** 53 - 2, 1.0, 10)]
    return cases


@pytest.mark.parametrize("case", range(len(_cases())))
def test_add_repeated_equals_loop(kd, oracle, case):
    W, x, L = _cases()[case]
    got = kd.load().kd_add_repeated(W, x, L)
    ref = oracle.add_loop_sum(W, x, L)
    assert got == ref or (math.isnan(got) and math.isnan(ref)), (W, x, L, got, ref)
