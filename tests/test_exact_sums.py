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


def _seq_loop(S, x, stop_on, thr):
    """The reference's per-sample loop, one double addition per step."""
    out = np.empty(x.size)
    for i, v in enumerate(x.tolist()):
        ns = S + v
        if stop_on and v != 0.0 and ns > thr:
            return i, S, out[:i], ns
        out[i] = ns
        S = ns
    return x.size, S, out, None


def _seq_cases():
    rng = np.random.default_rng(1)
    cases = []
    for k in range(120):
        m = int(rng.integers(1, 3000))
        kind = k % 6
        if kind == 0:      # decay-exponent-like increments from zero, with zero runs
            x = rng.exponential(1e-6, m) * (rng.uniform(size=m) < 0.8)
            S = 0.0
        elif kind == 1:    # scaled weights onto a large total
            x = rng.uniform(0.1, 10.0, m) * math.exp(rng.uniform(0, 300))
            S = float(x.sum() * rng.uniform(0, 50))
        elif kind == 2:    # binary fractions: many exact halves
            x = rng.integers(1, 64, m) * 2.0 ** int(rng.integers(-60, -40))
            S = float(rng.integers(1, 1 << 20)) * 2.0 ** -40
        elif kind == 3:    # mixed magnitudes, some beyond 2^52 units
            x = np.exp(rng.uniform(-40, 5, m))
            S = float(np.exp(rng.uniform(-10, 10)))
        elif kind == 4:    # sums that land exactly on powers of two
            x = np.full(m, 0.25)
            S = 1.0
        else:              # threshold crossing (the renormalization cut)
            x = rng.exponential(0.5, m) * (rng.uniform(size=m) < 0.9)
            S = 400.0 - float(rng.uniform(0, 0.5 * m * 0.5))
        cases.append((max(S, 0.0), x.astype(np.float64), kind == 5 or k % 7 == 0, 400.0))
    cases.append((0.0, np.zeros(50), True, 400.0))
    cases.append((2.0 ** -1070, np.full(40, 2.0 ** -1074), False, 0.0))
    cases.append((399.9, np.array([0.0, 0.05, 0.0, 0.05, 0.05, 1.0]), True, 400.0))
    return cases


@pytest.mark.parametrize("case", range(len(_seq_cases())))
def test_add_sequence_equals_loop(kd, case):
    """kd_add_sequence (the batched replay's parallel running sums, by binade
    rounds) against the plain loop: every partial sum, the consumed count and
    the sum that stops it, bit for bit."""
    import ctypes as C

    S0, x, stop_on, thr = _seq_cases()[case]
    n_ref, s_ref, out_ref, stop_ref = _seq_loop(S0, x, stop_on, thr)
    S = C.c_double(S0)
    stop = C.c_double(0.0)
    out = np.full(x.size, np.nan)
    dp = C.POINTER(C.c_double)
    n = kd.load().kd_add_sequence(C.byref(S), x.ctypes.data_as(dp), x.size, int(stop_on), thr,
                                  out.ctypes.data_as(dp), C.byref(stop))
    assert n == n_ref, (case, n, n_ref)
    assert S.value == s_ref, (case, S.value, s_ref)
    assert np.array_equal(out[:n].view(np.uint64), out_ref.view(np.uint64)), case
    if stop_ref is not None:
        assert stop.value == stop_ref, (case, stop.value, stop_ref)
