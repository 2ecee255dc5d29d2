"""GPU parity of the compress path against the CPU oracle.

Inputs are seeded numpy draws shared by both sides; the oracle runs the
reference's CPU MergingDigest semantics (tdigest.mojo:391-439 sweep, B
semantics of tdigest.mojo:223-315).
"""
import numpy as np
import pytest

from tests.parity import check_centroids, check_values, stable_sorted

pytestmark = pytest.mark.gpu


def _rng(seed):
    return np.random.default_rng(seed)


def test_device_math_bit_exact(kd, oracle):
    rng = _rng(1)
    xs = np.concatenate([rng.uniform(1e-300, 1, 200000), np.exp(rng.uniform(-700, 700, 200000)),
                         rng.uniform(0.5, 2.0, 100000), [1.0, 2.0, 0.5, 1e-15, 1 - 1e-15, 5e-324]])
    got = kd.gpu_math_eval("log", xs)
    ref = np.array([oracle.kd_log(float(x)) for x in xs])
    assert np.array_equal(got.view(np.int64), ref.view(np.int64))
    ys = np.concatenate([rng.uniform(-1, 1, 200000), [-1.0, 1.0, 0.0, 0.5, -0.5, 0.975, -0.975]])
    got = kd.gpu_math_eval("asin", ys)
    ref = np.array([oracle.kd_asin(float(y)) for y in ys])
    assert np.array_equal(got.view(np.int64), ref.view(np.int64))
    zs = np.concatenate([rng.uniform(0, 1, 200000), np.exp(rng.uniform(-700, 700, 100000))])
    got = kd.gpu_math_eval("sqrt", zs)
    assert np.array_equal(got.view(np.int64), np.sqrt(zs).view(np.int64))


@pytest.mark.parametrize("scale", [0, 1, 2, 3])
def test_scale_eval_bit_exact(kd, oracle, scale):
    rng = _rng(2 + scale)
    qs = np.concatenate([rng.uniform(0, 1, 100000), [0.0, 1.0, 1e-16, 1 - 1e-16, 0.5, 1e-15]])
    nrm = oracle.normalizer(scale, 100.0, 123456.0)
    got = kd.gpu_scale_eval(scale, nrm, qs)
    ref = np.array([oracle.k(scale, float(q), nrm) for q in qs])
    assert np.array_equal(got.view(np.int64), ref.view(np.int64))


def _oracle_compress(oracle, scale, prior_m, prior_w, new, w_uniform, nrm, stored_n):
    m = np.concatenate([prior_m, new])
    w = np.concatenate([prior_w, np.full(new.size, w_uniform)])
    sm, sw = stable_sorted(m, w)
    return oracle.sweep(scale, nrm, stored_n, sm, sw)


CASES = [
    # (n_new, n_prior, levels) — levels > 0 quantizes values (heavy ties)
    (1, 0, 0), (2, 0, 0), (3, 0, 0), (100, 0, 0), (4095, 0, 0), (4096, 0, 0), (4097, 0, 0),
    (10000, 0, 0), (100000, 0, 0), (300000, 0, 0), (1 << 20, 0, 0),
    (5000, 300, 0), (200000, 700, 0), (100, 2000, 0),
    (20000, 0, 37), (300000, 0, 37), (50000, 200, 20), (1000, 0, 1), (70001, 0, 2),
]


@pytest.mark.parametrize("scale", [0, 1, 2, 3])
@pytest.mark.parametrize("case", CASES, ids=[f"n{c[0]}_p{c[1]}_l{c[2]}" for c in CASES])
def test_compress_matches_oracle(kd, oracle, scale, case):
    n, n_prior, levels = case
    rng = _rng(hash((n, n_prior, levels, scale)) & 0xffffffff)
    new = rng.standard_normal(n) * 3.0 + 1.0
    if levels:
        new = np.floor(rng.uniform(0, levels, n))
    prior_m = np.sort(rng.standard_normal(n_prior) * 3.0 + 1.0)
    if levels:
        prior_m = np.sort(np.floor(rng.uniform(0, levels, n_prior)))
    prior_w = rng.integers(1, 6, n_prior).astype(np.float64)
    stored_n = float(prior_w.sum() + n)
    nrm = oracle.normalizer(scale, 100.0, stored_n)
    got_m, got_w = kd.gpu_compress(np.concatenate([prior_m, new]), prior_w, n_prior, 1.0, scale,
                                   nrm, stored_n)
    ref_m, ref_w = _oracle_compress(oracle, scale, prior_m, prior_w, new, 1.0, nrm, stored_n)
    scale_abs = float(np.abs(np.concatenate([prior_m, new])).max())
    check_centroids(got_m, got_w, ref_m, ref_w, scale_abs, True, tag=f"K{scale} {case}")


def test_compress_fractional_weights(kd, oracle):
    rng = _rng(11)
    for scale in range(4):
        n, n_prior = 50000, 500
        new = rng.standard_normal(n)
        prior_m = np.sort(rng.standard_normal(n_prior))
        prior_w = rng.uniform(0.1, 10, n_prior)
        wu = 1.7310234
        stored_n = float(prior_w.sum() + n * wu)
        nrm = oracle.normalizer(scale, 200.0, stored_n)
        got_m, got_w = kd.gpu_compress(np.concatenate([prior_m, new]), prior_w, n_prior, wu, scale,
                                       nrm, stored_n)
        ref_m, ref_w = _oracle_compress(oracle, scale, prior_m, prior_w, new, wu, nrm, stored_n)
        check_centroids(got_m, got_w, ref_m, ref_w, 5.0, False, tag=f"frac K{scale}")
