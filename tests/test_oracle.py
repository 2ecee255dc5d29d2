"""CPU tests that pin the oracle (oracle/kd_oracle.c).

Each check is a known answer, closed form or invariant taken from the
reference's own tests (cited file:line), or a hand-computed tiny case, chosen so
that a plausible slip in the restatement (an off-by-one in q_right, a dropped
weight, a wrong tie rule, a transposed operand) fails.  The reference's CPU
tests draw from Mojo's std.random, which cannot be regenerated here, so the
inputs are numpy draws with the same distributions and the same bounds.
"""
import math

import numpy as np
import pytest


def _sorted_cdf(sorted_data, v):
    return np.searchsorted(sorted_data, v, side="right") / sorted_data.size


# ------------------------------------------------ centroid (centroid.mojo)
def test_centroid_two_unit_samples(oracle):
    # test_centroid.mojo:14-18: (5, 7) -> mean 6, weight 2.  With K0 and a
    # tiny normalizer every record joins one cluster.
    m, w = oracle.sweep(0, 1e-9, 2.0, [5.0, 7.0], [1.0, 1.0])
    assert m.tolist() == [6.0] and w.tolist() == [2.0]


def test_centroid_weighted_mean_closed_form(oracle):
    # test_centroid.mojo:21-37 (values pre-sorted for the sweep).
    vals = np.array([1.0, 2.0, 10.0, 100.0])
    wts = np.array([3.0, 4.0, 1.0, 0.5])
    m, w = oracle.sweep(0, 1e-9, wts.sum(), vals, wts)
    assert abs(m[0] - (vals * wts).sum() / wts.sum()) <= 1e-12
    assert abs(w[0] - wts.sum()) <= 1e-12


def test_centroid_merge_7_6(oracle):
    # test_centroid.mojo:40-46: (4, 2) merged with (10, 3) -> 7.6.
    m, w = oracle.sweep(0, 1e-9, 5.0, [4.0, 10.0], [2.0, 3.0])
    assert abs(m[0] - 7.6) <= 1e-12 and w[0] == 5.0


# ------------------------------------------------ sweep (tdigest.mojo:391-439)
def test_sweep_k0_hand_computed(oracle):
    # K0, delta = 4 -> nrm 2, 10 unit samples 1..10: k_span = 2*(P_r-P_l)/10,
    # so a cluster takes candidates while P_r - P_l <= 5: {1..5}, {6..10}.
    nrm = oracle.normalizer(0, 4.0, 10.0)
    assert nrm == 2.0
    m, w = oracle.sweep(0, nrm, 10.0, np.arange(1.0, 11.0), np.ones(10))
    assert m.tolist() == [3.0, 8.0] and w.tolist() == [5.0, 5.0]


def test_sweep_k0_boundary_off_by_one(oracle):
    # 11 samples, delta = 4: P_r - P_l <= 5.5 -> clusters of 5, 5, 1.
    m, w = oracle.sweep(0, 2.0, 11.0, np.arange(1.0, 12.0), np.ones(11))
    assert w.tolist() == [5.0, 5.0, 1.0]
    assert m.tolist() == [3.0, 8.0, 11.0]


def test_sweep_tie_rule_absorbs_equal_means(oracle):
    # tdigest.mojo:429: candidate.mean == current.mean always merges, even when
    # the k budget is exhausted (K0 with delta = 2: every cluster would be a
    # singleton otherwise).
    x = np.array([1.0, 1.0, 1.0, 2.0, 3.0, 3.0])
    m, w = oracle.sweep(0, 1.0, 6.0, x, np.ones(6))
    # Budget: k_span = (P_r - P_l)/6 <= 1 always holds for K0 with nrm 1 ...
    assert w.sum() == 6.0
    m2, w2 = oracle.sweep(0, 0.1, 6.0, x, np.ones(6))
    # ... while nrm 0.1 leaves k_span = 0.1*(P_r-P_l)/6 <= 1 true too; use K2's
    # steep tails instead: K2 nrm 10 splits everything except exact ties.
    m3, w3 = oracle.sweep(2, 10.0, 6.0, x, np.ones(6))
    assert m3.tolist() == [1.0, 2.0, 3.0] and w3.tolist() == [3.0, 1.0, 2.0]


# ------------------------------------------------ scale (scale.mojo)
@pytest.mark.parametrize("scale", [0, 1, 2, 3])
def test_scale_round_trip(kd, oracle, scale):
    # test_scale.mojo:29-36 (compression 200, n 100000, atol 1e-6).
    nrm = oracle.normalizer(scale, 200.0, 100000.0)
    for q in (0.001, 0.01, 0.25, 0.5, 0.75, 0.99, 0.999):
        back = kd.load().kd_scale_q(scale, oracle.k(scale, q, nrm), nrm)
        assert abs(back - q) <= 1e-6


@pytest.mark.parametrize("scale", [0, 1, 2, 3])
def test_scale_monotone(oracle, scale):
    # test_scale.mojo:39-48.
    nrm = oracle.normalizer(scale, 200.0, 100000.0)
    prev = oracle.k(scale, 1e-6, nrm)
    for i in range(1, 200):
        cur = oracle.k(scale, i / 200.0, nrm)
        assert cur > prev
        prev = cur


@pytest.mark.parametrize("scale", [1, 2, 3])
def test_scale_tail_steepness(oracle, scale):
    # test_scale.mojo:51-63.
    nrm = oracle.normalizer(scale, 200.0, 100000.0)
    h = 1e-4
    d = lambda q: (oracle.k(scale, q + h, nrm) - oracle.k(scale, q - h, nrm)) / (2 * h)
    assert d(0.01) > d(0.5) and d(0.99) > d(0.5)


def test_scale_k0_linear(oracle):
    # test_scale.mojo:119-126.
    nrm = oracle.normalizer(0, 200.0, 100000.0)
    h = 1e-4
    d = lambda q: (oracle.k(0, q + h, nrm) - oracle.k(0, q - h, nrm)) / (2 * h)
    assert abs(d(0.01) - d(0.5)) <= 1e-9


def test_scale_closed_forms(oracle):
    # scale.mojo:86-184 evaluated by hand.
    assert oracle.normalizer(0, 100.0, 5.0) == 50.0
    assert oracle.normalizer(1, 100.0, 5.0) == 100.0 / (2 * math.pi)
    n = 1e6
    assert abs(oracle.normalizer(2, 100.0, n) - 100.0 / (4 * math.log(n / 100.0) + 24.0)) <= 1e-15
    assert abs(oracle.normalizer(3, 100.0, n) - 100.0 / (4 * math.log(n / 100.0) + 21.0)) <= 1e-15
    assert oracle.k(1, 0.5, 3.0) == 0.0
    assert abs(oracle.k(1, 1.0, 3.0) - 3.0 * math.pi / 2) <= 1e-15
    # K2/K3 clamp q to [1e-15, 1 - 1e-15] (scale.mojo:32-47).
    assert oracle.k(2, 0.0, 1.0) == oracle.k(2, 1e-15, 1.0)
    assert oracle.k(3, 1.0, 1.0) == oracle.k(3, 1.0 - 1e-15, 1.0)
    assert abs(oracle.k(3, 0.25, 2.0) - 2.0 * math.log(0.5)) <= 1e-15
    assert abs(oracle.k(3, 0.75, 2.0) + 2.0 * math.log(0.5)) <= 1e-15


def test_kd_math_within_one_ulp_of_glibc(oracle):
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.uniform(1e-300, 1, 30000), np.exp(rng.uniform(-700, 700, 30000))])
    mine = np.array([oracle.kd_log(float(x)) for x in xs])
    ulp = np.abs(mine.view(np.int64) - np.log(xs).view(np.int64))
    assert ulp.max() <= 1
    ys = rng.uniform(-1, 1, 60000)
    mine = np.array([oracle.kd_asin(float(y)) for y in ys])
    ulp = np.abs(mine.view(np.int64) - np.arcsin(ys).view(np.int64))
    assert ulp.max() <= 1
    assert oracle.kd_log(1.0) == 0.0 and oracle.kd_asin(0.0) == 0.0
    assert oracle.kd_asin(1.0) == math.pi / 2 and oracle.kd_asin(-1.0) == -math.pi / 2


# ------------------------------------------------ decay (decay.mojo, test_decay.mojo)
def test_decay_constructors(kd):
    assert not kd.Decay.none().is_active()
    assert kd.Decay.rate(0.5).rate_lambda == 0.5
    assert abs(kd.Decay.half_life(10.0).rate_lambda - math.log(2.0) / 10.0) <= 1e-12
    for h in (0.0, -1.0):
        with pytest.raises(ValueError):
            kd.Decay.half_life(h)


def test_decay_steady_state(oracle):
    # test_decay.mojo:108-134: effective total -> 1/(1-e^-lambda) within 5 %.
    h = 1000
    lam = math.log(2.0) / h
    d = oracle.Digest(1, 100.0, lam)
    d.add_loop(np.random.default_rng(1).standard_normal(12 * h))
    steady = 1.0 / (1.0 - math.exp(-lam))
    assert abs(d.total_weight() - steady) / steady < 0.05


def test_decay_seed_contribution_halves(oracle):
    # test_decay.mojo:137-160.
    lam = math.log(2.0) / 2000.0
    a, b = oracle.Digest(1, 100.0, lam), oracle.Digest(1, 100.0, lam)
    a.add_at(1000.0, 0.0)
    a.add_at(0.0, 2000.0)
    b.add_at(0.0, 2000.0)
    assert abs((a.total_weight() - b.total_weight()) - 0.5) <= 1e-6


def test_decay_add_at_total(oracle):
    # test_decay.mojo:203-211.
    d = oracle.Digest(1, 100.0, math.log(2.0) / 100.0)
    d.add_at(0.0, 0.0)
    d.add_at(0.0, 100.0)
    assert abs(d.total_weight() - 1.5) <= 1e-6


def test_decay_memory_bound_and_drift(oracle):
    # test_decay.mojo:166-178 (<= 6 delta centroids over 300k samples) and
    # 69-102 (drift).
    rng = np.random.default_rng(3)
    d = oracle.Digest(1, 100.0, math.log(2.0) / 5000.0)
    d.add_loop(rng.standard_normal(300000))
    d.merge()
    assert len(d) <= 600
    dec = oracle.Digest(1, 100.0, math.log(2.0) / 2000.0)
    plain = oracle.Digest(1, 100.0)
    x1, x2 = rng.standard_normal(20000), rng.standard_normal(20000) + 10
    for x in (x1, x2):
        dec.add_loop(x)
        plain.add_loop(x)
    assert abs(dec.quantile(0.5) - 10.0) < 1.0
    assert abs(plain.quantile(0.5) - 10.0) > 2.0


def test_zero_decay_equals_no_decay(oracle):
    # test_decay.mojo:184-197 (here: identical, not just within 1e-12).
    x = np.random.default_rng(4).uniform(size=20000)
    a, b = oracle.Digest(1, 100.0, 0.0), oracle.Digest(1, 100.0)
    a.add_loop(x)
    b.add_loop(x)
    for q in (0.001, 0.01, 0.5, 0.99, 0.999):
        assert a.quantile(q) == b.quantile(q)


# ------------------------------------------------ accuracy (test_tdigest.mojo)
def _gen(kind, n, rng):
    u = rng.uniform(size=n)
    if kind == "uniform":
        return u
    if kind == "normal":
        u1 = np.maximum(u, 1e-12)
        return np.sqrt(-2 * np.log(u1)) * np.cos(2 * np.pi * rng.uniform(size=n))
    if kind == "exponential":
        return -np.log(1 - np.minimum(u, 1 - 1e-12))
    return (1 - np.minimum(u, 1 - 1e-12)) ** (-1 / 1.5)


@pytest.mark.parametrize("kind,tail", [("uniform", 0.002), ("normal", 0.002),
                                       ("exponential", 0.002), ("pareto", 0.003)])
def test_rank_error(oracle, kind, tail):
    # test_tdigest.mojo:109-163: K1, delta 100, 50k samples.
    x = _gen(kind, 50000, np.random.default_rng(hash(kind) & 0xffff))
    s = np.sort(x)
    d = oracle.Digest(1, 100.0)
    d.add_loop(x)
    for q, tol in zip((0.001, 0.01, 0.5, 0.99, 0.999), (tail, tail, 0.01, tail, tail)):
        assert abs(_sorted_cdf(s, d.quantile(q)) - q) <= tol


def test_beats_histogram(oracle):
    # test_tdigest.mojo:169-192.
    x = np.random.default_rng(5).uniform(size=50000)
    d = oracle.Digest(1, 100.0)
    d.add_loop(x)
    true = np.sort(x)[int(0.999 * x.size)]
    assert abs(d.quantile(0.999) - true) * 10 < 0.01


def test_merge_equivalence(oracle):
    # test_tdigest.mojo:198-225.
    rng = np.random.default_rng(6)
    a, b = _gen("normal", 25000, rng), _gen("normal", 25000, rng)
    comb, da, db = oracle.Digest(2, 100.0), oracle.Digest(2, 100.0), oracle.Digest(2, 100.0)
    comb.add_loop(a)
    comb.add_loop(b)
    da.add_loop(a)
    db.add_loop(b)
    da.merge_into(db)
    for q in (0.01, 0.1, 0.5, 0.9, 0.99):
        assert abs(comb.quantile(q) - da.quantile(q)) <= 0.05
    assert abs(da.total_weight() - 50000) <= 1e-9


def test_compression_bound(oracle):
    # test_tdigest.mojo:231-243.
    d = oracle.Digest(1, 100.0)
    d.add_loop(np.random.default_rng(7).uniform(size=50000))
    d.merge()
    assert 10 < len(d) <= 600


def test_interpolation_edges(oracle):
    # test_tdigest.mojo:249-301.
    d = oracle.Digest(1, 100.0)
    d.add(42.0)
    assert d.quantile(0.0) == d.quantile(0.5) == d.quantile(1.0) == 42.0
    rng = np.random.default_rng(8)
    e = oracle.Digest(1, 100.0)
    e.add_loop(rng.uniform(size=1000))
    assert e.quantile(0.0) == e.state()["min"] and e.quantile(1.0) == e.state()["max"]
    for q in (-0.1, 1.1):
        with pytest.raises(oracle.OracleError):
            e.quantile(q)
    f = oracle.Digest(1, 100.0)
    f.add_loop(rng.uniform(size=5000))
    assert f.cdf(f.state()["min"] - 1) == 0.0 and f.cdf(f.state()["max"] + 1) == 1.0
    prev = f.cdf(0.0)
    for i in range(1, 50):
        cur = f.cdf(i / 50)
        assert cur >= prev - 1e-9
        prev = cur
    g = oracle.Digest(1, 100.0)
    g.add_loop(rng.uniform(size=50000))
    for q in (0.1, 0.25, 0.5, 0.75, 0.9):
        assert abs(g.cdf(g.quantile(q)) - q) <= 0.02


# ------------------------------------------------ edge cases (test_edge.mojo)
def test_edge_cases(oracle):
    d = oracle.Digest(1, 100.0)
    for call in (lambda: d.quantile(0.5), lambda: d.cdf(0.0)):
        with pytest.raises(oracle.OracleError):
            call()
    assert d.total_weight() == 0.0 and len(d) == 0
    d.add(7.0)
    assert (d.cdf(6.0), d.cdf(8.0), d.cdf(7.0)) == (0.0, 1.0, 0.5)
    same = oracle.Digest(1, 100.0)
    same.add_loop(np.full(10000, 3.0))
    assert same.quantile(0.999) == 3.0 and len(same) == 1
    bad = oracle.Digest(1, 100.0)
    bad.add(5.0)
    for v, w in ((math.nan, 1.0), (math.inf, 1.0), (-math.inf, 1.0), (1.0, 0.0), (1.0, -2.0)):
        with pytest.raises(oracle.OracleError):
            bad.add(v, w)
    assert bad.total_weight() == 1.0


def test_serde(oracle):
    # test_edge.mojo:139-180.
    rng = np.random.default_rng(1)
    d = oracle.Digest(2, 100.0, math.log(2.0) / 5000.0)
    d.add_loop(rng.uniform(size=20000))
    b = d.to_bytes()
    assert len(b) == 65 + 16 * len(d) and b[0] == 1
    r = oracle.Digest.from_bytes(b, 2)
    assert len(r) == len(d) and r.to_bytes() == b
    for q in (0.001, 0.01, 0.5, 0.99, 0.999):
        assert r.quantile(q) == d.quantile(q)
    with pytest.raises(oracle.OracleError):
        oracle.Digest.from_bytes(bytes([99]))
    with pytest.raises(oracle.OracleError):
        oracle.Digest.from_bytes(b[:-16])


def test_cross_process_merge(oracle):
    # test_edge.mojo:183-215.
    x = np.random.default_rng(3).uniform(size=20000)
    a, b = oracle.Digest(1, 100.0), oracle.Digest(1, 100.0)
    a.add_loop(x[0::2])
    b.add_loop(x[1::2])
    ra = oracle.Digest.from_bytes(a.to_bytes(), 1)
    rb = oracle.Digest.from_bytes(b.to_bytes(), 1)
    ra.merge_into(rb)
    c = oracle.Digest(1, 100.0)
    c.add_loop(x)
    for q in (0.01, 0.1, 0.5, 0.9, 0.99):
        assert abs(ra.quantile(q) - c.quantile(q)) <= 0.03
    assert abs(ra.total_weight() - 20000) <= 1e-6


# ------------------------------------------------ semantic pins
def test_add_batch_below_crossover_is_add_loop(oracle):
    # test_gpu_add_batch.mojo:369-386.
    x = np.random.default_rng(9).uniform(size=1000)
    a, b = oracle.Digest(1, 100.0), oracle.Digest(1, 100.0)
    a.add_loop(x)
    b.add_batch(x)
    a.merge()
    b.merge()
    assert a.centroids()[0].tolist() == b.centroids()[0].tolist()


def test_b_semantics_state(oracle):
    # tdigest.mojo:236-315: one clock tick of n, one scale factor, min/max of
    # the batch on a fresh digest, stored total = sum of compressed weights.
    x = np.random.default_rng(10).lognormal(0, 1.5, 70000)
    d = oracle.Digest(2, 200.0, math.log(2.0) / 1e5)
    d.add_batch_b(x)
    st = d.state()
    assert st["clock"] == 70000.0
    assert abs(st["log_scale"] - 70000.0 * math.log(2.0) / 1e5) <= 1e-15
    assert st["min"] == x.min() and st["max"] == x.max()
    m, w = d.centroids()
    assert st["stored_total"] == sum(w.tolist())
    sf = math.exp(st["log_scale"])
    assert abs(st["stored_total"] - 70000 * sf) <= 1e-9 * 70000 * sf


def test_grouped_is_per_key_b(oracle):
    # test_gpu_grouped.mojo:241-311: grouped equals per-key B batches.
    rng = np.random.default_rng(11)
    keys = rng.integers(0, 5, 30000).astype(np.uint32)
    vals = rng.standard_normal(keys.size)
    g = [oracle.Digest(2, 100.0) for _ in range(5)]
    oracle.grouped_add_batch(g, keys, vals, threads=3)
    for k in range(5):
        r = oracle.Digest(2, 100.0)
        r.add_batch_b(vals[keys == k])
        assert g[k].centroids()[1].tolist() == r.centroids()[1].tolist()
        assert g[k].centroids()[0].tolist() == r.centroids()[0].tolist()


def test_glibc_k_sensitivity(oracle):
    # The oracle's k() uses kd_math.h log/asin (bit-identical to the device).
    # With glibc's log/asin instead, the same workloads give the same
    # clusterings: the ulp-level choice of libm does not move any boundary here.
    rng = np.random.default_rng(12)
    for scale in range(4):
        x = rng.standard_normal(100000)
        a = oracle.Digest(scale, 100.0, variant="kd")
        b = oracle.Digest(scale, 100.0, variant="glibc")
        a.add_batch_b(x)
        b.add_batch_b(x)
        assert len(a) == len(b)
        assert np.array_equal(a.centroids()[1], b.centroids()[1])
