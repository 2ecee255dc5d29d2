"""Known-answer pins of the oracle's queries, decay bookkeeping, merge_into
and wire format (oracle/kd_oracle.c).

Every expected number below is worked out by hand from tiny digests whose
state is set directly (so no clustering is involved), and each case is placed
so that one plausible misreading of the reference fails it: a walk that
starts at 0 instead of w0/2, a fraction over the right centroid's weight
instead of the pair's half-sum, a tail that divides by w0 instead of w0/2, a
missing singleton snap, a missing rescale of the stored total, an un-rebased
or clock-dropping merge_into, a swapped header field, a strict prune floor.

The branches follow tdigest.mojo:485-538 (quantile), 540-587 (cdf),
324-377 (clock, renormalize, effective count), 441-458 (prune),
589-623 (merge_into) and 627-698 (serde v1).
"""
import math
import struct

import numpy as np
import pytest

TOL = 1e-12


def _digest(oracle, means, weights, vmin, vmax, stored_total=None, log_scale=0.0, clock=0.0,
            decay=0.0, scale=0, compression=1000.0, bmeans=(), bweights=()):
    d = oracle.Digest(scale, compression, decay)
    if stored_total is None:
        stored_total = float(sum(weights)) + float(sum(bweights))
    d.set_state(means, weights, list(bmeans), list(bweights), stored_total, log_scale, clock,
                vmin, vmax)
    return d


# Digest A: means 1, 2, 4 with weights 2, 2, 4 (n = 8), min 0, max 6.
#   half-weights 1, 1, 2; interior pair half-sums dw = 2 (1|2) and 3 (2|4);
#   the walk starts at w0/2 = 1 and the right tail begins at n - wL/2 = 6.
QUANTILE_A = [
    (0.0, 0.0),                # q <= 0 -> min
    (1.0, 6.0),                # q >= 1 -> max
    (0.05, 0.4),               # t = 0.4 < 1: 0 + (1 - 0) * (0.4 / 1)
    (0.1, 0.8),                # t = 0.8 < 1: 0.8
    (0.125, 1.0),              # t = 1 (not < 1): interior, frac 0 -> 1
    (0.25, 1.5),               # t = 2: (2 - 1) / 2 = 0.5 -> 1.5
    (0.375, 2.0),              # t = 3: frac 1 on the first pair -> 2
    (0.5, 2.0 + 2.0 / 3.0),    # t = 4: second pair, (4 - 3) / 3
    (0.75, 4.0),               # t = 6 = n - wL/2: interior, frac 1 -> 4
    (0.9, 5.2),                # t = 7.2 > 6: 4 + (6 - 4) * (1.2 / 2)
]

CDF_A = [
    (-1.0, 0.0),               # below min
    (7.0, 1.0),                # above max
    (0.0, 0.0),                # at min: (w0/2) * 0 / n
    (0.5, 0.0625),             # (2/2) * (0.5 / 1) / 8
    (1.0, 0.125),              # at m0: first pair, frac 0 -> (1 + 0) / 8
    (1.5, 0.25),               # (1 + 2 * 0.5) / 8
    (2.0, 0.375),              # (1 + 2 * 1) / 8
    (3.0, 0.5625),             # wsf 3, dw 3, frac 0.5: (3 + 1.5) / 8
    (4.0, 0.75),               # (3 + 3) / 8
    (5.0, 0.875),              # (8 - 2 + 2 * 0.5) / 8
    (6.0, 1.0),                # at max: (8 - 2 + 2) / 8
]


@pytest.mark.parametrize("q,want", QUANTILE_A)
def test_quantile_known_answers(oracle, q, want):
    d = _digest(oracle, [1.0, 2.0, 4.0], [2.0, 2.0, 4.0], 0.0, 6.0)
    assert abs(d.quantile(q) - want) <= TOL, (q, d.quantile(q), want)


@pytest.mark.parametrize("x,want", CDF_A)
def test_cdf_known_answers(oracle, x, want):
    d = _digest(oracle, [1.0, 2.0, 4.0], [2.0, 2.0, 4.0], 0.0, 6.0)
    assert abs(d.cdf(x) - want) <= TOL, (x, d.cdf(x), want)


def test_quantile_singleton_tails_snap(oracle):
    # Digest B: means 1, 2, 4, weights 1, 2, 1 (n = 4), min 0.5, max 5.  Both
    # tail centroids are singletons, so targets below w0/2 = 0.5 return min and
    # targets above n - wL/2 = 3.5 return max (interpolating would give 0.9 and
    # 4.6).
    d = _digest(oracle, [1.0, 2.0, 4.0], [1.0, 2.0, 1.0], 0.5, 5.0)
    assert d.quantile(0.1) == 0.5
    assert d.quantile(0.95) == 5.0
    # Interior between them: t = 2, wsf 0.5, dw 1.5 -> 1 + (2 - 1) * (1.5 / 1.5).
    assert abs(d.quantile(0.5) - 2.0) <= TOL
    # t = 1: frac (1 - 0.5) / 1.5 = 1/3 -> 1 + 1/3.
    assert abs(d.quantile(0.25) - (1.0 + 1.0 / 3.0)) <= TOL


def test_quantile_singleton_uses_effective_weight(oracle):
    # With decay, a stored weight of 2 at log_scale ln 2 is an effective single
    # sample (tdigest.mojo:475-483): the left tail snaps to min.  A stored-weight
    # test would interpolate to 0.5 + 0.5 * (0.8 / 1) = 0.9 instead.
    d = _digest(oracle, [1.0, 2.0, 4.0], [2.0, 4.0, 2.0], 0.5, 5.0, log_scale=math.log(2.0),
                decay=0.1)
    assert d.quantile(0.1) == 0.5
    assert d.quantile(0.95) == 5.0


def test_single_centroid_queries(oracle):
    # One centroid answers its mean for every q, q = 0 and 1 included (the
    # count test comes before the q <= 0 / q >= 1 rules); cdf gives 0 / 0.5 / 1
    # around the mean and 0 / 1 outside [min, max].
    d = _digest(oracle, [3.0], [5.0], 1.0, 9.0)
    for q in (0.0, 0.3, 1.0):
        assert d.quantile(q) == 3.0
    assert [d.cdf(x) for x in (0.5, 2.0, 3.0, 4.0, 9.5)] == [0.0, 0.0, 0.5, 1.0, 1.0]


def test_cdf_equal_means_pair(oracle):
    # Means 2, 2, 4 (weights 2, 2, 4, min 0, max 6): x = 2 reaches the first
    # pair with equal means, which answers (w0/2 + dw/2) / n = (1 + 1) / 8.
    d = _digest(oracle, [2.0, 2.0, 4.0], [2.0, 2.0, 4.0], 0.0, 6.0)
    assert d.cdf(2.0) == 0.25


def test_quantile_falls_through_to_last_mean(oracle):
    # A stored total above the centroid sum (a state only a wire blob can
    # carry) puts t = 90 past the whole interior walk (which ends at 6) but not
    # in the right tail (98): the loop falls through to the last mean.
    d = _digest(oracle, [1.0, 2.0, 4.0], [2.0, 2.0, 4.0], 0.0, 6.0, stored_total=100.0)
    assert d.quantile(0.9) == 4.0


# ------------------------------------------------------------ decay
def test_renormalize_closed_form_add_loop(oracle):
    # lambda = 0.01, 50,000 unit adds of one value: the exponent passes 400 at
    # the 40,000th tick and is rebased (tdigest.mojo:324-332, 359-371).  With
    # a single centroid nothing is pruned, so the effective total is the
    # geometric sum sum_{j<n} e^{-lambda j} = (1 - e^{-lambda n}) / (1 - e^{-lambda}).
    lam, n = 0.01, 50000
    d = oracle.Digest(1, 100.0, lam)
    d.add_loop(np.zeros(n))
    st = d.state()
    assert st["log_scale"] < 400.0
    assert abs(st["log_scale"] - lam * (n - 40000)) < 0.011
    want = (1.0 - math.exp(-lam * n)) / (1.0 - math.exp(-lam))
    assert abs(d.total_weight() - want) <= 1e-12 * want
    m, w = d.centroids()
    assert m.size == 1
    # The centroid weight was rebased too: stored weight equals stored total.
    assert abs(w[0] - st["stored_total"]) <= 1e-12 * st["stored_total"]


def test_renormalize_closed_form_batches(oracle):
    # The same with B-semantics batches of 1000 (one tick of 1000 per batch,
    # tdigest.mojo:258-282): 60 batches pass 400 at batch 41; the effective
    # total is 1000 * (1 - e^{-10 * 60}) / (1 - e^{-10}).
    lam, s, m = 0.01, 1000, 60
    d = oracle.Digest(1, 100.0, lam)
    for _ in range(m):
        d.add_batch_b(np.zeros(s))
    st = d.state()
    assert st["log_scale"] < 400.0 and st["clock"] == s * m
    want = s * (1.0 - math.exp(-lam * s * m)) / (1.0 - math.exp(-lam * s))
    assert abs(d.total_weight() - want) <= 1e-12 * want


def test_prune_floor_is_inclusive(oracle):
    # tdigest.mojo:441-458: keep a centroid whose stored weight is >= the floor
    # 1e-6 * e^{log_scale} (= 1e-6 at log_scale 0).  K0 with compression 1000
    # (normalizer 500) keeps the four records apart.
    d = _digest(oracle, [0.0, 1.0, 2.0], [1.0, 1e-6, 2.0], 0.0, 3.0, decay=0.5,
                bmeans=[3.0], bweights=[1.0])
    d.merge()
    m, w = d.centroids()
    assert m.tolist() == [0.0, 1.0, 2.0, 3.0] and w.tolist() == [1.0, 1e-6, 2.0, 1.0]
    assert d.state()["stored_total"] == 4.0 + 1e-6
    # Just below the floor the centroid goes, and its weight leaves the total.
    e = _digest(oracle, [0.0, 1.0, 2.0], [1.0, 0.5e-6, 2.0], 0.0, 3.0, decay=0.5,
                bmeans=[3.0], bweights=[1.0])
    st0 = e.state()["stored_total"]
    e.merge()
    m, w = e.centroids()
    assert m.tolist() == [0.0, 2.0, 3.0]
    assert e.state()["stored_total"] == st0 - 0.5e-6


def test_merge_into_decayed_rebases_both(oracle):
    # self: one centroid (0, stored 4) at log_scale ln 2 (effective 2), clock 10;
    # other: one centroid (10, stored 6) at log_scale ln 3 (effective 2), clock 20.
    # Both are rebased to log_scale 0 before folding (tdigest.mojo:605-609),
    # the clock becomes the later one (:622), min/max combine (:616-621).
    lam = 0.1
    a = _digest(oracle, [0.0], [4.0], -1.0, 1.0, log_scale=math.log(2.0), clock=10.0, decay=lam)
    b = _digest(oracle, [10.0], [6.0], 9.0, 11.0, log_scale=math.log(3.0), clock=20.0, decay=lam)
    a.merge_into(b)
    m, w = a.centroids()
    st = a.state()
    assert m.tolist() == [0.0, 10.0]
    np.testing.assert_allclose(w, [2.0, 2.0], rtol=TOL)
    assert abs(st["stored_total"] - 4.0) <= 4 * TOL
    assert st["log_scale"] == 0.0 and st["clock"] == 20.0
    assert (st["min"], st["max"]) == (-1.0, 11.0)
    # `other` is mutated: rebased in place.
    sb = b.state()
    assert sb["log_scale"] == 0.0 and abs(sb["stored_total"] - 2.0) <= 2 * TOL
    assert abs(b.centroids()[1][0] - 2.0) <= 2 * TOL
    # The later clock wins whichever side holds it.
    c = _digest(oracle, [0.0], [1.0], 0.0, 0.0, clock=30.0, decay=lam)
    e = _digest(oracle, [5.0], [1.0], 5.0, 5.0, clock=7.0, decay=lam)
    c.merge_into(e)
    assert c.state()["clock"] == 30.0


def test_merge_into_empty_self_takes_other_range(oracle):
    # A fresh digest has min = max = 0; folding into it copies the other's
    # range instead of combining with those zeros (tdigest.mojo:611-618).
    a = oracle.Digest(0, 1000.0)
    b = _digest(oracle, [10.0, 12.0], [1.0, 1.0], 9.0, 13.0, clock=2.0)
    a.merge_into(b)
    st = a.state()
    assert (st["min"], st["max"]) == (9.0, 13.0) and st["clock"] == 2.0
    assert a.centroids()[0].tolist() == [10.0, 12.0] and st["stored_total"] == 2.0


def test_merge_into_empty_other_is_noop(oracle):
    a = _digest(oracle, [1.0], [3.0], 1.0, 1.0, log_scale=2.0, clock=5.0, decay=0.1)
    before = a.to_bytes()
    a.merge_into(oracle.Digest(0, 1000.0, 0.1))
    assert a.to_bytes() == before  # log_scale 2 kept: no rebase without mass


# ------------------------------------------------------------ serde v1
def test_to_bytes_layout(oracle):
    # tdigest.mojo:634-651: u8 version 1, then little-endian f64 compression,
    # lambda, log_scale, clock, stored_total, min, max, count, then (mean,
    # weight) pairs.  Every header field holds a distinct value.
    d = _digest(oracle, [1.0, 5.0], [3.0, 7.0], -2.0, 9.0, log_scale=1.5, clock=77.0,
                decay=0.25, compression=123.0)
    b = d.to_bytes()
    want = bytes([1]) + struct.pack("<8d", 123.0, 0.25, 1.5, 77.0, 10.0, -2.0, 9.0, 2.0) \
        + struct.pack("<4d", 1.0, 3.0, 5.0, 7.0)
    assert b == want


def test_from_bytes_fields(oracle):
    # A hand-packed blob (with 5 trailing bytes, which are ignored) restores
    # every field into its own slot (tdigest.mojo:661-698).
    blob = bytes([1]) + struct.pack("<8d", 64.0, 0.5, 3.25, 41.0, 12.5, -7.0, 8.0, 2.0) \
        + struct.pack("<4d", -1.0, 4.5, 2.0, 8.0) + b"\x00junk"
    r = oracle.Digest.from_bytes(blob, 2)
    st = r.state()
    assert (st["compression"], st["decay"], st["log_scale"], st["clock"], st["stored_total"],
            st["min"], st["max"]) == (64.0, 0.5, 3.25, 41.0, 12.5, -7.0, 8.0)
    m, w = r.centroids()
    assert m.tolist() == [-1.0, 2.0] and w.tolist() == [4.5, 8.0]
    assert r.to_bytes() == blob[:-5]
    for bad in (b"", bytes([2]) + blob[1:], blob[:64], blob[:65 + 16 + 8]):
        with pytest.raises(oracle.OracleError):
            oracle.Digest.from_bytes(bad, 2)


def test_from_centroids_state(oracle):
    # tdigest.mojo:700-726: no decay, stored total = clock = sum of weights.
    r = oracle.Digest.from_centroids(1, 100.0, [1.0, 2.0, 3.0], [2.0, 0.5, 1.5], -1.0, 4.0)
    st = r.state()
    assert (st["stored_total"], st["clock"], st["log_scale"], st["decay"]) == (4.0, 4.0, 0.0, 0.0)
    assert (st["min"], st["max"], st["compression"]) == (-1.0, 4.0, 100.0)
    assert r.centroids()[1].tolist() == [2.0, 0.5, 1.5]


def test_queries_use_stored_total_under_decay(oracle):
    # Stored weights 2, 4, 2 (total 8) at log_scale ln 2: the target and the cdf
    # denominator are the stored total (tdigest.mojo:502, 508, 558), in which
    # the decay factor cancels.  q = 0.5 -> t = 4 = wsf 1 + dw 3 -> frac 1 -> 2;
    # cdf(2) = (1 + 3 * 1) / 8.
    d = _digest(oracle, [1.0, 2.0, 4.0], [2.0, 4.0, 2.0], 0.5, 5.0, log_scale=math.log(2.0),
                decay=0.1)
    assert abs(d.quantile(0.5) - 2.0) <= TOL
    assert d.cdf(2.0) == 0.5


def test_prune_floor_scales_with_exponent(oracle):
    # At log_scale 3 the stored floor is 1e-6 * e^3 (about 2.0e-5): a stored
    # weight of 1e-5 is below it and goes; 1e-4 stays (heavy neighbours keep
    # every record its own centroid).
    d = _digest(oracle, [0.0, 1.0, 2.0, 3.0], [1.0, 1e-5, 2.0, 1e-4], 0.0, 4.0, log_scale=3.0,
                decay=0.5, bmeans=[4.0], bweights=[1.0])
    d.merge()
    assert d.centroids()[0].tolist() == [0.0, 2.0, 3.0, 4.0]


def test_normalizer_uses_effective_count(oracle):
    # tdigest.mojo:410-412: the normalizer sees the effective total, the q
    # ratios the stored one.  The same records stored at 256x weight under
    # log_scale 8 ln 2 (effective weight 1 each) must cluster exactly as the
    # unscaled ones: K2's normalizer depends on n, so a stored-total
    # normalizer would differ by 4 * ln 256 in its denominator.
    rng = np.random.default_rng(21)
    x = np.sort(rng.standard_normal(3000))
    a = _digest(oracle, [], [], x[0], x[-1], decay=0.01, scale=2, compression=100.0,
                bmeans=x, bweights=np.ones_like(x))
    b = _digest(oracle, [], [], x[0], x[-1], decay=0.01, scale=2, compression=100.0,
                log_scale=8 * math.log(2.0), bmeans=x, bweights=np.full_like(x, 256.0))
    a.merge()
    b.merge()
    ma, wa = a.centroids()
    mb, wb = b.centroids()
    assert ma.size == mb.size and ma.size > 10
    assert np.array_equal(wa * 256.0, wb) and np.array_equal(ma, mb)


def test_add_at_earlier_timestamp_keeps_clock(oracle):
    # tdigest.mojo:172-175: an earlier timestamp is accepted without moving the
    # clock; the sample gets the current scale factor (effective weight 1).
    d = oracle.Digest(1, 100.0, 0.05)
    d.add_at(1.0, 10.0)
    d.add_at(2.0, 4.0)
    st = d.state()
    assert st["clock"] == 10.0 and abs(st["log_scale"] - 0.5) <= 1e-15
    assert abs(d.total_weight() - 2.0) <= TOL


def test_buffer_limit_is_int_compression_times_5(oracle):
    # tdigest.mojo:146, 349: compression 100.7 -> limit 500; the 500th add
    # merges the buffer.
    d = oracle.Digest(1, 100.7)
    d.add_loop(np.arange(499.0))
    assert d.nbuffer() == 499 and len(d) == 0
    d.add(499.0)
    assert d.nbuffer() == 0 and len(d) > 0


@pytest.mark.parametrize("session,cross", [(True, 40000), (False, 60000)])
def test_add_batch_crossover(oracle, session, cross):
    # tdigest.mojo:55-58, 199, 217: below the crossover add_batch is the add
    # loop (clock one tick per sample, buffer kept); at it, one B batch.
    x = np.random.default_rng(22).standard_normal(cross)
    lo, lo_ref = oracle.Digest(1, 100.0), oracle.Digest(1, 100.0)
    lo.add_batch(x[:-1], with_session=session)
    lo_ref.add_loop(x[:-1])
    assert lo.to_bytes() == lo_ref.to_bytes()
    hi, hi_ref = oracle.Digest(1, 100.0), oracle.Digest(1, 100.0)
    hi.add_batch(x, with_session=session)
    hi_ref.add_batch_b(x)
    assert hi.to_bytes() == hi_ref.to_bytes() and hi.nbuffer() == 0


def test_vector_helpers_match_scalar_calls(oracle):
    # The vector query / add_at helpers are loops over the scalar entry points.
    rng = np.random.default_rng(30)
    d = oracle.Digest(1, 100.0)
    d.add_loop(rng.standard_normal(5000))
    qs = rng.uniform(0, 1, 50)
    assert d.quantiles(qs).tolist() == [d.quantile(float(q)) for q in qs]
    xs = rng.normal(0, 2, 50)
    assert d.cdfs(xs).tolist() == [d.cdf(float(x)) for x in xs]
    a, b = oracle.Digest(2, 300.0, 1e-3), oracle.Digest(2, 300.0, 1e-3)
    v, t, w = rng.standard_t(3, 4000), np.cumsum(rng.exponential(1e-3, 4000)), rng.uniform(0.1, 10, 4000)
    a.add_at_many(v, t, w)
    for i in range(v.size):
        b.add_at(float(v[i]), float(t[i]), float(w[i]))
    assert a.to_bytes() == b.to_bytes()


def test_threaded_replays_match_sequential(oracle):
    # The threaded helpers give each digest its samples / requests in order,
    # whatever the thread count.
    rng = np.random.default_rng(31)
    nd = 9
    sizes = rng.integers(0, 3000, nd)
    off = np.concatenate([[0], np.cumsum(sizes)])
    v = rng.standard_t(3, off[-1])
    t = np.concatenate([np.cumsum(rng.exponential(1e-3, s)) for s in sizes])
    w = rng.uniform(0.1, 10, off[-1])
    a = [oracle.Digest(2, 300.0, 1e-3) for _ in range(nd)]
    b = [oracle.Digest(2, 300.0, 1e-3) for _ in range(nd)]
    oracle.batch_add_at(a, off, v, t, w, threads=4)
    for d in range(nd):
        b[d].add_at_many(v[off[d]:off[d + 1]], t[off[d]:off[d + 1]], w[off[d]:off[d + 1]])
    assert all(x.to_bytes() == y.to_bytes() for x, y in zip(a, b))
    keys = rng.integers(0, 5, 40)
    roff = np.concatenate([[0], np.cumsum(rng.integers(1, 500, 40))])
    data = rng.standard_normal(roff[-1])
    c = [oracle.Digest(1, 100.0, 1e-4) for _ in range(5)]
    e = [oracle.Digest(1, 100.0, 1e-4) for _ in range(5)]
    oracle.replay_ingests(c, keys, roff, data, threads=3)
    for r in range(40):
        e[keys[r]].add_batch_b(data[roff[r]:roff[r + 1]])
    assert all(x.to_bytes() == y.to_bytes() for x, y in zip(c, e))
