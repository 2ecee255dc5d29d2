"""Serving loop: every request's result is the same whether it runs alone or
inside a continuous batch, and equals the oracle replaying the key's requests
in arrival order (GroupedTDigest semantics, grouped.mojo:60-190: one batch
merge per ingest, tdigest.mojo:223-315; queries tdigest.mojo:485-587)."""
import numpy as np
import pytest

from tests.parity import QUERY_RTOL, check_centroids

pytestmark = pytest.mark.gpu

INGEST, QUANTILE, CDF = 0, 1, 2


def _workload(seed, nkeys, nreq, max_ingest):
    rng = np.random.default_rng(seed)
    reqs = []
    for _ in range(nreq):
        key = int(rng.zipf(1.4) - 1) % nkeys
        kind = int(rng.choice([INGEST, INGEST, QUANTILE, CDF]))
        if kind == INGEST:
            n = int(np.exp(rng.uniform(0.0, np.log(max_ingest))))
            data = key + rng.standard_t(3, n)
            if rng.uniform() < 0.1:
                data = np.round(data)  # ties
        elif kind == QUANTILE:
            data = rng.uniform(0.0, 1.0, int(rng.integers(1, 24)))
            data[0] = rng.choice([0.0, 1.0, 1e-6, 0.5])
        else:
            data = key + rng.normal(0.0, 3.0, int(rng.integers(1, 24)))
        reqs.append((kind, key, data))
    return reqs


def _take(kd, srv, i):
    try:
        return srv.result(i)
    except kd.KdigestError as e:
        assert e.code == -2, e  # KD_EEMPTY: query on an empty key
        return "empty"


def _run_mixed(kd, reqs, nkeys, scale, delta, decay, budget, rounds):
    srv = kd.DigestServer(nkeys, delta, decay, scale=scale, slot_capacity=2048, mem_budget=budget)
    ids = [srv.submit(k, key, d) for k, key, d in reqs]
    steps = 0
    while srv.pending:
        assert srv.step(rounds) > 0
        steps += 1
    out = [_take(kd, srv, i) for i in ids]
    assert all(o is not None for o in out)
    return srv, out, steps


def _run_alone(kd, reqs, nkeys, scale, delta, decay):
    srv = kd.DigestServer(nkeys, delta, decay, scale=scale, slot_capacity=2048, mem_budget=1)
    out = []
    for k, key, d in reqs:
        i = srv.submit(k, key, d)
        assert srv.step(1) == 1
        out.append(_take(kd, srv, i))
    return srv, out


def _same_bits(a, b):
    if isinstance(a, str) or isinstance(b, str):
        return a == b
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


CASES = [
    # scale, delta, decay lambda (5e-3 crosses the renormalize threshold)
    (1, 100.0, 0.0),
    (2, 200.0, 1e-5),
    (3, 50.0, 5e-3),
]


@pytest.mark.parametrize("scale,delta,lam", CASES)
def test_serve_mixed_equals_alone_and_oracle(kd, oracle, scale, delta, lam):
    nkeys = 37
    reqs = _workload(5 + scale, nkeys, 400, 30000)
    decay = kd.Decay.rate(lam)
    sc = [kd.ScaleK0, kd.ScaleK1, kd.ScaleK2, kd.ScaleK3][scale]
    mixed, got, steps = _run_mixed(kd, reqs, nkeys, sc, delta, decay, 2 << 20, 3)
    assert steps > 1  # the budget forces several continuous batches
    alone, ref_alone = _run_alone(kd, reqs, nkeys, sc, delta, decay)
    for i, (a, b) in enumerate(zip(got, ref_alone)):
        assert _same_bits(a, b), (i, reqs[i][0], reqs[i][1], a, b)

    digests = [oracle.Digest(scale, delta, lam) for _ in range(nkeys)]
    xmax = np.zeros(nkeys)
    for i, (kind, key, data) in enumerate(reqs):
        d = digests[key]
        if kind == INGEST:
            d.add_batch_b(data)
            xmax[key] = max(xmax[key], np.abs(data).max())
            continue
        if len(d) == 0:
            assert got[i] == "empty", (i, key)
            continue
        fn = d.quantile if kind == QUANTILE else d.cdf
        exp = np.array([fn(float(x)) for x in data])
        g = got[i]
        ok = np.abs(g - exp) <= QUERY_RTOL * np.maximum(1.0, np.abs(exp))
        assert ok.all(), (i, kind, key, data[~ok], g[~ok], exp[~ok])
    for key in range(nkeys):
        st, rs = mixed.key_state(key), digests[key].state()
        m, w = mixed.centroids(key)
        rm, rw = digests[key].centroids()
        check_centroids(m, w, rm, rw, xmax[key], lam == 0.0, f"key {key}")
        am, aw = alone.centroids(key)
        assert _same_bits(m, am) and _same_bits(w, aw), key
        for k in ("clock", "log_scale", "min", "max"):
            assert st[k] == rs[k], (key, k, st[k], rs[k])
        assert abs(st["stored_total"] - rs["stored_total"]) <= 1e-12 * max(1.0, abs(rs["stored_total"]))


def test_serve_admission_order_and_budget(kd):
    # Device memory of a batch: an ingest of n values 8 n + 24 (n + slot
    # capacity) bytes, a query 28 bytes per value.
    cap = 128
    need_a = 8 * 1000 + 24 * (1000 + cap)
    srv = kd.DigestServer(4, 100.0, slot_capacity=cap, mem_budget=need_a + 10)
    a = srv.submit_ingest(0, np.arange(1000.0))   # fills the budget by itself
    q = srv.submit_quantile(0, [0.5])             # waits: budget full
    b = srv.submit_ingest(1, np.arange(10.0))     # FIFO: waits behind q
    assert srv.step() == 1
    assert srv.result(a) is not None and srv.result(q) is None and srv.result(b) is None
    assert srv.pending == 2
    assert srv.step() == 2
    assert abs(srv.result(q)[0] - 499.5) < 1.0
    assert srv.result(b) is not None
    # One oversized request is still admitted on its own.
    big = srv.submit_ingest(2, np.ones(50000))
    assert srv.step() == 1 and srv.result(big) is not None
    assert srv.key_state(2)["stored_total"] == 50000.0


def test_serve_rejects_bad_requests(kd):
    srv = kd.DigestServer(3, 100.0)
    with pytest.raises(kd.KdigestError):
        srv.submit_ingest(3, [1.0])
    with pytest.raises(kd.KdigestError):
        srv.submit_ingest(0, [1.0, np.nan])
    with pytest.raises(kd.KdigestError):
        srv.submit_quantile(0, [1.5])
    with pytest.raises(kd.KdigestError):
        srv.submit(7, 0, [1.0])
    assert srv.pending == 0
    r = srv.submit_quantile(1, [0.5])
    assert srv.step() == 1
    with pytest.raises(kd.KdigestError):  # empty key
        srv.result(r)
    assert srv.pending == 0


def test_serve_submit_many_and_completions(kd):
    rng = np.random.default_rng(3)
    kinds = [0, 0, 1, 2, 0, 1, 2]
    keys = [1, 2, 1, 2, 1, 7, 1]
    datas = [rng.standard_normal(1000), rng.standard_normal(50), np.array([0.5, 0.9]),
             np.array([0.0, 1.0]), rng.standard_normal(3000) + 1.0, np.array([0.5]),
             np.array([-0.5, 0.5, 2.0])]
    off = np.cumsum([0] + [len(d) for d in datas])
    srv = kd.DigestServer(8, 100.0)
    with pytest.raises(kd.KdigestError):  # one bad request: nothing is queued
        srv.submit_many([0, 1], [0, 0], [0, 1, 2], [1.0, 2.0])
    assert srv.pending == 0
    ids = srv.submit_many(kinds, keys, off, np.concatenate(datas))
    while srv.pending:
        srv.step(4)
    got, t = srv.take_completed()
    assert sorted(got.tolist()) == sorted(ids.tolist())
    assert np.all(np.diff(t) >= 0) and t[0] >= 0.0 and t[-1] <= srv.clock()
    assert srv.take_completed()[0].size == 0
    one = kd.DigestServer(8, 100.0)
    for rid, k, key, d in zip(ids.tolist(), kinds, keys, datas):
        r1 = one.submit(k, key, d)
        assert one.step(1) == 1
        one.take_completed()
        if key == 7:
            for s_, r_ in ((srv, rid), (one, r1)):
                with pytest.raises(kd.KdigestError):
                    s_.result(r_)
            continue
        a, b = srv.result(rid), one.result(r1)
        assert a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64)), (k, key)


def test_serve_slot_overflow_fails_only_that_request(kd):
    srv = kd.DigestServer(4, 100.0, scale=kd.ScaleK1, slot_capacity=8)
    rng = np.random.default_rng(4)
    small = srv.submit_ingest(1, [1.0, 2.0, 3.0])
    big = srv.submit_ingest(0, rng.standard_normal(20000))  # about 50 centroids
    while srv.pending:
        srv.step(4)
    assert srv.result(small) is not None
    with pytest.raises(kd.KdigestError) as e:
        srv.result(big)
    assert e.value.code == -7  # KD_ENOMEM
    st = srv.key_state(0)
    assert st["centroids"] == 0 and st["stored_total"] == 0.0 and st["clock"] == 0.0
    assert srv.key_state(1)["stored_total"] == 3.0
    q = srv.submit_quantile(1, [0.5])
    srv.step(1)
    assert srv.result(q)[0] == 2.0


def test_serve_renormalize_then_slot_overflow(kd, oracle):
    """A request whose prepare renormalizes its key (the decay exponent passes
    400, tdigest.mojo:324-332) and whose merge then overflows the key's slot
    fails alone; the key is left exactly as before the request (weights,
    stored total, exponent, clock), so the next queries agree with the oracle
    digest that never saw the failed request."""
    lam = 0.01
    srv = kd.DigestServer(2, 100.0, kd.Decay.rate(lam), scale=kd.ScaleK1, slot_capacity=40,
                          mem_budget=64 << 20)
    rng = np.random.default_rng(12)
    first = rng.standard_normal(30)
    r = oracle.Digest(1, 100.0, lam)
    r.add_batch_b(first)
    a = srv.submit_ingest(0, first)
    srv.step()
    assert srv.result(a).size == 0
    m0, w0 = srv.centroids(0)
    st0 = srv.key_state(0)
    b = srv.submit_ingest(0, rng.standard_normal(50000))  # +500 to the exponent, > 40 centroids
    srv.step()
    with pytest.raises(kd.KdigestError) as e:
        srv.result(b)
    assert e.value.code == -7
    st, rs = srv.key_state(0), r.state()
    assert st == st0
    for k in ("log_scale", "clock", "min", "max"):
        assert st[k] == rs[k], (k, st[k], rs[k])
    assert abs(st["stored_total"] - rs["stored_total"]) <= 1e-12 * rs["stored_total"]
    m, w = srv.centroids(0)
    assert np.array_equal(m, m0) and np.array_equal(w, w0)
    rm, rw = r.centroids()
    check_centroids(m, w, rm, rw, 5.0, False, "kept")
    qs = np.array([0.0, 0.01, 0.1, 0.5, 0.9, 0.99, 1.0])
    xs = np.concatenate([first[:5], [-5.0, 0.0, 5.0]])
    iq = srv.submit_quantile(0, qs)
    ic = srv.submit_cdf(0, xs)
    srv.step()
    gq, gc = srv.result(iq), srv.result(ic)
    eq = r.quantiles(qs)
    ec = r.cdfs(xs)
    assert np.all(np.abs(gq - eq) <= QUERY_RTOL * np.maximum(1.0, np.abs(eq))), (gq, eq)
    assert np.all(np.abs(gc - ec) <= QUERY_RTOL), (gc, ec)
    # A later request that renormalizes (+450) and fits goes through as on the
    # CPU: values on a 0.5 grid (about 17 distinct) tie into few centroids.
    more = np.round(2.0 * rng.standard_normal(45000)) / 2.0
    r2 = oracle.Digest(1, 100.0, lam)
    r2.add_batch_b(first)
    r2.add_batch_b(more)
    assert len(r2) <= 40
    c = srv.submit_ingest(0, more)
    srv.step()
    assert srv.result(c).size == 0
    st, rs = srv.key_state(0), r2.state()
    for k in ("log_scale", "clock", "min", "max"):
        assert st[k] == rs[k], (k, st[k], rs[k])
    m, w = srv.centroids(0)
    rm, rw = r2.centroids()
    check_centroids(m, w, rm, rw, 5.0, False, "renormalized")


def _run_workers(kd, reqs, nkeys, scale, delta, decay, budget, rounds, seed):
    """Submits while the server's workers run: bursts of requests with short
    pauses between them, so later requests arrive while earlier ones are on
    the device; polls completions meanwhile."""
    import time

    srv = kd.DigestServer(nkeys, delta, decay, scale=scale, slot_capacity=2048, mem_budget=budget)
    rng = np.random.default_rng(seed)
    head = int(rng.integers(0, 8))  # a few queued before the start join the workers' queues
    ids = [srv.submit(k, key, d) for k, key, d in reqs[:head]]
    srv.start(rounds)
    seen = 0
    i = head
    while i < len(reqs):
        burst = int(rng.integers(1, 12))
        for k, key, d in reqs[i:i + burst]:
            ids.append(srv.submit(k, key, d))
        i += burst
        seen += srv.take_completed()[0].size
        time.sleep(float(rng.uniform(0.0, 0.002)))
    srv.stop()
    seen += srv.take_completed()[0].size
    assert seen == len(reqs) and srv.pending == 0
    out = [_take(kd, srv, j) for j in ids]
    assert all(o is not None for o in out)
    return srv, out


@pytest.mark.parametrize("scale,delta,lam", CASES)
def test_serve_workers_equal_alone(kd, scale, delta, lam):
    """Worker mode (queries on their own stream beside the ingest rounds,
    requests admitted as they arrive) gives every request the result it gets
    alone, bit for bit, and leaves every key in the same state."""
    nkeys = 37
    reqs = _workload(5 + scale, nkeys, 400, 30000)
    decay = kd.Decay.rate(lam)
    sc = [kd.ScaleK0, kd.ScaleK1, kd.ScaleK2, kd.ScaleK3][scale]
    alone, ref = _run_alone(kd, reqs, nkeys, sc, delta, decay)
    srv, got = _run_workers(kd, reqs, nkeys, sc, delta, decay, 2 << 20, 3, 40 + scale)
    for i, (a, b) in enumerate(zip(got, ref)):
        assert _same_bits(a, b), (i, reqs[i][0], reqs[i][1], a, b)
    st = srv.stats()
    assert st["admitted"] == len(reqs) and st["queued"] == 0
    for key in range(nkeys):
        assert srv.key_state(key) == alone.key_state(key), key
        m, w = srv.centroids(key)
        am, aw = alone.centroids(key)
        assert _same_bits(m, am) and _same_bits(w, aw), key


def test_serve_workers_guards_and_restart(kd):
    srv = kd.DigestServer(4, 100.0, scale=kd.ScaleK1, slot_capacity=8)
    rng = np.random.default_rng(9)
    srv.start(2)
    srv.start(2)  # already running: no effect
    with pytest.raises(kd.KdigestError):
        srv.step()
    with pytest.raises(kd.KdigestError):
        srv.key_state(0)
    with pytest.raises(kd.KdigestError):
        srv.centroids(0)
    e = srv.submit_quantile(2, [0.5])        # empty key
    small = srv.submit_ingest(1, [1.0, 2.0, 3.0])
    big = srv.submit_ingest(0, rng.standard_normal(20000))  # overflows its 8-centroid slot
    q = srv.submit_quantile(1, [0.5])
    srv.stop()
    srv.stop()  # not running: no effect
    with pytest.raises(kd.KdigestError) as ex:
        srv.result(e)
    assert ex.value.code == -2
    with pytest.raises(kd.KdigestError) as ex:
        srv.result(big)
    assert ex.value.code == -7
    assert srv.result(small).size == 0 and srv.result(q)[0] == 2.0
    assert srv.key_state(0)["stored_total"] == 0.0
    # Queued before a start, served after it; then the synchronous step again.
    a = srv.submit_ingest(2, [5.0, 6.0])
    srv.start(1)
    b = srv.submit_quantile(2, [0.0, 1.0])
    srv.stop()
    assert srv.result(a).size == 0 and srv.result(b).tolist() == [5.0, 6.0]
    c = srv.submit_cdf(2, [5.5])
    assert srv.step() == 1 and srv.result(c)[0] == 0.5
    del srv  # destroy while stopped
    other = kd.DigestServer(2, 100.0)
    other.start(1)
    other.submit_ingest(0, np.arange(100.0))
    del other  # destroy while running: drains and joins the workers
