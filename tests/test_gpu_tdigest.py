"""GPU parity of TDigest / GroupedTDigest / queries against the CPU oracle."""
import numpy as np
import pytest

from tests.parity import check_centroids, check_values

pytestmark = pytest.mark.gpu


def _cmp_digest(d, r, tag, integer_weights=True):
    m, w = d.centroids()
    rm, rw = r.centroids()
    scale_abs = max(abs(r.state()["min"]), abs(r.state()["max"]), 1e-300)
    check_centroids(m, w, rm, rw, scale_abs, integer_weights, tag)
    st, rs = d.state(), r.state()
    for k in ("log_scale", "clock", "min", "max"):
        assert st[k] == rs[k], (tag, k, st[k], rs[k])
    if integer_weights:
        assert st["stored_total"] == rs["stored_total"], (tag, st["stored_total"], rs["stored_total"])
    else:
        assert abs(st["stored_total"] - rs["stored_total"]) <= 1e-12 * abs(rs["stored_total"]), tag


@pytest.mark.parametrize("scale", [0, 1, 2, 3])
def test_add_batch_b_semantics(kd, oracle, scale):
    rng = np.random.default_rng(100 + scale)
    d = kd.TDigest(scale, 200.0)
    r = oracle.Digest(scale, 200.0)
    for b in range(4):
        x = rng.lognormal(0.0, 1.5, 70000 + 1000 * b)
        d.add_batch(x)
        r.add_batch(x)
        _cmp_digest(d, r, f"K{scale} batch {b}")
    qs = np.concatenate([rng.uniform(0, 1, 500), [0.0, 1.0, 1e-4, 1 - 1e-4]])
    got = d.quantile(qs)
    ref = np.array([r.quantile(float(q)) for q in qs])
    check_values(got, ref, 1.0, "quantile")
    xs = rng.lognormal(0.0, 1.5, 500)
    got = d.cdf(xs)
    ref = np.array([r.cdf(float(x)) for x in xs])
    check_values(got, ref, 1.0, "cdf")


@pytest.mark.parametrize("scale", [1, 2])
def test_add_batch_b_with_decay(kd, oracle, scale):
    rng = np.random.default_rng(200 + scale)
    lam = np.log(2.0) / 1e5
    d = kd.TDigest(scale, 200.0, kd.Decay.rate(lam))
    r = oracle.Digest(scale, 200.0, lam)
    for b in range(5):
        x = rng.lognormal(0.0, 1.5, 65536)
        d.add_batch(x)
        r.add_batch(x)
        _cmp_digest(d, r, f"decay K{scale} batch {b}", integer_weights=False)
    qs = rng.uniform(0, 1, 300)
    check_values(d.quantile(qs), np.array([r.quantile(float(q)) for q in qs]), 1.0, "quantile")


@pytest.mark.parametrize("scale", [1, 3])
def test_decay_repeated_values_keep_tie_merges(kd, oracle, scale):
    """Long clusters of one repeated value under fractional (decayed) weights
    keep that value as their exact mean, so a later batch holding the value
    ties with the stored centroid (tdigest.mojo:429) as on the CPU."""
    rng = np.random.default_rng(300 + scale)
    lam = 5e-3
    d = kd.TDigest(scale, 50.0, kd.Decay.rate(lam))
    r = oracle.Digest(scale, 50.0, lam)
    for b in range(6):
        x = np.round(rng.standard_t(3, 70000 + 997 * b))
        d.add_batch(x)
        r.add_batch(x)
        _cmp_digest(d, r, f"ties K{scale} batch {b}", integer_weights=False)


@pytest.mark.parametrize("scale", [0, 1, 2, 3])
def test_add_loop_matches_oracle(kd, oracle, scale):
    # Below the crossover add_batch is the per-sample path (tdigest.mojo:199-202);
    # its merges every 5*delta samples run on the GPU.
    rng = np.random.default_rng(300 + scale)
    d = kd.TDigest(scale, 100.0)
    r = oracle.Digest(scale, 100.0)
    x = rng.standard_normal(20000)
    for c in range(0, x.size, 4096):
        d.add_batch(x[c:c + 4096])
        r.add_batch(x[c:c + 4096])
    d.merge()
    r.merge()
    _cmp_digest(d, r, f"A K{scale}")
    for q in (0.001, 0.01, 0.5, 0.99, 0.999):
        check_values([d.quantile(q)], [r.quantile(q)], 1.0, f"q={q}")
    for x0 in (-3.0, 0.0, 3.0):
        check_values([d.cdf(x0)], [r.cdf(x0)], 1.0, f"cdf({x0})")


def test_add_at_decay_matches_oracle(kd, oracle):
    rng = np.random.default_rng(5)
    lam = 1e-3
    d = kd.TDigest(2, 300.0, kd.Decay.rate(lam))
    r = oracle.Digest(2, 300.0, lam)
    t = 0.0
    for i in range(6000):
        t += rng.exponential(1e-3)
        v = rng.standard_t(3)
        w = rng.uniform(0.1, 10)
        d.add_at(v, t, w)
        r.add_at(v, t, w)
    d.merge()
    r.merge()
    _cmp_digest(d, r, "add_at", integer_weights=False)


def test_merge_into_matches_oracle(kd, oracle):
    rng = np.random.default_rng(6)
    a, b = kd.TDigest(1, 100.0), kd.TDigest(1, 100.0)
    ra, rb = oracle.Digest(1, 100.0), oracle.Digest(1, 100.0)
    xa, xb = rng.standard_normal(70000), rng.standard_normal(3000) + 2
    a.add_batch(xa)
    ra.add_batch(xa)
    for v in xb:
        b.add(v)
        rb.add(v)
    a.merge_into(b)
    ra.merge_into(rb)
    _cmp_digest(a, ra, "merge_into")
    _cmp_digest(b, rb, "merge_into other")


def test_serde_round_trip_bytes_identical(kd, oracle):
    rng = np.random.default_rng(8)
    d = kd.TDigest(2, 100.0, kd.Decay.half_life(5000.0))
    r = oracle.Digest(2, 100.0, np.log(2.0) / 5000.0)
    x = rng.uniform(0, 1, 20000)
    for v in x:
        d.add(v)
        r.add(v)
    bd, br = d.to_bytes(), r.to_bytes()
    assert bd == br
    d2 = kd.TDigest.from_bytes(bd, kd.ScaleK2)
    assert d2.to_bytes() == bd
    assert len(d2) == len(d)


@pytest.mark.parametrize("decay", [0.0, np.log(2.0) / 500.0])
def test_grouped_accumulates_like_per_key_b(kd, oracle, decay):
    # test_gpu_grouped.mojo:241-311: per key equal to single-digest B batches.
    rng = np.random.default_rng(9)
    nk = 6
    g = kd.GroupedTDigest(nk, 100.0, kd.Decay.rate(decay), scale=kd.ScaleK2)
    refs = [oracle.Digest(2, 100.0, decay) for _ in range(nk)]
    for batch in range(5):
        keys, vals = [], []
        for i in range(400):
            for k in range(nk):
                if k == 1 and batch % 2 == 1:
                    continue
                if k == 2 and batch >= 2:
                    continue
                if k == 4 and i >= 10:
                    continue
                if k == 5 and batch == 3 and i >= 300:
                    continue
                keys.append(k)
                vals.append(rng.uniform() * (k + 1) + batch)
        if batch == 4:
            for _ in range(800):
                keys.append(3)
                vals.append(rng.uniform() * 4 + 4)
        keys = np.array(keys, np.uint32)
        vals = np.array(vals)
        g.add_batch(keys, vals)
        oracle.grouped_add_batch(refs, keys, vals)
    for k in range(nk):
        m, w = g.centroids(k)
        rm, rw = refs[k].centroids()
        check_centroids(m, w, rm, rw, 10.0, decay == 0.0, f"key {k}")
        for q in (0.0, 0.01, 0.25, 0.5, 0.75, 0.99, 1.0):
            check_values([g.quantile(k, q)], [refs[k].quantile(q)], 1.0, f"key {k} q {q}")


def test_grouped_many_keys_zipf(kd, oracle):
    rng = np.random.default_rng(10)
    nk = 3000
    p = 1.0 / np.arange(1, nk + 1) ** 1.1
    p /= p.sum()
    keys = rng.choice(nk, size=400000, p=p).astype(np.uint32)
    mu = rng.uniform(-100, 100, nk)
    vals = np.where(rng.uniform(size=keys.size) < 0.9, rng.standard_normal(keys.size) + mu[keys],
                    mu[keys] + (1 - rng.uniform(size=keys.size)) ** (-1 / 1.5))
    g = kd.GroupedTDigest(nk, 100.0, scale=kd.ScaleK3)
    g.add_batch(keys, vals)
    refs = [oracle.Digest(3, 100.0) for _ in range(nk)]
    oracle.grouped_add_batch(refs, keys, vals, threads=8)
    for k in range(0, nk, 7):
        m, w = g.centroids(k)
        rm, rw = refs[k].centroids()
        check_centroids(m, w, rm, rw, 200.0, True, f"key {k}")
    ks = np.arange(0, nk, 7)
    qk = np.repeat(ks, 4)
    qs = np.tile([1e-3, 0.01, 0.99, 0.999], ks.size)
    got = g.quantile(qk, qs)
    ref = np.array([refs[int(k)].quantile(float(q)) if len(refs[int(k)]) else np.nan
                    for k, q in zip(qk, qs)])
    ok = ~np.isnan(ref)
    check_values(got[ok], ref[ok], 100.0, "grouped quantile")


@pytest.mark.parametrize("nk,radix_bits", [(1, 27), (2, 27), (511, 27), (512, 27), (513, 27),
                                           (4097, 27), (1 << 18, 27), ((1 << 18) + 1, 27),
                                           ((1 << 20) + 3, 27), ((1 << 18) + 1, 18),
                                           (70_000, 12)])
def test_grouped_key_partition_sizes(kd, oracle, nk, radix_bits):
    """Key grouping at the digit-width edges of the partition (one radix pass
    up to 512 keys, two passes up to 2^18, three up to 2^27; the counting
    partition beyond the session's radix limit)."""
    rng = np.random.default_rng(nk + radix_bits)
    n = 300_001
    keys = (rng.zipf(1.3, n) - 1) % nk
    keys[:3] = [0, nk - 1, nk // 2]
    keys = keys.astype(np.uint32)
    vals = rng.standard_normal(n) + keys * 1e-3
    s = kd.GpuSession(0)
    s.set_radix_key_bits(radix_bits)
    g = kd.GroupedTDigest(nk, 100.0, scale=kd.ScaleK3, session=s)
    g.add_batch(keys, vals)
    cnt = np.bincount(keys, minlength=nk)
    chosen = sorted(set([0, nk - 1, nk // 2] + list(range(0, nk, max(1, nk // 40)))))
    for k in chosen:
        assert g.total_weight(k) == float(cnt[k]), (k, g.total_weight(k), cnt[k])
        if cnt[k] == 0:
            continue
        r = oracle.Digest(3, 100.0)
        r.add_batch_b(vals[keys == k])
        m, w = g.centroids(k)
        rm, rw = r.centroids()
        check_centroids(m, w, rm, rw, 300.0, True, f"nk {nk} key {k}")


@pytest.mark.parametrize("nk", [100, 5000, 300_000])
def test_grouped_rejects_bad_batches_without_change(kd, nk):
    # grouped.mojo:91-97: the whole batch is validated before any key changes.
    rng = np.random.default_rng(nk)
    g = kd.GroupedTDigest(nk, 100.0, scale=kd.ScaleK2)
    keys = rng.integers(0, nk, 50000).astype(np.uint32)
    vals = rng.standard_normal(50000)
    g.add_batch(keys, vals)
    before = [g.centroids(k) for k in range(0, nk, 7)]
    bad_key = keys.copy()
    bad_key[12345] = nk
    bad_val = vals.copy()
    bad_val[777] = np.inf
    nan_val = vals.copy()
    nan_val[-1] = np.nan
    for k_, v_ in ((bad_key, vals), (keys, bad_val), (keys, nan_val)):
        with pytest.raises(kd.KdigestError):
            g.add_batch(k_, v_)
    after = [g.centroids(k) for k in range(0, nk, 7)]
    for (m0, w0), (m1, w1) in zip(before, after):
        assert np.array_equal(m0, m1) and np.array_equal(w0, w1)
    g.add_batch(keys, vals)  # still usable
    assert g.total_weight(int(keys[0])) > 0


@pytest.mark.parametrize("nk,zipf,min_pairs", [(3001, 1.3, 32768), (1 << 17, 1.1, 1),
                                               (1 << 17, 3.0, 1 << 20), (1 << 18, 1.2, 1)])
def test_grouped_heavy_keys_identical(kd, oracle, nk, zipf, min_pairs):
    """Batches of 2^22 pairs and more place the keys holding many pairs in the
    first partition pass alone; grouping is identical with that split off
    (bit-identical centroids, two batches so the second merges; with 2^18
    keys the first pass has no digits to spare and the split stays off), the
    heaviest key matches the oracle, and a bad key in such a batch is still rejected
    without change."""
    rng = np.random.default_rng(nk + int(zipf * 10))
    n = (1 << 22) + 12345
    keys = ((rng.zipf(zipf, n) - 1) % nk).astype(np.uint32)
    vals = rng.standard_normal(n) * 2.0 + (keys % 13) * 0.5
    on = kd.GpuSession(0)
    on.set_heavy_keys(min_pairs)  # 1: the 255-key limit applies
    off = kd.GpuSession(0)
    off.set_heavy_keys(0)
    g = kd.GroupedTDigest(nk, 100.0, scale=kd.ScaleK2, session=on)
    h = kd.GroupedTDigest(nk, 100.0, scale=kd.ScaleK2, session=off)
    for half in (slice(0, n // 2), slice(n // 2, n)):
        g.add_batch(keys[half], vals[half])
        h.add_batch(keys[half], vals[half])
    cnt = np.bincount(keys, minlength=nk)
    top = [int(k) for k in np.argsort(cnt)[::-1][:300]]
    chosen = sorted(set(top + list(range(0, nk, max(1, nk // 200))) + [nk - 1]))
    for k in chosen:
        m, w = g.centroids(k)
        hm, hw = h.centroids(k)
        assert np.array_equal(m, hm) and np.array_equal(w, hw), k
        assert g.total_weight(k) == float(cnt[k]), k
    k = top[0]
    r = oracle.Digest(2, 100.0)
    sel = vals[keys == k]
    r.add_batch_b(sel[: np.count_nonzero(keys[: n // 2] == k)])
    r.add_batch_b(sel[np.count_nonzero(keys[: n // 2] == k):])
    m, w = g.centroids(k)
    rm, rw = r.centroids()
    check_centroids(m, w, rm, rw, 300.0, True, f"key {k}")
    bad = keys.copy()
    bad[n - 7] = nk
    before = g.centroids(top[0])
    with pytest.raises(kd.KdigestError):
        g.add_batch(bad, vals)
    after = g.centroids(top[0])
    assert np.array_equal(before[0], after[0]) and np.array_equal(before[1], after[1])


@pytest.mark.parametrize("items,chunks,nk", [(4, 1, 3001), (8, 3, 3001), (16, 7, 3001),
                                             (8, 4096, 3001), (4, 1 << 20, 3001),
                                             (8, 4096, 300_001), (4, 3, 300_001),
                                             (16, 1 << 20, 300_001)])
def test_grouped_partition_shapes_identical(kd, oracle, items, chunks, nk):
    """Every launch shape of the key partition (sub-tile width, chunks per
    pass: one chunk of many sub-tiles up to one sub-tile per chunk; two and
    three passes) groups the batch identically: centroids are bit-identical
    to the default shape, the heaviest key's are bit-identical to a
    one-key digest fed its values in batch order (the partition is stable),
    and they match the oracle."""
    rng = np.random.default_rng(items * 131 + chunks + nk)
    n = 250_003
    keys = ((rng.zipf(1.2, n) - 1) % nk).astype(np.uint32)
    vals = rng.standard_normal(n) * 3.0 + keys * 0.01
    s = kd.GpuSession(0)
    s.set_partition_shape(items, chunks)
    g = kd.GroupedTDigest(nk, 100.0, scale=kd.ScaleK3, session=s)
    g.add_batch(keys, vals)
    base = kd.GroupedTDigest(nk, 100.0, scale=kd.ScaleK3)
    base.add_batch(keys, vals)
    cnt = np.bincount(keys, minlength=nk)
    for k in list(range(0, nk, 97)) + [int(np.argmax(cnt))]:
        m, w = g.centroids(k)
        bm, bw = base.centroids(k)
        assert np.array_equal(m, bm) and np.array_equal(w, bw), k
        assert g.total_weight(k) == float(cnt[k])
    k = int(np.argmax(cnt))
    one = kd.GroupedTDigest(1, 100.0, scale=kd.ScaleK3)
    sel = vals[keys == k]
    one.add_batch(np.zeros(sel.size, dtype=np.uint32), sel)
    m, w = g.centroids(k)
    om, ow = one.centroids(0)
    assert np.array_equal(m, om) and np.array_equal(w, ow)
    r = oracle.Digest(3, 100.0)
    r.add_batch_b(sel)
    rm, rw = r.centroids()
    check_centroids(m, w, rm, rw, 300.0, True, f"key {k}")


@pytest.mark.parametrize("n,batch,lam,pre", [
    (1 << 20, 1 << 17, 0.0, 0),          # selection path, presorted samples
    ((1 << 20) + 12345, 1 << 17, 1e-5, 0),  # decay, a short last slice
    (300_000, 65_536, 2e-3, 37),         # samples buffered on the host first; renormalizes
    (50_000, 3000, 0.0, 0),              # one-block slices (no presort)
    (9000, 20_000, 1e-4, 5),             # one slice
])
def test_add_batches_device_equals_slices(kd, oracle, n, batch, lam, pre):
    """kd_tdigest_add_batches_device (consecutive batch merges, splitter
    samples of every slice sorted up front) gives the digest that
    add_batch_device gives slice by slice, bit for bit, and the oracle's."""
    import torch

    rng = np.random.default_rng(n + batch)
    x = np.exp(1.5 * rng.standard_normal(n))
    x[rng.integers(0, n, n // 50)] = 1.0  # ties
    head = rng.standard_normal(pre)
    dx = torch.from_numpy(x).cuda()
    a = kd.TDigest(kd.ScaleK2, 200.0, kd.Decay.rate(lam))
    b = kd.TDigest(kd.ScaleK2, 200.0, kd.Decay.rate(lam))
    r = oracle.Digest(2, 200.0, lam)
    for v in head:
        a.add(float(v))
        b.add(float(v))
        r.add(float(v))
    a.add_batches_device(dx, batch)
    for s0 in range(0, n, batch):
        b.add_batch_device(dx[s0:s0 + batch])
        r.add_batch_b(x[s0:s0 + batch])
    ma, wa = a.centroids()
    mb, wb = b.centroids()
    assert np.array_equal(ma.view(np.uint64), mb.view(np.uint64))
    assert np.array_equal(wa.view(np.uint64), wb.view(np.uint64))
    assert a.state() == b.state()
    rm, rw = r.centroids()
    check_centroids(ma, wa, rm, rw, float(x.max()), lam == 0.0, "batches")
    st, rs = a.state(), r.state()
    for k in ("clock", "log_scale", "min", "max", "stored_total"):
        assert st[k] == rs[k], (k, st[k], rs[k])


def test_add_batches_device_rejects_without_change(kd):
    import torch

    rng = np.random.default_rng(4)
    d = kd.TDigest(kd.ScaleK1, 100.0)
    d.add_batch_device(torch.from_numpy(rng.standard_normal(100_000)).cuda())
    before = (d.centroids(), d.state())
    x = rng.standard_normal(400_000)
    x[350_000] = np.nan  # in the last slice: nothing before it is merged either
    with pytest.raises(kd.KdigestError):
        d.add_batches_device(torch.from_numpy(x).cuda(), 1 << 17)
    (m0, w0), s0 = before
    m1, w1 = d.centroids()
    assert np.array_equal(m0, m1) and np.array_equal(w0, w1) and d.state() == s0
    with pytest.raises(kd.KdigestError):
        d.add_batches_device(torch.from_numpy(x[:10]).cuda(), 0)
