"""Compress by selection (no full sort of large batches) against the sorting
path and the oracle: the same clusters (counts and integer weights bit-exact,
means within 1e-12), and a fall back to sorting where a boundary may tie
(tdigest.mojo:429)."""
import numpy as np
import pytest

from tests.parity import check_centroids, check_values

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sessions(kd):
    sel = kd.GpuSession()
    srt = kd.GpuSession()
    srt.set_select(False)
    return sel, srt


def _data(kind, rng, n):
    if kind == "normal":
        return rng.standard_normal(n)
    if kind == "lognormal":
        return rng.lognormal(0.0, 1.5, n)
    if kind == "pareto":
        return (1.0 - rng.uniform(size=n)) ** (-1.0 / 1.5)
    if kind == "sorted":
        return np.sort(rng.standard_t(3, n))
    if kind == "reversed":
        return np.sort(rng.standard_normal(n))[::-1].copy()
    if kind == "straddle":
        # Three groups far apart (one around zero); spacing well above an ulp.
        return rng.uniform(-1e-3, 1e-3, n) + rng.choice([-1e3, 0.0, 1e3], n)
    if kind == "tiny":
        return rng.standard_normal(n) * 1e-200
    if kind == "periodic":
        base = rng.standard_normal(1024)
        return np.tile(base, n // 1024 + 1)[:n] + np.arange(n) * 1e-9
    raise ValueError(kind)


KINDS = ["normal", "lognormal", "pareto", "sorted", "reversed", "straddle", "tiny", "periodic"]


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("scale", [0, 1, 2, 3])
def test_select_matches_sort_and_oracle(kd, oracle, sessions, kind, scale):
    s_sel, s_srt = sessions
    rng = np.random.default_rng(hash((kind, scale)) % (2**32))
    before = s_sel.path_counts()
    a = kd.TDigest(scale, 150.0, session=s_sel)
    b = kd.TDigest(scale, 150.0, session=s_srt)
    r = oracle.Digest(scale, 150.0)
    for n in (70001, 5000, 250000):  # the second and third batches carry prior centroids
        x = _data(kind, rng, n)
        a.add_batch(x)
        b.add_batch(x)
        r.add_batch(x)
        scale_abs = float(np.abs(x).max())
        ma, wa = a.centroids()
        mb, wb = b.centroids()
        rm, rw = r.centroids()
        check_centroids(ma, wa, mb, wb, scale_abs, True, f"{kind} K{scale} n={n} sel vs sort")
        check_centroids(ma, wa, rm, rw, scale_abs, True, f"{kind} K{scale} n={n} sel vs oracle")
        sa, rs = a.state(), r.state()
        for k in ("min", "max", "stored_total"):
            assert sa[k] == rs[k], (kind, scale, n, k, sa[k], rs[k])
    after = s_sel.path_counts()
    assert after["selection"] > before["selection"], (before, after)
    qs = np.concatenate([rng.uniform(0, 1, 200), [0.0, 1e-4, 0.5, 1 - 1e-4, 1.0]])
    check_values(a.quantile(qs), np.array([r.quantile(float(q)) for q in qs]),
                 max(1.0, float(np.abs(x).max())), f"{kind} K{scale} quantile")


@pytest.mark.parametrize("scale", [0, 1, 2, 3])
def test_select_many_prior_centroids(kd, oracle, sessions, scale):
    """Compression 1500: batches merging into more than 512 prior
    centroids (the chain then reads their positions from memory instead of
    LDS) give the sort path's clusters and the oracle's."""
    s_sel, s_srt = sessions
    rng = np.random.default_rng(900 + scale)
    a = kd.TDigest(scale, 1500.0, session=s_sel)
    b = kd.TDigest(scale, 1500.0, session=s_srt)
    r = oracle.Digest(scale, 1500.0)
    big_prior = False
    sizes = []
    for n in (300001, 200000, 400000):
        x = rng.standard_t(4, n) * 10.0
        prior = len(a)
        sizes.append(prior)
        before = s_sel.path_counts()["selection"]
        a.add_batch(x)
        b.add_batch(x)
        r.add_batch(x)
        if prior > 512 and s_sel.path_counts()["selection"] > before:
            big_prior = True
        ma, wa = a.centroids()
        mb, wb = b.centroids()
        rm, rw = r.centroids()
        scale_abs = float(np.abs(x).max())
        check_centroids(ma, wa, mb, wb, scale_abs, True, f"K{scale} n={n} sel vs sort")
        check_centroids(ma, wa, rm, rw, scale_abs, True, f"K{scale} n={n} sel vs oracle")
    assert big_prior, f"no selection merge with more than 512 prior centroids: {sizes}"


@pytest.mark.parametrize("lam", [1e-5, 2e-3])
def test_select_with_decay(kd, oracle, sessions, lam):
    s_sel, s_srt = sessions
    rng = np.random.default_rng(77)
    a = kd.TDigest(2, 200.0, kd.Decay.rate(lam), session=s_sel)
    r = oracle.Digest(2, 200.0, lam)
    for i in range(6):
        x = rng.lognormal(0.0, 1.5, 65536 + 123 * i)
        a.add_batch(x)
        r.add_batch(x)
        ma, wa = a.centroids()
        rm, rw = r.centroids()
        check_centroids(ma, wa, rm, rw, float(x.max()), False, f"decay {lam} batch {i}")
        sa, rs = a.state(), r.state()
        for k in ("log_scale", "clock", "min", "max"):
            assert sa[k] == rs[k], (k, sa[k], rs[k])


@pytest.mark.parametrize("scale,compression", [(3, 100.0), (3, 300.0), (2, 200.0)])
def test_select_large_buckets(kd, oracle, sessions, scale, compression):
    """2^25 values in one batch: the 8191 splitters leave buckets of about
    4096 values, so many cluster boundaries fall in buckets selected by
    radix select rather than sorted, and the tail buckets hold ten to forty
    boundaries each (selected up to 16 at a time); a second batch merges
    into the result.  Same clusters as the sort path and the oracle."""
    s_sel, s_srt = sessions
    rng = np.random.default_rng(int(compression) + scale)
    a = kd.TDigest(scale, compression, session=s_sel)
    b = kd.TDigest(scale, compression, session=s_srt)
    r = oracle.Digest(scale, compression)
    for n in (1 << 25, (1 << 22) + 7):
        x = np.where(rng.uniform(size=n) < 0.9, rng.standard_normal(n),
                     (1.0 - rng.uniform(size=n)) ** (-1.0 / 1.5))
        before = s_sel.path_counts()["selection"]
        a.add_batch(x)
        b.add_batch(x)
        r.add_batch(x)
        assert s_sel.path_counts()["selection"] > before
        ma, wa = a.centroids()
        mb, wb = b.centroids()
        rm, rw = r.centroids()
        scale_abs = float(np.abs(x).max())
        check_centroids(ma, wa, mb, wb, scale_abs, True, f"K{scale} d{compression} n={n} sel vs sort")
        check_centroids(ma, wa, rm, rw, scale_abs, True, f"K{scale} d{compression} n={n} sel vs oracle")
        sa, rs = a.state(), r.state()
        for k in ("min", "max", "stored_total"):
            assert sa[k] == rs[k], (scale, compression, n, k, sa[k], rs[k])


def test_select_ties_fall_back_to_sort(kd, oracle, sessions):
    s_sel, _ = sessions
    rng = np.random.default_rng(5)
    before = s_sel.path_counts()
    a = kd.TDigest(1, 100.0, session=s_sel)
    r = oracle.Digest(1, 100.0)
    x = np.round(rng.standard_normal(100000) * 3.0)  # few distinct values: boundaries tie
    a.add_batch(x)
    r.add_batch(x)
    ma, wa = a.centroids()
    rm, rw = r.centroids()
    check_centroids(ma, wa, rm, rw, 20.0, True, "ties")
    after = s_sel.path_counts()
    assert after["sorted"] > before["sorted"], (before, after)


def test_select_grouped_zipf(kd, oracle, sessions):
    s_sel, s_srt = sessions
    rng = np.random.default_rng(12)
    nk = 400
    p = 1.0 / np.arange(1, nk + 1) ** 1.1
    p /= p.sum()
    keys = rng.choice(nk, size=600000, p=p).astype(np.uint32)
    mu = rng.uniform(-100, 100, nk)
    vals = np.where(rng.uniform(size=keys.size) < 0.9, rng.standard_normal(keys.size) + mu[keys],
                    mu[keys] + (1 - rng.uniform(size=keys.size)) ** (-1 / 1.5))
    ga = kd.GroupedTDigest(nk, 100.0, scale=kd.ScaleK3, session=s_sel)
    gb = kd.GroupedTDigest(nk, 100.0, scale=kd.ScaleK3, session=s_srt)
    for half in (slice(0, 300000), slice(300000, None)):
        ga.add_batch(keys[half], vals[half])
        gb.add_batch(keys[half], vals[half])
    refs = [oracle.Digest(3, 100.0) for _ in range(nk)]
    oracle.grouped_add_batch(refs, keys[:300000], vals[:300000], threads=8)
    oracle.grouped_add_batch(refs, keys[300000:], vals[300000:], threads=8)
    for k in range(nk):
        ma, wa = ga.centroids(k)
        mb, wb = gb.centroids(k)
        rm, rw = refs[k].centroids()
        check_centroids(ma, wa, mb, wb, 300.0, True, f"key {k} sel vs sort")
        check_centroids(ma, wa, rm, rw, 300.0, True, f"key {k} sel vs oracle")


@pytest.mark.parametrize("lam", [0.0, 2e-5])
def test_select_grouped_fallbacks(kd, oracle, lam):
    """Grouped batches whose large digests partly fall back from the
    selection path to sorting (few distinct values: boundaries tie): those
    digests are finished after the batch's sync, and every key (state
    included) equals the oracle's, before and after the fallen-back
    digests' keys take prior centroids."""
    rng = np.random.default_rng(int(lam * 1e6) + 31)
    s = kd.GpuSession(0)
    nk = 24
    g = kd.GroupedTDigest(nk, 100.0, kd.Decay.rate(lam), scale=kd.ScaleK1, session=s)
    refs = [oracle.Digest(1, 100.0, lam) for _ in range(nk)]
    for batch in range(2):
        # Keys 0-5 tied (selection fallbacks), 6-11 large and distinct
        # (selection), 12-23 small (one block).
        parts_k, parts_v = [], []
        for k in range(nk):
            n = 60000 + 1000 * k if k < 12 else 500 + 37 * k
            x = rng.standard_normal(n) * 3.0 + k
            if k < 6:
                x = np.round(x)
            parts_k.append(np.full(n, k, np.uint32))
            parts_v.append(x)
        keys = np.concatenate(parts_k)
        vals = np.concatenate(parts_v)
        perm = rng.permutation(keys.size)
        keys, vals = keys[perm], vals[perm]
        before = s.path_counts()
        g.add_batch(keys, vals)
        after = s.path_counts()
        oracle.grouped_add_batch(refs, keys, vals, threads=8)
        assert after["sorted"] > before["sorted"], (batch, before, after)
        assert after["selection"] > before["selection"], (batch, before, after)
        for k in range(nk):
            m, w = g.centroids(k)
            rm, rw = refs[k].centroids()
            check_centroids(m, w, rm, rw, 40.0, lam == 0.0, f"batch {batch} key {k}")
            st, rs = g.key_state(k), refs[k].state()
            for f in ("min", "max", "stored_total", "clock", "log_scale"):
                assert st[f] == rs[f], (batch, k, f, st[f], rs[f])


def test_grouped_csr_larger_than_last_batch(kd, oracle):
    """A batch of the same shape as the session's last one that ends with
    more centroids than it did (so the next CSR, sized from that batch,
    is gathered again with its exact size) equals the oracle."""
    s = kd.GpuSession(0)
    nk, n = 50, 200000
    rng = np.random.default_rng(8)
    keys = rng.integers(0, nk, n).astype(np.uint32)
    few = kd.GroupedTDigest(nk, 100.0, scale=kd.ScaleK2, session=s)
    few.add_batch(keys, np.round(rng.standard_normal(n)))  # few distinct values: few centroids
    g = kd.GroupedTDigest(nk, 100.0, scale=kd.ScaleK2, session=s)
    refs = [oracle.Digest(2, 100.0) for _ in range(nk)]
    vals = rng.standard_normal(n) + keys
    g.add_batch(keys, vals)
    oracle.grouped_add_batch(refs, keys, vals, threads=8)
    total = 0
    for k in range(nk):
        m, w = g.centroids(k)
        rm, rw = refs[k].centroids()
        check_centroids(m, w, rm, rw, 60.0, True, f"key {k}")
        total += m.size
    assert total > sum(few.centroid_count(k) for k in range(nk))


def test_one_block_path_matches_sort_path(kd, oracle):
    """Digests of at most 4096 merged records compressed in one block give
    the multi-kernel path's clusters bit for bit (same arithmetic, same
    order), with and without prior centroids and decay."""
    fused, multi = kd.GpuSession(), kd.GpuSession()
    multi.set_fused(False)
    rng = np.random.default_rng(21)
    for scale, lam in ((1, 0.0), (2, 1e-4), (3, 0.0), (0, 5e-3)):
        before = fused.path_counts()
        a = kd.GroupedTDigest(300, 100.0, kd.Decay.rate(lam), scale=scale, session=fused)
        b = kd.GroupedTDigest(300, 100.0, kd.Decay.rate(lam), scale=scale, session=multi)
        for _ in range(3):
            keys = (rng.zipf(1.2, 200000) % 300).astype(np.uint32)
            vals = rng.standard_t(3, keys.size) + keys
            vals[::7] = np.round(vals[::7])  # some ties
            a.add_batch(keys, vals)
            b.add_batch(keys, vals)
        for k in range(300):
            ma, wa = a.centroids(k)
            mb, wb = b.centroids(k)
            assert ma.shape == mb.shape, (scale, k)
            assert np.array_equal(ma.view(np.uint64), mb.view(np.uint64)), (scale, lam, k)
            assert np.array_equal(wa.view(np.uint64), wb.view(np.uint64)), (scale, lam, k)
        after = fused.path_counts()
        assert after["one_block"] > before["one_block"]


@pytest.mark.parametrize("tile_min,tiles_max", [(1024, 4096), (65536, 1), (16384, 256)])
def test_select_shapes_identical(kd, oracle, tile_min, tiles_max):
    """The tiling of the selection passes (many small tiles per digest up to
    one tile per digest) does not change any cluster."""
    rng = np.random.default_rng(tile_min + tiles_max)
    nk = 40
    keys = ((rng.zipf(1.1, 600_000) - 1) % nk).astype(np.uint32)
    vals = rng.standard_normal(keys.size) + keys * 2.0
    s = kd.GpuSession(0)
    s.set_select_shape(tile_min, tiles_max)
    g = kd.GroupedTDigest(nk, 100.0, scale=kd.ScaleK3, session=s)
    g.add_batch(keys, vals)
    base = kd.GroupedTDigest(nk, 100.0, scale=kd.ScaleK3)
    base.add_batch(keys, vals)
    assert s.path_counts()["selection"] > 0
    for k in range(nk):
        m, w = g.centroids(k)
        bm, bw = base.centroids(k)
        assert np.array_equal(m.view(np.uint64), bm.view(np.uint64)), k
        assert np.array_equal(w.view(np.uint64), bw.view(np.uint64)), k
    k = 0
    r = oracle.Digest(3, 100.0)
    r.add_batch_b(vals[keys == k])
    rm, rw = r.centroids()
    m, w = g.centroids(k)
    check_centroids(m, w, rm, rw, 300.0, True, f"key {k}")


def test_autotune_grouped_keeps_results(kd):
    """The autotuner times candidate launch shapes on a device batch, leaves
    the fastest on the session, and digests built afterwards equal those of
    the default shape bit for bit."""
    rng = np.random.default_rng(3)
    nk = 500
    keys = ((rng.zipf(1.1, 400_000) - 1) % nk).astype(np.uint32)
    vals = rng.standard_normal(keys.size) + keys * 0.5
    s = kd.GpuSession(0)
    shape = s.autotune_grouped(keys, vals, nk, scale=kd.ScaleK3, compression=100.0, reps=1)
    assert shape["items"] in (4, 8, 16) and shape["ms_after"] <= shape["ms_before"] * 1.000001
    g = kd.GroupedTDigest(nk, 100.0, scale=kd.ScaleK3, session=s)
    g.add_batch(keys, vals)
    # The comparison session fixes its own shape (the recorded choice is then
    # not applied to it).
    b = kd.GpuSession(0)
    b.set_partition_shape(8, 4096)
    b.set_select_shape(65536, 64)
    base = kd.GroupedTDigest(nk, 100.0, scale=kd.ScaleK3, session=b)
    base.add_batch(keys, vals)
    for k in range(nk):
        m, w = g.centroids(k)
        bm, bw = base.centroids(k)
        assert np.array_equal(m.view(np.uint64), bm.view(np.uint64)), k
        assert np.array_equal(w.view(np.uint64), bw.view(np.uint64)), k
    # Another session ingesting a batch of the same class runs with the
    # recorded shape; the pinned one keeps its own.
    fresh = kd.GpuSession(0)
    f = kd.GroupedTDigest(nk, 100.0, scale=kd.ScaleK3, session=fresh)
    f.add_batch(keys[: keys.size - 1000], vals[: keys.size - 1000])
    got = fresh.shape()
    for k in ("items", "chunks", "tile_min", "tiles_max", "overlap"):
        assert got[k] == shape[k], (k, got, shape)
    assert b.shape()["items"] == 8 and b.shape()["tile_min"] == 65536


def test_autotune_table_survives_save_and_load(kd, tmp_path):
    """A saved table, loaded into the process, gives a fresh session the
    recorded shape for that class."""
    path = tmp_path / "tune.txt"
    kd.GpuSession.autotune_save(str(path))
    assert kd.GpuSession.autotune_load(str(path)) >= 0


def test_overlap_keeps_results(kd):
    """Small digests compressed on the second stream while the selection
    path runs (set_overlap 1: after the selection count; 2: from the start)
    give the same digests bit for bit as on the session stream, over two
    batches (the second merges into the first's centroids)."""
    rng = np.random.default_rng(8)
    nk = 3001
    sessions = [kd.GpuSession(0) for _ in range(3)]
    for mode, s in enumerate(sessions):
        s.set_overlap(mode)
        assert s.shape()["overlap"] == mode
    with pytest.raises(kd.KdigestError):
        sessions[0].set_overlap(3)
    digs = [kd.GroupedTDigest(nk, 100.0, scale=kd.ScaleK3, session=s) for s in sessions]
    for _ in range(2):
        keys = ((rng.zipf(1.2, 2_000_000) - 1) % nk).astype(np.uint32)
        vals = rng.standard_normal(keys.size) * 5.0 + keys * 0.01
        for g in digs:
            g.add_batch(keys, vals)
    paths = sessions[1].path_counts()
    assert paths["selection"] > 0 and paths["one_block"] > 0, paths
    for k in range(nk):
        bm, bw = digs[0].centroids(k)
        for g in digs[1:]:
            m, w = g.centroids(k)
            assert np.array_equal(m.view(np.uint64), bm.view(np.uint64)), k
            assert np.array_equal(w.view(np.uint64), bw.view(np.uint64)), k
