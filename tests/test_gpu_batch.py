"""GPU parity of the batched per-sample path (add_at replay of many digests)
and of the merge_into tree, against the CPU oracle."""
import struct

import numpy as np
import pytest

from tests.parity import check_centroids

pytestmark = pytest.mark.gpu


def _streams(rng, sizes, int_weights):
    off = np.zeros(len(sizes) + 1, dtype=np.int64)
    off[1:] = np.cumsum(sizes)
    n = int(off[-1])
    vals = rng.standard_t(3, n)
    wts = rng.integers(1, 4, n).astype(np.float64) if int_weights else rng.uniform(0.1, 10, n)
    times = np.empty(n)
    for d in range(len(sizes)):
        a, b = off[d], off[d + 1]
        times[a:b] = np.cumsum(rng.exponential(1e-3, b - a))
    return off, vals, times, wts


def _oracle_digests(oracle, scale, delta, lam, off, vals, times, wts):
    out = []
    for d in range(off.size - 1):
        r = oracle.Digest(scale, delta, lam)
        for i in range(off[d], off[d + 1]):
            r.add_at(vals[i], times[i], wts[i])
        out.append(r)
    return out


def _compare(kd_b, refs, integer_weights, tag):
    for d, r in enumerate(refs):
        r.merge()
        m, w = kd_b.centroids(d)
        rm, rw = r.centroids()
        st, rs = kd_b.state(d), r.state()
        assert st["error"] == 0 and st["buffered"] == 0
        check_centroids(m, w, rm, rw, 20.0, integer_weights, f"{tag} digest {d}")
        # The exponent and the stored total are the reference's sequential
        # sums, bit for bit.
        for k in ("clock", "log_scale", "min", "max", "stored_total"):
            assert st[k] == rs[k], (tag, d, k, st[k], rs[k])


@pytest.mark.parametrize("lam,int_weights", [(0.0, True), (1e-3, False), (0.5, False)])
def test_batch_add_at_matches_oracle(kd, oracle, lam, int_weights):
    rng = np.random.default_rng(7)
    sizes = [0, 1, 2, 17, 1499, 1500, 1501, 3000, 4321, 12000, 250, 6000]
    off, vals, times, wts = _streams(rng, sizes, int_weights)
    if lam == 0.5:
        # Timestamps 0..n so that the decay exponent crosses 400 (renormalize).
        for d in range(len(sizes)):
            times[off[d]:off[d + 1]] = np.arange(off[d + 1] - off[d], dtype=np.float64)
    b = kd.BatchedDigests(len(sizes), 300.0, kd.Decay.rate(lam), scale=kd.ScaleK2)
    # Two calls: the buffer carries over between them.
    half = np.array([off[d] + (off[d + 1] - off[d]) // 2 for d in range(len(sizes))])
    off1 = np.zeros_like(off)
    off1[1:] = np.cumsum(half - off[:-1])
    idx1 = np.concatenate([np.arange(off[d], half[d]) for d in range(len(sizes))]).astype(np.int64)
    idx2 = np.concatenate([np.arange(half[d], off[d + 1]) for d in range(len(sizes))]).astype(np.int64)
    off2 = np.zeros_like(off)
    off2[1:] = np.cumsum(off[1:] - half)
    b.add_at(off1, vals[idx1], times[idx1], wts[idx1])
    b.add_at(off2, vals[idx2], times[idx2], wts[idx2])
    b.flush()
    refs = _oracle_digests(oracle, 2, 300.0, lam, off, vals, times, wts)
    _compare(b, refs, int_weights, f"lam={lam}")


@pytest.mark.parametrize("lam", [0.0, 1e-3])
def test_batch_merge_tree_matches_oracle(kd, oracle, lam):
    rng = np.random.default_rng(11)
    nd = 13
    sizes = rng.integers(0, 5000, nd)
    sizes[3] = 20000
    off, vals, times, wts = _streams(rng, sizes, lam == 0.0)
    b = kd.BatchedDigests(nd, 300.0, kd.Decay.rate(lam), scale=kd.ScaleK2)
    b.add_at(off, vals, times, wts)
    b.merge_tree()
    refs = _oracle_digests(oracle, 2, 300.0, lam, off, vals, times, wts)
    for r in refs:
        r.merge()
    stride = 1
    while stride < nd:
        for i in range(0, nd, 2 * stride):
            if i + stride < nd:
                refs[i].merge_into(refs[i + stride])
        stride *= 2
    m, w = b.centroids(0)
    rm, rw = refs[0].centroids()
    check_centroids(m, w, rm, rw, 50.0, lam == 0.0, "tree root")
    st, rs = b.state(0), refs[0].state()
    for k in ("clock", "min", "max"):
        assert st[k] == rs[k], (k, st[k], rs[k])
    t = b.to_tdigest(0)
    for q in (0.001, 0.5, 0.999):
        got, ref = t.quantile(q), refs[0].quantile(q)
        assert abs(got - ref) <= 1e-9 * max(1.0, abs(ref)), (q, got, ref)


def test_batch_rejects_bad_samples_without_change(kd):
    b = kd.BatchedDigests(2, 100.0, scale=kd.ScaleK1)
    off = np.array([0, 2, 4], np.int64)
    b.add_at(off, np.array([1.0, 2.0, 3.0, 4.0]), np.arange(4.0), np.ones(4))
    before = b.state(0)
    with pytest.raises(kd.KdigestError):
        b.add_at(off, np.array([1.0, np.nan, 3.0, 4.0]), np.arange(4.0), np.ones(4))
    assert b.state(0) == before


def test_batch_merge_tree_graph_replays_match_eager(kd):
    """The merge tree captured as a HIP graph, replayed on later calls with
    eager work (more add_at samples, state reads) in between: after every
    replay the digests equal those of an eager twin, bit for bit."""
    rng = np.random.default_rng(5)
    nd = 37
    g = kd.BatchedDigests(nd, 200.0, kd.Decay.rate(1e-3), scale=kd.ScaleK2)
    e = kd.BatchedDigests(nd, 200.0, kd.Decay.rate(1e-3), scale=kd.ScaleK2)
    e.set_graph(False)
    t0 = 0.0
    for rnd in range(4):
        sizes = rng.integers(0, 3000, nd)
        off, vals, times, wts = _streams(rng, sizes, False)
        times += t0
        t0 = float(times.max()) + 1.0
        for b in (g, e):
            b.add_at(off, vals, times, wts)
            b.merge_tree()
        for d in range(nd):
            mg, wg = g.centroids(d)
            me, we = e.centroids(d)
            assert np.array_equal(mg.view(np.uint64), me.view(np.uint64)), (rnd, d)
            assert np.array_equal(wg.view(np.uint64), we.view(np.uint64)), (rnd, d)
            assert g.state(d) == e.state(d), (rnd, d)


def _oracle_tree(refs):
    for r in refs:
        r.merge()
    stride = 1
    while stride < len(refs):
        for i in range(0, len(refs), 2 * stride):
            if i + stride < len(refs):
                refs[i].merge_into(refs[i + stride])
        stride *= 2
    return refs[0]


@pytest.mark.parametrize("lam", [0.0, 1e-3])
def test_batch_map_reduce_from_bytes(kd, oracle, lam):
    """Map-reduce finalize (test_edge.mojo:183-215): digests built elsewhere
    (here by the oracle) and serialized are loaded into a batch from their
    bytes, each exactly as serialized, and the merge_into tree reduces them
    as the oracle's tree does."""
    rng = np.random.default_rng(31)
    nd = 21
    refs = []
    for d in range(nd):
        r = oracle.Digest(2, 300.0, lam)
        n = 0 if d == 5 else int(rng.integers(1, 6000))
        vals = rng.standard_t(3, n) + d
        times = np.cumsum(rng.exponential(1e-3, n))
        wts = rng.integers(1, 4, n).astype(np.float64) if lam == 0.0 else rng.uniform(0.1, 10, n)
        for i in range(n):
            r.add_at(vals[i], times[i], wts[i])
        refs.append(r)
    blobs = [r.to_bytes() for r in refs]
    off = np.zeros(nd + 1, np.int64)
    off[1:] = np.cumsum([len(x) for x in blobs])
    b = kd.BatchedDigests.from_bytes_all(b"".join(blobs), off, scale=kd.ScaleK2)
    assert b.num_digests == nd and b.compression == 300.0 and b.decay.rate_lambda == lam
    for d in range(nd):
        m, w = b.centroids(d)
        rm, rw = refs[d].centroids()
        assert np.array_equal(m, rm) and np.array_equal(w, rw), d
        st, rs = b.state(d), refs[d].state()
        for k in ("stored_total", "log_scale", "clock", "min", "max"):
            assert st[k] == rs[k], (d, k)
        assert st["buffered"] == 0 and st["error"] == 0
    b.merge_tree()
    root = _oracle_tree(refs)
    m, w = b.centroids(0)
    rm, rw = root.centroids()
    check_centroids(m, w, rm, rw, 50.0, lam == 0.0, "map-reduce root")
    st, rs = b.state(0), root.state()
    for k in ("clock", "min", "max"):
        assert st[k] == rs[k], k
    assert abs(st["stored_total"] - rs["stored_total"]) <= 1e-12 * abs(rs["stored_total"])


def test_batch_from_bytes_rejects(kd, oracle):
    a = oracle.Digest(2, 100.0)
    a.add_batch_b(np.linspace(0.0, 1.0, 1000))
    c = oracle.Digest(2, 200.0)
    c.add_batch_b(np.linspace(0.0, 1.0, 1000))
    ba, bc = a.to_bytes(), c.to_bytes()
    # Another compression than digest 0's.
    with pytest.raises(kd.KdigestError) as e:
        kd.BatchedDigests.from_bytes_all(ba + bc, np.array([0, len(ba), len(ba) + len(bc)]),
                                         scale=kd.ScaleK2)
    assert e.value.code == -4
    # More centroids than a batched digest holds.
    big = bytearray(ba[:65])
    struct.pack_into("<d", big, 57, 2000.0)
    big += np.column_stack([np.arange(2000.0), np.ones(2000)]).astype("<f8").tobytes()
    with pytest.raises(kd.KdigestError) as e:
        kd.BatchedDigests.from_bytes_all(ba + bytes(big), np.array([0, len(ba), len(ba) + len(big)]),
                                         scale=kd.ScaleK2)
    assert e.value.code == -1
    # Truncated.
    with pytest.raises(kd.KdigestError) as e:
        kd.BatchedDigests.from_bytes_all(ba[:-3], np.array([0, len(ba)]), scale=kd.ScaleK2)
    assert e.value.code == -5


def test_batch_running_sums_with_exact_halves(kd, oracle):
    """The replay's decay exponent and stored total are the reference's
    sequential sums (tdigest.mojo:324-347) formed by the block in binade
    rounds; a step whose increment is exactly half a unit of the sum's last
    place rounds to the even sum, which depends on every step before it, and
    is added on its own.  Digest 0: lambda 0.75 and steps of 2^-43 at a clock
    near 350, so every decay increment is 1.5 units of the exponent's last
    place (about 262 to 400); digest 1 (no decay): a first weight of 2^30,
    then weights of 1.5 units of the total's last place.  Both exactly as the
    oracle, and the exponent crosses the renormalization threshold."""
    lam = 0.75
    n0 = 3000
    t = [350.0]
    for i in range(1, n0):
        t.append(t[-1] + (2.0 ** -43 if i % 50 else 3.5))
    times0 = np.array(t)
    vals0 = np.random.default_rng(2).standard_normal(n0)
    w0 = np.ones(n0)
    n1 = 2500
    w1 = np.full(n1, 1.5 * 2.0 ** -22)
    w1[0] = 2.0 ** 30
    w1[1::7] = 3.0 * 2.0 ** -23  # 1.5 units as well, written another way
    vals1 = np.random.default_rng(3).standard_normal(n1)
    times1 = np.arange(n1, dtype=np.float64)
    for d, (vals, times, wts, l) in enumerate([(vals0, times0, w0, lam), (vals1, times1, w1, 0.0)]):
        b = kd.BatchedDigests(1, 100.0, kd.Decay.rate(l), scale=kd.ScaleK1)
        off = np.array([0, vals.size], np.int64)
        b.add_at(off, vals, times, wts)
        b.flush()
        r = oracle.Digest(1, 100.0, l)
        for i in range(vals.size):
            r.add_at(vals[i], times[i], wts[i])
        r.merge()
        st, rs = b.state(0), r.state()
        for k in ("clock", "log_scale", "min", "max", "stored_total"):
            assert st[k] == rs[k], (d, k, st[k], rs[k])
        m, w = b.centroids(0)
        rm, rw = r.centroids()
        check_centroids(m, w, rm, rw, 5.0, False, f"digest {d}")
