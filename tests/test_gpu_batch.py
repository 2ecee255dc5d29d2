"""GPU parity of the batched per-sample path (add_at replay of many digests)
and of the merge_into tree, against the CPU oracle."""
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
        for k in ("clock", "log_scale", "min", "max"):
            assert st[k] == rs[k], (tag, d, k, st[k], rs[k])
        assert abs(st["stored_total"] - rs["stored_total"]) <= 1e-12 * abs(rs["stored_total"]), (tag, d)


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
