"""Dense ids for sparse uint32 keys on the device: first-occurrence order,
stable across batches, all-or-none at the capacity, and a grouped digest fed
through the map equals one fed the same ids computed on the host."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _first_order(keys, known=()):
    """Host reference: ids of new keys in order of first occurrence."""
    ids = {k: i for i, k in enumerate(known)}
    out = np.empty(len(keys), dtype=np.uint32)
    for i, k in enumerate(keys.tolist()):
        if k not in ids:
            ids[k] = len(ids)
        out[i] = ids[k]
    return out, list(ids.keys())


def test_keymap_first_occurrence_and_reverse(kd):
    rng = np.random.default_rng(3)
    pool = rng.choice(np.iinfo(np.uint32).max, 5000, replace=False).astype(np.uint32)
    pool[:3] = [0, 0xFFFFFFFF, 1]  # extremes are ordinary keys
    m = kd.KeyMap(10000)
    k1 = pool[rng.zipf(1.2, 200_000) % 3000]
    ids1 = m.map(k1)
    ref1, order = _first_order(k1)
    assert np.array_equal(ids1, ref1) and len(m) == len(order)
    # A second batch: known keys keep their ids, new ones follow.
    k2 = pool[rng.integers(0, 5000, 100_000)]
    ids2 = m.map(k2)
    ref2, order = _first_order(k2, order)
    assert np.array_equal(ids2, ref2) and len(m) == len(order)
    assert np.array_equal(m.keys(), np.array(order, dtype=np.uint32))
    # Lookup without insert: unknown keys are UNKNOWN, the map unchanged.
    probe = np.array([pool[0], 123456789, pool[4999]], dtype=np.uint32)
    known = set(order)
    got = m.map(probe, insert=False)
    for p, g in zip(probe.tolist(), got.tolist()):
        assert g == (order.index(p) if p in known else kd.KeyMap.UNKNOWN)
    assert len(m) == len(order)


def test_keymap_capacity_all_or_none(kd):
    m = kd.KeyMap(1000)
    a = np.arange(900, dtype=np.uint32) * 7919
    m.map(a)
    before = m.keys()
    # 200 new keys (with repeats) do not fit in the 100 left: nothing changes.
    b = np.concatenate([a[:50], np.arange(200, dtype=np.uint32) * 104729 + 3]).repeat(3)
    with pytest.raises(kd.KdigestError) as e:
        m.map(b)
    assert e.value.code == -7
    assert len(m) == 900 and np.array_equal(m.keys(), before)
    assert np.array_equal(m.map(a[:10], insert=False), np.arange(10, dtype=np.uint32))
    # Exactly the remaining capacity fits.
    c = np.arange(100, dtype=np.uint32) * 104729 + 3
    assert np.array_equal(m.map(c), np.arange(900, 1000, dtype=np.uint32))
    assert len(m) == 1000
    # A batch with far more distinct keys than the map holds fails cleanly.
    with pytest.raises(kd.KdigestError):
        m.map(np.arange(2_000_000, dtype=np.uint32) + 10_000_000)
    assert len(m) == 1000


def test_keymap_feeds_grouped_digest(kd):
    rng = np.random.default_rng(8)
    pool = rng.choice(2**32 - 1, 3000, replace=False).astype(np.uint32)
    m = kd.KeyMap(3000)
    g = kd.GroupedTDigest(3000, 100.0, scale=kd.ScaleK1)
    h = kd.GroupedTDigest(3000, 100.0, scale=kd.ScaleK1)
    order = []
    for b in range(3):
        keys = pool[rng.zipf(1.3, 300_000) % 3000]
        vals = rng.standard_normal(keys.size)
        g.add_batch(m.map(keys), vals)
        ids, order = _first_order(keys, order)
        h.add_batch(ids, vals)
    assert g.to_bytes_all()[0] == h.to_bytes_all()[0]
