"""Wire format of every key of a grouped digest at once: each key's bytes are
the per-digest layout (tdigest.mojo:627-652), byte-identical to the per-key
to_bytes; the restore gives back the same digests (and they keep ingesting
identically); a bad blob is rejected before anything is created."""
import struct

import numpy as np
import pytest

from tests.parity import check_centroids

pytestmark = pytest.mark.gpu


def _grouped(kd, nk, lam, seed, batches=2):
    rng = np.random.default_rng(seed)
    g = kd.GroupedTDigest(nk, 100.0, kd.Decay.rate(lam), scale=kd.ScaleK2)
    data = []
    for b in range(batches):
        n = 200_000
        # Keys above nk // 2 + 50 stay empty in the first batch; the last key
        # is never touched (an empty digest in the middle of the blobs).
        hi = min(nk // 2 + 50, nk - 1) if b == 0 else nk - 1
        keys = ((rng.zipf(1.3, n) - 1) % hi).astype(np.uint32)
        vals = rng.standard_normal(n) * 2.0 + keys * 0.01
        g.add_batch(keys, vals)
        data.append((keys, vals))
    return g, data


def _parse(blob):
    h = struct.unpack_from("<8d", blob, 1)
    c = int(h[7])
    pairs = np.frombuffer(blob, dtype="<f8", count=2 * c, offset=65).reshape(-1, 2) if c else np.zeros((0, 2))
    return h, pairs[:, 0], pairs[:, 1]


@pytest.mark.parametrize("lam", [0.0, 1e-5])
def test_wire_all_matches_per_key(kd, oracle, lam):
    nk = 1500
    g, data = _grouped(kd, nk, lam, 7)
    blob, off = g.to_bytes_all()
    assert off[0] == 0 and off[-1] == len(blob) == g.wire_size()
    for k in range(nk):
        assert blob[off[k]:off[k + 1]] == g.to_bytes(k), k
    # The empty key is a 65-byte header with count 0.
    assert off[nk] - off[nk - 1] == 65
    # Against the oracle fed the same per-key batches.
    for k in (0, 1, 2, 17, nk // 2, nk - 2):
        r = oracle.Digest(2, 100.0, lam)
        for keys, vals in data:
            sel = vals[keys == k]
            if sel.size:
                r.add_batch_b(sel)
        h, m, w = _parse(blob[off[k]:off[k + 1]])
        rh, rm, rw = _parse(r.to_bytes())
        assert h[0] == rh[0] and h[1] == rh[1], k  # compression, lambda
        assert h[7] == rh[7], k                      # count
        assert h[5] == rh[5] and h[6] == rh[6], k  # min, max
        # Stored total: bit-exact with integer weights; decayed (fractional)
        # weights are summed in another order, within 1e-12 relative.
        if lam == 0.0:
            assert h[4] == rh[4], k
        else:
            assert abs(h[4] - rh[4]) <= 1e-12 * abs(rh[4]), k
        if m.size:
            check_centroids(m, w, rm, rw, float(np.abs(m).max()) + 1.0, lam == 0.0, f"key {k}")


class _Hip:
    """Device memory from the HIP runtime the library itself loads."""

    def __init__(self):
        import ctypes as C

        self.C = C
        self.lib = C.CDLL("libamdhip64.so.7")
        self.ptrs = []

    def alloc(self, nbytes):
        p = self.C.c_void_p()
        assert self.lib.hipMalloc(self.C.byref(p), self.C.c_size_t(max(nbytes, 1))) == 0
        self.ptrs.append(p.value)
        return p.value

    def d2h(self, ptr, nbytes):
        buf = (self.C.c_uint8 * max(nbytes, 1))()
        assert self.lib.hipMemcpy(buf, self.C.c_void_p(ptr), self.C.c_size_t(nbytes), 2) == 0
        return bytes(buf)[:nbytes]

    def free(self):
        for p in self.ptrs:
            self.lib.hipFree(self.C.c_void_p(p))
        self.ptrs = []


def test_wire_device_matches_host(kd):
    g, _ = _grouped(kd, 777, 1e-5, 11)
    blob, off = g.to_bytes_all()
    hip = _Hip()
    try:
        n = g.wire_size()
        out = hip.alloc(n)
        doff = hip.alloc(8 * 778)
        g.to_bytes_device(out, doff, cap=n)
        assert hip.d2h(out, n) == blob
        assert np.array_equal(np.frombuffer(hip.d2h(doff, 8 * 778), dtype=np.int64), off)
        r = kd.GroupedTDigest.from_bytes_device(out, doff, scale=kd.ScaleK2, n=n, num_keys=777)
        assert r.compression == 100.0 and r.decay.rate_lambda == 1e-5
        rb, roff = r.to_bytes_all()
        assert rb == blob and np.array_equal(roff, off)
        # A buffer smaller than the wire size is refused.
        with pytest.raises(kd.KdigestError):
            g.to_bytes_device(out, doff, cap=n - 1)
    finally:
        hip.free()


@pytest.mark.parametrize("lam", [0.0, 1e-5])
def test_wire_round_trip_keeps_ingesting(kd, lam):
    nk = 2000
    g, _ = _grouped(kd, nk, lam, 23)
    blob, off = g.to_bytes_all()
    r = kd.GroupedTDigest.from_bytes_all(blob, off, scale=kd.ScaleK2)
    assert r.num_keys == nk and r.compression == 100.0 and r.decay.rate_lambda == lam
    for k in range(0, nk, 37):
        m, w = g.centroids(k)
        rm, rw = r.centroids(k)
        assert np.array_equal(m, rm) and np.array_equal(w, rw), k
        assert g.key_state(k) == r.key_state(k), k
    qs = np.linspace(0.0, 1.0, 33)
    for k in (0, 5, 99):
        assert np.array_equal(g.quantile(np.full(qs.size, k), qs), r.quantile(np.full(qs.size, k), qs))
    # The same next batch into both: identical digests, byte for byte.
    rng = np.random.default_rng(99)
    keys = rng.integers(0, nk, 300_000).astype(np.uint32)
    vals = rng.standard_normal(keys.size)
    g.add_batch(keys, vals)
    r.add_batch(keys, vals)
    assert g.to_bytes_all()[0] == r.to_bytes_all()[0]


def test_wire_restore_rejects_bad_blobs(kd):
    g, _ = _grouped(kd, 300, 0.0, 5, batches=1)
    blob, off = g.to_bytes_all()
    bad = bytearray(blob)
    # Truncated: the last key's bytes end one byte early.
    with pytest.raises(kd.KdigestError) as e:
        kd.GroupedTDigest.from_bytes_all(bytes(blob[:-1]), off, scale=kd.ScaleK2)
    assert e.value.code == -5
    # Offsets past the end.
    off2 = off.copy()
    off2[-1] += 8
    with pytest.raises(kd.KdigestError) as e:
        kd.GroupedTDigest.from_bytes_all(blob, off2, scale=kd.ScaleK2)
    assert e.value.code == -5
    # A bad version byte.
    bad[off[3]] = 7
    with pytest.raises(kd.KdigestError) as e:
        kd.GroupedTDigest.from_bytes_all(bytes(bad), off, scale=kd.ScaleK2)
    assert e.value.code == -4
    # Another compression for one key.
    bad = bytearray(blob)
    struct.pack_into("<d", bad, int(off[9]) + 1, 200.0)
    with pytest.raises(kd.KdigestError) as e:
        kd.GroupedTDigest.from_bytes_all(bytes(bad), off, scale=kd.ScaleK2)
    assert e.value.code == -4
    # A count that is not a whole number.
    bad = bytearray(blob)
    struct.pack_into("<d", bad, int(off[4]) + 57, 2.5)
    with pytest.raises(kd.KdigestError) as e:
        kd.GroupedTDigest.from_bytes_all(bytes(bad), off, scale=kd.ScaleK2)
    assert e.value.code == -4
    # The untouched blob still restores.
    r = kd.GroupedTDigest.from_bytes_all(blob, off, scale=kd.ScaleK2)
    assert r.to_bytes_all()[0] == blob


def test_wire_blobs_reduce_in_a_batch(kd, oracle):
    """The grouped wire buffer on the device loaded into a batch (one digest
    per key) and reduced by the merge_into tree, as the oracle reduces the
    same blobs."""
    nk = 40
    g, _ = _grouped(kd, nk, 0.0, 41)
    blob, off = g.to_bytes_all()
    hip = _Hip()
    try:
        n = g.wire_size()
        dd = hip.alloc(n)
        doff = hip.alloc(8 * (nk + 1))
        g.to_bytes_device(dd, doff, cap=n)
        b = kd.BatchedDigests.from_bytes_device(dd, doff, scale=kd.ScaleK2, n=n, num_digests=nk)
    finally:
        hip.free()
    for k in range(nk):
        m, w = b.centroids(k)
        gm, gw = g.centroids(k)
        assert np.array_equal(m, gm) and np.array_equal(w, gw), k
    b.merge_tree()
    refs = [oracle.Digest.from_bytes(blob[off[k]:off[k + 1]], 2) for k in range(nk)]
    stride = 1
    while stride < nk:
        for i in range(0, nk, 2 * stride):
            if i + stride < nk:
                refs[i].merge_into(refs[i + stride])
        stride *= 2
    m, w = b.centroids(0)
    rm, rw = refs[0].centroids()
    check_centroids(m, w, rm, rw, float(np.abs(rm).max()) + 1.0, True, "wire tree root")
