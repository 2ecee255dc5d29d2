"""GPU edge cases: digests wrapped from centroids, signed zeros as batch
extremes, checked device queries, and malformed wire blobs."""
import math
import struct

import numpy as np
import pytest

from tests.parity import check_centroids, check_values

pytestmark = pytest.mark.gpu


def _same_blob(got: bytes, ref: bytes, tag=""):
    """Wire blobs equal in every header byte and every centroid weight, means
    within the 1e-12 bar (long clusters may round their last bit
    differently)."""
    assert got[:65] == ref[:65], (tag, struct.unpack("<8d", got[1:65]), struct.unpack("<8d", ref[1:65]))
    assert len(got) == len(ref), tag
    g = np.frombuffer(got[65:], "<f8").reshape(-1, 2)
    r = np.frombuffer(ref[65:], "<f8").reshape(-1, 2)
    lo, hi = struct.unpack("<2d", ref[41:57])
    check_centroids(g[:, 0], g[:, 1], r[:, 0], r[:, 1], max(abs(lo), abs(hi)), True, tag)
    # Zero means carry the same sign as the reference's.
    z = r[:, 0] == 0.0
    assert np.array_equal(np.signbit(g[z, 0]), np.signbit(r[z, 0])), tag


def test_from_centroids_matches_oracle(kd, oracle):
    # tdigest.mojo:700-726: no decay, stored total = clock = sum of weights;
    # the wrapped digest answers queries and serializes as the oracle's.
    rng = np.random.default_rng(41)
    m = np.sort(rng.standard_normal(57)) * 3
    w = rng.integers(1, 40, m.size).astype(np.float64)
    w[3] = 0.5
    d = kd.TDigest._from_centroids(150.0, m, w, m[0] - 1.0, m[-1] + 2.0, kd.ScaleK2)
    r = oracle.Digest.from_centroids(2, 150.0, m, w, m[0] - 1.0, m[-1] + 2.0)
    assert d.to_bytes() == r.to_bytes()
    st = d.state()
    assert st["stored_total"] == st["clock"] == sum(w.tolist())
    qs = np.concatenate([rng.uniform(0, 1, 200), [0.0, 1.0, 1e-3, 1 - 1e-3]])
    check_values(d.quantile(qs), [r.quantile(float(q)) for q in qs], 1.0, "quantile")
    xs = np.concatenate([rng.normal(0, 4, 200), [m[0] - 1.0, m[-1] + 2.0, m[5]]])
    check_values(d.cdf(xs), [r.cdf(float(x)) for x in xs], 1.0, "cdf")
    # It keeps ingesting like any digest.
    x = rng.standard_normal(70000)
    d.add_batch(x)
    r.add_batch(x)
    _same_blob(d.to_bytes(), r.to_bytes(), "after a batch")


def _with_zeros(rng, n, first, second, sign):
    """n values of one sign (all >= 0 when sign > 0) with two zeros planted:
    `first` at position 17 and `second` at n // 2."""
    x = np.abs(rng.standard_normal(n)) * sign
    x[x == 0.0] = sign
    x[17] = first
    x[n // 2] = second
    return x


ZERO_ORDERS = [(-0.0, 0.0), (0.0, -0.0), (-0.0, -0.0)]


@pytest.mark.parametrize("n", [3000, 70000, 400000])
@pytest.mark.parametrize("first,second", ZERO_ORDERS)
@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_signed_zero_extremes_single(kd, oracle, n, first, second, sign):
    # tdigest.mojo:272-295: the batch min (max for non-positive data) is the
    # first zero in batch order; the bytes carry its sign.  3000 values run
    # the one-block path, 70000 and 400000 the selection path.
    rng = np.random.default_rng(int(n) + int(sign))
    x = _with_zeros(rng, n, first, second, sign)
    d = kd.TDigest(kd.ScaleK3, 100.0)
    r = oracle.Digest(3, 100.0)
    d.add_batch_device(__import__("torch").from_numpy(x).cuda())
    r.add_batch_b(x)
    ext = d.state()["min"] if sign > 0 else d.state()["max"]
    assert math.copysign(1.0, ext) == math.copysign(1.0, first)
    _same_blob(d.to_bytes(), r.to_bytes(), "single")


def test_signed_zero_extremes_grouped(kd, oracle):
    rng = np.random.default_rng(77)
    nk = 6
    sizes = [40, 3000, 9000, 70000, 200000, 5]
    keys, vals = [], []
    for k, s in enumerate(sizes):
        first, second = ZERO_ORDERS[k % 3]
        sign = 1.0 if k % 2 == 0 else -1.0
        keys.append(np.full(s, k, np.uint32))
        vals.append(_with_zeros(rng, s, first, second, sign) if s > 20 else
                    np.array([0.0, -0.0, 1.0, -0.0, 2.0]) * (1.0 if k % 2 == 0 else -1.0))
    keys = np.concatenate(keys)
    vals = np.concatenate(vals)
    perm = rng.permutation(keys.size)
    keys, vals = keys[perm], vals[perm]
    g = kd.GroupedTDigest(nk, 100.0, scale=kd.ScaleK3)
    g.add_batch(keys, vals)
    refs = [oracle.Digest(3, 100.0) for _ in range(nk)]
    oracle.grouped_add_batch(refs, keys, vals)
    for k in range(nk):
        _same_blob(g.to_bytes(k), refs[k].to_bytes(), f"key {k}")


@pytest.fixture(scope="module")
def small_grouped(kd):
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(3)
    keys = rng.integers(0, 3, 30000).astype(np.uint32)  # keys 0..2; key 3 stays empty
    vals = rng.standard_normal(keys.size) + keys
    g = kd.GroupedTDigest(4, 100.0, scale=kd.ScaleK1)
    g.add_batch(keys, vals)
    return torch, g


def test_grouped_quantile_device_matches_host(kd, small_grouped):
    torch, g = small_grouped
    rng = np.random.default_rng(4)
    k = rng.integers(0, 3, 5000).astype(np.int64)
    q = rng.uniform(0, 1, 5000)
    q[:3] = (0.0, 1.0, 0.5)
    out = torch.empty(5000, dtype=torch.float64, device="cuda")
    g.quantile_device(torch.from_numpy(k).cuda(), torch.from_numpy(q).cuda(), out)
    ref = g.quantile(k, q)
    assert np.array_equal(out.cpu().numpy(), ref)


@pytest.mark.parametrize("bad,code", [("key", -1), ("negkey", -1), ("q", -3), ("qnan", -3),
                                      ("empty", -2)])
def test_grouped_quantile_device_rejects(kd, small_grouped, bad, code):
    # The device checks every query; the bad ones come back NaN with the
    # reference's error (a key out of range, q outside [0, 1], an empty key),
    # the others are answered.
    torch, g = small_grouped
    k = np.array([0, 1, 2, 1], np.int64)
    q = np.array([0.1, 0.5, 0.9, 0.25])
    if bad == "key":
        k[1] = 4
    elif bad == "negkey":
        k[1] = -7
    elif bad == "q":
        q[1] = 1.5
    elif bad == "qnan":
        q[1] = math.nan
    else:
        k[1] = 3
    out = torch.zeros(4, dtype=torch.float64, device="cuda")
    with pytest.raises(kd.KdigestError) as e:
        g.quantile_device(torch.from_numpy(k).cuda(), torch.from_numpy(q).cuda(), out)
    assert e.value.code == code
    o = out.cpu().numpy()
    assert math.isnan(o[1])
    good = [0, 2, 3]
    assert np.array_equal(o[good], g.quantile(k[good], q[good]))


def test_grouped_quantile_device_argument_checks(kd, small_grouped):
    torch, g = small_grouped
    k = torch.zeros(4, dtype=torch.int64, device="cuda")
    q = torch.full((4,), 0.5, dtype=torch.float64, device="cuda")
    out = torch.empty(4, dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        g.quantile_device(k.to(torch.int32), q, out)
    with pytest.raises(ValueError):
        g.quantile_device(k[:3], q, out)
    with pytest.raises(ValueError):
        g.quantile_device(k, q, out[:2])
    with pytest.raises(ValueError):
        g.quantile_device(k.cpu(), q, out)
    g.quantile_device(k, q, out)
    assert np.array_equal(out.cpu().numpy(), g.quantile(np.zeros(4, np.int64), np.full(4, 0.5)))


def _blob(count, pairs):
    return bytes([1]) + struct.pack("<8d", 100.0, 0.0, 0.0, 3.0, 3.0, 1.0, 3.0, count) + \
        struct.pack(f"<{2 * len(pairs)}d", *[v for p in pairs for v in p])


@pytest.mark.parametrize("count,code", [(math.nan, -4), (-1.0, -4), (1.5, -4), (math.inf, -4),
                                        (1e300, -4), (3.0, -5), (1e15, -5)])
def test_from_bytes_rejects_bad_counts(kd, count, code):
    # The count field is a double: a whole number >= 0 whose pairs the blob
    # holds (tdigest.mojo:679-692 reads that many; short data raises
    # "truncated digest").
    blob = _blob(count, [(1.0, 1.0), (2.0, 2.0)])
    with pytest.raises(kd.KdigestError) as e:
        kd.TDigest.from_bytes(blob, kd.ScaleK1)
    assert e.value.code == code
    ok = kd.TDigest.from_bytes(_blob(2.0, [(1.0, 1.0), (2.0, 2.0)]), kd.ScaleK1)
    assert len(ok) == 2 and ok.quantile(0.0) == 1.0


@pytest.mark.parametrize("scale,delta", [(1, 100.0), (0, 2000.0)])
def test_grouped_device_queries_staged_equal_unstaged(kd, scale, delta):
    """A block of device queries that all name one key stages that key's
    centroids in LDS (keys of up to 512 centroids); the same queries in a
    shuffled order take the unstaged path.  Both give the same bits, for
    quantiles and cdfs, including keys of more than 512 centroids (ScaleK0,
    compression 2000) that a uniform block does not stage."""
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(scale + 7)
    nk = 24
    sizes = rng.integers(1, 200_000, nk)
    sizes[5] = 3          # tiny
    keys = np.repeat(np.arange(nk, dtype=np.uint32), sizes)
    vals = rng.standard_t(4, keys.size) + keys.astype(np.float64)
    perm = rng.permutation(keys.size)
    sc = [kd.ScaleK0, kd.ScaleK1, kd.ScaleK2, kd.ScaleK3][scale]
    g = kd.GroupedTDigest(nk, delta, scale=sc)
    g.add_batch(keys[perm], vals[perm])
    if scale == 0:
        assert max(len(g.centroids(k)[0]) for k in range(nk)) > 512
    per = 1000  # not a multiple of the block size: some blocks straddle two keys
    qk = np.repeat(np.arange(nk, dtype=np.int64), per)
    qs = rng.uniform(0, 1, qk.size)
    qs[::97] = 0.0
    qs[::89] = 1.0
    xs = np.concatenate([g.quantile(np.full(per, k, np.int64), rng.uniform(0, 1, per))
                         for k in range(nk)])
    shuf = rng.permutation(qk.size)
    for fn, arg in (("quantile_device", qs), ("cdf_device", xs)):
        if not hasattr(g, fn):
            continue
        dev = getattr(g, fn)
        a = torch.empty(qk.size, dtype=torch.float64, device="cuda")
        b = torch.empty(qk.size, dtype=torch.float64, device="cuda")
        dev(torch.from_numpy(qk).cuda(), torch.from_numpy(arg).cuda(), a)
        dev(torch.from_numpy(qk[shuf]).cuda(), torch.from_numpy(arg[shuf]).cuda(), b)
        got_sorted = a.cpu().numpy()
        got_shuf = np.empty(qk.size)
        got_shuf[shuf] = b.cpu().numpy()
        assert np.array_equal(got_sorted.view(np.uint64), got_shuf.view(np.uint64)), fn


def test_grouped_device_queries_empty_key_block(kd):
    """A block whose queries all name an empty key is not staged: every
    query comes back NaN with KD_EEMPTY (tdigest.mojo:485-487)."""
    torch = pytest.importorskip("torch")
    g = kd.GroupedTDigest(3, 100.0, scale=kd.ScaleK1)
    g.add_batch(np.array([0, 2, 2], np.uint32), np.array([1.0, 2.0, 3.0]))
    k = torch.full((300,), 1, dtype=torch.int64, device="cuda")
    q = torch.full((300,), 0.5, dtype=torch.float64, device="cuda")
    out = torch.zeros(300, dtype=torch.float64, device="cuda")
    with pytest.raises(kd.KdigestError) as e:
        g.quantile_device(k, q, out)
    assert e.value.code == -2
    assert np.isnan(out.cpu().numpy()).all()
