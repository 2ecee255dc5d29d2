"""The allocation-free grouped query (kd_grouped_quantile_device_noalloc)
gives exactly the results of kd_grouped_quantile_device, call after call, on
inputs that shrink between calls; it reports the same errors and refuses a
digest whose query scratch would not fit its fixed per-GPU array."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _digest(kd, rng, nk, n, lam):
    keys = np.minimum(rng.zipf(1.2, n) - 1, nk - 1).astype(np.uint32)
    vals = rng.standard_normal(n) * 3 + keys
    g = kd.GroupedTDigest(nk, 100.0, kd.Decay.rate(lam), scale=kd.ScaleK2)
    g.add_batch(keys, vals)
    live = np.unique(keys).astype(np.int64)
    return g, live


def _queries(torch, rng, live, nq):
    k = torch.from_numpy(rng.choice(live, nq)).cuda()
    q = rng.uniform(0, 1, nq)
    q[:4] = (0.0, 1.0, 1e-4, 0.5)
    return k, torch.from_numpy(q).cuda()


def test_noalloc_equals_allocating_entry(kd):
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(2024)
    big, big_live = _digest(kd, rng, 5000, 1 << 21, 1e-6)
    small, small_live = _digest(kd, rng, 300, 50000, 0.0)
    for g, live, nq in ((big, big_live, 300000), (small, small_live, 1000), (big, big_live, 77)):
        k, q = _queries(torch, rng, live, nq)
        ref = torch.empty(nq, dtype=torch.float64, device="cuda")
        got = torch.full((nq,), -1.0, dtype=torch.float64, device="cuda")
        g.quantile_device(k, q, ref)
        g.quantile_device_noalloc(k, q, got)
        r, o = ref.cpu().numpy(), got.cpu().numpy()
        assert np.array_equal(r.view(np.uint64), o.view(np.uint64)), nq


def test_noalloc_errors_match(kd):
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(7)
    g, live = _digest(kd, rng, 50, 20000, 0.0)
    k = torch.tensor([int(live[0]), 50, int(live[1])], dtype=torch.int64, device="cuda")
    q = torch.tensor([0.5, 0.5, 0.25], dtype=torch.float64, device="cuda")
    out = torch.zeros(3, dtype=torch.float64, device="cuda")
    with pytest.raises(kd.KdigestError) as e:
        g.quantile_device_noalloc(k, q, out)
    assert e.value.code == -1
    o = out.cpu().numpy()
    assert np.isnan(o[1]) and o[0] == g.quantile(int(live[0]), 0.5) and o[2] == g.quantile(int(live[1]), 0.25)
    q2 = torch.tensor([0.5, 1.5, 0.25], dtype=torch.float64, device="cuda")
    k2 = torch.tensor([int(live[0])] * 3, dtype=torch.int64, device="cuda")
    with pytest.raises(kd.KdigestError) as e:
        g.quantile_device_noalloc(k2, q2, out)
    assert e.value.code == -3


def test_noalloc_refuses_oversized_digest(kd):
    torch = pytest.importorskip("torch")
    # 32 bytes of key state per key: 3 million keys need 96 MB > 64 MiB.
    g = kd.GroupedTDigest(3_000_000, 100.0, scale=kd.ScaleK1)
    g.add_batch(np.array([5], np.uint32), np.array([1.0]))
    k = torch.tensor([5], dtype=torch.int64, device="cuda")
    q = torch.tensor([0.5], dtype=torch.float64, device="cuda")
    out = torch.full((1,), 7.0, dtype=torch.float64, device="cuda")
    with pytest.raises(kd.KdigestError) as e:
        g.quantile_device_noalloc(k, q, out)
    assert e.value.code == -1 and "scratch" in str(e.value)
    assert out.item() == 7.0  # nothing launched
    g.quantile_device(k, q, out)
    assert out.item() == 1.0
