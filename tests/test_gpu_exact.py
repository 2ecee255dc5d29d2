"""Long decayed streams of batch merges against the oracle, merge by merge.

With decay every stored weight is fractional, and the reference's boundary
test sums them in its running order (tdigest.mojo:414-435).  Near q = 1 that
sum can round past the stored total, where K1's asin(2q - 1) is NaN and the
last record becomes a cluster of its own.  A CPU emulation of a closed-form
prefix finds such a disagreement in about one merge in twenty on these K1
streams; the device must follow the CPU every time (centroid counts and
weights exact, means within the bar)."""
import math

import numpy as np
import pytest

from tests.parity import check_centroids

pytestmark = pytest.mark.gpu

STREAMS = [
    # scale, half-life (samples), largest batch, merges
    (1, 1e6, 4000, 320),     # one-block path
    (1, 1e4, 4000, 320),
    (1, 1e6, 65536, 160),    # selection path above 4096 values
    (2, 1e4, 65536, 120),
    (3, 1e4, 65536, 120),
]


@pytest.mark.parametrize("scale,half_life,maxn,merges", STREAMS)
def test_long_decayed_stream_matches_oracle(kd, oracle, scale, half_life, maxn, merges):
    import torch

    rng = np.random.default_rng(int(half_life) + maxn + scale)
    lam = math.log(2.0) / half_life
    d = kd.TDigest(scale, 100.0, kd.Decay.rate(lam))
    r = oracle.Digest(scale, 100.0, lam)
    for it in range(merges):
        nv = int(np.exp(rng.uniform(0, math.log(maxn))))
        v = rng.standard_normal(nv)
        d.add_batch_device(torch.from_numpy(v).cuda())
        r.add_batch_b(v)
        m, w = d.centroids()
        rm, rw = r.centroids()
        assert m.size == rm.size, (it, nv, m.size, rm.size)
        # Weights are the running sums the CPU forms: bit for bit, and so is
        # the stored total summed from them.
        check_centroids(m, w, rm, rw, 6.0, True, f"merge {it}")
        st, rs = d.state(), r.state()
        for k in ("clock", "log_scale", "min", "max", "stored_total"):
            assert st[k] == rs[k], (it, k, st[k], rs[k])
