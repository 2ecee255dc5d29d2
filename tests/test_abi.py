"""CPU checks of the C ABI: the library loads, exports every function that
include/kdigest_hip.h declares, the Python binding declares the same set, and
the host-side scale functions equal the oracle's (same kd_math.h code)."""
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "kdigest_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(kd_[a-z0-9_]+)\s*\(", src)))


def test_header_functions_exported(kd):
    import ctypes

    lib = ctypes.CDLL(kd.LIB_PATH)
    names = _declared()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_matches_header(kd):
    from kdigest_amd._lib import SIGNATURES

    assert sorted(SIGNATURES) == _declared()


def test_host_scale_functions_match_oracle(kd, oracle):
    L = kd.load()
    rng = np.random.default_rng(0)
    for scale in range(4):
        for n in (10.0, 1e4, 2.5e8):
            assert L.kd_scale_normalizer(scale, 100.0, n) == oracle.normalizer(scale, 100.0, n)
        nrm = oracle.normalizer(scale, 100.0, 1e6)
        for q in np.concatenate([rng.uniform(0, 1, 2000), [0.0, 1.0, 1e-16, 0.5]]):
            assert L.kd_scale_k(scale, float(q), nrm) == oracle.k(scale, float(q), nrm)


def test_python_api_surface(kd):
    # kdigest/__init__.mojo:10-25
    for name in ("TDigest", "GroupedTDigest", "GpuSession", "Decay", "ScaleK0", "ScaleK1",
                 "ScaleK2", "ScaleK3", "ScaleFunction"):
        assert hasattr(kd, name)
    for meth in ("add", "add_at", "add_batch", "merge", "quantile", "cdf", "merge_into",
                 "to_bytes", "from_bytes", "total_weight", "min", "max"):
        assert hasattr(kd.TDigest, meth)
    for meth in ("add_batch", "quantile", "cdf", "to_bytes", "centroid_count", "total_weight"):
        assert hasattr(kd.GroupedTDigest, meth)
