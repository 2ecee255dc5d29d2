"""ctypes binding of the CPU oracle (oracle/kd_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker.  The kdigest_amd package never
imports this module.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS: dict[str, C.CDLL] = {}

KO_OK, KO_EINVAL, KO_EEMPTY, KO_ERANGE, KO_ESERDE, KO_ETRUNC = 0, -1, -2, -3, -4, -5

_dp = C.POINTER(C.c_double)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def lib(variant: str = "kd") -> C.CDLL:
    """variant 'kd': k() through kd_math.h (bit-identical to the device);
    'glibc': k() through glibc log/asin (sensitivity check)."""
    if variant in _LIBS:
        return _LIBS[variant]
    name = "libkd_oracle.so" if variant == "kd" else "libkd_oracle_glibc.so"
    path = os.path.join(_HERE, "_build", name)
    if not os.path.exists(path):
        raise RuntimeError(f"oracle not built: {path} (run oracle/build.sh)")
    L = C.CDLL(path)
    vp, i64, d, i32 = C.c_void_p, C.c_int64, C.c_double, C.c_int
    sig = {
        "ko_new": (vp, [i32, d, d]),
        "ko_free": (None, [vp]),
        "ko_clone": (vp, [vp]),
        "ko_add": (i32, [vp, d, d]),
        "ko_add_at": (i32, [vp, d, d, d]),
        "ko_add_loop": (i32, [vp, _dp, i64]),
        "ko_add_batch_b": (i32, [vp, _dp, i64]),
        "ko_add_batch": (i32, [vp, _dp, i64, i32]),
        "ko_merge": (None, [vp]),
        "ko_quantile": (i32, [vp, d, _dp]),
        "ko_cdf": (i32, [vp, d, _dp]),
        "ko_merge_into": (i32, [vp, vp]),
        "ko_quantiles": (i32, [vp, _dp, i64, _dp]),
        "ko_cdfs": (i32, [vp, _dp, i64, _dp]),
        "ko_add_at_many": (i32, [vp, _dp, _dp, _dp, i64]),
        "ko_total_weight": (d, [vp]),
        "ko_to_bytes": (i64, [vp, C.c_char_p, i64]),
        "ko_from_bytes": (i32, [C.c_char_p, i64, i32, C.POINTER(vp)]),
        "ko_from_centroids": (vp, [i32, d, _dp, _dp, i64, d, d]),
        "ko_ncentroids": (i64, [vp]),
        "ko_nbuffer": (i64, [vp]),
        "ko_get_centroids": (None, [vp, _dp, _dp]),
        "ko_get_state": (None, [vp, _dp]),
        "ko_set_state": (None, [vp, _dp, _dp, i64, _dp, _dp, i64, _dp]),
        "ko_sweep": (i64, [i32, d, d, _dp, _dp, i64, _dp, _dp]),
        "ko_grouped_add_batch": (i32, [C.POINTER(vp), i64, C.POINTER(C.c_uint32), _dp, i64, i32]),
        "ko_batch_add_at": (i32, [C.POINTER(vp), i64, C.POINTER(i64), _dp, _dp, _dp, i32]),
        "ko_replay_ingests": (i32, [C.POINTER(vp), i64, C.POINTER(i64), C.POINTER(i64), _dp, i32]),
        "ko_k": (d, [i32, d, d]),
        "ko_add_loop_sum": (d, [d, d, i64]),
        "ko_normalizer": (d, [i32, d, d]),
        "ko_log": (d, [d]),
        "ko_asin": (d, [d]),
        "ko_exp": (d, [d]),
        "ko_rank_error": (d, [_dp, i64, d, d]),
    }
    for fn, (res, args) in sig.items():
        f = getattr(L, fn)
        f.restype = res
        f.argtypes = args
    _LIBS[variant] = L
    return L


class OracleError(Exception):
    pass


def _check(rc: int):
    if rc != KO_OK:
        raise OracleError(f"oracle error {rc}")


class Digest:
    """Oracle MergingDigest (tdigest.mojo:97-726)."""

    def __init__(self, scale: int = 1, compression: float = 100.0, decay: float = 0.0,
                 variant: str = "kd", _handle=None):
        self.L = lib(variant)
        self.scale = scale
        self.h = _handle if _handle is not None else self.L.ko_new(scale, compression, decay)

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            self.L.ko_free(h)
            self.h = None

    def clone(self) -> "Digest":
        return Digest(self.scale, _handle=self.L.ko_clone(self.h), variant="kd" if self.L is lib("kd") else "glibc")

    def add(self, v: float, w: float = 1.0):
        _check(self.L.ko_add(self.h, v, w))

    def add_at(self, v: float, t: float, w: float = 1.0):
        _check(self.L.ko_add_at(self.h, v, t, w))

    def add_loop(self, values):
        a = np.ascontiguousarray(values, dtype=np.float64)
        _check(self.L.ko_add_loop(self.h, _ptr(a), a.size))

    def add_batch_b(self, values):
        a = np.ascontiguousarray(values, dtype=np.float64)
        _check(self.L.ko_add_batch_b(self.h, _ptr(a), a.size))

    def add_batch(self, values, with_session: bool = False):
        a = np.ascontiguousarray(values, dtype=np.float64)
        _check(self.L.ko_add_batch(self.h, _ptr(a), a.size, int(with_session)))

    def merge(self):
        self.L.ko_merge(self.h)

    def quantile(self, q: float) -> float:
        out = C.c_double()
        _check(self.L.ko_quantile(self.h, q, C.byref(out)))
        return out.value

    def cdf(self, x: float) -> float:
        out = C.c_double()
        _check(self.L.ko_cdf(self.h, x, C.byref(out)))
        return out.value

    def quantiles(self, qs) -> np.ndarray:
        a = np.ascontiguousarray(qs, dtype=np.float64)
        out = np.empty(a.size)
        _check(self.L.ko_quantiles(self.h, _ptr(a), a.size, _ptr(out)))
        return out

    def cdfs(self, xs) -> np.ndarray:
        a = np.ascontiguousarray(xs, dtype=np.float64)
        out = np.empty(a.size)
        _check(self.L.ko_cdfs(self.h, _ptr(a), a.size, _ptr(out)))
        return out

    def add_at_many(self, values, times, weights):
        v = np.ascontiguousarray(values, dtype=np.float64)
        t = np.ascontiguousarray(times, dtype=np.float64)
        w = np.ascontiguousarray(weights, dtype=np.float64)
        _check(self.L.ko_add_at_many(self.h, _ptr(v), _ptr(t), _ptr(w), v.size))

    def merge_into(self, other: "Digest"):
        _check(self.L.ko_merge_into(self.h, other.h))

    def total_weight(self) -> float:
        return self.L.ko_total_weight(self.h)

    def __len__(self) -> int:
        return int(self.L.ko_ncentroids(self.h))

    def nbuffer(self) -> int:
        return int(self.L.ko_nbuffer(self.h))

    def centroids(self):
        n = len(self)
        m = np.empty(n)
        w = np.empty(n)
        if n:
            self.L.ko_get_centroids(self.h, _ptr(m), _ptr(w))
        return m, w

    def state(self) -> dict:
        o = np.empty(7)
        self.L.ko_get_state(self.h, _ptr(o))
        return dict(stored_total=o[0], log_scale=o[1], clock=o[2], min=o[3], max=o[4],
                    compression=o[5], decay=o[6])

    def set_state(self, means, weights, bmeans, bweights, stored_total, log_scale, clock, vmin, vmax):
        m = np.ascontiguousarray(means, np.float64)
        w = np.ascontiguousarray(weights, np.float64)
        bm = np.ascontiguousarray(bmeans, np.float64)
        bw = np.ascontiguousarray(bweights, np.float64)
        s = np.array([stored_total, log_scale, clock, vmin, vmax], np.float64)
        self.L.ko_set_state(self.h, _ptr(m), _ptr(w), m.size, _ptr(bm), _ptr(bw), bm.size, _ptr(s))

    def to_bytes(self) -> bytes:
        n = self.L.ko_to_bytes(self.h, None, 0)
        buf = C.create_string_buffer(n)
        self.L.ko_to_bytes(self.h, buf, n)
        return buf.raw

    @staticmethod
    def from_bytes(data: bytes, scale: int = 1) -> "Digest":
        L = lib("kd")
        h = C.c_void_p()
        _check(L.ko_from_bytes(data, len(data), scale, C.byref(h)))
        return Digest(scale, _handle=h.value)

    @staticmethod
    def from_centroids(scale, compression, means, weights, vmin, vmax) -> "Digest":
        L = lib("kd")
        m = np.ascontiguousarray(means, np.float64)
        w = np.ascontiguousarray(weights, np.float64)
        return Digest(scale, _handle=L.ko_from_centroids(scale, compression, _ptr(m), _ptr(w), m.size, vmin, vmax))


def sweep(scale, nrm, stored_n, means, weights, variant="kd"):
    L = lib(variant)
    m = np.ascontiguousarray(means, np.float64)
    w = np.ascontiguousarray(weights, np.float64)
    om = np.empty(max(m.size, 1))
    ow = np.empty(max(m.size, 1))
    c = L.ko_sweep(scale, nrm, stored_n, _ptr(m), _ptr(w), m.size, _ptr(om), _ptr(ow))
    return om[:c].copy(), ow[:c].copy()


def grouped_add_batch(digests, keys, values, threads: int = 1):
    """Per-key B semantics over a list of oracle digests (grouped.mojo:60-162)."""
    L = digests[0].L
    k = np.ascontiguousarray(keys, np.uint32)
    v = np.ascontiguousarray(values, np.float64)
    arr = (C.c_void_p * len(digests))(*[d.h for d in digests])
    _check(L.ko_grouped_add_batch(arr, len(digests), k.ctypes.data_as(C.POINTER(C.c_uint32)),
                                  _ptr(v), v.size, threads))


def k(scale, q, nrm, variant="kd"):
    return lib(variant).ko_k(scale, q, nrm)


def normalizer(scale, compression, n, variant="kd"):
    return lib(variant).ko_normalizer(scale, compression, n)


def kd_log(x):
    return lib("kd").ko_log(x)


def kd_asin(x):
    return lib("kd").ko_asin(x)


def rank_error(sorted_data: np.ndarray, v: float, q: float) -> float:
    a = np.ascontiguousarray(sorted_data, np.float64)
    return lib("kd").ko_rank_error(_ptr(a), a.size, v, q)


def kd_exp(x):
    return lib("kd").ko_exp(x)


def batch_add_at(digests, off, values, times, weights, threads: int = 1):
    """Digest d gets add_at of samples [off[d], off[d+1]) in order (tdigest.mojo:163-175)."""
    L = digests[0].L
    o = np.ascontiguousarray(off, np.int64)
    v = np.ascontiguousarray(values, np.float64)
    t = np.ascontiguousarray(times, np.float64)
    w = np.ascontiguousarray(weights, np.float64)
    arr = (C.c_void_p * len(digests))(*[d.h for d in digests])
    _check(L.ko_batch_add_at(arr, len(digests), o.ctypes.data_as(C.POINTER(C.c_int64)), _ptr(v),
                             _ptr(t), _ptr(w), threads))


def replay_ingests(digests, keys, off, data, threads: int = 1):
    """Ingest request r = one B batch merge of data[off[r]:off[r+1]] into
    digests[keys[r]], in arrival order per key (tdigest.mojo:223-315)."""
    L = digests[0].L
    k = np.ascontiguousarray(keys, np.int64)
    o = np.ascontiguousarray(off, np.int64)
    d = np.ascontiguousarray(data, np.float64)
    arr = (C.c_void_p * len(digests))(*[x.h for x in digests])
    _check(L.ko_replay_ingests(arr, k.size, k.ctypes.data_as(C.POINTER(C.c_int64)),
                               o.ctypes.data_as(C.POINTER(C.c_int64)), _ptr(d), threads))


def add_loop_sum(W: float, x: float, L: int) -> float:
    return lib("kd").ko_add_loop_sum(W, x, L)
