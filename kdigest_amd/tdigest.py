"""TDigest / GroupedTDigest / GpuSession: the reference's public API over the
HIP library.

Mirrors kdigest/__init__.mojo:10-25: scale functions ScaleK0..ScaleK3
(scale.mojo), Decay (decay.mojo), TDigest (tdigest.mojo), GroupedTDigest
(grouped.mojo) and GpuSession (gpu/session.mojo).  Every operation calls the C
ABI of include/kdigest_hip.h; errors surface as KdigestError, raised before
any state changes where the reference validates first.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from typing import Iterable, Sequence

import numpy as np

from ._lib import KdigestError, check, load

_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)
_u32p = C.POINTER(C.c_uint32)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _ptr(a: np.ndarray, t=_dp):
    return a.ctypes.data_as(t)


# ------------------------------------------------------------------ scales
class ScaleFunction:
    """k(q) scale function (scale.mojo:50-75); `id` selects it in the ABI."""

    id: int = -1
    name: str = ""

    def normalizer(self, compression: float, n: float) -> float:
        return load().kd_scale_normalizer(self.id, compression, n)

    def k(self, q: float, normalizer: float) -> float:
        return load().kd_scale_k(self.id, q, normalizer)

    def q(self, k: float, normalizer: float) -> float:
        return load().kd_scale_q(self.id, k, normalizer)

    def __repr__(self) -> str:
        return f"{self.name}()"

    def __eq__(self, other) -> bool:
        return isinstance(other, ScaleFunction) and other.id == self.id

    def __hash__(self) -> int:
        return self.id


class ScaleK0(ScaleFunction):
    """Linear: k = normalizer * q (scale.mojo:78-99)."""
    id, name = 0, "ScaleK0"


class ScaleK1(ScaleFunction):
    """k = normalizer * asin(2q - 1) (scale.mojo:102-124)."""
    id, name = 1, "ScaleK1"


class ScaleK2(ScaleFunction):
    """k = normalizer * log(q / (1 - q)) (scale.mojo:127-157)."""
    id, name = 2, "ScaleK2"


class ScaleK3(ScaleFunction):
    """Piecewise log scale (scale.mojo:160-192)."""
    id, name = 3, "ScaleK3"


def _scale(s) -> ScaleFunction:
    if isinstance(s, ScaleFunction):
        return s
    if isinstance(s, type) and issubclass(s, ScaleFunction):
        return s()
    if isinstance(s, int) and 0 <= s <= 3:
        return (ScaleK0, ScaleK1, ScaleK2, ScaleK3)[s]()
    raise ValueError(f"not a scale function: {s!r}")


# ------------------------------------------------------------------- decay
class Decay:
    """Exponential decay rate lambda (decay.mojo:23-51)."""

    def __init__(self, rate_lambda: float = 0.0):
        self.rate_lambda = float(rate_lambda)

    @staticmethod
    def none() -> "Decay":
        return Decay(0.0)

    @staticmethod
    def rate(lambda_: float) -> "Decay":
        return Decay(lambda_)

    @staticmethod
    def half_life(h: float) -> "Decay":
        if h <= 0.0:
            raise ValueError("half_life must be positive")
        return Decay(math.log(2.0) / h)

    def is_active(self) -> bool:
        return self.rate_lambda > 0.0

    def __repr__(self) -> str:
        return f"Decay(lambda={self.rate_lambda})"


# ----------------------------------------------------------------- session
class GpuSession:
    """One device, one HIP stream and reusable workspaces
    (gpu/session.mojo:19-27)."""

    def __init__(self, device: int = 0):
        L = load()
        h = C.c_void_p()
        check(L.kd_session_create(device, C.byref(h)))
        self._h = h.value
        self.device = device

    @property
    def handle(self):
        return self._h

    def synchronize(self) -> None:
        check(load().kd_session_sync(self._h))

    def profile(self, enable: bool = True) -> None:
        """Per-phase device timing with HIP events (clears the totals)."""
        check(load().kd_session_profile(self._h, int(enable)))

    def set_select(self, enable: bool = True) -> None:
        """Compress digests receiving more than 4096 new values by selection
        (default) or always by a full sort; both give the same clusters."""
        check(load().kd_session_set_select(self.handle, 1 if enable else 0))

    def set_fused(self, enable: bool = True) -> None:
        """Compress digests of at most 4096 merged records in one block each
        (default) or through the multi-kernel sort path; same clusters."""
        check(load().kd_session_set_fused(self.handle, 1 if enable else 0))

    def set_partition_shape(self, items: int = 8, chunks: int = 4096) -> None:
        """Launch shape of the grouped key partition: items per thread of a
        scatter sub-tile (4, 8, 16) and the target chunks per pass."""
        check(load().kd_session_set_partition_shape(self.handle, int(items), int(chunks)))

    def set_overlap(self, mode: int | bool = 1) -> None:
        """Where small digests are compressed: 0 (False, the default) on the
        session stream after the selection samples of large ones; 1 (True) on
        a second stream once the selection count is done; 2 on a second
        stream from the start of the compress."""
        check(load().kd_session_set_overlap(self.handle, int(mode)))

    def set_heavy_keys(self, min_pairs: int) -> None:
        """Keys of a large grouped batch estimated to hold at least min_pairs
        pairs are placed by the first partition pass alone (0: off)."""
        check(load().kd_session_set_heavy_keys(self.handle, int(min_pairs)))

    def set_radix_key_bits(self, bits: int = 27) -> None:
        """Grouped batches over at most 2^bits keys take the stable radix
        partition; more keys the counting partition."""
        check(load().kd_session_set_radix_key_bits(self.handle, int(bits)))

    def autotune_grouped(self, keys, values, num_keys: int, scale: int = 3,
                         compression: float = 100.0, reps: int = 2) -> dict:
        """Times the grouped ingest of a device-resident batch (torch tensors
        or device pointers) under candidate launch shapes and keeps the
        fastest on this session; returns the shape and both timings."""
        scale = _scale(scale).id
        shp = np.zeros(5, np.int64)
        ms = np.zeros(2)
        if isinstance(keys, np.ndarray):
            k = np.ascontiguousarray(keys, dtype=np.uint32)
            v = _f64(values)
            if k.size != v.size:
                raise ValueError("keys and values differ in length")
            check(load().kd_session_autotune_grouped_host(
                self.handle, scale, int(num_keys), float(compression), _ptr(k, _u32p), _ptr(v),
                int(k.size), int(reps), _ptr(shp, _i64p), _ptr(ms)))
        else:
            kn, vn = int(keys.numel()), int(values.numel())
            if kn != vn:
                raise ValueError("keys and values differ in length")
            check(load().kd_session_autotune_grouped(
                self.handle, scale, int(num_keys), float(compression), _device_ptr(keys),
                _device_ptr(values), kn, int(reps), _ptr(shp, _i64p), _ptr(ms)))
        return {"items": int(shp[0]), "chunks": int(shp[1]), "tile_min": int(shp[2]),
                "tiles_max": int(shp[3]), "overlap": int(shp[4]),
                "ms_before": float(ms[0]), "ms_after": float(ms[1])}

    def shape(self) -> dict:
        """The grouped-ingest launch shape this session runs with now."""
        o = np.zeros(5, np.int64)
        check(load().kd_session_shape(self.handle, _ptr(o, _i64p)))
        return {"items": int(o[0]), "chunks": int(o[1]), "tile_min": int(o[2]),
                "tiles_max": int(o[3]), "overlap": int(o[4])}

    @staticmethod
    def autotune_save(path: str) -> None:
        """Writes the process's recorded autotune choices to a text file."""
        check(load().kd_autotune_save(os.fsencode(path)))

    @staticmethod
    def autotune_load(path: str) -> int:
        """Merges recorded autotune choices from a file; returns how many."""
        n = C.c_int64()
        check(load().kd_autotune_load(os.fsencode(path), C.byref(n)))
        return n.value

    def set_select_shape(self, tile_min: int = 65536, tiles_max: int = 64) -> None:
        """Tiles of the selection compress: at least tile_min values per tile,
        at most tiles_max tiles per digest."""
        check(load().kd_session_set_select_shape(self.handle, int(tile_min), int(tiles_max)))

    def path_counts(self) -> dict:
        o = np.zeros(3, np.int64)
        check(load().kd_session_path_counts(self.handle, _ptr(o, _i64p)))
        return {"selection": int(o[0]), "sorted": int(o[1]), "one_block": int(o[2])}

    def phase_stats(self) -> dict:
        """{phase: (total_ms, launches, algorithmic_bytes)} since profile()."""
        import json

        L = load()
        need = L.kd_session_phase_stats(self._h, None, 0)
        if need < 0:
            check(int(need))
        buf = C.create_string_buffer(int(need))
        L.kd_session_phase_stats(self._h, buf, need)
        return {k: (v[0], int(v[1]), v[2]) for k, v in json.loads(buf.value.decode()).items()}

    def close(self) -> None:
        if getattr(self, "_h", None):
            load().kd_session_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_sessions: dict[int, GpuSession] = {}


def default_session(device: int = 0) -> GpuSession:
    s = _default_sessions.get(device)
    if s is None:
        s = GpuSession(device)
        _default_sessions[device] = s
    return s


def _device_ptr(t) -> int:
    """Device address of a torch tensor (or an int address)."""
    if isinstance(t, int):
        return t
    return int(t.data_ptr())


def _device_array(t, dtype_name: str, n: int | None, what: str) -> int:
    """Address of a contiguous device tensor of the given torch dtype name and
    element count; a wrong dtype, length, layout or device raises ValueError
    before anything reaches the library."""
    if str(t.dtype) != "torch." + dtype_name:
        raise ValueError(f"{what} must be a torch.{dtype_name} tensor (got {t.dtype})")
    if not t.is_cuda or not t.is_contiguous():
        raise ValueError(f"{what} must be a contiguous device tensor")
    if n is not None and t.numel() != n:
        raise ValueError(f"{what} has {t.numel()} elements, expected {n}")
    return int(t.data_ptr())


# ----------------------------------------------------------------- TDigest
class TDigest:
    """Merging t-digest with pluggable scale and decay (tdigest.mojo:97-726)."""

    def __init__(self, scale=ScaleK1, compression: float = 100.0, decay: Decay | None = None,
                 session: GpuSession | None = None, _handle=None):
        self._scale = _scale(scale)
        self.compression = float(compression)
        self.decay = decay if decay is not None else Decay.none()
        self.session = session or default_session()
        L = load()
        if _handle is None:
            h = C.c_void_p()
            check(L.kd_tdigest_create(self.session.handle, self._scale.id, self.compression,
                                      self.decay.rate_lambda, C.byref(h)))
            _handle = h.value
        self._h = _handle

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            load().kd_tdigest_destroy(h)
            self._h = None

    @property
    def scale(self) -> ScaleFunction:
        return self._scale

    # --- ingest
    def add(self, value: float, weight: float = 1.0) -> None:
        check(load().kd_tdigest_add(self._h, float(value), float(weight)))

    def add_at(self, value: float, timestamp: float, weight: float = 1.0) -> None:
        check(load().kd_tdigest_add_at(self._h, float(value), float(timestamp), float(weight)))

    def add_batch(self, values, session: GpuSession | None = None) -> None:
        """tdigest.mojo:177-221: per-sample path below the crossover (60000,
        or 40000 with a caller-held session), one batch merge above it."""
        a = _f64(values)
        check(load().kd_tdigest_add_batch(self._h, _ptr(a), a.size, int(session is not None)))

    def add_batch_device(self, values) -> None:
        """One batch merge of a device-resident float64 tensor."""
        n = int(values.numel())
        check(load().kd_tdigest_add_batch_device(self._h, _device_ptr(values), n))

    def add_batches_device(self, values, batch: int) -> None:
        """Consecutive batch merges of `batch` values each (the last one
        shorter) of a device-resident float64 tensor: the digest
        add_batch_device gives slice by slice, with every value validated
        first and the merge-independent work of all slices done up front."""
        n = int(values.numel())
        check(load().kd_tdigest_add_batches_device(self._h, _device_ptr(values), n, int(batch)))

    def merge(self) -> None:
        check(load().kd_tdigest_merge(self._h))

    # --- queries
    def quantile(self, q):
        scalar = np.isscalar(q)
        a = _f64(np.atleast_1d(q))
        out = np.empty(a.size)
        check(load().kd_tdigest_quantile(self._h, _ptr(a), a.size, _ptr(out)))
        return float(out[0]) if scalar else out

    def cdf(self, x):
        scalar = np.isscalar(x)
        a = _f64(np.atleast_1d(x))
        out = np.empty(a.size)
        check(load().kd_tdigest_cdf(self._h, _ptr(a), a.size, _ptr(out)))
        return float(out[0]) if scalar else out

    def merge_into(self, other: "TDigest") -> None:
        check(load().kd_tdigest_merge_into(self._h, other._h))

    # --- state
    def _state(self) -> np.ndarray:
        o = np.empty(9)
        check(load().kd_tdigest_state(self._h, _ptr(o)))
        return o

    def total_weight(self) -> float:
        return float(self._state()[7])

    def min(self) -> float:
        return float(self._state()[3])

    def max(self) -> float:
        return float(self._state()[4])

    def state(self) -> dict:
        o = self._state()
        return dict(stored_total=o[0], log_scale=o[1], clock=o[2], min=o[3], max=o[4],
                    compression=o[5], decay=o[6], total_weight=o[7], buffered=int(o[8]))

    def __len__(self) -> int:
        return int(load().kd_tdigest_size(self._h))

    def centroids(self):
        L = load()
        n = C.c_int64()
        check(L.kd_tdigest_centroids(self._h, None, None, 0, C.byref(n)))
        m = np.empty(n.value)
        w = np.empty(n.value)
        check(L.kd_tdigest_centroids(self._h, _ptr(m), _ptr(w), n.value, C.byref(n)))
        return m, w

    # --- serde (tdigest.mojo:627-698)
    def to_bytes(self) -> bytes:
        L = load()
        need = L.kd_tdigest_to_bytes(self._h, None, 0)
        if need < 0:
            check(int(need))
        buf = (C.c_uint8 * need)()
        check(0 if L.kd_tdigest_to_bytes(self._h, buf, need) == need else -6)
        return bytes(buf)

    @staticmethod
    def from_bytes(data: bytes, scale=ScaleK1, session: GpuSession | None = None) -> "TDigest":
        sc = _scale(scale)
        session = session or default_session()
        L = load()
        buf = (C.c_uint8 * max(len(data), 1)).from_buffer_copy(bytes(data) or b"\0")
        h = C.c_void_p()
        check(L.kd_tdigest_from_bytes(session.handle, buf, len(data), sc.id, C.byref(h)))
        st = np.empty(9)
        L.kd_tdigest_state(h.value, _ptr(st))
        return TDigest(sc, st[5], Decay.rate(st[6]), session, _handle=h.value)

    @staticmethod
    def _from_centroids(compression, means, weights, min_v, max_v, scale=ScaleK1,
                        session: GpuSession | None = None) -> "TDigest":
        sc = _scale(scale)
        session = session or default_session()
        m, w = _f64(means), _f64(weights)
        h = C.c_void_p()
        check(load().kd_tdigest_from_centroids(session.handle, sc.id, float(compression), _ptr(m),
                                               _ptr(w), m.size, float(min_v), float(max_v),
                                               C.byref(h)))
        return TDigest(sc, compression, Decay.none(), session, _handle=h.value)


# ---------------------------------------------------------- GroupedTDigest
class GroupedTDigest:
    """num_keys digests updated together from (key, value) batches
    (grouped.mojo:30-190).  State is resident in device memory."""

    def __init__(self, num_keys: int, compression: float = 100.0, decay: Decay | None = None,
                 scale=ScaleK1, session: GpuSession | None = None):
        self._scale = _scale(scale)
        self.num_keys = int(num_keys)
        self.compression = float(compression)
        self.decay = decay if decay is not None else Decay.none()
        self.session = session or default_session()
        h = C.c_void_p()
        check(load().kd_grouped_create(self.session.handle, self._scale.id, self.num_keys,
                                       self.compression, self.decay.rate_lambda, C.byref(h)))
        self._h = h.value

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            load().kd_grouped_destroy(h)
            self._h = None

    def add_batch(self, keys, values) -> None:
        k = np.asarray(keys)
        v = _f64(values)
        if k.shape[0] != v.shape[0]:
            raise ValueError("keys and values must have equal length")
        if k.size and (k.min() < 0 or k.max() >= self.num_keys):
            raise KdigestError(-1, "key out of range [0, num_keys)")
        k32 = np.ascontiguousarray(k, dtype=np.uint32)
        check(load().kd_grouped_add_batch(self._h, k32.ctypes.data_as(C.POINTER(C.c_uint32)),
                                          _ptr(v), v.size))

    def add_batch_device(self, keys, values) -> None:
        """Device-resident uint32 (or int32) keys and float64 values."""
        n = int(values.numel())
        if int(keys.numel()) != n:
            raise ValueError("keys and values must have equal length")
        check(load().kd_grouped_add_batch_device(self._h, _device_ptr(keys), _device_ptr(values), n))

    def _query(self, fn, keys, xs):
        k = np.ascontiguousarray(np.atleast_1d(keys), dtype=np.int64)
        x = _f64(np.atleast_1d(xs))
        if k.size == 1 and x.size > 1:
            k = np.full(x.size, k[0], dtype=np.int64)
        out = np.empty(x.size)
        check(fn(self._h, _ptr(k, _i64p), _ptr(x), x.size, _ptr(out)))
        return out

    def quantile(self, key, q):
        scalar = np.isscalar(key) and np.isscalar(q)
        out = self._query(load().kd_grouped_quantile, key, q)
        return float(out[0]) if scalar else out

    def cdf(self, key, x):
        scalar = np.isscalar(key) and np.isscalar(x)
        out = self._query(load().kd_grouped_cdf, key, x)
        return float(out[0]) if scalar else out

    def quantile_device(self, keys, qs, out) -> None:
        """quantile(keys[i], qs[i]) into out[i] for device tensors: keys int64,
        qs and out float64, all of one length.  A key out of range, a q outside
        [0, 1] or an empty key raises (the outputs of those queries are NaN)."""
        n = int(qs.numel())
        kp = _device_array(keys, "int64", n, "keys")
        qp = _device_array(qs, "float64", n, "qs")
        op = _device_array(out, "float64", n, "out")
        check(load().kd_grouped_quantile_device(self._h, kp, qp, n, op))

    def quantile_device_noalloc(self, keys, qs, out) -> None:
        """quantile_device without any device allocation (the library's fixed
        per-GPU scratch; see kd_grouped_quantile_device_noalloc)."""
        n = int(qs.numel())
        kp = _device_array(keys, "int64", n, "keys")
        qp = _device_array(qs, "float64", n, "qs")
        op = _device_array(out, "float64", n, "out")
        check(load().kd_grouped_quantile_device_noalloc(self._h, kp, qp, n, op))

    def to_bytes(self, key: int) -> bytes:
        L = load()
        need = L.kd_grouped_to_bytes(self._h, int(key), None, 0)
        if need < 0:
            check(int(need))
        buf = (C.c_uint8 * need)()
        L.kd_grouped_to_bytes(self._h, int(key), buf, need)
        return bytes(buf)

    def wire_size(self) -> int:
        """Bytes of every key's serialized digest (to_bytes_all)."""
        o = C.c_int64()
        check(load().kd_grouped_wire_size(self._h, C.byref(o)))
        return o.value

    def to_bytes_all(self):
        """Every key's to_bytes, concatenated in key order, and the offsets
        (num_keys + 1) of each key's bytes."""
        need = self.wire_size()
        buf = (C.c_uint8 * max(need, 1))()
        off = np.empty(self.num_keys + 1, dtype=np.int64)
        check(load().kd_grouped_to_bytes_all(self._h, buf, need, off.ctypes.data_as(_i64p)))
        return bytes(buf)[:need], off

    def to_bytes_device(self, out, off, cap: int | None = None) -> None:
        """to_bytes_all into device memory (torch tensors or addresses): out
        of at least wire_size() bytes (cap: its size when an address), off of
        num_keys + 1 int64."""
        cap = int(out.numel()) if cap is None else int(cap)
        check(load().kd_grouped_to_bytes_device(self._h, _device_ptr(out), cap, _device_ptr(off)))

    @classmethod
    def _adopt(cls, h, num_keys, scale, session):
        g = cls.__new__(cls)
        g._h = h
        g._scale = _scale(scale)
        g.num_keys = int(num_keys)
        g.session = session
        comp, lam = C.c_double(), C.c_double()
        check(load().kd_grouped_params(h, C.byref(comp), C.byref(lam)))
        g.compression = comp.value
        g.decay = Decay(lam.value)
        return g

    @classmethod
    def from_bytes_all(cls, data: bytes, off, scale=ScaleK1,
                       session: GpuSession | None = None) -> "GroupedTDigest":
        """A grouped digest restored from to_bytes_all's bytes and offsets."""
        session = session or default_session()
        off = np.ascontiguousarray(off, dtype=np.int64)
        nk = off.size - 1
        buf = (C.c_uint8 * max(len(data), 1)).from_buffer_copy(data if data else b"\0")
        h = C.c_void_p()
        sc = _scale(scale)
        check(load().kd_grouped_from_bytes_all(session.handle, sc.id, nk, buf, len(data),
                                               off.ctypes.data_as(_i64p), C.byref(h)))
        return cls._adopt(h.value, nk, sc, session)

    @classmethod
    def from_bytes_device(cls, data, off, scale=ScaleK1, session: GpuSession | None = None,
                          n: int | None = None,
                          num_keys: int | None = None) -> "GroupedTDigest":
        """A grouped digest restored from device memory (torch tensors, or
        addresses with n bytes of data and num_keys + 1 int64 offsets)."""
        session = session or default_session()
        nk = int(off.numel()) - 1 if num_keys is None else int(num_keys)
        n = int(data.numel()) if n is None else int(n)
        h = C.c_void_p()
        sc = _scale(scale)
        check(load().kd_grouped_from_bytes_device(session.handle, sc.id, nk, _device_ptr(data), n,
                                                  _device_ptr(off), C.byref(h)))
        return cls._adopt(h.value, nk, sc, session)

    def centroid_count(self, key: int) -> int:
        o = C.c_int64()
        check(load().kd_grouped_centroid_count(self._h, int(key), C.byref(o)))
        return o.value

    def total_weight(self, key: int) -> float:
        o = C.c_double()
        check(load().kd_grouped_total_weight(self._h, int(key), C.byref(o)))
        return o.value

    def centroids(self, key: int):
        L = load()
        n = C.c_int64()
        check(L.kd_grouped_centroids(self._h, int(key), None, None, 0, C.byref(n)))
        m = np.empty(n.value)
        w = np.empty(n.value)
        check(L.kd_grouped_centroids(self._h, int(key), _ptr(m), _ptr(w), n.value, C.byref(n)))
        return m, w

    def key_state(self, key: int) -> dict:
        o = np.empty(5)
        check(load().kd_grouped_key_state(self._h, int(key), _ptr(o)))
        return dict(stored_total=o[0], log_scale=o[1], clock=o[2], min=o[3], max=o[4])


# ------------------------------------------------------ operator functions
def gpu_compress(means, prior_weights, n_prior: int, w_uniform: float, scale, nrm: float,
                 stored_n: float, session: GpuSession | None = None):
    """gpu_compress_batch (gpu/add_batch.mojo:629-663) as the FP64 greedy
    sweep: returns (means, weights) of the compressed centroids."""
    session = session or default_session()
    m = _f64(means)
    w = _f64(prior_weights) if n_prior else np.zeros(1)
    n = m.size - n_prior
    om = np.empty(max(m.size, 1))
    ow = np.empty(max(m.size, 1))
    c = C.c_int64()
    check(load().kd_compress(session.handle, _scale(scale).id, float(nrm), float(stored_n), _ptr(m),
                             _ptr(w), float(w_uniform), int(n_prior), int(n), _ptr(om), _ptr(ow),
                             C.byref(c)))
    return om[:c.value].copy(), ow[:c.value].copy()


def _csr(parts_means: Sequence, parts_weights: Sequence):
    off = np.zeros(len(parts_means) + 1, dtype=np.int64)
    for i, p in enumerate(parts_means):
        off[i + 1] = off[i] + len(p)
    m = _f64(np.concatenate([np.asarray(p, np.float64) for p in parts_means])) if len(parts_means) else np.zeros(0)
    w = _f64(np.concatenate([np.asarray(p, np.float64) for p in parts_weights])) if len(parts_weights) else np.zeros(0)
    return off, m, w


def _gpu_query(fn, digests_means, digests_weights, states, q_digest, qs, session):
    session = session or default_session()
    off, m, w = _csr(digests_means, digests_weights)
    st = _f64(np.asarray(states, np.float64).reshape(-1))
    qd = np.ascontiguousarray(q_digest, dtype=np.int64)
    q = _f64(qs)
    out = np.empty(q.size)
    check(fn(session.handle, len(digests_means), _ptr(off, _i64p), _ptr(m), _ptr(w), _ptr(st),
             _ptr(qd, _i64p), _ptr(q), q.size, _ptr(out)))
    return out


def gpu_quantile(digests_means, digests_weights, states, q_digest, qs, session=None):
    """Batched gpu_quantile (gpu/query.mojo:320-331); states rows are
    (stored_total, min, max, eff_factor)."""
    return _gpu_query(load().kd_quantile, digests_means, digests_weights, states, q_digest, qs, session)


def gpu_cdf(digests_means, digests_weights, states, q_digest, xs, session=None):
    """Batched gpu_cdf (gpu/query.mojo:334-342)."""
    return _gpu_query(load().kd_cdf, digests_means, digests_weights, states, q_digest, xs, session)


def gpu_scale_eval(scale, nrm: float, qs, session=None) -> np.ndarray:
    """Device k(q) (gpu/scale_device.mojo:53-98), FP64."""
    session = session or default_session()
    q = _f64(qs)
    out = np.empty(q.size)
    check(load().kd_scale_eval(session.handle, _scale(scale).id, float(nrm), _ptr(q), q.size, _ptr(out)))
    return out


def gpu_math_eval(fn: str, xs, session=None) -> np.ndarray:
    session = session or default_session()
    x = _f64(xs)
    out = np.empty(x.size)
    check(load().kd_math_eval(session.handle, {"log": 0, "asin": 1, "sqrt": 2}[fn], _ptr(x), x.size,
                              _ptr(out)))
    return out


# ------------------------------------------------------ batched digests
class KeyMap:
    """Dense ids 0..size-1 for sparse uint32 keys, held on the device: a key
    gets the next id the first time it is mapped with insert, in order of
    first occurrence in its batch (for GroupedTDigest over sparse keys)."""

    UNKNOWN = 0xFFFFFFFF

    def __init__(self, capacity: int, session: GpuSession | None = None):
        self.capacity = int(capacity)
        self.session = session or default_session()
        h = C.c_void_p()
        check(load().kd_keymap_create(self.session.handle, self.capacity, C.byref(h)))
        self._h = h.value

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            load().kd_keymap_destroy(h)
            self._h = None

    def __len__(self) -> int:
        o = C.c_int64()
        check(load().kd_keymap_size(self._h, C.byref(o)))
        return o.value

    def map(self, keys, insert: bool = True) -> np.ndarray:
        """Dense ids of host keys (UNKNOWN for keys not in the map when
        insert is off)."""
        k = np.ascontiguousarray(keys, dtype=np.uint32)
        out = np.empty(k.size, dtype=np.uint32)
        u32 = C.POINTER(C.c_uint32)
        check(load().kd_keymap_map(self._h, k.ctypes.data_as(u32), k.size, out.ctypes.data_as(u32),
                                   1 if insert else 0))
        return out

    def map_device(self, keys, ids, insert: bool = True, n: int | None = None) -> None:
        """Dense ids of device keys into device ids (uint32, torch tensors or
        addresses with n keys)."""
        n = int(keys.numel()) if n is None else int(n)
        check(load().kd_keymap_map_device(self._h, _device_ptr(keys), n, _device_ptr(ids),
                                          1 if insert else 0))

    def keys(self) -> np.ndarray:
        """The key of every id, in id order."""
        n = len(self)
        out = np.empty(max(n, 1), dtype=np.uint32)
        check(load().kd_keymap_keys(self._h, out.ctypes.data_as(C.POINTER(C.c_uint32)), n))
        return out[:n]


class BatchedDigests:
    """num_digests independent TDigest fed sample by sample with add_at
    (tdigest.mojo:163-175) and reduced by a pairwise merge_into tree
    (tdigest.mojo:589-623); state resident in device memory, one GPU
    workgroup per digest."""

    def __init__(self, num_digests: int, compression: float = 100.0, decay: Decay | None = None,
                 scale=ScaleK1, session: GpuSession | None = None):
        self._scale = _scale(scale)
        self.num_digests = int(num_digests)
        self.compression = float(compression)
        self.decay = decay if decay is not None else Decay.none()
        self.session = session or default_session()
        h = C.c_void_p()
        check(load().kd_batch_create(self.session.handle, self._scale.id, self.num_digests,
                                     self.compression, self.decay.rate_lambda, C.byref(h)))
        self._h = h.value

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            load().kd_batch_destroy(h)
            self._h = None

    def add_at(self, off, values, times, weights) -> None:
        """Digest d gets samples [off[d], off[d+1]) in order (host arrays)."""
        o = np.ascontiguousarray(off, dtype=np.int64)
        v, t, w = _f64(values), _f64(times), _f64(weights)
        check(load().kd_batch_add_at(self._h, _ptr(o, _i64p), _ptr(v), _ptr(t), _ptr(w)))

    def add_at_device(self, off, values, times, weights) -> None:
        check(load().kd_batch_add_at_device(self._h, _device_ptr(off), _device_ptr(values),
                                            _device_ptr(times), _device_ptr(weights)))

    def flush(self) -> None:
        check(load().kd_batch_flush(self._h))

    @classmethod
    def _adopt(cls, h, num_digests, scale, session):
        b = cls.__new__(cls)
        b._h = h
        b._scale = _scale(scale)
        b.num_digests = int(num_digests)
        b.session = session
        comp, lam = C.c_double(), C.c_double()
        check(load().kd_batch_params(h, C.byref(comp), C.byref(lam)))
        b.compression = comp.value
        b.decay = Decay(lam.value)
        return b

    @classmethod
    def from_bytes_all(cls, data: bytes, off, scale=ScaleK1,
                       session: GpuSession | None = None) -> "BatchedDigests":
        """Digests loaded from their serialized bytes (digest d at
        data[off[d]:off[d+1]]), for merge_tree to reduce (map-reduce
        finalize)."""
        session = session or default_session()
        off = np.ascontiguousarray(off, dtype=np.int64)
        nd = off.size - 1
        buf = (C.c_uint8 * max(len(data), 1)).from_buffer_copy(data if data else b"\0")
        h = C.c_void_p()
        sc = _scale(scale)
        check(load().kd_batch_from_bytes_all(session.handle, sc.id, nd, buf, len(data),
                                             off.ctypes.data_as(_i64p), C.byref(h)))
        return cls._adopt(h.value, nd, sc, session)

    @classmethod
    def from_bytes_device(cls, data, off, scale=ScaleK1, session: GpuSession | None = None,
                          n: int | None = None,
                          num_digests: int | None = None) -> "BatchedDigests":
        """from_bytes_all from device memory (torch tensors, or addresses
        with n bytes of data and num_digests + 1 int64 offsets)."""
        session = session or default_session()
        nd = int(off.numel()) - 1 if num_digests is None else int(num_digests)
        n = int(data.numel()) if n is None else int(n)
        h = C.c_void_p()
        sc = _scale(scale)
        check(load().kd_batch_from_bytes_device(session.handle, sc.id, nd, _device_ptr(data), n,
                                                _device_ptr(off), C.byref(h)))
        return cls._adopt(h.value, nd, sc, session)

    def set_graph(self, enable: bool = True) -> None:
        """Run merge_tree as a captured HIP graph (default) or as eager launches."""
        check(load().kd_batch_set_graph(self._h, 1 if enable else 0))

    def merge_tree(self) -> None:
        check(load().kd_batch_merge_tree(self._h))

    def state(self, d: int) -> dict:
        o = np.empty(7)
        check(load().kd_batch_state(self._h, int(d), _ptr(o)))
        return dict(stored_total=o[0], log_scale=o[1], clock=o[2], min=o[3], max=o[4],
                    buffered=int(o[5]), error=int(o[6]))

    def centroids(self, d: int):
        L = load()
        n = C.c_int64()
        check(L.kd_batch_centroids(self._h, int(d), None, None, 0, C.byref(n)))
        m, w = np.empty(n.value), np.empty(n.value)
        check(L.kd_batch_centroids(self._h, int(d), _ptr(m), _ptr(w), n.value, C.byref(n)))
        return m, w

    def to_tdigest(self, d: int) -> TDigest:
        h = C.c_void_p()
        check(load().kd_batch_to_tdigest(self._h, int(d), C.byref(h)))
        return TDigest(self._scale, self.compression, self.decay, self.session, _handle=h.value)


class DigestServer:
    """Continuous-batching server over num_keys per-key digests kept in device
    memory (GroupedTDigest semantics per key, grouped.mojo:60-190).

    Requests are ingest batches (one batch merge each, tdigest.mojo:223-315)
    and quantile / cdf batches (tdigest.mojo:485-587), each for one key.  A
    step admits queued requests in arrival order while the device memory
    their batch occupies (staged data and merge workspace) fits the memory
    budget and runs them as one batch; every request's result is the
    one it would have alone, in its key's arrival order."""

    INGEST, QUANTILE, CDF = 0, 1, 2

    def __init__(self, num_keys: int, compression: float = 100.0, decay: Decay | None = None,
                 scale=ScaleK1, slot_capacity: int = 1024, mem_budget: int = 64 << 20,
                 session: GpuSession | None = None):
        self._scale = _scale(scale)
        self.num_keys = int(num_keys)
        self.compression = float(compression)
        self.decay = decay if decay is not None else Decay.none()
        self.session = session or default_session()
        self._n = {}
        h = C.c_void_p()
        check(load().kd_server_create(self.session.handle, self._scale.id, self.num_keys,
                                      self.compression, self.decay.rate_lambda, int(slot_capacity),
                                      int(mem_budget), C.byref(h)))
        self._h = h.value

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            load().kd_server_destroy(h)
            self._h = None

    def submit(self, kind: int, key: int, data) -> int:
        a = _f64(data)
        rid = C.c_int64()
        check(load().kd_server_submit(self._h, int(kind), int(key), _ptr(a), a.size, C.byref(rid)))
        self._n[rid.value] = a.size if kind != self.INGEST else 0
        return rid.value

    def submit_many(self, kinds, keys, off, data) -> np.ndarray:
        """Request r: kinds[r], keys[r], data[off[r]:off[r+1]]; returns the ids."""
        kn = np.ascontiguousarray(kinds, dtype=np.int32)
        ky = np.ascontiguousarray(keys, dtype=np.int64)
        of = np.ascontiguousarray(off, dtype=np.int64)
        d = _f64(data)
        ids = np.empty(kn.size, np.int64)
        check(load().kd_server_submit_many(self._h, kn.size, _ptr(kn, C.POINTER(C.c_int)),
                                           _ptr(ky, _i64p), _ptr(of, _i64p), _ptr(d), _ptr(ids, _i64p)))
        sizes = np.diff(of)
        for r, rid in enumerate(ids.tolist()):
            self._n[rid] = int(sizes[r]) if kn[r] != self.INGEST else 0
        return ids

    def take_completed(self, cap: int = 1 << 16):
        """(ids, completion times on the server clock) since the last call."""
        ids = np.empty(cap, np.int64)
        t = np.empty(cap)
        n = C.c_int64()
        check(load().kd_server_take_completed(self._h, _ptr(ids, _i64p), _ptr(t), cap, C.byref(n)))
        return ids[:n.value], t[:n.value]

    def clock(self) -> float:
        t = C.c_double()
        check(load().kd_server_clock(self._h, C.byref(t)))
        return t.value

    def discard(self, rid: int) -> None:
        """Forget a completed request without reading its result."""
        try:
            self.result(rid)
        except KdigestError:
            pass

    def submit_ingest(self, key: int, values) -> int:
        return self.submit(self.INGEST, key, values)

    def submit_quantile(self, key: int, qs) -> int:
        return self.submit(self.QUANTILE, key, qs)

    def submit_cdf(self, key: int, xs) -> int:
        return self.submit(self.CDF, key, xs)

    def step(self, max_rounds: int = 4) -> int:
        done = C.c_int64()
        check(load().kd_server_step(self._h, int(max_rounds), C.byref(done)))
        return done.value

    def start(self, max_rounds: int = 4) -> None:
        """Serve requests on two worker threads as they are submitted: queries
        run beside ingest rounds, each key still sees its requests in
        submission order (results equal those of step()).  step(), key_state()
        and centroids() raise until stop()."""
        check(load().kd_server_start(self._h, int(max_rounds)))

    def stop(self) -> None:
        """Waits until every submitted request completed, then ends the
        workers; raises the first device error a worker met."""
        check(load().kd_server_stop(self._h))

    def result(self, rid: int):
        """Query values (np.ndarray), an empty array for a finished ingest, or
        None while the request is pending.  A finished request is forgotten
        once its result has been read; a query that found its key empty raises
        KdigestError (code KD_EEMPTY) here."""
        n = self._n[rid]
        out = np.empty(max(n, 1))
        done = C.c_int()
        rc = load().kd_server_result(self._h, int(rid), _ptr(out), n, C.byref(done))
        if done.value:
            del self._n[rid]
        check(rc)
        if not done.value:
            return None
        return out[:n]

    @property
    def pending(self) -> int:
        p = C.c_int64()
        check(load().kd_server_pending(self._h, C.byref(p)))
        return p.value

    def stats(self) -> dict:
        o = np.zeros(6, np.int64)
        check(load().kd_server_stats(self._h, _ptr(o, _i64p)))
        return dict(steps=int(o[0]), admitted=int(o[1]), admitted_bytes=int(o[2]),
                    max_step_bytes=int(o[3]), ingest_rounds=int(o[4]), queued=int(o[5]))

    def key_state(self, key: int) -> dict:
        o = np.empty(6)
        check(load().kd_server_key_state(self._h, int(key), _ptr(o)))
        return dict(stored_total=o[0], log_scale=o[1], clock=o[2], min=o[3], max=o[4],
                    centroids=int(o[5]))

    def centroids(self, key: int):
        L = load()
        n = C.c_int64()
        check(L.kd_server_key_centroids(self._h, int(key), None, None, 0, C.byref(n)))
        m, w = np.empty(n.value), np.empty(n.value)
        check(L.kd_server_key_centroids(self._h, int(key), _ptr(m), _ptr(w), n.value, C.byref(n)))
        return m, w
