"""kdigest_amd — the t-digest ingest / compress / query / merge path of
jbrownuf07/kdigest on MI355X (gfx950), FP64 throughout.

Public API (kdigest/__init__.mojo:10-25): ScaleK0..ScaleK3, Decay, TDigest,
GroupedTDigest, GpuSession; operator functions gpu_compress, gpu_quantile,
gpu_cdf, gpu_scale_eval.  All compute runs in libkdigest_hip.so (HIP kernels
behind the C ABI of include/kdigest_hip.h).
"""
from ._lib import KdigestError, LIB_PATH, load
from .tdigest import (
    BatchedDigests,
    Decay,
    DigestServer,
    GpuSession,
    GroupedTDigest,
    KeyMap,
    ScaleFunction,
    ScaleK0,
    ScaleK1,
    ScaleK2,
    ScaleK3,
    TDigest,
    default_session,
    gpu_cdf,
    gpu_compress,
    gpu_math_eval,
    gpu_quantile,
    gpu_scale_eval,
)

__all__ = [
    "BatchedDigests", "DigestServer",
    "Decay", "GpuSession", "GroupedTDigest", "KdigestError", "KeyMap", "LIB_PATH", "ScaleFunction",
    "ScaleK0", "ScaleK1", "ScaleK2", "ScaleK3", "TDigest", "default_session", "gpu_cdf",
    "gpu_compress", "gpu_math_eval", "gpu_quantile", "gpu_scale_eval", "load",
]
