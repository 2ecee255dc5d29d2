"""Parity helpers shared by the GPU tests.

Bars (stated once, used everywhere):
  * centroid counts and sorted orders: exact;
  * centroid weights: exact when all weights are integers (no decay), else
    within W_RTOL relative;
  * centroid means: within MEAN_RTOL relative, with an absolute floor of
    MEAN_RTOL * max|x| over the digest's inputs (means of clusters straddling
    zero have no meaningful relative error);
  * quantile / cdf: within QUERY_RTOL relative (absolute floor QUERY_RTOL *
    max(1, |x|max) for quantiles, QUERY_RTOL for cdf values in [0, 1]).
"""
from __future__ import annotations

import numpy as np

MEAN_RTOL = 1e-12
W_RTOL = 1e-12
QUERY_RTOL = 1e-9


def stable_sorted(means, weights):
    m = np.asarray(means, np.float64)
    w = np.asarray(weights, np.float64)
    o = np.argsort(m, kind="stable")
    return m[o], w[o]


def check_centroids(got_m, got_w, ref_m, ref_w, scale_abs: float, integer_weights: bool, tag=""):
    got_m, got_w, ref_m, ref_w = map(np.asarray, (got_m, got_w, ref_m, ref_w))
    assert got_m.size == ref_m.size, f"{tag}: centroid count {got_m.size} != {ref_m.size}"
    if integer_weights:
        bad = np.nonzero(got_w != ref_w)[0]
        assert bad.size == 0, f"{tag}: weights differ at {bad[:5]}: {got_w[bad[:5]]} vs {ref_w[bad[:5]]}"
    else:
        np.testing.assert_allclose(got_w, ref_w, rtol=W_RTOL, atol=0, err_msg=f"{tag}: weights")
    scale_abs = max(float(scale_abs), 1e-300)
    err = np.abs(got_m - ref_m)
    tol = MEAN_RTOL * np.maximum(np.abs(ref_m), scale_abs)
    bad = np.nonzero(err > tol)[0]
    assert bad.size == 0, (f"{tag}: means differ at {bad[:5]}: {got_m[bad[:5]]} vs {ref_m[bad[:5]]} "
                           f"(err {err[bad[:5]]})")
    return float((err / np.maximum(np.abs(ref_m), scale_abs)).max()) if err.size else 0.0


def check_values(got, ref, floor: float, tag=""):
    got, ref = np.asarray(got), np.asarray(ref)
    err = np.abs(got - ref)
    tol = QUERY_RTOL * np.maximum(np.abs(ref), floor)
    bad = np.nonzero(~(err <= tol))[0]
    assert bad.size == 0, f"{tag}: {bad.size} values differ, first {got[bad[:4]]} vs {ref[bad[:4]]}"
