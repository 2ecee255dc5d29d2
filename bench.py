"""Benchmark of the digest ingest / query path on one MI355X.

Headline workload (value): grouped per-key digests, 100,000 uint32 keys with
Zipf(s=1.1) frequencies, 2^28 (key, FP64 value) pairs with per-key values
0.9*N(mu_k, 1) + 0.1*(mu_k + Pareto(alpha=1.5, x_m=1)), mu_k ~ U(-100, 100);
ScaleK3, compression 100, no decay.  One step = one GroupedTDigest.add_batch
of the whole device-resident batch into fresh digests (key grouping, per-key
sort, greedy-sweep compress, finish).  Secondary lines in the same JSON:
single-digest ingest (2^26 lognormal(0, 1.5) values in 64 batches of 2^20,
ScaleK2, compression 200, half-life 1e5); the 1024 tail quantiles per key of
the grouped digests; the same grouped batch through plain one-digest-per-launch
compresses (timed on a stratified key sample, extrapolated); 8192 decayed
digests fed by add_at and reduced by a merge_into tree; and an open-loop
serving run over one million device-resident keys.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N] [--keys K]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s
GROUPED_BYTES_PER_SAMPLE = 28.0  # read (u32 key + f64) once, write + read the sorted f64 once


def gen_grouped(n: int, num_keys: int, seed: int, dev):
    import torch

    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    ranks = torch.arange(1, num_keys + 1, dtype=torch.float64, device=dev)
    p = ranks ** -1.1
    cdf = torch.cumsum(p, 0)
    cdf /= cdf[-1].clone()
    perm = torch.randperm(num_keys, generator=g, device=dev)
    mu = torch.rand(num_keys, generator=g, dtype=torch.float64, device=dev) * 200.0 - 100.0
    keys = torch.empty(n, dtype=torch.int32, device=dev)
    vals = torch.empty(n, dtype=torch.float64, device=dev)
    chunk = 1 << 25
    for a in range(0, n, chunk):
        b = min(n, a + chunk)
        u = torch.rand(b - a, generator=g, dtype=torch.float64, device=dev)
        r = torch.searchsorted(cdf, u).clamp_(max=num_keys - 1)
        k = perm[r]
        keys[a:b] = k.to(torch.int32)
        z = torch.randn(b - a, generator=g, dtype=torch.float64, device=dev)
        t = torch.rand(b - a, generator=g, dtype=torch.float64, device=dev)
        par = (1.0 - torch.rand(b - a, generator=g, dtype=torch.float64, device=dev)) ** (-1.0 / 1.5)
        vals[a:b] = torch.where(t < 0.9, mu[k] + z, mu[k] + par)
    torch.cuda.synchronize()
    return keys, vals


RANK_QS = (0.001, 0.01, 0.5, 0.99, 0.999)


def tail_quantiles(per_key: int = 1024) -> np.ndarray:
    """The per-key tail quantiles of the grouped workload: per_key / 2
    log-spaced points in [1e-4, 1e-2] and their mirror images near 1."""
    lo = np.logspace(-4, -2, per_key // 2)
    return np.concatenate([lo, 1.0 - lo[::-1]])


def host_threads() -> int:
    """Worker threads for host-side oracle runs: the cores this process may use,
    at most 16 (one GPU's share of the host on the measurement boxes)."""
    return max(1, min(16, len(os.sched_getaffinity(0))))


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _parse_blobs(blob: np.ndarray, off: np.ndarray):
    """Per-key headers (num_keys x 8 doubles, as raw uint64 bit patterns) and the
    concatenated (mean, weight) pairs of consecutive wire blobs."""
    nk = off.size - 1
    starts = off[:-1].astype(np.int64)
    hidx = starts[:, None] + np.arange(65, dtype=np.int64)[None, :]
    hdr = np.ascontiguousarray(blob[hidx][:, 1:]).view("<u8").reshape(nk, 8)
    mask = np.ones(blob.size, bool)
    mask[hidx.ravel()] = False
    pairs = np.ascontiguousarray(blob[mask]).view("<f8").reshape(-1, 2)
    return hdr, pairs


def grouped_parity(kd, g, keys, vals, nk: int, refs=None, threads: int | None = None,
                   stride: int = 100, scale: int = 3, compression: float = 100.0):
    """Compares every key of the grouped digest g (built from keys, vals on the
    device) with the oracle's per-key batch merges of the same pairs: header
    fields (stored total, min, max, clock, count) bit for bit, weights bit for
    bit, means within 1e-12 relative (floor 1e-12 * max|x| of the key), Sum w
    equal to the key's pair count and means sorted; then, on the 64 largest
    keys and every `stride`-th key by size, the 1024 tail quantiles and the
    quantiles at RANK_QS within 1e-9 relative, and the rank error of both
    estimates against an exact sort of the key's values (on the device).
    `refs` may hold the oracle digests already built; returns (report, refs)."""
    import torch

    from oracle import oracle as O

    failures = []
    t_all = time.perf_counter()
    dev = vals.device
    nbytes = g.wire_size()
    wout = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    woff = torch.empty(nk + 1, dtype=torch.int64, device=dev)
    g.to_bytes_device(wout, woff)
    gblob, goff = wout.cpu().numpy(), woff.cpu().numpy()
    del wout, woff
    t_oracle = None
    threads = threads or host_threads()
    if refs is None:
        kh = keys.cpu().numpy().astype(np.uint32)
        vh = vals.cpu().numpy()
        refs = [O.Digest(scale, compression) for _ in range(nk)]
        t0 = time.perf_counter()
        O.grouped_add_batch(refs, kh, vh, threads=threads)
        t_oracle = time.perf_counter() - t0
        del kh, vh
    oparts = [r.to_bytes() for r in refs]
    olen = np.fromiter((len(b) for b in oparts), np.int64, nk)
    ooff = np.zeros(nk + 1, np.int64)
    ooff[1:] = np.cumsum(olen)
    oblob = np.frombuffer(b"".join(oparts), np.uint8)
    del oparts
    gh, gp = _parse_blobs(gblob, goff)
    oh, op = _parse_blobs(oblob, ooff)
    bad_hdr = np.nonzero(np.any(gh != oh, axis=1))[0]
    names = ("compression", "lambda", "log_scale", "clock", "stored_total", "min", "max", "count")
    for k in bad_hdr[:5].tolist():
        diff = {names[i]: (gh[k, i:i + 1].view("<f8")[0], oh[k, i:i + 1].view("<f8")[0])
                for i in range(8) if gh[k, i] != oh[k, i]}
        failures.append(f"key {k}: header {diff}")
    cnt = torch.bincount(keys.to(torch.int64), minlength=nk).cpu().numpy()
    ghd = gh.view("<f8")
    counts = ghd[:, 7].astype(np.int64)
    rep = {"keys_compared": nk, "centroids": int(counts.sum()), "headers_equal": bad_hdr.size == 0,
           "mean_bar": 1e-12, "query_bar": 1e-9}
    sumw_ok = bool(np.array_equal(ghd[:, 4], cnt.astype(np.float64)))
    rep["stored_total_equals_pairs"] = sumw_ok
    if not sumw_ok:
        failures.append("stored total differs from the key's pair count")
    if bad_hdr.size == 0:
        kid = np.repeat(np.arange(nk), counts)
        wsum = np.bincount(kid, weights=gp[:, 1], minlength=nk)
        rep["weights_equal"] = bool(np.array_equal(gp[:, 1].view("<u8"), op[:, 1].view("<u8")))
        if not rep["weights_equal"]:
            failures.append("centroid weights differ")
        rep["sum_w_equals_pairs"] = bool(np.array_equal(wsum, cnt.astype(np.float64)))
        if not rep["sum_w_equals_pairs"]:
            failures.append("sum of centroid weights differs from the pair count")
        same_key = kid[1:] == kid[:-1]
        rep["means_sorted"] = bool(np.all(gp[1:, 0][same_key] >= gp[:-1, 0][same_key]))
        if not rep["means_sorted"]:
            failures.append("centroid means not sorted within a key")
        scale_k = np.maximum(np.abs(ghd[:, 5]), np.abs(ghd[:, 6]))
        floor = np.maximum(np.abs(op[:, 0]), np.repeat(scale_k, counts))
        rel = np.abs(gp[:, 0] - op[:, 0]) / np.maximum(floor, 1e-300)
        rep["max_mean_rel_err"] = float(rel.max()) if rel.size else 0.0
        rep["means_bit_identical_frac"] = float(np.mean(gp[:, 0].view("<u8") == op[:, 0].view("<u8")))
        if rep["max_mean_rel_err"] > 1e-12:
            failures.append(f"mean relative error {rep['max_mean_rel_err']:.3g} > 1e-12")
    # Queries and rank error on a stratified key sample.
    by_size = np.argsort(-cnt, kind="stable")
    sample = np.unique(np.concatenate([by_size[:64], by_size[64::stride]]))
    sample = sample[cnt[sample] > 0]
    tq = tail_quantiles()
    qrow = np.concatenate([tq, np.array(RANK_QS)])
    qk = torch.from_numpy(np.repeat(sample.astype(np.int64), qrow.size)).to(dev)
    qv = torch.from_numpy(np.tile(qrow, sample.size)).to(dev)
    qo = torch.empty_like(qv)
    g.quantile_device(qk, qv, qo)
    gq = qo.cpu().numpy().reshape(sample.size, qrow.size)
    del qk, qv, qo
    oq = np.stack([refs[int(k)].quantiles(qrow) for k in sample])
    qrel = np.abs(gq - oq) / np.maximum(np.abs(oq), 1.0)
    rep["query_keys"] = int(sample.size)
    rep["queries_compared"] = int(gq.size)
    rep["max_quantile_rel_err"] = float(qrel.max())
    rep["quantiles_bit_identical_frac"] = float(np.mean(gq.view("<u8") == oq.view("<u8")))
    if rep["max_quantile_rel_err"] > 1e-9:
        failures.append(f"quantile relative error {rep['max_quantile_rel_err']:.3g} > 1e-9")
    # Rank error: fraction of the key's values <= the estimate, minus q
    # (test_tdigest.mojo:93-103), exact counts from the key's sorted values.
    kk = keys.to(torch.int64)
    vs, idx = torch.sort(vals)
    ks = kk[idx]
    del idx
    ks, idx2 = torch.sort(ks, stable=True)
    vs = vs[idx2]
    del idx2, ks, kk
    koff = np.zeros(nk + 1, np.int64)
    koff[1:] = np.cumsum(cnt)
    nq = len(RANK_QS)
    est = torch.from_numpy(np.concatenate([gq[:, -nq:], oq[:, -nq:]], axis=1)).to(dev)
    ranks = np.empty((sample.size, 2 * nq))
    for i, k in enumerate(sample.tolist()):
        seg = vs[koff[k]:koff[k + 1]]
        ranks[i] = (torch.searchsorted(seg, est[i], right=True).cpu().numpy() / cnt[k])
    del vs, est
    qarr = np.array(RANK_QS)
    err_g = np.abs(ranks[:, :nq] - qarr)
    err_o = np.abs(ranks[:, nq:] - qarr)
    big = cnt[sample] >= 100000
    rep["rank_error"] = {
        f"q={q}": {"gpu_max": float(err_g[:, j].max()), "gpu_mean": float(err_g[:, j].mean()),
                   "oracle_max": float(err_o[:, j].max()), "oracle_mean": float(err_o[:, j].mean()),
                   "gpu_mean_keys_ge_1e5": float(err_g[big, j].mean()) if big.any() else None,
                   "oracle_mean_keys_ge_1e5": float(err_o[big, j].mean()) if big.any() else None}
        for j, q in enumerate(RANK_QS)}
    rep["rank_error_note"] = ("|#(x <= estimate)/n - q| per key against the key's sorted values, over "
                              f"the {sample.size} sampled keys ({int(big.sum())} with >= 1e5 values)")
    rep["oracle_s"] = t_oracle
    rep["oracle_threads"] = threads
    rep["check_s"] = time.perf_counter() - t_all
    rep["failures"] = failures
    return rep, refs


def plain_cpu(kd, session):
    """Plain single digest, ScaleK1, compression 100, no decay: 2^20 N(0,1)
    values (numpy seed 0) added in chunks of 4096 (below the crossover: the
    per-sample add path, tdigest.mojo:199-202) on the oracle (ns per add,
    merges included, one thread) and on the library (its merges on the GPU);
    quantiles at RANK_QS, cdf at -3, 0, 3, and the rank error of both against
    the sorted values."""
    from oracle import oracle as O

    x = np.random.default_rng(0).standard_normal(1 << 20)
    chunk = 4096
    best = None
    for _ in range(3):
        d = O.Digest(1, 100.0)
        t0 = time.perf_counter()
        for a in range(0, x.size, chunk):
            d.add_batch(x[a:a + chunk])
        d.merge()
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
    g = kd.TDigest(kd.ScaleK1, 100.0, session=session)
    t0 = time.perf_counter()
    for a in range(0, x.size, chunk):
        g.add_batch(x[a:a + chunk])
    g.merge()
    dt_gpu = time.perf_counter() - t0
    qs = np.array(RANK_QS)
    xs = np.array([-3.0, 0.0, 3.0])
    oq, oc = d.quantiles(qs), d.cdfs(xs)
    gq, gc = g.quantile(qs), g.cdf(xs)
    srt = np.sort(x)
    rank = lambda v: np.searchsorted(srt, v, side="right") / srt.size
    return {"workload": "one digest, ScaleK1, compression 100, 2^20 N(0,1) values (numpy seed 0) "
                        "added in chunks of 4096 (per-sample add path)",
            "cpu_ns_per_add": best / x.size * 1e9, "cpu_threads": 1,
            "reference_published_ns_per_add": 88.0,
            "reference_note": "M1 Max, one thread, 1M U(0,1) adds (bench/Benchmark.md:102)",
            "gpu_library_ms": dt_gpu * 1e3,
            "gpu_note": "the same adds through the library: host buffer, merges every 500 samples on "
                        "the GPU (latency-bound; shown for parity, not speed)",
            "quantiles": {"q": list(RANK_QS), "oracle": oq.tolist(), "gpu": gq.tolist()},
            "cdf": {"x": xs.tolist(), "oracle": oc.tolist(), "gpu": gc.tolist()},
            "max_query_rel_diff": float(max(np.max(np.abs(gq - oq) / np.maximum(np.abs(oq), 1.0)),
                                            np.max(np.abs(gc - oc)))),
            "rank_error_oracle": (np.abs(rank(oq) - qs)).tolist(),
            "rank_error_gpu": (np.abs(rank(gq) - qs)).tolist(),
            "centroids": {"oracle": len(d), "gpu": len(g)}}


def traffic_bytes(phase: str):
    """HBM bytes per launch of a kernel group, from the committed rocprof
    FETCH_SIZE / WRITE_SIZE measurement (profiles/traffic.json), or None."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    e = t.get(phase)
    return None if e is None else e.get("bytes_per_launch")


def cpu_baseline_grouped(keys_h: np.ndarray, vals_h: np.ndarray, num_keys: int, threads: int = 1):
    """The CPU oracle (restatement of the reference MergingDigest) on a
    bounded sample of the grouped workload; keys split over `threads`."""
    from oracle import oracle as O

    digests = [O.Digest(3, 100.0) for _ in range(num_keys)]
    t0 = time.perf_counter()
    O.grouped_add_batch(digests, keys_h, vals_h, threads=threads)
    dt = time.perf_counter() - t0
    return keys_h.size / dt, dt


def plain_per_key(kd, session, keys, vals, nk, ms_grouped, dev):
    """The same grouped batch as one TDigest batch merge per key (one compress
    launch sequence per key).  All of the 64 largest keys are timed; the rest
    on every STRIDE-th key by size rank, scaled by STRIDE."""
    import torch

    order = torch.argsort(keys.to(torch.int64), stable=False)
    sv = vals[order]
    cnt = torch.bincount(keys.to(torch.int64), minlength=nk).cpu().numpy()
    off = np.zeros(nk + 1, np.int64)
    off[1:] = np.cumsum(cnt)
    by_size = np.argsort(-cnt, kind="stable")
    stride = 97
    picked = [(int(k), 1.0) for k in by_size[:64]]
    picked += [(int(k), float(stride)) for k in by_size[64::stride]]

    def one(k):
        d = kd.TDigest(kd.ScaleK3, 100.0, session=session)
        if cnt[k] > 0:
            d.add_batch_device(sv[off[k]:off[k + 1]])
        return d

    for k, _ in picked[:8]:
        one(k)
    session.synchronize()
    est = 0.0
    timed = 0.0
    for k, w in picked:
        t0 = time.perf_counter()
        one(k)
        session.synchronize()
        dt = time.perf_counter() - t0
        est += w * dt
        timed += dt
    del sv, order
    return {"workload": "one TDigest.add_batch_device (batch merge) per key of the same batch",
            "keys_timed": len(picked), "timed_s": timed,
            "sample": f"the 64 largest keys, then every {stride}th key by size, weighted by {stride}",
            "est_ms": est * 1e3, "speedup_grouped": est * 1e3 / ms_grouped}


def batched_decay(kd, session, dev, nd=8192, seed=3, with_parity=True):
    """8192 digests, geometric sizes (mean 32768, at most 2^22), Student-t(3)
    values, U(0.1, 10) weights, add_at timestamps with Exp(1e3/s) gaps,
    lambda 1e-3/s, ScaleK2, compression 300; flush; merge_into tree to one."""
    import torch

    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    p = 1.0 / 32768.0
    u = torch.rand(nd, generator=g, dtype=torch.float64, device=dev)
    sizes = torch.ceil(torch.log1p(-u) / math.log1p(-p)).clamp_(1, 1 << 22).to(torch.int64)
    off = torch.zeros(nd + 1, dtype=torch.int64, device=dev)
    off[1:] = torch.cumsum(sizes, 0)
    n = int(off[-1].item())
    z = torch.randn(n, generator=g, dtype=torch.float64, device=dev)
    chi = torch.zeros(n, dtype=torch.float64, device=dev)
    for _ in range(3):
        chi += torch.randn(n, generator=g, dtype=torch.float64, device=dev) ** 2
    vals = z / torch.sqrt(chi / 3.0)
    del z, chi
    w = 0.1 + 9.9 * torch.rand(n, generator=g, dtype=torch.float64, device=dev)
    gaps = -torch.log1p(-torch.rand(n, generator=g, dtype=torch.float64, device=dev)) * 1e-3
    cs = torch.cumsum(gaps, 0)
    del gaps
    seg = torch.repeat_interleave(torch.arange(nd, device=dev), sizes)
    base = torch.cat([torch.zeros(1, dtype=torch.float64, device=dev), cs])[off[:-1]]
    times = cs - base[seg]
    del cs, seg, base
    torch.cuda.synchronize()

    def run():
        b = kd.BatchedDigests(nd, 300.0, kd.Decay.rate(1e-3), scale=kd.ScaleK2, session=session)
        b.add_at_device(off, vals, times, w)
        b.flush()
        b.merge_tree()
        session.synchronize()
        return b

    run()
    t0 = time.perf_counter()
    b = run()
    dt = time.perf_counter() - t0
    st = b.state(0)
    out = {"workload": f"{nd} digests, geometric sizes mean 32768 (max 2^22), {n} samples, "
                       "Student-t(3), weights U(0.1,10), add_at with Exp(1e3/s) gaps, "
                       "lambda 1e-3/s, ScaleK2, compression 300, flush + merge_into tree",
           "samples": n, "max_digest": int(sizes.max().item()), "ms": dt * 1e3,
           "samples_per_s": n / dt, "bytes_per_sample": 56.0,
           "roofline_frac": 56.0 * n / dt / 1e9 / HBM_PEAK_GBS,
           "root_centroids": len(b.centroids(0)[0]), "root_total": st["stored_total"]}
    del b
    if with_parity:
        out["parity"] = batched_parity(kd, session, nd, off, vals, times, w)
    return out


def _cmp_digest_arrays(m, w, st, rm, rw, rs):
    """(max mean rel err, max weight rel err, state fields equal) or None when
    the centroid counts differ."""
    if m.size != rm.size:
        return None
    floor = 1e-12 * max(abs(rs["min"]), abs(rs["max"]), 1e-300)
    me = float(np.max(np.abs(m - rm) / np.maximum(np.abs(rm), floor))) if m.size else 0.0
    we = float(np.max(np.abs(w - rw) / np.abs(rw))) if w.size else 0.0
    eq = all(st[k] == rs[k] for k in ("clock", "log_scale", "min", "max"))
    te = abs(st["stored_total"] - rs["stored_total"]) / max(abs(rs["stored_total"]), 1e-300)
    return me, we, eq, te


def batched_parity(kd, session, nd, off, vals, times, w):
    """The batched decayed digests against the oracle: every digest replayed
    with add_at on the host (all host threads: the all-cores CPU baseline; a
    bounded prefix of digests on one thread: the single-thread baseline); the
    8 largest digests and every 256th are compared after the flush, then the
    merge_into tree roots."""
    from oracle import oracle as O

    offh, vh, th, wh = off.cpu().numpy(), vals.cpu().numpy(), times.cpu().numpy(), w.cpu().numpy()
    sizes = np.diff(offh)
    b = kd.BatchedDigests(nd, 300.0, kd.Decay.rate(1e-3), scale=kd.ScaleK2, session=session)
    b.add_at_device(off, vals, times, w)
    b.flush()
    session.synchronize()
    pick = np.unique(np.concatenate([np.argsort(-sizes, kind="stable")[:8], np.arange(0, nd, 256)]))
    got = {int(d): (b.centroids(int(d)), b.state(int(d))) for d in pick}
    b.merge_tree()
    root_m, root_w = b.centroids(0)
    root_st = b.state(0)
    del b
    thr = host_threads()
    refs = [O.Digest(2, 300.0, 1e-3) for _ in range(nd)]
    t0 = time.perf_counter()
    O.batch_add_at(refs, offh, vh, th, wh, threads=thr)
    for r in refs:
        r.merge()
    t_all = time.perf_counter() - t0
    # One thread on the first digests holding about 2^25 samples.
    k1 = int(np.searchsorted(offh, 1 << 25)) + 1
    k1 = min(k1, nd)
    one = [O.Digest(2, 300.0, 1e-3) for _ in range(k1)]
    t0 = time.perf_counter()
    O.batch_add_at(one, offh[:k1 + 1], vh, th, wh, threads=1)
    for r in one:
        r.merge()
    t_one = time.perf_counter() - t0
    del one
    worst = {"mean": 0.0, "weight": 0.0, "total": 0.0}
    failures = []
    for d, ((m, ww), st) in got.items():
        rm, rw = refs[d].centroids()
        c = _cmp_digest_arrays(m, ww, st, rm, rw, refs[d].state())
        if c is None:
            failures.append(f"digest {d}: {m.size} centroids, oracle {rm.size}")
            continue
        worst["mean"] = max(worst["mean"], c[0])
        worst["weight"] = max(worst["weight"], c[1])
        worst["total"] = max(worst["total"], c[3])
        if not c[2]:
            failures.append(f"digest {d}: state differs")
    stride = 1
    while stride < nd:
        for i in range(0, nd, 2 * stride):
            if i + stride < nd:
                refs[i].merge_into(refs[i + stride])
        stride *= 2
    rm, rw = refs[0].centroids()
    croot = _cmp_digest_arrays(root_m, root_w, root_st, rm, rw, refs[0].state())
    if croot is None:
        failures.append(f"tree root: {root_m.size} centroids, oracle {rm.size}")
    n = int(offh[-1])
    return {"digests_compared": int(pick.size), "max_mean_rel_err": worst["mean"],
            "max_weight_rel_err": worst["weight"], "max_stored_total_rel_err": worst["total"],
            "tree_root": None if croot is None else
            {"centroids": int(rm.size), "max_mean_rel_err": croot[0], "max_weight_rel_err": croot[1],
             "state_equal": croot[2], "stored_total_rel_err": croot[3]},
            "failures": failures,
            "cpu_baseline": {"one_thread": {"value": float(offh[k1]) / t_one, "unit": "samples/s",
                                            "cores": 1, "kind": "port", "s": t_one,
                                            "sample": f"the first {k1} digests ({int(offh[k1])} samples), "
                                                      "add_at per sample then merge"},
                             "all_cores": {"value": n / t_all, "unit": "samples/s", "cores": thr,
                                           "s": t_all, "sample": "all digests, add_at then merge "
                                                                 "(the tree not included)"}}}


def _serve_requests(rng, nk, rate, seconds, cdf, perm, pool):
    """An open-loop request schedule: Poisson arrivals, Zipf keys, request
    kinds and sizes as the serving workload states."""
    nreq = int(rate * seconds)
    arr = np.cumsum(rng.exponential(1.0 / rate, nreq))
    keys = perm[np.minimum(np.searchsorted(cdf, rng.uniform(size=nreq)), nk - 1)].astype(np.int64)
    kinds = rng.choice([0, 0, 1, 2], nreq).astype(np.int32)
    sizes = np.where(kinds == 0, np.exp(rng.uniform(0, np.log(65536.0), nreq)),
                     np.exp(rng.uniform(0, np.log(4096.0), nreq))).astype(np.int64)
    sizes = np.maximum(sizes, 1)
    off = np.zeros(nreq + 1, np.int64)
    off[1:] = np.cumsum(sizes)
    shift = (keys % 1000) / 10.0
    data = np.empty(off[-1])
    for r in range(nreq):
        a, b, m = off[r], off[r + 1], sizes[r]
        if kinds[r] == 0:
            p0 = rng.integers(0, pool.size - m)
            data[a:b] = pool[p0:p0 + m] + shift[r]
        elif kinds[r] == 1:
            data[a:b] = rng.uniform(0.0, 1.0, m)
        else:
            data[a:b] = shift[r] + pool[:m]
    return arr, keys, kinds, sizes, off, data


def _serve_run(srv, reqs, watch=None, workers=True):
    """Submits each request once its arrival time has passed; returns
    per-request latency (completion minus arrival, s), the wall time and, for
    the requests on keys in `watch`, their outcomes (query values, "empty",
    "enomem" or None for a finished ingest).  With `workers` the server's
    worker threads admit and run requests as they are submitted (this thread
    only submits and collects completions); otherwise this thread steps the
    server whenever requests are pending."""
    import kdigest_amd as kd

    arr, keys, kinds, sizes, off, data = reqs
    nreq = arr.size
    lat = np.full(nreq, np.nan)
    kept = {}
    first_id = None
    i = 0
    ndone = 0
    if workers:
        srv.start(4)
    t0 = time.perf_counter()
    c0 = srv.clock()
    while ndone < nreq:
        now = time.perf_counter() - t0
        j = int(np.searchsorted(arr, now, side="right"))
        if j > i:
            ids = srv.submit_many(kinds[i:j], keys[i:j], off[i:j + 1] - off[i], data[off[i]:off[j]])
            if first_id is None:
                first_id = int(ids[0])
            i = j
        if workers:
            cids, ct = srv.take_completed()
            if cids.size == 0:
                nxt = arr[i] - now if i < nreq else 1e-4
                time.sleep(max(0.0, min(nxt, 1e-4)))
        else:
            if srv.pending == 0:
                if i < nreq:
                    time.sleep(max(0.0, min(arr[i] - now, 1e-3)))
                continue
            srv.step(4)
            cids, ct = srv.take_completed()
        for rid, tc in zip(cids.tolist(), ct.tolist()):
            r = rid - first_id
            lat[r] = (tc - c0) - arr[r]
            if watch is not None and int(keys[r]) in watch:
                try:
                    res = srv.result(rid)
                    kept[r] = res if kinds[r] != 0 else None
                except kd.KdigestError as e:
                    kept[r] = "empty" if e.code == -2 else "enomem"
            else:
                srv.discard(rid)
        ndone += len(cids)
    wall = time.perf_counter() - t0
    if workers:
        srv.stop()
    return lat, wall, kept


def serving(kd, session, seconds=2.0, rate=20000.0, nk=1_000_000, seed=5, warm_seconds=0.5,
            with_parity=True, workers=True):
    """Open-loop serving: Poisson arrivals at `rate` req/s for `seconds`,
    keys Zipf(1.05) over 1M device-resident digests (ScaleK1, compression
    100, half-life 1e6 samples); half the requests ingest log-uniform 1..65536
    N(0,1) values shifted per key, a quarter ask quantiles and a quarter cdf
    values (log-uniform 1..4096 per request).  Requests are submitted when
    their arrival time has passed; latency = completion (server clock, when
    the worker batch or step that finished the request ends) minus scheduled
    arrival.  With `workers` the server's two worker threads serve requests
    as they arrive (queries beside ingest rounds); otherwise this thread
    steps it.  A warm-up run of the same mix (its own draws, `warm_seconds`
    long) goes first and is not measured."""
    rng = np.random.default_rng(seed)
    ranks = np.arange(1, nk + 1, dtype=np.float64)
    cdf = np.cumsum(ranks ** -1.05)
    cdf /= cdf[-1]
    perm = rng.permutation(nk)
    pool = rng.standard_normal(1 << 22)
    warm = _serve_requests(rng, nk, rate, warm_seconds, cdf, perm, pool)
    reqs = _serve_requests(rng, nk, rate, seconds, cdf, perm, pool)
    srv = kd.DigestServer(nk, 100.0, kd.Decay.half_life(1e6), scale=kd.ScaleK1, slot_capacity=128,
                          mem_budget=64 << 20, session=session)
    arr, keys, kinds, sizes, off, data = reqs
    # Watched keys (their outcomes are kept for the oracle replay): the four
    # with the most requests and 28 others drawn at random.
    uk, uc = np.unique(keys, return_counts=True)
    hot = uk[np.argsort(-uc, kind="stable")[:4]]
    rest = np.setdiff1d(uk[uc >= 2], hot)
    watch = set(int(k) for k in np.concatenate([hot, rng.choice(rest, min(28, rest.size),
                                                                replace=False)]))
    if not with_parity:
        watch = None
    _, _, warm_kept = _serve_run(srv, warm, watch, workers)
    s0 = srv.stats()
    lat, wall, kept = _serve_run(srv, reqs, watch, workers)
    nreq = arr.size
    lat *= 1e3
    st = srv.stats()
    st = {k: (v - s0[k] if k in ("steps", "admitted", "admitted_bytes", "ingest_rounds") else v)
          for k, v in st.items()}
    ing = kinds == 0
    q = int(nreq * 0.25)
    return {"workload": f"{nk} keys Zipf(1.05), ScaleK1, compression 100, half-life 1e6; "
                        f"Poisson {rate:.0f} req/s for {seconds:.1f} s after a {warm_seconds:.1f} s "
                        "warm-up, 1/2 ingest (log-uniform 1..65536), 1/4 quantile, 1/4 cdf "
                        "(log-uniform 1..4096)",
            "requests": nreq, "wall_s": wall, "requests_per_s": nreq / wall,
            "offered_per_s": rate,
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "max_ms": float(np.max(lat)),
            "p99_ms_by_quarter": [float(np.percentile(lat[i * q:(i + 1) * q], 99)) for i in range(4)],
            "ingest_p99_ms": float(np.percentile(lat[ing], 99)),
            "query_p99_ms": float(np.percentile(lat[~ing], 99)),
            "ingest_samples_per_s": float(sizes[ing].sum()) / wall,
            "queries_per_s": float(sizes[~ing].sum()) / wall,
            "admission": st, "mem_budget_bytes": 64 << 20,
            "parity": serve_parity(srv, warm, warm_kept, reqs, kept, watch) if with_parity else None}


def serve_parity(srv, warm, warm_kept, reqs, kept, watch):
    """Oracle replay of the serving request log: every ingest (warm-up and
    measured run, arrival order, one B batch merge per request) on all host
    threads (the CPU baseline), and for the watched keys every query answer
    and the final digest compared with the oracle at the same point of the
    replay."""
    from oracle import oracle as O

    def ingest_log(rq):
        arr, keys, kinds, sizes, off, data = rq
        sel = np.nonzero(kinds == 0)[0]
        return keys[sel], sel, off, data

    # All host threads, every ingest of both runs (dense ids of touched keys).
    kw, sw, ow, dw = ingest_log(warm)
    km, sm, om, dm = ingest_log(reqs)
    all_keys = np.concatenate([kw, km])
    ids, dense = np.unique(all_keys, return_inverse=True)
    pieces = [dw[ow[r]:ow[r + 1]] for r in sw] + [dm[om[r]:om[r + 1]] for r in sm]
    lens = np.array([p.size for p in pieces], np.int64)
    roff = np.zeros(lens.size + 1, np.int64)
    roff[1:] = np.cumsum(lens)
    rdata = np.concatenate(pieces)
    thr = host_threads()
    refs = [O.Digest(1, 100.0, math.log(2.0) / 1e6) for _ in range(ids.size)]
    t0 = time.perf_counter()
    O.replay_ingests(refs, dense, roff, rdata, threads=thr)
    t_all = time.perf_counter() - t0
    del refs, rdata, pieces
    # Watched keys: replay in arrival order, answering queries on the way.
    lam = math.log(2.0) / 1e6
    ref = {k: O.Digest(1, 100.0, lam) for k in watch}
    worst_q, nq, failures = 0.0, 0, []
    for rq, kp, measured in ((warm, warm_kept, False), (reqs, kept, True)):
        arr, keys, kinds, sizes, off, data = rq
        for r in np.nonzero(np.isin(keys, list(watch)))[0]:
            k, x = int(keys[r]), data[off[r]:off[r + 1]]
            out = kp.get(int(r), "missing")
            if kinds[r] == 0:
                if out != "enomem":
                    ref[k].add_batch_b(x)
                continue
            if not measured:
                continue
            d = ref[k]
            if len(d) == 0 and d.nbuffer() == 0:
                if not (isinstance(out, str) and out == "empty"):
                    failures.append(f"request {r}: expected an empty-key error")
                continue
            exp = d.quantiles(x) if kinds[r] == 1 else d.cdfs(x)
            if isinstance(out, str):
                failures.append(f"request {r}: {out}")
                continue
            err = np.abs(out - exp) / np.maximum(np.abs(exp), 1.0)
            worst_q = max(worst_q, float(err.max()))
            nq += x.size
    worst_m, worst_w = 0.0, 0.0
    for k in watch:
        m, w = srv.centroids(k)
        st = srv.key_state(k)
        rm, rw = ref[k].centroids()
        c = _cmp_digest_arrays(m, w, st, rm, rw, ref[k].state())
        if c is None:
            failures.append(f"key {k}: {m.size} centroids, oracle {rm.size}")
            continue
        worst_m, worst_w = max(worst_m, c[0]), max(worst_w, c[1])
        if not c[2]:
            failures.append(f"key {k}: state differs")
    n_ing = int(lens.size)
    return {"watched_keys": len(watch), "queries_compared": nq, "max_query_rel_err": worst_q,
            "max_mean_rel_err": worst_m, "max_weight_rel_err": worst_w, "failures": failures,
            "cpu_baseline": {"value": n_ing / t_all, "unit": "ingest requests/s", "cores": thr,
                             "kind": "port", "s": t_all,
                             "samples_per_s": float(roff[-1]) / t_all,
                             "sample": f"all {n_ing} ingest requests of the warm-up and measured "
                                       "runs, replayed in arrival order (keys over host threads)"}}


def single_digest(kd, session, dev, single_n: int, with_parity: bool = True):
    """One digest fed 2^20-value device batches (ScaleK2, compression 200,
    half-life 1e5 samples): one warm run, then one timed run."""
    import torch

    gen = torch.Generator(device=dev)
    gen.manual_seed(1)
    xs = torch.exp(1.5 * torch.randn(single_n, generator=gen, dtype=torch.float64, device=dev))
    batch = 1 << 20

    def run_single(one_call):
        d = kd.TDigest(kd.ScaleK2, 200.0, kd.Decay.half_life(1e5), session=session)
        if one_call:
            # The same 64 batch merges in one call: validation and every
            # batch's splitter sample done up front.
            d.add_batches_device(xs, batch)
        else:
            for a in range(0, single_n, batch):
                d.add_batch_device(xs[a:a + batch])
        return d

    def timed(one_call):
        run_single(one_call)
        session.synchronize()
        t0 = time.perf_counter()
        d = run_single(one_call)
        session.synchronize()
        return d, time.perf_counter() - t0

    d_loop, dt_loop = timed(False)
    d, dt = timed(True)
    m0, w0 = d_loop.centroids()
    m1, w1 = d.centroids()
    same = bool(np.array_equal(m0.view(np.uint64), m1.view(np.uint64))
                and np.array_equal(w0.view(np.uint64), w1.view(np.uint64))
                and d_loop.state() == d.state())
    out = {"workload": f"one digest, {single_n} lognormal(0,1.5) FP64 values in batches "
                       f"of {batch}, ScaleK2, compression 200, half-life 1e5",
           "samples_per_s": single_n / dt, "ms": dt * 1e3, "centroids": len(d),
           "batches": (single_n + batch - 1) // batch,
           "entry": "TDigest.add_batches_device (the batch merges in one call)",
           "per_call_loop": {"samples_per_s": single_n / dt_loop, "ms": dt_loop * 1e3,
                             "entry": "one TDigest.add_batch_device call per batch"},
           "one_call_equals_loop": same,
           "roofline_frac": 24.0 * single_n / dt / 1e9 / HBM_PEAK_GBS}
    if with_parity:
        out["queries_and_parity"] = single_parity(kd, d, xs, batch)
    return out


def single_parity(kd, d, xs, batch: int):
    """The single digest's query batch (4096 q ~ U[0,1] plus 1e-4 and 1 - 1e-4;
    4096 cdf values at sampled data values) timed on the GPU and compared with
    the oracle fed the same batches (B semantics, one thread: the CPU
    baseline), plus the digest state and centroids, and the decay-weighted
    rank error of both against the sorted data."""
    import torch

    from oracle import oracle as O

    rng = np.random.default_rng(1)
    qs = np.concatenate([rng.uniform(0, 1, 4096), [1e-4, 1 - 1e-4]])
    xh = xs.cpu().numpy()
    xq = xh[rng.integers(0, xh.size, 4096)]
    d.quantile(qs[:8])
    t0 = time.perf_counter()
    gq = d.quantile(qs)
    gc = d.cdf(xq)
    t_q = time.perf_counter() - t0
    lam = math.log(2.0) / 1e5
    r = O.Digest(2, 200.0, lam)
    t0 = time.perf_counter()
    for a in range(0, xh.size, batch):
        r.add_batch_b(xh[a:a + batch])
    t_cpu = time.perf_counter() - t0
    oq, oc = r.quantiles(qs), r.cdfs(xq)
    st, rs = d.state(), r.state()
    m, w = d.centroids()
    rm, rw = r.centroids()
    same_count = m.size == rm.size
    mean_err = float(np.max(np.abs(m - rm) / np.maximum(np.abs(rm), 1e-12 * max(abs(rs["min"]),
                                                                                   abs(rs["max"]))))) \
        if same_count else None
    w_err = float(np.max(np.abs(w - rw) / np.abs(rw))) if same_count else None
    # Decay-weighted rank error: sample i of batch b weighs e^(-lambda (T - T_b)).
    nb = (xh.size + batch - 1) // batch
    bw = np.exp(-lam * batch * (nb - 1 - np.arange(nb)))
    xt = torch.from_numpy(xh)
    order = torch.argsort(xt)
    sx = xt[order].numpy()
    sw = torch.from_numpy(np.repeat(bw, batch)[:xh.size])[order].numpy()
    cw = np.cumsum(sw) / sw.sum()
    q5 = np.array(RANK_QS)

    def wrank(v):
        i = np.searchsorted(sx, v, side="right")
        return np.where(i > 0, cw[np.maximum(i - 1, 0)], 0.0)

    return {"query_batch": {"quantiles": int(qs.size), "cdfs": int(xq.size), "ms": t_q * 1e3,
                            "queries_per_s": (qs.size + xq.size) / t_q,
                            "note": "host arrays in and out, one call each"},
            "cpu_baseline": {"value": xh.size / t_cpu, "unit": "samples/s", "cores": 1,
                             "kind": "port", "s": t_cpu,
                             "sample": "the whole workload, same batches, oracle B-semantics merges"},
            "centroids": {"gpu": int(m.size), "oracle": int(rm.size)},
            "state_equal": {k: bool(st[k] == rs[k]) for k in ("clock", "log_scale", "min", "max")},
            "stored_total_rel_diff": abs(st["stored_total"] - rs["stored_total"]) / rs["stored_total"],
            "max_mean_rel_err": mean_err, "max_weight_rel_err": w_err,
            "max_quantile_rel_err": float(np.max(np.abs(gq - oq) / np.maximum(np.abs(oq), 1.0))),
            "max_cdf_abs_err": float(np.max(np.abs(gc - oc))),
            "weighted_rank_error": {"q": list(RANK_QS),
                                    "gpu": np.abs(wrank(d.quantile(q5)) - q5).tolist(),
                                    "oracle": np.abs(wrank(r.quantiles(q5)) - q5).tolist(),
                                    "note": "each sample weighted by its batch's decay factor at "
                                            "the end of the stream"}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1 << 28)
    ap.add_argument("--keys", type=int, default=100000)
    ap.add_argument("--single-n", type=int, default=1 << 26)
    ap.add_argument("--cpu-sample", type=int, default=1 << 26)
    ap.add_argument("--serve-seconds", type=float, default=2.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-single", action="store_true")
    ap.add_argument("--no-queries", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-wire", action="store_true")
    ap.add_argument("--no-parity", action="store_true",
                    help="skip every parity / oracle comparison leg")
    ap.add_argument("--no-plain-cpu", action="store_true")
    ap.add_argument("--batched-digests", type=int, default=8192)
    ap.add_argument("--no-keymap", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--shape", type=str, default=None,
                    help="fixed grouped launch shape items,chunks,tile_min,tiles_max,overlap "
                         "(implies --no-autotune)")
    ap.add_argument("--radix-key-bits", type=int, default=None,
                    help="grouped batches over more than 2^bits keys take the counting partition "
                         "(default: the library's 27)")
    ap.add_argument("--only-batched", action="store_true",
                    help="only the batched decayed digests (config 4) line, no grouped run")
    ap.add_argument("--only-single", action="store_true",
                    help="measure only the single-digest line (for its kernel profile)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run one 1 GiB device-to-device copy first (calibrates rocprof byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch

    import kdigest_amd as kd

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    session = kd.default_session(0)
    if args.only_single:
        single = single_digest(kd, session, dev, args.single_n, with_parity=False)
        print(json.dumps({"metric": "single-digest ingest samples/s", "value": single["samples_per_s"],
                          "unit": "samples/s", "n_gpus": 1, "higher_is_better": True,
                          "dtype": "f64", "data": "synthetic (torch Philox, seeded), device-resident",
                          "config": {"workload": single["workload"]}, "single_digest": single}))
        return
    if args.only_batched:
        batched = batched_decay(kd, session, dev, nd=args.batched_digests, with_parity=False)
        print(json.dumps({"metric": "batched decayed digests samples/s", "value": batched["samples_per_s"],
                          "unit": "samples/s", "n_gpus": 1, "higher_is_better": True,
                          "dtype": "f64", "data": "synthetic (torch Philox, seeded), device-resident",
                          "config": {"workload": batched["workload"]}, "batched_decay": batched}))
        return
    if args.radix_key_bits is not None:
        session.set_radix_key_bits(args.radix_key_bits)
    n, nk = args.n, args.keys
    print(f"[bench] generating {n} grouped pairs over {nk} keys", file=sys.stderr, flush=True)
    keys, vals = gen_grouped(n, nk, 0, dev)
    if args.copy_probe:
        a = torch.empty(1 << 27, dtype=torch.float64, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)  # 1 GiB read + 1 GiB written by one blit kernel
        torch.cuda.synchronize()
        del a, b

    # ------------------------------------------------ grouped ingest (value)
    tuned = None
    if args.shape:
        it, ch, tmin, tmax, ov = (int(x) for x in args.shape.split(","))
        session.set_partition_shape(it, ch)
        session.set_select_shape(tmin, tmax)
        session.set_overlap(int(ov))
        tuned = {"items": it, "chunks": ch, "tile_min": tmin, "tiles_max": tmax, "overlap": int(ov),
                 "fixed": True}
    elif not args.no_autotune:
        # Launch shape picked by the library's autotuner on this batch (not
        # timed; every shape builds identical digests).
        tuned = session.autotune_grouped(keys, vals, nk, scale=kd.ScaleK3, compression=100.0)
        print(f"[bench] autotuned shape {tuned}", file=sys.stderr, flush=True)
    for w in range(args.warmup):
        g = kd.GroupedTDigest(nk, 100.0, scale=kd.ScaleK3, session=session)
        g.add_batch_device(keys, vals)
        del g
        print(f"[bench] warmup {w} done", file=sys.stderr, flush=True)
    digs = [kd.GroupedTDigest(nk, 100.0, scale=kd.ScaleK3, session=session) for _ in range(args.steps)]
    p0 = session.path_counts()
    torch.cuda.synchronize()
    session.synchronize()
    t0 = time.perf_counter()
    for g in digs:
        g.add_batch_device(keys, vals)
    session.synchronize()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    p1 = session.path_counts()
    paths = {k: (p1[k] - p0[k]) // args.steps for k in p1}
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = n * args.steps / (t1 - t0)
    print(f"[bench] grouped ingest {value / 1e9:.3f} G samples/s, {ms_per_step:.2f} ms/step",
          file=sys.stderr, flush=True)
    # One more step, outside the timed region, with per-phase HIP events.
    prof_steps = 2
    session.profile(True)
    for _ in range(prof_steps):
        gp = kd.GroupedTDigest(nk, 100.0, scale=kd.ScaleK3, session=session)
        gp.add_batch_device(keys, vals)
    session.synchronize()
    phases = session.phase_stats()
    session.profile(False)
    del gp
    # Dominant device kernel group with algorithmic bytes (HIP events on the
    # session stream around each launch).
    dev_phases = {k: v for k, v in phases.items() if not k.startswith("host:") and v[2] > 0}
    dom_name, (dom_ms, dom_launches, dom_bytes) = max(dev_phases.items(), key=lambda kv: kv[1][0])
    dom_avg_ms = dom_ms / max(dom_launches, 1)
    dom_gbs = (dom_bytes / max(dom_launches, 1)) / (dom_avg_ms * 1e-3) / 1e9 if dom_avg_ms > 0 else 0.0
    pipe_gbs = GROUPED_BYTES_PER_SAMPLE * n / (ms_per_step * 1e-3) / 1e9
    dom_share = dom_ms / prof_steps / ms_per_step

    # ------------------------------------------------ tail quantiles per key
    queries = None
    if not args.no_queries:
        g = digs[-1]
        per_key = 1024
        qk = torch.arange(nk, dtype=torch.int64, device=dev).repeat_interleave(per_key)
        lo = torch.logspace(-4, -2, per_key // 2, dtype=torch.float64, device=dev)
        qv = torch.cat([lo, 1.0 - lo.flip(0)]).repeat(nk)
        out = torch.empty_like(qv)
        g.quantile_device(qk, qv, out)
        session.synchronize()
        reps = 3
        t0 = time.perf_counter()
        for _ in range(reps):
            g.quantile_device(qk, qv, out)
        session.synchronize()
        dt = (time.perf_counter() - t0) / reps
        # Single-request latency: one key, 16 quantiles, host round trip.
        lat = []
        rng = np.random.default_rng(0)
        for i in range(20):  # warm-up requests (first-call staging), not timed
            g.quantile(np.full(16, int(rng.integers(0, nk))), np.linspace(0.001, 0.999, 16))
        for i in range(200):
            k = int(rng.integers(0, nk))
            s = time.perf_counter()
            g.quantile(np.full(16, k), np.linspace(0.001, 0.999, 16))
            lat.append(time.perf_counter() - s)
        lat = np.array(lat) * 1e6
        queries = {"queries": int(qv.numel()), "queries_per_s": qv.numel() / dt,
                   "ms": dt * 1e3, "request_p50_us": float(np.percentile(lat, 50)),
                   "request_p99_us": float(np.percentile(lat, 99)),
                   "request": "16 quantiles of one key, host arrays, synchronous"}
        del qk, qv, out
        print(f"[bench] queries {queries['queries_per_s'] / 1e9:.3f} G/s", file=sys.stderr, flush=True)
    # ------------------------------------------------ parity at full size
    parity = None
    oracle_all = None
    if not args.no_parity:
        parity, refs = grouped_parity(kd, digs[-1], keys, vals, nk)
        oracle_all = (parity["oracle_s"], parity["oracle_threads"])
        del refs
        print(f"[bench] parity: {len(parity['failures'])} failures, max mean rel err "
              f"{parity.get('max_mean_rel_err')}, max quantile rel err {parity['max_quantile_rel_err']}",
              file=sys.stderr, flush=True)
    # ------------------------------------------------ wire format, every key
    wire = None
    if not args.no_wire:
        g = digs[-1]
        nbytes = g.wire_size()
        wout = torch.empty(nbytes, dtype=torch.uint8, device=dev)
        woff = torch.empty(nk + 1, dtype=torch.int64, device=dev)
        g.to_bytes_device(wout, woff)
        r = kd.GroupedTDigest.from_bytes_device(wout, woff, scale=kd.ScaleK3, session=session)
        del r
        reps = 3
        t0 = time.perf_counter()
        for _ in range(reps):
            g.to_bytes_device(wout, woff)
        session.synchronize()
        t_ser = (time.perf_counter() - t0) / reps
        t0 = time.perf_counter()
        for _ in range(reps):
            r = kd.GroupedTDigest.from_bytes_device(wout, woff, scale=kd.ScaleK3, session=session)
            del r
        session.synchronize()
        t_de = (time.perf_counter() - t0) / reps
        # Map-reduce finalize: the same blobs loaded as independent digests
        # and reduced to one by the merge_into tree.
        session.synchronize()
        t0 = time.perf_counter()
        mb = kd.BatchedDigests.from_bytes_device(wout, woff, scale=kd.ScaleK3, session=session)
        session.synchronize()
        t_load = time.perf_counter() - t0
        t0 = time.perf_counter()
        mb.merge_tree()
        session.synchronize()
        t_tree = time.perf_counter() - t0
        root_c = len(mb.centroids(0)[0])
        del mb
        wire = {"workload": f"every key of the {nk}-key grouped digest, per-digest wire layout, "
                            "device to device",
                "bytes": nbytes, "to_bytes_ms": t_ser * 1e3, "from_bytes_ms": t_de * 1e3,
                "to_bytes_gb_s": nbytes / t_ser / 1e9, "from_bytes_gb_s": nbytes / t_de / 1e9,
                "map_reduce": {"workload": f"the {nk} blobs loaded as independent digests and "
                                          "reduced by the merge_into tree (first call: "
                                          "graph capture included)",
                               "load_ms": t_load * 1e3, "tree_ms": t_tree * 1e3,
                               "root_centroids": root_c}}
        del wout, woff
        print(f"[bench] wire {wire['to_bytes_gb_s']:.1f} / {wire['from_bytes_gb_s']:.1f} GB/s",
              file=sys.stderr, flush=True)
    del digs

    # ------------------------------------------------ plain per-key launches
    plain = None
    if not args.no_plain:
        plain = plain_per_key(kd, session, keys, vals, nk, ms_per_step, dev)
        print(f"[bench] plain per-key est {plain['est_ms']:.1f} ms "
              f"({plain['speedup_grouped']:.1f}x)", file=sys.stderr, flush=True)

    # ------------------------------------------------ CPU baseline
    cpu = None
    if not args.no_cpu:
        m = min(args.cpu_sample, n)
        kh = keys[:m].cpu().numpy().astype(np.uint32)
        vh = vals[:m].cpu().numpy()
        rate, dt = cpu_baseline_grouped(kh, vh, nk)
        if oracle_all is not None:
            dt_mt, nthr = oracle_all
            all_cores = {"value": n / dt_mt, "cores": nthr, "s": dt_mt,
                         "note": f"the whole {n}-pair batch (the parity run's oracle), keys "
                                 "partitioned over host threads"}
        else:
            nthr = host_threads()
            rate_mt, dt_mt = cpu_baseline_grouped(kh, vh, nk, threads=nthr)
            all_cores = {"value": rate_mt, "cores": nthr, "s": dt_mt,
                         "note": "same sample, keys partitioned over host threads"}
        cpu = {"value": rate, "unit": "samples/s", "cores": 1, "kind": "port",
               "sample": f"first {m} pairs of the same grouped workload, oracle per-key batch merge "
                         "(C restatement of the reference MergingDigest, gcc -O3 "
                         f"-ffp-contract=off), {dt:.1f} s",
               "all_cores": all_cores,
               "host": {"cpu_model": cpu_model(), "nproc": os.cpu_count(),
                        "usable_cores": len(os.sched_getaffinity(0)),
                        "threads_note": "multi-threaded runs use at most 16 threads, one GPU's "
                                        "share of the host"},
               "gpu_over_cpu": value / rate}
        del kh, vh
    # ------------------------------------------------ key map (sparse keys)
    keymap = None
    if not args.no_keymap:
        ids = torch.empty_like(keys)
        km = kd.KeyMap(nk, session=session)
        session.synchronize()
        t0 = time.perf_counter()
        km.map_device(keys, ids, insert=True)
        t_new = time.perf_counter() - t0
        t0 = time.perf_counter()
        km.map_device(keys, ids, insert=True)
        t_known = time.perf_counter() - t0
        keymap = {"workload": f"the {n} grouped keys through a {nk}-key map (uint32 keys to dense "
                              "ids): first batch inserting every key, then the same batch with every "
                              "key known",
                  "keys": len(km), "first_ms": t_new * 1e3, "known_ms": t_known * 1e3,
                  "known_keys_per_s": n / t_known}
        del ids, km
        print(f"[bench] key map {keymap['known_keys_per_s'] / 1e9:.1f} G keys/s", file=sys.stderr,
              flush=True)
    del keys, vals
    torch.cuda.empty_cache()

    # ------------------------------------------------ single digest
    single = None
    if not args.no_single:
        single = single_digest(kd, session, dev, args.single_n, with_parity=not args.no_parity)
        print(f"[bench] single {single['samples_per_s'] / 1e9:.3f} G samples/s", file=sys.stderr, flush=True)

    # ------------------------------------------------ batched, decayed, merge tree
    batched = None
    if not args.no_batched:
        batched = batched_decay(kd, session, dev, nd=args.batched_digests,
                                with_parity=not args.no_parity)
        torch.cuda.empty_cache()
        print(f"[bench] batched {batched['samples_per_s'] / 1e9:.3f} G samples/s", file=sys.stderr,
              flush=True)

    # ------------------------------------------------ plain single digest (CPU)
    plain_single = None
    if not args.no_plain_cpu:
        plain_single = plain_cpu(kd, session)
        print(f"[bench] plain single digest: oracle {plain_single['cpu_ns_per_add']:.1f} ns/add",
              file=sys.stderr, flush=True)

    # ------------------------------------------------ serving
    serve = None
    if not args.no_serve:
        serve = serving(kd, session, seconds=args.serve_seconds, with_parity=not args.no_parity)
        print(f"[bench] serving (workers) p50 {serve['p50_ms']:.2f} ms p99 {serve['p99_ms']:.2f} ms",
              file=sys.stderr, flush=True)
        # The same arrivals served by stepping the server from this thread
        # (one continuous batch at a time), for comparison.
        step_mode = serving(kd, session, seconds=args.serve_seconds, with_parity=False,
                            workers=False)
        serve["step_mode"] = {k: step_mode[k] for k in
                              ("requests_per_s", "p50_ms", "p99_ms", "max_ms", "ingest_p99_ms",
                               "query_p99_ms", "admission")}
        print(f"[bench] serving (step) p50 {step_mode['p50_ms']:.2f} ms "
              f"p99 {step_mode['p99_ms']:.2f} ms", file=sys.stderr, flush=True)

    line = {
        "metric": "ingest samples/s (% HBM peak) single + 100k-key grouped; p99 query latency",
        "value": value,
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (torch Philox / numpy, seeded), device-resident before timing",
        "config": {"workload": f"grouped ingest: {nk} uint32 keys Zipf(1.1), {n} (key, f64) pairs, "
                               "0.9*N(mu_k,1)+0.1*(mu_k+Pareto(1.5)), ScaleK3, compression 100",
                   "keys": nk, "pairs": n},
        "roofline": {"bound": "hbm", "achieved": dom_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": dom_gbs / HBM_PEAK_GBS, "traffic": traffic_bytes(dom_name),
                     "kernel": dom_name, "kernel_avg_ms": dom_avg_ms,
                     "share_of_step": dom_share,
                     "algorithmic_bytes_per_launch": dom_bytes / max(dom_launches, 1)},
        "pipeline_roofline": {"bytes_per_sample": GROUPED_BYTES_PER_SAMPLE, "achieved": pipe_gbs,
                              "frac": pipe_gbs / HBM_PEAK_GBS},
        "phases_ms_per_step": {k: v[0] / prof_steps for k, v in phases.items()},
        "phases_note": f"HIP events per kernel group, {prof_steps} extra steps outside the timed region",
        "compress_paths": paths,
        "autotuned_shape": tuned,
        "cpu_baseline": cpu,
        "parity": parity,
        "plain_per_key": plain,
        "plain_single_cpu": plain_single,
        "single_digest": single,
        "grouped_queries": queries,
        "grouped_wire_format": wire,
        "key_map": keymap,
        "batched_decay": batched,
        "serving": serve,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
