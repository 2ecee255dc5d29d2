"""Benchmark of the digest ingest / query path on one MI355X.

Headline workload (value): grouped per-key digests, 100,000 uint32 keys with
Zipf(s=1.1) frequencies, 2^28 (key, FP64 value) pairs with per-key values
0.9*N(mu_k, 1) + 0.1*(mu_k + Pareto(alpha=1.5, x_m=1)), mu_k ~ U(-100, 100);
ScaleK3, compression 100, no decay.  One step = one GroupedTDigest.add_batch
of the whole device-resident batch into fresh digests (key grouping, per-key
sort, greedy-sweep compress, finish).  Secondary lines in the same JSON:
single-digest ingest (2^26 lognormal(0, 1.5) values in 64 batches of 2^20,
ScaleK2, compression 200, half-life 1e5) and the 1024 tail quantiles per key
of the grouped digests.

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


def cpu_baseline_grouped(keys_h: np.ndarray, vals_h: np.ndarray, num_keys: int):
    """The CPU oracle (restatement of the reference MergingDigest) on a
    bounded sample of the grouped workload, one thread."""
    from oracle import oracle as O

    digests = [O.Digest(3, 100.0) for _ in range(num_keys)]
    t0 = time.perf_counter()
    O.grouped_add_batch(digests, keys_h, vals_h, threads=1)
    dt = time.perf_counter() - t0
    return keys_h.size / dt, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1 << 28)
    ap.add_argument("--keys", type=int, default=100000)
    ap.add_argument("--single-n", type=int, default=1 << 26)
    ap.add_argument("--cpu-sample", type=int, default=1 << 24)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-single", action="store_true")
    ap.add_argument("--no-queries", action="store_true")
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
    n, nk = args.n, args.keys
    print(f"[bench] generating {n} grouped pairs over {nk} keys", file=sys.stderr, flush=True)
    keys, vals = gen_grouped(n, nk, 0, dev)

    # ------------------------------------------------ grouped ingest (value)
    for w in range(args.warmup):
        g = kd.GroupedTDigest(nk, 100.0, scale=kd.ScaleK3, session=session)
        g.add_batch_device(keys, vals)
        del g
        print(f"[bench] warmup {w} done", file=sys.stderr, flush=True)
    digs = [kd.GroupedTDigest(nk, 100.0, scale=kd.ScaleK3, session=session) for _ in range(args.steps)]
    session.profile(True)
    torch.cuda.synchronize()
    session.synchronize()
    t0 = time.perf_counter()
    for g in digs:
        g.add_batch_device(keys, vals)
    session.synchronize()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    phases = session.phase_stats()
    session.profile(False)
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = n * args.steps / (t1 - t0)
    print(f"[bench] grouped ingest {value / 1e9:.3f} G samples/s, {ms_per_step:.2f} ms/step",
          file=sys.stderr, flush=True)
    dom = max(phases.items(), key=lambda kv: kv[1][0])
    dom_name, (dom_ms, dom_launches, dom_bytes) = dom
    dom_avg_ms = dom_ms / max(dom_launches, 1)
    dom_gbs = (dom_bytes / max(dom_launches, 1)) / (dom_avg_ms * 1e-3) / 1e9 if dom_avg_ms > 0 else 0.0
    pipe_gbs = GROUPED_BYTES_PER_SAMPLE * n / (ms_per_step * 1e-3) / 1e9

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
        print(f"[bench] queries {queries['queries_per_s'] / 1e9:.3f} G/s", file=sys.stderr, flush=True)

    # ------------------------------------------------ single digest
    single = None
    if not args.no_single:
        gen = torch.Generator(device=dev)
        gen.manual_seed(1)
        xs = torch.exp(1.5 * torch.randn(args.single_n, generator=gen, dtype=torch.float64, device=dev))
        batch = 1 << 20

        def run_single():
            d = kd.TDigest(kd.ScaleK2, 200.0, kd.Decay.half_life(1e5), session=session)
            for a in range(0, args.single_n, batch):
                d.add_batch_device(xs[a:a + batch])
            return d

        run_single()
        session.synchronize()
        t0 = time.perf_counter()
        d = run_single()
        session.synchronize()
        dt = time.perf_counter() - t0
        single = {"workload": f"one digest, {args.single_n} lognormal(0,1.5) FP64 values in batches "
                              f"of {batch}, ScaleK2, compression 200, half-life 1e5",
                  "samples_per_s": args.single_n / dt, "ms": dt * 1e3, "centroids": len(d),
                  "roofline_frac": 24.0 * args.single_n / dt / 1e9 / HBM_PEAK_GBS}
        print(f"[bench] single {single['samples_per_s'] / 1e9:.3f} G samples/s", file=sys.stderr, flush=True)

    # ------------------------------------------------ CPU baseline
    cpu = None
    if not args.no_cpu:
        m = min(args.cpu_sample, n)
        kh = keys[:m].cpu().numpy().astype(np.uint32)
        vh = vals[:m].cpu().numpy()
        rate, dt = cpu_baseline_grouped(kh, vh, nk)
        cpu = {"value": rate, "unit": "samples/s", "cores": 1, "kind": "port",
               "sample": f"first {m} pairs of the same grouped workload, oracle per-key batch merge "
                         f"(C restatement of the reference MergingDigest), {dt:.1f} s"}

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
        "data": "synthetic (torch Philox, seeded), device-resident before timing",
        "config": {"workload": f"grouped ingest: {nk} uint32 keys Zipf(1.1), {n} (key, f64) pairs, "
                               "0.9*N(mu_k,1)+0.1*(mu_k+Pareto(1.5)), ScaleK3, compression 100",
                   "keys": nk, "pairs": n},
        "roofline": {"bound": "hbm", "achieved": dom_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": dom_gbs / HBM_PEAK_GBS, "traffic": None, "kernel": dom_name,
                     "kernel_avg_ms": dom_avg_ms},
        "pipeline_roofline": {"bytes_per_sample": GROUPED_BYTES_PER_SAMPLE, "achieved": pipe_gbs,
                              "frac": pipe_gbs / HBM_PEAK_GBS},
        "phases_ms_per_step": {k: v[0] / args.steps for k, v in phases.items()},
        "cpu_baseline": cpu,
        "single_digest": single,
        "grouped_queries": queries,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
