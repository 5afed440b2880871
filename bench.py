#!/usr/bin/env python3
"""Benchmark of the TGT+/SABA hot path on one MI355X.

Default workload (the metric's SABA leg at its single-GPU scale): SABA 1-way
walks on a seeded Chung-Lu graph, n = 2^20 vertices, ~10M undirected edges
(~20M CSR entries), exponent 2.2, 10 hops, 16,384 walks per vertex, 8-lane
bouquets on the ordered grid, uint32 counters + uint64 checksum.  One step =
one full pass (counter clear + walks + reduction), the window of the
reference's t_rw_s (src/rw_experiment_saba.cpp:184-211).  The graph is
resident in HBM before timing starts.

value = transitions/s = n * n_walks * depth / seconds per step
(rw_experiment_saba.cpp:214-215).  One JSON line on stdout.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu] [--no-tgtp]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "SABA transitions/s + tgtp edges/s at eps 0.01, batched vs plain, 1 GPU"
HBM_PEAK_GBS = 8000.0
BYTES_PER_TRANSITION = 20  # row bounds 8 + neighbour 4 + counter RMW 8


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def host_threads() -> int:
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return max(1, min(16, len(os.sched_getaffinity(0))))
    except AttributeError:
        return max(1, min(16, os.cpu_count() or 1))


def traffic_for(kernel: str, workload: str):
    """HBM bytes per launch from the committed rocprofv3 PMC summary, if any."""
    p = ROOT / "profiles" / "pmc_traffic.json"
    if not p.exists():
        return None
    try:
        d = json.loads(p.read_text())
        return d.get(f"{kernel}@{workload}", {}).get("bytes_per_launch")
    except (ValueError, OSError):
        return None


def bench_saba(args):
    from tgtp_amd import Context, workloads as WL

    depth, n_walks = args.depth, args.walks
    t0 = time.time()
    g = WL.chung_lu_1m()
    log(f"[bench] graph n={g.n} m={g.m} generated in {time.time() - t0:.1f}s")
    ctx = Context(0)
    dg = ctx.upload(g)
    trans = g.n * n_walks * depth  # reference definition, all n vertices
    for i in range(args.warmup):
        _, vc, cs, st = ctx.saba_walks(dg, depth, n_walks, want_counters=False)
        log(f"[bench] warmup {i}: {st['kernel_ms']:.1f} ms checksum={cs}")
    walk_ms = []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        _, vc, cs, st = ctx.saba_walks(dg, depth, n_walks, want_counters=False)
        walk_ms.append(st["walk_ms"])
    elapsed = time.perf_counter() - t_start  # each call ends with a stream synchronize
    dg.free()
    ctx.close()
    ms_per_step = 1e3 * elapsed / args.steps
    value = trans / (elapsed / args.steps)
    avg_walk_s = sum(walk_ms) / len(walk_ms) / 1e3
    achieved = BYTES_PER_TRANSITION * trans / avg_walk_s / 1e9
    workload = (f"saba_walks chung-lu n=2^20 beta=2.2 m={g.m} depth={depth} "
                f"n_walks={n_walks}")
    log(f"[bench] {args.steps} steps: {ms_per_step:.2f} ms/step, walk kernel "
        f"{avg_walk_s * 1e3:.2f} ms, {value:.3e} transitions/s, checksum={cs}")
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic_for("saba_walk_kernel", "cl1m_d10_w16384"),
            "kernel": "saba_walk_kernel", "kernel_ms": round(avg_walk_s * 1e3, 3),
            "bytes_per_unit": BYTES_PER_TRANSITION, "units_per_launch": trans}
    cfg = {"workload": workload, "graph": WL.describe(WL.CL1M), "n": g.n, "m": g.m,
           "depth": depth, "n_walks": n_walks, "visit_count": vc, "checksum": cs}
    return value, ms_per_step, roof, cfg, g


def bench_tgtp(args):
    """tgtp() on the n=4096 Chung-Lu graph (omega=128, gamma=10): edges/s =
    CSR entries / seconds of one call with graph and eigenpairs resident in
    HBM (benchmark.cpp:40-42 times tgtp() only).  Returns a dict of legs."""
    import numpy as np

    import oracle
    from oracle.exact import exact_resistance
    from tgtp_amd import (MODE_FAITHFUL, MODE_FAST_BF16, MODE_FAST_FP32, Context,
                          eigenprep as E, workloads as WL)

    g = WL.chung_lu_4096()
    t0 = time.time()
    lam, X = E.eigens_f32(*E.compute_eigens(g, 128))
    R, _ = exact_resistance(g.idx, g.adj)
    log(f"[bench] tgtp graph n={g.n} m={g.m}, eigenpairs + exact oracle in {time.time() - t0:.1f}s")
    out = {"graph": WL.describe(WL.CL4096), "n": g.n, "m": g.m, "omega": 128, "gamma": 10}
    threads = host_threads()
    modes = ((MODE_FAITHFUL, "faithful"), (MODE_FAST_FP32, "fast_fp32"),
             (MODE_FAST_BF16, "fast_bf16"))
    with Context(0) as ctx:
        dg = ctx.upload(g)
        for eps in (0.01, 0.05):
            leg = {}
            ref = oracle.tgtp(g.idx, g.adj, lam, X, eps, 10, threads=threads)
            for mode, name in modes:
                s, st = ctx.score(dg, lam, X, eps, 10, mode=mode)  # warmup + eigen upload
                times, kms = [], []
                for _ in range(max(3, args.steps)):
                    t1 = time.perf_counter()
                    s, st = ctx.score(dg, lam, X, eps, 10, mode=mode)
                    times.append(time.perf_counter() - t1)
                    kms.append(st["kernel_ms"])
                best = min(times)
                err = np.abs(s.astype(np.float64) - R)
                leg[name] = {"edges_per_s": g.m / best, "seconds": best,
                             "kernel_ms": min(kms), "taus_ms": st["taus_ms"],
                             "diffuse_ms": st["diffuse_ms"], "host_fixups": st["host_fixups"],
                             "bitwise_equal_cpu": bool(np.array_equal(s.view(np.uint32),
                                                                      ref.view(np.uint32))),
                             "vs_exact_fp64": {"max_abs": float(err.max()),
                                               "mean_abs": float(err.mean()),
                                               "max_rel": float((err / R).max()),
                                               "mean_rel": float((err / R).mean()),
                                               "within_eps": bool(err.max() <= eps)}}
            # PCIe-inclusive: host arrays in, scores out (tgtp_score)
            t1 = time.perf_counter()
            ctx.score(g, lam, X, eps, 10)
            leg["faithful"]["e2e_with_h2d_edges_per_s"] = g.m / (time.perf_counter() - t1)
            ctimes = []
            for _ in range(3):
                t1 = time.perf_counter()
                oracle.tgtp(g.idx, g.adj, lam, X, eps, 10, threads=threads)
                ctimes.append(time.perf_counter() - t1)
            leg["cpu_baseline"] = {"edges_per_s": g.m / min(ctimes), "seconds": min(ctimes),
                                   "cores": threads, "kind": "port",
                                   "sample": "full run, min of 3 (oracle C++ OpenMP, "
                                             "schedule(dynamic) over sources)"}
            out[f"eps_{eps}"] = leg
            log(f"[bench] tgtp eps={eps}: " + ", ".join(
                f"{nm} {leg[nm]['edges_per_s']:.3e} edges/s (max err {leg[nm]['vs_exact_fp64']['max_abs']:.4f})"
                for _, nm in modes) + f"; cpu {leg['cpu_baseline']['edges_per_s']:.3e}")
        dg.free()
    return out


def cpu_saba_baseline(g, depth, n_walks, sample_sources: int):
    """The oracle's OpenMP walk (schedule(dynamic) over sources, per-thread
    counters) on evenly spread source slices of the same workload."""
    import oracle

    threads = host_threads()
    slices = 32
    per = max(1, sample_sources // slices)
    step = g.n // slices
    oracle.saba_walks(g.idx, g.adj, depth, n_walks, 0, min(g.n, 64), threads=threads, chunk=8)
    t0 = time.perf_counter()
    total = 0
    for k in range(slices):
        lo = k * step
        hi = min(g.n, lo + per)
        oracle.saba_walks(g.idx, g.adj, depth, n_walks, lo, hi, threads=threads, chunk=8)
        total += (hi - lo) * n_walks * depth
    dt = time.perf_counter() - t0
    return {"value": total / dt, "unit": "transitions/s", "cores": threads, "kind": "port",
            "sample": f"{slices} slices x {per} sources (of {g.n}), depth {depth}, "
                      f"{n_walks} walks/vertex = {total:.3e} transitions in {dt:.1f}s; "
                      f"oracle C++ OpenMP dynamic chunk 8, uint32 counters"}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawTextHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--depth", type=int, default=10)
    ap.add_argument("--walks", type=int, default=16384)
    ap.add_argument("--cpu-sample-sources", type=int, default=8192)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tgtp", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the multi-GPU "
              "design (source sharding + one counter all-reduce) is described in DESIGN.md "
              "and is not built", file=sys.stderr)
        return 2
    value, ms_per_step, roof, cfg, g = bench_saba(args)
    line = {"metric": METRIC, "value": value, "unit": "transitions/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
            "data": "synthetic (seeded Chung-Lu generator, regenerated on the box)",
            "config": cfg, "roofline": roof}
    if not args.no_cpu:
        log("[bench] cpu baseline ...")
        line["cpu_baseline"] = cpu_saba_baseline(g, args.depth, args.walks,
                                                 args.cpu_sample_sources)
        line["gpu_over_cpu"] = value / line["cpu_baseline"]["value"]
    del g
    if not args.no_tgtp:
        line["tgtp"] = bench_tgtp(args)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
