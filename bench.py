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
(rw_experiment_saba.cpp:214-215).  One JSON line on stdout.  Further
objects in the line: the host CPU baseline of the walks, `mt_walks` (the
MT19937 comparator walks), `tgtp` (n=4096 Chung-Lu, FAITHFUL and FAST),
`tgtp_batch` (256 graphs, batched vs plain) and, with --precision, the
512 x 512 rewired grid.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu] [--no-mt]
                    [--no-tgtp] [--no-batch] [--precision]
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
    # one walk pass = source chunks x (saba_walk_kernel, bucket_count_kernel);
    # the unit of "launch" here is the pass, timed over all its kernels
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic_for("saba_walk_pass", "cl1m_d10_w16384"),
            "kernel": "saba_walk_kernel + bucket_count_kernel (one pass)",
            "kernel_ms": round(avg_walk_s * 1e3, 3),
            "bytes_per_unit": BYTES_PER_TRANSITION, "units_per_launch": trans}
    cfg = {"workload": workload, "graph": WL.describe(WL.CL1M), "n": g.n, "m": g.m,
           "depth": depth, "n_walks": n_walks, "visit_count": vc, "checksum": cs}
    return value, ms_per_step, roof, cfg, g


def bench_tgtp(args):
    """tgtp() on the n=4096 Chung-Lu graph (omega=128, gamma=10): edges/s =
    CSR entries / seconds of one call with graph and eigenpairs resident in
    HBM (benchmark.cpp:40-42 times tgtp() only).  Errors are against the
    FP64 exact scores (tgtp_amd.exact_sc, dense grounded-Laplacian
    Cholesky).  Returns a dict of legs."""
    import numpy as np

    from tgtp_amd import (MODE_FAITHFUL, MODE_FAST_BF16, MODE_FAST_FP32, Context,
                          eigenprep as E, exact_sc, workloads as WL)

    g = WL.chung_lu_4096()
    t0 = time.time()
    lam, X = E.eigens_f32(*E.compute_eigens(g, 128))
    R = exact_sc.exact_scores(g, "cuda:0")
    log(f"[bench] tgtp graph n={g.n} m={g.m}, eigenpairs + exact scores in {time.time() - t0:.1f}s")
    out = {"graph": WL.describe(WL.CL4096), "n": g.n, "m": g.m, "omega": 128, "gamma": 10}
    modes = ((MODE_FAITHFUL, "faithful"), (MODE_FAST_FP32, "fast_fp32"),
             (MODE_FAST_BF16, "fast_bf16"))
    with Context(0) as ctx:
        dg = ctx.upload(g)
        for eps in (0.01, 0.05):
            leg = {}
            faithful_scores = None
            for mode, name in modes:
                s, st = ctx.score(dg, lam, X, eps, 10, mode=mode)  # warmup + eigen upload
                times, kms = [], []
                for _ in range(max(3, args.steps)):
                    t1 = time.perf_counter()
                    s, st = ctx.score(dg, lam, X, eps, 10, mode=mode)
                    times.append(time.perf_counter() - t1)
                    kms.append(st["kernel_ms"])
                if mode == MODE_FAITHFUL:
                    faithful_scores = s
                best = min(times)
                leg[name] = {"edges_per_s": g.m / best, "seconds": best,
                             "kernel_ms": min(kms), "taus_ms": st["taus_ms"],
                             "diffuse_ms": st["diffuse_ms"], "host_fixups": st["host_fixups"],
                             "vs_exact_fp64": error_stats(s, R, eps)}
            # PCIe-inclusive: host arrays in, scores out (tgtp_score)
            t1 = time.perf_counter()
            ctx.score(g, lam, X, eps, 10)
            leg["faithful"]["e2e_with_h2d_edges_per_s"] = g.m / (time.perf_counter() - t1)
            if not args.no_cpu:
                leg["cpu_baseline"] = cpu_tgtp_baseline([g], [(lam, X)], eps, [faithful_scores])
            out[f"eps_{eps}"] = leg
            log(f"[bench] tgtp eps={eps}: " + ", ".join(
                f"{nm} {leg[nm]['edges_per_s']:.3e} edges/s (max err {leg[nm]['vs_exact_fp64']['max_abs']:.4f})"
                for _, nm in modes) + (f"; cpu {leg['cpu_baseline']['edges_per_s']:.3e}"
                                       if "cpu_baseline" in leg else ""))
        dg.free()
    return out


def _timed(fn) -> float:
    t1 = time.perf_counter()
    fn()
    return time.perf_counter() - t1


def error_stats(scores, exact, eps, pos=None):
    import numpy as np

    s = scores.astype(np.float64) if pos is None else scores[pos].astype(np.float64)
    err = np.abs(s - exact)
    rel = err / exact
    return {"max_abs": float(err.max()), "mean_abs": float(err.mean()),
            "max_rel": float(rel.max()), "mean_rel": float(rel.mean()),
            "within_eps": bool(err.max() <= eps), "edges_checked": int(err.shape[0])}


def cpu_tgtp_baseline(graphs, eigens, eps, gpu_faithful=None, sampled=None, reps=3):
    """The oracle's tgtp (C++ OpenMP over sources, the reference's
    algorithm) on the host cores, min of `reps` per graph; when the GPU
    faithful scores are given, whether they equal the oracle's bit for bit;
    when (positions, exact) per graph are given, the reference algorithm's
    own error on those edges."""
    import numpy as np

    import oracle

    threads = host_threads()
    total_m, total_s, equal = 0, 0.0, True
    est, ex = [], []
    for k, (g, (lam, X)) in enumerate(zip(graphs, eigens)):
        best = None
        for _ in range(reps):
            t1 = time.perf_counter()
            ref = oracle.tgtp(g.idx, g.adj, lam, X, eps, 10, threads=threads)
            dt = time.perf_counter() - t1
            best = dt if best is None else min(best, dt)
        if gpu_faithful is not None and gpu_faithful[k] is not None:
            equal &= bool(np.array_equal(ref.view(np.uint32), gpu_faithful[k].view(np.uint32)))
        if sampled is not None:
            est.append(ref[sampled[k][0]])
            ex.append(sampled[k][1])
        total_m += g.m
        total_s += best
    out = {"edges_per_s": total_m / total_s, "seconds": total_s, "cores": threads,
           "kind": "port",
           "sample": f"{len(graphs)} graph(s), {total_m} CSR entries, min of {reps} per graph "
                     "(oracle C++ OpenMP, schedule(dynamic) over sources)"}
    if gpu_faithful is not None:
        out["gpu_faithful_bitwise_equal"] = equal
    if sampled is not None:
        out["reference_algorithm_vs_exact_fp64_sampled"] = error_stats(
            np.concatenate(est), np.concatenate(ex), eps)
    return out


def bench_batch(args):
    """Batched tgtp on 256 seeded Chung-Lu graphs (n log-uniform in
    [1024, 65536], exponent U(2.1, 3.0), mean degree U(8, 40)), omega=128,
    gamma=10: one autotuned, graph-captured batch run vs one plain
    tgtp_score_dev call per graph, both with graphs and eigenpairs resident
    in HBM and scores copied back to the host.  edges/s = total CSR entries /
    seconds.  Errors are against FP64 exact resistances of sampled edges
    (block PCG, tgtp_amd.exact_sc)."""
    import numpy as np
    import torch

    from tgtp_amd import MODE_FAST_BF16, MODE_FAST_FP32, Context, eigenprep as E, exact_sc
    from tgtp_amd import workloads as WL

    t0 = time.time()
    graphs = WL.batch_graphs(args.batch_count)
    total_m = sum(g.m for g in graphs)
    log(f"[bench] batch: {len(graphs)} graphs, sum n={sum(g.n for g in graphs)} "
        f"sum m={total_m}, generated in {time.time() - t0:.1f}s")
    t0 = time.time()
    eig = [E.eigens_f32(*e) for e in E.compute_eigens_many(
        graphs, 128, device="cuda:0",
        progress=lambda done, tot: log(f"[bench] batch: eigenpairs {done}/{tot} "
                                       f"({time.time() - t0:.1f}s)"))]
    t_eig = time.time() - t0
    log(f"[bench] batch: eigenpairs (batched block Krylov, FP64 on the GPU) in {t_eig:.1f}s")
    t0 = time.time()
    rng = np.random.default_rng(7)
    samples = [np.sort(rng.choice(g.m, size=min(args.batch_sample_edges, g.m), replace=False))
               for g in graphs]
    exact = [exact_sc.sampled_scores(g, p, "cuda:0") for g, p in zip(graphs, samples)]
    torch.cuda.empty_cache()
    log(f"[bench] batch: exact resistances of {sum(len(p) for p in samples)} sampled edges "
        f"in {time.time() - t0:.1f}s")
    out = {"workload": "256 Chung-Lu graphs " + WL.describe(
               {k: v for k, v in WL.BATCH256.items() if k != "count"}),
           "graphs": len(graphs), "sum_n": sum(g.n for g in graphs), "sum_m": total_m,
           "omega": 128, "gamma": 10, "eigen_prep_s": round(t_eig, 1)}
    reps = 2
    fast_out = None
    with Context(0) as ctx:
        for eps in (0.05, 0.01):
            leg = {}
            for mode, name in ((MODE_FAST_FP32, "fast_fp32"), (MODE_FAST_BF16, "fast_bf16")):
                plain_s, plain_out, plain_each = 0.0, [], []
                for gi, (g, e) in enumerate(zip(graphs, eig)):
                    if gi % 64 == 63:
                        log(f"[bench] batch eps={eps} {name}: plain {gi + 1}/{len(graphs)}")
                    dg = ctx.upload(g)
                    ctx.score(dg, *e, eps, 10, mode=mode)  # eigen upload + plan
                    best = None
                    for _ in range(1):
                        t1 = time.perf_counter()
                        s, _ = ctx.score(dg, *e, eps, 10, mode=mode)
                        dt = time.perf_counter() - t1
                        best = dt if best is None else min(best, dt)
                    plain_s += best
                    plain_each.append(best)
                    plain_out.append(s)
                    dg.free()
                t1 = time.perf_counter()
                b = ctx.batch(graphs, eig, eps, 10, mode=mode, autotune=True, graph=True)
                t_create = time.perf_counter() - t1
                outs, st = b.run()
                times, kms = [], []
                for _ in range(reps):
                    t1 = time.perf_counter()
                    outs, st = b.run()
                    times.append(time.perf_counter() - t1)
                    kms.append(st["kernel_ms"])
                cfg = b.config()
                b.close()
                if mode == MODE_FAST_FP32:
                    fast_out = outs
                best = min(times)
                bit_equal = all(np.array_equal(a.view(np.uint32), p.view(np.uint32))
                                for a, p in zip(outs, plain_out))
                est = np.concatenate([o[p] for o, p in zip(outs, samples)])
                leg[name] = {"batched_edges_per_s": total_m / best, "batched_seconds": best,
                             "batched_kernel_ms": min(kms),
                             "plain_edges_per_s": total_m / plain_s, "plain_seconds": plain_s,
                             "batched_over_plain": plain_s / best,
                             "create_autotune_capture_s": round(t_create, 2),
                             "autotuned": cfg, "bitwise_equal_plain": bit_equal,
                             "vs_exact_fp64_sampled": error_stats(est, np.concatenate(exact),
                                                                  eps)}
                log(f"[bench] batch eps={eps} {name}: batched {total_m / best:.3e} edges/s "
                    f"({best * 1e3:.1f} ms), plain {total_m / plain_s:.3e} "
                    f"({plain_s * 1e3:.1f} ms), x{plain_s / best:.2f}, cfg {cfg}, "
                    f"max err {leg[name]['vs_exact_fp64_sampled']['max_abs']:.4f}")
                if mode == MODE_FAST_FP32:
                    # the batch's small graphs alone, where per-call costs dominate
                    small = [i for i, g in enumerate(graphs) if g.n <= args.batch_small_n]
                    if small:
                        sm = sum(graphs[i].m for i in small)
                        sp = sum(plain_each[i] for i in small)
                        b = ctx.batch([graphs[i] for i in small], [eig[i] for i in small], eps,
                                      10, mode=mode, autotune=True, graph=True)
                        b.run()
                        sb = min(_timed(b.run) for _ in range(reps))
                        b.close()
                        leg[name]["small_graphs"] = {
                            "graphs": len(small), "max_n": args.batch_small_n, "sum_m": sm,
                            "batched_edges_per_s": sm / sb, "plain_edges_per_s": sm / sp,
                            "batched_over_plain": sp / sb}
                        log(f"[bench] batch eps={eps} {name}: {len(small)} graphs with n <= "
                            f"{args.batch_small_n}: batched {sm / sb:.3e} edges/s, plain "
                            f"{sm / sp:.3e}, x{sp / sb:.2f}")
            if not args.no_cpu:
                every = max(1, len(graphs) // args.batch_cpu_graphs)
                sub = list(range(0, len(graphs), every))
                leg["cpu_baseline"] = cpu_tgtp_baseline(
                    [graphs[i] for i in sub], [eig[i] for i in sub], eps,
                    sampled=[(samples[i], exact[i]) for i in sub], reps=1)
                leg["cpu_baseline"]["sample"] += f"; every {every}th graph of the batch"
                fe = np.concatenate([fast_out[i][samples[i]] for i in sub])
                leg["fast_fp32_vs_exact_fp64_same_graphs"] = error_stats(
                    fe, np.concatenate([exact[i] for i in sub]), eps)
                log(f"[bench] batch eps={eps} cpu {leg['cpu_baseline']['edges_per_s']:.3e} edges/s; "
                    f"max err on the cpu graphs: reference algorithm "
                    f"{leg['cpu_baseline']['reference_algorithm_vs_exact_fp64_sampled']['max_abs']:.4f}, "
                    f"fast_fp32 {leg['fast_fp32_vs_exact_fp64_same_graphs']['max_abs']:.4f}")
            out[f"eps_{eps}"] = leg
    return out


def bench_precision(args):
    """Precision form of tgtp: the 512 x 512 grid with 5% of edges rewired
    (n = 262,144), omega = 128, gamma = 10, eps 0.01 and 0.005, FAST_FP32
    and FAST_BF16 (projection in FP32 / BF16 with FP32 accumulation, FP64
    score sums); max / mean relative error against FP64 exact resistances of
    sampled edges (block PCG).  No CPU leg: the reference algorithm needs
    ~2.4e13 pushed edges here (hours on the host)."""
    import numpy as np
    import torch

    from tgtp_amd import MODE_FAST_BF16, MODE_FAST_FP32, Context, eigenprep as E, exact_sc
    from tgtp_amd import workloads as WL

    os.environ["TGTP_PROGRESS"] = "1"  # calls of minutes: the library reports each subset
    side = args.precision_side
    g = WL.G.grid_rewired(side, side, 0.05, seed=1)
    t0 = time.time()
    lam, X = E.eigens_f32(*E.compute_eigens_many(
        [g], 128, device="cuda:0",
        on_cycle=lambda c, r: log(f"[bench] precision eigen cycle {c}: residual {r:.2e} "
                                  f"({time.time() - t0:.0f}s)") if c % 5 == 0 else None)[0])
    t_eig = time.time() - t0
    rng = np.random.default_rng(3)
    pos = np.sort(rng.choice(g.m, size=min(args.precision_sample_edges, g.m), replace=False))
    R = exact_sc.sampled_scores(g, pos, "cuda:0")
    torch.cuda.empty_cache()
    log(f"[bench] precision grid {side}x{side}: n={g.n} m={g.m}, eigenpairs {t_eig:.1f}s")
    out = {"workload": f"grid {side}x{side}, 5% rewired, seed 1", "n": g.n, "m": g.m,
           "omega": 128, "gamma": 10, "eigen_prep_s": round(t_eig, 1)}
    with Context(0) as ctx:
        dg = ctx.upload(g)
        for eps in (0.01, 0.005):
            leg = {}
            for mode, name in ((MODE_FAST_FP32, "fast_fp32"), (MODE_FAST_BF16, "fast_bf16")):
                t1 = time.perf_counter()
                s, st = ctx.score(dg, lam, X, eps, 10, mode=mode)
                dt = time.perf_counter() - t1
                leg[name] = {"edges_per_s": g.m / dt, "seconds": dt,
                             "taus_ms": st["taus_ms"], "diffuse_ms": st["diffuse_ms"],
                             "vs_exact_fp64_sampled": error_stats(s, R, eps, pos)}
                e = leg[name]["vs_exact_fp64_sampled"]
                log(f"[bench] precision eps={eps} {name}: {dt:.1f}s, max rel {e['max_rel']:.4f} "
                    f"mean rel {e['mean_rel']:.5f} max abs {e['max_abs']:.4f}")
            out[f"eps_{eps}"] = leg
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


def bench_mt(args, g_big=None):
    """MT19937 comparator walks (rw_experiment.cpp / rw_experiment_simd.cpp)
    on the GPU beside the oracle on the host: the n=4096 walk config in full
    (5 hops, 2048 walks, counters compared bit for bit) and, when given, 16
    evenly spread slices of 4096 sources of the n=2^20 graph (10 hops, 16384
    walks)."""
    import numpy as np

    import oracle
    from tgtp_amd import Context, WALK_MT_REGULAR, WALK_MT_SIMD, workloads as WL

    threads = host_threads()
    ctx = Context(0)
    out = {}
    g = WL.chung_lu_4096()
    dg = ctx.upload(g)
    leg = {"workload": f"chung-lu n=4096 m={g.m} depth=5 n_walks=2048, all sources"}
    for name, v in (("regular", WALK_MT_REGULAR), ("simd", WALK_MT_SIMD)):
        ctx.mt_walks(dg, 5, 2048, v)
        best, cnt, vc = 1e30, None, 0
        for _ in range(3):
            cnt, vc, cs, st = ctx.mt_walks(dg, 5, 2048, v)
            best = min(best, st["kernel_ms"])
        t0 = time.perf_counter()
        ref = oracle.mt_walks(g.idx, g.adj, 5, 2048, variant=v, threads=threads)
        tc = time.perf_counter() - t0
        leg[name] = {"gpu_transitions_per_s": vc / (best / 1e3), "gpu_ms": round(best, 3),
                     "cpu_transitions_per_s": vc / tc, "cpu_cores": threads,
                     "bitwise_equal_oracle": bool(np.array_equal(cnt, ref))}
        log(f"[bench] mt {name} n=4096: gpu {vc / (best / 1e3):.3e} transitions/s, "
            f"cpu ({threads} threads) {vc / tc:.3e}, equal={leg[name]['bitwise_equal_oracle']}")
    out["cl4096_d5_w2048"] = leg
    dg.free()
    if g_big is not None:
        dg = ctx.upload(g_big)
        step, per = g_big.n // 16, 4096
        leg = {"workload": f"chung-lu n=2^20 m={g_big.m} depth=10 n_walks=16384, "
                           f"16 slices x {per} sources"}
        for name, v in (("regular", WALK_MT_REGULAR), ("simd", WALK_MT_SIMD)):
            ms, total = 0.0, 0
            for k in range(16):
                _, vc, _, st = ctx.mt_walks(dg, 10, 16384, v, k * step, k * step + per,
                                           want_counters=False)
                ms += st["walk_ms"]
                total += vc
            leg[name] = {"gpu_transitions_per_s": total / (ms / 1e3), "gpu_walk_ms": round(ms, 1)}
            log(f"[bench] mt {name} n=2^20 sample: gpu {total / (ms / 1e3):.3e} transitions/s")
        out["cl1m_d10_w16384_sample"] = leg
        dg.free()
    ctx.close()
    return out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawTextHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--depth", type=int, default=10)
    ap.add_argument("--walks", type=int, default=16384)
    ap.add_argument("--cpu-sample-sources", type=int, default=32768)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tgtp", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-saba", action="store_true", help="tgtp legs only (profiling)")
    ap.add_argument("--no-mt", action="store_true", help="skip the MT19937 comparator walks")
    ap.add_argument("--precision", action="store_true",
                    help="add the 512x512 rewired-grid precision leg (minutes)")
    ap.add_argument("--precision-side", type=int, default=512)
    ap.add_argument("--precision-sample-edges", type=int, default=256)
    ap.add_argument("--batch-count", type=int, default=256)
    ap.add_argument("--batch-sample-edges", type=int, default=32)
    ap.add_argument("--batch-cpu-graphs", type=int, default=16)
    ap.add_argument("--batch-small-n", type=int, default=8192)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the multi-GPU "
              "form (sources sharded across GPUs, one all-reduce of the walk counters) "
              "is not built", file=sys.stderr)
        return 2
    if args.no_saba:  # profiling runs of the tgtp legs only; no headline value
        line = {"metric": METRIC, "value": None, "unit": "transitions/s", "n_gpus": 1}
        if not args.no_tgtp:
            line["tgtp"] = bench_tgtp(args)
        if not args.no_batch:
            line["tgtp_batch"] = bench_batch(args)
        if args.precision:
            line["tgtp_precision"] = bench_precision(args)
        print(json.dumps(line), flush=True)
        return 0
    value, ms_per_step, roof, cfg, g = bench_saba(args)
    line = {"metric": METRIC, "value": value, "unit": "transitions/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
            "data": "synthetic (seeded Chung-Lu generator, regenerated at run time)",
            "config": cfg, "roofline": roof}
    if not args.no_cpu:
        log("[bench] cpu baseline ...")
        line["cpu_baseline"] = cpu_saba_baseline(g, args.depth, args.walks,
                                                 args.cpu_sample_sources)
        line["gpu_over_cpu"] = value / line["cpu_baseline"]["value"]
    if not args.no_mt:
        line["mt_walks"] = bench_mt(args, g)
    del g
    if not args.no_tgtp:
        line["tgtp"] = bench_tgtp(args)
    if not args.no_batch:
        line["tgtp_batch"] = bench_batch(args)
    if args.precision:
        line["tgtp_precision"] = bench_precision(args)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
