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
objects in the line: `saba_depth15` (the same graph at 15 hops), the host
CPU baseline of the walks (AVX2 + OpenMP restatement), `mt_walks` (the
MT19937 comparator walks), `tgtp` (n=4096 Chung-Lu, FAITHFUL and FAST),
`tgtp_batch` (256 graphs, batched vs plain, with the round kernel's
roofline) and `tgtp_precision` (the 512 x 512 rewired grid on sampled
sources).

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu] [--no-mt]
                    [--no-tgtp] [--no-batch] [--no-precision] [--no-depth15]
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


def record_bytes_per_transition(depth: int, gen_hops: int = 2) -> float:
    """Walker-record bytes the partitioned engine moves per transition: the
    generation round (gen_hops hops) writes the records of the next round,
    each later round reads its records and writes the next round's, once.
    A record into the round that takes hop h holds 3 words {vertex, secret,
    walker id}, 2 {vertex, secret} into the round of the last hop and 1 (the
    vertex) into the last round, which only counts arrivals."""
    g = min(gen_hops, depth)

    def words_into(h: int) -> int:
        return 1 if h == depth else 2 if h + 1 == depth else 3

    return sum(2 * 4 * words_into(h) for h in range(g, depth + 1)) / depth


T_START = time.perf_counter()


def log(msg: str) -> None:
    print(f"{msg} [t={time.perf_counter() - T_START:.0f}s]", file=sys.stderr, flush=True)


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


def bench_saba(args, g=None, depth=None, steps=None, warmup=None):
    """SABA walk passes on the n=2^20 graph with the CSR resident in HBM.
    Returns (value, ms_per_step, roofline, config, graph)."""
    from tgtp_amd import Context, workloads as WL

    depth = args.depth if depth is None else depth
    steps = args.steps if steps is None else steps
    warmup = args.warmup if warmup is None else warmup
    n_walks = args.walks
    if g is None:
        t0 = time.time()
        g = WL.chung_lu_1m()
        log(f"[bench] graph n={g.n} m={g.m} generated in {time.time() - t0:.1f}s")
    ctx = Context(0)
    dg = ctx.upload(g)
    trans = g.n * n_walks * depth  # reference definition, all n vertices
    for i in range(warmup):
        _, vc, cs, st = ctx.saba_walks(dg, depth, n_walks, want_counters=False)
        log(f"[bench] depth {depth} warmup {i}: {st['kernel_ms']:.1f} ms checksum={cs}")
    walk_ms = []
    t_start = time.perf_counter()
    for _ in range(steps):
        _, vc, cs, st = ctx.saba_walks(dg, depth, n_walks, want_counters=False)
        walk_ms.append(st["walk_ms"])
    elapsed = time.perf_counter() - t_start  # each call ends with a stream synchronize
    dg.free()
    ctx.close()
    ms_per_step = 1e3 * elapsed / steps
    value = trans / (elapsed / steps)
    avg_walk_s = sum(walk_ms) / len(walk_ms) / 1e3
    achieved = BYTES_PER_TRANSITION * trans / avg_walk_s / 1e9
    workload = (f"saba_walks chung-lu n=2^20 beta=2.2 m={g.m} depth={depth} "
                f"n_walks={n_walks}")
    log(f"[bench] depth {depth}, {steps} steps: {ms_per_step:.2f} ms/step, walk kernel "
        f"{avg_walk_s * 1e3:.2f} ms, {value:.3e} transitions/s, checksum={cs}")
    # The library walks graphs whose rows and adjacency exceed 64 MiB with the
    # partitioned engine (saba_part.hip: per source chunk a two-hop
    # generation round, depth-2 middle rounds and a counting round of
    # part_round_kernel, plus the fallback replay) unless TGTP_WALK_PART=0;
    # smaller graphs with the direct kernel (saba_walk.hip: saba_walk_kernel
    # + bucket_count_kernel per chunk).  The unit of "launch" here is the
    # pass, timed over all its kernels with HIP events on the library's
    # stream.
    part = (os.environ.get("TGTP_WALK_PART", "-1") != "0"
            and 4 * g.m + 8 * g.n >= 64 * 2**20)
    if part:
        kernel = ("part_round_kernel<0|1|2> + part_fixup_kernel + block listing (one pass; "
                  "the middle rounds part_round_kernel<1> are 84-88% of it)")
        traffic = traffic_for("saba_part_pass", f"cl1m_d{depth}_w{n_walks}")
    else:
        kernel = "saba_walk_kernel + bucket_count_kernel (one pass)"
        traffic = traffic_for("saba_walk_pass", f"cl1m_d{depth}_w{n_walks}")
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "engine": "partitioned" if part else "direct", "kernel": kernel,
            "kernel_ms": round(avg_walk_s * 1e3, 3),
            "bytes_per_unit": BYTES_PER_TRANSITION, "units_per_launch": trans}
    cfg = {"workload": workload, "graph": WL.describe(WL.CL1M), "n": g.n, "m": g.m,
           "depth": depth, "n_walks": n_walks, "visit_count": vc, "checksum": cs}
    return value, ms_per_step, roof, cfg, g


def stream_copy_rate():
    """Device-to-device copy rate (bytes read + written per second, GB/s) of
    a 4 GiB buffer, best of 5, timed with events on the copy's stream."""
    import torch

    a = torch.empty(1 << 30, dtype=torch.int32, device="cuda:0")
    a.fill_(1)
    b = torch.empty_like(a)
    s = torch.cuda.current_stream()
    best = None
    for _ in range(6):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s)
        b.copy_(a)
        e1.record(s)
        e1.synchronize()
        ms = e0.elapsed_time(e1)
        best = ms if best is None else min(best, ms)
    del a, b
    torch.cuda.empty_cache()
    return 2 * (4 << 30) / (best / 1e3) / 1e9


def gather_roof(mib: float):
    """The binding roof of the walk's memory pattern, measured on this GPU by
    tgtp_amd/bin/gather_roof (a child process): chains of dependent random
    8-byte loads over a buffer of `mib` MiB, best over 1-8 waves per SIMD and
    1-4 chains per lane.  Returns loads/s (None if the tool is missing)."""
    import subprocess

    tool = ROOT / "tgtp_amd" / "bin" / "gather_roof"
    if not tool.exists():
        return None
    r = subprocess.run([str(tool), "--mib", f"{mib:.0f}", "--hops", "256", "--reps", "3"],
                       capture_output=True, text=True, timeout=120)
    if r.returncode != 0:
        log(f"[bench] gather_roof failed: {r.stderr.strip()[:200]}")
        return None
    best = 0.0
    for line in r.stdout.splitlines():
        f = line.split()
        if len(f) == 4 and f[0].isdigit():
            best = max(best, float(f[2]))
    return best or None


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


def reference_algorithm_sampled(graphs, eigens, samples, exact, eps):
    """The reference algorithm's own error on the sampled edges of every
    graph: the oracle (bit-exact to the reference build, reference tau rule)
    runs the rows of both endpoints of each sampled edge, score = row(pos) +
    row(reverse position) as aesc.h:532-550 sums them."""
    import numpy as np

    import oracle

    threads = host_threads()
    est = []
    for g, (lam, X), p in zip(graphs, eigens, samples):
        rows_of = np.repeat(np.arange(g.n, dtype=np.uint32), np.diff(g.idx.astype(np.int64)))
        src = np.unique(np.concatenate([rows_of[p], g.adj[p]]))
        taus = oracle.taus(g.idx, g.adj, lam, X, np.float32(eps / 2), threads=threads)
        rows = oracle.tgtp_sources(g.idx, g.adj, taus, eps, 10, src, threads=threads)
        est.append(rows[p] + rows[g.reverse_index()[p]])
    return error_stats(np.concatenate(est), np.concatenate(exact), eps)


def bench_batch(args):
    """Batched tgtp on 256 seeded Chung-Lu graphs (n log-uniform in
    [1024, 65536], exponent U(2.1, 3.0), mean degree U(8, 40)), omega=128,
    gamma=10: one autotuned, graph-captured batch run vs one plain
    tgtp_score_dev call per graph, both with graphs and eigenpairs resident
    in HBM and scores copied back to the host, each side the minimum of 3
    timed runs (the maximum is reported too).  edges/s = total CSR entries /
    seconds.  Errors are against FP64 exact resistances of 32 sampled edges
    per graph (block PCG, tgtp_amd.exact_sc), beside the reference
    algorithm's own error on the same edges."""
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
    cache = Path(args.eig_cache) if args.eig_cache else None
    if cache is not None and cache.exists():
        # eigenpairs of an earlier run of this workload (profiling runs: no
        # torch work in the profiled process)
        with np.load(cache) as z:
            lam_all, X_all = z["lam"], z["X"]
        eig, off = [], 0
        for k, g in enumerate(graphs):
            eig.append((lam_all[k], np.ascontiguousarray(X_all[off:off + g.n])))
            off += g.n
    else:
        eig = [E.eigens_f32(*e) for e in E.compute_eigens_many(
            graphs, 128, device="cuda:0",
            progress=lambda done, tot: log(f"[bench] batch: eigenpairs {done}/{tot} "
                                           f"({time.time() - t0:.1f}s)"))]
        if cache is not None:
            np.savez(cache, lam=np.stack([lam for lam, _ in eig]),
                     X=np.concatenate([X for _, X in eig]))
    t_eig = time.time() - t0
    log(f"[bench] batch: eigenpairs (batched block Krylov, FP64 on the GPU) in {t_eig:.1f}s")
    t0 = time.time()
    rng = np.random.default_rng(7)
    samples = [np.sort(rng.choice(g.m, size=min(args.batch_sample_edges, g.m), replace=False))
               for g in graphs]
    exact = None
    if args.batch_exact:
        exact = [exact_sc.sampled_scores(g, p, "cuda:0") for g, p in zip(graphs, samples)]
        torch.cuda.empty_cache()
        log(f"[bench] batch: exact resistances of {sum(len(p) for p in samples)} sampled edges "
            f"in {time.time() - t0:.1f}s")
    out = {"workload": "256 Chung-Lu graphs " + WL.describe(
               {k: v for k, v in WL.BATCH256.items() if k != "count"}),
           "graphs": len(graphs), "sum_n": sum(g.n for g in graphs), "sum_m": total_m,
           "omega": 128, "gamma": 10, "eigen_prep_s": round(t_eig, 1),
           "tau_rule": "bounded (FAST default; the reference's rule leaves eps on bridge edges)"}
    reps = 3
    fast_out = None
    roof = None
    with Context(0) as ctx:
        modes = [(m, nm) for m, nm in ((MODE_FAST_FP32, "fast_fp32"), (MODE_FAST_BF16, "fast_bf16"))
                 if nm[5:] in args.batch_modes.split(",")]
        for eps in [float(x) for x in args.batch_eps.split(",")]:
            leg = {}
            for mode, name in modes:
                plain_min, plain_max, plain_out, plain_each = 0.0, 0.0, [], []
                for gi, (g, e) in enumerate(zip(graphs, eig) if args.batch_plain else []):
                    if gi % 64 == 63:
                        log(f"[bench] batch eps={eps} {name}: plain {gi + 1}/{len(graphs)}")
                    dg = ctx.upload(g)
                    ctx.score(dg, *e, eps, 10, mode=mode)  # eigen upload + plan
                    ts = []
                    for _ in range(reps):
                        t1 = time.perf_counter()
                        s, _ = ctx.score(dg, *e, eps, 10, mode=mode)
                        ts.append(time.perf_counter() - t1)
                    plain_min += min(ts)
                    plain_max += max(ts)
                    plain_each.append(min(ts))
                    plain_out.append(s)
                    dg.free()
                if not args.batch_plain:
                    plain_min = plain_max = float("nan")
                t1 = time.perf_counter()
                b = ctx.batch(graphs, eig, eps, 10, mode=mode, autotune=True,
                              graph=bool(args.batch_graph))
                t_create = time.perf_counter() - t1
                outs, st = b.run()
                times, kms = [], []
                for _ in range(reps):
                    t1 = time.perf_counter()
                    outs, st = b.run()
                    times.append(time.perf_counter() - t1)
                    kms.append(st["kernel_ms"])
                cfg = b.config()
                if mode == MODE_FAST_FP32 and eps == 0.01:
                    prof = b.profile()
                    avg_ms = prof["round_ms"] / max(1, prof["round_launches"])
                    per_launch = prof["round_bytes"] / max(1, prof["round_launches"])
                    achieved = prof["round_bytes"] / (prof["round_ms"] / 1e3) / 1e9
                    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                            "traffic": traffic_for("blk_round4_kernel", "batch256_eps0.01"),
                            "kernel": "blk_round4_kernel<16,4> (diffusion round, FAST_FP32)",
                            "kernel_ms": round(avg_ms, 4), "launches": prof["round_launches"],
                            "bytes_per_launch": round(per_launch),
                            "work_unit": "per source-round: 4 B per edge + 4 B per vertex of its "
                                         "graph; per block-round: 4 B per edge + 8 B per vertex",
                            "profile": prof}
                    log(f"[bench] batch eps=0.01 fast_fp32 round kernel: {prof['round_launches']} "
                        f"launches, {avg_ms:.3f} ms avg, {achieved:.0f} GB/s algorithmic")
                b.close()
                if mode == MODE_FAST_FP32:
                    fast_out = outs
                best, worst = min(times), max(times)
                bit_equal = (all(np.array_equal(a.view(np.uint32), p.view(np.uint32))
                                 for a, p in zip(outs, plain_out)) if args.batch_plain else None)
                est = np.concatenate([o[p] for o, p in zip(outs, samples)])
                errs = (error_stats(est, np.concatenate(exact), eps) if exact is not None
                        else None)
                leg[name] = {"batched_edges_per_s": total_m / best, "batched_seconds": best,
                             "batched_seconds_max": worst, "batched_kernel_ms": min(kms),
                             "plain_edges_per_s": total_m / plain_min,
                             "plain_seconds": plain_min, "plain_seconds_max": plain_max,
                             "batched_over_plain": plain_min / best,
                             "batched_over_plain_range": [plain_min / worst, plain_max / best],
                             "timing": f"min (and max) of {reps} runs per side",
                             "create_autotune_capture_s": round(t_create, 2),
                             "autotuned": cfg, "bitwise_equal_plain": bit_equal,
                             "vs_exact_fp64_sampled": errs}
                log(f"[bench] batch eps={eps} {name}: batched {total_m / best:.3e} edges/s "
                    f"({best * 1e3:.1f}-{worst * 1e3:.1f} ms), plain {total_m / plain_min:.3e} "
                    f"({plain_min * 1e3:.1f}-{plain_max * 1e3:.1f} ms), x{plain_min / best:.3f}, "
                    f"cfg {cfg}, max err {errs['max_abs'] if errs else float('nan'):.4f}")
                if mode == MODE_FAST_FP32 and args.batch_plain:
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
            if not args.no_cpu and exact is not None:
                t1 = time.time()
                leg["reference_algorithm_vs_exact_fp64_sampled"] = reference_algorithm_sampled(
                    graphs, eig, samples, exact, eps)
                log(f"[bench] batch eps={eps}: reference algorithm on every sampled edge, max err "
                    f"{leg['reference_algorithm_vs_exact_fp64_sampled']['max_abs']:.4f} "
                    f"({time.time() - t1:.1f}s)")
                every = max(1, len(graphs) // args.batch_cpu_graphs)
                sub = list(range(0, len(graphs), every))
                leg["cpu_baseline"] = cpu_tgtp_baseline(
                    [graphs[i] for i in sub], [eig[i] for i in sub], eps,
                    sampled=[(samples[i], exact[i]) for i in sub], reps=reps)
                leg["cpu_baseline"]["sample"] += f"; every {every}th graph of the batch"
                if fast_out is not None:
                    fe = np.concatenate([fast_out[i][samples[i]] for i in sub])
                    leg["fast_fp32_vs_exact_fp64_same_graphs"] = error_stats(
                        fe, np.concatenate([exact[i] for i in sub]), eps)
                log(f"[bench] batch eps={eps} cpu {leg['cpu_baseline']['edges_per_s']:.3e} edges/s")
            out[f"eps_{eps}"] = leg
    if roof is not None:
        out["roofline"] = roof
    return out


def bench_precision(args):
    """Precision form of tgtp: the 512 x 512 grid with 5% of edges rewired
    (n = 262,144, m = 1,046,528), omega = 128, gamma = 10, eps 0.01 and
    0.005, FAST_FP32 and FAST_BF16 (projection in FP32 / BF16 with FP32
    accumulation, FP64 score sums).  A full run takes ~3 minutes per leg, so
    a seeded uniform sample of source ranges (4 x 64 sources) runs through
    tgtp_score_sources; per-source work is independent (aesc.h:456-527), so
    sources/s extrapolates to the whole graph.  Errors: max / mean relative
    error of the edges whose endpoints are both sampled against FP64 exact
    resistances (block PCG).  CPU leg: the oracle (the reference algorithm)
    on the same sources, eps 0.01.  --precision-full adds full-graph runs."""
    import numpy as np
    import torch

    import oracle
    from tgtp_amd import MODE_FAST_BF16, MODE_FAST_FP32, Context, eigenprep as E, exact_sc
    from tgtp_amd import workloads as WL

    side = args.precision_side
    g = WL.G.grid_rewired(side, side, 0.05, seed=1)
    t0 = time.time()
    lam, X = E.eigens_f32(*E.compute_eigens_many([g], 128, device="cuda:0")[0])
    t_eig = time.time() - t0
    rng = np.random.default_rng(11)
    n_rng = max(1, args.precision_sources // 64)
    starts = np.sort(rng.choice(g.n // 64, size=n_rng, replace=False)) * 64
    ranges = [(int(x), int(x) + 64) for x in starts]
    mask = np.zeros(g.n, dtype=bool)
    for b, e in ranges:
        mask[b:e] = True
    rows_of = np.repeat(np.arange(g.n), np.diff(g.idx.astype(np.int64)))
    both = np.nonzero(mask[rows_of] & mask[g.adj])[0]
    R = exact_sc.sampled_scores(g, both, "cuda:0")
    torch.cuda.empty_cache()
    n_src = int(mask.sum())
    log(f"[bench] precision grid {side}x{side}: n={g.n} m={g.m}, eigenpairs {t_eig:.1f}s, "
        f"{n_src} sampled sources, {both.shape[0]} edges with both ends sampled")
    out = {"workload": f"grid {side}x{side}, 5% rewired, seed 1", "n": g.n, "m": g.m,
           "omega": 128, "gamma": 10, "eigen_prep_s": round(t_eig, 1),
           "sample": f"{n_rng} seeded ranges of 64 sources ({n_src} sources), "
                     f"{both.shape[0]} edges with both endpoints sampled",
           "source_ranges": ranges}
    with Context(0) as ctx:
        dg = ctx.upload(g)
        for eps in (0.01, 0.005):
            leg = {}
            for mode, name in ((MODE_FAST_FP32, "fast_fp32"), (MODE_FAST_BF16, "fast_bf16")):
                ctx.score_sources(dg, lam, X, eps, ranges, mode=mode)
                best, sc = None, None
                for _ in range(3):
                    t1 = time.perf_counter()
                    _, sc, st = ctx.score_sources(dg, lam, X, eps, ranges, mode=mode)
                    dt = time.perf_counter() - t1
                    best = dt if best is None else min(best, dt)
                err = np.abs(sc[both].astype(np.float64) - R)
                leg[name] = {"sources_per_s": n_src / best, "seconds": best,
                             "kernel_ms": st["kernel_ms"], "taus_ms": st["taus_ms"],
                             "extrapolated_full_s": g.n / (n_src / best),
                             "vs_exact_fp64_sampled": error_stats(sc, R, eps, both)}
                e = leg[name]["vs_exact_fp64_sampled"]
                log(f"[bench] precision eps={eps} {name}: {best * 1e3:.0f} ms for {n_src} sources, "
                    f"max rel {e['max_rel']:.2e} mean rel {e['mean_rel']:.2e} max abs "
                    f"{e['max_abs']:.2e}")
                if args.precision_full:
                    t1 = time.perf_counter()
                    s_full, _ = ctx.score(dg, lam, X, eps, 10, mode=mode)
                    dtf = time.perf_counter() - t1
                    leg[name]["full_graph"] = {"edges_per_s": g.m / dtf, "seconds": dtf,
                                               "vs_exact_fp64_sampled": error_stats(s_full, R,
                                                                                    eps, both)}
            if eps == 0.01 and not args.no_cpu:
                src = np.nonzero(mask)[0].astype(np.uint32)
                threads = host_threads()
                taus = oracle.taus(g.idx, g.adj, lam, X, np.float32(eps / 2), threads=threads)
                t1 = time.perf_counter()
                oracle.tgtp_sources(g.idx, g.adj, taus, eps, 10, src, threads=threads)
                dtc = time.perf_counter() - t1
                leg["cpu_baseline"] = {"sources_per_s": n_src / dtc, "seconds": dtc,
                                       "extrapolated_full_s": g.n / (n_src / dtc),
                                       "cores": threads, "kind": "port",
                                       "sample": f"the same {n_src} sources (oracle C++ OpenMP, "
                                                 "schedule(dynamic) over sources)"}
                log(f"[bench] precision eps={eps} cpu: {n_src / dtc:.1f} sources/s on {threads} "
                    f"threads (full graph ~{g.n / (n_src / dtc):.0f} s)")
            out[f"eps_{eps}"] = leg
        dg.free()
    return out


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_saba_baseline(g, depth, n_walks):
    """The AVX2 + OpenMP SABA restatement of the oracle (bouquets of 8 walkers
    in one AVX2 register, per-thread counters, equal to the scalar oracle bit
    for bit) on evenly spread source slices of the same workload: one thread,
    then all host threads with the reference's schedule(dynamic, 1024)
    (rw_experiment_saba.cpp:194) and with a chunk of 16.  value = the best."""
    import numpy as np

    import oracle

    threads = host_threads()
    step = g.n // 32

    def timed(per, th, chunk, slices=32):
        # evenly spread slices of `per` consecutive sources
        t0 = time.perf_counter()
        total = 0
        stride = g.n // slices
        for k in range(slices):
            lo = k * stride
            hi = min(g.n, lo + per)
            oracle.saba_walks_avx2(g.idx, g.adj, depth, n_walks, lo, hi, threads=th, chunk=chunk)
            total += (hi - lo) * n_walks * depth
        dt = time.perf_counter() - t0
        return total / dt, total, dt

    # a source slice checked against the scalar oracle first
    a = oracle.saba_walks_avx2(g.idx, g.adj, depth, n_walks, 0, 64, threads=threads)
    b = oracle.saba_walks(g.idx, g.adj, depth, n_walks, 0, 64, threads=threads)
    legs = {}
    v1, t1, d1 = timed(16, 1, 1024)
    legs["avx2_1thread"] = {"transitions_per_s": v1, "sources": 32 * 16, "seconds": round(d1, 2)}
    # the reference's chunk of 1024 sources needs slices of many chunks to
    # keep every thread busy: 2 slices of 16 chunks (the same 32,768 sources)
    vr, tr, dr = timed(16384, threads, 1024, slices=2)
    legs[f"avx2_{threads}threads_dynamic1024"] = {"transitions_per_s": vr, "sources": 2 * 16384,
                                                  "seconds": round(dr, 2)}
    vt, tt, dt = timed(1024, threads, 16)
    legs[f"avx2_{threads}threads_dynamic16"] = {"transitions_per_s": vt, "sources": 32 * 1024,
                                                "seconds": round(dt, 2)}
    best = max(vr, vt)
    log(f"[bench] cpu walks (AVX2): 1 thread {v1:.3e}, {threads} threads {vr:.3e} (chunk 1024) / "
        f"{vt:.3e} (chunk 16)")
    return {"value": best, "unit": "transitions/s", "cores": threads, "kind": "port",
            "sample": f"32,768 of {g.n} sources (2 slices of 16,384 for dynamic,1024; 32 "
                      f"evenly spread slices of 1,024 for dynamic,16), depth {depth}, {n_walks} "
                      f"walks/vertex = {tt:.3e} transitions; AVX2 + OpenMP restatement of the "
                      f"oracle (orc_saba_walks_avx2), best of the two schedules on {threads} "
                      f"threads",
            "cpu_model": cpu_model(), "compiler": "g++ -std=c++20 -O2 -fopenmp, AVX2 target",
            "equal_scalar_oracle": bool(np.array_equal(a, b)), "legs": legs}


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
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tgtp", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-saba", action="store_true", help="tgtp legs only (profiling)")
    ap.add_argument("--no-mt", action="store_true", help="skip the MT19937 comparator walks")
    ap.add_argument("--no-precision", action="store_true",
                    help="skip the 512x512 rewired-grid precision leg (sampled sources)")
    ap.add_argument("--precision-full", action="store_true",
                    help="also score the whole 512x512 grid in each precision leg (minutes)")
    ap.add_argument("--precision-side", type=int, default=512)
    ap.add_argument("--precision-sources", type=int, default=256)
    ap.add_argument("--no-depth15", action="store_true", help="skip the 15-hop walk leg")
    ap.add_argument("--batch-count", type=int, default=256)
    ap.add_argument("--batch-sample-edges", type=int, default=32)
    ap.add_argument("--batch-cpu-graphs", type=int, default=16)
    ap.add_argument("--batch-small-n", type=int, default=8192)
    ap.add_argument("--batch-eps", default="0.05,0.01", help="comma-separated eps of the batch legs")
    ap.add_argument("--batch-modes", default="fp32,bf16", help="batch modes: fp32, bf16")
    ap.add_argument("--batch-graph", type=int, default=1,
                    help="0: run the batch uncaptured (counter collection runs)")
    ap.add_argument("--batch-exact", type=int, default=1,
                    help="0: skip the FP64 exact resistances of the sampled edges")
    ap.add_argument("--eig-cache", default="",
                    help="load the batch's eigenpairs from this .npz if it exists, else save them")
    ap.add_argument("--batch-plain", type=int, default=1,
                    help="0: skip the per-graph plain runs (profiling the batched run alone)")
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
        if not args.no_precision:
            line["tgtp_precision"] = bench_precision(args)
        print(json.dumps(line), flush=True)
        return 0
    value, ms_per_step, roof, cfg, g = bench_saba(args)
    roof_lps = gather_roof(g.m * 8 / 2**20)
    if roof_lps:
        # the direct kernel's bound: one random cache line per hop
        roof["random_gather_roof"] = {
            "what": "dependent random 8-byte loads over a buffer of the packed adjacency's "
                    f"size ({g.m * 8 / 2**20:.0f} MiB), best over 1-8 waves/SIMD and 1-4 "
                    "chains per lane (tgtp_amd/bin/gather_roof, measured in this run); "
                    "the direct kernel runs at 1.1-1.2x this rate",
            "loads_per_s": roof_lps, "transitions_per_s": value, "frac": round(value / roof_lps, 4)}
        log(f"[bench] random-gather roof {roof_lps:.3e} loads/s; walk at "
            f"{value / roof_lps:.2f} of it")
    copy_gbs = stream_copy_rate()
    if copy_gbs:
        # the partitioned engine's bound: each walker's record is written and
        # read once per round (12 bytes; 8 and 4 into the last two rounds)
        rb = record_bytes_per_transition(cfg["depth"])
        bound = copy_gbs * 1e9 / rb
        roof["record_stream_roof"] = {
            "what": "a 4 GiB device-to-device copy measured in this run (read + write bytes "
                    f"per second) over the partitioned engine's {rb:g} bytes of records per "
                    "transition (12-byte records read and written once per round, the two-hop "
                    "generation round writing one, 8-byte {vertex, secret} records into the "
                    "round of the last hop, 4-byte vertices into the last round)",
            "copy_GBps": round(copy_gbs, 1), "bytes_per_transition": rb,
            "transitions_per_s": bound, "frac": round(value / bound, 4)}
        log(f"[bench] record-stream roof {bound:.3e} transitions/s ({copy_gbs:.0f} GB/s copy); "
            f"walk at {value / bound:.2f} of it")
    line = {"metric": METRIC, "value": value, "unit": "transitions/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
            "data": "synthetic (seeded Chung-Lu generator, regenerated at run time)",
            "config": cfg, "roofline": roof}
    if not args.no_depth15:
        v15, ms15, roof15, cfg15, _ = bench_saba(args, g=g, depth=15, steps=3, warmup=1)
        line["saba_depth15"] = {"value": v15, "unit": "transitions/s", "ms_per_step": ms15,
                                "steps": 3, "warmup": 1, "config": cfg15, "roofline": roof15}
    if not args.no_cpu:
        log("[bench] cpu baseline ...")
        line["cpu_baseline"] = cpu_saba_baseline(g, args.depth, args.walks)
        line["gpu_over_cpu"] = value / line["cpu_baseline"]["value"]
    if not args.no_mt:
        line["mt_walks"] = bench_mt(args, g)
    del g
    if not args.no_tgtp:
        line["tgtp"] = bench_tgtp(args)
    if not args.no_batch:
        line["tgtp_batch"] = bench_batch(args)
    if not args.no_precision:
        line["tgtp_precision"] = bench_precision(args)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
