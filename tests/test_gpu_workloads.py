"""The tgtp workloads at full size on the GPU, through the C ABI.

* Precision form: the 512 x 512 grid with 5% of edges rewired (n = 262,144,
  m = 1,046,528), omega = 128, gamma = 10.  A full run takes minutes, so
  four seeded ranges of 64 sources are scored with tgtp_score_sources:
  FAITHFUL rows equal the oracle's rows of the same sources bit for bit
  (eps 0.01); FAST_FP32 / FAST_BF16 scores of the edges whose endpoints are
  both sampled are within eps of the FP64 exact resistances at eps 0.01 and
  0.005.
* Batched form: 256 seeded Chung-Lu graphs, n log-uniform in [1024, 65536],
  exponent U(2.1, 3.0), mean degree U(8, 40), omega = 128, gamma = 10: the
  captured batch equals per-graph calls bit for bit on every graph, and on
  32 sampled edges per graph FAST (bounded tau rule) is within eps of the
  FP64 exact resistances at eps 0.05 and 0.01, where the reference's tau
  rule leaves eps on some bridge edges.
"""
import os
import sys

import numpy as np
import pytest

import oracle
from tgtp_amd import (MODE_FAITHFUL, MODE_FAST_BF16, MODE_FAST_FP32, TAU_REFERENCE, Context,
                      eigenprep as E, exact_sc, workloads as WL)

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _threads():
    try:
        return max(1, min(16, len(os.sched_getaffinity(0))))
    except AttributeError:
        return 8


@pytest.fixture(scope="module")
def grid512():
    g = WL.grid_512()
    lam, X = E.eigens_f32(*E.compute_eigens_many([g], 128, device="cuda:0")[0])
    rng = np.random.default_rng(11)
    starts = np.sort(rng.choice(g.n // 64, size=4, replace=False)) * 64
    ranges = [(int(s), int(s) + 64) for s in starts]
    mask = np.zeros(g.n, dtype=bool)
    for b, e in ranges:
        mask[b:e] = True
    rows_of = np.repeat(np.arange(g.n), np.diff(g.idx.astype(np.int64)))
    both = np.nonzero(mask[rows_of] & mask[g.adj])[0]
    R = exact_sc.sampled_scores(g, both, "cuda:0")
    _say(f"grid512: n={g.n} m={g.m} ranges={ranges} both-sampled edges={both.shape[0]}")
    return g, lam, X, ranges, mask, rows_of, both, R


@pytest.mark.parametrize("mode,name", [(MODE_FAST_FP32, "fp32"), (MODE_FAST_BF16, "bf16")])
def test_precision_grid_fast_within_eps(gpu_ctx, grid512, mode, name):
    g, lam, X, ranges, _, _, both, R = grid512
    dg = gpu_ctx.upload(g)
    try:
        for eps in (0.01, 0.005):
            _, sc, st = gpu_ctx.score_sources(dg, lam, X, eps, ranges, mode=mode)
            err = np.abs(sc[both].astype(np.float64) - R)
            _say(f"  grid512 {name} eps={eps}: {st['kernel_ms']:.0f} ms for 256 sources, max abs "
                 f"{err.max():.2e}, max rel {(err / R).max():.2e}, mean rel {(err / R).mean():.2e}")
            assert np.isfinite(sc[both]).all()
            assert err.max() <= eps
    finally:
        dg.free()


def test_precision_grid_faithful_rows_equal_oracle(gpu_ctx, grid512):
    g, lam, X, ranges, mask, rows_of, _, _ = grid512
    dg = gpu_ctx.upload(g)
    try:
        rows, _, st = gpu_ctx.score_sources(dg, lam, X, 0.01, ranges, mode=MODE_FAITHFUL)
    finally:
        dg.free()
    src = np.nonzero(mask)[0].astype(np.uint32)
    taus = oracle.taus(g.idx, g.adj, lam, X, np.float32(0.005))
    ref = oracle.tgtp_sources(g.idx, g.adj, taus, 0.01, 10, src, threads=_threads())
    rp = mask[rows_of]
    _say(f"  grid512 faithful: {st['kernel_ms']:.0f} ms for {src.shape[0]} sources")
    assert np.array_equal(rows[rp].view(np.uint32), ref[rp].view(np.uint32))


@pytest.fixture(scope="module")
def batch256():
    graphs = WL.batch_graphs()
    eig = [E.eigens_f32(*e) for e in E.compute_eigens_many(graphs, 128, device="cuda:0")]
    rng = np.random.default_rng(7)
    samples = [np.sort(rng.choice(g.m, size=min(32, g.m), replace=False)) for g in graphs]
    exact = [exact_sc.sampled_scores(g, p, "cuda:0") for g, p in zip(graphs, samples)]
    _say(f"batch256: sum n={sum(g.n for g in graphs)} sum m={sum(g.m for g in graphs)}")
    return graphs, eig, samples, exact


def _sampled_err(outs, samples, exact):
    est = np.concatenate([o[p] for o, p in zip(outs, samples)]).astype(np.float64)
    return np.abs(est - np.concatenate(exact))


@pytest.mark.parametrize("eps,mode,name", [(0.05, MODE_FAST_FP32, "fp32"),
                                           (0.01, MODE_FAST_FP32, "fp32"),
                                           (0.01, MODE_FAST_BF16, "bf16")])
def test_batch256_equals_per_graph_and_within_eps(gpu_ctx, batch256, eps, mode, name):
    graphs, eig, samples, exact = batch256
    b = gpu_ctx.batch(graphs, eig, eps, 10, mode=mode, graph=True)
    try:
        b.run()
        outs, st = b.run()  # a replay of the captured run
    finally:
        b.close()
    for k, (g, e) in enumerate(zip(graphs, eig)):
        dg = gpu_ctx.upload(g)
        s, _ = gpu_ctx.score(dg, *e, eps, 10, mode=mode)
        dg.free()
        assert np.array_equal(s.view(np.uint32), outs[k].view(np.uint32)), f"graph {k}"
    err = _sampled_err(outs, samples, exact)
    _say(f"  batch256 {name} eps={eps}: {st['kernel_ms']:.0f} ms, sampled max abs {err.max():.4f}, "
         f"mean abs {err.mean():.2e}")
    assert err.max() <= eps


def test_batch256_reference_tau_rule_leaves_eps(gpu_ctx, batch256):
    """The same batch with the reference's tau rule: the bridge edges whose
    top-omega tail alone exceeds eps/2 get tau = 1 (aesc.h:130), and the
    sampled error exceeds eps at eps 0.05."""
    graphs, eig, samples, exact = batch256
    outs, _ = gpu_ctx.score_batch(graphs, eig, 0.05, 10, mode=MODE_FAST_FP32 | TAU_REFERENCE)
    err = _sampled_err(outs, samples, exact)
    _say(f"  batch256 reference tau rule eps=0.05: sampled max abs {err.max():.4f}")
    assert err.max() > 0.05
