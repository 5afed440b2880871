"""FAST_FP32 / FAST_BF16 tgtp on the GPU, through the C ABI.

The fast modes change the arithmetic (synchronous diffusion, MFMA tau
projection, FP64 row sums), so they are checked against the FP64 exact
resistances (the accuracy oracle) and against the faithful path, not
bitwise.  Tolerances are written in each test.
"""
import sys

import numpy as np
import pytest

from oracle.exact import exact_resistance
import tau_rules as TR
from tgtp_amd import MODE_FAITHFUL, MODE_FAST_BF16, MODE_FAST_FP32, TAU_REFERENCE
from tgtp_amd import eigenprep as E
from tgtp_amd import graph as G
from tgtp_amd import workloads as W

pytestmark = pytest.mark.gpu
MODES = [(MODE_FAST_FP32, "fp32"), (MODE_FAST_BF16, "bf16")]


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def cl4096():
    g = W.chung_lu_4096()
    lam, X = E.eigens_f32(*E.compute_eigens(g, 128))
    R, _ = exact_resistance(g.idx, g.adj)
    return g, lam, X, R


@pytest.mark.parametrize("eps", [0.05, 0.01])
@pytest.mark.parametrize("mode,name", MODES)
def test_fast_within_eps_of_exact(gpu_ctx, cl4096, eps, mode, name):
    """|score - R_exact| <= eps on every CSR entry (the accuracy bound of
    paper_accuracy.py:69-73), deterministic across calls."""
    g, lam, X, R = cl4096
    _say(f"fast {name} cl4096 eps={eps}")
    dg = gpu_ctx.upload(g)
    s1, st = gpu_ctx.score(dg, lam, X, eps, 10, mode=mode)
    s2, _ = gpu_ctx.score(dg, lam, X, eps, 10, mode=mode)
    f, _ = gpu_ctx.score(dg, lam, X, eps, 10, mode=MODE_FAITHFUL)
    dg.free()
    assert np.array_equal(s1.view(np.uint32), s2.view(np.uint32)), "not deterministic"
    err = np.abs(s1.astype(np.float64) - R)
    ferr = np.abs(f.astype(np.float64) - R)
    rel = err / R
    _say(f"  {name}: max abs {err.max():.5f} mean {err.mean():.6f} max rel {rel.max():.4f} "
         f"mean rel {rel.mean():.5f} | faithful max abs {ferr.max():.5f} | kernel "
         f"{st['kernel_ms']:.2f} ms (taus {st['taus_ms']:.2f})")
    assert err.max() <= eps
    # no worse than the faithful path's own error by more than eps/2
    assert err.mean() <= ferr.mean() + eps / 2


@pytest.mark.parametrize("mode,name", MODES)
def test_mfma_taus_match_faithful(gpu_ctx, cl4096, mode, name):
    """With the reference's refinement rule the projection changes only
    rounding: taus equal the faithful ones except where the bound sits on an
    integer boundary (<= 1% FP32, <= 5% BF16 of CSR positions)."""
    g, lam, X, _ = cl4096
    for eps in (0.05, 0.01):
        tf, _ = gpu_ctx.taus(g, lam, X, np.float32(eps / 2), MODE_FAITHFUL)
        tq, _ = gpu_ctx.taus(g, lam, X, np.float32(eps / 2), mode | TAU_REFERENCE)
        diff = np.mean(tf != tq)
        _say(f"  taus {name} eps={eps}: {diff * 100:.3f}% differ, max |d| {np.abs(tf - tq).max()}")
        assert diff <= (0.01 if mode == MODE_FAST_FP32 else 0.05)
        assert np.abs(tf - tq).max() <= 2


@pytest.fixture(scope="module")
def batch_bridges():
    """Batched-workload graphs 29 and 156: edges whose top-omega tail alone
    exceeds eps/2 (the reference's rule gives them tau = 1)."""
    gs = W.batch_graphs(157)
    return [(gs[i], *E.eigens_f32(*E.compute_eigens(gs[i], 128))) for i in (29, 156)]


@pytest.mark.parametrize("mode,name", MODES)
def test_mfma_taus_bounded_rule(gpu_ctx, cl4096, batch_bridges, mode, name):
    """The default (bounded) rule: taus equal the numpy restatement of that
    rule except at rounding boundaries (FP32 <= 1% of positions, BF16 <= 5%),
    and every tau keeps the FP64 truncation bound within the projection's
    rounding (delta and upsilon carry ~1e-6 relative error in FP32, ~1e-2 in
    BF16)."""
    slack = 1e-3 if mode == MODE_FAST_FP32 else 5e-2
    cases = [(g, lam, X) for g, lam, X in batch_bridges] + [cl4096[:3]]
    for g, lam, X in cases:
        for eps in (0.05, 0.01):
            tq, _ = gpu_ctx.taus(g, lam, X, np.float32(eps / 2), mode)
            tn = TR.taus(g, lam, X, eps / 2, bounded=True)
            diff = np.mean(tq != tn)
            bound = TR.truncation_bound(g, lam, X, tq)
            _say(f"  bounded taus {name} n={g.n} eps={eps}: {diff * 100:.3f}% differ from the "
                 f"restatement, max bound {bound.max():.5f} (eps/2 {eps / 2})")
            assert diff <= (0.01 if mode == MODE_FAST_FP32 else 0.05)
            assert bound.max() <= eps / 2 * (1 + slack)


def test_fast_symmetric_output_and_trace(gpu_ctx):
    g = G.chung_lu(1500, 2.5, 12.0, seed=31)
    lam, X = E.eigens_f32(*E.compute_eigens(g, 32))
    dg = gpu_ctx.upload(g)
    s, tr, st = gpu_ctx.score_trace(dg, lam, X, 0.05, 10, mode=MODE_FAST_FP32)
    dg.free()
    rev = g.reverse_index()
    assert np.array_equal(s, s[rev])
    assert tr["ell"].min() >= 1 and np.all(tr["q_size"] >= 1)


def test_fast_state_in_hbm(gpu_ctx):
    """n beyond the LDS-resident limit: per-source state in HBM.  This sparse
    graph has bridge edges (R = 1) that the tau truncation itself misses --
    the reference's own scores are 0.17 off there -- so the bound is the
    faithful error plus eps."""
    import oracle

    g = G.chung_lu(12000, 2.5, 10.0, seed=8)
    lam, X = E.eigens_f32(*E.compute_eigens(g, 32))
    R, _ = exact_resistance(g.idx, g.adj)
    ferr = np.abs(oracle.tgtp(g.idx, g.adj, lam, X, 0.05, 10) - R)
    for mode, name in MODES:
        s, st = gpu_ctx.score(g, lam, X, 0.05, 10, mode=mode)
        err = np.abs(s - R)
        _say(f"  cl12000 {name}: max abs {err.max():.5f} mean {err.mean():.6f} "
             f"(faithful {ferr.max():.5f} / {ferr.mean():.6f})")
        assert err.max() <= ferr.max() + 0.05
        assert err.mean() <= ferr.mean() + 0.01
        assert np.mean(err > 0.05) <= np.mean(ferr > 0.05) + 0.001


def test_fast_on_rewired_grid_tracks_faithful(gpu_ctx):
    """Slow-mixing graph with long walks: the fast modes' errors against the
    exact resistances stay within the faithful path's own error plus eps."""
    g = G.grid_rewired(24, 24, 0.05, seed=2)
    lam, X = E.eigens_f32(*E.compute_eigens(g, 32))
    R, _ = exact_resistance(g.idx, g.adj)
    f, _ = gpu_ctx.score(g, lam, X, 0.05, 10, mode=MODE_FAITHFUL)
    ferr = np.abs(f - R)
    for mode, name in MODES:
        s, _ = gpu_ctx.score(g, lam, X, 0.05, 10, mode=mode)
        err = np.abs(s - R)
        _say(f"  grid24 {name}: max abs {err.max():.4f} (faithful {ferr.max():.4f}) "
             f"mean {err.mean():.5f} (faithful {ferr.mean():.5f})")
        assert err.max() <= ferr.max() + 0.05
        assert err.mean() <= ferr.mean() + 0.025


def test_fast_tile_orders_agree(gpu_ctx, monkeypatch):
    """Both orders of the diffusion tiles (per-XCD runs of blocks, and
    groups of 8 tiles of one block across the XCDs, used for blocks above
    65,536 vertices) compute the same scores."""
    g = G.chung_lu(5000, 2.3, 10.0, seed=31)
    lam, X = E.eigens_f32(*E.compute_eigens(g, 32))
    out = []
    for mode in ("0", "1"):
        monkeypatch.setenv("TGTP_FAST_TILE_MODE", mode)
        s, _ = gpu_ctx.score(g, lam, X, 0.02, 10, mode=MODE_FAST_FP32)
        out.append(s)
    assert np.array_equal(out[0].view(np.uint32), out[1].view(np.uint32))


def test_fast_subsets_do_not_change_scores(gpu_ctx, monkeypatch):
    """Sources advance in blocks of 64 whose states share subset buffers; the
    scores do not depend on how blocks are grouped into subsets."""
    g = G.chung_lu(3000, 2.4, 14.0, seed=21)
    lam, X = E.eigens_f32(*E.compute_eigens(g, 32))
    for mode, name in MODES:
        a, _ = gpu_ctx.score(g, lam, X, 0.02, 10, mode=mode)
        monkeypatch.setenv("TGTP_FAST_STATE_MIB", "1")  # one block per subset
        b, _ = gpu_ctx.score(g, lam, X, 0.02, 10, mode=mode)
        monkeypatch.delenv("TGTP_FAST_STATE_MIB")
        _say(f"  subsets {name}: equal={np.array_equal(a.view(np.uint32), b.view(np.uint32))}")
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_fast_round_forms_agree(gpu_ctx, monkeypatch):
    """The diffusion round with one row per wave and with four rows per wave
    (16 lanes per row, 8 or 16 gathers per step) sum every source in the
    same order: without the hub-row kernel the scores are equal bit for
    bit, also with partly filled blocks (n = 5000 leaves 8 sources in the
    last block) and sources that stop at different rounds.  With rows above
    a degree threshold summed in 16 segments (hub workgroups) the three
    forms still agree bit for bit; against the sequential sums the
    diffusion differs by FP32 rounding, which can move a walk count's
    ceiling and so the sampled part of a score: the scores stay close
    (mean change < 1e-3, max < eps)."""
    cases = [(G.chung_lu(5000, 2.3, 10.0, seed=31), 0.02),
             (G.grid_rewired(40, 40, 0.05, seed=3), 0.05)]
    for g, eps in cases:
        lam, X = E.eigens_f32(*E.compute_eigens(g, 32))
        monkeypatch.setenv("TGTP_FAST_HUB_DEG", "0")
        seq = []
        for form in ("0", "1", "2"):
            monkeypatch.setenv("TGTP_FAST_ROUND", form)
            s, _ = gpu_ctx.score(g, lam, X, eps, 10, mode=MODE_FAST_FP32)
            seq.append(s)
        for s in seq[1:]:
            assert np.array_equal(seq[0].view(np.uint32), s.view(np.uint32))
        for hub in ("128", "6"):  # 6: most rows of these graphs are hubs
            monkeypatch.setenv("TGTP_FAST_HUB_DEG", hub)
            out = []
            for form in ("0", "1", "2"):
                monkeypatch.setenv("TGTP_FAST_ROUND", form)
                out.append(gpu_ctx.score(g, lam, X, eps, 10, mode=MODE_FAST_FP32)[0])
            for s in out[1:]:
                assert np.array_equal(out[0].view(np.uint32), s.view(np.uint32))
            d = np.abs(out[0] - seq[0])
            _say(f"  hub rows above degree {hub}: mean change {d.mean():.2e} max {d.max():.2e}")
            assert d.mean() < 1e-3 and d.max() < eps


def test_fast_walk_sums_independent_of_small_bq(gpu_ctx, monkeypatch):
    """A position's 2-way walk sum has one order whichever path runs it (one
    8-lane group for few bouquets, eight groups for many): scores are equal
    bit for bit for any small_bq, the knob the batch autotuner sets."""
    g = G.chung_lu(3000, 2.4, 14.0, seed=21)
    lam, X = E.eigens_f32(*E.compute_eigens(g, 32))
    out = []
    for sbq in ("0", "8", "100000"):
        monkeypatch.setenv("TGTP_FAST_SMALL_BQ", sbq)
        out.append(gpu_ctx.score(g, lam, X, 0.01, 10, mode=MODE_FAST_FP32)[0])
    for s in out[1:]:
        assert np.array_equal(out[0].view(np.uint32), s.view(np.uint32))


@pytest.mark.parametrize("mode", [MODE_FAITHFUL, MODE_FAST_FP32])
def test_device_graph_eigenpair_cache_keys(gpu_ctx, mode):
    """tgtp_score_dev re-uploads the eigenpairs (and rebuilds a FAST plan)
    when either the vectors or the values pointer changes: new eigenvalues
    with the same vectors give the scores of a fresh device graph."""
    g = G.chung_lu(600, 2.5, 9.0, seed=31)
    lam, X = E.eigens_f32(*E.compute_eigens(g, 16))
    lam2 = lam.copy()
    lam2[1:] *= 0.9  # other eigenvalues, same vectors array
    dg = gpu_ctx.upload(g)
    try:
        s1, _ = gpu_ctx.score(dg, lam, X, 0.05, 10, mode=mode)
        s2, _ = gpu_ctx.score(dg, lam2, X, 0.05, 10, mode=mode)
        s3, _ = gpu_ctx.score(dg, lam, X, 0.05, 10, mode=mode)
    finally:
        dg.free()
    f2, _ = gpu_ctx.score(g, lam2, X, 0.05, 10, mode=mode)
    assert s1.tobytes() == s3.tobytes()
    assert s2.tobytes() == f2.tobytes()
    assert s1.tobytes() != s2.tobytes()
