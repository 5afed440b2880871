"""Faithful TGT+ on the GPU vs the reference and the oracle, through the C ABI.

* taus: equal to the reference build's calculate_taus (golden fixtures).
* scores: bit-identical to the oracle's SABA-mode pipeline, whose every stage
  except the 2-way walk's fallback draws is pinned against the reference
  build (test_oracle_golden) -- compared stage by stage via tgtp_score_trace.
* accuracy of the n=4096 config against the FP64 exact resistances.
"""
import sys

import numpy as np
import pytest

import oracle
from conftest import GOLDEN_CASES, load_golden
from tgtp_amd import MODE_FAITHFUL, MODE_FAST_FP32, WALK2_MT19937, TgtpError
from tgtp_amd import eigenprep as E
from tgtp_amd import graph as G
from tgtp_amd import workloads as W

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _compare_stages(g, gpu_tr, orc_tr):
    deg = np.diff(g.idx.astype(np.int64))
    for key in ("taus", "ell", "q_size", "gmax", "gmin", "edge_max"):
        a, b = gpu_tr[key], orc_tr[key]
        bad = np.flatnonzero(a.view(np.uint32) != b.view(np.uint32)) if a.dtype.kind == "f" \
            else np.flatnonzero(a != b)
        assert bad.size == 0, f"{key}: {bad.size} differ, first at {bad[:5]}: {a[bad[:5]]} vs {b[bad[:5]]}"
    rest_pos = (orc_tr["taus"] - np.repeat(orc_tr["ell"], deg)) > 0
    for key in ("nr", "walk_sum"):
        a, b = gpu_tr[key][rest_pos], orc_tr[key][rest_pos]
        bad = np.flatnonzero(a.view(np.uint32 if a.dtype.itemsize == 4 else np.uint64)
                             != b.view(np.uint32 if b.dtype.itemsize == 4 else np.uint64))
        assert bad.size == 0, f"{key}: {bad.size} differ at {np.flatnonzero(rest_pos)[bad[:5]]}"
    a, b = gpu_tr["preds_row"], orc_tr["preds_row"]
    bad = np.flatnonzero(a.view(np.uint32) != b.view(np.uint32))
    assert bad.size == 0, f"preds_row: {bad.size} differ"


def _run(ctx, g, lam, X, eps, gamma, label):
    _say(f"tgtp case {label}: n={g.n} m={g.m} omega={lam.shape[0]} eps={eps} gamma={gamma}")
    dg = ctx.upload(g)
    try:
        s, tr, st = ctx.score_trace(dg, lam, X, eps, gamma)
    finally:
        dg.free()
    return s, tr, st


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_faithful_matches_reference_taus_and_oracle_scores(gpu_ctx, name):
    z = load_golden(name)
    g = G.Graph(z["idx"], z["adj"])
    for k, eps in enumerate(z["epsilons"]):
        s, tr, st = _run(gpu_ctx, g, z["lam"], z["X"], float(eps), int(z["gamma"]), name)
        assert np.array_equal(tr["taus"], z[f"taus_{k}"]), "taus differ from the reference"
        os_, otr = oracle.tgtp(z["idx"], z["adj"], z["lam"], z["X"], float(eps),
                               int(z["gamma"]), trace=True)
        _compare_stages(g, tr, otr)
        assert np.array_equal(s.view(np.uint32), os_.view(np.uint32))


@pytest.mark.parametrize("eps", [0.05, 0.01])
def test_faithful_chunglu4096_bitwise(gpu_ctx, eps):
    g = W.chung_lu_4096()
    lam, X = E.eigens_f32(*E.compute_eigens(g, 128))
    s, tr, st = _run(gpu_ctx, g, lam, X, eps, 10, f"cl4096 eps={eps}")
    os_, otr = oracle.tgtp(g.idx, g.adj, lam, X, eps, 10, trace=True)
    _compare_stages(g, tr, otr)
    assert np.array_equal(s.view(np.uint32), os_.view(np.uint32))
    _say(f"  kernel_ms={st['kernel_ms']:.2f} taus_ms={st['taus_ms']:.2f} "
         f"diffuse_ms={st['diffuse_ms']:.2f} host_fixups={st['host_fixups']} "
         f"exact_selections={st['work_items']}")


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_mt19937_walks_equal_reference_baseline(gpu_ctx, name):
    """The reference's baseline tgtp (aesc_baseline.h: one std::mt19937 seeded
    42, draws in source and position order, one float chain per position):
    scores equal the reference build's own outputs bit for bit."""
    z = load_golden(name)
    g = G.Graph(z["idx"], z["adj"])
    for k, eps in enumerate(z["epsilons"]):
        _say(f"mt19937 walks {name}: eps={float(eps)}")
        s, st = gpu_ctx.score(g, z["lam"], z["X"], float(eps), int(z["gamma"]),
                              mode=MODE_FAITHFUL | WALK2_MT19937)
        ref = z[f"scores_mt_{k}"]
        bad = np.flatnonzero(s.view(np.uint32) != ref.view(np.uint32))
        assert bad.size == 0, f"{bad.size} scores differ, first at {bad[:5]}"


@pytest.mark.parametrize("eps", [0.05, 0.01])
def test_mt19937_walks_chunglu4096_equal_oracle(gpu_ctx, eps):
    g = W.chung_lu_4096()
    lam, X = E.eigens_f32(*E.compute_eigens(g, 128))
    s, st = gpu_ctx.score(g, lam, X, eps, 10, mode=MODE_FAITHFUL | WALK2_MT19937)
    ref = oracle.tgtp(g.idx, g.adj, lam, X, eps, 10, walk_mode=oracle.WALK_MT19937)
    assert np.array_equal(s.view(np.uint32), ref.view(np.uint32))
    _say(f"  mt19937 walks eps={eps}: kernel_ms={st['kernel_ms']:.2f}")


def test_mt19937_walks_refused_outside_faithful(gpu_ctx):
    g = G.cycle(15)
    lam, X = E.eigens_f32(*E.compute_eigens(g, 6))
    with pytest.raises(TgtpError, match="mode"):
        gpu_ctx.score(g, lam, X, 0.05, 10, mode=MODE_FAST_FP32 | WALK2_MT19937)


def test_faithful_symmetric_graphs_tie_selection(gpu_ctx):
    """Vertex-transitive graphs make pvec tie everywhere, so the top-gamma set
    is decided by std::nth_element's permutation."""
    for g, omega, eps, gamma in [(G.petersen(), 4, 0.01, 3), (G.cycle(21), 5, 0.02, 4),
                                 (G.grid_rewired(12, 12, 0.0, seed=1), 16, 0.05, 10),
                                 (G.complete(7), 3, 0.01, 2)]:
        lam, X = E.eigens_f32(*E.compute_eigens(g, omega))
        s, tr, st = _run(gpu_ctx, g, lam, X, eps, gamma, f"sym n={g.n}")
        os_, otr = oracle.tgtp(g.idx, g.adj, lam, X, eps, gamma, trace=True)
        _compare_stages(g, tr, otr)
        assert np.array_equal(s.view(np.uint32), os_.view(np.uint32))


def test_faithful_gamma_one_and_duplicates(gpu_ctx, tmp_path):
    # duplicate edges and a self-loop exercise the duplicate-run push and the
    # first-occurrence reverse positions
    p = tmp_path / "g.txt"
    base = G.chung_lu(400, 2.5, 9.0, seed=21)
    base.write_txt(p)
    with open(p, "a") as f:
        f.write("0 1\n0 1\n5 5\n7 9\n7 9\n")
    g = G.load_txt(p)
    lam, X = E.eigens_f32(*E.compute_eigens(g, 24))
    for gamma in (1, 10):
        s, tr, st = _run(gpu_ctx, g, lam, X, 0.05, gamma, f"dups gamma={gamma}")
        os_, otr = oracle.tgtp(g.idx, g.adj, lam, X, 0.05, gamma, trace=True)
        _compare_stages(g, tr, otr)
        assert np.array_equal(s.view(np.uint32), os_.view(np.uint32))


def test_faithful_global_memory_variant(gpu_ctx):
    """n above the LDS-resident limit: per-source state in HBM."""
    g = G.chung_lu(14000, 2.5, 8.0, seed=3)
    lam, X = E.eigens_f32(*E.compute_eigens(g, 16))
    s, tr, st = _run(gpu_ctx, g, lam, X, 0.2, 10, "cl14000 global")
    os_ = oracle.tgtp(g.idx, g.adj, lam, X, 0.2, 10)
    assert np.array_equal(s.view(np.uint32), os_.view(np.uint32))


def test_accuracy_vs_exact_chunglu4096(gpu_ctx):
    from oracle.exact import exact_resistance

    g = W.chung_lu_4096()
    lam, X = E.eigens_f32(*E.compute_eigens(g, 128))
    R, info = exact_resistance(g.idx, g.adj)
    dg = gpu_ctx.upload(g)
    for eps in (0.05, 0.01):
        s, st = gpu_ctx.score(dg, lam, X, eps, 10)
        err = np.abs(s.astype(np.float64) - R)
        _say(f"  eps={eps}: max abs err {err.max():.5f} mean {err.mean():.6f} "
             f"max rel {(err / R).max():.4f}")
        assert err.max() <= eps
    dg.free()


def test_invalid_inputs_rejected(gpu_ctx):
    g = G.Graph.from_edges(6, [(0, 1), (1, 2), (2, 0), (3, 4)])  # vertex 5 isolated
    lam = np.array([1.0, 0.5, 0.2], dtype=np.float64)
    X = np.ones((6, 3), dtype=np.float32)
    with pytest.raises(TgtpError, match="isolated"):
        gpu_ctx.score(g, lam, X, 0.05, 10)
    g2 = G.cycle(9)
    X2 = np.ones((9, 3), dtype=np.float32)
    with pytest.raises(TgtpError, match="gamma"):
        gpu_ctx.score(g2, lam, X2, 0.05, 0)
    with pytest.raises(TgtpError, match="epsilon"):
        gpu_ctx.score(g2, lam, X2, 1.5, 10)
    with pytest.raises(TgtpError, match="omega"):
        gpu_ctx.score(g2, lam[:1], X2[:, :1], 0.05, 10)
    with pytest.raises(TgtpError, match="unknown tgtp mode"):
        gpu_ctx.taus(g2, lam, X2, 0.025, 7)
    assert MODE_FAITHFUL == 0
    # a caller's CSR missing a reverse edge (the reference's loader always
    # stores both; its lower_bound would run past the row), or with an
    # unsorted row: refused before any kernel indexes by them
    idx = np.array([0, 2, 4, 6, 7], dtype=np.uint32)
    adj = np.array([1, 2, 0, 2, 0, 1, 1], dtype=np.uint32)  # 3 -> 1, but not 1 -> 3
    X4 = np.ones((4, 3), dtype=np.float32)
    with pytest.raises(TgtpError, match="both directions"):
        gpu_ctx.score(G.Graph(idx, adj), lam, X4, 0.05, 10)
    adj_u = np.array([2, 1, 0, 2, 0, 1], dtype=np.uint32)  # triangle, row 0 unsorted
    with pytest.raises(TgtpError, match="sorted"):
        gpu_ctx.score(G.Graph(idx[:4], adj_u), lam, X4[:3], 0.05, 10)
