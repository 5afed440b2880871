"""SABA 1-way walks on the GPU vs the oracle: counters bit-exact.
Calls go through the C ABI (tgtp_saba_walks)."""
import json
import sys

import numpy as np
import pytest

import oracle
from conftest import GOLDEN
from tgtp_amd import TgtpError
from tgtp_amd import graph as G
from tgtp_amd import workloads as W

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _sums(c):
    c = c.astype(np.uint64)
    return int(c.sum()), int((np.arange(1, c.shape[0] + 1, dtype=np.uint64) * c).sum())


def _check(ctx, g, depth, W_, lo=0, hi=None, label=""):
    hi = g.n if hi is None else hi
    _say(f"saba case {label}: n={g.n} m={g.m} depth={depth} W={W_} src=[{lo},{hi})")
    dg = ctx.upload(g)
    try:
        cnt, vc, cs, st = ctx.saba_walks(dg, depth, W_, lo, hi)
    finally:
        dg.free()
    ref = oracle.saba_walks(g.idx, g.adj, depth, W_, lo, hi)
    assert np.array_equal(cnt, ref), f"{label}: {int((cnt != ref).sum())} counters differ"
    assert (vc, cs) == _sums(ref)
    return cnt, st


@pytest.mark.parametrize("case", json.loads((GOLDEN / "saba_kats.json").read_text())["cases"],
                         ids=lambda c: f"{c['graph']}{c['n']}")
def test_known_answers_fallback_forced(gpu_ctx, case):
    g = G.cycle(case["n"]) if case["graph"] == "cycle" else G.complete(case["n"])
    dg = gpu_ctx.upload(g)
    _, vc, cs, _ = gpu_ctx.saba_walks(dg, case["depth"], case["n_walks"])
    dg.free()
    assert (vc, cs) == (case["visit_count"], case["checksum"])


@pytest.mark.parametrize("depth", [5, 15])
def test_chunglu4096_walks_bit_exact(gpu_ctx, depth):
    g = W.chung_lu_4096()
    cnt, st = _check(gpu_ctx, g, depth, 2048, label=f"cl4096 d{depth}")
    assert st["transitions"] == g.n * 2048 * depth


@pytest.mark.parametrize("W_", [8, 24, 64, 136])
def test_bouquets_spanning_sources_in_a_wave(gpu_ctx, W_):
    # W/8 not a multiple of 8: one wave holds bouquets of several sources
    g = G.chung_lu(3000, 2.5, 12.0, seed=9)
    _check(gpu_ctx, g, 7, W_, 13, 2999, label=f"cl3000 W{W_}")


def test_isolated_vertices_and_partial_ranges(gpu_ctx):
    # vertices 0, 20 and 41 have no edges
    edges = [(i, (i * 7 + 3) % 40) for i in range(1, 40) if i % 5 != 0 and (i * 7 + 3) % 40 not in (0, 20)]
    g = G.Graph.from_edges(42, [e for e in edges if 20 not in e] + [(1, 2)])
    assert g.degree(0) == 0 and g.degree(20) == 0 and g.degree(41) == 0
    _check(gpu_ctx, g, 12, 64, 3, 37, label="isolated")


def test_degree_powers_of_two_long_walks(gpu_ctx):
    # C9 / K5 / a 4-regular torus: every bouquet hits the fallback
    for g, L in [(G.cycle(9), 60), (G.complete(5), 33), (G.grid_rewired(16, 16, 0.0, seed=1), 40)]:
        _check(gpu_ctx, g, L, 256, label=f"pow2 n={g.n}")


def test_counters_beyond_lds_staging(gpu_ctx):
    # n above the LDS-staged vertex range: mixed LDS / global counters
    g = G.chung_lu(50000, 2.2, 15.0, seed=4)
    _check(gpu_ctx, g, 10, 128, label="cl50000")


def test_rewired_grid(gpu_ctx):
    _check(gpu_ctx, G.grid_rewired(64, 64, 0.05, seed=1), 10, 256, label="grid64")


@pytest.mark.parametrize("packed,hot", [("0", "32768"), ("1", "0"), ("0", "100"), ("1", "4096")])
def test_layouts_and_lds_staging_agree(gpu_ctx, monkeypatch, packed, hot):
    """The packed {id|start|degree} walk layout and the plain CSR give the
    same counters, for any LDS-staged vertex range."""
    monkeypatch.setenv("TGTP_WALK_PACKED", packed)
    monkeypatch.setenv("TGTP_WALK_HOT", hot)
    _check(gpu_ctx, W.chung_lu_4096(), 9, 512, 5, 4000, label=f"packed={packed} hot={hot}")


@pytest.mark.parametrize("knobs", [
    {"TGTP_WALK_BUCKETS": "0"},                             # direct atomics
    {"TGTP_WALK_REC_MIB": "1"},                             # many source chunks
    {"TGTP_WALK_REC_PCT": "5"},                             # regions overflow: atomics
    {"TGTP_WALK_HOT": "1000", "TGTP_WALK_PACKED": "0"},     # 2 buckets, plain layout
])
def test_bucketed_counting_paths(gpu_ctx, monkeypatch, knobs):
    """Cold visits staged per vertex-range bucket and counted by the LDS
    histogram pass give the same counters as direct atomics, through source
    chunking and region overflow."""
    for k, v in knobs.items():
        monkeypatch.setenv(k, v)
    _check(gpu_ctx, G.chung_lu(70000, 2.3, 12.0, seed=9), 7, 64, label=f"buckets {knobs}")


def test_deterministic_repeat(gpu_ctx):
    g = W.chung_lu_4096()
    dg = gpu_ctx.upload(g)
    a = gpu_ctx.saba_walks(dg, 10, 512)
    b = gpu_ctx.saba_walks(dg, 10, 512)
    dg.free()
    assert np.array_equal(a[0], b[0]) and a[1:3] == b[1:3]


def test_invalid_arguments_rejected(gpu_ctx):
    g = G.cycle(9)
    dg = gpu_ctx.upload(g)
    with pytest.raises(TgtpError, match="multiple of 8"):
        gpu_ctx.saba_walks(dg, 5, 12)
    with pytest.raises(TgtpError, match="source range"):
        gpu_ctx.saba_walks(dg, 5, 8, 0, 10)
    dg.free()


def test_zero_walks_give_zero_counters(gpu_ctx):
    """n_walks = 0 walks nothing and returns zero counters and sums, as the
    reference's driver does (rw_experiment_saba.cpp:160, 186), also right
    after a call that left counts in the context's buffers."""
    g = G.cycle(9)
    dg = gpu_ctx.upload(g)
    try:
        gpu_ctx.saba_walks(dg, 5, 8)
        cnt, vc, cs, _ = gpu_ctx.saba_walks(dg, 5, 0)
    finally:
        dg.free()
    assert vc == 0 and cs == 0 and not cnt.any()


@pytest.fixture(scope="module")
def cl1m():
    return W.chung_lu_1m()


@pytest.mark.parametrize("depth", [10, 15])
def test_full_scale_chunglu_1m(gpu_ctx, cl1m, depth):
    """Walk config at full size (n=2^20, ~20M CSR entries, 10 and 15 hops,
    16384 walks/vertex): size-independent invariants on the whole run (visit
    count n*W*L, checksum of the counters, the counters of four disjoint
    source quarters summing to the whole run's), bit-exact counters on
    source slices the oracle finishes in seconds."""
    g = cl1m
    _say(f"cl1m n={g.n} m={g.m} depth={depth}")
    dg = gpu_ctx.upload(g)
    try:
        cnt, vc, cs, st = gpu_ctx.saba_walks(dg, depth, 16384)
        assert vc == g.n * 16384 * depth and _sums(cnt) == (vc, cs)
        acc = np.zeros(g.n, dtype=np.uint64)
        q = g.n // 4
        for k in range(4):
            part, _, _, _ = gpu_ctx.saba_walks(dg, depth, 16384, k * q, (k + 1) * q)
            acc += part
        assert np.array_equal(acc, cnt.astype(np.uint64)), "quarters do not sum to the run"
        for lo, hi in [(0, 384), (g.n // 2, g.n // 2 + 512), (g.n - 640, g.n)]:
            part, pvc, pcs, _ = gpu_ctx.saba_walks(dg, depth, 16384, lo, hi)
            ref = oracle.saba_walks(g.idx, g.adj, depth, 16384, lo, hi)
            assert np.array_equal(part, ref), f"slice [{lo},{hi})"
            assert (pvc, pcs) == _sums(ref)
    finally:
        dg.free()


@pytest.mark.parametrize("packed,hot", [("0", "32768"), ("1", "32768"), ("1", "0")])
def test_walker_stops_at_vertex_without_neighbours(gpu_ctx, monkeypatch, packed, hot):
    """A non-symmetric CSR through the C ABI: walkers that reach a vertex
    without neighbours stop and leave their bouquet's zero check, as in the
    oracle (itself pinned by a brute-force restatement,
    tests/test_oracle_pinning.py), on both walk layouts and counter paths."""
    from test_oracle_pinning import SABA_SINK_ADJ, SABA_SINK_IDX
    monkeypatch.setenv("TGTP_WALK_PACKED", packed)
    monkeypatch.setenv("TGTP_WALK_HOT", hot)
    g = G.Graph(SABA_SINK_IDX, SABA_SINK_ADJ)
    for depth, W_ in ((12, 16), (40, 64), (60, 256)):
        _check(gpu_ctx, g, depth, W_, label=f"sink packed={packed} hot={hot}")
