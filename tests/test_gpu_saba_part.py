"""The partitioned SABA walk engine (hop-synchronous rounds over vertex
partitions, bouquet fallback repaired by a replay pass) vs the oracle:
counters bit-exact.  TGTP_WALK_PART=1 makes the engine required (the call
fails if it does not take the pass), so every case here runs it; small
TGTP_WALK_PART_VMAX values split small graphs into many partitions.
Calls go through the C ABI (tgtp_saba_walks)."""
import json

import numpy as np
import pytest

import oracle
from conftest import GOLDEN
from test_gpu_saba import _check, _say, _sums
from tgtp_amd import TgtpError
from tgtp_amd import graph as G
from tgtp_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(params=["1", "0"], ids=["tagged", "lookup"])
def part(monkeypatch, request):
    """Engine required; the neighbours' partitions read from the tagged
    adjacency (the default up to 2^24 vertices) or looked up in the partition
    bounds (the path of larger graphs)."""
    monkeypatch.setenv("TGTP_WALK_PART", "1")
    monkeypatch.setenv("TGTP_WALK_PART_VMAX", "64")
    monkeypatch.setenv("TGTP_WALK_PART_TAGS", request.param)
    return monkeypatch


@pytest.mark.parametrize("depth,W_,vmax", [(5, 2048, "64"), (15, 512, "32"), (10, 256, "4096")])
def test_chunglu4096_partitioned(gpu_ctx, part, depth, W_, vmax):
    part.setenv("TGTP_WALK_PART_VMAX", vmax)
    _check(gpu_ctx, W.chung_lu_4096(), depth, W_, label=f"part cl4096 d{depth} vmax{vmax}")


@pytest.mark.parametrize("case", json.loads((GOLDEN / "saba_kats.json").read_text())["cases"],
                         ids=lambda c: f"{c['graph']}{c['n']}")
def test_known_answers_fallback_forced(gpu_ctx, part, case):
    """C9 / K5: the bouquet fallback fires in most bouquets, so nearly every
    count passes through the replay pass."""
    g = G.cycle(case["n"]) if case["graph"] == "cycle" else G.complete(case["n"])
    part.setenv("TGTP_WALK_PART_VMAX", "2")
    dg = gpu_ctx.upload(g)
    try:
        _, vc, cs, _ = gpu_ctx.saba_walks(dg, case["depth"], case["n_walks"])
    finally:
        dg.free()
    assert (vc, cs) == (case["visit_count"], case["checksum"])


def test_degree_powers_of_two_long_walks(gpu_ctx, part):
    part.setenv("TGTP_WALK_PART_VMAX", "8")
    for g, L in [(G.cycle(9), 60), (G.complete(5), 33), (G.grid_rewired(16, 16, 0.0, seed=1), 40)]:
        _check(gpu_ctx, g, L, 256, label=f"part pow2 n={g.n}")


@pytest.mark.parametrize("W_", [8, 24, 136])
def test_bouquets_spanning_sources(gpu_ctx, part, W_):
    _check(gpu_ctx, G.chung_lu(3000, 2.5, 12.0, seed=9), 7, W_, 13, 2999, label=f"part cl3000 W{W_}")


def test_isolated_vertices_and_partial_ranges(gpu_ctx, part):
    edges = [(i, (i * 7 + 3) % 40) for i in range(1, 40) if i % 5 != 0 and (i * 7 + 3) % 40 not in (0, 20)]
    g = G.Graph.from_edges(42, [e for e in edges if 20 not in e] + [(1, 2)])
    part.setenv("TGTP_WALK_PART_VMAX", "4")
    _check(gpu_ctx, g, 12, 64, 3, 37, label="part isolated")


def test_walker_stops_at_vertex_without_neighbours(gpu_ctx, part):
    """Non-symmetric CSR: walkers that reach a vertex without neighbours stop
    and leave their bouquet's zero check, in the rounds and in the replay."""
    from test_oracle_pinning import SABA_SINK_ADJ, SABA_SINK_IDX
    part.setenv("TGTP_WALK_PART_VMAX", "2")
    g = G.Graph(SABA_SINK_IDX, SABA_SINK_ADJ)
    for depth, W_ in ((12, 16), (40, 64), (60, 256)):
        _check(gpu_ctx, g, depth, W_, label=f"part sink d{depth}")


@pytest.mark.parametrize("knobs", [
    {"TGTP_WALK_PART_CHUNK": "7"},                      # strided source chunks
    {"TGTP_WALK_PART_PER_CU": "1"},                     # one workgroup per CU
    {"TGTP_WALK_PARTS": "1", "TGTP_WALK_PART_VMAX": "8192"},  # a single partition
    {"TGTP_WALK_PARTS": "256", "TGTP_WALK_PART_VMAX": "16"},  # the most sort bins
    {"TGTP_WALK_PART_GEN_HOPS": "1"},                   # one hop in the generation round
    {"TGTP_WALK_PART_GEN_HOPS": "4"},                   # four
    {"TGTP_WALK_PART_VMAX": "512"},                     # arrival table of 256 slots
    {"TGTP_WALK_PART_BALANCE": "0"},                    # partition p to workgroups b = p mod 8
])
def test_engine_shapes(gpu_ctx, part, knobs):
    for k, v in knobs.items():
        part.setenv(k, v)
    _check(gpu_ctx, G.chung_lu(3000, 2.3, 14.0, seed=5), 9, 64, label=f"part {knobs}")


@pytest.mark.parametrize("depth,gen_hops", [(1, "2"), (2, "2"), (3, "2"), (4, "4"), (5, "4")])
def test_round_feeding_the_last(gpu_ctx, part, depth, gen_hops):
    """Record widths near the end of a walk: the round whose records the
    last round reads writes vertices alone, the one before it {vertex,
    secret}.  The generation round writes the vertices when it takes every
    hop (depth 1 and 2 at two generation hops, depth 4 at four) and the
    2-word records when one hop follows it (depth 3 at two, depth 5 at
    four); middle rounds otherwise."""
    part.setenv("TGTP_WALK_PART_GEN_HOPS", gen_hops)
    _check(gpu_ctx, G.chung_lu(3000, 2.3, 14.0, seed=5), depth, 64,
           label=f"part last-feed d{depth} gen{gen_hops}")


@pytest.mark.parametrize("W_,slots", [(6144, "1"), (12288, "1"), (16384, "1"), (16384, "0")])
def test_generation_first_hop_counting(gpu_ctx, part, W_, slots):
    """Two-hop generation round with walks per source of at least a tile
    (6,144): first-hop arrivals counted per neighbour slot of the tile's one
    or two sources (slots "1"), or in the LDS table ("0")."""
    part.setenv("TGTP_WALK_PARTS", "4")
    part.setenv("TGTP_WALK_PART_VMAX", "4096")
    part.setenv("TGTP_WALK_PART_GEN_SLOTS", slots)
    _check(gpu_ctx, G.chung_lu(3000, 2.3, 14.0, seed=5), 6, W_, 0, 700, label=f"part gen W{W_}")


def test_counts_beyond_one_block_per_partition(gpu_ctx, part):
    """Many blocks per partition and per workgroup: 70,000 vertices, 128
    partitions of at most 1024 vertices, 2^27 walks over 4 chunks."""
    part.setenv("TGTP_WALK_PART_VMAX", "1024")
    part.setenv("TGTP_WALK_PART_CHUNK", "17500")
    _check(gpu_ctx, G.chung_lu(70000, 2.3, 12.0, seed=9), 7, 1920, label="part cl70000")


def test_hub_arrivals_beyond_16_bits(gpu_ctx, part):
    """Star of 2,001 vertices, 16,384 walks from every vertex: half of each
    round's 32.8M walkers stand on the hub, about 256,000 per workgroup copy,
    so the hub's 16-bit LDS counter is added into the copy and cleared many
    times within one partition (at most 65,535 arrivals between two)."""
    g = G.Graph.from_edges(2001, [(0, i) for i in range(1, 2001)])
    cnt, _ = _check(gpu_ctx, g, 5, 16384, label="part star2001")
    assert cnt[0] > 4 * 65535 * 64


def test_required_engine_refuses_too_many_ranges(gpu_ctx, part):
    part.setenv("TGTP_WALK_PART_VMAX", "64")
    g = G.chung_lu(50000, 2.2, 15.0, seed=4)  # > 256 ranges of 64 vertices
    dg = gpu_ctx.upload(g)
    try:
        with pytest.raises(TgtpError, match="partitioned engine required"):
            gpu_ctx.saba_walks(dg, 5, 8)
    finally:
        dg.free()


def test_auto_falls_back_to_direct_kernel(gpu_ctx, monkeypatch):
    """Automatic choice on a graph the engine does not take (too many
    ranges): the direct kernel runs and the counts are exact."""
    monkeypatch.setenv("TGTP_WALK_PART", "-1")
    monkeypatch.setenv("TGTP_WALK_PART_MIN_MIB", "0")
    monkeypatch.setenv("TGTP_WALK_PART_VMAX", "64")
    _check(gpu_ctx, G.chung_lu(50000, 2.2, 15.0, seed=4), 6, 64, label="auto fallback")


def test_engines_agree_at_full_scale(gpu_ctx, monkeypatch):
    """Walk config graph (n=2^20, ~20M CSR entries), 10 hops, 16384 walks on
    a spread of 4096 sources: the partitioned engine (the default at this
    size) and the direct kernel give the same counters."""
    g = W.chung_lu_1m()
    _say(f"engines cl1m n={g.n} m={g.m}")
    dg = gpu_ctx.upload(g)
    try:
        monkeypatch.setenv("TGTP_WALK_PART", "1")
        a = gpu_ctx.saba_walks(dg, 10, 16384, 200000, 204096)
        monkeypatch.setenv("TGTP_WALK_PART", "0")
        b = gpu_ctx.saba_walks(dg, 10, 16384, 200000, 204096)
    finally:
        dg.free()
    assert np.array_equal(a[0], b[0]) and a[1:3] == b[1:3]
    assert a[1] == 4096 * 16384 * 10 and _sums(a[0]) == a[1:3]
    ref = oracle.saba_walks(g.idx, g.adj, 10, 16384, 200000, 200256)
    monkeypatch.setenv("TGTP_WALK_PART", "1")
    dg = gpu_ctx.upload(g)
    try:
        c = gpu_ctx.saba_walks(dg, 10, 16384, 200000, 200256)
        monkeypatch.setenv("TGTP_WALK_PART_TAGS", "0")
        d = gpu_ctx.saba_walks(dg, 10, 16384, 200000, 200256)
    finally:
        dg.free()
    assert np.array_equal(c[0], ref) and np.array_equal(d[0], ref)
