"""MT19937 comparator walks on the GPU (tgtp_mt_walks, through the C ABI):
counters bit-exact against the reference's own scalar_randomwalk fixtures
(regular) and the oracle (simd)."""
import sys

import numpy as np
import pytest

import oracle
from conftest import MT_WALK_CASES, load_mt_walk_golden
from tgtp_amd import WALK_MT_REGULAR, WALK_MT_SIMD, TgtpError
from tgtp_amd import graph as G

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _sums(c):
    c = c.astype(np.uint64)
    return int(c.sum()), int((np.arange(1, c.shape[0] + 1, dtype=np.uint64) * c).sum())


@pytest.mark.parametrize("name", MT_WALK_CASES)
def test_regular_matches_reference_fixtures(gpu_ctx, name):
    z = load_mt_walk_golden(name)
    g = G.Graph(z["idx"], z["adj"])
    dg = gpu_ctx.upload(g)
    try:
        for depth, n_walks in z["runs"]:
            _say(f"mt regular {name}: n={g.n} depth={depth} W={n_walks}")
            cnt, vc, cs, _ = gpu_ctx.mt_walks(dg, int(depth), int(n_walks), WALK_MT_REGULAR)
            ref = z[f"counts_d{depth}_w{n_walks}"]
            assert np.array_equal(cnt, ref), f"{int((cnt != ref).sum())} counters differ"
            assert (vc, cs) == _sums(ref)
    finally:
        dg.free()


@pytest.mark.parametrize("name", MT_WALK_CASES)
def test_simd_matches_oracle(gpu_ctx, name):
    z = load_mt_walk_golden(name)
    g = G.Graph(z["idx"], z["adj"])
    dg = gpu_ctx.upload(g)
    try:
        for depth, n_walks in z["runs"]:
            n_walks = int(n_walks) - int(n_walks) % 8 or 8
            _say(f"mt simd {name}: n={g.n} depth={depth} W={n_walks}")
            cnt, vc, cs, _ = gpu_ctx.mt_walks(dg, int(depth), n_walks, WALK_MT_SIMD)
            ref = oracle.mt_walks(g.idx, g.adj, int(depth), n_walks, variant=1)
            assert np.array_equal(cnt, ref), f"{int((cnt != ref).sum())} counters differ"
            assert (vc, cs) == _sums(ref)
    finally:
        dg.free()


def test_ring_window_and_source_ranges(gpu_ctx):
    """Depths that leave fewer than 256 walks per batch (the draw ring holds
    4096 outputs), partial batches, and source sub-ranges."""
    z = load_mt_walk_golden("chunglu2000")
    g = G.Graph(z["idx"], z["adj"])
    dg = gpu_ctx.upload(g)
    try:
        for variant, depth, n_walks, lo, hi in ((0, 23, 300, 0, 2000), (1, 23, 312, 0, 2000),
                                               (0, 434, 19, 100, 900), (1, 434, 24, 5, 77),
                                               (0, 3472, 2, 0, 50), (1, 1, 8, 1999, 2000),
                                               (0, 7, 100, 300, 300), (0, 5, 700, 0, 2000)):
            _say(f"mt ring case variant={variant} depth={depth} W={n_walks} src=[{lo},{hi})")
            cnt, vc, cs, _ = gpu_ctx.mt_walks(dg, depth, n_walks, variant, lo, hi)
            ref = oracle.mt_walks(g.idx, g.adj, depth, n_walks, variant=variant,
                                  src_begin=lo, src_end=hi)
            assert np.array_equal(cnt, ref)
            assert (vc, cs) == _sums(ref)
    finally:
        dg.free()


def test_errors(gpu_ctx):
    g = G.cycle(9)
    dg = gpu_ctx.upload(g)
    try:
        with pytest.raises(TgtpError, match="multiple of 8"):
            gpu_ctx.mt_walks(dg, 3, 12, WALK_MT_SIMD)
        with pytest.raises(TgtpError, match="depth above"):
            gpu_ctx.mt_walks(dg, 435, 8, WALK_MT_SIMD)
        with pytest.raises(TgtpError, match="depth above"):
            gpu_ctx.mt_walks(dg, 3473, 1, WALK_MT_REGULAR)
        with pytest.raises(TgtpError, match="variant"):
            gpu_ctx.mt_walks(dg, 3, 8, 7)
        with pytest.raises(TgtpError, match="source range"):
            gpu_ctx.mt_walks(dg, 3, 8, WALK_MT_REGULAR, 0, 10)
    finally:
        dg.free()


@pytest.mark.parametrize("variant", [WALK_MT_REGULAR, WALK_MT_SIMD])
def test_walker_stops_at_vertex_without_neighbours(gpu_ctx, variant):
    """A non-symmetric CSR through the C ABI: walkers that reach vertex 2
    (no neighbours) stop there, as in the oracle (itself pinned by a
    brute-force restatement, tests/test_mt_walks.py)."""
    from test_mt_walks import SINK_ADJ, SINK_IDX
    g = G.Graph(SINK_IDX, SINK_ADJ)
    dg = gpu_ctx.upload(g)
    try:
        for depth, n_walks in ((7, 24), (40, 256)):
            _say(f"mt variant {variant} sink graph: depth={depth} W={n_walks}")
            cnt, vc, cs, _ = gpu_ctx.mt_walks(dg, depth, n_walks, variant)
            ref = oracle.mt_walks(SINK_IDX, SINK_ADJ, depth, n_walks, variant=variant)
            assert np.array_equal(cnt, ref)
            assert (vc, cs) == _sums(ref)
    finally:
        dg.free()
