"""tgtp_mt_walks_noalloc (the entry point that allocates nothing) against
tgtp_mt_walks, through the C ABI: equal counters, visit counts and checksums
on two calls in a row, the second on a smaller graph (the shared scratch
keeps the last call's contents, so every call must clear it), for both
comparator variants; a graph beyond the scratch is refused before any
kernel runs."""
import numpy as np
import pytest

import oracle
from tgtp_amd import WALK_MT_REGULAR, WALK_MT_SIMD, TgtpError
from tgtp_amd import graph as G
from tgtp_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("variant", [WALK_MT_REGULAR, WALK_MT_SIMD])
def test_noalloc_equals_mt_walks_two_calls(gpu_ctx, variant):
    cases = [(W.chung_lu_4096(), 5, 2048, 0, None), (G.chung_lu(600, 2.6, 9.0, seed=5), 9, 64, 17, 533)]
    for g, depth, walks, lo, hi in cases:
        hi = g.n if hi is None else hi
        dg = gpu_ctx.upload(g)
        try:
            ref = gpu_ctx.mt_walks(dg, depth, walks, variant, lo, hi)
            got = gpu_ctx.mt_walks(dg, depth, walks, variant, lo, hi, noalloc=True)
        finally:
            dg.free()
        assert np.array_equal(got[0], ref[0]), f"n={g.n}: counters differ"
        assert got[1:3] == ref[1:3], f"n={g.n}: visit count / checksum differ"
        if g.n <= 600:
            assert np.array_equal(got[0], oracle.mt_walks(g.idx, g.adj, depth, walks, variant,
                                                          lo, hi))


def test_noalloc_refuses_graph_beyond_scratch(gpu_ctx):
    n = (1 << 22) + 1  # a cycle, one vertex more than the scratch holds
    v = np.arange(n, dtype=np.int64)
    nb = np.sort(np.stack([(v - 1) % n, (v + 1) % n], axis=1), axis=1)
    g = G.Graph(np.arange(0, 2 * n + 1, 2, dtype=np.uint32), nb.reshape(-1).astype(np.uint32))
    dg = gpu_ctx.upload(g)
    try:
        with pytest.raises(TgtpError, match="at most 4194304"):
            gpu_ctx.mt_walks(dg, 3, 8, WALK_MT_REGULAR, 0, 16, noalloc=True)
        # the allocating form still takes it
        _, vc, _, _ = gpu_ctx.mt_walks(dg, 3, 8, WALK_MT_REGULAR, 0, 16)
        assert vc == 16 * 8 * 3
    finally:
        dg.free()
