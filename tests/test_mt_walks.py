"""The oracle's MT19937 comparator walks (rw_experiment.cpp /
rw_experiment_simd.cpp) against the reference's own scalar_randomwalk
(fixtures written by tests/golden/make_golden.py) and the C++ standard's
known answer for std::mt19937.  CPU only."""
import numpy as np
import pytest

import oracle
from conftest import MT_WALK_CASES, load_mt_walk_golden


def test_mt19937_standard_known_answer():
    # [rand.predef]: the 10000th consecutive invocation of a default-
    # constructed mt19937 (seed 5489) produces 4123659995
    assert oracle.mt19937_nth(5489, 10000) == 4123659995
    assert oracle.mt19937_nth(5489, 1) == 3499211612  # first output, seed 5489


@pytest.mark.parametrize("name", MT_WALK_CASES)
def test_regular_walks_match_reference(name):
    z = load_mt_walk_golden(name)
    for depth, n_walks in z["runs"]:
        c = oracle.mt_walks(z["idx"], z["adj"], int(depth), int(n_walks), variant=0)
        assert np.array_equal(c, z[f"counts_d{depth}_w{n_walks}"]), (depth, n_walks)


@pytest.mark.parametrize("name", MT_WALK_CASES)
def test_simd_walks_invariants(name):
    """simd draws output 8b+j for walk 8b+j at hop 0, exactly what the
    regular walk draws: at depth 1 both variants agree; every walk counts
    one visit per hop from every source with an edge."""
    z = load_mt_walk_golden(name)
    idx, adj = z["idx"], z["adj"]
    deg = np.diff(idx.astype(np.int64))
    for n_walks in (8, 40):
        a = oracle.mt_walks(idx, adj, 1, n_walks, variant=0)
        b = oracle.mt_walks(idx, adj, 1, n_walks, variant=1)
        assert np.array_equal(a, b)
    for depth in (3, 7):
        c = oracle.mt_walks(idx, adj, depth, 16, variant=1)
        assert int(c.sum()) == int((deg > 0).sum()) * 16 * depth
        d = oracle.mt_walks(idx, adj, depth, 16, variant=0)
        assert not np.array_equal(c, d) or depth == 1 or len(idx) < 12


def test_walks_thread_independent_and_ranges_add_up():
    z = load_mt_walk_golden("chunglu2000")
    idx, adj = z["idx"], z["adj"]
    for variant in (0, 1):
        full = oracle.mt_walks(idx, adj, 5, 24, variant=variant, threads=1)
        assert np.array_equal(full, oracle.mt_walks(idx, adj, 5, 24, variant=variant, threads=4))
        parts = sum(oracle.mt_walks(idx, adj, 5, 24, variant=variant, src_begin=a, src_end=b)
                    for a, b in ((0, 700), (700, 1333), (1333, 2000)))
        assert np.array_equal(full, parts)


def test_simd_rejects_partial_bouquets():
    z = load_mt_walk_golden("cycle15")
    with pytest.raises(ValueError):
        oracle.mt_walks(z["idx"], z["adj"], 3, 12, variant=1)


def _twist_in_phases(mt, chunk=256):
    """The GPU kernel's twist (mt_walk.hip): words [0, 227), [227, 454),
    [454, 623) and 623 as four phases, each in chunks whose reads all happen
    before the chunk's writes."""
    N, M = 624, 397
    up, lo, a = np.uint32(0x80000000), np.uint32(0x7FFFFFFF), np.uint32(0x9908B0DF)
    for p_lo, p_hi in ((0, N - M), (N - M, 2 * (N - M)), (2 * (N - M), N - 1), (N - 1, N)):
        for c in range(p_lo, p_hi, chunk):
            i = np.arange(c, min(c + chunk, p_hi))
            y = (mt[i] & up) | (mt[(i + 1) % N] & lo)
            new = mt[(i + M) % N] ^ (y >> np.uint32(1)) ^ np.where(y & np.uint32(1), a, np.uint32(0))
            mt[i] = new.astype(np.uint32)


def _temper(y):
    y = y ^ (y >> np.uint32(11))
    y = y ^ ((y << np.uint32(7)) & np.uint32(0x9D2C5680))
    y = y ^ ((y << np.uint32(15)) & np.uint32(0xEFC60000))
    return y ^ (y >> np.uint32(18))


@pytest.mark.parametrize("seed", [5489, 42, 42 + 4095])
@pytest.mark.parametrize("chunk", [256, 64])  # the kernel's workgroup, and multi-chunk phases
def test_phased_twist_matches_std_mt19937(seed, chunk):
    mt = np.zeros(624, dtype=np.uint32)
    x = seed
    mt[0] = x
    for i in range(1, 624):  # std::mt19937::seed
        x = (1812433253 * (x ^ (x >> 30)) + i) & 0xFFFFFFFF
        mt[i] = x
    outs = []
    for _ in range(17):
        _twist_in_phases(mt, chunk)
        outs.append(_temper(mt.copy()))
    outs = np.concatenate(outs)
    for k in (1, 2, 227, 228, 454, 455, 623, 624, 625, 1248, 1249, 5000, 10000):
        assert int(outs[k - 1]) == oracle.mt19937_nth(seed, k), k


def _mt_stream(seed, count):
    mt = np.zeros(624, dtype=np.uint32)
    x = seed
    mt[0] = x
    for i in range(1, 624):
        x = (1812433253 * (x ^ (x >> 30)) + i) & 0xFFFFFFFF
        mt[i] = x
    outs = []
    while sum(o.shape[0] for o in outs) < count:
        _twist_in_phases(mt)
        outs.append(_temper(mt.copy()))
    return np.concatenate(outs)[:count]


# a non-symmetric CSR: 0 -> {1, 2}, 1 -> {0}, 2 -> {} (no neighbours),
# 3 -> {0, 4}, 4 -> {3}
SINK_IDX = np.array([0, 2, 3, 3, 5, 6], dtype=np.uint32)
SINK_ADJ = np.array([1, 2, 0, 0, 4, 3], dtype=np.uint32)


def sink_walks_brute_force(depth, n_walks, variant):
    """The walk rule restated on the draws of each source's own stream: a
    walker at a vertex without neighbours stops; its draws stay taken."""
    n = SINK_IDX.shape[0] - 1
    cnt = np.zeros(n, dtype=np.uint32)
    for s in range(n):
        if SINK_IDX[s + 1] == SINK_IDX[s]:
            continue
        draws = _mt_stream(42 + s, n_walks * depth)
        for w in range(n_walks):
            cur = s
            for h in range(depth):
                o = w * depth + h if variant == 0 else (w // 8) * 8 * depth + 8 * h + w % 8
                d = int(SINK_IDX[cur + 1] - SINK_IDX[cur])
                if d == 0:
                    break
                cur = int(SINK_ADJ[SINK_IDX[cur] + int(draws[o]) % d])
                cnt[cur] += 1
    return cnt


@pytest.mark.parametrize("variant", [0, 1])
def test_walker_stops_at_vertex_without_neighbours(variant):
    exp = sink_walks_brute_force(7, 24, variant)
    got = oracle.mt_walks(SINK_IDX, SINK_ADJ, 7, 24, variant=variant)
    assert np.array_equal(got, exp)
    assert exp[2] > 0  # some walkers did reach the vertex without neighbours
