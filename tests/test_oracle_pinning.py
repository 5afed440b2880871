"""Pins the oracle against things other than itself: recorded known answers
of the reference, published hash test vectors, closed forms and invariants.
CPU only."""
import json
import math

import numpy as np
import pytest

import oracle
from conftest import GOLDEN
from tgtp_amd import graph as G


def _checksum(c):
    c = c.astype(np.uint64)
    return int(c.sum()), int((np.arange(1, c.shape[0] + 1, dtype=np.uint64) * c).sum())


@pytest.mark.parametrize("case", json.loads((GOLDEN / "saba_kats.json").read_text())["cases"],
                         ids=lambda c: f"{c['graph']}{c['n']}")
def test_saba_known_answers(case):
    """Recorded outputs of the reference's saba_randomwalk; both graphs have
    power-of-two degrees, so every bouquet ends in the SplitMix32 fallback."""
    g = G.cycle(case["n"]) if case["graph"] == "cycle" else G.complete(case["n"])
    c = oracle.saba_walks(g.idx, g.adj, case["depth"], case["n_walks"], threads=1)
    assert _checksum(c) == (case["visit_count"], case["checksum"])


def test_fmix32_published_vectors():
    # MurmurHash3_x86_32 of the empty key is fmix32(seed): SMHasher vectors
    assert oracle.fmix32(0) == 0
    assert oracle.fmix32(1) == 0x514E28B7
    assert oracle.fmix32(0xFFFFFFFF) == 0x81F16F39


def test_fmix32_inverse_with_published_constants():
    """fmix32 is a bijection; undoing it with the published modular inverses
    of its two multipliers pins both multipliers and all three shifts."""
    inv1, inv2 = 0xA5CB9243, 0x7ED1B41D
    assert (0x85EBCA6B * inv1) % 2**32 == 1
    assert (0xC2B2AE35 * inv2) % 2**32 == 1

    def unfmix(h):
        h ^= h >> 16
        h = (h * inv2) % 2**32
        h ^= (h >> 13) ^ (h >> 26)
        h = (h * inv1) % 2**32
        h ^= h >> 16
        return h

    rng = np.random.default_rng(7)
    for x in rng.integers(0, 2**32, size=2000, dtype=np.uint64).tolist():
        assert unfmix(oracle.fmix32(x)) == x


def test_visit_count_invariant_and_isolated_vertices():
    # vertices 3 and 7 isolated: skipped as sources, never reached
    g = G.Graph.from_edges(9, [(0, 1), (1, 2), (2, 0), (4, 5), (5, 6), (6, 8), (8, 4), (1, 4)])
    deg = g.degree()
    assert deg[3] == 0 and deg[7] == 0
    for depth, w in [(5, 64), (15, 8), (33, 256)]:
        c = oracle.saba_walks(g.idx, g.adj, depth, w)
        assert int(c.sum()) == int((deg > 0).sum()) * w * depth
        assert c[3] == 0 and c[7] == 0


def test_star_closed_form():
    """Leaves have degree 1, so every step from a leaf returns to the hub and
    walks alternate hub/leaf whatever the hash picks."""
    n, W = 17, 64
    g = G.Graph.from_edges(n, [(0, i) for i in range(1, n)])
    for L in (1, 4, 7):
        c = oracle.saba_walks(g.idx, g.adj, L, W)
        hub = W * (L // 2) + (n - 1) * W * ((L + 1) // 2)
        assert int(c[0]) == hub
        assert int(c.sum()) == n * W * L


def test_thread_count_and_chunk_do_not_change_counters():
    g = G.chung_lu(3000, 2.5, 12.0, seed=9)
    a = oracle.saba_walks(g.idx, g.adj, 10, 128, threads=1)
    b = oracle.saba_walks(g.idx, g.adj, 10, 128, threads=4, chunk=7)
    assert np.array_equal(a, b)


def test_source_ranges_add_up():
    g = G.chung_lu(1500, 2.5, 10.0, seed=2)
    full = oracle.saba_walks(g.idx, g.adj, 6, 64)
    part = oracle.saba_walks(g.idx, g.adj, 6, 64, 0, 700).astype(np.uint64)
    part += oracle.saba_walks(g.idx, g.adj, 6, 64, 700, 1500)
    assert np.array_equal(full.astype(np.uint64), part)


def test_walk2_hashed_phase_equals_two_one_way_walks():
    """With all degrees odd no secret can reach zero, so the 2-way walk never
    falls back and each walker follows the 1-way walk from its endpoint on
    the same grid.  Integer-valued pvec keeps the float sum exact, so
    x = sum_u pvec[u] * (count_s[u] - count_v[u])."""
    g = G.petersen()  # 3-regular
    pvec = np.arange(1, g.n + 1, dtype=np.float32)
    for s, v, nr, L in [(0, 1, 16, 5), (0, 5, 64, 9), (3, 8, 8, 30)]:
        cs = oracle.saba_walk_source(g.idx, g.adj, s, L, nr).astype(np.int64)
        cv = oracle.saba_walk_source(g.idx, g.adj, v, L, nr).astype(np.int64)
        x = oracle.walk2_saba(g.idx, g.adj, s, v, nr, L, pvec)
        assert x == float((pvec.astype(np.int64) * (cs - cv)).sum())


def test_walk2_identical_endpoints_cancel_before_fallback():
    g = G.petersen()
    pvec = np.linspace(0.1, 1.0, g.n, dtype=np.float32)
    assert oracle.walk2_saba(g.idx, g.adj, 4, 4, 32, 12, pvec) == 0.0


def test_walk2_fallback_draws_both_walkers_from_one_state():
    """aesc.h:339-340: in the fallback both walkers draw from one SplitMix32
    state, one after the other.  From identical endpoints the hashed phase
    keeps the two walkers together (x stays 0: on C9 every secret reaches 0
    within 32 hops, before which the sum is exactly 0); once the bouquet
    falls back they take different draws and part, so x != 0.  Two equal
    copies of the state (a plausible misreading) would keep x == 0.  Which
    walker takes the first of the two draws stays unpinned: swapping them
    only negates x here."""
    g = G.cycle(9)
    pvec = np.linspace(0.1, 1.0, g.n, dtype=np.float32)
    for nr, L in [(64, 60), (16, 40), (8, 33)]:
        assert oracle.walk2_saba(g.idx, g.adj, 4, 4, nr, 20, pvec) == 0.0
        assert oracle.walk2_saba(g.idx, g.adj, 4, 4, nr, L, pvec) != 0.0


def test_pow_square_compiles_to_product():
    """aesc.h:498 squares with std::pow(x, 2).  g++ folds that to x*x (the
    correctly rounded square), so an IEEE build of the reference -- and this
    oracle -- never calls glibc pow, which is NOT always correctly rounded
    (e.g. pow(9413.715629082753, 2) != x*x).  The device computes x*x too.
    This pins that no pow call survives in the oracle or the reference build."""
    import shutil
    import subprocess

    x = 9413.715629082753
    assert math.pow(x, 2) != x * x  # glibc pow differs here: the fold matters
    objdump = shutil.which("objdump")
    if objdump is None:
        pytest.skip("objdump not available")
    bins = [oracle.LIB_PATH] + ([oracle.REF_TGTP_MT] if oracle.REF_TGTP_MT.exists() else [])
    for b in bins:
        text = subprocess.run([objdump, "-d", str(b)], capture_output=True, text=True).stdout
        assert "<pow@plt>" not in text and "<log@plt>" in text


def test_exact_closed_forms():
    from oracle.exact import exact_resistance

    for n in (5, 8):
        g = G.complete(n)
        R, _ = exact_resistance(g.idx, g.adj)
        assert np.allclose(R, 2.0 / n, rtol=0, atol=1e-12)
    for n in (9, 15):
        g = G.cycle(n)
        R, _ = exact_resistance(g.idx, g.adj)
        assert np.allclose(R, (n - 1) / n, rtol=0, atol=1e-12)
    g = G.petersen()
    R, _ = exact_resistance(g.idx, g.adj)
    assert np.allclose(R, 0.6, rtol=0, atol=1e-12)


def test_exact_foster_and_cg_agree():
    from oracle.exact import exact_resistance, sampled_cg

    g = G.grid_rewired(20, 20, 0.05, seed=4)
    R, info = exact_resistance(g.idx, g.adj)
    assert info["foster_error"] < 1e-8
    assert abs(R.sum() / 2 - (g.n - 1)) < 1e-8
    pos = np.arange(0, g.m, 97)
    Rc = sampled_cg(g.idx, g.adj, pos)
    assert np.allclose(Rc, R[pos], rtol=1e-9, atol=1e-11)


# a non-symmetric CSR: 0 -> {1, 2, 3, 4}, 1 -> {0, 5}, 2 -> {0, 5}, 3 -> {0},
# 4 -> {} and 6 -> {} (no neighbours), 5 -> {1, 2, 6}
SABA_SINK_IDX = np.array([0, 4, 6, 8, 9, 9, 12, 12], dtype=np.uint32)
SABA_SINK_ADJ = np.array([1, 2, 3, 4, 0, 5, 0, 5, 0, 1, 2, 6], dtype=np.uint32)


def _py_fmix32(h):
    h ^= h >> 16
    h = (h * 0x85EBCA6B) & 0xFFFFFFFF
    h ^= h >> 13
    h = (h * 0xC2B2AE35) & 0xFFFFFFFF
    return h ^ (h >> 16)


def _py_splitmix32(s):
    s = (s + 0x9E3779B9) & 0xFFFFFFFF
    z = s ^ (s >> 16)
    z = (z * 0x21F0AAAD) & 0xFFFFFFFF
    z ^= z >> 15
    z = (z * 0x735A2D97) & 0xFFFFFFFF
    return z ^ (z >> 15), s


def saba_sink_brute_force(idx, adj, depth, n_walks):
    """rw_experiment_saba.cpp:65-145 restated, with the rule for a walker
    that reaches a vertex without neighbours: it stops moving and counting
    and leaves its bouquet's zero check.  Returns (counts, bouquets that fell
    back, walkers stopped, hops where only a stopped walker's secret was 0)."""
    n = idx.shape[0] - 1
    deg = lambda v: int(idx[v + 1] - idx[v])  # noqa: E731
    cnt = np.zeros(n, dtype=np.uint32)
    step = 0xFFFFFFFF // n_walks
    fell = stopped = dead_zero = 0
    for s in range(n):
        if deg(s) == 0:
            continue
        for ii in range(0, n_walks, 8):
            cur = [s] * 8
            sec = [((ii + b + 1) * step) & 0xFFFFFFFF for b in range(8)]
            st = list(sec)
            live = [True] * 8
            fb = False
            for _ in range(depth):
                live = [live[b] and deg(cur[b]) > 0 for b in range(8)]
                if not fb and any(live[b] and sec[b] == 0 for b in range(8)):
                    fb = True
                    fell += 1
                elif not fb and any(not live[b] and sec[b] == 0 for b in range(8)):
                    dead_zero += 1
                for b in range(8):
                    if not live[b]:
                        continue
                    d = deg(cur[b])
                    if not fb:
                        k = ((sec[b] ^ _py_fmix32(cur[b])) * d) >> 32
                        sec[b] = (sec[b] * d) & 0xFFFFFFFF
                    else:
                        r, st[b] = _py_splitmix32(st[b])
                        k = r % d
                    cur[b] = int(adj[idx[cur[b]] + k])
                    cnt[cur[b]] += 1
            stopped += sum(not x for x in live)
    return cnt, fell, stopped, dead_zero


@pytest.mark.parametrize("depth,n_walks,falls,dead_zero", [
    (12, 16, False, False), (40, 64, True, False), (60, 256, True, True)])
def test_saba_walker_stops_at_vertex_without_neighbours(depth, n_walks, falls, dead_zero):
    exp, fell, stopped, dz = saba_sink_brute_force(SABA_SINK_IDX, SABA_SINK_ADJ, depth, n_walks)
    got = oracle.saba_walks(SABA_SINK_IDX, SABA_SINK_ADJ, depth, n_walks, threads=1)
    assert np.array_equal(got, exp)
    # the rules each case exercises
    assert stopped > 0 and (fell > 0) == falls and (dz > 0) == dead_zero


@pytest.mark.parametrize("case", ["c9", "k5", "sink", "cl", "grid"])
def test_avx2_walks_equal_scalar(case):
    """The AVX2 host baseline gives the scalar oracle's counters bit for bit,
    also where the whole-bouquet fallback runs (C9, K5), where walkers stop
    at a vertex without neighbours (non-symmetric CSR) and on larger graphs,
    on several threads and chunk sizes."""
    if case == "c9":
        g, depth, W = G.cycle(9), 40, 64
    elif case == "k5":
        g, depth, W = G.complete(5), 20, 64
    elif case == "sink":
        g, depth, W = G.Graph(SABA_SINK_IDX, SABA_SINK_ADJ), 40, 64
    elif case == "cl":
        g, depth, W = G.chung_lu(3000, 2.3, 14.0, seed=4), 15, 256
    else:
        g, depth, W = G.grid_rewired(40, 40, 0.05, seed=2), 30, 128
    ref = oracle.saba_walks(g.idx, g.adj, depth, W)
    try:
        for threads, chunk in ((1, 1024), (4, 7), (8, 1)):
            got = oracle.saba_walks_avx2(g.idx, g.adj, depth, W, threads=threads, chunk=chunk)
            assert np.array_equal(got, ref), f"{case} threads={threads} chunk={chunk}"
    except RuntimeError:
        pytest.skip("no AVX2 on this host")


def walk2_brute_force(idx, adj, s, v, nr, length, pvec, swap=False):
    """randomwalk_2way_avx2 (aesc.h:241-391) restated lane by lane from the
    reference's text: both walkers' grid secrets start equal (:259-266); the
    whole bouquet leaves the hashed phase when any lane of either secret is
    zero (:272-277); the fallback draws two SplitMix32 outputs per lane from
    one state, the first for the walker from s, the second for the walker
    from v (:339-340; swap=True gives the other order); each lane adds
    pvec[cur] - pvec[cur2] in float32 and the lanes are summed 0..7
    (:320-330, 386-390).  Returns (x, whether some bouquet fell back)."""
    deg = lambda u: int(idx[u + 1] - idx[u])  # noqa: E731
    f32 = np.float32
    step = (0xFFFFFFFF // nr) & 0xFFFFFFFF
    x = [f32(0.0)] * 8
    fell = False
    for i in range(0, nr, 8):
        c1, c2 = [s] * 8, [v] * 8
        s1 = [((i + b + 1) * step) & 0xFFFFFFFF for b in range(8)]
        s2 = list(s1)
        st = list(s1)
        hop = 0
        while hop < length and not any(a == 0 for a in s1 + s2):
            for b in range(8):
                d1, d2 = deg(c1[b]), deg(c2[b])
                k1 = ((s1[b] ^ _py_fmix32(c1[b])) * d1) >> 32
                k2 = ((s2[b] ^ _py_fmix32(c2[b])) * d2) >> 32
                c1[b] = int(adj[idx[c1[b]] + k1])
                c2[b] = int(adj[idx[c2[b]] + k2])
                s1[b] = (s1[b] * d1) & 0xFFFFFFFF
                s2[b] = (s2[b] * d2) & 0xFFFFFFFF
                x[b] = f32(x[b] + f32(pvec[c1[b]] - pvec[c2[b]]))
            hop += 1
        fell |= hop < length
        for _ in range(hop, length):
            for b in range(8):
                r1, st[b] = _py_splitmix32(st[b])
                r2, st[b] = _py_splitmix32(st[b])
                if swap:
                    r1, r2 = r2, r1
                c1[b] = int(adj[idx[c1[b]] + r1 % deg(c1[b])])
                c2[b] = int(adj[idx[c2[b]] + r2 % deg(c2[b])])
                x[b] = f32(x[b] + f32(pvec[c1[b]] - pvec[c2[b]]))
    total = f32(0.0)
    for b in range(8):
        total = f32(total + x[b])
    return total, fell


def _star(leaves):
    edges = [(0, k) for k in range(1, leaves + 1)]
    return G.Graph.from_edges(leaves + 1, edges)


def _hypercube(dim):
    n = 1 << dim
    return G.Graph.from_edges(n, [(u, u ^ (1 << b)) for u in range(n) for b in range(dim)
                                  if u < u ^ (1 << b)])


@pytest.mark.parametrize("graph", ["star16", "q4"])
def test_walk2_fallback_draw_order_pinned(graph):
    """Graphs of power-of-two degrees drive every secret to zero (the secret
    is multiplied by the degree each hop), so every bouquet reaches the
    fallback.  On the 16-leaf star one walker is always at the hub and the
    other at a leaf (a forced step), so the hub walker's draw -- the first or
    the second of the hop -- decides the leaf it visits; on the 4-cube both
    draw.  The oracle equals the restatement of the reference's order, and
    the restatement with the two draws swapped differs from it, so a swap in
    the oracle would fail here."""
    g = _star(16) if graph == "star16" else _hypercube(4)
    rng = np.random.default_rng(17)
    pvec = rng.random(g.n).astype(np.float32)
    pairs = [(0, 3), (5, 0)] if graph == "star16" else [(0, 1), (6, 14)]
    swapped_differs = 0
    for s, v in pairs:
        for nr in (8, 64):
            exp, fell = walk2_brute_force(g.idx, g.adj, s, v, nr, 40, pvec)
            got = oracle.walk2_saba(g.idx, g.adj, s, v, nr, 40, pvec)
            assert fell
            assert np.float32(got).view(np.uint32) == exp.view(np.uint32), (s, v, nr)
            alt, _ = walk2_brute_force(g.idx, g.adj, s, v, nr, 40, pvec, swap=True)
            swapped_differs += int(alt != exp)
    assert swapped_differs == 2 * len(pairs)


def test_simd_comparator_hop_order_discriminates():
    """The SIMD comparator draws 8 outputs per hop, one per lane in lane
    order (rw_experiment_simd.cpp:55-57): hop h of bouquet b takes draws
    b*8*depth + 8h + lane.  The oracle equals the brute-force restatement of
    that order on the non-symmetric test graph (test_mt_walks) at depth 7;
    here the restatement with the regular comparator's order (walk-major,
    w*depth + h) gives different counts on a Chung-Lu graph at depth 6, so
    the order is what the equality checks."""
    from test_mt_walks import _mt_stream

    g = G.chung_lu(60, 2.4, 5.0, seed=9)
    depth, W = 6, 16

    def brute(simd_order):
        cnt = np.zeros(g.n, dtype=np.uint32)
        for s in range(g.n):
            draws = _mt_stream(42 + s, W * depth)
            for w in range(W):
                cur = s
                for h in range(depth):
                    o = (w // 8) * 8 * depth + 8 * h + w % 8 if simd_order else w * depth + h
                    d = int(g.idx[cur + 1] - g.idx[cur])
                    cur = int(g.adj[g.idx[cur] + int(draws[o]) % d])
                    cnt[cur] += 1
        return cnt

    got = oracle.mt_walks(g.idx, g.adj, depth, W, variant=1)
    assert np.array_equal(got, brute(True))
    assert not np.array_equal(got, brute(False))
