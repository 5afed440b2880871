"""CPU model of the partitioned walk engine's decomposition
(tgtp_amd/csrc/saba_part.hip): every lane walks all its hops with its own
secret, ignoring the bouquet-wide zero check; a bouquet is marked when a
lane starts at secret 0 or its secret becomes 0 while hops remain; a replay
then walks each marked bouquet in lockstep to the hop t* at which the
reference's check fires (rw_experiment_saba.cpp:83-85) and, for every lane,
subtracts the independent visits from t* on and adds those of the
SplitMix32 fallback (rw_experiment_saba.cpp:117-143).  The model's counters
must equal the oracle's on graphs where few, most or all bouquets fall back;
the GPU tests check the engine itself against the oracle.  CPU only."""
import numpy as np
import pytest

import oracle
from test_oracle_pinning import SABA_SINK_ADJ, SABA_SINK_IDX
from tgtp_amd import graph as G

M32 = 0xFFFFFFFF


def _fmix32(h):
    h ^= h >> 16
    h = (h * 0x85EBCA6B) & M32
    h ^= h >> 13
    h = (h * 0xC2B2AE35) & M32
    return h ^ (h >> 16)


def _splitmix32(state):
    state = (state + 0x9E3779B9) & M32
    z = state ^ (state >> 16)
    z = (z * 0x21F0AAAD) & M32
    z ^= z >> 15
    z = (z * 0x735A2D97) & M32
    return state, z ^ (z >> 15)


def engine_model(idx, adj, depth, n_walks, lo, hi, mistake=None):
    """Counters by the engine's decomposition; also the numbers of marked
    and of repaired bouquets.  `mistake` plants an error in the repair (to
    show that the comparison detects it): "state" starts the SplitMix32
    walks from the secret at t* instead of the initial one, "late" repairs
    from t* + 1, "keep" adds the fallback visits without removing the
    independent ones."""
    idx = [int(x) for x in idx]
    adj = [int(x) for x in adj]
    n = len(idx) - 1
    deg = [idx[v + 1] - idx[v] for v in range(n)]
    counts = np.zeros(n, dtype=np.int64)
    step = M32 // n_walks

    def hashed(cur, sec):
        d = deg[cur]
        k = ((sec ^ _fmix32(cur)) * d) >> 32
        return adj[idx[cur] + k], (sec * d) & M32

    marked = []
    for s in range(lo, hi):
        for i0 in range(0, n_walks, 8):
            mark = False
            for b in range(8):  # the rounds: every lane on its own
                cur, sec = s, ((i0 + b + 1) * step) & M32
                mark |= sec == 0
                for h in range(depth):
                    if deg[cur] == 0:
                        break  # a walker at a vertex without neighbours stops
                    cur, ns = hashed(cur, sec)
                    mark |= ns == 0 and sec != 0 and h + 1 < depth
                    sec = ns
                    counts[cur] += 1
            if mark:
                marked.append((s, i0))

    repaired = 0
    for s, i0 in marked:  # the replay pass
        cur = [s] * 8
        sec = [((i0 + b + 1) * step) & M32 for b in range(8)]
        st = list(sec)
        live = [True] * 8
        t_star = None
        for h in range(depth):
            live = [live[b] and deg[cur[b]] != 0 for b in range(8)]
            if any(live[b] and sec[b] == 0 for b in range(8)):
                t_star = h
                break
            for b in range(8):
                if live[b]:
                    cur[b], sec[b] = hashed(cur[b], sec[b])
        if t_star is None:
            continue  # marked, but the check never fires (the lane stopped)
        repaired += 1
        if mistake == "late" and t_star + 1 < depth:
            for b in range(8):
                if live[b] and deg[cur[b]] != 0:
                    cur[b], sec[b] = hashed(cur[b], sec[b])
            t_star += 1
        for b in range(8):
            c, x = cur[b], sec[b]
            for _ in range(t_star, depth):  # independent visits: out
                if deg[c] == 0:
                    break
                c, x = hashed(c, x)
                counts[c] -= 0 if mistake == "keep" else 1
            c, state, alive = cur[b], (sec[b] if mistake == "state" else st[b]), live[b]
            for _ in range(t_star, depth):  # SplitMix32 visits: in
                alive = alive and deg[c] != 0
                if not alive:
                    break
                state, r = _splitmix32(state)
                c = adj[idx[c] + r % deg[c]]
                counts[c] += 1
    if mistake is None:
        assert counts.min() >= 0
    return counts, len(marked), repaired


@pytest.mark.parametrize("name,make,depth,n_walks,lo,hi,falls", [
    ("cycle9", lambda: G.cycle(9), 40, 64, 0, 9, "all"),
    ("complete5", lambda: G.complete(5), 20, 64, 0, 5, "all"),
    ("torus16", lambda: G.grid_rewired(16, 16, 0.0, seed=1), 40, 16, 0, 256, "all"),
    ("grid16_rewired_d15", lambda: G.grid_rewired(16, 16, 0.1, seed=2), 15, 64, 0, 256, "some"),
    ("chunglu300_d10", lambda: G.chung_lu(300, 2.3, 8.0, seed=3), 10, 256, 40, 140, "any"),
])
def test_engine_decomposition_equals_oracle(name, make, depth, n_walks, lo, hi, falls):
    g = make()
    got, n_marked, n_repaired = engine_model(g.idx, g.adj, depth, n_walks, lo, hi)
    ref = oracle.saba_walks(g.idx, g.adj, depth, n_walks, lo, hi)
    assert np.array_equal(got, ref.astype(np.int64)), f"{name}: {int((got != ref).sum())} counters differ"
    bouquets = (hi - lo) * n_walks // 8
    if falls == "all":
        assert n_repaired == bouquets
    elif falls == "some":
        assert 0 < n_repaired < bouquets, (n_repaired, bouquets)


@pytest.mark.parametrize("depth,n_walks", [(12, 16), (40, 64)])
def test_engine_decomposition_with_stopped_walkers(depth, n_walks):
    """Non-symmetric CSR: walkers reaching a vertex without neighbours stop
    and leave the bouquet's zero check; a bouquet marked by such a lane may
    need no repair."""
    idx, adj = SABA_SINK_IDX, SABA_SINK_ADJ
    n = idx.shape[0] - 1
    got, n_marked, n_repaired = engine_model(idx, adj, depth, n_walks, 0, n)
    ref = oracle.saba_walks(idx, adj, depth, n_walks, 0, n)
    assert np.array_equal(got, ref.astype(np.int64))
    assert n_marked >= n_repaired
    if depth >= 40:
        assert n_repaired > 0


@pytest.mark.parametrize("mistake", ["state", "late", "keep"])
def test_engine_decomposition_detects_a_wrong_repair(mistake):
    """The comparison has teeth: each planted repair mistake changes the
    counters on graphs where every bouquet falls back."""
    for g, depth in ((G.complete(5), 20), (G.cycle(9), 40)):
        ref = oracle.saba_walks(g.idx, g.adj, depth, 64, 0, g.n).astype(np.int64)
        got, _, _ = engine_model(g.idx, g.adj, depth, 64, 0, g.n, mistake=mistake)
        assert not np.array_equal(got, ref), f"{mistake} on n={g.n} went unnoticed"
