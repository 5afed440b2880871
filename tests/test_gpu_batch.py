"""Batched tgtp (tgtp_batch_*): a batch of graphs of uneven size scored as
one disjoint-union run must give each graph exactly the scores the per-graph
call gives with the same launch knobs, including through replays of the
captured HIP graph.  Through the C ABI."""
import sys

import numpy as np
import pytest

from oracle.exact import exact_resistance
from tgtp_amd import MODE_FAITHFUL, MODE_FAST_BF16, MODE_FAST_FP32, TgtpError
from tgtp_amd import eigenprep as E
from tgtp_amd import graph as G

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def small_batch():
    """Uneven sizes, exponents and degrees; one graph above the LDS limit,
    one slow-mixing grid with long walks."""
    rng = np.random.default_rng(5)
    graphs = []
    for k in range(9):
        n = int(2 ** rng.uniform(8, 11.5))
        graphs.append(G.chung_lu(n, float(rng.uniform(2.1, 3.0)), float(rng.uniform(6, 30)),
                                 seed=100 + k))
    graphs.append(G.grid_rewired(20, 20, 0.05, seed=3))
    graphs.append(G.chung_lu(9000, 2.5, 9.0, seed=77))
    eig = [E.eigens_f32(*E.compute_eigens(g, 24)) for g in graphs]
    return graphs, eig


def _per_graph(ctx, graphs, eig, eps, mode):
    out = []
    for g, (lam, X) in zip(graphs, eig):
        s, _ = ctx.score(g, lam, X, eps, 10, mode=mode)
        out.append(s)
    return out


@pytest.mark.parametrize("mode,name", [(MODE_FAST_FP32, "fp32"), (MODE_FAST_BF16, "bf16")])
def test_batch_equals_per_graph(gpu_ctx, small_batch, mode, name):
    graphs, eig = small_batch
    _say(f"batch {name}: {len(graphs)} graphs, n = {[g.n for g in graphs]}")
    ref = _per_graph(gpu_ctx, graphs, eig, 0.05, mode)
    outs, st = gpu_ctx.score_batch(graphs, eig, 0.05, 10, mode=mode)
    for k, (a, b) in enumerate(zip(outs, ref)):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), f"graph {k}"
    _say(f"  kernel_ms {st['kernel_ms']:.2f}")


def test_captured_graph_replays(gpu_ctx, small_batch):
    """Replays of the captured run, with host work and other launches in
    between, equal the uncaptured run every time."""
    graphs, eig = small_batch
    plain = gpu_ctx.batch(graphs, eig, 0.05, 10, mode=MODE_FAST_FP32)
    ref, _ = plain.run()
    plain.close()
    b = gpu_ctx.batch(graphs, eig, 0.05, 10, mode=MODE_FAST_FP32, graph=True)
    for rep in range(3):
        outs, st = b.run()
        for k, (a, r) in enumerate(zip(outs, ref)):
            assert np.array_equal(a.view(np.uint32), r.view(np.uint32)), f"replay {rep} graph {k}"
        gpu_ctx.score(graphs[0], *eig[0], 0.05, 10, mode=MODE_FAST_FP32)  # interleaved work
        _say(f"  replay {rep}: {st['kernel_ms']:.2f} ms")
    b.close()


@pytest.mark.parametrize("streams", ["1", "2"])
def test_captured_replays_on_several_subsets(gpu_ctx, small_batch, monkeypatch, streams):
    """Small state subsets (1 MiB) split the batch into many subsets: on two
    streams the capture is one graph per stream share, replayed with the
    fork / join events between them.  Every replay, with host work and other
    launches in between, equals the uncaptured run bit for bit."""
    graphs, eig = small_batch
    monkeypatch.setenv("TGTP_FAST_STATE_MIB", "1")
    monkeypatch.setenv("TGTP_BATCH_STREAMS", streams)
    plain = gpu_ctx.batch(graphs, eig, 0.05, 10, mode=MODE_FAST_FP32)
    ref, _ = plain.run()
    plain.close()
    b = gpu_ctx.batch(graphs, eig, 0.05, 10, mode=MODE_FAST_FP32, graph=True)
    for rep in range(3):
        outs, st = b.run()
        for k, (a, r) in enumerate(zip(outs, ref)):
            assert np.array_equal(a.view(np.uint32), r.view(np.uint32)), f"replay {rep} graph {k}"
        gpu_ctx.score(graphs[1], *eig[1], 0.05, 10, mode=MODE_FAST_FP32)  # interleaved work
        _say(f"  streams={streams} replay {rep}: {st['kernel_ms']:.2f} ms")
    b.close()
    # and equal to the per-graph calls
    for k, (g, e) in enumerate(zip(graphs, eig)):
        s, _ = gpu_ctx.score(g, *e, 0.05, 10, mode=MODE_FAST_FP32)
        assert np.array_equal(s.view(np.uint32), ref[k].view(np.uint32)), f"graph {k}"


def test_autotuned_batch_accuracy(gpu_ctx, small_batch):
    graphs, eig = small_batch
    b = gpu_ctx.batch(graphs, eig, 0.05, 10, mode=MODE_FAST_FP32, autotune=True, graph=True)
    cfg = b.config()
    outs, st = b.run()
    b.close()
    _say(f"  autotuned config {cfg}, run {st['kernel_ms']:.2f} ms, tuning {st['taus_ms']:.0f} ms")
    assert cfg["light_degree"] in (16, 32, 64) and 1 <= cfg["slots"] <= cfg["max_slots"]
    import oracle

    for g, (lam, X), s in zip(graphs, eig, outs):
        R, _ = exact_resistance(g.idx, g.adj)
        ferr = np.abs(oracle.tgtp(g.idx, g.adj, lam, X, 0.05, 10) - R)
        err = np.abs(s - R)
        assert err.max() <= ferr.max() + 0.05
        assert err.mean() <= ferr.mean() + 0.01


def test_faithful_batch_is_bitwise_per_graph(gpu_ctx, small_batch):
    graphs, eig = small_batch
    sub = graphs[:4]
    ref = _per_graph(gpu_ctx, sub, eig[:4], 0.05, MODE_FAITHFUL)
    outs, _ = gpu_ctx.score_batch(sub, eig[:4], 0.05, 10, mode=MODE_FAITHFUL)
    for a, b in zip(outs, ref):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_batch_rejects_mixed_inputs(gpu_ctx, small_batch):
    graphs, eig = small_batch
    lam, X = eig[0]
    with pytest.raises(TgtpError, match="one omega"):
        gpu_ctx.score_batch(graphs[:2], [eig[0], (lam[:16], X[:, :16])], 0.05, 10)
    with pytest.raises(TgtpError, match="epsilon"):
        gpu_ctx.score_batch(graphs[:2], eig[:2], 1.5, 10)
    # an id past its own graph (a valid id of the union) is refused per graph
    g0 = graphs[0]
    adj = g0.adj.copy()
    adj[-1] = g0.n
    with pytest.raises(TgtpError, match="graph 0: neighbour id >= n"):
        gpu_ctx.score_batch([G.Graph(g0.idx, adj), graphs[1]], eig[:2], 0.05, 10)
    # a missing reverse edge is refused for the union as for one graph
    c = G.cycle(12)
    assert c.adj[c.idx[0]:c.idx[1]].tolist() == [1, 11]
    adj = c.adj.copy()
    adj[c.idx[0] + 1] = 10  # 0 -> 10, but 10 -/-> 0
    e3 = (np.array([1.0, 0.5, 0.2]), np.ones((12, 3), dtype=np.float32))
    with pytest.raises(TgtpError, match="both directions"):
        gpu_ctx.score_batch([G.Graph(c.idx, adj), c], [e3, e3], 0.05, 10)


def test_batch_on_two_streams_equals_per_graph(gpu_ctx, small_batch, monkeypatch):
    """Small state buffers split the batch into many subsets, which a batch
    runs alternately on two streams (captured as a fork-join graph); the
    scores still equal the per-graph ones bit for bit."""
    graphs, eig = small_batch
    monkeypatch.setenv("TGTP_FAST_STATE_MIB", "8")
    ref = []
    for g, (lam, X) in zip(graphs, eig):
        s, _ = gpu_ctx.score(g, lam, X, 0.05, 10, mode=MODE_FAST_FP32)
        ref.append(s)
    b = gpu_ctx.batch(graphs, eig, 0.05, 10, mode=MODE_FAST_FP32, graph=True)
    for rep in range(2):
        outs, _ = b.run()
        for k, (a, r) in enumerate(zip(outs, ref)):
            assert np.array_equal(a.view(np.uint32), r.view(np.uint32)), f"replay {rep} graph {k}"
    b.close()
