"""tgtp_amd.exact_sc (the exact_sc tool, src/exact_sc.cpp) against the
oracle's FP64 Cholesky resistances and closed forms."""
import numpy as np
import pytest

from oracle.exact import exact_resistance
from tgtp_amd import exact_sc as X
from tgtp_amd import graph as G


@pytest.mark.parametrize("dev", ["cpu", pytest.param("cuda:0", marks=pytest.mark.gpu)])
def test_closed_forms(dev):
    assert np.allclose(X.exact_scores(G.petersen(), dev), 0.6, atol=1e-12)
    assert np.allclose(X.exact_scores(G.complete(9), dev), 2.0 / 9, atol=1e-12)
    assert np.allclose(X.exact_scores(G.cycle(11), dev), 10.0 / 11, atol=1e-12)


@pytest.mark.parametrize("dev", ["cpu", pytest.param("cuda:0", marks=pytest.mark.gpu)])
def test_dense_and_sampled_match_oracle(dev):
    g = G.chung_lu(700, 2.3, 6.0, seed=3)
    R, _ = exact_resistance(g.idx, g.adj)
    r = X.exact_scores(g, dev)
    assert np.abs(r - R).max() < 1e-10
    # Foster: sum over undirected edges of R = n - 1 (both directions in CSR)
    assert abs(r.sum() / 2 - (g.n - 1)) < 1e-7
    pos = np.random.default_rng(1).choice(g.m, 90, replace=False)
    s = X.sampled_scores(g, pos, dev, chunk=32)
    assert np.abs(s - R[pos]).max() < 1e-9


def test_rejects_isolated_and_tiny():
    g = G.Graph.from_edges(4, [(0, 1), (1, 2)])
    with pytest.raises(ValueError, match="isolated"):
        X.exact_scores(g, "cpu")


def test_cli_writes_scores_bin(tmp_path):
    g = G.petersen()
    p = tmp_path / "graph.txt"
    g.write_txt(p)
    assert X.main([str(p), str(tmp_path / "s.bin"), "--device", "cpu"]) == 0
    s = np.fromfile(tmp_path / "s.bin", dtype=np.float32)
    assert s.shape == (g.m,) and np.allclose(s, 0.6)
