"""The block Krylov eigensolver of tgtp_amd.eigenprep against LAPACK's
dense eigh on the same normalized adjacency."""
import numpy as np
import pytest

from tgtp_amd import eigenprep as E
from tgtp_amd import graph as G


def _check(w1, X1, w, X):
    assert np.abs(w1 - w).max() < 1e-9
    cos = np.abs((X1 * X).sum(0)) / (np.linalg.norm(X1, axis=0) * np.linalg.norm(X, axis=0))
    # eigenvectors of simple eigenvalues agree up to sign
    assert np.abs(cos - 1).max() < 1e-6


@pytest.mark.parametrize("dev", ["cpu", pytest.param("cuda:0", marks=pytest.mark.gpu)])
def test_block_krylov_matches_dense(dev):
    g = G.chung_lu(1500, 2.4, 10.0, seed=11)
    w, X = E.compute_eigens(g, 40)            # dense LAPACK (n <= 2500)
    w1, X1 = E.compute_eigens_device(g, 40, device=dev)
    _check(w1, X1, w, X)


def test_block_krylov_small_blocks_and_restarts():
    # a basis of 96 columns forces several thick restarts
    import torch

    g = G.grid_rewired(20, 30, 0.1, seed=2)
    P, _, _ = E.normalized_adjacency(g)
    w, V = np.linalg.eigh(P.toarray())
    order = np.argsort(-np.abs(w), kind="stable")[:24]
    th, Y = E.block_krylov_eigs(E._torch_csr(P, "cpu"), 24, block=16, basis=96, keep=48)
    assert np.abs(th.numpy() - w[order]).max() < 1e-9
    PY = P @ Y.numpy()
    assert np.abs(PY - Y.numpy() * th.numpy()).max() < 1e-8


@pytest.mark.gpu
def test_device_eigens_large_sparse():
    g = G.chung_lu(12000, 2.8, 30.0, seed=4)
    w, X = E.compute_eigens(g, 128)           # ARPACK
    w1, X1 = E.compute_eigens_device(g, 128, device="cuda:0")
    assert np.abs(w1 - w).max() < 1e-9


@pytest.mark.parametrize("dev", ["cpu", pytest.param("cuda:0", marks=pytest.mark.gpu)])
def test_batched_size_class_matches_per_graph(dev):
    """Graphs of different sizes solved as one padded block-diagonal class
    give each graph its own eigenpairs."""
    gs = [G.chung_lu(n, b, d, seed=s) for n, b, d, s in
          [(1100, 2.5, 10.0, 1), (1300, 2.2, 30.0, 2), (1500, 2.9, 8.0, 3), (2400, 2.5, 20.0, 4)]]
    res = E.compute_eigens_many(gs, 48, device=dev)
    for g, (w1, X1) in zip(gs, res):
        _check(w1, X1, *E.compute_eigens(g, 48))
