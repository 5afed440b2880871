"""FP64 exact spanning edge centrality (effective resistance) -- the accuracy
oracle.  TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates src/exact_sc.cpp:24-49 and scripts/paper_accuracy.py:13-54: ground
vertex n-1 of each connected component (paper_accuracy.py:37; exact_sc.cpp
grounds vertex 0 -- the resistances do not depend on the choice), factor the
grounded Laplacian with LAPACK Cholesky in FP64, R(u,v) = G[u,u] + G[v,v] -
2 G[u,v] for every CSR entry.  Unlike exact_sc (exact_sc.cpp:47) the result
stays FP64.

Pinned by: closed forms on edge-transitive graphs (K_n: 2/n, odd C_n:
(n-1)/n, Petersen: 3/5), Foster's theorem sum_undirected R = n - #components
(paper_accuracy.py:49-52), and the solve residual on probe vectors
(paper_accuracy.py:40-43).

`sampled_cg` is the large-n form (dense is O(n^2) memory): FP64 conjugate
gradients on the grounded Laplacian for a stated sample of edges.
"""
from __future__ import annotations

import numpy as np


def _components(idx, adj):
    import scipy.sparse as sp
    from scipy.sparse.csgraph import connected_components

    n = idx.shape[0] - 1
    A = sp.csr_matrix((np.ones(adj.shape[0]), adj.astype(np.int64), idx.astype(np.int64)),
                      shape=(n, n))
    return connected_components(A, directed=False)


def exact_resistance(idx, adj, check: bool = True):
    """R float64[m] in CSR order plus a validation dict."""
    from scipy.linalg import cho_factor, cho_solve

    idx = np.asarray(idx, dtype=np.int64)
    adj = np.asarray(adj, dtype=np.int64)
    n = idx.shape[0] - 1
    m = adj.shape[0]
    row = np.repeat(np.arange(n), np.diff(idx))
    ncomp, labels = _components(idx, adj)
    R = np.zeros(m)
    residuals = []
    for c in range(ncomp):
        nodes = np.flatnonzero(labels == c)
        k = nodes.shape[0]
        local = np.full(n, -1, dtype=np.int64)
        local[nodes] = np.arange(k)
        pos = np.flatnonzero(labels[row] == c)
        if k == 1:
            continue
        Lc = np.zeros((k, k))
        u, v = local[row[pos]], local[adj[pos]]
        np.add.at(Lc, (u, v), -1.0)
        np.add.at(Lc, (u, u), 1.0)
        # self-loops add +1 and -1 on the diagonal: no net effect, as in a Laplacian
        B = Lc[:-1, :-1]
        fac = cho_factor(B, lower=True)
        G = np.zeros((k, k))
        G[:-1, :-1] = cho_solve(fac, np.eye(k - 1))
        if check:
            probe = np.random.default_rng(42).normal(size=(k - 1, 4))
            residuals.append(float(np.max(np.abs(B @ cho_solve(fac, probe) - probe))))
        R[pos] = G[u, u] + G[v, v] - 2.0 * G[u, v]
    info = {"vertices": n, "directed_edges": m, "components": int(ncomp),
            "max_solve_residual": max(residuals) if residuals else 0.0}
    if check:
        loops = row == adj
        foster = abs(R[~loops].sum() / 2.0 - (n - ncomp))
        info["foster_error"] = float(foster)
        if not np.all(np.isfinite(R)) or info["max_solve_residual"] > 1e-6 or foster > 1e-5 * max(1, n):
            raise RuntimeError("exact reference failed residual/Foster checks")
    return R, info


def sampled_cg(idx, adj, positions, tol: float = 1e-12, maxiter: int = 100000):
    """R(u,v) for the CSR positions given, by FP64 CG on the Laplacian
    grounded at vertex n-1 (graph assumed connected)."""
    import scipy.sparse as sp
    from scipy.sparse.linalg import cg

    idx = np.asarray(idx, dtype=np.int64)
    adj = np.asarray(adj, dtype=np.int64)
    n = idx.shape[0] - 1
    row = np.repeat(np.arange(n), np.diff(idx))
    A = sp.csr_matrix((np.ones(adj.shape[0]), adj, idx), shape=(n, n))
    Lap = (sp.diags(np.asarray(A.sum(axis=1)).ravel()) - A).tocsr()
    B = Lap[:-1, :-1].tocsr()
    diag = B.diagonal()
    M = sp.diags(1.0 / diag)
    out = np.zeros(len(positions))
    for i, p in enumerate(positions):
        u, v = int(row[p]), int(adj[p])
        b = np.zeros(n)
        b[u] += 1.0
        b[v] -= 1.0
        x, info = cg(B, b[:-1], rtol=tol, atol=0.0, maxiter=maxiter, M=M)
        if info != 0:
            raise RuntimeError(f"CG did not converge for position {p}")
        xf = np.append(x, 0.0)
        out[i] = xf[u] - xf[v]
    return out
