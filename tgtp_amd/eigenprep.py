"""Top-omega eigenpairs of the normalized adjacency, in the layout the hot
path consumes (EigenVecs<float>, src/aesc.h:43-81).

Follows src/eigenprep.h:95-189: unique undirected edges (self-loops once),
P = D^-1/2 A D^-1/2 with networkx row sums, the omega eigenpairs of largest
magnitude in FP64, vectors back-transformed by sqrt(2m / deg_i), pairs
stably sorted by descending |lambda|, text rows "lambda x_0 .. x_{n-1}" at
max_digits10.  The solver is LAPACK (dense, small n) or ARPACK Lanczos
(scipy), not Spectra: eigenpairs agree with the reference's only to solver
tolerance, so every parity check feeds one eigen file to both sides.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

from .graph import Graph


def normalized_adjacency(g: Graph):
    import scipy.sparse as sp

    # unique undirected edges incl. self-loops once (eigenprep.h:166-169)
    u = np.repeat(np.arange(g.n, dtype=np.int64), np.diff(g.idx.astype(np.int64)))
    v = g.adj.astype(np.int64)
    keep = u <= v
    pairs = np.unique(np.stack([u[keep], v[keep]], axis=1), axis=0)
    a, b = pairs[:, 0], pairs[:, 1]
    rs = np.zeros(g.n)
    np.add.at(rs, a, 1.0)
    off = a != b
    np.add.at(rs, b[off], 1.0)
    val = 1.0 / np.sqrt(rs[a] * rs[b])
    rows = np.concatenate([a, b[off]])
    cols = np.concatenate([b, a[off]])
    vals = np.concatenate([val, val[off]])
    P = sp.csr_matrix((vals, (rows, cols)), shape=(g.n, g.n))
    return P, rs, len(pairs)


def compute_eigens(g: Graph, omega: int, dense_limit: int = 2500, tol: float = 1e-10,
                   seed: int = 1):
    """(values f64[omega], vectors f64[n, omega]) sorted by descending |lambda|."""
    P, rs, m_unique = normalized_adjacency(g)
    n = g.n
    if not (2 <= omega <= n):
        raise ValueError("omega must satisfy 2 <= omega <= n")
    if n <= dense_limit or omega >= n - 1:
        w, V = np.linalg.eigh(P.toarray())
    else:
        from scipy.sparse.linalg import eigsh

        v0 = np.random.default_rng(seed).standard_normal(n)
        ncv = min(n, max(2 * omega + 1, 20))
        w, V = eigsh(P, k=omega, which="LM", tol=tol, ncv=ncv, v0=v0)
    order = np.argsort(-np.abs(w), kind="stable")[:omega]
    w, V = w[order], V[:, order]
    scale = np.where(rs > 0, np.sqrt(2.0 * m_unique) / np.sqrt(np.maximum(rs, 1e-300)), 0.0)
    X = V * scale[:, None]
    return w.astype(np.float64), X


def eigens_f32(values: np.ndarray, vectors: np.ndarray):
    """The in-memory form tgtp consumes: lambda f64[omega], X f32[n, omega]."""
    return (np.ascontiguousarray(values, dtype=np.float64),
            np.ascontiguousarray(vectors, dtype=np.float32))


def write_eigens_txt(path, values: np.ndarray, vectors32: np.ndarray) -> None:
    """sorted_eigens_<omega>.txt rows (datasets/README.md:10-12).  lambda is
    written round-trip exact for double; vector entries with 9 significant
    digits, which parse (as float, aesc.h:60) back to exactly vectors32, so a
    run from the file and an in-memory run see the same inputs."""
    vectors32 = np.asarray(vectors32, dtype=np.float32)
    with open(path, "w") as f:
        for i in range(values.shape[0]):
            f.write(repr(float(values[i])))
            f.write(" " + " ".join("%.9g" % x for x in vectors32[:, i].tolist()))
            f.write("\n")


def cached_eigens(g: Graph, omega: int, cache: Path | None = None, **kw):
    """compute_eigens with an .npz cache keyed by the caller."""
    if cache is not None and Path(cache).exists():
        z = np.load(cache)
        return z["values"], z["vectors"]
    w, X = compute_eigens(g, omega, **kw)
    if cache is not None:
        np.savez(cache, values=w, vectors=X)
    return w, X
