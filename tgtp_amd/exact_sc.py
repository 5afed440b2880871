"""Exact spanning-edge centrality in FP64 on a torch device: the reference's
exact_sc tool (src/exact_sc.cpp), which the accuracy statements of the
approximate scores are measured against.

SC(u, v) is the effective resistance between the endpoints of a unit-weight
edge.  `exact_scores` follows exact_sc.cpp:21-45: grounded Laplacian B
(vertex 0 removed; duplicates count once per CSR entry, a self-loop adds 2
to the degree and -2 to its own entry, so it cancels), G = B^-1 by dense
Cholesky, R(0, v) = G[v, v] and R(u, v) = G[u, u] + G[v, v] - 2 G[u, v],
range-checked like exact_sc.cpp:39.  It keeps FP64 where exact_sc writes
float32 (exact_sc.cpp:40).

Dense Cholesky is O(n^3); `sampled_scores` gives R for chosen CSR entries
of larger graphs by block conjugate gradients on L x = e_u - e_v (Jacobi
preconditioner, FP64, consistent singular system: the right-hand side sums
to zero), R = (e_u - e_v)^T x.

    python -m tgtp_amd.exact_sc graph.txt outfile.bin [--device cuda:0]
writes exact_sc's scores.bin (float32 x m, CSR order).
"""
from __future__ import annotations

import argparse
import sys

import numpy as np

from .graph import Graph


def _rows(g: Graph) -> np.ndarray:
    return np.repeat(np.arange(g.n, dtype=np.int64), np.diff(g.idx.astype(np.int64)))


def _check_graph(g: Graph) -> None:
    if g.n < 2:
        raise ValueError("graph needs at least two vertices")
    if (np.diff(g.idx.astype(np.int64)) == 0).any():
        raise ValueError("isolated vertex")


def exact_scores(g: Graph, device: str = "cuda:0") -> np.ndarray:
    """FP64 effective resistance of every CSR entry (exact_sc.cpp:21-45)."""
    import torch

    _check_graph(g)
    n = g.n
    k = n - 1
    u = _rows(g)
    v = g.adj.astype(np.int64)
    deg = np.diff(g.idx.astype(np.int64)).astype(np.float64)
    B = torch.zeros((k, k), dtype=torch.float64, device=device)
    B[torch.arange(k, device=device), torch.arange(k, device=device)] = \
        torch.from_numpy(deg[1:]).to(device)
    mask = (u > 0) & (v > 0)
    uu = torch.from_numpy(u[mask] - 1).to(device)
    vv = torch.from_numpy(v[mask] - 1).to(device)
    B.index_put_((uu, vv), torch.full((uu.shape[0],), -1.0, dtype=torch.float64, device=device),
                 accumulate=True)
    L, info = torch.linalg.cholesky_ex(B)
    if int(info) != 0:
        raise RuntimeError("grounded Laplacian is singular; graph must be connected")
    del B
    Gm = torch.cholesky_inverse(L)
    del L
    if not bool(torch.isfinite(Gm).all()):
        raise RuntimeError("nonfinite pseudoinverse")
    diag = torch.cat([torch.zeros(1, dtype=torch.float64, device=device), Gm.diagonal()])
    ut = torch.from_numpy(u).to(device)
    vt = torch.from_numpy(v).to(device)
    both = (ut > 0) & (vt > 0)
    cross = torch.zeros(ut.shape[0], dtype=torch.float64, device=device)
    cross[both] = Gm[ut[both] - 1, vt[both] - 1]
    r = (diag[ut] + diag[vt] - 2.0 * cross).cpu().numpy()
    if not ((r >= 0.0).all() and (r <= 1.0 + 1e-9).all()):
        raise RuntimeError("effective resistance out of range")
    return r


def sampled_scores(g: Graph, pos: np.ndarray, device: str = "cuda:0", rtol: float = 1e-11,
                   max_iter: int = 100000, chunk: int = 512) -> np.ndarray:
    """FP64 effective resistance of the CSR entries `pos` by block PCG."""
    import torch

    _check_graph(g)
    n = g.n
    pos = np.asarray(pos, dtype=np.int64)
    u_all = _rows(g)
    deg = np.diff(g.idx.astype(np.int64)).astype(np.float64)
    rows = np.concatenate([np.arange(n, dtype=np.int64), u_all])
    cols = np.concatenate([np.arange(n, dtype=np.int64), g.adj.astype(np.int64)])
    vals = np.concatenate([deg, -np.ones(g.m)])
    import scipy.sparse as sp

    Lh = sp.csr_matrix((vals, (rows, cols)), shape=(n, n))
    Lh.sum_duplicates()
    L = torch.sparse_csr_tensor(torch.from_numpy(Lh.indptr.astype(np.int64)),
                                torch.from_numpy(Lh.indices.astype(np.int64)),
                                torch.from_numpy(Lh.data), size=(n, n), dtype=torch.float64,
                                device=device)
    dl = np.asarray(Lh.diagonal())
    minv = torch.from_numpy(np.where(dl > 0, 1.0 / np.where(dl > 0, dl, 1.0), 0.0)).to(device)
    out = np.empty(pos.shape[0], dtype=np.float64)
    for c0 in range(0, pos.shape[0], chunk):
        p = pos[c0:c0 + chunk]
        s = p.shape[0]
        us = u_all[p]
        vs = g.adj[p].astype(np.int64)
        Bm = torch.zeros((n, s), dtype=torch.float64, device=device)
        cols_t = torch.arange(s, device=device)
        Bm[torch.from_numpy(us).to(device), cols_t] += 1.0
        Bm[torch.from_numpy(vs).to(device), cols_t] -= 1.0
        X = torch.zeros_like(Bm)
        R = Bm.clone()
        Z = R * minv[:, None]
        Pd = Z.clone()
        rz = (R * Z).sum(0)
        bn = torch.linalg.vector_norm(Bm, dim=0).clamp_min(1e-300)
        for _ in range(max_iter):
            AP = L @ Pd
            pap = (Pd * AP).sum(0)
            alpha = torch.where(pap > 0, rz / pap.clamp_min(1e-300), torch.zeros_like(pap))
            X += Pd * alpha
            R -= AP * alpha
            if float((torch.linalg.vector_norm(R, dim=0) / bn).max()) < rtol:
                break
            Z = R * minv[:, None]
            rz_new = (R * Z).sum(0)
            beta = torch.where(rz > 0, rz_new / rz.clamp_min(1e-300), torch.zeros_like(rz))
            rz = rz_new
            Pd = Z + Pd * beta
        else:
            raise RuntimeError("block PCG did not converge")
        out[c0:c0 + s] = (Bm * X).sum(0).cpu().numpy()
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="exact_sc", description="exact spanning-edge centrality")
    ap.add_argument("graph")
    ap.add_argument("outfile")
    ap.add_argument("--device", default="cuda:0")
    a = ap.parse_args(argv)
    from .graph import load_txt

    try:
        g = load_txt(a.graph)
        r = exact_scores(g, a.device)
        r.astype(np.float32).tofile(a.outfile)
    except (ValueError, RuntimeError, OSError) as e:
        print(f"exact_sc: {e}", file=sys.stderr)
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
