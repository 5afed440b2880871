"""Vectorised numpy restatement of the tau refinement (calculate_taus,
src/aesc.h:129-216) with the reference's rule or the bounded rule of the
FAST modes, and the truncation bound a tau must satisfy.

Test infrastructure: the reference rule is pinned against oracle.taus (which
equals the reference build's calculate_taus); the bounded rule differs from
it only where a refinement step's top-omega tail delta already reaches
eps/2 (aesc.h:130 then bounds against eps/2 alone; the bounded rule skips the
step).

For an edge (u, v) with y = e_u/sqrt(d_u) - e_v/sqrt(d_v) and the
eigenpairs (lambda_i, v_i) of D^-1/2 A D^-1/2, the effective resistance is
sum_i (y.v_i)^2 / (1 - lambda_i) (i >= 1) and the series truncated after tau
steps misses sum_i (y.v_i)^2 lambda_i^(tau+1) / (1 - lambda_i).  With X the
back-transformed vectors (x_ui = v_i(u) sqrt(2 m_u / d_u), m_u undirected
edges) and m = 2 m_u CSR entries, (y.v_i)^2 = D_i / m, D_i = (x_ui - x_vi)^2.
For odd tau the missed part is at most
    top(tau)  = sum_{i in [1, omega-2]} D_i lambda_i^(tau+1) / (1 - lambda_i) / m
  + rest(tau) = (1/d_u + 1/d_v - 2/(d_u d_v) - upsilon) lam_w^(tau+1) / (1 - lam_w^2)
with upsilon = sum_{[1, omega-2]} D_i (1 + lambda_i) / m, lam_w = lambda_{omega-1};
or, bounding every component by lambda_1,
    (1/d_u + 1/d_v - 2/(d_u d_v)) lambda_1^(tau+1) / (1 - lambda_1^2).
"""
from __future__ import annotations

import numpy as np


def _edges(g):
    rows = np.repeat(np.arange(g.n, dtype=np.int64), np.diff(g.idx.astype(np.int64)))
    und = np.where(g.adj.astype(np.int64) >= rows)[0]
    return rows, und


def _get_tau(du, dv, eps_half32, lamba, delta32, ups32):
    """get_tau (aesc.h:129-135) vectorised: float eps - float delta, the
    rest in double."""
    ed = (eps_half32 - delta32).astype(np.float32)
    ed = np.where(ed > 0.0, ed, eps_half32).astype(np.float64)
    x = (1.0 / du + 1.0 / dv - 2.0 / du / dv - ups32.astype(np.float64)) / ed / (1.0 - lamba * lamba)
    a = np.log(np.maximum(x, 1.0))
    b = np.log(1.0 / abs(lamba))
    t = np.maximum(np.ceil(a / b - 1), 1.0).astype(np.int64)
    return np.where(t % 2 == 0, t + 1, t)


def taus(g, lam, X, eps_half: float, bounded: bool) -> np.ndarray:
    """int32[m] taus for both directions of every edge, with the reference's
    operand types and summation order (aesc.h:137-162: float accumulators,
    float squared differences, double eigenvalue terms)."""
    lam = np.asarray(lam, dtype=np.float64)
    X = np.asarray(X, dtype=np.float32)
    om = lam.shape[0]
    rows, und = _edges(g)
    U, V = rows[und], g.adj[und].astype(np.int64)
    deg = np.diff(g.idx.astype(np.int64)).astype(np.float64)
    du, dv = deg[V], deg[U]  # get_tau's "du" is the degree of v (aesc.h:129, 186-189)
    e32 = np.float32(eps_half)
    m32 = np.float32(g.m)
    sq = [(X[U, i] - X[V, i]) * (X[U, i] - X[V, i]) for i in range(om)]  # float32
    ups = np.zeros(U.shape[0], dtype=np.float32)
    for i in range(1, om - 1):
        ups = (ups.astype(np.float64) + sq[i].astype(np.float64) * (1 + lam[i])).astype(np.float32)
    ups = (ups / m32).astype(np.float32)
    tau = _get_tau(du, dv, e32, lam[1], np.zeros_like(ups), np.zeros_like(ups))
    t = np.ones_like(tau)
    live = np.ones(tau.shape[0], dtype=bool)
    et = lam * lam  # eigen_t_1, double (aesc.h:179-181, 211-213)
    while live.any():
        delta = np.zeros(U.shape[0], dtype=np.float32)
        for i in range(1, om - 1):
            term = (sq[i] * np.float32(et[i])).astype(np.float64) / (1 - lam[i])
            delta = (delta.astype(np.float64) + term).astype(np.float32)
        delta = (delta / m32).astype(np.float32)
        tp = _get_tau(du, dv, e32, lam[om - 1], delta, ups)
        skip = live & bounded & ~((e32 - delta).astype(np.float32) > 0.0)
        acc = live & ~skip & (t <= tp) & (tp < tau)
        tau = np.where(acc, tp, tau)
        t = np.where(acc | skip, t + 2, t)
        live = acc | (skip & (t < tau))
        et = et * lam * lam
    out = np.zeros(g.m, dtype=np.int32)
    out[und] = tau
    # the reverse position: first occurrence of u in v's row (aesc.h:207-209)
    for k in range(und.shape[0]):
        v, u = V[k], U[k]
        lo = g.idx[v] + np.searchsorted(g.adj[g.idx[v]:g.idx[v + 1]], u)
        out[lo] = tau[k]
    return out


def truncation_bound(g, lam, X, tau_arr) -> np.ndarray:
    """FP64 bound on the part of each undirected edge's series missed after
    tau steps (the smaller of the two forms above), per undirected edge in
    the order of positions with adj >= row."""
    lam = np.asarray(lam, dtype=np.float64)
    om = lam.shape[0]
    rows, und = _edges(g)
    U, V = rows[und], g.adj[und].astype(np.int64)
    deg = np.diff(g.idx.astype(np.int64)).astype(np.float64)
    tau = np.asarray(tau_arr)[und].astype(np.int64)
    li = lam[1:om - 1]
    D = (np.asarray(X, dtype=np.float64)[U, 1:om - 1] - np.asarray(X, dtype=np.float64)[V, 1:om - 1]) ** 2
    mass = 1.0 / deg[U] + 1.0 / deg[V] - 2.0 / deg[U] / deg[V]
    ups = D @ (1.0 + li) / g.m
    top = np.array([D[k] @ (li ** (tau[k] + 1) / (1.0 - li)) for k in range(D.shape[0])]) / g.m
    lw, l1 = lam[om - 1], lam[1]
    rest = np.maximum(mass - ups, 0.0) * np.abs(lw) ** (tau + 1) / (1.0 - lw * lw)
    first = mass * np.abs(l1) ** (tau + 1) / (1.0 - l1 * l1)
    return np.minimum(top + rest, first)
