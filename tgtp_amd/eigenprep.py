"""Top-omega eigenpairs of the normalized adjacency, in the layout the hot
path consumes (EigenVecs<float>, src/aesc.h:43-81).

Follows src/eigenprep.h:95-189: unique undirected edges (self-loops once),
P = D^-1/2 A D^-1/2 with networkx row sums, the omega eigenpairs of largest
magnitude in FP64, vectors back-transformed by sqrt(2m / deg_i), pairs
stably sorted by descending |lambda|, text rows "lambda x_0 .. x_{n-1}" at
max_digits10.  The solver is LAPACK (dense, small n), ARPACK Lanczos
(scipy), or on a torch device FP64 block Lanczos with thick restart
(compute_eigens_device, compute_eigens_many), not Spectra: eigenpairs agree
with the reference's only to solver tolerance, so every parity check feeds
one eigen file to both sides.
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
    if not (2 <= omega < n):  # eigenprep.h:99-103
        raise ValueError(f"omega must satisfy 2 <= omega < n; got {omega} for n={n}")
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


def _torch_csr(P, device):
    import torch

    P = P.tocsr()
    return torch.sparse_csr_tensor(torch.from_numpy(P.indptr.astype(np.int64)),
                                   torch.from_numpy(P.indices.astype(np.int64)),
                                   torch.from_numpy(P.data.astype(np.float64)),
                                   size=P.shape, dtype=torch.float64, device=device)


def _tmm(A, B, chunk: int = 4096):
    """A^T B for tall A (.., n, p), B (.., n, q) with n >> p, q: the long
    dimension is cut into chunks whose partial products run as one batched
    GEMM and are then summed (split-K).  One GEMM over all n rows keeps few
    workgroups busy on such shapes: at n = 262,144, p = 1536, q = 64 it took
    28 ms against 1 ms split."""
    n = A.shape[-2]
    S = n // chunk
    if S < 2:
        return A.mT @ B
    h = S * chunk
    lead = A.shape[:-2]
    As = A[..., :h, :].reshape(*lead, S, chunk, A.shape[-1])
    Bs = B[..., :h, :].reshape(*lead, S, chunk, B.shape[-1])
    out = (As.mT @ Bs).sum(dim=-3)
    if h < n:
        out = out + A[..., h:, :].mT @ B[..., h:, :]
    return out


def _orth_against(Z, Q):
    # classical Gram-Schmidt, twice (CGS2): one pass loses orthogonality when
    # Z is nearly inside span(Q)
    for _ in range(2):
        Z = Z - Q @ _tmm(Q, Z)
    return Z


def block_krylov_eigs(P, omega: int, tol: float = 1e-9, block: int = 64, basis: int = 768,
                      keep: int | None = None, seed: int = 1, max_cycles: int = 200):
    """The omega eigenpairs of largest |lambda| of the symmetric sparse P
    (torch CSR, float64, on any device): block Lanczos with full
    reorthogonalisation and thick restart (Stathopoulos, Saad & Wu 1998;
    Zhou & Saad 2008, block form).

    The basis grows block by block (W = P Q_j, CGS2 against the basis, QR)
    to `basis` columns; Rayleigh-Ritz on H = Q^T P Q then keeps the `keep`
    Ritz vectors of largest |theta| together with P times them, and
    expansion continues from the next Krylov block, which the kept Ritz
    vectors' residuals span.  Stops when every wanted Ritz pair has
    ||P y - theta y|| < tol.  Returns (theta[omega], Y[n, omega]) as
    float64 tensors on P's device, ordered by descending |theta|.
    """
    import torch

    n = P.shape[0]
    dev = P.device
    if not (2 <= omega < n):  # eigenprep.h:99-103
        raise ValueError(f"omega must satisfy 2 <= omega < n; got {omega} for n={n}")
    keep = keep or omega + max(omega, 64)
    basis = min(basis, n)
    block = min(block, max(1, basis - omega))
    keep = max(omega, min(keep, basis - block))
    gen = np.random.default_rng(seed)
    Q = torch.empty((n, basis), dtype=torch.float64, device=dev)
    W = torch.empty_like(Q)
    nxt, _ = torch.linalg.qr(torch.from_numpy(gen.standard_normal((n, block))).to(dev))
    c = 0
    for _cycle in range(max_cycles):
        while c + nxt.shape[1] <= basis:
            k = nxt.shape[1]
            Q[:, c:c + k] = nxt
            W[:, c:c + k] = P @ nxt
            c += k
            nxt, R = torch.linalg.qr(_orth_against(W[:, c - k:c], Q[:, :c]))
            d = R.diagonal().abs()
            weak = d < 1e-10 * max(float(d.max()), 1e-300)
            if bool(weak.any()):
                # an invariant subspace was reached: continue from random
                # directions outside the basis
                nxt = nxt.clone()
                nxt[:, weak] = torch.from_numpy(
                    gen.standard_normal((n, int(weak.sum())))).to(dev)
                nxt, _ = torch.linalg.qr(_orth_against(nxt, Q[:, :c]))
            if c + k > basis:
                break
        H = _tmm(Q[:, :c], W[:, :c])
        H = 0.5 * (H + H.T)
        th, S = torch.linalg.eigh(H)
        order = torch.argsort(-th.abs(), stable=True)
        th, S = th[order], S[:, order[:keep]]
        Y = Q[:, :c] @ S
        PY = W[:, :c] @ S
        res = torch.linalg.vector_norm(PY[:, :omega] - Y[:, :omega] * th[:omega], dim=0)
        if float(res.max()) < tol or c == n:
            return th[:omega].clone(), Y[:, :omega].clone()
        Q[:, :keep] = Y
        W[:, :keep] = PY
        c = keep
        nxt, _ = torch.linalg.qr(_orth_against(nxt, Q[:, :c]))
    raise RuntimeError(f"block Krylov eigensolver: no convergence in {max_cycles} cycles")


def _cholqr2(Z):
    """Orthonormal basis of the columns of each Z[i] (k, n, b) by two
    Cholesky-QR passes (GEMM-shaped, unlike Householder QR).  Returns (Q,
    ok[k]); ok is False where a Gram matrix was not positive definite."""
    import torch

    ok = torch.ones(Z.shape[0], dtype=torch.bool, device=Z.device)
    for _ in range(2):
        G = _tmm(Z, Z)
        L, info = torch.linalg.cholesky_ex(0.5 * (G + G.mT))
        bad = info != 0
        ok &= ~bad
        if bool(bad.any()):
            eye = torch.eye(G.shape[-1], dtype=G.dtype, device=G.device)
            L = torch.where(bad[:, None, None], eye, L)
        # Z L^-T through the small explicit inverse: one batched GEMM (a
        # batched triangular solve on the tall side needs a large workspace)
        eye = torch.eye(G.shape[-1], dtype=G.dtype, device=G.device).expand_as(L)
        Linv_t = torch.linalg.solve_triangular(L, eye, upper=False).mT.contiguous()
        Z = Z @ Linv_t
    return Z, ok


def batched_block_krylov_eigs(P, k: int, n_pad: int, n_real, omega: int, tol: float = 1e-9,
                              block: int = 64, basis: int = 768, keep: int | None = None,
                              seed: int = 1, max_cycles: int = 300, on_cycle=None,
                              accept_tol: float = 1e-6):
    """block_krylov_eigs for k graphs at once.  P is the block-diagonal
    union (torch CSR, float64) of the k normalized adjacencies, each padded
    to n_pad vertices with empty rows (eigenvalue 0, never among the wanted
    ones); n_real[i] is graph i's vertex count.  Every product on the basis
    is one batched GEMM over the k graphs and one SpMM with the union, so a
    size class of graphs costs about as many launches as one graph.
    on_cycle(cycle, worst residual) is called after each Rayleigh-Ritz step.
    A graph not below `tol` after max_cycles is accepted if its residual is
    below accept_tol (the vectors are stored in FP32, whose rounding is
    ~6e-8 relative; clustered spectra such as large grids converge slowly),
    otherwise this raises.  Returns (theta[k, omega], Y[k, n_pad, omega])."""
    import torch

    dev = P.device
    keep = keep or omega + max(omega // 2, 64)
    basis = min(basis, min(n_real))
    block = min(block, max(1, basis - keep))
    gen = np.random.default_rng(seed)
    mask = torch.zeros((k, n_pad, 1), dtype=torch.float64, device=dev)
    for i, ni in enumerate(n_real):
        mask[i, :ni] = 1.0

    def rand(cols):
        return torch.from_numpy(gen.standard_normal((k, n_pad, cols))).to(dev) * mask

    Q = torch.empty((k, n_pad, basis), dtype=torch.float64, device=dev)
    W = torch.empty_like(Q)
    out_th = torch.zeros((k, omega), dtype=torch.float64, device=dev)
    out_Y = torch.zeros((k, n_pad, omega), dtype=torch.float64, device=dev)
    done = torch.zeros(k, dtype=torch.bool, device=dev)

    def orth(Z, c):
        for _ in range(2):
            if c:
                Z = Z - Q[:, :, :c] @ _tmm(Q[:, :, :c], Z)
        Z, ok = _cholqr2(Z)
        if not bool(ok.all()):
            # an invariant subspace was reached: random directions instead
            Z = torch.where(ok[:, None, None], Z, rand(Z.shape[-1]))
            for _ in range(2):
                if c:
                    Z = Z - Q[:, :, :c] @ _tmm(Q[:, :, :c], Z)
            Z, _ = _cholqr2(Z)
        return Z

    nxt = orth(rand(block), 0)
    c = 0
    for _cycle in range(max_cycles):
        while c + block <= basis:
            Q[:, :, c:c + block] = nxt
            W[:, :, c:c + block] = (P @ nxt.reshape(k * n_pad, block)).reshape(k, n_pad, block)
            c += block
            nxt = orth(W[:, :, c - block:c].clone(), c)
        H = _tmm(Q[:, :, :c], W[:, :, :c])
        H = 0.5 * (H + H.mT)
        th, S = torch.linalg.eigh(H)
        order = torch.argsort(-th.abs(), dim=1, stable=True)
        th = torch.gather(th, 1, order)
        S = torch.gather(S, 2, order[:, None, :keep].expand(-1, c, -1))
        Y = Q[:, :, :c] @ S
        PY = W[:, :, :c] @ S
        res = torch.linalg.vector_norm(PY[:, :, :omega] - Y[:, :, :omega] * th[:, None, :omega],
                                       dim=1).amax(dim=1)
        if on_cycle is not None:
            on_cycle(_cycle, float(res.max()))
        newly = (res < tol) & ~done
        if bool(newly.any()):
            out_th[newly] = th[newly, :omega]
            out_Y[newly] = Y[newly, :, :omega]
            done |= newly
        if bool(done.all()):
            return out_th, out_Y
        if _cycle == max_cycles - 1:
            if bool((res[~done] < accept_tol).all()):
                out_th[~done] = th[~done, :omega]
                out_Y[~done] = Y[~done, :, :omega]
                return out_th, out_Y
            break
        Q[:, :, :keep] = Y
        W[:, :, :keep] = PY
        c = keep
        nxt = orth(nxt, c)
    raise RuntimeError(f"batched block Krylov eigensolver: no convergence in {max_cycles} cycles")


def compute_eigens_many(graphs, omega: int, device: str = "cuda:0", tol: float = 1e-9,
                        seed: int = 1, max_pad: float = 1.5, mem_bytes: float = 8e9,
                        progress=None, on_cycle=None, basis: int | None = None):
    """compute_eigens for many graphs on a torch device: graphs are grouped
    into size classes (n within max_pad of the class's smallest, basis
    memory within mem_bytes) and each class runs one batched block Krylov
    solve on the union of its graphs.  Returns [(values, vectors)] in the
    order of `graphs`."""
    import scipy.sparse as sp
    import torch

    order = sorted(range(len(graphs)), key=lambda i: graphs[i].n)
    out = [None] * len(graphs)
    pos = 0
    while pos < len(order):
        n0 = graphs[order[pos]].n
        cls = [order[pos]]
        while pos + len(cls) < len(order):
            nxt_i = order[pos + len(cls)]
            n_pad = graphs[nxt_i].n
            if n_pad > max_pad * n0 or (len(cls) + 1) * n_pad * 768 * 16 > mem_bytes:
                break
            cls.append(nxt_i)
        pos += len(cls)
        n_pad = max(graphs[i].n for i in cls)
        mats, extras = [], []
        for i in cls:
            Pi, rs, mu = normalized_adjacency(graphs[i])
            pad = n_pad - graphs[i].n
            mats.append(sp.block_diag([Pi, sp.csr_matrix((pad, pad))]) if pad else Pi)
            extras.append((rs, mu))
        U = sp.block_diag(mats, format="csr")
        # a larger basis for large graphs: fewer restarts on clustered spectra
        nb = basis or (768 if n_pad <= 131072 else 1536)
        th, Y = batched_block_krylov_eigs(_torch_csr(U, device), len(cls), n_pad,
                                          [graphs[i].n for i in cls], omega, tol=tol, seed=seed,
                                          on_cycle=on_cycle, basis=nb)
        th = th.cpu().numpy()
        Y = Y.cpu().numpy()
        for j, i in enumerate(cls):
            rs, mu = extras[j]
            ni = graphs[i].n
            scale = np.where(rs > 0, np.sqrt(2.0 * mu) / np.sqrt(np.maximum(rs, 1e-300)), 0.0)
            out[i] = (th[j].astype(np.float64), Y[j, :ni] * scale[:, None])
        del U, Y
        torch.cuda.empty_cache() if str(device).startswith("cuda") else None
        if progress:
            progress(pos, len(graphs))
    return out


def compute_eigens_device(g: Graph, omega: int, device: str = "cuda:0", tol: float = 1e-9,
                          seed: int = 1, dense_limit: int = 512):
    """compute_eigens with the eigensolver on a torch device (an MI355X for
    the batched workloads, where 256 host ARPACK runs would take minutes):
    same P, same ordering and back-transform.  The device is required
    explicitly; nothing falls back to another one."""
    import torch

    P, rs, m_unique = normalized_adjacency(g)
    n = g.n
    if not (2 <= omega < n):  # eigenprep.h:99-103
        raise ValueError(f"omega must satisfy 2 <= omega < n; got {omega} for n={n}")
    if n <= dense_limit or omega >= n // 2:
        w, V = torch.linalg.eigh(torch.from_numpy(P.toarray()).to(device))
        order = torch.argsort(-w.abs(), stable=True)[:omega]
        w, V = w[order], V[:, order]
    else:
        return compute_eigens_many([g], omega, device=device, tol=tol, seed=seed)[0]
    w = w.cpu().numpy()
    V = V.cpu().numpy()
    scale = np.where(rs > 0, np.sqrt(2.0 * m_unique) / np.sqrt(np.maximum(rs, 1e-300)), 0.0)
    return w.astype(np.float64), V * scale[:, None]


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
