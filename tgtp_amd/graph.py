"""Host CSR graphs: loading, generation, eigenpair files.

`Graph` mirrors fuser::graph_t (src/graph.h:10-28): `idx` uint32[n+1],
`adj` uint32[m], sorted rows, both directions, duplicates kept.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from . import _lib as L


@dataclass
class Graph:
    idx: np.ndarray  # uint32 [n+1]
    adj: np.ndarray  # uint32 [m]

    @property
    def n(self) -> int:
        return int(self.idx.shape[0] - 1)

    @property
    def m(self) -> int:
        return int(self.adj.shape[0])

    def degree(self, v=None):
        d = np.diff(self.idx.astype(np.int64))
        return d if v is None else int(d[v])

    def begin(self, v: int) -> int:
        return int(self.idx[v])

    def end(self, v: int) -> int:
        return int(self.idx[v + 1])

    def struct(self) -> L.GraphStruct:
        self.idx = np.ascontiguousarray(self.idx, dtype=np.uint32)
        self.adj = np.ascontiguousarray(self.adj, dtype=np.uint32)
        return L.GraphStruct(self.n, self.m, L.ptr(self.idx, C.c_uint32),
                             L.ptr(self.adj, C.c_uint32))

    def csr_struct(self) -> L.CsrStruct:
        s = self.struct()
        return L.CsrStruct(s.n, s.m, s.idx, s.adj)

    def reverse_index(self) -> np.ndarray:
        rev = np.empty(self.m, dtype=np.uint32)
        cs = self.csr_struct()
        L.lib().tgtp_csr_reverse_index(C.byref(cs), L.ptr(rev, C.c_uint32))
        return rev

    def write_txt(self, path) -> None:
        cs = self.csr_struct()
        rc = L.lib().tgtp_write_graph_txt(C.byref(cs), str(path).encode())
        if rc != 0:
            raise OSError(f"cannot write {path}")

    def undirected_edges(self) -> np.ndarray:
        """(u, v) pairs with u < v, one per CSR entry (self-loops dropped)."""
        u = np.repeat(np.arange(self.n, dtype=np.uint32), np.diff(self.idx.astype(np.int64)))
        keep = u < self.adj
        return np.stack([u[keep], self.adj[keep]], axis=1)

    @staticmethod
    def from_edges(n: int, edges) -> "Graph":
        """CSR with both directions of each (u, v), rows sorted (graph.h order)."""
        e = np.asarray(edges, dtype=np.int64).reshape(-1, 2)
        src = np.concatenate([e[:, 0], e[:, 1]])
        dst = np.concatenate([e[:, 1], e[:, 0]])
        order = np.lexsort((dst, src))
        src, dst = src[order], dst[order]
        idx = np.zeros(n + 1, dtype=np.uint32)
        np.add.at(idx, src + 1, 1)
        idx = np.cumsum(idx, dtype=np.uint64).astype(np.uint32)
        return Graph(idx, dst.astype(np.uint32))


def _take_csr(cs: L.CsrStruct) -> Graph:
    n, m = int(cs.n), int(cs.m)
    idx = np.ctypeslib.as_array(cs.idx, shape=(n + 1,)).copy()
    adj = np.ctypeslib.as_array(cs.adj, shape=(max(m, 1),))[:m].copy()
    L.lib().tgtp_csr_free(C.byref(cs))
    return Graph(idx, adj)


def load_txt(path) -> Graph:
    """graph_t::load_txt (src/graph.h:29-97); raises ValueError / OSError
    where the reference throws runtime_error."""
    cs = L.CsrStruct()
    err = C.create_string_buffer(512)
    rc = L.lib().tgtp_load_txt(str(path).encode(), C.byref(cs), err, 512)
    if rc == L.TGTP_E_IO:
        raise OSError(err.value.decode())
    if rc != 0:
        raise ValueError(err.value.decode())
    return _take_csr(cs)


def save_bin(g: Graph, path) -> None:
    """Binary CSR cache (tgtp_save_csr_bin)."""
    cs = g.csr_struct()
    if L.lib().tgtp_save_csr_bin(C.byref(cs), str(path).encode()) != 0:
        raise OSError(f"cannot write {path}")


def load_bin(path) -> Graph:
    """Binary CSR cache (tgtp_load_csr_bin); ValueError for a damaged file."""
    cs = L.CsrStruct()
    err = C.create_string_buffer(512)
    rc = L.lib().tgtp_load_csr_bin(str(path).encode(), C.byref(cs), err, 512)
    if rc == L.TGTP_E_IO:
        raise OSError(err.value.decode())
    if rc != 0:
        raise ValueError(err.value.decode())
    return _take_csr(cs)


def load_txt_cached(path, cache=None) -> Graph:
    """load_txt through a binary cache (default: path + ".csr"), rebuilt when
    missing, older than the text or damaged."""
    path = Path(path)
    cache = Path(cache) if cache is not None else path.with_name(path.name + ".csr")
    if cache.exists() and cache.stat().st_mtime >= path.stat().st_mtime:
        try:
            return load_bin(cache)
        except ValueError:
            pass
    g = load_txt(path)
    save_bin(g, cache)
    return g


def chung_lu(n: int, beta: float, mean_deg: float, w_cap: float = 0.0, seed: int = 1) -> Graph:
    cs = L.CsrStruct()
    rc = L.lib().tgtp_gen_chung_lu(n, beta, mean_deg, w_cap, seed, C.byref(cs))
    if rc != 0:
        raise ValueError(f"tgtp_gen_chung_lu failed ({rc})")
    return _take_csr(cs)


def grid_rewired(rows: int, cols: int, p_rewire: float, seed: int = 1) -> Graph:
    cs = L.CsrStruct()
    rc = L.lib().tgtp_gen_grid_rewired(rows, cols, p_rewire, seed, C.byref(cs))
    if rc != 0:
        raise ValueError(f"tgtp_gen_grid_rewired failed ({rc})")
    return _take_csr(cs)


def load_eigens_txt(path, n: int, omega: int):
    """EigenVecs<float>::load_txt (src/aesc.h:51-64) -> (values f64[omega],
    vectors f32[n, omega] vertex-major)."""
    lam = np.empty(omega, dtype=np.float64)
    X = np.empty((n, omega), dtype=np.float32)
    err = C.create_string_buffer(512)
    rc = L.lib().tgtp_load_eigens_txt(str(path).encode(), n, omega, L.ptr(lam, C.c_double),
                                      L.ptr(X, C.c_float), err, 512)
    if rc == L.TGTP_E_IO:
        raise OSError(err.value.decode())
    if rc != 0:
        raise ValueError(err.value.decode())
    return lam, X


def save_eigens_bin(path, values: np.ndarray, vectors: np.ndarray) -> None:
    """Binary eigenpair cache (tgtp_save_eigens_bin): values f64[omega],
    vectors f32[n, omega] vertex-major, as load_eigens_txt returns them."""
    lam = np.ascontiguousarray(values, dtype=np.float64)
    X = np.ascontiguousarray(vectors, dtype=np.float32)
    if X.ndim != 2 or X.shape[1] != lam.shape[0]:
        raise ValueError("vectors must be [n, omega] for omega values")
    rc = L.lib().tgtp_save_eigens_bin(str(path).encode(), X.shape[0], lam.shape[0],
                                      L.ptr(lam, C.c_double), L.ptr(X, C.c_float))
    if rc == L.TGTP_E_IO:
        raise OSError(f"cannot write {path}")
    if rc != 0:
        raise ValueError("eigenpairs must be finite")


def load_eigens_bin(path, n: int, omega: int):
    """tgtp_load_eigens_bin: the cache must hold exactly n x omega pairs."""
    lam = np.empty(omega, dtype=np.float64)
    X = np.empty((n, omega), dtype=np.float32)
    err = C.create_string_buffer(512)
    rc = L.lib().tgtp_load_eigens_bin(str(path).encode(), n, omega, L.ptr(lam, C.c_double),
                                      L.ptr(X, C.c_float), err, 512)
    if rc == L.TGTP_E_IO:
        raise OSError(err.value.decode())
    if rc != 0:
        raise ValueError(err.value.decode())
    return lam, X


def load_eigens_txt_cached(path, n: int, omega: int, cache=None):
    """load_eigens_txt through a binary cache (default: path + ".bin"),
    rebuilt when missing, older than the text, damaged or of other sizes."""
    path = Path(path)
    cache = Path(cache) if cache is not None else path.with_name(path.name + ".bin")
    if cache.exists() and cache.stat().st_mtime >= path.stat().st_mtime:
        try:
            return load_eigens_bin(cache, n, omega)
        except ValueError:
            pass
    lam, X = load_eigens_txt(path, n, omega)
    save_eigens_bin(cache, lam, X)
    return lam, X


def check_aesc_graph(g: Graph, omega: int) -> None:
    """AESCGraph constructor checks (src/aesc.h:87-89)."""
    cs = g.csr_struct()
    err = C.create_string_buffer(256)
    if L.lib().tgtp_check_aesc_graph(C.byref(cs), omega, err, 256) != 0:
        raise ValueError(err.value.decode())


def cycle(n: int) -> Graph:
    return Graph.from_edges(n, [(i, (i + 1) % n) for i in range(n)])


def complete(n: int) -> Graph:
    return Graph.from_edges(n, [(i, j) for i in range(n) for j in range(i + 1, n)])


def petersen() -> Graph:
    outer = [(i, (i + 1) % 5) for i in range(5)]
    spokes = [(i, i + 5) for i in range(5)]
    inner = [(5 + i, 5 + (i + 2) % 5) for i in range(5)]
    return Graph.from_edges(10, outer + spokes + inner)
