"""Python view of the oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this package, and only as the checker.  The product (tgtp_amd/) never
imports it.  Function-level provenance and pinning: oracle/oracle.h.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
REF_TGTP_MT = HERE / "_ref" / "ref_tgtp_mt"

WALK_SABA = 0
WALK_MT19937 = 1

_lib = None


class Trace(C.Structure):
    _fields_ = [("taus", C.POINTER(C.c_int32)), ("nr", C.POINTER(C.c_uint64)),
                ("walk_sum", C.POINTER(C.c_float)), ("preds_row", C.POINTER(C.c_float)),
                ("ell", C.POINTER(C.c_int32)), ("edge_max", C.POINTER(C.c_float)),
                ("gmax", C.POINTER(C.c_float)), ("gmin", C.POINTER(C.c_float)),
                ("q_size", C.POINTER(C.c_uint32))]


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} not built; run `make -C oracle`")
        L = C.CDLL(str(LIB_PATH))
        P = C.POINTER
        u32p = P(C.c_uint32)
        L.orc_fmix32.argtypes = [C.c_uint32]
        L.orc_fmix32.restype = C.c_uint32
        L.orc_splitmix32.argtypes = [u32p]
        L.orc_splitmix32.restype = C.c_uint32
        L.orc_saba_walks.argtypes = [C.c_uint64, u32p, u32p, C.c_uint32, C.c_uint32, C.c_uint32,
                                     C.c_uint64, u32p, C.c_int, C.c_int]
        L.orc_saba_walks_avx2.argtypes = [C.c_uint64, u32p, u32p, C.c_uint32, C.c_uint32,
                                          C.c_uint32, C.c_uint64, u32p, C.c_int, C.c_int]
        L.orc_mt_walks.argtypes = [C.c_uint64, u32p, u32p, C.c_uint32, C.c_uint32, C.c_uint32,
                                   C.c_uint64, C.c_int, u32p, C.c_int]
        L.orc_mt19937_nth.argtypes = [C.c_uint32, C.c_uint64]
        L.orc_mt19937_nth.restype = C.c_uint32
        L.orc_saba_walk_source.argtypes = [u32p, u32p, C.c_uint32, C.c_uint32, C.c_uint64, u32p]
        L.orc_saba_walk_source.restype = None
        L.orc_walk2_saba.argtypes = [u32p, u32p, C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint32,
                                     P(C.c_float)]
        L.orc_walk2_saba.restype = C.c_float
        L.orc_taus.argtypes = [C.c_uint64, C.c_uint64, u32p, u32p, C.c_int, P(C.c_double),
                               P(C.c_float), C.c_float, P(C.c_int32), C.c_int]
        L.orc_tgtp.argtypes = [C.c_uint64, C.c_uint64, u32p, u32p, C.c_int, P(C.c_double),
                               P(C.c_float), C.c_double, C.c_double, C.c_int, C.c_int, C.c_int,
                               P(C.c_float), P(Trace)]
        L.orc_tgtp_rows.argtypes = [C.c_uint64, C.c_uint64, u32p, u32p, P(C.c_int32), C.c_double,
                                    C.c_double, C.c_int, C.c_uint32, C.c_uint32, C.c_int,
                                    P(C.c_float)]
        L.orc_tgtp_sources.argtypes = [C.c_uint64, C.c_uint64, u32p, u32p, P(C.c_int32),
                                       C.c_double, C.c_double, C.c_int, u32p, C.c_uint64,
                                       C.c_int, P(C.c_float)]
        L.orc_load_txt.argtypes = [C.c_char_p, P(C.c_uint64), P(C.c_uint64), P(u32p), P(u32p)]
        L.orc_load_eigens.argtypes = [C.c_char_p, C.c_uint64, C.c_int, P(C.c_double),
                                      P(C.c_float)]
        L.orc_free.argtypes = [C.c_void_p]
        L.orc_free.restype = None
        L.orc_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def fmix32(h: int) -> int:
    return int(lib().orc_fmix32(h))


def splitmix32(state: int):
    s = C.c_uint32(state)
    r = lib().orc_splitmix32(C.byref(s))
    return int(r), int(s.value)


def saba_walks(idx, adj, depth: int, n_walks: int, src_begin: int = 0, src_end=None,
               threads: int = 0, chunk: int = 1024) -> np.ndarray:
    idx, adj = _u32(idx), _u32(adj)
    n = idx.shape[0] - 1
    src_end = n if src_end is None else src_end
    cnt = np.zeros(n, dtype=np.uint32)
    rc = lib().orc_saba_walks(n, _p(idx, C.c_uint32), _p(adj, C.c_uint32), src_begin, src_end,
                              depth, n_walks, _p(cnt, C.c_uint32), threads, chunk)
    if rc != 0:
        raise ValueError("n_walks must be a multiple of 8 within the 32-bit grid domain")
    return cnt


def saba_walks_avx2(idx, adj, depth: int, n_walks: int, src_begin: int = 0, src_end=None,
                    threads: int = 0, chunk: int = 1024) -> np.ndarray:
    """saba_walks with AVX2 bouquets (the host CPU baseline); raises
    RuntimeError where the CPU has no AVX2."""
    idx, adj = _u32(idx), _u32(adj)
    n = idx.shape[0] - 1
    src_end = n if src_end is None else src_end
    cnt = np.zeros(n, dtype=np.uint32)
    rc = lib().orc_saba_walks_avx2(n, _p(idx, C.c_uint32), _p(adj, C.c_uint32), src_begin,
                                   src_end, depth, n_walks, _p(cnt, C.c_uint32), threads, chunk)
    if rc == -2:
        raise RuntimeError("this CPU has no AVX2")
    if rc != 0:
        raise ValueError("n_walks must be a positive multiple of 8; graph below 2^31 entries")
    return cnt


def mt_walks(idx, adj, depth: int, n_walks: int, variant: int = 0, src_begin: int = 0,
             src_end=None, threads: int = 0) -> np.ndarray:
    """MT19937 comparator walks (rw_experiment.cpp / rw_experiment_simd.cpp)."""
    idx, adj = _u32(idx), _u32(adj)
    n = idx.shape[0] - 1
    src_end = n if src_end is None else src_end
    cnt = np.zeros(n, dtype=np.uint32)
    rc = lib().orc_mt_walks(n, _p(idx, C.c_uint32), _p(adj, C.c_uint32), src_begin, src_end,
                            depth, n_walks, variant, _p(cnt, C.c_uint32), threads)
    if rc != 0:
        raise ValueError("unknown variant, or n_walks not a multiple of 8 for simd")
    return cnt


def mt19937_nth(seed: int, k: int) -> int:
    return int(lib().orc_mt19937_nth(seed, k))


def saba_walk_source(idx, adj, src: int, depth: int, n_walks: int) -> np.ndarray:
    idx, adj = _u32(idx), _u32(adj)
    cnt = np.zeros(idx.shape[0] - 1, dtype=np.uint32)
    lib().orc_saba_walk_source(_p(idx, C.c_uint32), _p(adj, C.c_uint32), src, depth, n_walks,
                               _p(cnt, C.c_uint32))
    return cnt


def walk2_saba(idx, adj, s: int, v: int, nr: int, length: int, pvec) -> float:
    idx, adj = _u32(idx), _u32(adj)
    pvec = np.ascontiguousarray(pvec, dtype=np.float32)
    return float(lib().orc_walk2_saba(_p(idx, C.c_uint32), _p(adj, C.c_uint32), s, v, nr, length,
                                      _p(pvec, C.c_float)))


def taus(idx, adj, lam, X, eps_half: float, threads: int = 0) -> np.ndarray:
    idx, adj = _u32(idx), _u32(adj)
    lam = np.ascontiguousarray(lam, dtype=np.float64)
    X = np.ascontiguousarray(X, dtype=np.float32)
    n, m = idx.shape[0] - 1, adj.shape[0]
    out = np.zeros(m, dtype=np.int32)
    rc = lib().orc_taus(n, m, _p(idx, C.c_uint32), _p(adj, C.c_uint32), lam.shape[0],
                        _p(lam, C.c_double), _p(X, C.c_float), float(np.float32(eps_half)),
                        _p(out, C.c_int32), threads)
    if rc != 0:
        raise ValueError("orc_taus failed")
    return out


def tgtp(idx, adj, lam, X, epsilon: float, gamma: int = 10, delta=None,
         walk_mode: int = WALK_SABA, threads: int = 0, trace: bool = False):
    """Scores float32[m]; with trace=True also a dict of per-stage arrays."""
    idx, adj = _u32(idx), _u32(adj)
    lam = np.ascontiguousarray(lam, dtype=np.float64)
    X = np.ascontiguousarray(X, dtype=np.float32)
    n, m = idx.shape[0] - 1, adj.shape[0]
    delta = 1.0 / n if delta is None else delta
    scores = np.zeros(m, dtype=np.float32)
    tr = None
    arrays = {}
    if trace:
        arrays = dict(taus=np.zeros(m, np.int32), nr=np.zeros(m, np.uint64),
                      walk_sum=np.zeros(m, np.float32), preds_row=np.zeros(m, np.float32),
                      ell=np.zeros(n, np.int32), edge_max=np.zeros(n, np.float32),
                      gmax=np.zeros(n, np.float32), gmin=np.zeros(n, np.float32),
                      q_size=np.zeros(n, np.uint32))
        ct = {np.int32: C.c_int32, np.uint64: C.c_uint64, np.float32: C.c_float,
              np.uint32: C.c_uint32}
        tr = Trace(**{k: _p(a, ct[a.dtype.type]) for k, a in arrays.items()})
    rc = lib().orc_tgtp(n, m, _p(idx, C.c_uint32), _p(adj, C.c_uint32), lam.shape[0],
                        _p(lam, C.c_double), _p(X, C.c_float), epsilon, delta, gamma, walk_mode,
                        threads, _p(scores, C.c_float), C.byref(tr) if tr else None)
    if rc != 0:
        raise ValueError(lib().orc_last_error().decode())
    return (scores, arrays) if trace else scores


def tgtp_rows(idx, adj, taus_arr, epsilon: float, gamma: int, src_begin: int, src_end: int,
              delta=None, threads: int = 0) -> np.ndarray:
    idx, adj = _u32(idx), _u32(adj)
    t = np.ascontiguousarray(taus_arr, dtype=np.int32)
    n, m = idx.shape[0] - 1, adj.shape[0]
    delta = 1.0 / n if delta is None else delta
    out = np.zeros(m, dtype=np.float32)
    lib().orc_tgtp_rows(n, m, _p(idx, C.c_uint32), _p(adj, C.c_uint32), _p(t, C.c_int32),
                        epsilon, delta, gamma, src_begin, src_end, threads, _p(out, C.c_float))
    return out


def tgtp_sources(idx, adj, taus_arr, epsilon: float, gamma: int, sources, delta=None,
                 threads: int = 0) -> np.ndarray:
    """Rows (preds before symmetrization) of the listed sources only; other
    entries are 0.  score(pos) = rows[pos] + rows[rev(pos)] for an edge whose
    endpoints are both listed."""
    idx, adj = _u32(idx), _u32(adj)
    t = np.ascontiguousarray(taus_arr, dtype=np.int32)
    src = _u32(sources)
    n, m = idx.shape[0] - 1, adj.shape[0]
    delta = 1.0 / n if delta is None else delta
    out = np.zeros(m, dtype=np.float32)
    rc = lib().orc_tgtp_sources(n, m, _p(idx, C.c_uint32), _p(adj, C.c_uint32),
                                _p(t, C.c_int32), epsilon, delta, gamma, _p(src, C.c_uint32),
                                src.shape[0], threads, _p(out, C.c_float))
    if rc != 0:
        raise ValueError("orc_tgtp_sources: gamma < 1 or a source id >= n")
    return out


def load_txt(path):
    n, m = C.c_uint64(), C.c_uint64()
    pi, pa = C.POINTER(C.c_uint32)(), C.POINTER(C.c_uint32)()
    if lib().orc_load_txt(str(path).encode(), C.byref(n), C.byref(m), C.byref(pi),
                          C.byref(pa)) != 0:
        raise ValueError(lib().orc_last_error().decode())
    idx = np.ctypeslib.as_array(pi, shape=(n.value + 1,)).copy()
    adj = np.ctypeslib.as_array(pa, shape=(max(m.value, 1),))[:m.value].copy()
    lib().orc_free(C.cast(pi, C.c_void_p))
    lib().orc_free(C.cast(pa, C.c_void_p))
    return idx, adj


def load_eigens(path, n: int, omega: int):
    lam = np.zeros(omega, np.float64)
    X = np.zeros((n, omega), np.float32)
    if lib().orc_load_eigens(str(path).encode(), n, omega, _p(lam, C.c_double),
                             _p(X, C.c_float)) != 0:
        raise ValueError(lib().orc_last_error().decode())
    return lam, X
