"""Context / device-graph wrappers over the C ABI (include/tgtp_hip.h)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .graph import Graph


def _check(ctx_ptr, rc: int, what: str) -> None:
    if rc != L.TGTP_OK:
        lib = L.lib()
        msg = lib.tgtp_ctx_last_error(ctx_ptr).decode() if ctx_ptr else ""
        raise L.TgtpError(rc, f"{what}: {lib.tgtp_strerror(rc).decode()}: {msg}")


class DeviceGraph:
    """A CSR resident in HBM (tgtp_graph_upload)."""

    def __init__(self, ctx: "Context", g: Graph):
        self.ctx = ctx
        self.n, self.m = g.n, g.m
        self.host = g
        h = C.c_void_p()
        gs = g.struct()
        _check(ctx.ptr, L.lib().tgtp_graph_upload(ctx.ptr, C.byref(gs), C.byref(h)),
               "tgtp_graph_upload")
        self.ptr = h

    def free(self) -> None:
        if self.ptr:
            L.lib().tgtp_graph_free(self.ctx.ptr, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Batch:
    """A batch of graphs scored together (tgtp_batch_*).  Keeps the host
    arrays alive for the batch's lifetime."""

    def __init__(self, ctx: "Context", graphs, eigens, epsilon: float, gamma: int = 10,
                 delta=None, mode: int = L.MODE_FAST_FP32, autotune: bool = False,
                 graph: bool = False):
        self.ctx = ctx
        self.graphs = list(graphs)
        k = len(self.graphs)
        self._keep = []
        gs = (L.GraphStruct * k)()
        es = (L.EigensStruct * k)()
        ps = (L.ParamsStruct * k)()
        for i, (g, (lam, X)) in enumerate(zip(self.graphs, eigens)):
            gs[i] = g.struct()
            e, keep = Context._eigens(lam, X)
            es[i] = e
            self._keep.append((g, keep))
            ps[i] = L.ParamsStruct(epsilon, (1.0 / g.n) if delta is None else delta, gamma, mode)
        self._structs = (gs, es, ps)
        h = C.c_void_p()
        flags = (L.BATCH_AUTOTUNE if autotune else 0) | (L.BATCH_GRAPH if graph else 0)
        _check(ctx.ptr, L.lib().tgtp_batch_create(ctx.ptr, k, gs, es, ps, flags, C.byref(h)),
               "tgtp_batch_create")
        self.ptr = h

    def run(self):
        """-> (list of float32 score arrays, stats)."""
        outs = [np.zeros(g.m, dtype=np.float32) for g in self.graphs]
        arr = (C.POINTER(C.c_float) * len(outs))(*[L.ptr(o, C.c_float) for o in outs])
        st = L.StatsStruct()
        _check(self.ctx.ptr, L.lib().tgtp_batch_run(self.ctx.ptr, self.ptr, arr, C.byref(st)),
               "tgtp_batch_run")
        return outs, st.as_dict()

    def profile(self) -> dict:
        """tgtp_batch_profile: per-kernel times of one uncaptured run and the
        round kernel's algorithmic bytes."""
        pr = L.ProfileStruct()
        _check(self.ctx.ptr, L.lib().tgtp_batch_profile(self.ctx.ptr, self.ptr, C.byref(pr)),
               "tgtp_batch_profile")
        return pr.as_dict()

    def config(self) -> dict:
        v = [C.c_uint32() for _ in range(4)]
        rc = L.lib().tgtp_batch_config(self.ptr, *[C.byref(x) for x in v])
        if rc != L.TGTP_OK:
            return {}
        return dict(zip(("light_degree", "small_bouquets", "slots", "max_slots"),
                        [x.value for x in v]))

    def close(self) -> None:
        if self.ptr:
            L.lib().tgtp_batch_destroy(self.ctx.ptr, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    """One HIP stream + scratch on one device (tgtp_ctx)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        rc = L.lib().tgtp_ctx_create(device, C.byref(h))
        if rc != L.TGTP_OK:
            raise L.TgtpError(rc, f"tgtp_ctx_create({device}): "
                                  f"{L.lib().tgtp_strerror(rc).decode()}")
        self.ptr = h
        self.device = device

    def close(self) -> None:
        if self.ptr:
            L.lib().tgtp_ctx_destroy(self.ptr)
            self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, g: Graph) -> DeviceGraph:
        return DeviceGraph(self, g)

    def saba_walks(self, dg: DeviceGraph, depth: int, n_walks: int, src_begin: int = 0,
                   src_end: int | None = None, want_counters: bool = True):
        """-> (counters uint32[n] or None, visit_count, checksum, stats dict)."""
        src_end = dg.n if src_end is None else src_end
        cnt = np.zeros(dg.n, dtype=np.uint32) if want_counters else None
        vc, cs = C.c_uint64(), C.c_uint64()
        st = L.StatsStruct()
        rc = L.lib().tgtp_saba_walks(self.ptr, dg.ptr, src_begin, src_end, depth, n_walks,
                                     L.ptr(cnt, C.c_uint32) if cnt is not None else None,
                                     C.byref(vc), C.byref(cs), C.byref(st))
        _check(self.ptr, rc, "tgtp_saba_walks")
        return cnt, int(vc.value), int(cs.value), st.as_dict()

    def mt_walks(self, dg: DeviceGraph, depth: int, n_walks: int, variant: int = L.WALK_MT_REGULAR,
                 src_begin: int = 0, src_end: int | None = None, want_counters: bool = True,
                 noalloc: bool = False):
        """MT19937 comparator walks (tgtp_mt_walks, or tgtp_mt_walks_noalloc
        with noalloc=True) -> (counters uint32[n] or None, visit_count,
        checksum, stats dict)."""
        src_end = dg.n if src_end is None else src_end
        cnt = np.zeros(dg.n, dtype=np.uint32) if want_counters else None
        vc, cs = C.c_uint64(), C.c_uint64()
        st = L.StatsStruct()
        fn = L.lib().tgtp_mt_walks_noalloc if noalloc else L.lib().tgtp_mt_walks
        rc = fn(self.ptr, dg.ptr, src_begin, src_end, depth, n_walks, variant,
                L.ptr(cnt, C.c_uint32) if cnt is not None else None,
                C.byref(vc), C.byref(cs), C.byref(st))
        _check(self.ptr, rc, "tgtp_mt_walks_noalloc" if noalloc else "tgtp_mt_walks")
        return cnt, int(vc.value), int(cs.value), st.as_dict()

    @staticmethod
    def _eigens(lam: np.ndarray, X: np.ndarray):
        lam = np.ascontiguousarray(lam, dtype=np.float64)
        X = np.ascontiguousarray(X, dtype=np.float32)
        return L.EigensStruct(lam.shape[0], L.ptr(lam, C.c_double), L.ptr(X, C.c_float)), (lam, X)

    def taus(self, g: Graph, lam, X, eps_half: float, mode: int = L.MODE_FAITHFUL):
        es, keep = self._eigens(lam, X)
        out = np.zeros(g.m, dtype=np.int32)
        st = L.StatsStruct()
        gs = g.struct()
        rc = L.lib().tgtp_taus(self.ptr, C.byref(gs), C.byref(es), float(np.float32(eps_half)),
                               mode, L.ptr(out, C.c_int32), C.byref(st))
        _check(self.ptr, rc, "tgtp_taus")
        del keep
        return out, st.as_dict()

    def score(self, g, lam, X, epsilon: float, gamma: int = 10, delta: float | None = None,
              mode: int = L.MODE_FAITHFUL):
        """tgtp() scores float32[m] (CSR order) and stats.  `g` is a Graph
        (copied to HBM per call) or a DeviceGraph."""
        es, keep = self._eigens(lam, X)
        n = g.n
        ps = L.ParamsStruct(epsilon, (1.0 / n) if delta is None else delta, gamma, mode)
        out = np.zeros(g.m, dtype=np.float32)
        st = L.StatsStruct()
        if isinstance(g, DeviceGraph):
            rc = L.lib().tgtp_score_dev(self.ptr, g.ptr, C.byref(es), C.byref(ps),
                                        L.ptr(out, C.c_float), C.byref(st))
        else:
            gs = g.struct()
            rc = L.lib().tgtp_score(self.ptr, C.byref(gs), C.byref(es), C.byref(ps),
                                    L.ptr(out, C.c_float), C.byref(st))
        _check(self.ptr, rc, "tgtp_score")
        del keep
        return out, st.as_dict()

    def batch(self, graphs, eigens, epsilon: float, gamma: int = 10, delta=None,
              mode: int = L.MODE_FAST_FP32, autotune: bool = False, graph: bool = False) -> Batch:
        return Batch(self, graphs, eigens, epsilon, gamma, delta, mode, autotune, graph)

    def score_batch(self, graphs, eigens, epsilon: float, gamma: int = 10, delta=None,
                    mode: int = L.MODE_FAST_FP32):
        """One-shot tgtp_score_batch -> (list of score arrays, stats)."""
        graphs = list(graphs)
        k = len(graphs)
        gs = (L.GraphStruct * k)()
        es = (L.EigensStruct * k)()
        ps = (L.ParamsStruct * k)()
        keep = []
        for i, (g, (lam, X)) in enumerate(zip(graphs, eigens)):
            gs[i] = g.struct()
            e, kp = self._eigens(lam, X)
            es[i] = e
            keep.append(kp)
            ps[i] = L.ParamsStruct(epsilon, (1.0 / g.n) if delta is None else delta, gamma, mode)
        outs = [np.zeros(g.m, dtype=np.float32) for g in graphs]
        arr = (C.POINTER(C.c_float) * k)(*[L.ptr(o, C.c_float) for o in outs])
        st = L.StatsStruct()
        _check(self.ptr, L.lib().tgtp_score_batch(self.ptr, k, gs, es, ps, arr, C.byref(st)),
               "tgtp_score_batch")
        del keep
        return outs, st.as_dict()

    def score_sources(self, dg: DeviceGraph, lam, X, epsilon: float, ranges, gamma: int = 10,
                      delta: float | None = None, mode: int = L.MODE_FAST_FP32):
        """tgtp_score_sources over half-open source ranges [(begin, end), ...]:
        (rows, scores, stats); rows are NaN where the source is not listed,
        scores NaN where the source or the target is not listed."""
        es, keep = self._eigens(lam, X)
        n, m = dg.n, dg.m
        ps = L.ParamsStruct(epsilon, (1.0 / n) if delta is None else delta, gamma, mode)
        rg = np.ascontiguousarray(np.asarray(ranges, dtype=np.uint32).reshape(-1))
        rows = np.full(m, np.nan, dtype=np.float32)
        scores = np.full(m, np.nan, dtype=np.float32)
        st = L.StatsStruct()
        rc = L.lib().tgtp_score_sources(self.ptr, dg.ptr, C.byref(es), C.byref(ps),
                                        rg.shape[0] // 2, L.ptr(rg, C.c_uint32),
                                        L.ptr(rows, C.c_float), L.ptr(scores, C.c_float),
                                        C.byref(st))
        _check(self.ptr, rc, "tgtp_score_sources")
        del keep
        return rows, scores, st.as_dict()

    def score_trace(self, dg: DeviceGraph, lam, X, epsilon: float, gamma: int = 10,
                    delta: float | None = None, mode: int = L.MODE_FAITHFUL):
        """tgtp_score_trace: scores plus the per-stage arrays (same keys as
        the oracle's trace)."""
        es, keep = self._eigens(lam, X)
        n, m = dg.n, dg.m
        ps = L.ParamsStruct(epsilon, (1.0 / n) if delta is None else delta, gamma, mode)
        out = np.zeros(m, dtype=np.float32)
        tr = dict(taus=np.zeros(m, np.int32), nr=np.zeros(m, np.uint64),
                  walk_sum=np.zeros(m, np.float32), preds_row=np.zeros(m, np.float32),
                  ell=np.zeros(n, np.int32), edge_max=np.zeros(n, np.float32),
                  gmax=np.zeros(n, np.float32), gmin=np.zeros(n, np.float32),
                  q_size=np.zeros(n, np.uint32))
        ct = {np.int32: C.c_int32, np.uint64: C.c_uint64, np.float32: C.c_float,
              np.uint32: C.c_uint32}
        args = [L.ptr(tr[k], ct[tr[k].dtype.type]) for k in
                ("taus", "nr", "walk_sum", "preds_row", "ell", "edge_max", "gmax", "gmin",
                 "q_size")]
        st = L.StatsStruct()
        rc = L.lib().tgtp_score_trace(self.ptr, dg.ptr, C.byref(es), C.byref(ps),
                                      L.ptr(out, C.c_float), C.byref(st), *args)
        _check(self.ptr, rc, "tgtp_score_trace")
        del keep
        return out, tr, st.as_dict()
