"""ctypes view of libtgtp_hip.so (include/tgtp_hip.h, include/tgtp_graph_io.h).

The library is built in-tree by tgtp_amd.build; importing this module never
compiles anything, and raises if the library is missing so that no caller can
silently fall back to CPU code.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libtgtp_hip.so"

TGTP_OK = 0
TGTP_E_ARG = -1
TGTP_E_HIP = -2
TGTP_E_NOMEM = -3
TGTP_E_IO = -4
TGTP_E_UNSUPPORTED = -5

MODE_FAITHFUL = 0
MODE_FAST_FP32 = 1
MODE_FAST_BF16 = 2
# FAST modes: the reference's tau refinement rule (aesc.h:130) instead of the bounded one
TAU_REFERENCE = 0x100
# FAITHFUL with the MT19937 2-way walks of the reference baseline (aesc_baseline.h:363-379)
WALK2_MT19937 = 0x200

WALK_MT_REGULAR = 0
WALK_MT_SIMD = 1

BATCH_AUTOTUNE = 1
BATCH_GRAPH = 2


class TgtpError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code


class CsrStruct(C.Structure):
    _fields_ = [("n", C.c_uint64), ("m", C.c_uint64),
                ("idx", C.POINTER(C.c_uint32)), ("adj", C.POINTER(C.c_uint32))]


class GraphStruct(C.Structure):
    _fields_ = [("n", C.c_uint64), ("m", C.c_uint64),
                ("idx", C.POINTER(C.c_uint32)), ("adj", C.POINTER(C.c_uint32))]


class EigensStruct(C.Structure):
    _fields_ = [("omega", C.c_int32), ("values", C.POINTER(C.c_double)),
                ("vectors", C.POINTER(C.c_float))]


class ParamsStruct(C.Structure):
    _fields_ = [("epsilon", C.c_double), ("delta", C.c_double),
                ("gamma", C.c_int32), ("mode", C.c_int32)]


class StatsStruct(C.Structure):
    _fields_ = [("kernel_ms", C.c_double), ("e2e_ms", C.c_double),
                ("walk_ms", C.c_double), ("taus_ms", C.c_double),
                ("diffuse_ms", C.c_double), ("transitions", C.c_uint64),
                ("work_items", C.c_uint64), ("host_fixups", C.c_uint64)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class ProfileStruct(C.Structure):
    _fields_ = [("total_ms", C.c_double), ("taus_ms", C.c_double), ("round_ms", C.c_double),
                ("round_launches", C.c_uint64), ("round_bytes", C.c_uint64),
                ("stop_ms", C.c_double), ("stop_launches", C.c_uint64),
                ("post_ms", C.c_double), ("post_launches", C.c_uint64),
                ("source_rounds", C.c_uint64)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


# every symbol the two public headers declare
EXPORTED = [
    "tgtp_ctx_create", "tgtp_ctx_destroy", "tgtp_ctx_last_error", "tgtp_strerror",
    "tgtp_graph_upload", "tgtp_graph_free", "tgtp_saba_walks", "tgtp_mt_walks",
    "tgtp_mt_walks_noalloc", "tgtp_taus",
    "tgtp_score", "tgtp_score_dev", "tgtp_score_trace", "tgtp_score_sources",
    "tgtp_batch_create", "tgtp_batch_run", "tgtp_batch_destroy", "tgtp_batch_config",
    "tgtp_batch_profile",
    "tgtp_score_batch",
    "tgtp_load_txt", "tgtp_csr_free", "tgtp_save_csr_bin", "tgtp_load_csr_bin",
    "tgtp_load_eigens_txt", "tgtp_save_eigens_bin", "tgtp_load_eigens_bin",
    "tgtp_check_aesc_graph", "tgtp_gen_chung_lu", "tgtp_gen_grid_rewired", "tgtp_write_graph_txt",
    "tgtp_csr_reverse_index",
]

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is not built; run `python -m tgtp_amd.build`")
    # One HIP runtime per process.  torch ships its own libamdhip64 /
    # libhsa-runtime64 (SONAMEs libamdhip64.so.7 / libhsa-runtime64.so.1,
    # but torch asks for them by unversioned names), so if this library
    # loaded /opt/rocm's copies first, a later torch import would bring a
    # second HSA runtime into the process and torch would find no GPU.
    # Importing torch first makes this library bind to the runtime torch
    # already loaded (the dynamic loader matches our versioned NEEDED
    # entries against those SONAMEs).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(str(LIB_PATH))
    P = C.POINTER
    vp = C.c_void_p
    L.tgtp_ctx_create.argtypes = [C.c_int, P(vp)]
    L.tgtp_ctx_destroy.argtypes = [vp]
    L.tgtp_ctx_destroy.restype = None
    L.tgtp_ctx_last_error.argtypes = [vp]
    L.tgtp_ctx_last_error.restype = C.c_char_p
    L.tgtp_strerror.argtypes = [C.c_int]
    L.tgtp_strerror.restype = C.c_char_p
    L.tgtp_graph_upload.argtypes = [vp, P(GraphStruct), P(vp)]
    L.tgtp_graph_free.argtypes = [vp, vp]
    L.tgtp_graph_free.restype = None
    L.tgtp_saba_walks.argtypes = [vp, vp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                  P(C.c_uint32), P(C.c_uint64), P(C.c_uint64), P(StatsStruct)]
    L.tgtp_mt_walks.argtypes = [vp, vp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int32,
                                P(C.c_uint32), P(C.c_uint64), P(C.c_uint64), P(StatsStruct)]
    L.tgtp_mt_walks_noalloc.argtypes = [vp, vp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int32,
                                        P(C.c_uint32), P(C.c_uint64), P(C.c_uint64), P(StatsStruct)]
    L.tgtp_taus.argtypes = [vp, P(GraphStruct), P(EigensStruct), C.c_float, C.c_int32,
                            P(C.c_int32), P(StatsStruct)]
    L.tgtp_score.argtypes = [vp, P(GraphStruct), P(EigensStruct), P(ParamsStruct),
                             P(C.c_float), P(StatsStruct)]
    L.tgtp_score_dev.argtypes = [vp, vp, P(EigensStruct), P(ParamsStruct),
                                 P(C.c_float), P(StatsStruct)]
    L.tgtp_score_trace.argtypes = [vp, vp, P(EigensStruct), P(ParamsStruct), P(C.c_float),
                                   P(StatsStruct), P(C.c_int32), P(C.c_uint64), P(C.c_float),
                                   P(C.c_float), P(C.c_int32), P(C.c_float), P(C.c_float),
                                   P(C.c_float), P(C.c_uint32)]
    L.tgtp_score_sources.argtypes = [vp, vp, P(EigensStruct), P(ParamsStruct), C.c_uint32,
                                     P(C.c_uint32), P(C.c_float), P(C.c_float), P(StatsStruct)]
    L.tgtp_batch_create.argtypes = [vp, C.c_int32, P(GraphStruct), P(EigensStruct),
                                    P(ParamsStruct), C.c_int32, P(vp)]
    L.tgtp_batch_run.argtypes = [vp, vp, P(P(C.c_float)), P(StatsStruct)]
    L.tgtp_batch_destroy.argtypes = [vp, vp]
    L.tgtp_batch_destroy.restype = None
    L.tgtp_batch_profile.argtypes = [vp, vp, P(ProfileStruct)]
    L.tgtp_batch_config.argtypes = [vp, P(C.c_uint32), P(C.c_uint32), P(C.c_uint32),
                                    P(C.c_uint32)]
    L.tgtp_score_batch.argtypes = [vp, C.c_int32, P(GraphStruct), P(EigensStruct),
                                   P(ParamsStruct), P(P(C.c_float)), P(StatsStruct)]
    L.tgtp_load_txt.argtypes = [C.c_char_p, P(CsrStruct), C.c_char_p, C.c_size_t]
    L.tgtp_csr_free.argtypes = [P(CsrStruct)]
    L.tgtp_csr_free.restype = None
    L.tgtp_load_eigens_txt.argtypes = [C.c_char_p, C.c_uint64, C.c_int32, P(C.c_double),
                                       P(C.c_float), C.c_char_p, C.c_size_t]
    L.tgtp_save_eigens_bin.argtypes = [C.c_char_p, C.c_uint64, C.c_int32, P(C.c_double),
                                       P(C.c_float)]
    L.tgtp_load_eigens_bin.argtypes = [C.c_char_p, C.c_uint64, C.c_int32, P(C.c_double),
                                       P(C.c_float), C.c_char_p, C.c_size_t]
    L.tgtp_check_aesc_graph.argtypes = [P(CsrStruct), C.c_int32, C.c_char_p, C.c_size_t]
    L.tgtp_gen_chung_lu.argtypes = [C.c_uint64, C.c_double, C.c_double, C.c_double,
                                    C.c_uint64, P(CsrStruct)]
    L.tgtp_gen_grid_rewired.argtypes = [C.c_uint32, C.c_uint32, C.c_double, C.c_uint64,
                                        P(CsrStruct)]
    L.tgtp_write_graph_txt.argtypes = [P(CsrStruct), C.c_char_p]
    L.tgtp_csr_reverse_index.argtypes = [P(CsrStruct), P(C.c_uint32)]
    _lib = L
    return L


def ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))
