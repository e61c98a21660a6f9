"""ctypes binding of libwmd.so (the C ABI declared in include/wmd.h).

There is no fallback: if the library is missing or fails to load, every
entry point raises.  The library must be built first (build.py).
"""
from __future__ import annotations

import ctypes as C
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libwmd.so")

WMD_OK = 0
WMD_ERR_ARG = -1
WMD_ERR_HIP = -2
WMD_ERR_NONFINITE = -3
WMD_ERR_STATE = -4
WMD_ERR_UNSUPPORTED = -5
STATUS_NAMES = {0: "WMD_OK", -1: "WMD_ERR_ARG", -2: "WMD_ERR_HIP", -3: "WMD_ERR_NONFINITE",
                -4: "WMD_ERR_STATE", -5: "WMD_ERR_UNSUPPORTED"}

VARIANT_RESIDENT = 0
VARIANT_FUSED = 1
VARIANT_NONFUSED = 2

DOMAIN_LINEAR = 0
DOMAIN_LOG = 1

CDIST_FP32 = 0
CDIST_MFMA = 1

WARM_UNIFORM = 0
WARM_LEARNED = 1
WARM_GIVEN = 2
ACCEL_NONE = 0
ACCEL_ANDERSON = 1
SHARE_AUTO = 0
SHARE_OFF = 1
SHARE_ON = 2
MLP_HIDDEN = 32
MLP_FEATURES = 15

# Every symbol include/wmd.h declares.
EXPORTS = ("wmd_options_default", "wmd_sinkhorn", "wmd_sinkhorn_nonfused", "wmd_sinkhorn_ex",
           "wmd_context_create", "wmd_context_destroy", "wmd_context_set_corpus",
           "wmd_context_set_queries", "wmd_context_run", "wmd_context_result",
           "wmd_context_device_result", "wmd_context_run_query_noalloc", "wmd_num_shapes",
           "wmd_shape_info", "wmd_shape_domains",
           "wmd_last_error",
           "wmd_mlp_param_count", "wmd_context_mlp_set", "wmd_context_mlp_get",
           "wmd_train_options_default", "wmd_context_mlp_train", "wmd_context_duals",
           "wmd_mlp_loss_grad", "wmd_context_set_warm", "wmd_context_warm_features",
           "wmd_context_warm_predictions",
           "wmd_context_autotune",
           "wmd_tuning_dims", "wmd_context_get_tuning", "wmd_context_set_tuning",
           "wmd_mtx_header", "wmd_read_mtx", "wmd_vecs_dims", "wmd_read_vecs", "wmd_read_r",
           "wmd_write_scores", "wmd_write_mtx", "wmd_write_vecs", "wmd_write_r",
           "wmd_context_topk", "wmd_prepare_corpus", "wmd_prepared_dims", "wmd_prepared_copy",
           "wmd_prepared_text", "wmd_prepared_free")
TEXT_CSV = 0
TEXT_LINES = 1
TEXT_TOKENS = 0
TEXT_DOCS = 1


class Options(C.Structure):
    _fields_ = [("max_iter", C.c_int32), ("variant", C.c_int32), ("tol", C.c_double),
                ("check_every", C.c_int32), ("shape", C.c_int32), ("summarize", C.c_int32),
                ("domain", C.c_int32), ("cdist", C.c_int32), ("warm_start", C.c_int32),
                ("export_duals", C.c_int32), ("accel", C.c_int32), ("omega", C.c_double),
                ("graph", C.c_int32), ("share", C.c_int32), ("reserved", C.c_int32 * 2)]


class Stats(C.Structure):
    _fields_ = [("ms_total", C.c_double), ("ms_cdist", C.c_double), ("ms_solve", C.c_double),
                ("pairs", C.c_int64), ("iters_max", C.c_int32), ("launches", C.c_int32),
                ("iters_mean", C.c_double), ("resid_max", C.c_double),
                ("variant", C.c_int32), ("fallbacks", C.c_int32), ("ms_warm", C.c_double),
                ("shared", C.c_int32), ("distinct", C.c_int32), ("streamed_docs", C.c_int32),
                ("reserved", C.c_int32 * 1)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class TrainOptions(C.Structure):
    _fields_ = [("epochs", C.c_int32), ("pairs_per_batch", C.c_int32), ("lr", C.c_double),
                ("weight_decay", C.c_double), ("seed", C.c_uint64), ("standardize", C.c_int32),
                ("reserved", C.c_int32 * 7)]


class WMDError(RuntimeError):
    def __init__(self, status: int, where: str, detail: str):
        self.status = status
        super().__init__(f"{where}: {STATUS_NAMES.get(status, status)}: {detail}")


_lib = None


def load():
    """Load libwmd.so and declare its prototypes (raises if absent)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run the package's build.py "
                          "(hipcc, gfx950) -- there is no CPU fallback")
    L = C.CDLL(LIB_PATH)
    P, U64 = C.c_void_p, C.c_uint64
    L.wmd_options_default.argtypes = [C.POINTER(Options)]
    L.wmd_options_default.restype = None
    ref_sig = [P, P, P, C.c_double, C.c_int, U64, U64, U64, P, P, P]
    L.wmd_sinkhorn.argtypes = ref_sig
    L.wmd_sinkhorn_nonfused.argtypes = ref_sig
    L.wmd_sinkhorn_ex.argtypes = [P, P, P, C.c_double, U64, U64, U64, P, P, P,
                                  C.POINTER(Options), C.POINTER(Stats)]
    L.wmd_context_create.argtypes = [C.c_int, C.POINTER(C.c_void_p)]
    L.wmd_context_destroy.argtypes = [P]
    L.wmd_context_destroy.restype = None
    L.wmd_context_set_corpus.argtypes = [P, P, P, U64, U64, P, P, P, U64]
    L.wmd_context_set_queries.argtypes = [P, U64, P, P, P]
    L.wmd_context_run.argtypes = [P, C.c_double, C.POINTER(Options), C.POINTER(Stats)]
    L.wmd_context_result.argtypes = [P, P, P, P]
    L.wmd_context_run_query_noalloc.argtypes = [P, C.c_int32, P, P, C.c_double,
                                                C.POINTER(Options), P, P, P]
    L.wmd_context_device_result.argtypes = [P]
    L.wmd_context_device_result.restype = C.c_void_p
    L.wmd_num_shapes.argtypes = []
    L.wmd_shape_domains.argtypes = [C.c_int]
    L.wmd_shape_info.argtypes = [C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                 C.POINTER(C.c_int32)]
    L.wmd_last_error.argtypes = []
    L.wmd_last_error.restype = C.c_char_p
    L.wmd_mlp_param_count.argtypes = []
    L.wmd_mlp_param_count.restype = C.c_uint64
    L.wmd_context_mlp_set.argtypes = [P, P, U64]
    L.wmd_context_mlp_get.argtypes = [P, P, U64]
    L.wmd_train_options_default.argtypes = [C.POINTER(TrainOptions)]
    L.wmd_train_options_default.restype = None
    L.wmd_context_mlp_train.argtypes = [P, C.POINTER(TrainOptions), P]
    L.wmd_context_duals.argtypes = [P, P]
    L.wmd_context_set_warm.argtypes = [P, P]
    L.wmd_mlp_loss_grad.argtypes = [C.c_int, U64, P, P, U64, P, P, P, P]
    L.wmd_context_warm_features.argtypes = [P, P]
    L.wmd_context_warm_predictions.argtypes = [P, P]
    L.wmd_context_autotune.argtypes = [P, C.c_double, C.POINTER(Options), C.c_int32,
                                       C.POINTER(C.c_double)]
    L.wmd_tuning_dims.argtypes = [C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
    L.wmd_context_get_tuning.argtypes = [P, C.c_int32, P, U64]
    L.wmd_context_set_tuning.argtypes = [P, C.c_int32, P, U64]
    S = C.c_char_p
    L.wmd_mtx_header.argtypes = [S, P, P, P]
    L.wmd_read_mtx.argtypes = [S, U64, U64, U64, P, P, P]
    L.wmd_vecs_dims.argtypes = [S, P, P]
    L.wmd_read_vecs.argtypes = [S, U64, U64, P]
    L.wmd_read_r.argtypes = [S, U64, P]
    L.wmd_write_scores.argtypes = [S, P, U64]
    L.wmd_write_mtx.argtypes = [S, U64, U64, P, P, P]
    L.wmd_write_vecs.argtypes = [S, U64, U64, P]
    L.wmd_write_r.argtypes = [S, U64, P]
    L.wmd_context_topk.argtypes = [P, C.c_int32, P, P]
    L.wmd_prepare_corpus.argtypes = [S, S, C.c_int, U64, U64, C.POINTER(C.c_void_p)]
    L.wmd_prepared_dims.argtypes = [P, P, P, P, P]
    L.wmd_prepared_copy.argtypes = [P, P, P, P, P, P, P]
    L.wmd_prepared_text.argtypes = [P, C.c_int, P, U64, P]
    L.wmd_prepared_free.argtypes = [P]
    L.wmd_prepared_free.restype = None
    for name in ("wmd_prepare_corpus", "wmd_prepared_dims", "wmd_prepared_copy",
                 "wmd_prepared_text",
                 "wmd_sinkhorn", "wmd_sinkhorn_nonfused", "wmd_sinkhorn_ex",
                 "wmd_context_create", "wmd_context_set_corpus", "wmd_context_set_queries",
                 "wmd_context_run", "wmd_context_result", "wmd_context_run_query_noalloc",
                 "wmd_num_shapes", "wmd_shape_info",
                 "wmd_shape_domains",
                 "wmd_context_mlp_set", "wmd_context_mlp_get", "wmd_context_mlp_train",
                 "wmd_context_duals", "wmd_mlp_loss_grad", "wmd_context_set_warm",
                 "wmd_context_warm_features", "wmd_context_warm_predictions",
                 "wmd_context_autotune", "wmd_tuning_dims", "wmd_context_get_tuning",
                 "wmd_context_set_tuning", "wmd_mtx_header", "wmd_read_mtx", "wmd_vecs_dims",
                 "wmd_read_vecs", "wmd_read_r", "wmd_write_scores", "wmd_write_mtx",
                 "wmd_write_vecs", "wmd_write_r", "wmd_context_topk"):
        getattr(L, name).restype = C.c_int
    _lib = L
    return L


def check(status: int, where: str) -> int:
    if status != WMD_OK:
        raise WMDError(status, where, load().wmd_last_error().decode(errors="replace"))
    return status


def default_options(**over) -> Options:
    o = Options()
    load().wmd_options_default(C.byref(o))
    for k, v in over.items():
        setattr(o, k, v)
    return o
