"""ctypes binding of libggms.so (the C ABI declared in include/ggms.h).

The library must have been built (ggms-distributed_amd/build.py or
__graft_entry__.build()); loading fails loudly otherwise -- there is no CPU
fallback.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libggms.so")

GGMS_MAX_PASS_STATS = 64
KERNEL_KINDS = ["plan0", "count0", "countS", "colsum", "plan", "compact0", "compactS", "finish", "other"]


class Opts(C.Structure):
    _fields_ = [
        ("num_buckets", C.c_uint32),
        ("num_pivots", C.c_uint32),
        ("sample_size", C.c_uint32),
        ("block_threads", C.c_uint32),
        ("blocks_per_cu", C.c_uint32),
        ("splitter", C.c_uint32),
        ("learned_pieces", C.c_uint32),
        ("max_passes", C.c_uint32),
        ("seed", C.c_uint64),
        ("stream", C.c_void_p),
        ("profile", C.c_uint32),
        ("schedule", C.c_uint32),
        ("device_passes", C.c_uint32),
        ("first_pass", C.c_uint32),
        ("reserved", C.c_uint32 * 4),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("passes", C.c_uint32),
        ("finished_segments", C.c_uint32),
        ("n_in", C.c_uint64 * GGMS_MAX_PASS_STATS),
        ("active", C.c_uint32 * GGMS_MAX_PASS_STATS),
        ("n_kept", C.c_uint64 * GGMS_MAX_PASS_STATS),
        ("buckets", C.c_uint32 * GGMS_MAX_PASS_STATS),
        ("count_ms", C.c_float * GGMS_MAX_PASS_STATS),
        ("compact_ms", C.c_float * GGMS_MAX_PASS_STATS),
        ("kernel_ms", C.c_float * 9),
        ("kernel_launches", C.c_uint32 * 9),
    ]


class RangeU64(C.Structure):
    _fields_ = [
        ("lo", C.c_uint64),
        ("hi", C.c_uint64),
        ("base", C.c_uint32),
        ("nb", C.c_uint32),
        ("sh", C.c_uint32),
        ("mult", C.c_uint32),
    ]


GGMS_BATCH_CLASSES = 24
GGMS_BATCH_MAX_KS = 2048


class BatchClass(C.Structure):
    _fields_ = [
        ("num_buckets", C.c_uint32),
        ("sample_size", C.c_uint32),
        ("pieces", C.c_uint32),
        ("rebucket", C.c_uint32),
    ]


class BatchOpts(C.Structure):
    _fields_ = [
        ("workgroups", C.c_uint32),
        ("block_threads", C.c_uint32),
        ("use_graph", C.c_uint32),
        ("reserved", C.c_uint32),
        ("seed", C.c_uint64),
        ("classes", BatchClass * GGMS_BATCH_CLASSES),
    ]


# (name, restype, argtypes) of every symbol include/ggms.h declares.
_P = C.c_void_p
_U32P = C.POINTER(C.c_uint32)
SIGNATURES = {
    "ggms_opts_default": (None, [C.POINTER(Opts)]),
    "ggms_status_str": (C.c_char_p, [C.c_int]),
    "ggms_ctx_create": (C.c_int, [C.c_int, C.POINTER(_P)]),
    "ggms_ctx_destroy": (None, [_P]),
    "ggms_ctx_stats": (C.c_int, [_P, C.POINTER(Stats)]),
    "ggms_workspace_bytes": (C.c_uint64, [C.c_int, C.c_uint64, C.c_uint32, C.POINTER(Opts)]),
    "ggms_select": (C.c_int, [_P, C.c_int, C.c_int, _P, C.c_uint64, _U32P, C.c_uint32, _P, C.POINTER(Opts)]),
    "ggms_select_device": (C.c_int, [_P, C.c_int, C.c_int, _P, C.c_uint64, _P, C.c_uint32, _P, C.POINTER(Opts)]),
    "ggms_debug_first_pass": (C.c_int, [_P, C.c_int, _P, C.c_uint64, C.POINTER(Opts), C.POINTER(RangeU64), _U32P,
                                        _U32P, _U32P]),
    "ggms_first_pass_tiling": (C.c_int, [C.c_uint64, C.c_uint32, _U32P, _U32P]),
    "ggms_select_topk": (C.c_int, [_P, C.c_int, C.c_int, _P, C.c_uint64, C.c_uint32, _P, C.POINTER(Opts)]),
    "ggms_ctx_reserve": (C.c_int, [_P, C.c_int, C.c_uint64, C.c_uint32, C.POINTER(Opts)]),
    "ggms_ctx_device_status": (C.c_int, [_P]),
    "ggms_select_noalloc": (C.c_int, [C.c_int, C.c_int, _P, C.c_uint64, _P, C.c_uint32, _P, C.POINTER(Opts)]),
    "ggms_last_wrapper_status": (C.c_int, []),
}
SIGNATURES["ggms_generate"] = (C.c_int, [C.c_int, C.c_int, _P, C.c_uint64, C.c_uint64, C.c_double, C.c_double, _P])
SIGNATURES["ggms_mixture_params"] = (C.c_int, [C.c_uint64, C.POINTER(C.c_int), _P, _P])
SIGNATURES["ggms_generate_ks"] = (C.c_int, [C.c_int, C.c_uint32, C.c_uint64, C.c_uint64, _U32P])
SIGNATURES["ggms_mlp_param_count"] = (C.c_uint32, [])
SIGNATURES["ggms_set_learned_weights"] = (C.c_int, [C.c_int, _P, C.c_uint32])
SIGNATURES["ggms_clear_learned_weights"] = (C.c_int, [C.c_int])
SIGNATURES["ggms_mlp_forward"] = (C.c_int, [_P, _P, C.c_uint32, _P, _P])
SIGNATURES["ggms_mlp_workspace_floats"] = (C.c_uint64, [C.c_uint32])
SIGNATURES["ggms_mlp_loss_grad"] = (C.c_int, [_P, _P, _P, C.c_uint32, _P, _P, _P, _P])
SIGNATURES["ggms_mlp_adam_step"] = (C.c_int, [_P, _P, _P, _P, C.c_uint32, C.c_float, C.c_float, C.c_float, C.c_float, _P])
SIGNATURES["ggms_batch_opts_default"] = (None, [C.POINTER(BatchOpts)])
SIGNATURES["ggms_batch_default_classes"] = (None, [C.POINTER(BatchClass)])
SIGNATURES["ggms_batch_plan_create"] = (C.c_int, [C.c_int, C.c_int, _P, C.c_uint32, C.POINTER(C.c_uint64), _P, _U32P,
                                                  _P, C.POINTER(BatchOpts), C.POINTER(_P)])
SIGNATURES["ggms_batch_plan_launch"] = (C.c_int, [_P, _P])
SIGNATURES["ggms_batch_plan_status"] = (C.c_int, [_P, _U32P])
SIGNATURES["ggms_batch_plan_destroy"] = (None, [_P])
SIGNATURES["ggms_select_batched"] = (C.c_int, [C.c_int, C.c_int, _P, C.c_uint32, C.POINTER(C.c_uint64), _P, _U32P,
                                               _P, C.POINTER(BatchOpts)])
for _alg in ("iterativeSMOSWrapper", "bucketMultiselectWrapper"):
    for _suf in ("f32", "f64", "u32", "i32"):
        SIGNATURES[f"ggms_{_alg}_{_suf}"] = (C.c_int, [_P, C.c_int, _U32P, C.c_int, _P, C.c_int, C.c_int])
for _alg in ("distributedSMOSWrapper", "distributedBucketMultiselectWrapper"):
    for _suf in ("f32", "f64", "u32", "i32"):
        SIGNATURES[f"ggms_{_alg}_{_suf}"] = (C.c_int, [_P, C.c_int, _U32P, C.c_int, _P, C.c_int, C.c_int, C.c_int])

_lib = None


def load(path: str = LIB_PATH) -> C.CDLL:
    """Load libggms.so and attach prototypes.  Raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"libggms.so not built at {path}; run ggms-distributed_amd/build.py")
    # torch ships its own HIP runtime with the same soname (libamdhip64.so.7).
    # Whichever is loaded first serves the whole process, so when torch is
    # present it is loaded first: torch and this library then share one
    # runtime (and one null stream), which is what torch needs to find the GPU.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def extra_signature(name: str, restype, argtypes):
    """Register a symbol added by a later module (generators, batched, MLP)."""
    SIGNATURES[name] = (restype, argtypes)
    if _lib is not None:
        fn = getattr(_lib, name)
        fn.restype = restype
        fn.argtypes = argtypes
