"""ctypes loader of the CPU oracle (test infrastructure).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, as the checker / CPU baseline.  See oracle.cpp for what each
function restates and how it is pinned.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")

DT = {np.dtype(np.float32): 0, np.dtype(np.float64): 1, np.dtype(np.uint32): 2, np.dtype(np.int32): 3}
DIST = {"uniform": 0, "normal": 1, "half_normal": 2, "lognormal": 3, "mixture": 4, "near_constant": 5,
        "constant": 6, "cauchy": 7, "mod_index": 8, "ascending": 9}
KDIST = {"uniform_random": 0, "uniform": 1, "normal": 2, "cluster": 3, "sectioned": 4, "distinct_sorted": 5,
         "even": 6, "window": 7}

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        lib = C.CDLL(LIB)
        P = C.c_void_p
        U32P = C.POINTER(C.c_uint32)
        lib.oracle_sort_and_choose.argtypes = [C.c_int, P, C.c_uint64, U32P, C.c_uint32, P, C.c_int]
        lib.oracle_nth_multiselect.argtypes = [C.c_int, P, C.c_uint64, U32P, C.c_uint32, P, C.c_int]
        lib.oracle_keys.argtypes = [C.c_int, P, C.c_uint64, P]
        lib.oracle_philox4x32_10.argtypes = [U32P, U32P, U32P]
        lib.oracle_detmath.argtypes = [C.c_int, P, C.c_uint64, P, P]
        lib.oracle_generate.argtypes = [C.c_int, C.c_int, C.c_uint64, C.c_uint64, C.c_double, C.c_double, P]
        lib.oracle_generate_ks.argtypes = [C.c_int, C.c_uint32, C.c_uint64, C.c_uint64, U32P]
        lib.oracle_mixture_params.argtypes = [C.c_uint64, C.POINTER(C.c_int), P, P]
        lib.oracle_map_bucket.argtypes = [P, C.c_uint32, C.c_uint64]
        lib.oracle_map_bucket.restype = C.c_uint32
        lib.oracle_first_pass_hist.argtypes = [C.c_int, P, C.c_uint64, P, C.c_uint32, C.c_uint32, U32P]
        lib.oracle_mlp_forward.argtypes = [P, P, C.c_uint32, P]
        lib.oracle_mlp_loss_grad.argtypes = [P, P, P, C.c_uint32, P]
        lib.oracle_mlp_loss_grad.restype = C.c_double
        _lib = lib
    return _lib


def _u32p(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def sort_and_choose(x: np.ndarray, ks, threads: int = 1) -> np.ndarray:
    """k-th largest (k = 1 is the max) by full sort, as the reference oracle."""
    x = np.ascontiguousarray(x)
    ks = np.ascontiguousarray(np.asarray(ks, dtype=np.uint32))
    out = np.empty(len(ks), dtype=x.dtype)
    rc = load().oracle_sort_and_choose(DT[x.dtype], x.ctypes.data, x.size, _u32p(ks), len(ks), out.ctypes.data,
                                       threads)
    if rc == 1:
        raise ValueError("k outside [1, n]")
    if rc == 2:
        raise ValueError("input holds a NaN")
    return out


def nth_multiselect(x: np.ndarray, ks, threads: int = 1) -> np.ndarray:
    x = np.ascontiguousarray(x)
    ks = np.ascontiguousarray(np.asarray(ks, dtype=np.uint32))
    out = np.empty(len(ks), dtype=x.dtype)
    rc = load().oracle_nth_multiselect(DT[x.dtype], x.ctypes.data, x.size, _u32p(ks), len(ks), out.ctypes.data,
                                       threads)
    if rc:
        raise ValueError("bad k")
    return out


def keys(x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x)
    out = np.empty(x.size, dtype=np.uint64 if x.dtype == np.float64 else np.uint32)
    load().oracle_keys(DT[x.dtype], x.ctypes.data, x.size, out.ctypes.data)
    return out


def philox(ctr, key) -> np.ndarray:
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    load().oracle_philox4x32_10(_u32p(c), _u32p(k), _u32p(o))
    return o


def detmath(fn: str, x: np.ndarray):
    x = np.ascontiguousarray(x, dtype=np.float64)
    o1 = np.empty_like(x)
    o2 = np.empty_like(x)
    code = {"log": 0, "exp": 1, "sqrt": 2, "sincos2pi": 3}[fn]
    load().oracle_detmath(code, x.ctypes.data, x.size, o1.ctypes.data, o2.ctypes.data)
    return (o1, o2) if code == 3 else o1


def generate(dtype, dist: str, n: int, seed: int, p0: float = 0.0, p1: float = 0.0) -> np.ndarray:
    dtype = np.dtype(dtype)
    out = np.empty(n, dtype=dtype)
    load().oracle_generate(DT[dtype], DIST[dist], n, seed, p0, p1, out.ctypes.data)
    return out


def generate_ks(kdist: str, nk: int, n: int, seed: int) -> np.ndarray:
    out = np.zeros(nk, dtype=np.uint32)
    rc = load().oracle_generate_ks(KDIST[kdist], nk, n, seed, _u32p(out))
    if rc:
        raise ValueError("bad k distribution")
    return out


def mixture_params(seed: int):
    m = C.c_int()
    mu = np.zeros(8)
    sg = np.zeros(8)
    load().oracle_mixture_params(seed, C.byref(m), mu.ctypes.data, sg.ctypes.data)
    return m.value, mu, sg


RANGE_DTYPE = np.dtype([("lo", "<u8"), ("hi", "<u8"), ("base", "<u4"), ("nb", "<u4"), ("sh", "<u4"), ("mult", "<u4")])


def map_bucket(ranges: np.ndarray, key: int) -> int:
    r = np.ascontiguousarray(ranges.astype(RANGE_DTYPE))
    return load().oracle_map_bucket(r.ctypes.data, len(r), int(key))


def first_pass_hist(x: np.ndarray, ranges: np.ndarray, B: int) -> np.ndarray:
    x = np.ascontiguousarray(x)
    r = np.ascontiguousarray(ranges.astype(RANGE_DTYPE))
    h = np.zeros(B, dtype=np.uint32)
    rc = load().oracle_first_pass_hist(DT[x.dtype], x.ctypes.data, x.size, r.ctypes.data, len(r), B, _u32p(h))
    if rc:
        raise ValueError("bucket outside histogram")
    return h


def mlp_forward(w: np.ndarray, x: np.ndarray) -> np.ndarray:
    w = np.ascontiguousarray(w, dtype=np.float32)
    x = np.ascontiguousarray(x, dtype=np.float32)
    nb = x.shape[0]
    out = np.empty(nb, dtype=np.float32)
    load().oracle_mlp_forward(w.ctypes.data, x.ctypes.data, nb, out.ctypes.data)
    return out


def mlp_loss_grad(w: np.ndarray, x: np.ndarray, y: np.ndarray):
    w = np.ascontiguousarray(w, dtype=np.float32)
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.ascontiguousarray(y, dtype=np.float32)
    g = np.empty_like(w)
    loss = load().oracle_mlp_loss_grad(w.ctypes.data, x.ctypes.data, y.ctypes.data, x.shape[0], g.ctypes.data)
    return loss, g
