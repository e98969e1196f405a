"""Python host mirror of the multiselect interface (over the C ABI).

Mirrors the reference's entry points (DistributedSMOS/iterativeSMOS.cuh:218-220,
DistributedSMOS/bucketMultiselect.cuh:105-107, DistributedSMOS/distributedSMOS.hpp:22-25):
a device vector, its length, k ranks (k = 1 is the largest) and an output array
of k-th values.  Device memory comes from torch tensors on a ROCm device;
torch is plumbing only -- every computation runs in libggms.so's HIP kernels.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np

from . import _lib
from ._lib import Opts, RangeU64, Stats

__all__ = [
    "GGMSError", "Context", "Opts", "Stats", "default_opts", "select", "select_device", "topk",
    "iterativeSMOSWrapper", "bucketMultiselectWrapper", "distributedSMOSWrapper", "distributedBucketMultiselectWrapper",
    "DTYPES", "ALGS", "SPLITTERS", "DISTS", "KDISTS", "generate", "generate_ks", "mixture_params",
]

OK, EINVAL, ENAN, ENOMEM, EDEVICE, ENOTSUP = range(6)
DTYPES = {"float32": 0, "float64": 1, "uint32": 2, "int32": 3}
ALGS = {"bucket": 0, "iterative": 1, "distributed": 2, "distributed_bucket": 3}
SPLITTERS = {"linear": 0, "learned": 1, "sketch": 2}
DISTS = {"uniform": 0, "normal": 1, "half_normal": 2, "lognormal": 3, "mixture": 4, "near_constant": 5,
         "constant": 6, "cauchy": 7, "mod_index": 8, "ascending": 9}
KDISTS = {"uniform_random": 0, "uniform": 1, "normal": 2, "cluster": 3, "sectioned": 4, "distinct_sorted": 5,
          "even": 6, "window": 7}
_NP = {0: np.float32, 1: np.float64, 2: np.uint32, 3: np.int32}


class GGMSError(RuntimeError):
    def __init__(self, status: int, what: str = ""):
        self.status = status
        name = _lib.load().ggms_status_str(status).decode()
        super().__init__(f"{what}: {name}" if what else name)


def _check(status: int, what: str):
    if status != OK:
        raise GGMSError(status, what)


def default_opts(**kw) -> Opts:
    o = Opts()
    _lib.load().ggms_opts_default(C.byref(o))
    for k, v in kw.items():
        if k == "splitter" and isinstance(v, str):
            v = SPLITTERS[v]
        setattr(o, k, v)
    return o


def _dtype_code(t) -> int:
    name = str(t).replace("torch.", "")
    if name not in DTYPES:
        raise TypeError(f"unsupported dtype {t}")
    return DTYPES[name]


class Context:
    """Owns a device workspace reused across calls (one per host thread)."""

    def __init__(self, device: int = 0):
        self.lib = _lib.load()
        h = C.c_void_p()
        _check(self.lib.ggms_ctx_create(device, C.byref(h)), "ggms_ctx_create")
        self.h = h
        self.device = device

    def close(self):
        if getattr(self, "h", None):
            self.lib.ggms_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def stats(self) -> dict:
        s = Stats()
        _check(self.lib.ggms_ctx_stats(self.h, C.byref(s)), "ggms_ctx_stats")
        p = s.passes
        return {
            "passes": p,
            "finished_segments": s.finished_segments,
            "n_in": [int(s.n_in[i]) for i in range(p)],
            "active": [int(s.active[i]) for i in range(p)],
            "n_kept": [int(s.n_kept[i]) for i in range(p)],
            "buckets": [int(s.buckets[i]) for i in range(p)],
            "count_ms": [float(s.count_ms[i]) for i in range(p)],
            "compact_ms": [float(s.compact_ms[i]) for i in range(p)],
            "kernel_ms": {k: float(s.kernel_ms[i]) for i, k in enumerate(_lib.KERNEL_KINDS)},
            "kernel_launches": {k: int(s.kernel_launches[i]) for i, k in enumerate(_lib.KERNEL_KINDS)},
        }

    def select(self, x, ks: Sequence[int], alg: str = "iterative", opts: Optional[Opts] = None,
               n: Optional[int] = None) -> np.ndarray:
        """k-th largest values of device tensor x for every k in ks (host result)."""
        dt = _dtype_code(x.dtype)
        if not x.is_cuda:
            raise ValueError("x must be a device tensor")
        x = x.contiguous()
        n = x.numel() if n is None else n
        ks_np = np.ascontiguousarray(np.asarray(ks, dtype=np.uint32))
        out = np.empty(len(ks_np), dtype=_NP[dt])
        o = opts if opts is not None else default_opts()
        if o.splitter == SPLITTERS["learned"]:
            _ensure_learned(self.device)
        st = self.lib.ggms_select(self.h, ALGS[alg], dt, C.c_void_p(x.data_ptr()), n,
                                  ks_np.ctypes.data_as(_lib._U32P), len(ks_np), out.ctypes.data_as(C.c_void_p),
                                  C.byref(o))
        _check(st, "ggms_select")
        return out

    def topk(self, x, K: int, alg: str = "iterative", opts: Optional[Opts] = None) -> np.ndarray:
        """The K largest values of device tensor x, largest first (host result)."""
        dt = _dtype_code(x.dtype)
        x = x.contiguous()
        out = np.empty(K, dtype=_NP[dt])
        o = opts if opts is not None else default_opts()
        if o.splitter == SPLITTERS["learned"]:
            _ensure_learned(self.device)
        st = self.lib.ggms_select_topk(self.h, ALGS[alg], dt, C.c_void_p(x.data_ptr()), x.numel(), K,
                                       out.ctypes.data_as(C.c_void_p), C.byref(o))
        _check(st, "ggms_select_topk")
        return out

    def select_device(self, x, d_ks, d_out, alg: str = "iterative", opts: Optional[Opts] = None):
        """Device-resident ks (uint32 tensor) and outputs (tensor of x's dtype)."""
        dt = _dtype_code(x.dtype)
        o = opts if opts is not None else default_opts()
        if o.splitter == SPLITTERS["learned"]:
            _ensure_learned(self.device)
        st = self.lib.ggms_select_device(self.h, ALGS[alg], dt, C.c_void_p(x.data_ptr()), x.numel(),
                                         C.c_void_p(d_ks.data_ptr()), d_ks.numel(), C.c_void_p(d_out.data_ptr()),
                                         C.byref(o))
        _check(st, "ggms_select_device")
        return d_out

    def first_pass(self, x, opts: Optional[Opts] = None, max_ranges: int = 4096, max_buckets: int = 20000):
        """Range table and histogram of the first bucketing pass (introspection)."""
        dt = _dtype_code(x.dtype)
        o = opts if opts is not None else default_opts()
        rr = (RangeU64 * max_ranges)()
        hist = np.zeros(max_buckets, dtype=np.uint32)
        nr = C.c_uint32(max_ranges)
        nb = C.c_uint32(max_buckets)
        st = self.lib.ggms_debug_first_pass(self.h, dt, C.c_void_p(x.data_ptr()), x.numel(), C.byref(o), rr,
                                            C.byref(nr), hist.ctypes.data_as(_lib._U32P), C.byref(nb))
        _check(st, "ggms_debug_first_pass")
        ranges = np.zeros(nr.value, dtype=[("lo", "<u8"), ("hi", "<u8"), ("base", "<u4"), ("nb", "<u4"),
                                           ("sh", "<u4"), ("mult", "<u4")])
        C.memmove(ranges.ctypes.data, rr, nr.value * C.sizeof(RangeU64))
        return ranges, hist[: nb.value].copy()


_learned_installed: set = set()


def _ensure_learned(device: int):
    """Install the committed learned-splitter weights on first use (if present)."""
    if device in _learned_installed:
        return
    from . import learned
    if learned.install(device):
        _learned_installed.add(device)


def set_learned_weights(w, device: int = 0):
    from . import learned
    learned.install(device, w=np.asarray(w, dtype=np.float32))
    _learned_installed.add(device)


_default_ctx: dict = {}


def _ctx(device: int = 0) -> Context:
    if device not in _default_ctx:
        _default_ctx[device] = Context(device)
    return _default_ctx[device]


def select(x, ks, alg: str = "iterative", opts: Optional[Opts] = None) -> np.ndarray:
    return _ctx(x.device.index or 0).select(x, ks, alg, opts)


def select_device(x, d_ks, d_out, alg: str = "iterative", opts: Optional[Opts] = None):
    return _ctx(x.device.index or 0).select_device(x, d_ks, d_out, alg, opts)


def topk(x, K: int, alg: str = "iterative", opts: Optional[Opts] = None) -> np.ndarray:
    return _ctx(x.device.index or 0).topk(x, K, alg, opts)


_SUF = {0: "f32", 1: "f64", 2: "u32", 3: "i32"}


def _wrapper(name, d_vector, length, kVals_ori, numKs, outputs, blocks, threads, *extra):
    lib = _lib.load()
    dt = _dtype_code(d_vector.dtype)
    ks = np.ascontiguousarray(np.asarray(kVals_ori, dtype=np.uint32))
    if outputs.dtype != _NP[dt]:
        raise TypeError("outputs dtype must match the vector")
    fn = getattr(lib, f"ggms_{name}_{_SUF[dt]}")
    return fn(C.c_void_p(d_vector.data_ptr()), int(length), ks.ctypes.data_as(_lib._U32P), int(numKs),
              outputs.ctypes.data_as(C.c_void_p), int(blocks), int(threads), *extra)


def iterativeSMOSWrapper(d_vector, length, kVals_ori, numKs, outputs, blocks=0, threads=1024) -> int:
    """IterativeSMOS::iterativeSMOSWrapper (DistributedSMOS/iterativeSMOS.cu:1579-1602); returns a status."""
    return _wrapper("iterativeSMOSWrapper", d_vector, length, kVals_ori, numKs, outputs, blocks, threads)


def bucketMultiselectWrapper(d_vector, length, kVals_ori, numKs, outputs, blocks=0, threads=1024) -> int:
    """BucketMultiselect::bucketMultiselectWrapper (DistributedSMOS/bucketMultiselect.cu:852-886)."""
    return _wrapper("bucketMultiselectWrapper", d_vector, length, kVals_ori, numKs, outputs, blocks, threads)


def distributedSMOSWrapper(d_vector, length, kVals_ori, numKs, outputs, blocks=0, threads=1024, rank=0) -> int:
    """DistributedSMOS::distributedSMOSWrapper with one rank (DistributedSMOS/distributedSMOS.cpp:1208-1267)."""
    return _wrapper("distributedSMOSWrapper", d_vector, length, kVals_ori, numKs, outputs, blocks, threads,
                    int(rank))


def distributedBucketMultiselectWrapper(d_vector, length, kVals_ori, numKs, outputs, blocks=0, threads=1024,
                                        rank=0) -> int:
    """DistributedBucketMultiselect::distributedBucketMultiselectWrapper with one rank
    (DistributedSMOS/distributedBucketMultiselect.cpp:929-1018)."""
    return _wrapper("distributedBucketMultiselectWrapper", d_vector, length, kVals_ori, numKs, outputs, blocks,
                    threads, int(rank))


def generate(dtype: str, dist: str, n: int, seed: int, p0: float = 0.0, p1: float = 0.0, device: int = 0):
    """Seeded synthetic vector generated on the device (returns a torch tensor)."""
    import torch
    tt = {"float32": torch.float32, "float64": torch.float64, "uint32": torch.uint32, "int32": torch.int32}[dtype]
    x = torch.empty(n, dtype=tt, device=f"cuda:{device}")
    st = _lib.load().ggms_generate(DTYPES[dtype], DISTS[dist], C.c_void_p(x.data_ptr()), n, seed, p0, p1, None)
    _check(st, "ggms_generate")
    torch.cuda.synchronize(x.device)
    return x


def generate_ks(kdist: str, nk: int, n: int, seed: int) -> np.ndarray:
    """k-rank list (host)."""
    out = np.zeros(nk, dtype=np.uint32)
    _check(_lib.load().ggms_generate_ks(KDISTS[kdist], nk, n, seed, out.ctypes.data_as(_lib._U32P)), "ggms_generate_ks")
    return out


def mixture_params(seed: int):
    m = C.c_int()
    mu = np.zeros(8)
    sg = np.zeros(8)
    _check(_lib.load().ggms_mixture_params(seed, C.byref(m), mu.ctypes.data_as(C.c_void_p),
                                           sg.ctypes.data_as(C.c_void_p)), "ggms_mixture_params")
    return m.value, mu, sg
