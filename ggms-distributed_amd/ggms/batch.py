"""Batched selection: many independent problems in one persistent launch.

A problem has the meaning of one reference call
(IterativeSMOS::iterativeSMOSWrapper, DistributedSMOS/iterativeSMOS.cu:1579-1612):
values, k ranks (k = 1 is the largest) -> the k-th largest values.  A batch
concatenates the problems' values (one device array, one dtype) and their ks;
offsets / koffsets (num_problems + 1 entries) delimit them.

    plan = BatchPlan(values, offsets, ks, koffsets, classes=autotune.load())
    plan.launch(); out = plan.result()      # numpy, in ks order
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _lib
from ._lib import BatchClass, BatchOpts

MAX_KS = _lib.GGMS_BATCH_MAX_KS
CLASSES = _lib.GGMS_BATCH_CLASSES


def default_opts(classes: Optional[dict] = None, **kw) -> BatchOpts:
    """BatchOpts with library defaults; classes maps class index -> dict(num_buckets, sample_size, pieces[, rebucket])."""
    o = BatchOpts()
    _lib.load().ggms_batch_opts_default(C.byref(o))
    for c, v in (classes or {}).items():
        c = int(c)
        o.classes[c].num_buckets = int(v["num_buckets"])
        o.classes[c].sample_size = int(v["sample_size"])
        o.classes[c].pieces = int(v["pieces"])
        o.classes[c].rebucket = int(v.get("rebucket", 0))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def size_class(n: int) -> int:
    """Class of a problem of n values: ceil(log2 n), the last class taking every larger size."""
    c = int(n - 1).bit_length() if n > 1 else 0
    return min(c, CLASSES - 1)


class BatchPlan:
    """A batch bound to device buffers; launch() may be repeated (hipGraph replay with use_graph)."""

    def __init__(self, values, offsets, ks, koffsets, classes: Optional[dict] = None, workgroups: int = 0,
                 block_threads: int = 0, use_graph: bool = False, seed: int = 0x5EED, opts: BatchOpts = None):
        import torch
        from . import DTYPES, _check, _dtype_code
        self.lib = _lib.load()
        self.values = values
        self.t = _dtype_code(values.dtype)
        self.device = values.device.index or 0
        self.offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
        self.koffsets = np.ascontiguousarray(koffsets, dtype=np.uint32)
        self.nprob = len(self.offsets) - 1
        nks = int(self.koffsets[-1])
        if torch.is_tensor(ks):
            self.d_ks = ks.to(device=values.device, dtype=torch.uint32).contiguous()
        else:
            self.d_ks = torch.from_numpy(np.ascontiguousarray(ks, dtype=np.uint32)).to(values.device)
        self.out = torch.empty(max(nks, 1), dtype=values.dtype, device=values.device)
        o = opts if opts is not None else default_opts(classes, workgroups=workgroups, block_threads=block_threads,
                                                       use_graph=int(use_graph), seed=seed)
        h = C.c_void_p()
        st = self.lib.ggms_batch_plan_create(
            self.device, self.t, C.c_void_p(values.data_ptr()), self.nprob,
            self.offsets.ctypes.data_as(C.POINTER(C.c_uint64)), C.c_void_p(self.d_ks.data_ptr()),
            self.koffsets.ctypes.data_as(_lib._U32P), C.c_void_p(self.out.data_ptr()), C.byref(o), C.byref(h))
        _check(st, "ggms_batch_plan_create")
        self.h = h

    def launch(self, stream=None):
        from . import _check
        _check(self.lib.ggms_batch_plan_launch(self.h, stream), "ggms_batch_plan_launch")

    def status(self) -> np.ndarray:
        import torch
        torch.cuda.synchronize(self.values.device)
        s = np.zeros(self.nprob, dtype=np.uint32)
        self.lib.ggms_batch_plan_status(self.h, s.ctypes.data_as(_lib._U32P))
        return s

    def result(self, check: bool = True) -> np.ndarray:
        from . import GGMSError
        s = self.status()
        if check and s.any():
            bad = int(np.flatnonzero(s)[0])
            raise GGMSError(int(s[bad]), f"batched problem {bad}")
        return self.out[: int(self.koffsets[-1])].cpu().numpy()

    def close(self):
        if getattr(self, "h", None):
            self.lib.ggms_batch_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def select_batched(values, offsets, ks, koffsets, classes: Optional[dict] = None, **kw) -> np.ndarray:
    """One-shot batched selection through ggms_select_batched (plan, launch, wait, destroy)."""
    import torch
    from . import _check, _dtype_code
    lib = _lib.load()
    offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
    koffsets = np.ascontiguousarray(koffsets, dtype=np.uint32)
    d_ks = torch.from_numpy(np.ascontiguousarray(ks, dtype=np.uint32)).to(values.device)
    out = torch.empty(max(int(koffsets[-1]), 1), dtype=values.dtype, device=values.device)
    o = default_opts(classes, **kw)
    st = lib.ggms_select_batched(values.device.index or 0, _dtype_code(values.dtype), C.c_void_p(values.data_ptr()),
                                 len(offsets) - 1, offsets.ctypes.data_as(C.POINTER(C.c_uint64)),
                                 C.c_void_p(d_ks.data_ptr()), koffsets.ctypes.data_as(_lib._U32P),
                                 C.c_void_p(out.data_ptr()), C.byref(o))
    _check(st, "ggms_select_batched")
    return out[: int(koffsets[-1])].cpu().numpy()


def make_batch(num_problems: int, log2_lo: int, log2_hi: int, max_ks: int, seed: int, dtype: str = "float32",
               dists=("uniform", "normal", "lognormal"), device: int = 0):
    """Synthetic batch: sizes log-uniform in [2^log2_lo, 2^log2_hi], each problem one of `dists`
    (a name, or (name, p0, p1); seeded per problem), 1..max_ks uniform random ranks per problem.

    Returns (values tensor, offsets, ks, koffsets, sizes)."""
    import torch
    from . import DISTS, DTYPES, _check
    rng = np.random.default_rng(seed)
    sizes = np.floor(2.0 ** rng.uniform(log2_lo, log2_hi, num_problems)).astype(np.int64)
    sizes = np.clip(sizes, 1, 2 ** log2_hi)
    offsets = np.zeros(num_problems + 1, dtype=np.uint64)
    offsets[1:] = np.cumsum(sizes)
    nks = rng.integers(1, max_ks + 1, num_problems)
    koffsets = np.zeros(num_problems + 1, dtype=np.uint32)
    koffsets[1:] = np.cumsum(nks)
    ks = np.concatenate([rng.integers(1, s + 1, k) for s, k in zip(sizes, nks)]).astype(np.uint32)
    tt = {"float32": torch.float32, "float64": torch.float64, "uint32": torch.uint32, "int32": torch.int32}[dtype]
    x = torch.empty(int(offsets[-1]), dtype=tt, device=f"cuda:{device}")
    esz = x.element_size()
    lib = _lib.load()
    which = rng.integers(0, len(dists), num_problems)
    seeds = rng.integers(0, 2 ** 62, num_problems)
    for p in range(num_problems):
        ptr = x.data_ptr() + int(offsets[p]) * esz
        spec = dists[which[p]]
        name, p0, p1 = (spec, 0.0, 0.0) if isinstance(spec, str) else spec
        st = lib.ggms_generate(DTYPES[dtype], DISTS[name], C.c_void_p(ptr), int(sizes[p]), int(seeds[p]),
                               float(p0), float(p1), None)
        _check(st, "ggms_generate")
    torch.cuda.synchronize(x.device)
    return x, offsets, ks, koffsets, sizes
