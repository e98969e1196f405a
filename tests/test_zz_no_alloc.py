"""The allocation-free entry point (ggms_select_noalloc: a fixed per-device
scratch array, no allocation) through the C ABI, against the existing
ggms_select_device on the same inputs, bit-exact, on two calls in a row, the
second with a smaller input; and the refusal of a call larger than the scratch."""
import ctypes as C

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint64 if a.dtype.itemsize == 8 else np.uint32)


def _noalloc(lib, alg, dt, x, d_ks, d_out, opts=None):
    import ggms
    return lib.ggms_select_noalloc(ggms.ALGS[alg], ggms.DTYPES[dt], C.c_void_p(x.data_ptr()), x.numel(),
                                   C.c_void_p(d_ks.data_ptr()), d_ks.numel(), C.c_void_p(d_out.data_ptr()),
                                   C.byref(opts) if opts is not None else None)


@pytest.mark.parametrize("dt,alg", [("float32", "iterative"), ("float64", "bucket"), ("uint32", "distributed")])
def test_noalloc_matches_select_device(dt, alg):
    import torch
    import ggms
    lib = ggms._lib.load()
    ctx = ggms.Context(0)
    tt = {"float32": torch.float32, "float64": torch.float64, "uint32": torch.uint32}[dt]
    for call, (n, nk, dist, seed) in enumerate((((1 << 22) + 3, 3000, "normal", 11), (150_001, 700, "mixture", 12))):
        x = ggms.generate(dt, "uniform" if dt == "uint32" else dist, n, seed, 5000.0 if dt == "uint32" else 0.0)
        ks = ggms.generate_ks("uniform_random", nk, n, seed + 1)
        d_ks = torch.from_numpy(ks.view(np.int32)).cuda()
        d_ref = torch.empty(nk, dtype=tt, device="cuda")
        ctx.select_device(x, d_ks, d_ref, alg)
        d_out = torch.zeros(nk, dtype=tt, device="cuda")
        print(f"noalloc call {call}: {dt} {alg} n={n} nk={nk}", flush=True)
        assert _noalloc(lib, alg, dt, x, d_ks, d_out) == ggms.OK
        torch.cuda.synchronize()
        got, ref = d_out.cpu().numpy(), d_ref.cpu().numpy()
        assert np.array_equal(_bits(got), _bits(ref)), f"call {call}: differs from ggms_select_device"
        want = oracle.sort_and_choose(x.cpu().numpy(), ks, threads=16)
        assert np.array_equal(_bits(got), _bits(want)), f"call {call}: differs from sort-and-choose"
    ctx.close()


def test_noalloc_refuses_calls_beyond_its_scratch():
    import torch
    import ggms
    lib = ggms._lib.load()
    n = 1 << 26
    o = ggms.default_opts(schedule=1)
    need = lib.ggms_workspace_bytes(0, n, 10, C.byref(o))
    assert need + 256 > (96 << 20)
    x = torch.zeros(n, dtype=torch.float32, device="cuda")
    d_ks = torch.ones(10, dtype=torch.int32, device="cuda")
    d_out = torch.zeros(10, dtype=torch.float32, device="cuda")
    assert _noalloc(lib, "iterative", "float32", x, d_ks, d_out) == ggms.ENOMEM
    # a bad k is reported, the scratch stays usable
    x = ggms.generate("float32", "normal", 100_000, 3)
    ks = np.array([5, 100_001], dtype=np.uint32)
    d_ks = torch.from_numpy(ks.view(np.int32)).cuda()
    d_out = torch.zeros(2, dtype=torch.float32, device="cuda")
    assert _noalloc(lib, "iterative", "float32", x, d_ks, d_out) == ggms.EINVAL
    ks = np.array([5, 77], dtype=np.uint32)
    d_ks = torch.from_numpy(ks.view(np.int32)).cuda()
    assert _noalloc(lib, "iterative", "float32", x, d_ks, d_out) == ggms.OK
    want = oracle.sort_and_choose(x.cpu().numpy(), ks)
    assert np.array_equal(_bits(d_out.cpu().numpy()), _bits(want))
