"""The device-driven schedule (every pass enqueued at once, no host read between
kernels) against the CPU sort-and-choose oracle, bit-exact, and the same call
captured in a hipGraph and replayed on new values in the same buffers.  It
replaces the reference's host-driven iteration loop
(DistributedSMOS/iterativeSMOS.cu:1348-1512), which copies counts and pivots
between host and device on every pass."""
import ctypes as C

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

DEVICE = 1


def _bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint64 if a.dtype.itemsize == 8 else np.uint32)


@pytest.mark.parametrize("dt,dist,n,kdist,nk,alg", [
    ("float32", "normal", (1 << 22) + 5, "uniform_random", 1001, "iterative"),
    ("float32", "half_normal", 1 << 22, "even", 8192, "iterative"),
    ("float32", "near_constant", 1 << 21, "window", 8192, "iterative"),
    ("float64", "normal", 1 << 21, "uniform_random", 500, "distributed"),
    ("uint32", "uniform", 1 << 22, "even", 8192, "bucket"),
    ("int32", "uniform", 20000, "uniform_random", 300, "iterative"),
    ("float32", "mixture", 1 << 22, "uniform_random", 20000, "iterative"),
])
def test_device_schedule_exact(gpu_ctx, dt, dist, n, kdist, nk, alg):
    import ggms
    p0, p1 = (1024.0, 0.0) if dt == "uint32" else ((1.0, 0.01) if dist == "near_constant" else (0.0, 0.0))
    x = ggms.generate(dt, dist, n, 31, p0, p1)
    ks = ggms.generate_ks(kdist, nk, n, 32)
    want = oracle.sort_and_choose(x.cpu().numpy(), ks, threads=16)
    for passes in (0, 2, 5):
        got = gpu_ctx.select(x, ks, alg, ggms.default_opts(schedule=DEVICE, device_passes=passes))
        bad = np.nonzero(_bits(got) != _bits(want))[0]
        assert bad.size == 0, f"passes={passes}: {bad.size} wrong, first k={ks[bad[0]]}"


def test_device_schedule_errors(gpu_ctx):
    import ggms
    import torch
    v = np.random.default_rng(3).standard_normal(100_000).astype(np.float32)
    v[777] = np.nan
    with pytest.raises(ggms.GGMSError) as e:
        gpu_ctx.select(torch.from_numpy(v).cuda(), [5, 9], "iterative", ggms.default_opts(schedule=DEVICE))
    assert e.value.status == ggms.ENAN
    x = ggms.generate("float32", "normal", 100_000, 3)
    with pytest.raises(ggms.GGMSError) as e:
        gpu_ctx.select(x, [5, 100_001], "iterative", ggms.default_opts(schedule=DEVICE))
    assert e.value.status == ggms.EINVAL
    got = gpu_ctx.select(x, [5, 9], "iterative", ggms.default_opts(schedule=DEVICE))
    want = oracle.sort_and_choose(x.cpu().numpy(), np.array([5, 9], dtype=np.uint32))
    assert np.array_equal(_bits(got), _bits(want))


def test_device_schedule_graph_replays(gpu_ctx):
    """ggms_select_device with the device-driven schedule captured once in a
    hipGraph (through torch's graph capture on a side stream) and replayed on
    new values written into the same input buffer; every replay checked
    against the oracle, with eager work and synchronisation between replays."""
    import torch
    import ggms
    lib = ggms._lib.load()
    n, nk = 1 << 23, 1001
    x = ggms.generate("float32", "normal", n, 1)
    ks = ggms.generate_ks("uniform_random", nk, n, 2)
    d_ks = torch.from_numpy(ks.view(np.int32)).cuda()
    d_out = torch.empty(nk, dtype=torch.float32, device="cuda")
    o = ggms.default_opts(schedule=DEVICE)
    assert lib.ggms_ctx_reserve(gpu_ctx.h, 0, n, nk, C.byref(o)) == ggms.OK
    # eager call first (reference result and kernel attributes)
    gpu_ctx.select_device(x, d_ks, d_out, "iterative", o)
    torch.cuda.synchronize()
    assert lib.ggms_ctx_device_status(gpu_ctx.h) == ggms.OK
    s = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s, capture_error_mode="thread_local"):
        oc = ggms.default_opts(schedule=DEVICE, stream=torch.cuda.current_stream().cuda_stream)
        st = lib.ggms_select_device(gpu_ctx.h, 1, 0, C.c_void_p(x.data_ptr()), n, C.c_void_p(d_ks.data_ptr()), nk,
                                    C.c_void_p(d_out.data_ptr()), C.byref(oc))
    assert st == ggms.OK
    for rep in range(4):
        seed = 100 + rep
        assert lib.ggms_generate(0, ggms.DISTS["mixture" if rep % 2 else "normal"], C.c_void_p(x.data_ptr()), n, seed,
                                 0.0, 0.0, None) == 0
        torch.cuda.synchronize()
        d_out.fill_(0)
        g.replay()
        torch.cuda.synchronize()
        assert lib.ggms_ctx_device_status(gpu_ctx.h) == ggms.OK
        want = oracle.sort_and_choose(x.cpu().numpy(), ks, threads=16)
        got = d_out.cpu().numpy()
        assert np.array_equal(_bits(got), _bits(want)), f"replay {rep}"
        # eager work between replays: a host-driven call on the same context
        eager = gpu_ctx.select(x, ks[:101], "iterative")
        assert np.array_equal(_bits(eager), _bits(want[:101]))
    # a capture that would have to allocate is refused, not run: a fresh context
    # reserved for 2^20 values, asked for 2^24 under capture
    small = ggms.Context(0)
    assert lib.ggms_ctx_reserve(small.h, 0, 1 << 20, nk, C.byref(o)) == ggms.OK
    g2 = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g2, stream=s, capture_error_mode="thread_local"):
        big = ggms.default_opts(schedule=DEVICE, stream=torch.cuda.current_stream().cuda_stream)
        st2 = lib.ggms_select_device(small.h, 1, 0, C.c_void_p(x.data_ptr()), 1 << 24, C.c_void_p(d_ks.data_ptr()),
                                     nk, C.c_void_p(d_out.data_ptr()), C.byref(big))
    assert st2 == ggms.ENOMEM
    small.close()
