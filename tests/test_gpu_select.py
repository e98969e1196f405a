"""GPU parity: every selection entry point against the CPU sort-and-choose oracle.

Bit-exact comparison (value bits, so -0.0 vs +0.0 counts).  Mirrors the
reference's compare harness (DistributedSMOS/compareIterativeSMOS.cu:186-210),
which flags any output that differs from sort-and-choose.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

ALGS = ["bucket", "iterative", "distributed", "distributed_bucket"]
NP = {"float32": np.float32, "float64": np.float64, "uint32": np.uint32, "int32": np.int32}


def _host(x):
    return x.cpu().numpy()


def _bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint64 if a.dtype.itemsize == 8 else np.uint32)


def _check(ctx, x, ks, alg, opts=None, label=""):
    got = ctx.select(x, ks, alg, opts)
    want = oracle.sort_and_choose(_host(x), ks, threads=8)
    bad = np.nonzero(_bits(got) != _bits(want))[0]
    assert bad.size == 0, (f"{label} alg={alg}: {bad.size} of {len(ks)} wrong; first k={ks[bad[0]]} "
                           f"got {got[bad[0]]!r} want {want[bad[0]]!r}")
    return got


def test_generators_match_oracle_bits(gpu_ctx):
    import ggms
    n = 200_003
    cases = [("float32", "uniform", 0, 0), ("float64", "uniform", 0, 0), ("float32", "normal", 0, 0),
             ("float64", "normal", 0, 0), ("float32", "half_normal", 0, 0), ("float32", "lognormal", 0, 0),
             ("float32", "mixture", 0, 0), ("float64", "mixture", 0, 0), ("float32", "near_constant", 1.0, 0.01),
             ("float32", "cauchy", 0, 0), ("uint32", "uniform", 1024, 0), ("uint32", "uniform", 0, 0),
             ("int32", "uniform", 0, 0), ("uint32", "mod_index", 100, 0), ("float64", "ascending", 0, 0)]
    for dt, dist, p0, p1 in cases:
        for seed in (1, 0xDEADBEEF12345):
            g = _host(ggms.generate(dt, dist, n, seed, p0, p1))
            c = oracle.generate(NP[dt], dist, n, seed, p0, p1)
            assert np.array_equal(_bits(g), _bits(c)), f"{dt} {dist} seed={seed}"


def test_first_pass_histogram_matches_oracle(gpu_ctx):
    import ggms
    for dt, dist, p0, p1 in [("float32", "normal", 0, 0), ("float64", "uniform", 0, 0),
                             ("uint32", "uniform", 1024, 0), ("float32", "near_constant", 1.0, 0.01),
                             ("int32", "uniform", 0, 0)]:
        x = ggms.generate(dt, dist, (1 << 20) + 7, 11, p0, p1)
        for splitter, alg_opts in [("linear", {}), ("sketch", {"learned_pieces": 64})]:
            o = ggms.default_opts(splitter=splitter, **alg_opts)
            ranges, hist = gpu_ctx.first_pass(x, o)
            # ranges partition the key space in order
            assert ranges["lo"][0] == 0
            assert np.all(ranges["lo"][1:] == ranges["hi"][:-1] + 1)
            kmax = np.iinfo(np.uint64 if dt == "float64" else np.uint32).max
            assert int(ranges["hi"][-1]) == kmax
            assert np.all(ranges["base"][1:] == ranges["base"][:-1] + ranges["nb"][:-1])
            want = oracle.first_pass_hist(_host(x), ranges, len(hist))
            assert np.array_equal(hist, want), f"{dt} {dist} {splitter}"
            assert int(hist.sum()) == x.numel()


@pytest.mark.parametrize("alg", ALGS)
@pytest.mark.parametrize("dt", ["float32", "float64", "uint32", "int32"])
def test_small_and_ragged(gpu_ctx, alg, dt):
    import ggms
    rng = np.random.default_rng(5)
    for n in (1, 2, 3, 63, 64, 65, 1000, 8191, 16384, 16385, 65537, 300_001):
        x = ggms.generate(dt, "normal" if dt.startswith("float") else "uniform", n, n * 7 + 1)
        ks = rng.integers(1, n + 1, size=min(37, 2 * n), dtype=np.uint32)
        ks[0] = 1
        ks[-1] = n
        _check(gpu_ctx, x, ks, alg, label=f"{dt} n={n}")


def test_config1_bucket_uniform_2p20(gpu_ctx):
    """N = 2^20 fp32 U[0,1), 101 distinct sorted ranks, bucketMultiselect."""
    import ggms
    n = 1 << 20
    x = ggms.generate("float32", "uniform", n, 2024)
    ks = ggms.generate_ks("distinct_sorted", 101, n, 2024)
    assert len(np.unique(ks)) == 101
    for alg in ALGS:
        _check(gpu_ctx, x, ks, alg, label="config1")


def test_known_answers(gpu_ctx):
    """v[i] = i + 1: the k-th largest is n - k + 1 (DistributedSMOS/iterativeSMOS_copy.cu:2312-2360)."""
    import torch
    for n in (1 << 20, 3_000_017):
        x = torch.arange(1, n + 1, dtype=torch.int32, device="cuda")
        ks = np.array([i * 100 + 1 for i in range(1001) if i * 100 + 1 <= n], dtype=np.uint32)
        for alg in ALGS:
            got = gpu_ctx.select(x, ks, alg)
            assert np.array_equal(got, (n - ks.astype(np.int64) + 1).astype(np.int32)), alg


def test_flat_and_extremes(gpu_ctx):
    import ggms
    import torch
    n = 1 << 20
    x = ggms.generate("float32", "constant", n, 0, 3.5)
    ks = np.array([1, 2, n // 2, n - 1, n], dtype=np.uint32)
    for alg in ALGS:
        assert np.all(gpu_ctx.select(x, ks, alg) == np.float32(3.5))
    x = ggms.generate("float32", "normal", n, 3)
    h = _host(x)
    got = gpu_ctx.select(x, [1, n], "iterative")
    assert got[0] == h.max() and got[1] == h.min()
    # signed zeros and infinities order as keys: -inf < -0.0 < +0.0 < +inf
    v = np.zeros(50_000, dtype=np.float32)
    v[::3] = -0.0
    v[7] = np.inf
    v[8] = -np.inf
    xt = torch.from_numpy(v).cuda()
    ks = np.arange(1, v.size + 1, 997, dtype=np.uint32)
    for alg in ALGS:
        _check(gpu_ctx, xt, ks, alg, label="signed zeros")


def test_errors(gpu_ctx):
    import ggms
    import torch
    x = ggms.generate("float32", "normal", 100_000, 1)
    with pytest.raises(ggms.GGMSError) as e:
        gpu_ctx.select(x, [0], "iterative")
    assert e.value.status == ggms.EINVAL
    with pytest.raises(ggms.GGMSError) as e:
        gpu_ctx.select(x, [100_001], "iterative")
    assert e.value.status == ggms.EINVAL
    v = np.random.default_rng(0).standard_normal(100_000).astype(np.float32)
    v[1234] = np.nan
    with pytest.raises(ggms.GGMSError) as e:
        gpu_ctx.select(torch.from_numpy(v).cuda(), [5], "iterative")
    assert e.value.status == ggms.ENAN


@pytest.mark.parametrize("nk", [8192, 8193, 20000])
def test_setup_paths(gpu_ctx, nk):
    """The setup step runs inside the first-pass planning kernel up to 8192 ks and as its own
    kernel above: both answer exactly and both reject a k outside [1, n] (last position)."""
    import ggms
    n = 1 << 20
    x = ggms.generate("float32", "normal", n, 77)
    ks = ggms.generate_ks("uniform_random", nk, n, 78)
    for alg in ("iterative", "bucket"):
        _check(gpu_ctx, x, ks, alg, label=f"nk={nk}")
    bad = ks.copy()
    bad[-1] = n + 1
    with pytest.raises(ggms.GGMSError) as e:
        gpu_ctx.select(x, bad, "iterative")
    assert e.value.status == ggms.EINVAL
    _check(gpu_ctx, x, ks, "iterative", label=f"nk={nk} after EINVAL")


def test_ulp_neighbours_regression(gpu_ctx):
    """Floats one ULP apart must be separated (DistributedSMOS/pivots and slopes:3-12 traces the
    reference failing to separate 1.101046334952116e-02 from 1.101046241819858e-02)."""
    import torch
    base = np.array([1.101046334952116012573242187500e-02, 1.101046241819858551025390625000e-02,
                     1.183493062853813171386718750000e-02, 1.183493155986070632934570312500e-02], dtype=np.float32)
    rng = np.random.default_rng(9)
    v = rng.random(1 << 21).astype(np.float32) * 0.0185
    pos = rng.choice(v.size, 4000, replace=False)
    v[pos] = np.tile(base, 1000)
    xt = torch.from_numpy(v).cuda()
    ks = rng.integers(1, v.size + 1, 2000, dtype=np.uint32)
    srt = np.sort(v)
    extra = [v.size - np.searchsorted(srt, b) for b in base]
    ks = np.concatenate([ks, np.array(extra, dtype=np.uint32)])
    for alg in ALGS:
        _check(gpu_ctx, xt, ks, alg, label="ulp")


@pytest.mark.parametrize("dt,dist,p0,p1", [("float32", "normal", 0, 0), ("float64", "normal", 0, 0),
                                           ("uint32", "uniform", 1024, 0), ("float32", "half_normal", 0, 0),
                                           ("float32", "near_constant", 1.0, 0.01), ("float32", "mixture", 0, 0),
                                           ("float32", "cauchy", 0, 0), ("int32", "uniform", 0, 0)])
@pytest.mark.parametrize("kdist,nk", [("uniform_random", 1001), ("even", 8192), ("window", 8192),
                                      ("cluster", 500), ("normal", 300)])
def test_medium_distributions(gpu_ctx, dt, dist, p0, p1, kdist, nk):
    import ggms
    n = (1 << 22) + 13
    x = ggms.generate(dt, dist, n, 77, p0, p1)
    ks = ggms.generate_ks(kdist, nk, n, 78)
    _check(gpu_ctx, x, ks, "iterative", label=f"{dt}/{dist}/{kdist}")


def test_reference_wrappers(gpu_ctx):
    import ggms
    n = 1 << 20
    x = ggms.generate("float64", "normal", n, 5)
    ks = ggms.generate_ks("uniform_random", 101, n, 6)
    want = oracle.sort_and_choose(_host(x), ks)
    for fn in (ggms.iterativeSMOSWrapper, ggms.bucketMultiselectWrapper, ggms.distributedSMOSWrapper,
               ggms.distributedBucketMultiselectWrapper):
        out = np.zeros(len(ks), dtype=np.float64)
        assert fn(x, n, ks, len(ks), out, 256, 1024) == 0
        assert np.array_equal(_bits(out), _bits(want)), fn.__name__
    # input untouched
    assert np.array_equal(_bits(_host(x)), _bits(_host(ggms.generate("float64", "normal", n, 5))))


def test_device_entry_point(gpu_ctx):
    import ggms
    import torch
    n = 5_000_000
    x = ggms.generate("float32", "lognormal", n, 8)
    ks = ggms.generate_ks("uniform_random", 777, n, 9)
    d_ks = torch.from_numpy(ks.astype(np.int32)).cuda().view(torch.uint32) if hasattr(torch, "uint32") else None
    d_out = torch.empty(len(ks), dtype=torch.float32, device="cuda")
    gpu_ctx.select_device(x, d_ks, d_out, "iterative")
    want = oracle.sort_and_choose(_host(x), ks, threads=8)
    assert np.array_equal(_bits(d_out.cpu().numpy()), _bits(want))


def test_seed_independence(gpu_ctx):
    """Results do not depend on the sample (the splitter only changes the work)."""
    import ggms
    n = 3_000_000
    x = ggms.generate("float32", "mixture", n, 21)
    ks = ggms.generate_ks("uniform_random", 257, n, 22)
    want = oracle.sort_and_choose(_host(x), ks, threads=8)
    for seed in (1, 2, 3, 12345):
        for pieces, splitter in ((17, "linear"), (64, "sketch")):
            o = ggms.default_opts(seed=seed, splitter=splitter, num_pivots=pieces, learned_pieces=pieces)
            got = gpu_ctx.select(x, ks, "iterative", o)
            assert np.array_equal(_bits(got), _bits(want)), (seed, splitter)
    for b in (16, 64, 512, 4096, 16384):
        got = gpu_ctx.select(x, ks, "iterative", ggms.default_opts(num_buckets=b))
        assert np.array_equal(_bits(got), _bits(want)), b


@pytest.mark.parametrize("dt,dist", [("uint32", "uniform"), ("float32", "normal"), ("float64", "lognormal")])
def test_many_passes(gpu_ctx, dt, dist):
    """Tiny bucket budgets force long re-bucketing chains (every pass-to-pass hand-off is exercised)."""
    import ggms
    n = 8_000_003  # > 16 x 16 x 32768: at least three passes with 16 buckets
    x = ggms.generate(dt, dist, n, 31)
    ks = ggms.generate_ks("uniform_random", 3000, n, 32)
    want = oracle.sort_and_choose(_host(x), ks, threads=8)
    for b in (16, 24, 40):
        got = gpu_ctx.select(x, ks, "iterative", ggms.default_opts(num_buckets=b))
        st = gpu_ctx.stats()
        assert np.array_equal(_bits(got), _bits(want)), (b, st["passes"])
        assert st["passes"] >= 3


@pytest.mark.parametrize("dt", ["uint32", "float64"])
def test_finish_layouts(gpu_ctx, dt):
    """The finish holds a segment as keys (at most half its capacity), as
    half-width offsets from the segment's low key (narrow key interval), or, in a
    second launch, as keys at full capacity (wide and large).  Sixteen first-pass
    buckets over uniform values give segments of ~n/16 keys with wide intervals;
    the tiny path puts everything in one whole-range segment."""
    import ggms
    cap = 32768 if dt == "uint32" else 16384
    for n, dist, b in ((cap // 2, "uniform", 0), (cap // 2 + 1, "uniform", 0), (cap, "uniform", 0),
                       (int(cap * 9.5), "uniform", 16), (int(cap * 9.5), "normal", 16), (3 * cap, "uniform", 0)):
        x = ggms.generate(dt, dist, n, 41 + n)
        ks = ggms.generate_ks("uniform_random", 40, n, 42 + n)
        want = oracle.sort_and_choose(_host(x), ks, threads=8)
        got = gpu_ctx.select(x, ks, "iterative", ggms.default_opts(num_buckets=b) if b else None)
        assert np.array_equal(_bits(got), _bits(want)), (n, dist, b)


def test_length_cap_rejected_before_any_device_work(gpu_ctx):
    """Lengths above the reference's int range (2^31 - 1) are refused with EINVAL
    by every single-vector entry point, before anything touches the vector."""
    import ctypes as C
    import ggms
    import torch
    lib = ggms._lib.load()
    x = torch.zeros(16, dtype=torch.float32, device="cuda")
    ks = np.array([1], dtype=np.uint32)
    out = np.zeros(1, dtype=np.float32)
    for n in (1 << 31, (1 << 32) - 1, 1 << 33):
        st = lib.ggms_select(gpu_ctx.h, 1, 0, C.c_void_p(x.data_ptr()), n, ks.ctypes.data_as(ggms._lib._U32P), 1,
                             out.ctypes.data_as(C.c_void_p), None)
        assert st == ggms.EINVAL, n
    with pytest.raises(ggms.GGMSError):
        gpu_ctx.select(x, [1], "iterative", n=1 << 31)


def test_reference_shim_program(gpu_ctx):
    """A C++ program written against include/ggms.hpp (the reference's template
    call shape) links libggms.so and runs on the GPU: (T)1 and the known answers
    v[i] = i + 1 for all three wrappers and four types, (T)0 and
    ggms_last_wrapper_status() == GGMS_EINVAL for a bad k."""
    import os
    import shutil
    import subprocess
    import tempfile
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    libdir = os.path.join(root, "ggms-distributed_amd", "lib")
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "shim_main")
        r = subprocess.run([gxx, "-std=c++17", "-O1", "-D__HIP_PLATFORM_AMD__", "-I", "/opt/rocm/include",
                            "-I", os.path.join(root, "include"), os.path.join(root, "tests", "shim", "shim_main.cpp"),
                            "-o", exe, "-L", libdir, "-lggms", "-L", "/opt/rocm/lib", "-lamdhip64",
                            f"-Wl,-rpath,{libdir}", "-Wl,-rpath,/opt/rocm/lib"], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-3000:]
        r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
        print(r.stdout)
        assert r.returncode == 0 and "shim ok" in r.stdout, r.stdout + r.stderr


@pytest.mark.parametrize("K", [1, 1024, 65536])
def test_topk(gpu_ctx, K):
    """Top-k select (the reference's timeBucketTopkselect feeds ks = 1..K,
    DistributedSMOS/iterativeSMOSTimingFunctions.cu:306-333): the K largest values,
    largest first, through the iterative engine and the one-pass bucketMultiselect
    (whose top segment is k-dense: it is sorted from global memory)."""
    import ggms
    n = 1 << 26
    x = ggms.generate("float32", "normal", n, 4242)
    want = oracle.sort_and_choose(_host(x), np.arange(1, K + 1, dtype=np.uint32), threads=16)
    assert np.all(np.diff(want.astype(np.float64)) <= 0)
    for alg in ("iterative", "bucket"):
        got = gpu_ctx.topk(x, K, alg)
        bad = np.nonzero(_bits(got) != _bits(want))[0]
        assert bad.size == 0, f"{alg} K={K}: {bad.size} wrong, first at {bad[0]}"
    assert gpu_ctx.topk(x, 0).size == 0
    with pytest.raises(ggms.GGMSError):
        gpu_ctx.topk(x[:100], 101)


@pytest.mark.parametrize("dt,dist,kdist,nk", [("float32", "normal", "window", 8192), ("float64", "normal", "even", 4000),
                                              ("uint32", "uniform", "uniform_random", 3000),
                                              ("float32", "half_normal", "window", 8192)])
def test_oversized_segments_finish_from_global_memory(gpu_ctx, dt, dist, kdist, nk):
    """Segments larger than the LDS finish capacity: the one-pass bucketMultiselect
    finishes every k-bucket whatever its size, and an iterative run capped at two
    passes force-finishes its pass-1 k-buckets; both go through the global-memory
    finish (radix select for <= 8 ks, a stable radix sort for k-dense segments)."""
    import ggms
    n = (1 << 24) + 77
    x = ggms.generate(dt, dist, n, 99)
    ks = ggms.generate_ks(kdist, nk, n, 100)
    want = oracle.sort_and_choose(_host(x), ks, threads=16)
    for alg, o in (("bucket", ggms.default_opts(num_buckets=256)), ("distributed_bucket", ggms.default_opts(num_buckets=512)),
                   ("iterative", ggms.default_opts(num_buckets=64, max_passes=2))):
        got = gpu_ctx.select(x, ks, alg, o)
        bad = np.nonzero(_bits(got) != _bits(want))[0]
        assert bad.size == 0, f"{alg}: {bad.size} wrong, first k={ks[bad[0]]}"
        print(f"{dt}/{dist}/{kdist} {alg}: exact", flush=True)


@pytest.mark.parametrize("dt,dist,p0", [("float32", "mixture", 0), ("float64", "normal", 0), ("uint32", "uniform", 1024),
                                        ("float32", "near_constant", 1.0), ("int32", "uniform", 0)])
def test_first_pass_layouts(gpu_ctx, dt, dist, p0):
    """The iterative forms with each first-pass layout: MERGED (every k-bucket
    re-bucketed together in pass 2) and SEGMENTED (every k-bucket compacted into
    a segment of its own by the two-level grouped compaction; those within the
    finish capacity finished at once, the larger ones re-bucketed segment by
    segment), with few and many ks, small bucket budgets (long re-bucketing
    chains below the first pass) and a two-pass cap (forced finish)."""
    import ggms
    n = (1 << 23) + 5
    x = ggms.generate(dt, dist, n, 41, p0, 0.01 if dist == "near_constant" else 0)
    for kdist, nk in (("uniform_random", 1001), ("even", 8192), ("window", 5000)):
        ks = ggms.generate_ks(kdist, nk, n, 42 + nk)
        want = oracle.sort_and_choose(_host(x), ks, threads=16)
        for alg in ("iterative", "distributed"):
            for fp, extra in ((1, {}), (2, {}), (2, {"num_buckets": 64}), (2, {"num_buckets": 512, "max_passes": 2})):
                got = gpu_ctx.select(x, ks, alg, ggms.default_opts(first_pass=fp, **extra))
                bad = np.nonzero(_bits(got) != _bits(want))[0]
                assert bad.size == 0, f"{kdist}/{nk} {alg} first_pass={fp} {extra}: {bad.size} wrong, first k={ks[bad[0]]}"
        print(f"{dt}/{dist}/{kdist}: exact", flush=True)


@pytest.mark.parametrize("dt", ["float32", "float64", "int32"])
def test_dominant_value_counted_in_registers(gpu_ctx, dt):
    """A value holding ~40 % of the keys beside values of both signs (so the
    first-pass cell table covers the whole key space, no window): the planning
    kernel reports its one-key range's bucket and the first-pass count tallies
    that bucket in registers.  Both first-pass layouts, few and many ks, every
    answer against sort-and-choose (including ranks inside the dominant run)."""
    import ggms
    import torch
    rng = np.random.default_rng(77)
    n = (1 << 22) + 3
    if dt == "int32":
        v = rng.integers(-(1 << 30), 1 << 30, n, dtype=np.int64).astype(np.int32)
        v[rng.random(n) < 0.4] = -12345
    else:
        v = rng.standard_normal(n).astype(NP[dt])
        v[rng.random(n) < 0.4] = NP[dt](0.75)
    x = torch.from_numpy(v).cuda()
    for kdist, nk in (("uniform_random", 1001), ("even", 8192)):
        ks = ggms.generate_ks(kdist, nk, n, 5 + nk)
        want = oracle.sort_and_choose(v, ks, threads=16)
        for fp in (1, 2):
            for sp in ("linear", "sketch"):
                got = gpu_ctx.select(x, ks, "iterative", ggms.default_opts(first_pass=fp, splitter=sp))
                bad = np.nonzero(_bits(got) != _bits(want))[0]
                assert bad.size == 0, f"{dt} {kdist}/{nk} first_pass={fp} {sp}: {bad.size} wrong, first k={ks[bad[0]]}"
    print(f"{dt}: exact", flush=True)
