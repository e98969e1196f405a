"""GPU parity at the benchmark configurations' full sizes, against the CPU
sort-and-choose oracle (DistributedSMOS/iterativeSMOSTimingFunctions.cu:75-120),
bit-exact:

  * the headline workload: 2^28 fp32 N(0,1) with 1001 uniform random ranks,
    under the linear, sketch and learned splitters;
  * the batched workload: 4096 problems with sizes log-uniform in
    [2^10, 2^22], the committed autotuned class table, replayed from a hipGraph,
    every problem of every replay checked;
  * adversarial inputs at 2^26 (uint32 in [0, 1023]; fp32 |N(0,1)|; fp32 where
    99 % of positions hold 1.0f), each with 8192 evenly spaced ranks and 8192
    ranks clustered in a 1 % window;
  * 2^28 fp32 mixture of Gaussians with lognormal tails, 8192 random ranks;
  * 2^27 fp64 N(0,1) with 1001 random ranks.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint64 if a.dtype.itemsize == 8 else np.uint32)


def _check_all(ctx, x, ks_list, algs, splitters=("linear",), label=""):
    import ggms
    host = x.cpu().numpy()
    ks_all = np.concatenate(ks_list)
    want_all = oracle.sort_and_choose(host, ks_all, threads=16)
    at = 0
    for ks in ks_list:
        want = want_all[at:at + len(ks)]
        at += len(ks)
        for alg in algs:
            for sp in splitters:
                got = ctx.select(x, ks, alg, ggms.default_opts(splitter=sp))
                bad = np.nonzero(_bits(got) != _bits(want))[0]
                assert bad.size == 0, f"{label} {alg}/{sp}: {bad.size} wrong, first k={ks[bad[0]]}"
                print(f"{label} {alg}/{sp}: {len(ks)} ks exact", flush=True)


@pytest.mark.parametrize("dt,dist,p0,p1", [("uint32", "uniform", 1024.0, 0.0), ("float32", "half_normal", 0, 0),
                                           ("float32", "near_constant", 1.0, 0.01)])
def test_adversarial_2p26(gpu_ctx, dt, dist, p0, p1):
    import ggms
    n = 1 << 26
    x = ggms.generate(dt, dist, n, 303, p0, p1)
    ks = [ggms.generate_ks("even", 8192, n, 304), ggms.generate_ks("window", 8192, n, 305)]
    _check_all(gpu_ctx, x, ks, ("iterative", "bucket"), ("linear", "sketch"), f"{dt}/{dist}")


def test_mixture_2p28_8192ks(gpu_ctx):
    import ggms
    n = 1 << 28
    x = ggms.generate("float32", "mixture", n, 505)
    ks = [ggms.generate_ks("uniform_random", 8192, n, 506)]
    _check_all(gpu_ctx, x, ks, ("iterative",), ("linear", "sketch", "learned"), "mixture 2^28")


def test_fp64_normal_2p27_1001ks(gpu_ctx):
    import ggms
    n = 1 << 27
    x = ggms.generate("float64", "normal", n, 207)
    ks = [ggms.generate_ks("uniform_random", 1001, n, 208)]
    _check_all(gpu_ctx, x, ks, ("iterative", "distributed"), ("linear",), "fp64 normal 2^27")


def test_fp64_normal_2p27_8192ks(gpu_ctx):
    """8-byte keys with many ks on a large input: the segmented first pass with its
    large-call defaults (16384 buckets, one tile per CU, 128 quantile pieces)."""
    import ggms
    n = 1 << 27
    x = ggms.generate("float64", "normal", n, 209)
    ks = [ggms.generate_ks("uniform_random", 8192, n, 210)]
    _check_all(gpu_ctx, x, ks, ("iterative", "distributed"), ("linear", "sketch"), "fp64 normal 2^27 8192 ks")


def test_headline_2p28_normal_1001ks(gpu_ctx):
    import ggms
    n = 1 << 28
    x = ggms.generate("float32", "normal", n, 20240609)
    ks = [ggms.generate_ks("uniform_random", 1001, n, 20240610)]
    _check_all(gpu_ctx, x, ks, ("iterative",), ("linear", "sketch", "learned"), "normal 2^28 1001 ks")


def test_batched_4096_problems_autotuned_graph(gpu_ctx):
    """The batched benchmark workload at full size: every problem of two graph
    replays against the oracle."""
    from concurrent.futures import ThreadPoolExecutor
    import torch
    from ggms import autotune
    from ggms.batch import BatchPlan, make_batch
    x, off, ks, koff, sizes = make_batch(4096, 10, 22, 512, seed=20240613)
    classes, bt = autotune.load()
    assert classes is not None, "the committed tuning table is missing"
    plan = BatchPlan(x, off, ks, koff, classes=classes, block_threads=bt, use_graph=True)
    host = x.cpu().numpy()

    def want(p):
        return oracle.sort_and_choose(host[int(off[p]):int(off[p + 1])], ks[koff[p]:koff[p + 1]], threads=1)

    with ThreadPoolExecutor(16) as ex:
        wants = list(ex.map(want, range(len(sizes))))
    print(f"batched: {len(sizes)} problems, {int(off[-1])} values, {len(ks)} ks; oracle done", flush=True)
    for rep in range(2):
        plan.launch()
        torch.cuda.synchronize()
        got = plan.result()
        for p in range(len(sizes)):
            g = got[koff[p]:koff[p + 1]]
            bad = np.nonzero(_bits(g) != _bits(wants[p]))[0]
            assert bad.size == 0, f"replay {rep} problem {p} (n={sizes[p]}): {bad.size} wrong"
        print(f"batched replay {rep}: all {len(sizes)} problems exact", flush=True)
    plan.close()
