"""Learned splitter: HIP training kernels against the CPU oracle, exactness of
the learned first pass, and what the trained network gains."""
import os

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

MP = 32 * 65 + 32 + 32 + 1


def _bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint64 if a.dtype.itemsize == 8 else np.uint32)


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def test_mlp_kernels_match_oracle(gpu_ctx):
    import ctypes as C
    import torch
    import ggms
    lib = ggms._lib.load()
    P = lambda t: C.c_void_p(t.data_ptr())
    rng = np.random.default_rng(3)
    w = (rng.standard_normal(MP) * 0.3).astype(np.float32)
    for nb in (1, 37, 5000):
        x = rng.random((nb, 65)).astype(np.float32)
        y = rng.random(nb).astype(np.float32)
        wd, xd, yd = _dev(w), _dev(x), _dev(y)
        out = torch.empty(nb, dtype=torch.float32, device="cuda")
        assert lib.ggms_mlp_forward(P(wd), P(xd), nb, P(out), None) == 0
        assert np.allclose(out.cpu().numpy(), oracle.mlp_forward(w, x), atol=2e-6, rtol=1e-5)
        g = torch.empty(MP, dtype=torch.float32, device="cuda")
        loss = torch.empty(1, dtype=torch.float32, device="cuda")
        work = torch.empty(int(lib.ggms_mlp_workspace_floats(nb)), dtype=torch.float32, device="cuda")
        assert lib.ggms_mlp_loss_grad(P(wd), P(xd), P(yd), nb, P(g), P(loss), P(work), None) == 0
        l_ref, g_ref = oracle.mlp_loss_grad(w, x, y)
        assert abs(loss.item() - l_ref) <= 1e-5 + 1e-4 * abs(l_ref)
        gg = g.cpu().numpy()
        scale = np.max(np.abs(g_ref)) + 1e-12
        assert np.max(np.abs(gg - g_ref)) <= 2e-4 * scale, np.max(np.abs(gg - g_ref)) / scale
    # one Adam step against numpy
    gr = rng.standard_normal(MP).astype(np.float32)
    m0 = rng.standard_normal(MP).astype(np.float32) * 0.1
    v0 = rng.random(MP).astype(np.float32) * 0.1
    wd, gd, md, vd = _dev(w), _dev(gr), _dev(m0), _dev(v0)
    assert lib.ggms_mlp_adam_step(P(wd), P(gd), P(md), P(vd), 3, 1e-3, 0.9, 0.999, 1e-8, None) == 0
    m1 = 0.9 * m0 + 0.1 * gr
    v1 = 0.999 * v0 + 0.001 * gr * gr
    w1 = w - 1e-3 * (m1 / (1 - 0.9 ** 3)) / (np.sqrt(v1 / (1 - 0.999 ** 3)) + 1e-8)
    assert np.allclose(wd.cpu().numpy(), w1, atol=1e-6, rtol=1e-5)


def test_learned_features_match_device_window():
    """The dataset's feature window is the one the kernel builds (pure host check of the mirror)."""
    from ggms.learned import window_features
    s = np.sort(np.random.default_rng(1).integers(0, 2**32, 8192, dtype=np.uint64)).astype(np.uint32)
    X, ip = window_features(s, np.array([0.0, 0.5, 0.999999, 1.0 / 256]))
    assert X.shape == (4, 65) and np.all((X[:, :64] >= 0) & (X[:, :64] <= 1))
    assert np.all(X[:, 0] == 0) and np.all(X[:, 63] == 1)
    assert ip[0] == 0 and ip[2] == 8190


@pytest.mark.parametrize("dt,dist", [("float32", "mixture"), ("float64", "normal"), ("uint32", "uniform"),
                                     ("float32", "near_constant")])
def test_learned_splitter_is_exact_with_any_weights(gpu_ctx, dt, dist):
    import ggms
    n = 3_000_017
    x = ggms.generate(dt, dist, n, 5, *((1.0, 0.01) if dist == "near_constant" else ()))
    ks = ggms.generate_ks("uniform_random", 999, n, 6)
    want = oracle.sort_and_choose(x.cpu().numpy(), ks, threads=8)
    for wseed in (1, 2):
        ggms.set_learned_weights((np.random.default_rng(wseed).standard_normal(MP) * 2).astype(np.float32))
        for pieces in (64, 256, 512):
            got = gpu_ctx.select(x, ks, "iterative", ggms.default_opts(splitter="learned", learned_pieces=pieces))
            assert np.array_equal(_bits(got), _bits(want)), (wseed, pieces)
    ggms._learned_installed.discard(0)


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                                    "ggms-distributed_amd", "weights", "splitter_mlp.json")),
                    reason="trained weights not committed yet")
def test_trained_network_gain(gpu_ctx):
    """Mixtures of 1-8 Gaussians + lognormal tails, 2^13 ks: the learned first pass keeps fewer
    elements than the reference linear splitter and than the sketch splitter (same sample and
    pieces, no network: what the trained network itself gains), with identical exact outputs."""
    import ggms
    from ggms import learned
    learned.install(0)
    ggms._learned_installed.add(0)
    n = 1 << 24
    kept = {"linear": [], "sketch": [], "learned": []}
    for seed in (101, 102, 103):
        x = ggms.generate("float32", "mixture", n, seed)
        ks = ggms.generate_ks("uniform_random", 8192, n, seed + 1)
        want = oracle.sort_and_choose(x.cpu().numpy(), ks, threads=8)
        for sp in kept:
            got = gpu_ctx.select(x, ks, "iterative", ggms.default_opts(splitter=sp))
            assert np.array_equal(_bits(got), _bits(want)), sp
            kept[sp].append(gpu_ctx.stats()["n_kept"][0])
    tot = {k: sum(v) for k, v in kept.items()}
    print("first-pass survivors:", tot)
    assert tot["learned"] < tot["linear"]
    # the network's own contribution: same sample and piece count as the sketch,
    # pivots predicted instead of read off the sample (a small margin: ~0.2 %)
    assert tot["learned"] < tot["sketch"]
