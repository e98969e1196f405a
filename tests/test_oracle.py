"""Pins the CPU oracle (no GPU): published vectors, closed forms, brute force,
library routines (numpy, libm, scipy) and invariants, chosen so that a dropped
term, wrong sign, wrong index or transposed operand fails one of them."""
import json
import math
import os

import numpy as np
import pytest

import oracle

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "golden.json")))


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint64 if a.dtype.itemsize == 8 else np.uint32)


# ---------------------------------------------------------------- Philox
def test_philox_published_kats():
    for v in GOLD["philox4x32_10"]:
        assert list(oracle.philox(v["ctr"], v["key"])) == v["out"]


# ---------------------------------------------------------------- keys
SPECIALS32 = np.array([-np.inf, -3.4028235e38, -1.0, -1.1754944e-38, -1e-45, -0.0, 0.0, 1e-45, 1.1754944e-38,
                       0.5, 1.0, 3.4028235e38, np.inf], dtype=np.float32)


def test_keys_total_order_over_specials():
    k = oracle.keys(SPECIALS32)
    assert np.all(np.diff(k.astype(np.int64)) > 0)
    k64 = oracle.keys(SPECIALS32.astype(np.float64))
    assert np.all(np.diff(k64[1:].astype(object)) != 0)
    assert all(int(a) < int(b) for a, b in zip(k64[:-1], k64[1:]))


def test_keys_agree_with_numpy_sort():
    rng = np.random.default_rng(0)
    x = (rng.standard_normal(100_000) * 1e3).astype(np.float32)
    assert np.array_equal(np.sort(x), x[np.argsort(oracle.keys(x), kind="stable")])
    y = rng.integers(-2**31, 2**31, 50_000, dtype=np.int64).astype(np.int32)
    assert np.array_equal(np.sort(y), y[np.argsort(oracle.keys(y), kind="stable")])


# ---------------------------------------------------------------- selection
@pytest.mark.parametrize("case", range(len(GOLD["small_selections"])))
def test_golden_small_selections(case):
    c = GOLD["small_selections"][case]
    dt = np.float32 if c["dtype"] == "float32" else np.uint32
    x = np.array(c["values"], dtype=dt)
    got = oracle.sort_and_choose(x, c["ks"])
    assert np.array_equal(got, np.array(c["want"], dtype=dt))


def test_known_answer_identity_vector():
    """v[i] = i + 1 => k-th largest is n - k + 1 (DistributedSMOS/iterativeSMOS_copy.cu:2312-2352)."""
    n = 1 << 20
    x = np.arange(1, n + 1, dtype=np.int32)
    ks = np.arange(1, n + 1, 997, dtype=np.uint32)
    assert np.array_equal(oracle.sort_and_choose(x, ks), (n - ks.astype(np.int64) + 1).astype(np.int32))
    assert np.array_equal(oracle.nth_multiselect(x, ks, threads=4), (n - ks.astype(np.int64) + 1).astype(np.int32))


@pytest.mark.parametrize("dt", [np.float32, np.float64, np.uint32, np.int32])
def test_sort_and_choose_vs_numpy(dt):
    rng = np.random.default_rng(1)
    n = 200_001
    if np.dtype(dt).kind == "f":
        x = rng.standard_normal(n).astype(dt)
    else:
        x = rng.integers(np.iinfo(dt).min, np.iinfo(dt).max, n, dtype=np.int64).astype(dt)
    ks = rng.integers(1, n + 1, 333).astype(np.uint32)
    want = np.sort(x)[n - ks.astype(np.int64)]
    for threads in (1, 4):
        assert np.array_equal(bits(oracle.sort_and_choose(x, ks, threads)), bits(want))
        assert np.array_equal(bits(oracle.nth_multiselect(x, ks, threads)), bits(want))


def test_brute_force_tiny_all_k():
    rng = np.random.default_rng(2)
    for n in range(1, 40):
        x = rng.integers(0, 6, n).astype(np.float32) * np.float32(0.5) - np.float32(1.0)
        x[rng.random(n) < 0.2] = -0.0
        ks = np.arange(1, n + 1, dtype=np.uint32)
        srt = sorted(x.tolist(), key=lambda v: (v, math.copysign(1.0, v)))  # -0.0 before +0.0
        want = np.array([srt[n - k] for k in ks], dtype=np.float32)
        assert np.array_equal(bits(oracle.sort_and_choose(x, ks)), bits(want))


def test_constant_extremes_and_errors():
    x = np.full(1000, 2.5, dtype=np.float64)
    assert np.all(oracle.sort_and_choose(x, [1, 500, 1000]) == 2.5)
    y = np.random.default_rng(3).standard_normal(5000).astype(np.float32)
    got = oracle.sort_and_choose(y, [1, 5000])
    assert got[0] == y.max() and got[1] == y.min()
    with pytest.raises(ValueError):
        oracle.sort_and_choose(y, [0])
    with pytest.raises(ValueError):
        oracle.sort_and_choose(y, [5001])
    y[10] = np.nan
    with pytest.raises(ValueError):
        oracle.sort_and_choose(y, [1])


# ---------------------------------------------------------------- deterministic math
def test_detmath_accuracy_against_libm():
    rng = np.random.default_rng(4)
    u = rng.random(200_000) * (1 - 2**-53) + 2**-53
    lg = oracle.detmath("log", u)
    assert np.max(np.abs(lg - np.log(u)) / np.abs(np.log(u))) < 4e-15
    e = rng.uniform(-40, 40, 200_000)
    assert np.max(np.abs(oracle.detmath("exp", e) / np.exp(e) - 1)) < 4e-15
    s = rng.uniform(0, 100, 200_000)
    assert np.max(np.abs(oracle.detmath("sqrt", s) / np.sqrt(s) - 1)) < 4e-16
    sn, cs = oracle.detmath("sincos2pi", rng.random(200_000))
    assert np.max(np.abs(sn ** 2 + cs ** 2 - 1)) < 4e-15
    t = rng.random(200_000)
    sn, cs = oracle.detmath("sincos2pi", t)
    assert np.max(np.abs(sn - np.sin(2 * np.pi * t))) < 2e-15
    assert np.max(np.abs(cs - np.cos(2 * np.pi * t))) < 2e-15


# ---------------------------------------------------------------- generators
def test_generators_distributions():
    from scipy import stats
    n = 1 << 20
    u = oracle.generate(np.float64, "uniform", n, 5)
    assert stats.kstest(u, "uniform").pvalue > 1e-3 and u.min() >= 0 and u.max() < 1
    u32 = oracle.generate(np.float32, "uniform", n, 5)
    assert np.all((u32 * 2**24) == np.floor(u32 * 2**24))  # 24-bit fractions
    z = oracle.generate(np.float64, "normal", n, 6)
    assert stats.kstest(z, "norm").pvalue > 1e-3
    assert abs(z.mean()) < 5 / math.sqrt(n) and abs(z.std() - 1) < 5 / math.sqrt(n)
    h = oracle.generate(np.float64, "half_normal", n, 6)
    assert np.array_equal(h, np.abs(z))
    ln = oracle.generate(np.float64, "lognormal", n, 6)
    assert np.allclose(np.log(ln), z, rtol=1e-13, atol=1e-13)
    nc = oracle.generate(np.float32, "near_constant", n, 7, 1.0, 0.01)
    frac = np.mean(nc != 1.0)
    assert abs(frac - 0.01) < 0.001
    m, mu, sg = oracle.mixture_params(9)
    assert 1 <= m <= 8 and np.all((mu >= -100) & (mu <= 100)) and np.all((sg >= 0.01) & (sg <= 10))
    mix = oracle.generate(np.float64, "mixture", n, 9)
    assert np.isfinite(mix).all()
    r = oracle.generate(np.uint32, "uniform", n, 10, 1024)
    assert r.max() < 1024 and abs(np.bincount(r, minlength=1024).std() / (n / 1024) - math.sqrt(1024 / n)) < 0.01
    assert np.array_equal(oracle.generate(np.uint32, "mod_index", 1000, 0, 100), np.arange(1000) % 100)
    assert np.array_equal(oracle.generate(np.float64, "ascending", 1000, 0), np.arange(1000.0))
    c = oracle.generate(np.float64, "cauchy", n, 11)
    assert stats.kstest(c[np.isfinite(c)], "cauchy").pvalue > 1e-3


def test_generators_seed_and_index_structure():
    a = oracle.generate(np.float64, "normal", 1000, 1)
    b = oracle.generate(np.float64, "normal", 2000, 1)
    assert np.array_equal(a, b[:1000])  # element i depends only on (seed, i)
    assert not np.array_equal(a, oracle.generate(np.float64, "normal", 1000, 2))


# ---------------------------------------------------------------- k generators
def test_k_generators_formulas():
    n, nk = 1 << 20, 1001
    ku = oracle.generate_ks("uniform", nk, n, 0)
    want = [1] + [max(1, min(n, int(np.float32(np.float32(i) / np.float32(nk - 1)) * n))) for i in range(1, nk - 1)] + [n]
    assert list(ku) == want
    ke = oracle.generate_ks("even", 8192, n, 0)
    assert list(ke) == [-(-(i + 1) * n // 8192) for i in range(8192)]
    # generateKUniformRandom (DistributedSMOS/generateProblems.cu:584-604): k = (uint)(u * (float)N)
    # clamped into [1, N], with u a float in [0, 1) -- here the top 24 bits of a Philox4x32-10 word
    # (stream 4, counter i >> 2, word i & 3; key = seed ^ "ggms"), the generator pinned above by the
    # Random123 known-answer vectors
    for n_u, seed in ((n, 7), (1 << 28, 20240610), (1000, 3)):
        kr = oracle.generate_ks("uniform_random", nk, n_u, seed)
        key = [(seed & 0xFFFFFFFF) ^ 0x67676D73, seed >> 32]
        words = [oracle.philox([i, 0, 4, 0], key) for i in range((nk + 3) // 4)]
        u = np.array([np.float32(words[i >> 2][i & 3] >> 8) * np.float32(1.0 / 16777216.0) for i in range(nk)],
                     dtype=np.float32)
        raw = (u * np.float32(n_u)).astype(np.float64).astype(np.int64)
        assert list(kr) == list(np.clip(raw, 1, n_u)), n_u
        assert kr.min() >= 1 and kr.max() <= n_u
    for kd in ("uniform_random", "normal", "cluster", "sectioned", "window"):
        k = oracle.generate_ks(kd, nk, n, 3)
        assert k.min() >= 1 and k.max() <= n, kd
    kn = oracle.generate_ks("normal", 5000, n, 3)
    assert kn.min() >= n // 2  # (|z| + 7) * n / 14 >= n/2
    ks = oracle.generate_ks("sectioned", 100, n, 4)
    assert np.all(np.diff(ks.astype(np.int64)) == 1)
    kd = oracle.generate_ks("distinct_sorted", 101, n, 5)
    assert len(np.unique(kd)) == 101 and np.all(np.diff(kd.astype(np.int64)) > 0)
    kw = oracle.generate_ks("window", 8192, n, 6)
    assert kw.max() - kw.min() < max(8192, n // 100)
    kc = oracle.generate_ks("cluster", 90, n, 7)
    for g in range(10):  # groups of 9 consecutive ranks around a random centre
        grp = kc[9 * g:9 * g + 9].astype(np.int64)
        assert np.all(np.diff(grp) == 1)


# ---------------------------------------------------------------- bucket map
def _range(lo, hi, nb, base=0):
    w = hi - lo
    if w < nb:
        return (lo, hi, base, w + 1, 0, 0)
    bl = w.bit_length()
    sh = bl - 32 if bl > 32 else 0
    wt = w >> sh
    return (lo, hi, base, nb, sh, (nb << 32) // (wt + 1))


def test_bucket_map_brute_force_properties():
    rng = np.random.default_rng(12)
    for trial in range(60):
        lo = int(rng.integers(0, 2**40))
        w = int(rng.choice([0, 1, 5, 63, 64, 1000, 4095, 2**20 + 3]))
        nb = int(rng.choice([1, 2, 3, 7, 64, 512]))
        r = np.array([_range(0, lo - 1, 1, 0), _range(lo, lo + w, nb, 1), _range(lo + w + 1, 2**64 - 1, 1, 1 + min(nb, w + 1))],
                     dtype=oracle.RANGE_DTYPE) if lo > 0 else None
        if r is None:
            continue
        nbr = int(r[1]["nb"])
        keys = np.unique(np.concatenate([np.arange(lo, lo + min(w, 5000) + 1), rng.integers(lo, lo + w + 1, 2000)]))
        b = np.array([oracle.map_bucket(r, int(k)) for k in keys])
        assert np.all(np.diff(b) >= 0)                      # monotone
        assert b.min() >= 1 and b.max() <= nbr              # inside the range's buckets
        if w < 5000:
            assert set(b.tolist()) == set(range(1, 1 + nbr))  # every bucket of a small range is hit
        assert oracle.map_bucket(r, lo - 1) == 0 and oracle.map_bucket(r, lo + w + 1) == 1 + nbr


def test_first_pass_hist_counts_everything():
    x = np.random.default_rng(13).standard_normal(100_000).astype(np.float32)
    r = np.array([_range(0, 2**31 - 1, 4, 0), _range(2**31, 2**32 - 1, 8, 4)], dtype=oracle.RANGE_DTYPE)
    h = oracle.first_pass_hist(x, r, 12)
    assert h.sum() == x.size
    assert h[:4].sum() == np.sum(np.signbit(x))


# ---------------------------------------------------------------- MLP
def test_mlp_gradient_finite_differences():
    rng = np.random.default_rng(14)
    P = 32 * 65 + 32 + 32 + 1
    w = (rng.standard_normal(P) * 0.3).astype(np.float32)
    x = rng.random((16, 65)).astype(np.float32)
    y = rng.random(16).astype(np.float32)
    loss, g = oracle.mlp_loss_grad(w, x, y)
    assert np.allclose(oracle.mlp_forward(w, x), 1 / (1 + np.exp(-_np_forward_logit(w, x))), atol=1e-6)
    idx = rng.choice(P, 60, replace=False)
    for i in idx:
        e = np.zeros(P, dtype=np.float32)
        h = 1e-2
        e[i] = h
        lp, _ = oracle.mlp_loss_grad(w + e, x, y)
        lm, _ = oracle.mlp_loss_grad(w - e, x, y)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) < 2e-4 + 0.05 * abs(fd), (i, fd, g[i])


def _np_forward_logit(w, x):
    W1 = w[:32 * 65].reshape(32, 65).astype(np.float64)
    b1 = w[32 * 65:32 * 65 + 32].astype(np.float64)
    W2 = w[32 * 65 + 32:32 * 65 + 64].astype(np.float64)
    b2 = float(w[-1])
    return np.maximum(x.astype(np.float64) @ W1.T + b1, 0) @ W2 + b2
