"""Learned splitter: dataset construction, training on the GPU with the
library's own MLP kernels, and weight management.

Pairs are (sorted-sample window, target rank) -> position of the exact k-th
value between the two sample keys that bracket the target; the exact k-th
values come from the plain method (linear splitter, iterativeSMOS) run on
synthetic problems.  The feature construction mirrors the first-pass kernel's
window (splitter.hpp, learned_pivot) exactly: 64 consecutive sorted sample keys
centred on position r*(S-1), normalised to [0, 1] by the window's extremes,
plus the target's fractional position in the window.

    python -m ggms.learned --out ggms-distributed_amd/weights/splitter_mlp.json
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import time

import numpy as np

from . import _lib

MI, MH = 65, 32
MP = MH * MI + MH + MH + 1
WIN = 64
DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "weights",
                               "splitter_mlp.json")


def _keys(v: np.ndarray) -> np.ndarray:
    if v.dtype == np.float32:
        u = v.view(np.uint32)
        return np.where(u >> 31, ~u, u | np.uint32(0x80000000)).astype(np.uint32)
    if v.dtype == np.float64:
        u = v.view(np.uint64)
        return np.where(u >> 63, ~u, u | np.uint64(0x8000000000000000)).astype(np.uint64)
    if v.dtype == np.int32:
        return (v.view(np.uint32) ^ np.uint32(0x80000000)).astype(np.uint32)
    return v.astype(np.uint32)


def window_features(s: np.ndarray, r: np.ndarray):
    """Features of target rank fractions r over sorted sample keys s (uint32/uint64).

    Returns (X [len(r), 65] float32, ip [len(r)] int64) exactly as the device
    computes them."""
    S = len(s)
    pos = r.astype(np.float64) * float(S - 1)
    ip = pos.astype(np.int64)
    ip = np.clip(ip, 0, S - 2)
    fr = pos - ip
    w0 = np.clip(ip - WIN // 2 + 1, 0, S - WIN)
    idx = w0[:, None] + np.arange(WIN)[None, :]
    win = s[idx]
    klo = win[:, :1]
    span = (win[:, -1:] - klo).astype(np.float64)
    rel = (win - klo).astype(np.float64)
    X = np.zeros((len(r), MI), dtype=np.float32)
    ok = span[:, 0] > 0
    X[ok, :WIN] = (rel[ok] / span[ok]).astype(np.float32)
    X[:, WIN] = (((ip - w0).astype(np.float64) + fr) / float(WIN - 1)).astype(np.float32)
    return X, ip


def build_problem(ctx, dist: str, n: int, seed: int, S: int, pieces: int, dtype: str = "float32",
                  extra_ranks: int = 0, rng=None):
    """One synthetic problem -> (X, y) training pairs, labels from the plain method."""
    import ggms
    x = ggms.generate(dtype, dist, n, seed, *((1.0, 0.01) if dist == "near_constant" else ()))
    rng = rng or np.random.default_rng(seed)
    host = x.cpu().numpy()
    samp = np.sort(_keys(host[rng.integers(0, n, S)]))
    r = np.arange(1, pieces, dtype=np.float64) / pieces
    if extra_ranks:
        r = np.concatenate([r, rng.random(extra_ranks)])
    k_small = np.clip(np.floor(r * (n - 1)).astype(np.int64) + 1, 1, n)
    ks = (n - k_small + 1).astype(np.uint32)
    vals = ctx.select(x, ks, "iterative", ggms.default_opts(splitter="linear"))
    truth = _keys(vals)
    X, ip = window_features(samp, r)
    a = samp[ip].astype(np.float64)
    b = samp[ip + 1].astype(np.float64)
    keep = b > a
    y = np.clip((truth.astype(np.float64) - a) / np.where(keep, b - a, 1.0), 0.0, 1.0).astype(np.float32)
    return X[keep], y[keep]


DISTS = ["mixture", "normal", "lognormal", "half_normal", "uniform", "cauchy"]


def build_dataset(n_problems: int = 96, n: int = 1 << 21, S: int = 8192, pieces: int = 256, seed0: int = 1000,
                  extra_ranks: int = 768, log=print):
    import ggms
    ctx = ggms.Context(0)
    Xs, ys = [], []
    t0 = time.time()
    for p in range(n_problems):
        dist = DISTS[p % len(DISTS)] if p % 2 else "mixture"
        X, y = build_problem(ctx, dist, n, seed0 + p, S, pieces, extra_ranks=extra_ranks,
                             rng=np.random.default_rng(seed0 + 7919 * p))
        Xs.append(X)
        ys.append(y)
        if (p + 1) % 16 == 0:
            log(f"dataset: {p + 1}/{n_problems} problems, {sum(len(v) for v in ys)} pairs, {time.time() - t0:.1f}s")
    return np.concatenate(Xs), np.concatenate(ys)


def init_weights(seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    w = np.zeros(MP, dtype=np.float32)
    w[:MH * MI] = rng.standard_normal(MH * MI) * np.sqrt(2.0 / MI)
    w[MH * MI + MH:MH * MI + 2 * MH] = rng.standard_normal(MH) * np.sqrt(1.0 / MH)
    return w


def train(X: np.ndarray, y: np.ndarray, epochs: int = 40, batch: int = 4096, lr: float = 2e-3, seed: int = 0,
          w0: np.ndarray = None, log=print):
    """Adam on the GPU with the library's forward/backward/update kernels."""
    import torch
    lib = _lib.load()
    dev = "cuda:0"
    n = len(y)
    Xd = torch.from_numpy(np.ascontiguousarray(X)).to(dev)
    yd = torch.from_numpy(np.ascontiguousarray(y)).to(dev)
    w = torch.from_numpy(init_weights(seed) if w0 is None else w0.copy()).to(dev)
    g = torch.zeros(MP, dtype=torch.float32, device=dev)
    m = torch.zeros_like(g)
    v = torch.zeros_like(g)
    loss = torch.zeros(1, dtype=torch.float32, device=dev)
    work = torch.zeros(int(lib.ggms_mlp_workspace_floats(batch)), dtype=torch.float32, device=dev)
    rng = np.random.default_rng(seed)
    step = 0
    P = lambda t: C.c_void_p(t.data_ptr())
    hist = []
    for ep in range(epochs):
        perm = torch.from_numpy(rng.permutation(n)).to(dev)
        tot = 0.0
        nbt = 0
        for i in range(0, n - batch + 1, batch):
            idx = perm[i:i + batch]
            xb = Xd.index_select(0, idx).contiguous()
            yb = yd.index_select(0, idx).contiguous()
            assert lib.ggms_mlp_loss_grad(P(w), P(xb), P(yb), batch, P(g), P(loss), P(work), None) == 0
            step += 1
            assert lib.ggms_mlp_adam_step(P(w), P(g), P(m), P(v), step, lr, 0.9, 0.999, 1e-8, None) == 0
            if i % (batch * 16) == 0:
                tot += float(loss.item())
                nbt += 1
        hist.append(tot / max(nbt, 1))
        if ep % 5 == 0 or ep == epochs - 1:
            log(f"epoch {ep}: loss {hist[-1]:.6f}")
    return w.cpu().numpy(), hist


def predict(w: np.ndarray, X: np.ndarray) -> np.ndarray:
    import torch
    lib = _lib.load()
    wd = torch.from_numpy(np.ascontiguousarray(w, dtype=np.float32)).cuda()
    Xd = torch.from_numpy(np.ascontiguousarray(X, dtype=np.float32)).cuda()
    out = torch.empty(len(X), dtype=torch.float32, device="cuda")
    assert lib.ggms_mlp_forward(C.c_void_p(wd.data_ptr()), C.c_void_p(Xd.data_ptr()), len(X),
                                C.c_void_p(out.data_ptr()), None) == 0
    return out.cpu().numpy()


def save(w: np.ndarray, path: str, meta: dict):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        json.dump({"layout": "W1[32][65], b1[32], W2[32], b2", "params": [float(v) for v in w], "meta": meta}, f)


def load(path: str = DEFAULT_WEIGHTS) -> np.ndarray:
    with open(path) as f:
        d = json.load(f)
    w = np.asarray(d["params"], dtype=np.float32)
    if w.size != MP:
        raise ValueError(f"{path}: expected {MP} parameters, got {w.size}")
    return w


def install(device: int = 0, path: str = DEFAULT_WEIGHTS, w: np.ndarray = None) -> bool:
    """Install weights for the learned splitter of `device` (default file if present)."""
    if w is None:
        if not os.path.exists(path):
            return False
        w = load(path)
    w = np.ascontiguousarray(w, dtype=np.float32)
    st = _lib.load().ggms_set_learned_weights(device, C.c_void_p(w.ctypes.data), w.size)
    if st != 0:
        raise RuntimeError(f"ggms_set_learned_weights failed: {st}")
    return True


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    ap.add_argument("--problems", type=int, default=96)
    ap.add_argument("--log2n", type=int, default=21)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--seed", type=int, default=0)
    args = ap.parse_args()
    t0 = time.time()
    X, y = build_dataset(args.problems, 1 << args.log2n)
    nv = len(y) // 10
    Xv, yv, Xt, yt = X[:nv], y[:nv], X[nv:], y[nv:]
    print(f"pairs: train {len(yt)}, validation {len(yv)} ({time.time() - t0:.1f}s)", flush=True)
    w, hist = train(Xt, yt, epochs=args.epochs, seed=args.seed, log=lambda s: print(s, flush=True))
    pv = predict(w, Xv)
    mse = float(np.mean(0.5 * (pv - yv) ** 2))
    base = float(np.mean(0.5 * (_interp_baseline(Xv) - yv) ** 2))
    print(json.dumps({"val_loss": mse, "val_loss_interpolation": base, "epochs": args.epochs}), flush=True)
    save(w, args.out, {"val_loss": mse, "val_loss_interpolation": base, "pairs": int(len(yt)),
                       "problems": args.problems, "log2n": args.log2n, "epochs": args.epochs,
                       "train_loss": hist})


def _interp_baseline(X):
    """Linear interpolation between the two bracketing sample keys (fractional position)."""
    pos = X[:, 64] * (WIN - 1)
    return (pos - np.floor(pos)).astype(np.float32)


if __name__ == "__main__":
    main()
