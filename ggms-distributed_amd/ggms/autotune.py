"""Autotuner for batched selection: per size class, the first-pass bucket
count, sample size and pieces, plus the workgroup size of the launch.

A size class c holds the problems with ceil(log2(n)) == c.  For every class
the tuner builds a synthetic batch of that class (the batched workload's value
mix: uniform / normal / lognormal(0,1), 1..512 random ks), times each
candidate with the persistent kernel replayed from a hipGraph, cross-checks
the winner's outputs on a few problems against the library's other engine
(the multi-workgroup iterativeSMOS path, one call per problem), and writes the
table as JSON:

    python -m ggms.autotune --out ggms-distributed_amd/weights/batch_tuning.json
"""
from __future__ import annotations

import argparse
import itertools
import json
import os
import time

import numpy as np

from . import batch as B

DEFAULT_TABLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "weights",
                             "batch_tuning.json")
BUCKETS = (1024, 2048, 4096, 8048)
SAMPLES = (1024, 2048)
PIECES = (32, 64)
REBUCKET = (8192,)
THREADS = (512, 1024)
# classes up to 14 (<= 16384 values) are finished in LDS without a bucketing
# pass, so their parameters never matter
FIRST_TUNED_CLASS = 15


def load(path: str = DEFAULT_TABLE):
    """(classes dict, block_threads) from a tuning table; (None, 0) if there is none."""
    if not os.path.exists(path):
        return None, 0
    with open(path) as f:
        d = json.load(f)
    return {int(c): v for c, v in d["classes"].items()}, int(d.get("block_threads", 0))


def _time_plan(plan, reps: int) -> float:
    import torch
    plan.launch()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        plan.launch()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def class_batch(c: int, problems: int, seed: int):
    """Batch whose sizes are log-uniform inside class c: (2^(c-1), 2^c]."""
    lo = max(c - 1, 0) + 1e-9
    return B.make_batch(problems, lo, c, 512, seed)


def check(x, off, ks, koff, got, problems):
    """The batched kernel's answers for `problems` must equal those of the
    iterativeSMOS engine (an independent set of kernels) on the same values."""
    from . import select
    for p in problems:
        a, b = int(off[p]), int(off[p + 1])
        want = select(x[a:b], ks[koff[p]:koff[p + 1]], "iterative")
        if not np.array_equal(got[koff[p]:koff[p + 1]].view(np.uint32), want.view(np.uint32)):
            raise AssertionError(f"autotune winner disagrees with the iterative engine on problem {p}")


def tune(classes=range(10, 23), problems: int = 512, reps: int = 5, seed: int = 77, log=print):
    table = {}
    timings = {}
    for c in classes:
        x, off, ks, koff, sizes = class_batch(c, problems, seed + c)
        best = None
        res = []
        for nb, s, p, rb in itertools.product(BUCKETS, SAMPLES, PIECES, REBUCKET):
            if p + 1 > s:
                continue
            cls = {c: dict(num_buckets=nb, sample_size=s, pieces=p, rebucket=rb)}
            plan = B.BatchPlan(x, off, ks, koff, classes=cls, use_graph=True)
            ms = _time_plan(plan, reps)
            plan.close()
            res.append((ms, nb, s, p, rb))
            if best is None or ms < best[0]:
                best = (ms, nb, s, p, rb)
        ms, nb, s, p, rb = best
        cls = {c: dict(num_buckets=nb, sample_size=s, pieces=p, rebucket=rb)}
        plan = B.BatchPlan(x, off, ks, koff, classes=cls)
        plan.launch()
        check(x, off, ks, koff, plan.result(), range(0, problems, max(problems // 8, 1)))
        plan.close()
        table[c] = dict(num_buckets=nb, sample_size=s, pieces=p, rebucket=rb)
        elems = float(off[-1])
        timings[c] = dict(best_ms=ms, elements=elems, gelem_per_s=elems / ms / 1e6,
                          worst_ms=max(r[0] for r in res))
        log(f"class {c}: best {table[c]} {ms:.3f} ms ({elems / ms / 1e6:.1f} G elem/s), worst {timings[c]['worst_ms']:.3f} ms")
        del x
    return table, timings


def tune_threads(table, problems: int = 4096, seed: int = 5, reps: int = 3, log=print):
    """Workgroup size for the whole batched workload (sizes 2^10..2^22)."""
    x, off, ks, koff, sizes = B.make_batch(problems, 10, 22, 512, seed)
    out = {}
    for t in THREADS:
        plan = B.BatchPlan(x, off, ks, koff, classes=table, block_threads=t, use_graph=True)
        out[t] = _time_plan(plan, reps)
        plan.close()
        log(f"block_threads {t}: {out[t]:.3f} ms")
    return min(out, key=out.get), out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_TABLE)
    ap.add_argument("--problems", type=int, default=512)
    ap.add_argument("--lo", type=int, default=FIRST_TUNED_CLASS)
    ap.add_argument("--hi", type=int, default=22)
    args = ap.parse_args()
    import torch
    t0 = time.time()
    table, timings = tune(range(args.lo, args.hi + 1), args.problems, log=lambda s: print(s, flush=True))
    bt, bt_ms = tune_threads(table, log=lambda s: print(s, flush=True))
    os.makedirs(os.path.dirname(args.out), exist_ok=True)
    with open(args.out, "w") as f:
        json.dump({"device": torch.cuda.get_device_name(0), "block_threads": bt, "workgroups": 0,
                   "classes": {str(c): v for c, v in table.items()},
                   "timings": {str(c): v for c, v in timings.items()},
                   "block_threads_ms": {str(k): v for k, v in bt_ms.items()},
                   "tuning_seconds": time.time() - t0}, f, indent=1)
    print(f"wrote {args.out} ({time.time() - t0:.0f} s)", flush=True)


if __name__ == "__main__":
    main()
