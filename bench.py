"""Benchmark: iterativeSMOS on one MI355X.

Workload (the configuration the headline metric is quoted on): N = 2^28 fp32
values ~ Normal(0, 1) from the seeded generator, 1001 k ranks drawn uniformly
from [1, N]; the whole vector is resident in HBM before timing starts.  One
step = one multiselect call (ks in host memory, outputs copied back to host),
as timed by the reference (DistributedSMOS/iterativeSMOSTimingFunctions.cu:142-151).

Prints ONE JSON line.  value = elements/s of the linear-splitter (reference
method) form; the learned-splitter form is reported beside it.

    python bench.py [--gpus 1] [--steps 10] [--warmup 2] [--log2n 28] [--nk 1001]
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ggms-distributed_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

HBM_PEAK_GBS = 8000.0


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` (read + written, scaled by the copy calibration) from the
    committed byte-counter summary of this workload; (None, None) when it is absent."""
    path = os.path.join(ROOT, "profiles", "iterative_2p28_f32_normal_1001ks_hbm_bytes.md")
    if not os.path.exists(path):
        return None, None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 5 and cells[0] == f"`{kernel}`":
            try:
                return int((float(cells[3]) + float(cells[4])) * 2 ** 20), os.path.relpath(path, ROOT)
            except ValueError:
                return None, None
    return None, None


def bench_other_configs(ctx, args):
    """The remaining benchmark inputs, timed like the headline (linear splitter, iterativeSMOS):
    2^26 adversarial inputs with 8192 evenly spaced or clustered ranks, a 2^28 fp32 Gaussian
    mixture with 8192 random ranks, and 2^27 fp64 N(0,1) with 1001 random ranks.  Exactness at
    these sizes is checked by tests/test_gpu_configs.py."""
    import numpy as np
    import torch
    import ggms
    lin = ("linear",)
    cases = [("uint32 U[0,1023]", "uint32", "uniform", 26, 1024.0, 0.0, "even", 8192, lin),
             ("uint32 U[0,1023]", "uint32", "uniform", 26, 1024.0, 0.0, "window", 8192, lin),
             ("fp32 |N(0,1)|", "float32", "half_normal", 26, 0.0, 0.0, "even", 8192, lin),
             ("fp32 99% 1.0f", "float32", "near_constant", 26, 1.0, 0.01, "even", 8192, lin),
             ("fp32 99% 1.0f", "float32", "near_constant", 26, 1.0, 0.01, "window", 8192, lin),
             ("fp32 mixture", "float32", "mixture", 28, 0.0, 0.0, "uniform_random", 8192,
              ("linear", "sketch", "learned")),
             ("fp64 N(0,1)", "float64", "normal", 27, 0.0, 0.0, "uniform_random", 1001, lin)]
    out = []
    for label, dt, dist, lg, p0, p1, kd, nk, splitters in cases:
        n = 1 << lg
        x = ggms.generate(dt, dist, n, args.seed + lg, p0, p1)
        ks = ggms.generate_ks(kd, nk, n, args.seed + nk)
        ref = None
        ms_linear = None
        for sp in splitters:
            o = ggms.default_opts(splitter=sp)
            got = ctx.select(x, ks, "iterative", o)
            ref = got if ref is None else ref
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            reps = 5
            for _ in range(reps):
                ctx.select(x, ks, "iterative", o)
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t0) / reps * 1e3
            ms_linear = ms if sp == "linear" else ms_linear
            ctx.select(x, ks, "iterative", ggms.default_opts(splitter=sp, profile=1))
            st = ctx.stats()
            out.append({"workload": f"{label}, N=2^{lg}, {nk} ks {kd}", "splitter": sp, "ms_per_step": round(ms, 4),
                        "speedup_vs_linear": round(ms_linear / ms, 4) if ms_linear else None,
                        "elements_per_s": round(n / (ms * 1e-3), 1), "passes": st["passes"],
                        "n_kept": st["n_kept"], "buckets": st["buckets"],
                        "identical_to_linear": bool((got.view(np.uint8) == ref.view(np.uint8)).all()),
                        "kernel_ms": {k: round(v, 4) for k, v in st["kernel_ms"].items() if v > 0}})
        del x
    return out


def bench_batched(args):
    """Batched workload: 4096 problems, sizes log-uniform in [2^10, 2^22], each
    uniform / normal / lognormal(0,1) fp32 (seeded per problem), 1..512 random
    ks; one persistent launch per batch, replayed from a hipGraph, per-class
    parameters from the autotuner's table."""
    import numpy as np
    import torch
    import oracle
    from ggms import autotune
    from ggms.batch import BatchPlan, make_batch
    x, off, ks, koff, sizes = make_batch(4096, 10, 22, 512, seed=args.seed + 4)
    classes, bt = autotune.load()
    plan = BatchPlan(x, off, ks, koff, classes=classes, block_threads=bt, use_graph=True)
    for _ in range(max(args.warmup, 1)):
        plan.launch()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    steps = max(args.steps, 1)
    e0.record()
    for _ in range(steps):
        plan.launch()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    got = plan.result()
    host = x.cpu().numpy()
    rng = np.random.default_rng(args.seed)
    sample = sorted(rng.choice(len(sizes), 48, replace=False).tolist() + [int(np.argmax(sizes))])
    exact = True
    for p in sample:
        a, b = int(off[p]), int(off[p + 1])
        want = oracle.sort_and_choose(host[a:b], ks[koff[p]:koff[p + 1]], threads=8)
        exact &= bool(np.array_equal(want.view(np.uint32), got[koff[p]:koff[p + 1]].view(np.uint32)))
    elems = int(off[-1])
    plan.close()
    return {"workload": "4096 problems, sizes log-uniform [2^10, 2^22], fp32 uniform/normal/lognormal(0,1), "
                        "1..512 random ks each; one persistent launch (hipGraph replay)",
            "problems": len(sizes), "elements": elems, "ks": int(koff[-1]), "ms_per_batch": round(ms, 4),
            "value": round(elems / (ms * 1e-3), 1), "unit": "elements/s",
            "input_GBs": round(elems * 4 / (ms * 1e6), 1), "block_threads": bt or 1024,
            "autotuned": classes is not None,
            "exact_vs_sort_and_choose_sample": exact, "checked_problems": len(sample)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--log2n", type=int, default=28)
    ap.add_argument("--nk", type=int, default=1001)
    ap.add_argument("--dist", default="normal")
    ap.add_argument("--seed", type=int, default=20240609)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-check", action="store_true", help="skip the full-size exactness check")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-selection leg")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one plain device copy of the input (known bytes: 4N read, 4N written) so "
                         "hardware byte counters can be scaled")
    ap.add_argument("--no-forms", action="store_true", help="skip the learned and sketch forms")
    ap.add_argument("--no-extra", action="store_true", help="skip the adversarial / mixture / fp64 workloads")
    ap.add_argument("--ref-csv", metavar="DIR", default=None,
                    help="also write reference-format result rows (one per step, every algorithm beside a GPU "
                         "sort-and-choose) and their Summary / Overview / Ratio files into DIR")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: the multi-GPU form (one histogram all-reduce "
                                   "per pass) is described but not built; use --gpus 1"}))
        sys.exit(2)

    import numpy as np
    import torch
    import ggms

    dev = torch.device("cuda:0")
    n = 1 << args.log2n
    x = ggms.generate("float32", args.dist, n, args.seed)
    ks = ggms.generate_ks("uniform_random", args.nk, n, args.seed + 1)
    torch.cuda.synchronize()
    ctx = ggms.Context(0)
    if args.copy_calibration:
        y = torch.empty_like(x)
        y.copy_(x)
        torch.cuda.synchronize()
        del y

    def timed(opts, steps, warmup):
        for _ in range(warmup):
            ctx.select(x, ks, "iterative", opts)
        torch.cuda.synchronize()
        times = []
        t_all0 = time.perf_counter()
        for _ in range(steps):
            t0 = time.perf_counter()
            out = ctx.select(x, ks, "iterative", opts)
            times.append(time.perf_counter() - t0)
        torch.cuda.synchronize()
        t_all = time.perf_counter() - t_all0
        return out, times, t_all

    lin_opts = ggms.default_opts(splitter="linear")
    out_lin, t_lin, tall_lin = timed(lin_opts, args.steps, args.warmup)
    ms_per_step = 1000.0 * tall_lin / args.steps
    value = n / (tall_lin / args.steps)

    # profiled calls: per-kernel durations from HIP events recorded around every
    # launch on the launch stream, averaged over PROFILED calls
    PROFILED = 5
    sums = {}
    for _ in range(PROFILED):
        ctx.select(x, ks, "iterative", ggms.default_opts(splitter="linear", profile=1))
        st = ctx.stats()
        for p in range(st["passes"]):
            for kind, key in (("count", "count_ms"), ("compact", "compact_ms")):
                sums[(kind, p)] = sums.get((kind, p), 0.0) + st[key][p]
        for k, v in st["kernel_ms"].items():
            sums[k] = sums.get(k, 0.0) + v
    bpe = 4
    passes = []
    pass_kernels = []  # (name, algorithmic bytes per launch, average ms)
    for p in range(st["passes"]):
        n_in = st["n_in"][p]
        kept = st["n_kept"][p]
        c_ms = sums[("count", p)] / PROFILED
        k_ms = sums[("compact", p)] / PROFILED
        c_name = "k_count0<float>" if p == 0 else ("k_count1<unsigned int>" if p == 1 else "k_countS<unsigned int>")
        k_name = "k_compact0<float>" if p == 0 else ("k_compact1<unsigned int>" if p == 1 else "k_compactS<unsigned int>")
        pass_kernels.append((c_name, n_in * bpe, c_ms, p))
        pass_kernels.append((k_name, (n_in + kept) * bpe, k_ms, p))
        passes.append({
            "pass": p, "n_in": n_in, "n_kept": kept, "active": st["active"][p], "buckets": st["buckets"][p],
            "count_ms": round(c_ms, 4), "compact_ms": round(k_ms, 4),
            "count_GBs": round(n_in * bpe / (c_ms * 1e6), 1) if c_ms > 0 else None,
            "compact_GBs": round((n_in + kept) * bpe / (k_ms * 1e6), 1) if k_ms > 0 else None,
        })
    # dominant kernel: the longest bucketing-pass launch; algorithmic bytes per
    # launch: count n_p * 4, compaction (n_p + n_{p+1}) * 4
    dom = max(pass_kernels, key=lambda t: t[2])
    d_name, d_bytes, d_ms, _ = dom
    achieved = d_bytes / (d_ms * 1e6)
    traffic, traffic_src = measured_traffic(d_name) if args.log2n == 28 else (None, None)
    roofline = {"bound": "hbm", "kernel": d_name, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "traffic_source": traffic_src,
                "algorithmic_bytes_per_launch": d_bytes, "avg_launch_ms": round(d_ms, 4),
                "timing": f"HIP events around each launch on its stream, mean of {PROFILED} calls",
                "pass_kernels": [{"kernel": nm, "pass": p, "algorithmic_bytes": b, "avg_ms": round(ms, 4),
                                  "GBs": round(b / (ms * 1e6), 1) if ms > 0 else None,
                                  "frac": round(b / (ms * 1e6) / HBM_PEAK_GBS, 4) if ms > 0 else None}
                                 for nm, b, ms, p in pass_kernels]}
    # plain streaming references on the same buffer: a torch sum (reads 4N bytes) and a device copy
    # (reads and writes 4N bytes), timed with events
    probe = {}
    for name, fn, nbytes in (("torch_sum_read", lambda: torch.sum(x), n * bpe),
                             ("device_copy", lambda: y_probe.copy_(x), 2 * n * bpe)):
        y_probe = torch.empty_like(x) if name == "device_copy" else None
        fn()
        torch.cuda.synchronize()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record()
        for _ in range(5):
            fn()
        ev1.record()
        torch.cuda.synchronize()
        probe[name + "_GBs"] = round(nbytes * 5 / (ev0.elapsed_time(ev1) * 1e6), 1)
        del y_probe
    roofline["streaming_probes"] = probe

    result = {
        "metric": "elements/s + HBM GB/s, 2^28 fp32 1001 ks: learned vs linear splitters",
        "value": round(value, 1),
        "unit": "elements/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (Philox4x32-10 seeded generator, Box-Muller normal)",
        "config": {"workload": f"iterativeSMOS, N=2^{args.log2n} fp32 {args.dist}(0,1), {args.nk} ks uniform "
                               f"random in [1,N], linear splitter (the reference's 17 pivots over a "
                               f"{2048 if n >= 1 << 27 and args.nk * 4 < 14336 else 1024}-key sample, the reference "
                               f"draws 1024; 14336 first-pass buckets, the reference uses 8192: see "
                               f"linear_8192_buckets)",
                   "n": n, "nk": args.nk, "seed": args.seed, "alg": "iterative", "splitter": "linear"},
        "roofline": roofline,
        "passes": passes,
        "kernel_ms": {k: round(sums[k] / PROFILED, 4) for k in st["kernel_ms"] if sums[k] > 0},
        "step_ms_median": round(1000 * statistics.median(t_lin), 4),
    }

    # learned-splitter form (same engine, MLP pivots) and its no-network ablation
    for form in () if args.no_forms else ("learned", "sketch"):
        try:
            f_opts = ggms.default_opts(splitter=form)
            out_f, t_f, tall_f = timed(f_opts, args.steps, args.warmup)
            ctx.select(x, ks, "iterative", ggms.default_opts(splitter=form, profile=1))
            sf = ctx.stats()
            result[form] = {"value": round(n / (tall_f / args.steps), 1),
                            "ms_per_step": round(1000.0 * tall_f / args.steps, 4),
                            "speedup_vs_linear": round(tall_lin / tall_f, 4),
                            "identical_outputs": bool(np.array_equal(out_f.view(np.uint32),
                                                                     out_lin.view(np.uint32))),
                            "n_kept": sf["n_kept"], "buckets": sf["buckets"],
                            "kernel_ms": {k: round(v, 4) for k, v in sf["kernel_ms"].items() if v > 0}}
        except ggms.GGMSError as e:
            result[form] = {"value": None, "note": str(e)}

    # the device-driven schedule (every pass enqueued at once, no host read
    # between kernels): as a blocking call, and captured once in a hipGraph
    # (ggms_select_device) and replayed
    if not args.no_forms:
        dev_opts = ggms.default_opts(splitter="linear", schedule=1)
        out_d, t_d, tall_d = timed(dev_opts, args.steps, args.warmup)
        entry = {"value": round(n / (tall_d / args.steps), 1), "ms_per_step": round(1000.0 * tall_d / args.steps, 4),
                 "speedup_vs_host_schedule": round(tall_lin / tall_d, 4),
                 "identical_outputs": bool(np.array_equal(out_d.view(np.uint32), out_lin.view(np.uint32)))}
        try:
            import ctypes as C
            lib = ggms._lib.load()
            d_ks = torch.from_numpy(ks.view(np.int32)).cuda()
            d_out = torch.empty(len(ks), dtype=torch.float32, device="cuda")
            assert lib.ggms_ctx_reserve(ctx.h, 0, n, len(ks), C.byref(dev_opts)) == 0
            ctx.select_device(x, d_ks, d_out, "iterative", dev_opts)
            torch.cuda.synchronize()
            gs = torch.cuda.Stream()
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g, stream=gs, capture_error_mode="thread_local"):
                oc = ggms.default_opts(splitter="linear", schedule=1, stream=torch.cuda.current_stream().cuda_stream)
                rc = lib.ggms_select_device(ctx.h, 1, 0, C.c_void_p(x.data_ptr()), n, C.c_void_p(d_ks.data_ptr()),
                                            len(ks), C.c_void_p(d_out.data_ptr()), C.byref(oc))
            assert rc == 0
            for _ in range(args.warmup):
                g.replay()
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(args.steps):
                g.replay()
            e1.record()
            torch.cuda.synchronize()
            gms = e0.elapsed_time(e1) / args.steps
            entry["graph_replay"] = {"ms_per_step": round(gms, 4), "value": round(n / (gms * 1e-3), 1),
                                     "status_ok": lib.ggms_ctx_device_status(ctx.h) == 0,
                                     "identical_outputs": bool(np.array_equal(d_out.cpu().numpy().view(np.uint32),
                                                                              out_lin.view(np.uint32))),
                                     "note": "outputs stay in device memory (no host copy)"}
            del g
        except Exception as e:  # reported, never fatal for the headline
            entry["graph_replay"] = {"error": repr(e)}
        result["device_schedule"] = entry

    # the same linear splitter at the reference's bucket count
    if not args.no_forms:
        o8 = ggms.default_opts(splitter="linear", num_buckets=8192)
        out8, t8, tall8 = timed(o8, args.steps, args.warmup)
        ctx.select(x, ks, "iterative", ggms.default_opts(splitter="linear", num_buckets=8192, profile=1))
        s8 = ctx.stats()
        result["linear_8192_buckets"] = {"value": round(n / (tall8 / args.steps), 1),
                                         "ms_per_step": round(1000.0 * tall8 / args.steps, 4),
                                         "identical_outputs": bool(np.array_equal(out8.view(np.uint32),
                                                                                  out_lin.view(np.uint32))),
                                         "n_kept": s8["n_kept"], "buckets": s8["buckets"]}

    if not args.no_batch:
        result["batched"] = bench_batched(args)
    if not args.no_extra:
        result["other_configs"] = bench_other_configs(ctx, args)

    if not args.no_check:
        import oracle
        want = oracle.sort_and_choose(x.cpu().numpy(), ks, threads=min(16, os.cpu_count() or 1))
        result["exact_vs_sort_and_choose"] = bool(np.array_equal(want.view(np.uint32), out_lin.view(np.uint32)))

    if not args.no_cpu:
        # CPU sort-and-choose on the same input and ks: order-preserving keys,
        # __gnu_parallel::sort, pick sorted[n - k]; median of 5 after 1 warm-up
        import oracle
        threads = min(16, os.cpu_count() or 1)
        xs = x.cpu().numpy()
        oracle.sort_and_choose(xs, ks, threads=threads)  # warm-up
        tt = []
        for _ in range(5):
            t0 = time.perf_counter()
            oracle.sort_and_choose(xs, ks, threads=threads)
            tt.append(time.perf_counter() - t0)
        t_nth = []
        for _ in range(3):
            t0 = time.perf_counter()
            oracle.nth_multiselect(xs, ks, threads=threads)
            t_nth.append(time.perf_counter() - t0)
        result["cpu_baseline"] = {
            "value": round(n / statistics.median(tt), 1), "unit": "elements/s", "cores": threads,
            "kind": "port",
            "sample": f"the same 2^{args.log2n} input and {args.nk} ks; order-preserving keys + "
                      f"__gnu_parallel::sort + pick; median of 5 after 1 warm-up",
            "ms_per_call": round(1000 * statistics.median(tt), 1),
            "nth_element_multiselect_value": round(n / statistics.median(t_nth), 1),
            "host_cpus": os.cpu_count(),
        }
        del xs
    if args.ref_csv:
        from ggms import refbench
        path = refbench.sweep("float32", args.dist, "uniform_random", args.log2n, args.log2n, args.nk, 1, args.nk,
                              args.steps, ("bucket", "iterative", "distributed", "distributed_bucket"),
                              args.ref_csv, seed=args.seed, log=lambda s: print(s, file=sys.stderr, flush=True))
        base = os.path.basename(path)
        result["ref_csv"] = {"file": base, "summaries": [os.path.basename(q) for q in
                                                         refbench.summarize(base, args.ref_csv)]}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
