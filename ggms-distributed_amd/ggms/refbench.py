"""Reference-format benchmark output: CSV rows, file names and the summary files
of the reference's test harness, so results from this library sit beside the
reference's published data.

Row (DistributedSMOS/runTestsDistributedSMOS.cpp:301-303, :427, :461):

    size,ranks,numKs,distribution,kDistribution,(algorithm,ms,)*errorFlag

(the single-GPU compare harness, DistributedSMOS/compareIterativeSMOS.cu:136-138,
:184, :208, writes the same row without the ranks field).  Times are printed
like a C++ ostream prints a float (6 significant digits).  The error flag
counts the ks where an algorithm's value differs from the first algorithm's.

File name (DistributedSMOS/compareIterativeSMOS.cu:355-359): snprintf into 128
bytes of "%s %s k-dist:%s 2^%d to 2^%d (%d:%d:%d) %d-tests on %s at %s" with the
type, distribution and k-distribution names, the size powers, the k sweep,
the test count, the host name and the asctime-style start time.

Summary files (DistributedSMOS/readDistriSMOSoutputfile.cpp:405-460) for every
result file whose name starts with a base name:
    Summary<base>.csv       size,numKs,(mean,max,min,won,)* tests,errors
    Overview<base>.csv      size,(mean,max,min,won,)* tests,errors
    RatioSummary<base>.csv  size,numKs,(name,mean,max,min,)*   ratio = time / first algorithm's time
    RatioOverview<base>.csv size,(name,mean,max,min,)*
aggregated in single precision as the reference does.  The reference's reader
takes the single-GPU row (no ranks field); this one reads both forms.

Sweeps run the library's engines on seeded synthetic inputs; the first column
("Sort and Choose Multiselect") sorts the whole vector on the GPU and picks,
as the reference's timeSortAndChooseMultiselect does with thrust::sort
(DistributedSMOS/iterativeSMOSTimingFunctions.cu:75-120), and every other
algorithm's values are compared with it for the error flag.

    python -m ggms.refbench sweep --type float --dist normal --kdist uniform_random \\
        --start 21 --stop 26 --startK 101 --jumpK 100 --stopK 1001 --tests 25 --out DIR
    python -m ggms.refbench summarize "BASE" --dir DIR
"""
from __future__ import annotations

import argparse
import os
import socket
import time
from typing import Dict, Iterable, List, Sequence, Tuple

import numpy as np

# reference names (DistributedSMOS/generateProblems.cu:78-79, :444-447, :512-513, :729-730)
TYPE_NAMES = {"float32": "float", "float64": "double", "uint32": "uint"}
DIST_NAMES = {
    "float32": {"uniform": "UNIFORM FLOATS", "normal": "NORMAL FLOATS", "half_normal": "HALF NORMAL FLOATS",
                "cauchy": "CAUCHY FLOAT", "constant": "ALL ONES FLOATS", "lognormal": "LOGNORMAL FLOATS",
                "mixture": "GAUSSIAN MIXTURE FLOATS", "near_constant": "NEAR CONSTANT FLOATS"},
    "float64": {"uniform": "UNIFORM DOUBLES", "normal": "NORMAL DOUBLES", "half_normal": "HALF NORMAL DOUBLES",
                "lognormal": "LOGNORMAL DOUBLES", "mixture": "GAUSSIAN MIXTURE DOUBLES"},
    "uint32": {"uniform": "UNIFORM UNSIGNED INTEGERS", "ascending": "SORTED UINTS", "mod_index": "MOD UINTS"},
}
KDIST_NAMES = {"uniform_random": "Uniform Random Ks", "uniform": "Uniform Ks", "normal": "Normal Random Ks",
               "cluster": "Cluster Ks", "sectioned": "Sectioned Ks", "distinct_sorted": "Distinct Sorted Ks",
               "even": "Even Ks", "window": "Window Ks"}
SORT_AND_CHOOSE = "Sort and Choose Multiselect"
ALG_NAMES = {"bucket": "Bucket Multiselect", "iterative": "IterativeSMOS", "distributed": "DistributedSMOS",
             "distributed_bucket": "DistributedBucketMultiSelect"}


def cfloat(v: float) -> str:
    """A float printed like a default C++ ostream (%g, 6 significant digits)."""
    return f"{float(np.float32(v)):g}"


def row(size: int, num_ks: int, dist_name: str, kdist_name: str, times: Sequence[Tuple[str, float]], errors: int,
        ranks: int = None) -> str:
    """One result line (without the newline)."""
    head = [str(size)] + ([str(ranks)] if ranks is not None else []) + [str(num_ks), dist_name, kdist_name]
    body = []
    for name, ms in times:
        body += [name, cfloat(ms)]
    return ",".join(head + body + [str(errors)])


def file_name(type_name: str, dist_name: str, kdist_name: str, start_pow: int, stop_pow: int, start_k: int,
              jump_k: int, stop_k: int, tests: int, host: str, when: str) -> str:
    """The harness's result file name, cut to 127 characters like its 128-byte snprintf."""
    s = (f"{type_name} {dist_name} k-dist:{kdist_name} 2^{start_pow} to 2^{stop_pow} "
         f"({start_k}:{jump_k}:{stop_k}) {tests}-tests on {host} at {when}")
    return s[:127]


def human_time(t: float = None) -> str:
    """asctime-style local time without the newline ("Wed Jun 12 08:50:28 2024")."""
    return time.asctime(time.localtime(t))


def parse_row(line: str):
    """(size, numKs, dist, kdist, [(name, ms)], errors) from a row of either form."""
    f = line.strip().split(",")
    f = [x for x in f if x != ""]
    # the distributed form carries the rank count as the second field: both it
    # and numKs are numbers, the single-GPU form's third field is a name
    if f[2].strip().isdigit():
        size, num_ks, dist, kd, rest = int(f[0]), int(f[2]), f[3], f[4], f[5:]
    else:
        size, num_ks, dist, kd, rest = int(f[0]), int(f[1]), f[2], f[3], f[4:]
    errors = int(rest[-1])
    pairs = [(rest[i], float(rest[i + 1])) for i in range(0, len(rest) - 1, 2)]
    return size, num_ks, dist, kd, pairs, errors


def _stats(vals: Iterable[float]):
    """mean, max, min of float32 values accumulated in float32 (the reference's float arrays)."""
    total = np.float32(0)
    mx = np.float32(0)
    mn = np.float32(100000)
    n = 0
    for v in vals:
        v = np.float32(v)
        total = np.float32(total + v)
        mx = max(mx, v)
        mn = min(mn, v)
        n += 1
    return np.float32(total / np.float32(n)), mx, mn, n


def summarize(base: str, directory: str = ".") -> List[str]:
    """Writes the four summary files for every result file in `directory` whose
    name starts with `base`; returns their paths."""
    rows = []
    for f in sorted(os.listdir(directory)):
        if not f.startswith(base) or f.startswith(("Summary", "Overview", "Ratio")):
            continue
        with open(os.path.join(directory, f)) as fh:
            for line in fh:
                if line.strip() and line[0].isdigit():
                    rows.append(parse_row(line))
    if not rows:
        raise ValueError(f"no result files start with {base!r} in {directory}")
    names = [n for n, _ in rows[0][4]]
    na = len(names)

    def times(r):
        return [np.float32(t) for _, t in r[4]]

    def ratios(r):
        t = times(r)
        return [np.float64(x) / np.float64(t[0]) for x in t]

    out = {k: [] for k in ("Summary", "Overview", "RatioSummary", "RatioOverview")}
    groups_sk: Dict[Tuple[int, int], list] = {}
    groups_s: Dict[int, list] = {}
    for r in rows:
        groups_sk.setdefault((r[0], r[1]), []).append(r)
        groups_s.setdefault(r[0], []).append(r)

    def timing_fields(grp):
        won = [0] * na
        for r in grp:
            t = times(r)
            won[int(np.argmin(t))] += 1
        f = []
        for j in range(na):
            mean, mx, mn, _ = _stats(times(r)[j] for r in grp)
            f += [cfloat(mean), cfloat(mx), cfloat(mn), str(won[j])]
        return f + [str(len(grp)), str(sum(r[5] for r in grp))]

    def ratio_fields(grp):
        f = []
        for j in range(na):
            v = [ratios(r)[j] for r in grp]
            f += [names[j], f"{np.mean(v):g}", f"{max(v):g}", f"{min(v):g}"]
        return f

    for (size, k), grp in sorted(groups_sk.items()):
        out["Summary"].append(",".join([str(size), str(k)] + timing_fields(grp)))
        out["RatioSummary"].append(",".join([str(size), str(k)] + ratio_fields(grp)) + ",")
    for size, grp in sorted(groups_s.items()):
        out["Overview"].append(",".join([str(size)] + timing_fields(grp)))
        out["RatioOverview"].append(",".join([str(size)] + ratio_fields(grp)) + ",")
    paths = []
    for kind, lines in out.items():
        p = os.path.join(directory, f"{kind}{base}.csv")
        with open(p, "w") as fh:
            fh.write("\n".join(lines) + "\n")
        paths.append(p)
    return paths


def _sort_and_choose_gpu(x, ks: np.ndarray):
    """The reference's sort-and-choose on the GPU: sort ascending, pick sorted[n - k]."""
    import torch
    srt, _ = torch.sort(x)
    idx = torch.from_numpy((x.numel() - ks.astype(np.int64))).to(x.device)
    return srt[idx].cpu().numpy()


def sweep(dtype: str, dist: str, kdist: str, start_pow: int, stop_pow: int, start_k: int, jump_k: int, stop_k: int,
          tests: int, algs: Sequence[str] = ("bucket", "iterative"), out_dir: str = ".", ranks: int = None,
          seed: int = 20240609, log=print) -> str:
    """Runs the harness's loops (sizes 2^start..2^stop, numKs start_k..stop_k by
    jump_k, `tests` seeded inputs each) and appends one row per test to the
    reference-named result file in out_dir; returns its path."""
    import torch
    from . import Context, generate, generate_ks
    ctx = Context(0)
    name = file_name(TYPE_NAMES[dtype], DIST_NAMES[dtype][dist], KDIST_NAMES[kdist], start_pow, stop_pow, start_k,
                     jump_k, stop_k, tests, socket.gethostname(), human_time())
    path = os.path.join(out_dir, name)
    os.makedirs(out_dir, exist_ok=True)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    with open(path, "a") as fh:
        for p in range(start_pow, stop_pow + 1):
            n = 1 << p
            for nk in range(start_k, stop_k + 1, jump_k):
                for t in range(tests):
                    s = seed + 1000003 * p + 1009 * nk + t
                    x = generate(dtype, dist, n, s)
                    ks = generate_ks(kdist, nk, n, s + 1)
                    results = []
                    for a in (SORT_AND_CHOOSE,) + tuple(algs):
                        torch.cuda.synchronize()
                        ev0.record()
                        vals = _sort_and_choose_gpu(x, ks) if a == SORT_AND_CHOOSE else ctx.select(x, ks, a)
                        ev1.record()
                        torch.cuda.synchronize()
                        results.append((a, ev0.elapsed_time(ev1), vals))
                    ref = results[0][2].view(np.uint8).reshape(len(ks), -1)
                    errors = sum(int(np.count_nonzero((v.view(np.uint8).reshape(len(ks), -1) != ref).any(axis=1)))
                                 for _, _, v in results[1:])
                    line = row(n, nk, DIST_NAMES[dtype][dist], KDIST_NAMES[kdist],
                               [(a if a == SORT_AND_CHOOSE else ALG_NAMES[a], ms) for a, ms, _ in results], errors,
                               ranks)
                    fh.write(line + "\n")
                    fh.flush()
                    log(line)
                    del x
    ctx.close()
    return path


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("sweep")
    s.add_argument("--type", default="float", choices=["float", "double", "uint"])
    s.add_argument("--dist", default="uniform")
    s.add_argument("--kdist", default="uniform_random")
    s.add_argument("--start", type=int, default=21)
    s.add_argument("--stop", type=int, default=26)
    s.add_argument("--startK", type=int, default=101)
    s.add_argument("--jumpK", type=int, default=100)
    s.add_argument("--stopK", type=int, default=1001)
    s.add_argument("--tests", type=int, default=25)
    s.add_argument("--algs", default="bucket,iterative")
    s.add_argument("--ranks", type=int, default=None, help="write the distributed row form with this rank count")
    s.add_argument("--out", default=".")
    m = sub.add_parser("summarize")
    m.add_argument("base")
    m.add_argument("--dir", default=".")
    a = ap.parse_args()
    if a.cmd == "sweep":
        dt = {"float": "float32", "double": "float64", "uint": "uint32"}[a.type]
        print(sweep(dt, a.dist, a.kdist, a.start, a.stop, a.startK, a.jumpK, a.stopK, a.tests,
                    tuple(a.algs.split(",")), a.out, a.ranks))
    else:
        for p in summarize(a.base, a.dir):
            print(p)


if __name__ == "__main__":
    main()
