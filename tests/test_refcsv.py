"""The reference-format benchmark writer (ggms.refbench) against the
reference's own published result files (tests/golden/refcsv.json, made by
tests/golden/make_refcsv_fixture.py from "DistributedSMOS Test Result"):
file names, row format, and the summary files of
DistributedSMOS/readDistriSMOSoutputfile.cpp:405-460.  No GPU."""
import json
import os
import re

import numpy as np

from ggms import refbench as R

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = json.load(open(os.path.join(HERE, "golden", "refcsv.json")))


def test_file_names_match_published():
    cases = [
        (("float", "NORMAL FLOATS", "Uniform Ks", 21, 26, 101, 100, 1001, 25, "zermelo.cs.grinnell.",
          "Wed Jun 12 08:50:28 2024"),
         "float NORMAL FLOATS k-dist:Uniform Ks 2^21 to 2^26 (101:100:1001) 25-tests on zermelo.cs.grinnell. at "
         "Wed Jun 12 08:50:28 2024"),
        # 128-byte snprintf: the name is cut at 127 characters
        (("uint", "UNIFORM UNSIGNED INTEGERS", "Uniform Ks", 17, 26, 101, 900, 1001, 25, "delamain.cs.grinnell",
          "Sat Jun  8 20:00:00 2024"),
         "uint UNIFORM UNSIGNED INTEGERS k-dist:Uniform Ks 2^17 to 2^26 (101:900:1001) 25-tests on "
         "delamain.cs.grinnell at Sat Jun  8 20:"),
        (("double", "UNIFORM DOUBLES", "Uniform Random Ks", 21, 26, 101, 100, 501, 25, "zermelo.cs.grinnell.",
          "Mon Jun 10 22:09:11 2024"),
         "double UNIFORM DOUBLES k-dist:Uniform Random Ks 2^21 to 2^26 (101:100:501) 25-tests on zermelo.cs.grinnell. "
         "at Mon Jun 10 22:09"),
    ]
    for args, want in cases:
        assert R.file_name(*args) == want
        assert want in FIX["names"]
    pat = re.compile(r"^(float|double|uint) .+ k-dist:.+ 2\^\d+ to 2\^\d+ \(\d+:\d+:\d+\) \d+-tests on ")
    assert all(pat.match(n) and len(n) <= 127 for n in FIX["names"])
    assert len(R.human_time(0)) == 24


def test_rows_round_trip_published_lines():
    """Every published line parses and is written back character for character
    (the distributed form with the rank count, C++ float printing)."""
    n = 0
    for lines in FIX["rows"].values():
        for line in lines:
            size, nk, dist, kd, pairs, err = R.parse_row(line)
            ranks = int(line.split(",")[1])
            assert R.row(size, nk, dist, kd, pairs, err, ranks=ranks) == line
            n += 1
    assert n == 180
    # the single-GPU form (no ranks field) parses the same
    one = R.row(1 << 20, 101, "UNIFORM FLOATS", "Uniform Random Ks",
                [("Sort and Choose Multiselect", 3.25), ("IterativeSMOS", 0.125)], 0)
    assert one == "1048576,101,UNIFORM FLOATS,Uniform Random Ks,Sort and Choose Multiselect,3.25,IterativeSMOS,0.125,0"
    assert R.parse_row(one)[:2] == (1 << 20, 101)


def test_summaries_of_published_rows(tmp_path):
    name, lines = next(iter(FIX["rows"].items()))
    (tmp_path / name).write_text("\n".join(lines) + "\n")
    base = name[:40]
    paths = R.summarize(base, str(tmp_path))
    assert [os.path.basename(p) for p in paths] == [f"{k}{base}.csv" for k in
                                                    ("Summary", "Overview", "RatioSummary", "RatioOverview")]
    rows = [R.parse_row(l) for l in lines]
    t = np.array([[ms for _, ms in r[4]] for r in rows], dtype=np.float64)
    keys = sorted({(r[0], r[1]) for r in rows})
    summary = [l.split(",") for l in open(paths[0]).read().split("\n") if l]
    assert len(summary) == len(keys)
    for (size, k), got in zip(keys, summary):
        sel = np.array([(r[0], r[1]) == (size, k) for r in rows])
        assert (int(got[0]), int(got[1])) == (size, k)
        won = np.bincount(np.argmin(t[sel], axis=1), minlength=t.shape[1])
        for j in range(t.shape[1]):
            mean, mx, mn, w = (float(got[2 + 4 * j]), float(got[3 + 4 * j]), float(got[4 + 4 * j]),
                               int(got[5 + 4 * j]))
            assert abs(mean - t[sel, j].mean()) <= 1e-5 * t[sel, j].mean() + 1e-3
            assert mx == float(np.float32(t[sel, j].max())) or abs(mx - t[sel, j].max()) < 1e-3 * mx
            assert abs(mn - t[sel, j].min()) < 1e-3 * mn
            assert w == won[j]
        assert int(got[-2]) == sel.sum() and int(got[-1]) == sum(r[5] for r, s in zip(rows, sel) if s)
    ratio = [l.split(",") for l in open(paths[2]).read().split("\n") if l]
    for (size, k), got in zip(keys, ratio):
        sel = np.array([(r[0], r[1]) == (size, k) for r in rows])
        r = t[sel] / t[sel, :1]
        assert got[2] == "Sort and Choose Multiselect" and abs(float(got[3]) - 1.0) < 1e-9
        assert abs(float(got[7]) - r[:, 1].mean()) <= 1e-5 * r[:, 1].mean()
