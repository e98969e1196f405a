"""Writes tests/golden/refcsv.json from the reference's published benchmark
result files (data, not source): every result file name under
"DistributedSMOS Test Result/Four machines/results" and the first 60 rows of
three of them.  The reference harness wrote these files
(DistributedSMOS/runTestsDistributedSMOS.cpp:301-303,427,461 for the rows,
DistributedSMOS/compareDistributedSMOS.cpp / compareIterativeSMOS.cu:355-359 for
the names); tests/test_refcsv.py checks the library's reference-format writer
against them.

    python tests/golden/make_refcsv_fixture.py   (needs the reference checkout)
"""
import json
import os

REF = "/root/reference/DistributedSMOS Test Result/Four machines/results"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "refcsv.json")
PICK = ("float NORMAL FLOATS k-dist:Uniform Ks 2^21 to 2^26", "uint UNIFORM UNSIGNED INTEGERS k-dist:Uniform Ks 2^17",
        "double UNIFORM DOUBLES k-dist:Uniform Random Ks 2^21")


def main():
    names = sorted(os.listdir(REF))
    rows = {}
    for p in PICK:
        f = next(n for n in names if n.startswith(p))
        with open(os.path.join(REF, f)) as fh:
            rows[f] = [fh.readline().rstrip("\n") for _ in range(60)]
    with open(OUT, "w") as fh:
        json.dump({"source": "DistributedSMOS Test Result/Four machines/results", "names": names, "rows": rows},
                  fh, indent=1)
    print(f"wrote {OUT}: {len(names)} names, {sum(len(v) for v in rows.values())} rows")


if __name__ == "__main__":
    main()
