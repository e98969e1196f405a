"""Writes tests/golden/golden.json.

Sources of every vector:
  * philox4x32_10: the published Random123 known-answer vectors for
    Philox4x32-10 (Salmon et al., SC'11; kat_vectors of the Random123
    distribution), typed in from the publication.
  * ulp_neighbours: the four fp32 values of DistributedSMOS/pivots and slopes:3
    (two pairs one ULP apart that the reference failed to separate).
  * small_selections: tiny inputs with their k-th largest values computed by a
    pure-Python sort (k = 1 is the largest; the reference's sort-and-choose,
    DistributedSMOS/iterativeSMOSTimingFunctions.cu:67-73).
Run:  python tests/golden/make_golden.py
"""
import json
import os
import random
import struct

HERE = os.path.dirname(os.path.abspath(__file__))


def f32(v):
    return struct.unpack("<f", struct.pack("<f", v))[0]


def main():
    g = {}
    g["philox4x32_10"] = [
        {"ctr": [0, 0, 0, 0], "key": [0, 0], "out": [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]},
        {"ctr": [0xFFFFFFFF] * 4, "key": [0xFFFFFFFF] * 2, "out": [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]},
        {"ctr": [0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], "key": [0xA4093822, 0x299F31D0],
         "out": [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]},
    ]
    g["ulp_neighbours"] = ["1.101046334952116012573242187500e-02", "1.101046241819858551025390625000e-02",
                           "1.183493062853813171386718750000e-02", "1.183493155986070632934570312500e-02"]
    rnd = random.Random(1234)
    cases = []
    for n in (1, 2, 5, 17, 64, 100):
        vals = [f32(rnd.uniform(-10, 10)) for _ in range(n)]
        if n > 4:
            vals[1] = vals[3]  # a duplicate
        ks = sorted({rnd.randint(1, n) for _ in range(min(n, 9))} | {1, n})
        srt = sorted(vals)
        cases.append({"dtype": "float32", "values": vals, "ks": ks, "want": [srt[n - k] for k in ks]})
    for n in (3, 50):
        vals = [rnd.randint(0, 7) for _ in range(n)]
        ks = list(range(1, n + 1))
        srt = sorted(vals)
        cases.append({"dtype": "uint32", "values": vals, "ks": ks, "want": [srt[n - k] for k in ks]})
    g["small_selections"] = cases
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(g, f, indent=1)


if __name__ == "__main__":
    main()
