"""Extract a small fixture from the reference's shipped initial-state output
(output_8/resultsIMEX0_<rank>.dat, written by Euler::Output, src/Euler.cpp:792-973, six
significant digits, block-packed X,Y,Z,T,D,P,U,V,Ma,Y1..Y7,L,WaitTime over i fastest).

Run where the reference is mounted:
    python tests/golden/make_output8_fixture.py /root/reference/output_8
Writes tests/golden/output8_initial.json: per sampled point T, D, P, U and Y1..Y7.
"""
import json
import os
import sys

import numpy as np

NAMES = "X,Y,Z,T,D,P,U,V,Ma,Y1,Y2,Y3,Y4,Y5,Y6,Y7,L,WaitTime".split(",")
HERE = os.path.dirname(os.path.abspath(__file__))


def read(path):
    lines = open(path).read().split("\n")
    zone = lines[1].replace("\t", " ").split()
    ni, nj, nk = (int(z.split("=")[1]) for z in zone[1:4])
    vals = np.array([float(x) for x in lines[3:] if x.strip()])
    n = ni * nj * nk
    return {k: vals[i * n:(i + 1) * n] for i, k in enumerate(NAMES)}


def main(src):
    rng = np.random.default_rng(0)
    pts = []
    for rank in range(8):
        d = read(os.path.join(src, f"resultsIMEX0_{rank}.dat"))
        n = len(d["T"])
        # stratify: post-shock, pre-shock air, bubble core, bubble edge
        cls = np.where(d["U"] > 0, 0, np.where(d["Y6"] >= 0.999, 2, np.where(d["Y6"] > 0, 3, 1)))
        for c in range(4):
            idx = np.flatnonzero(cls == c)
            if len(idx):
                for i in rng.choice(idx, size=min(6, len(idx)), replace=False):
                    pts.append({k: float(d[k][i]) for k in ("X", "Y", "T", "D", "P", "U", "V",
                                                            "Y1", "Y2", "Y3", "Y4", "Y5", "Y6", "Y7")}
                               | {"rank": rank, "class": int(c)})
    out = {"source": "output_8/resultsIMEX0_*.dat", "digits": 6, "points": pts}
    with open(os.path.join(HERE, "output8_initial.json"), "w") as f:
        json.dump(out, f, indent=0)
    print(len(pts), "points")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/output_8")
