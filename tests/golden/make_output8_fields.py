"""Copy the reference's shipped initial-state output as a whole-field fixture.

The reference's 8-rank IMEX run (m_block 4 x 2 x 1 on the 241 x 81 x 2 node grid of
input/grid-3dRSBI.txt) wrote its state at time 0 to output_8/resultsIMEX0_<rank>.dat through
Euler::Output (src/Euler.cpp:792-973): 61 x 41 x 2 points per rank, 18 variables in block
packing, six significant digits. Those eight data files are stored gzipped under
tests/golden/output8/, unchanged, with the deck values that produced them (input/ctrl.txt,
input/grid-3dRSBI.txt, input/input-3dRSBI.txt) in decks.json.

Run where the reference is mounted:
    python tests/golden/make_output8_fields.py /root/reference
"""
import gzip
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "output8")


def deck(path):
    """First two whitespace tokens of each line that is neither blank nor a comment."""
    kv = {}
    for line in open(path):
        t = line.split()
        if len(t) >= 2 and not t[0].startswith("#"):
            kv[t[0]] = t[1]
    return kv


def main(ref):
    os.makedirs(OUT, exist_ok=True)
    for rank in range(8):
        name = f"resultsIMEX0_{rank}.dat"
        with open(os.path.join(ref, "output_8", name), "rb") as f:
            data = f.read()
        with gzip.GzipFile(os.path.join(OUT, name + ".gz"), "wb", compresslevel=9, mtime=0) as g:
            g.write(data)
    ctrl = deck(os.path.join(ref, "input", "ctrl.txt"))
    decks = {"ctrl": {k: ctrl[k] for k in ("Final-Time", "cfl", "dt", "count", "difference", "time-adv",
                                           "reaction", "m_block_x", "m_block_y", "m_block_z")},
             "grid": deck(os.path.join(ref, "input", "grid-3dRSBI.txt")),
             "initialization": deck(os.path.join(ref, "input", "input-3dRSBI.txt"))}
    with open(os.path.join(OUT, "decks.json"), "w") as f:
        json.dump(decks, f, indent=1)
    print(json.dumps(decks))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
