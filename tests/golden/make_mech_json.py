"""Regenerate eu3d_amd/data/h2_ja919.json from the reference's mechanism and thermo files.

Run in a container where the reference is mounted (it only reads two data files as text):
    python tests/golden/make_mech_json.py /root/reference/input
"""
import os
import sys

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", ".."))
from eu3d_amd.mech import parse_mechanism, save_json  # noqa: E402

if __name__ == "__main__":
    src = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/input"
    m = parse_mechanism(os.path.join(src, "H2_JA919.txt"), os.path.join(src, "thermo.dat"))
    save_json(m)
    print("species", m.species, "ns", m.ns, "nr", m.nr, "tb", m.tb)
