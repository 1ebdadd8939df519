"""Emit eu3d_amd/csrc/mech_gen.h: the mechanism as constexpr tables for the HIP kernels.

The kinetics kernels are specialised at compile time to one mechanism so that the loops over
species and reactions unroll into straight-line register code and every zero stoichiometric
coefficient disappears (a runtime-indexed table would force the per-cell arrays into scratch
memory). eu3d_create() checks that the mechanism handed over the C ABI equals these tables
bit for bit and refuses any other.
"""
import os
import sys

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), ".."))
from eu3d_amd.mech import load_default  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc", "mech_gen.h")


def _d(x: float) -> str:
    r = repr(float(x))
    if "e" not in r and "." not in r and "inf" not in r and "nan" not in r:
        r += ".0"
    return r


def _arr1(v):
    return "{" + ", ".join(_d(x) for x in v) + "}"


def _arr2(m):
    return "{\n" + ",\n".join("    " + _arr1(row) for row in m) + "}"


def _iarr2(m):
    return "{\n" + ",\n".join("    {" + ", ".join(str(int(x)) for x in row) + "}" for row in m) + "}"


def generate(path: str = OUT) -> str:
    m = load_default()
    for a in (m.stoi_f, m.stoi_b):
        assert (a == a.astype(int)).all(), "stoichiometry must be integral"
    src = f"""// Generated by eu3d_amd/gen_mech.py from eu3d_amd/data/h2_ja919.json. Do not edit.
// Mechanism: {m.ns} species ({', '.join(m.species)}), {m.nr} reactions, {m.tb} third-body rows.
#pragma once
namespace mech {{
constexpr int NS = {m.ns};
constexpr int NR = {m.nr};
constexpr double MW[NS] = {_arr1(m.mw)};
// Stoichiometry (s, r): forward (reactants) and backward (products).
constexpr int SF[NS][NR] = {_iarr2(m.stoi_f)};
constexpr int SB[NS][NR] = {_iarr2(m.stoi_b)};
// Third-body efficiencies (s, r); rows of non third-body reactions are zero.
constexpr double TB[NS][NR] = {_arr2(m.react_tb)};
// Arrhenius A [cm^3-mol-s units], temperature exponent b, activation energy E [cal/mol].
constexpr double AF[NR] = {_arr1(m.af)};
constexpr double BF[NR] = {_arr1(m.bf)};
constexpr double EAF[NR] = {_arr1(m.eaf)};
// NASA-9 coefficients (k, s): a1..a7 b1 b2, low (T < 1000 K) and high range.
constexpr double C0[9][NS] = {_arr2(m.coeff0)};
constexpr double C1[9][NS] = {_arr2(m.coeff1)};
}}  // namespace mech
"""
    with open(path, "w") as f:
        f.write(src)
    return path


if __name__ == "__main__":
    print(generate())
