"""H2-air mechanism and NASA-9 thermo data: parsers for the reference's text formats
and the packed constant set the kernels consume.

Parse semantics follow the reference reader exactly:
  * mechanism  ``src/Reaction.cpp:150-283`` (``seek_to_line`` at ``:122-137`` skips n lines;
    the size line is tab-split and ``stoi``-parsed; stoichiometry rows are tab-split ``stoi``;
    Arrhenius rows are whitespace tokens ``idx A b E``; third-body rows are read into ``int``,
    so ``16.`` becomes 16; the molar-ratio line is read and otherwise unused),
  * thermo     ``src/Reaction.cpp:285-321`` (block j starts after 15 + 6j lines and holds 19
    whitespace numbers: a1..a7 b1 b2 low range, the same for the high range, then Mw),
  * ``Ri = R / Mw`` with R = 8.31434 (``src/Reaction.cpp:353-354``, ``include/Reaction.hpp:57``).

Species order is the mechanism's: H, O, H2O, OH, O2, H2, H2O2, HO2, N2
(``input/H2_JA919.txt:5``).
"""
from __future__ import annotations

import json
import os
import re
from dataclasses import dataclass, field

import numpy as np

R_UNIV = 8.31434        # include/Reaction.hpp:57, include/Flowfield.hpp:54
RU_CAL = 1.987          # include/Reaction.hpp:58
P0 = 101325.0           # include/Reaction.hpp:59

DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
DEFAULT_JSON = os.path.join(DATA_DIR, "h2_ja919.json")


def _stoi(tok: str) -> int:
    """std::stoi: skip leading whitespace, optional sign, then decimal digits."""
    m = re.match(r"\s*([+-]?\d+)", tok)
    if not m:
        raise ValueError(f"stoi: invalid argument {tok!r}")
    return int(m.group(1))


def _istream_int(tok: str) -> int:
    """``stringstream >> int``: leading integer part of the token (``16.`` -> 16)."""
    m = re.match(r"\s*([+-]?\d+)", tok)
    return int(m.group(1)) if m else 0


def _istream_double(tok: str) -> float:
    """``stringstream >> double`` on one token: longest leading float literal."""
    m = re.match(r"\s*([+-]?(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?)", tok)
    return float(m.group(1)) if m else 0.0


def _lines(path: str) -> list[str]:
    with open(path, "rb") as f:
        raw = f.read().decode("latin-1")
    # std::getline splits on '\n' only; a trailing '\r' stays in the line.
    return raw.split("\n")


def _tokens_from(lines: list[str], start: int):
    for ln in lines[start:]:
        for tok in ln.split():
            yield tok


@dataclass
class Mechanism:
    ns: int
    nr: int
    tb: int
    species: list
    stoi_f: np.ndarray       # (ns, nr)
    stoi_b: np.ndarray       # (ns, nr)
    af: np.ndarray           # (nr,)
    bf: np.ndarray           # (nr,)
    eaf: np.ndarray          # (nr,)  cal/mol
    react_tb: np.ndarray     # (ns, nr) third-body efficiencies (0 for non third-body reactions)
    coeff0: np.ndarray       # (9, ns) low-range a1..a7 b1 b2
    coeff1: np.ndarray       # (9, ns) high-range
    mw: np.ndarray           # (ns,) g/mol
    mr_init: np.ndarray = field(default_factory=lambda: np.zeros(0))

    @property
    def ri(self) -> np.ndarray:
        return R_UNIV / self.mw

    def index(self, name: str) -> int:
        return self.species.index(name)

    def to_dict(self) -> dict:
        return {
            "ns": self.ns, "nr": self.nr, "tb": self.tb, "species": list(self.species),
            "stoi_f": self.stoi_f.tolist(), "stoi_b": self.stoi_b.tolist(),
            "af": self.af.tolist(), "bf": self.bf.tolist(), "eaf": self.eaf.tolist(),
            "react_tb": self.react_tb.tolist(),
            "coeff0": self.coeff0.tolist(), "coeff1": self.coeff1.tolist(),
            "mw": self.mw.tolist(), "mr_init": self.mr_init.tolist(),
        }

    @staticmethod
    def from_dict(d: dict) -> "Mechanism":
        a = lambda k: np.asarray(d[k], dtype=np.float64)
        return Mechanism(ns=int(d["ns"]), nr=int(d["nr"]), tb=int(d["tb"]),
                         species=list(d["species"]), stoi_f=a("stoi_f"), stoi_b=a("stoi_b"),
                         af=a("af"), bf=a("bf"), eaf=a("eaf"), react_tb=a("react_tb"),
                         coeff0=a("coeff0"), coeff1=a("coeff1"), mw=a("mw"),
                         mr_init=a("mr_init"))


def parse_mechanism(mech_path: str, thermo_path: str) -> Mechanism:
    """Read a mechanism + thermo pair with the reference's fixed line offsets."""
    L = _lines(mech_path)
    # size line: after skipping 2 lines (src/Reaction.cpp:163-179)
    items = L[2].split("\t")
    vals = [_stoi(t) for t in items if t != ""]
    ns, nr, tb = vals[0], vals[1], vals[2]
    species = [t for t in L[4].replace("\r", "").split("\t") if t.strip()][:ns]
    stoi_f = np.zeros((ns, nr))
    stoi_b = np.zeros((ns, nr))
    line = 2 + 6                                    # :187
    for j in range(nr):                             # :190-212
        for row, dst in ((L[line + 2 * j], stoi_f), (L[line + 2 * j + 1], stoi_b)):
            for i, item in enumerate(row.split("\t")):
                if i < ns:
                    dst[i, j] = _stoi(item)
    line = line + 2 * nr + 2                        # :216
    toks = _tokens_from(L, line)
    af = np.zeros(nr); bf = np.zeros(nr); eaf = np.zeros(nr)
    for j in range(nr):                             # :219-236
        b = [_istream_double(next(toks)) for _ in range(4)]
        af[j], bf[j], eaf[j] = b[1], b[2], b[3]
    line = line + nr + 3                            # :240
    toks = _tokens_from(L, line)
    react_tb = np.zeros((ns, nr))
    for _ in range(tb):                             # :243-261
        a = [_istream_int(next(toks)) for _ in range(ns + 1)]
        for k in range(ns):
            react_tb[k, a[0] - 1] = a[k + 1]
    line = line + tb + 2                            # :265
    toks = _tokens_from(L, line)
    mr = np.array([_istream_double(next(toks)) for _ in range(ns)])

    T = _lines(thermo_path)
    coeff0 = np.zeros((9, ns)); coeff1 = np.zeros((9, ns)); mw = np.zeros(ns)
    line = 15                                       # :289
    for j in range(ns):                             # :292-320
        toks = _tokens_from(T, line)
        aa = [_istream_double(next(toks)) for _ in range(19)]
        coeff0[:, j] = aa[0:9]
        coeff1[:, j] = aa[9:18]
        mw[j] = aa[18]
        line += 6
    return Mechanism(ns=ns, nr=nr, tb=tb, species=species, stoi_f=stoi_f, stoi_b=stoi_b,
                     af=af, bf=bf, eaf=eaf, react_tb=react_tb, coeff0=coeff0, coeff1=coeff1,
                     mw=mw, mr_init=mr)


def write_text(mech: Mechanism, mech_path: str, thermo_path: str) -> None:
    """Write a mechanism + thermo pair in the line layout parse_mechanism (and the reference
    reader) expects; every number is written so that it reads back exactly."""
    ns, nr = mech.ns, mech.nr
    tb_rows = [r for r in range(nr) if np.any(mech.react_tb[:, r] != 0)]
    L = ["! mechanism: species, reactions, third-body rows", "NS\tNR\tTB", f"{ns}\t{nr}\t{len(tb_rows)}", "!",
         "\t".join(mech.species), "!", "! stoichiometry: forward row, backward row per reaction",
         "\t".join(mech.species)]
    for r in range(nr):
        L.append("\t".join(f"{int(v)}." for v in mech.stoi_f[:, r]))
        L.append("\t".join(f"{int(v)}." for v in mech.stoi_b[:, r]))
    L += ["!", "! reaction A b E[cal/mol]"]
    L += [f"{r + 1}\t{float(mech.af[r])!r}\t{float(mech.bf[r])!r}\t{float(mech.eaf[r])!r}" for r in range(nr)]
    L += ["!", "! third-body efficiencies", "!"]
    L += [f"{r + 1}\t" + "\t".join(f"{int(v)}." for v in mech.react_tb[:, r]) for r in tb_rows]
    L += ["!", "! initial molar ratios"]
    mr = mech.mr_init if len(mech.mr_init) == ns else np.zeros(ns)
    L.append("\t".join(repr(float(v)) for v in mr))
    with open(mech_path, "w") as f:
        f.write("\n".join(L) + "\n")
    T = ["! NASA-9 fits: per species a1..a7 b1 b2 (low range), the same (high range), Mw"] + ["!"] * 13
    T.append(mech.species[0])
    for j in range(ns):
        c = [*mech.coeff0[:, j], *mech.coeff1[:, j], mech.mw[j]]
        rows = [c[0:5], c[5:9], c[9:14], c[14:18], c[18:19]]
        T += ["  ".join(repr(float(v)) for v in row) for row in rows]
        T.append(mech.species[j + 1] if j + 1 < ns else "END")
    with open(thermo_path, "w") as f:
        f.write("\n".join(T) + "\n")


def load_default() -> Mechanism:
    """The packed 9-species / 19-reaction H2-air set (parsed from input/H2_JA919.txt and
    input/thermo.dat with :func:`parse_mechanism`)."""
    with open(DEFAULT_JSON) as f:
        return Mechanism.from_dict(json.load(f))


def save_json(mech: Mechanism, path: str = DEFAULT_JSON) -> None:
    with open(path, "w") as f:
        json.dump(mech.to_dict(), f, indent=1)
        f.write("\n")
