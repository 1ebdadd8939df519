"""Mutation check of the oracle's pins (CPU only, test infrastructure).

Each mutant below is one plausible mistake in oracle/oracle.c: a dropped term, a flipped sign, a
wrong threshold or unit. The script builds each mutant into its own liboracle.so in a work
directory, runs the CPU test suite against it (through EU3D_ORACLE_LIB) and reports which test
fails first. Every mutant must be caught; the exit status is the number that were not.

    python oracle/check_mutants.py [--workdir DIR]
"""
import argparse
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)

# (name, reference lines the restated code follows, exact text in oracle.c, replacement)
MUTANTS = [
    ("imex: drop the chemical source w*dt", "src/Flowfield.cpp:1224",
     "f->RHS[c * nc + s] + f->CMS[c * S + s] * dt;", "f->RHS[c * nc + s];"),
    ("imex: 1 + dt*MD instead of 1 - dt*MD", "src/Flowfield.cpp:1227",
     "(1.0 - dt * md)", "(1.0 + dt * md)"),
    ("diagonalized: sign of E/(R T^2) in WJH2", "src/Reaction.cpp:800",
     "m->bf[r] / T_local + m->eaf[r]", "m->bf[r] / T_local - m->eaf[r]"),
    ("diagonalized: WJH1 without its backward term", "src/Reaction.cpp:798",
     "ip.KB[r] * sb * ip.RR_B[r] / ip.Di[s];", "0.0 * sb;"),
    ("diagonalized: drop the third-body term", "src/Reaction.cpp:807",
     "ip.RR[r] * m->react_tb[s * NR + r] / m->mw[s])", "0.0)"),
    ("prediction: Y threshold 1e-3 instead of 1e-6", "src/Reaction.cpp:652",
     "f->Yi[c * S + s] >= 1e-6", "f->Yi[c * S + s] >= 1e-3"),
    ("update_after_cs: 0.9 R inside gamma", "src/Flowfield.cpp:1280",
     "(f->Cp[c3] - R_UNIV * orc_sum", "(f->Cp[c3] - 0.9 * R_UNIV * orc_sum"),
    ("update_after_cs: H without E", "src/Flowfield.cpp:1276",
     "f->H[c3] = f->P[c3] / D + Enew;", "f->H[c3] = f->P[c3] / D;"),
    ("update_after_cs: Ma without w", "src/Flowfield.cpp:1282",
     "f->Ma[c3] = sqrt(pow(U, 2) + pow(V, 2) + pow(W, 2))", "f->Ma[c3] = sqrt(pow(U, 2) + pow(V, 2))"),
    ("get_temp: half the gas constant", "src/Flowfield.cpp:1556",
     "- orc_sum(2, Y, m->ri, ns) * T0;", "- 0.5 * orc_sum(2, Y, m->ri, ns) * T0;"),
    ("partial_t: FIXED h_s scale 1e2", "src/Flowfield.cpp:1590",
     "Hi = Hi * 1e3;", "Hi = Hi * 1e2;"),
    ("trapezoid: dt P instead of dt P / 2", "src/Reaction.cpp:481,587",
     "((1.0 - dt / 2.0 * P)", "((1.0 - dt * P)"),
    ("kinetics: k_b = k_f K_c", "src/Reaction.cpp:527",
     "ip->KB[r] = ip->KF[r] / ip->Kc[r];", "ip->KB[r] = ip->KF[r] * ip->Kc[r];"),
    ("kinetics: sign of dnu in K_p", "src/Reaction.cpp:525",
     "pow((P0_PA * 1e-6), Xr)", "pow((P0_PA * 1e-6), -Xr)"),
]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workdir", default=None)
    args = ap.parse_args()
    work = args.workdir or tempfile.mkdtemp(prefix="oracle_mutants_")
    os.makedirs(work, exist_ok=True)
    src = open(os.path.join(HERE, "oracle.c")).read()
    missed = 0
    for n, (name, ref, old, new) in enumerate(MUTANTS):
        if src.count(old) != 1:
            print(f"[{name}] text not found exactly once in oracle.c", flush=True)
            missed += 1
            continue
        d = os.path.join(work, f"m{n:02d}")
        os.makedirs(d, exist_ok=True)
        c = os.path.join(d, "oracle.c")
        open(c, "w").write(src.replace(old, new))
        lib = os.path.join(d, "liboracle.so")
        subprocess.run(["gcc", "-std=c11", "-O2", "-fno-fast-math", "-ffp-contract=off", "-fPIC", "-fopenmp",
                        "-w", "-I", HERE, "-shared", "-o", lib, c, "-lm"], check=True)
        env = dict(os.environ, EU3D_ORACLE_LIB=lib)
        r = subprocess.run([sys.executable, "-m", "pytest", os.path.join(ROOT, "tests"), "-x", "-q", "-m", "not gpu and not slow",
                            "-p", "no:cacheprovider"], cwd=ROOT, env=env, stdout=subprocess.PIPE,
                           stderr=subprocess.STDOUT, text=True)
        failed = [l for l in r.stdout.splitlines() if l.startswith("FAILED")]
        if r.returncode != 0 and failed:
            print(f"caught   [{name}] ({ref}) by {failed[0].split()[1]}", flush=True)
        else:
            print(f"MISSED   [{name}] ({ref})", flush=True)
            missed += 1
    print(f"{len(MUTANTS) - missed} of {len(MUTANTS)} mutants caught")
    return missed


if __name__ == "__main__":
    sys.exit(main())
