"""The C++ host layer (eu3d_amd/host): deck, grid and mechanism readers on the CPU, and the
deck-driven driver (the reference's Euler flow over the C ABI) against the oracle on the GPU."""
import ctypes as C
import json
import os
import subprocess
import tempfile

import numpy as np
import pytest

from tests.common import compare, say

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUN = os.path.join(ROOT, "eu3d_amd", "eu3d_run")
HOST_LIB = os.path.join(ROOT, "eu3d_amd", "libeu3d_host.so")
REF_INPUT = "/root/reference/input"


def _need_driver():
    if not os.path.exists(RUN):
        pytest.skip("eu3d_run not built (run __graft_entry__.build())")


def _mech_files(tmp_path):
    from eu3d_amd import mech
    m, t = str(tmp_path / "mech.txt"), str(tmp_path / "thermo.dat")
    mech.write_text(mech.load_default(), m, t)
    return m, t


def _decks(tmp_path, grid, reaction="IMEX", cfl=0.05, extra=""):
    m, t = _mech_files(tmp_path)
    g = tmp_path / "grid.txt"
    g.write_text("#grid\n" + "\n".join(f"{k}\t{v}" for k, v in grid.items()) + "\n")
    c = tmp_path / "ctrl.txt"
    c.write_text(f"""#The grid file
gridname\t{g}

#chemical reaction kinetic model
reaction-model\t{m}
Thermofile\t{t}

#The final time
Final-Time\t1.0
cfl\t{cfl}
dt\t3e-8
count\t1
difference\tMUSCL_2
time-adv\tEE
reaction\t{reaction}
{extra}""")
    return str(c)


def _driver(args):
    """eu3d_run_main (the eu3d_run program's flow in libeu3d_host.so) called in this process, with
    its stdout captured: no child process is started from a process whose GPU runtime is active."""
    lib = C.CDLL(HOST_LIB)
    lib.eu3d_run_main.argtypes = [C.c_int, C.POINTER(C.c_char_p)]
    argv = (C.c_char_p * (len(args) + 2))(b"eu3d_run", *[a.encode() for a in args], None)
    libc = C.CDLL(None)
    with tempfile.TemporaryFile() as f:
        libc.fflush(None)
        saved = os.dup(1)
        os.dup2(f.fileno(), 1)
        try:
            rc = lib.eu3d_run_main(len(args) + 1, argv)
            libc.fflush(None)
        finally:
            os.dup2(saved, 1)
            os.close(saved)
        f.seek(0)
        out = f.read().decode()
    return rc, out


def _print_decks(ctrl):
    rc, out = _driver([ctrl, "--print-decks"])
    assert rc == 0, out
    return json.loads(out)


def _gpu_visible():
    """A GPU is visible (counted without initialising a runtime). On such a host this process may
    already hold a GPU runtime from the GPU tests, and a child spawned from it has crashed inside
    fork/exec, so the program test runs only where no GPU is."""
    try:
        import torch
        return torch.cuda.device_count() > 0
    except Exception:
        return False


def test_driver_program_prints_decks(tmp_path):
    """The eu3d_run program forwards its command line to eu3d_run_main (CPU only: parse, no GPU)."""
    _need_driver()
    if _gpu_visible():
        pytest.skip("GPU host: no child process from a process that may hold a GPU runtime; "
                    "the in-process deck tests cover the parse")
    ctrl = _decks(tmp_path, GRID_1D)
    r = subprocess.run([RUN, ctrl, "--print-decks"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert json.loads(r.stdout) == _print_decks(ctrl)


GRID_1D = {"total_ni": 160, "xa": 0.0, "xb": 159 * 5e-5, "total_nj": 2, "ya": 0.0, "yb": 5e-5,
           "total_nk": 2, "za": 0.0, "zb": 5e-5, "bc": 2}


def test_deck_reader_semantics(tmp_path):
    """Values are kept as 6-significant-digit text, tokens carry over to a line with fewer tokens,
    '#' lines are skipped, and a non-numeric value reads as 0 through getdoubleParameter."""
    _need_driver()
    ctrl = _decks(tmp_path, GRID_1D, extra="precise\t0.123456789\nlonely\n#comment\tline\nword\tabc\n")
    d = _print_decks(ctrl)
    deck, control = d["deck"], d["control"]
    assert deck["precise"] == "0.123457"
    assert deck["Final-Time"] == "1" and deck["cfl"] == "0.05" and deck["dt"] == "3e-08"
    assert deck["lonely"] == "0.123457"          # value carried over from the line before
    assert "#comment" not in deck and deck["word"] == "abc"
    assert control["cfl"] == 0.05 and control["dt"] == 0.0   # dt is read only when cfl == 0
    assert control["Reaction_Sch"] == 1 and control["Diff_Sch"] == 1 and control["TimeAdv_Sch"] == 0


def test_mesh_follows_reference_rule(tmp_path):
    """Mesh::Initial + Boundary: nodes xa + (i - bc) dx, ghosts mirrored through the end nodes."""
    _need_driver()
    grid = {"total_ni": 241, "xa": 0.0, "xb": 0.17, "total_nj": 81, "ya": 0.0, "yb": 0.08,
            "total_nk": 2, "za": 0.0, "zb": 0.02, "bc": 2}
    d = _print_decks(_decks(tmp_path, grid))["mesh"]
    assert (d["ni"], d["nj"], d["nk"], d["bc"]) == (241, 81, 2, 2)
    for key, n, a, b in (("xnode", 241, 0.0, 0.17), ("ynode", 81, 0.0, 0.08), ("znode", 2, 0.0, 0.02)):
        dx = (b - a) / (n - 1)
        x = np.zeros(n + 4)
        x[2:n + 2] = a + np.arange(n) * dx
        if n >= 3:
            x[0], x[1] = 2 * x[2] - x[4], 2 * x[2] - x[3]
            x[n + 2], x[n + 3] = 2 * x[n + 1] - x[n], 2 * x[n + 1] - x[n - 1]
        else:   # two nodes: nothing to mirror through, the spacing is extended
            x[0], x[1] = x[2] - 2 * dx, 2 * x[2] - x[3]
            x[n + 2], x[n + 3] = 2 * x[n + 1] - x[n], x[n + 1] + 2 * dx
        assert np.array_equal(np.array(d[key]), x), key
    # the 1-D decks' two-node axes: ghosts extend the spacing (nothing to mirror through)
    d1 = _print_decks(_decks(tmp_path, GRID_1D))["mesh"]
    assert np.allclose(np.array(d1["ynode"]), np.array([-2, -1, 0, 1, 2, 3]) * 5e-5, rtol=0, atol=1e-20)


def test_mechanism_reader_matches(tmp_path):
    """The C++ ReactionRead gives the packed mechanism exactly, from the text files written by
    mech.write_text and, where the reference tree is present, from its own input files."""
    _need_driver()
    from eu3d_amd import mech
    ref = mech.load_default()
    cases = [_decks(tmp_path, GRID_1D)]
    if os.path.exists(os.path.join(REF_INPUT, "H2_JA919.txt")):
        p = tmp_path / "ref"
        p.mkdir()
        g = tmp_path / "grid.txt"
        c = p / "ctrl.txt"
        c.write_text(f"gridname\t{g}\nreaction-model\t{REF_INPUT}/H2_JA919.txt\nThermofile\t{REF_INPUT}/thermo.dat\n"
                     "Final-Time\t1\ncfl\t0.5\ncount\t1\ndifference\tMUSCL_2\ntime-adv\tEE\nreaction\tIMEX\n")
        cases.append(str(c))
    for ctrl in cases:
        m = _print_decks(ctrl)["mech"]
        assert (m["NS"], m["NR"], m["TB"]) == (ref.ns, ref.nr, ref.tb)
        assert m["species"] == list(ref.species)
        for k, want in (("stoi_f", ref.stoi_f), ("stoi_b", ref.stoi_b), ("react_tb", ref.react_tb), ("af", ref.af),
                        ("bf", ref.bf), ("eaf", ref.eaf), ("coeff0", ref.coeff0), ("coeff1", ref.coeff1),
                        ("mw", ref.mw)):
            assert np.array_equal(np.array(m[k]).reshape(want.shape), want), (ctrl, k)


def _ic_file(tmp_path, prims):
    p = tmp_path / "ic.bin"
    with open(p, "wb") as f:
        for k in ("T", "P", "U", "V", "W", "YI"):
            np.ascontiguousarray(prims[k], dtype=np.float64).tofile(f)
    return str(p)


def _zn_prims(ni, n_driver=10):
    """The 1-D detonation IC on the deck's 2 x 2 transverse nodes (uniform in y and z)."""
    from eu3d_amd import configs
    wl = configs.zn1d(ni=ni, n_driver=n_driver)
    out = {}
    for k, v in wl.prims.items():
        out[k] = np.repeat(np.repeat(v[:, 2:3, 2:3], 6, axis=1), 6, axis=2) if v.ndim == 3 else \
            np.repeat(np.repeat(v[:, 2:3, 2:3, :], 6, axis=1), 6, axis=2)
    return out


def _run(ctrl, ic, args, tmp_path):
    dump = str(tmp_path / "dump.bin")
    cmd = [ctrl, "--ic", ic, "--out", str(tmp_path), "--dump", dump, *args]
    say("run", " ".join(cmd))
    rc, out = _driver(cmd)
    assert rc == 0, out
    return dump


def _oracle(decks, prims, nsteps, **step):
    import oracle
    m = decks["mesh"]
    o = oracle.OracleFlow(np.array(m["xnode"]), np.array(m["ynode"]), np.array(m["znode"]), quirks=0)
    for k, v in prims.items():
        o.set(k, v)
    o.field_initial()
    o.step(nsteps, **step)
    return o


def _check_dump(dump, o, tol=1e-10):
    n3 = o.get("T").size
    raw = np.fromfile(dump, dtype=np.float64)
    cs, T = raw[:n3 * 13].reshape(o.get("CS").shape), raw[n3 * 13:].reshape(o.get("T").shape)

    class G:   # the driver's result seen through the comparison helper's interface
        def get(self, k):
            return {"CS": cs, "T": T}[k]
    compare(G(), o, ["CS", "T"], tol=tol)


@pytest.mark.gpu
def test_driver_stepwise_trapezoid_adaptive(tmp_path):
    """Euler::Computing method by method (CFL, ghosts, advection, Update_after_Adv, adaptive
    trapezoid) on the 1-D detonation tube equals the oracle's composition after 30 steps, and
    Output writes the Tecplot block file."""
    _need_driver()
    ctrl = _decks(tmp_path, GRID_1D, reaction="Trapezoid_Adaptive", cfl=0.05)
    decks = _print_decks(ctrl)
    prims = _zn_prims(160)
    dump = _run(ctrl, _ic_file(tmp_path, prims), ["--bc", "2,1,1,1,1,1", "--steps", "30"], tmp_path)
    o = _oracle(decks, prims, 30, cfl=0.05, bc_kinds=(2, 1, 1, 1, 1, 1), rk=1, chem=3)
    _check_dump(dump, o)
    tec = tmp_path / "SeriesTrapezoid_Adaptive0_0.dat"
    lines = tec.read_text().splitlines()
    assert lines[0].startswith("Variables = X,Y,Z,T,D,P,U,V,Ma,Y1")
    assert lines[1] == "ZONE I=160\tJ=2\tK=2" and lines[2] == "datapacking=block"
    assert len(lines) == 3 + 18 * 160 * 2 * 2


@pytest.mark.gpu
def test_driver_fused_rk2(tmp_path):
    """The same tube run as one eu3d_step call with SSP-RK2 stages equals the oracle's RK2 step."""
    _need_driver()
    ctrl = _decks(tmp_path, GRID_1D, reaction="Trapezoid_Adaptive", cfl=0.05)
    decks = _print_decks(ctrl)
    prims = _zn_prims(160)
    dump = _run(ctrl, _ic_file(tmp_path, prims), ["--bc", "2,1,1,1,1,1", "--steps", "30", "--fused", "2"], tmp_path)
    o = _oracle(decks, prims, 30, cfl=0.05, bc_kinds=(2, 1, 1, 1, 1, 1), rk=2, chem=3)
    _check_dump(dump, o)


@pytest.mark.gpu
def test_driver_stepwise_imex_bubble(tmp_path):
    """Euler::Computing with IMEX (Diagonalized + Update_IMEX) on a small shock-bubble case, with
    the reference's time records (src/Euler.cpp:984-1099) written beside the results."""
    _need_driver()
    from tests.common import small_bubble
    n = (24, 12, 10)
    grid = {"total_ni": n[0], "xa": 0.0, "xb": (n[0] - 1) * 1e-4, "total_nj": n[1], "ya": 0.0,
            "yb": (n[1] - 1) * 1e-4, "total_nk": n[2], "za": 0.0, "zb": (n[2] - 1) * 1e-4, "bc": 2}
    ctrl = _decks(tmp_path, grid, reaction="IMEX", cfl=0.3)
    decks = _print_decks(ctrl)
    wl = small_bubble()
    dump = _run(ctrl, _ic_file(tmp_path, wl.prims), ["--bc", "1,1,2,2,2,2", "--steps", "5", "--time-record"],
                tmp_path)
    o = _oracle(decks, wl.prims, 5, cfl=0.3, bc_kinds=(1, 1, 2, 2, 2, 2), rk=1, chem=1)
    _check_dump(dump, o)
    rec = tmp_path / "output_1" / "time_record"

    def keyed(path):
        return dict(line.split(":", 1) for line in path.read_text().splitlines())

    total = keyed(rec / "total" / "Total_Time_0.txt")
    assert list(total) == ["myid", "FileReader", "Mpi_Initial", "MeshProcess", "InputRead", "ReactionRead",
                           "FieldInitial", "MPI_Barrier", "Computing", "Output"]
    comp = keyed(rec / "compute" / "Compute_Time_0.txt")
    assert list(comp)[:6] == ["myid", "FieldBoundary", "Mpi_Boundary", "Advection", "Reaction", "Output"]
    assert list(comp)[6:8] == ["Diagnalized", "Update_IMEX"] and "FileBoundary" in comp
    loop = float(total["Computing"])
    parts = sum(float(comp[k]) for k in ("FieldBoundary", "Mpi_Boundary", "Advection", "Reaction", "Output"))
    assert loop > 0 and 0 < parts <= loop
    assert abs(float(comp["Diagnalized"]) + float(comp["Update_IMEX"]) - float(comp["Reaction"])) <= 0.5 * loop
    cmp_ = (rec / "compare" / "IMEX_mpi_1_omp_1.txt").read_text().split("\t")
    assert float(cmp_[0]) == loop
    for f in ("All_wait.txt", "All_react.txt", "All_advert.txt", "All_DLB.txt"):
        assert (rec / "all" / f).exists()


@pytest.mark.gpu
def test_driver_reference_start_writes_shipped_output(tmp_path):
    """The reference's own run from its own decks: the 3dRSBI grid (241 x 81 x 2 nodes), its shock
    deck and 4 x 2 x 1 MPI blocks, started without an initial-condition file (ReactionInitial +
    FieldInitial on the device). One GPU holds the whole grid, and Output(0) at the first step
    writes one file per block as the reference's ranks do. Each of the eight files equals the
    reference's output_8/resultsIMEX0_<rank>.dat line for line (header, 18 x 5002 values)."""
    _need_driver()
    import gzip
    gold = os.path.join(ROOT, "tests", "golden", "output8")
    decks = json.load(open(os.path.join(gold, "decks.json")))
    init = tmp_path / "input-3dRSBI.txt"
    init.write_text("".join(f"{k}\t{v}\n" for k, v in decks["initialization"].items()))
    c = decks["ctrl"]
    extra = f"initialization\t{init}\n" + "".join(f"m_block_{a}\t{c[f'm_block_{a}']}\n" for a in "xyz")
    ctrl = _decks(tmp_path, decks["grid"], reaction=c["reaction"], cfl=float(c["cfl"]), extra=extra)
    out = tmp_path / "run"
    out.mkdir()
    cmd = [ctrl, "--out", str(out), "--steps", "1"]
    say("run", " ".join(cmd))
    rc, text = _driver(cmd)
    assert rc == 0, text
    for rank in range(8):
        mine = (out / "output_8" / f"resultsIMEX0_{rank}.dat").read_text().split("\n")
        with gzip.open(os.path.join(gold, f"resultsIMEX0_{rank}.dat.gz"), "rt") as f:
            ref = f.read().split("\n")
        assert len(mine) == len(ref), (rank, len(mine), len(ref))
        bad = [i for i, (a, b) in enumerate(zip(mine, ref)) if a != b]
        assert not bad, (rank, len(bad), [(i, mine[i], ref[i]) for i in bad[:5]])
