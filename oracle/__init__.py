"""CPU oracle (TEST INFRASTRUCTURE ONLY): ctypes wrapper of oracle/liboracle.so.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this package, to
check the HIP path against it. The product (eu3d_amd) never imports it.

``OracleFlow`` mirrors ``eu3d_amd.api.Context`` method for method so a parity test runs the
same sequence of reference calls on both and compares the fields.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# EU3D_ORACLE_LIB points the oracle at another build of oracle.c (oracle/check_mutants.py uses it
# to load deliberately broken copies and confirm the pins fail).
LIB_PATH = os.environ.get("EU3D_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")

Q_REF, Q_FIXED = 0x7FF, 0
FIELD3 = ["U", "V", "W", "P", "D", "T", "C", "Ma", "Wav", "Rgas", "Cp", "H", "E", "Gamma"]
FIELD4_NS = ["Mc", "Yi", "Di"]
FIELD4_NC = ["CS", "F", "G", "Q", "RHS", "PT"]
FIELD4_NS2 = ["CMS", "MD"]
# product field names -> oracle member names
NAME = {"CS": "CS", "T": "T", "E": "E", "U": "U", "V": "V", "W": "W", "P": "P", "D": "D", "H": "H",
        "GAMMA": "Gamma", "C": "C", "MA": "Ma", "RGAS": "Rgas", "CP": "Cp", "WAV": "Wav", "YI": "Yi",
        "DI": "Di", "MC": "Mc", "PT": "PT", "RHS": "RHS", "F": "F", "G": "G", "Q": "Q", "CMS": "CMS",
        "MD": "MD", "NCHEM": "Nchem"}

_dp = C.POINTER(C.c_double)


class _Mech(C.Structure):
    _fields_ = [("ns", C.c_int), ("nr", C.c_int), ("stoi_f", _dp), ("stoi_b", _dp), ("af", _dp),
                ("bf", _dp), ("eaf", _dp), ("react_tb", _dp), ("coeff0", _dp), ("coeff1", _dp),
                ("mw", _dp), ("ri", _dp)]


class _Flow(C.Structure):
    _fields_ = ([("ni", C.c_int), ("nj", C.c_int), ("nk", C.c_int), ("bc", C.c_int), ("ns", C.c_int),
                 ("xnode", _dp), ("ynode", _dp), ("znode", _dp)]
                + [(n, _dp) for n in FIELD3] + [(n, _dp) for n in FIELD4_NS]
                + [(n, _dp) for n in FIELD4_NC] + [(n, _dp) for n in FIELD4_NS2]
                + [("Nchem", C.POINTER(C.c_int)), ("dt", C.c_double), ("time", C.c_double),
                   ("quirks", C.c_uint32), ("nchem_max", C.c_int)])


_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    lib = C.CDLL(LIB_PATH)
    M, Fp = C.POINTER(_Mech), C.POINTER(_Flow)
    for fn in ["orc_update_after_cs", "orc_get_partial_t", "orc_ee", "orc_update_after_adv", "orc_explicit",
               "orc_diagonalized", "orc_update_imex", "orc_field_from_prims"]:
        getattr(lib, fn).argtypes = [Fp, M] if fn != "orc_ee" else [Fp]
    lib.orc_advection.argtypes = [Fp, C.c_int, C.c_int]
    lib.orc_cfl.argtypes = [Fp, C.c_double, C.c_double]
    lib.orc_fill_ghosts.argtypes = [Fp, M, C.POINTER(C.c_int)]
    lib.orc_trapezoid_full.argtypes = [Fp, M, C.c_double]
    lib.orc_trapezoid_prediction.argtypes = [Fp, M, C.c_double]
    lib.orc_trapezoid_adaptive.argtypes = [Fp, M, C.c_double]
    for fn in ["orc_getcpi", "orc_gethi", "orc_getsi"]:
        getattr(lib, fn).argtypes = [C.c_double, C.c_double, C.c_int, M, C.c_uint32]
        getattr(lib, fn).restype = C.c_double
    lib.orc_get_temp.argtypes = [M, C.c_uint32, C.c_double, C.c_double, _dp, C.POINTER(C.c_int)]
    lib.orc_get_temp.restype = C.c_double
    lib.orc_sum.argtypes = [C.c_int, _dp, _dp, C.c_int]
    lib.orc_sum.restype = C.c_double
    lib.orc_muscl.argtypes = [C.c_int, C.c_int, Fp, _dp, _dp, C.c_int]
    lib.orc_set_num_threads.argtypes = [C.c_int]
    lib.orc_check_div_const.argtypes = [C.c_double, C.c_long, C.c_ulonglong]
    lib.orc_check_div_const.restype = C.c_long
    lib.orc_check_div_fast.argtypes = [C.c_double, C.c_long, C.c_ulonglong]
    lib.orc_check_div_fast.restype = C.c_long
    _lib = lib
    return lib


def _p(a):
    return a.ctypes.data_as(_dp)


class OracleMech:
    def __init__(self, mech):
        self.m = mech
        self.keep = [np.ascontiguousarray(x, dtype=np.float64) for x in
                     (mech.stoi_f, mech.stoi_b, mech.af, mech.bf, mech.eaf, mech.react_tb,
                      mech.coeff0, mech.coeff1, mech.mw, mech.ri)]
        self.s = _Mech(mech.ns, mech.nr, *[_p(k) for k in self.keep])

    def ref(self):
        return C.byref(self.s)


class OracleFlow:
    """Reference-layout CPU state with the same methods as eu3d_amd.api.Context."""

    def __init__(self, xnode, ynode, znode, bc=2, mech=None, quirks=Q_FIXED):
        from eu3d_amd.mech import load_default
        self.lib = load()
        self.mech = mech or load_default()
        self.om = OracleMech(self.mech)
        ns = self.ns = self.mech.ns
        self.bc = bc
        self.xnode = np.ascontiguousarray(xnode, dtype=np.float64)
        self.ynode = np.ascontiguousarray(ynode, dtype=np.float64)
        self.znode = np.ascontiguousarray(znode, dtype=np.float64)
        self.ni, self.nj, self.nk = (len(a) - 2 * bc for a in (self.xnode, self.ynode, self.znode))
        sh = self.shape3 = (self.ni + 2 * bc, self.nj + 2 * bc, self.nk + 2 * bc)
        self.a = {}
        for n in FIELD3:
            self.a[n] = np.zeros(sh)
        for n in FIELD4_NS + FIELD4_NS2:
            self.a[n] = np.zeros(sh + (ns,))
        for n in FIELD4_NC:
            self.a[n] = np.zeros(sh + (ns + 4,))
        self.a["Nchem"] = np.ones(sh, dtype=np.int32)
        f = self.f = _Flow()
        f.ni, f.nj, f.nk, f.bc, f.ns = self.ni, self.nj, self.nk, bc, ns
        f.xnode, f.ynode, f.znode = _p(self.xnode), _p(self.ynode), _p(self.znode)
        for n in FIELD3 + FIELD4_NS + FIELD4_NC + FIELD4_NS2:
            setattr(f, n, _p(self.a[n]))
        f.Nchem = self.a["Nchem"].ctypes.data_as(C.POINTER(C.c_int))
        f.quirks = quirks
        f.nchem_max = 0
        f.dt = 0.0
        f.time = 0.0

    @property
    def quirks(self):
        return self.f.quirks

    def set_quirks(self, q):
        self.f.quirks = q

    def _F(self):
        return C.byref(self.f)

    # fields
    def get(self, name):
        a = self.a[NAME[name]]
        return a.astype(np.float64) if name == "NCHEM" else a.copy()

    def set(self, name, arr):
        a = self.a[NAME[name]]
        a[...] = np.asarray(arr).astype(a.dtype)

    def set_time(self, dt, time=0.0):
        self.f.dt, self.f.time = dt, time

    def get_time(self):
        return self.f.dt, self.f.time

    # reference calls
    def field_initial(self):
        self.lib.orc_field_from_prims(self._F(), self.om.ref())
        self.lib.orc_get_partial_t(self._F(), self.om.ref())

    def cfl(self, cfl, final_time):
        self.lib.orc_cfl(self._F(), cfl, final_time)

    def fill_ghosts(self, kinds):
        k = (C.c_int * 6)(*kinds)
        self.lib.orc_fill_ghosts(self._F(), self.om.ref(), k)

    def advection(self, diff=1, limiter=0):
        self.lib.orc_advection(self._F(), 2 if diff == 1 else 1, limiter)

    def advection_fused(self, diff=1, limiter=0):
        self.advection(diff, limiter)

    def update_after_adv(self):
        self.lib.orc_update_after_adv(self._F(), self.om.ref())

    def update_after_cs(self):
        self.lib.orc_update_after_cs(self._F(), self.om.ref())

    def partial_t(self):
        self.lib.orc_get_partial_t(self._F(), self.om.ref())

    def explicit(self):
        self.lib.orc_explicit(self._F(), self.om.ref())

    def diagonalized(self):
        self.lib.orc_diagonalized(self._F(), self.om.ref())

    def update_imex(self):
        self.lib.orc_update_imex(self._F(), self.om.ref())

    def trapezoid(self):
        self.lib.orc_trapezoid_full(self._F(), self.om.ref(), self.f.dt)

    def trapezoid_prediction(self, nchem_max=0):
        self.f.nchem_max = nchem_max if nchem_max > 0 else 10000
        self.lib.orc_trapezoid_prediction(self._F(), self.om.ref(), self.f.dt)

    def trapezoid_adaptive(self, nchem_max=0):
        self.f.nchem_max = nchem_max if nchem_max > 0 else 10000
        self.lib.orc_trapezoid_adaptive(self._F(), self.om.ref(), self.f.dt)

    def check_finite(self):
        sl = (slice(self.bc, self.bc + self.ni), slice(self.bc, self.bc + self.nj), slice(self.bc, self.bc + self.nk))
        ok = np.isfinite(self.a["D"][sl]).all() and np.isfinite(self.a["T"][sl]).all()
        return 0 if ok else -4

    def step(self, nsteps=1, cfl=0.5, final_time=1e30, bc_kinds=(1,) * 6, diff=1, limiter=0, rk=1, chem=1,
             nchem_max=0):
        """The same composition of reference calls as eu3d_step (src/Euler.cpp:266-333)."""
        for _ in range(nsteps):
            if cfl > 0:
                self.cfl(cfl, final_time)
            else:
                self.f.time += self.f.dt
            self.fill_ghosts(bc_kinds)
            if rk == 2:
                save = self.a["CS"].copy()
                sl = (slice(self.bc, self.bc + self.ni), slice(self.bc, self.bc + self.nj),
                      slice(self.bc, self.bc + self.nk))
                self.advection(diff, limiter)
                self.update_after_adv()
                self.fill_ghosts(bc_kinds)
                self.advection(diff, limiter)
                cs = self.a["CS"]
                cs[sl] = 0.5 * save[sl] + 0.5 * (cs[sl] + self.a["RHS"][sl])
                self.update_after_cs()
                if chem == 2:
                    self.trapezoid()
                    self.explicit()
                elif chem == 3:
                    self.trapezoid_adaptive(nchem_max)
            else:
                self.advection(diff, limiter)
                if chem == 1:
                    self.diagonalized()
                    self.update_imex()
                elif chem == 2:
                    self.update_after_adv()
                    self.trapezoid()
                    self.explicit()
                elif chem == 3:
                    self.update_after_adv()
                    self.trapezoid_adaptive(nchem_max)
                else:
                    self.update_after_adv()

    # scalar helpers
    def getcpi(self, T, R, sp):
        return self.lib.orc_getcpi(T, R, sp, self.om.ref(), self.f.quirks)

    def gethi(self, T, R, sp):
        return self.lib.orc_gethi(T, R, sp, self.om.ref(), self.f.quirks)

    def getsi(self, T, R, sp):
        return self.lib.orc_getsi(T, R, sp, self.om.ref(), self.f.quirks)


def set_threads(n: int):
    load().orc_set_num_threads(n)
