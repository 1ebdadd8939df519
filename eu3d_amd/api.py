"""ctypes binding of the C ABI in include/eu3d.h (libeu3d.so, built for gfx950).

The binding is plumbing: every computation runs in the HIP library. It loads the in-tree
library and raises if it is missing; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .mech import Mechanism, load_default

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libeu3d.so")

# quirk bits (include/eu3d.h)
Q_F1, Q_F1B, Q_F2, Q_F3, Q_F6 = 1, 2, 4, 8, 16
Q_C1, Q_C2, Q_C3, Q_C4, Q_C6, Q_C7 = 32, 64, 128, 256, 512, 1024
QUIRKS_REF = 0x7FF
QUIRKS_FIXED = 0

BC_RSBI, BC_OUTFLOW, BC_REFLECT, BC_PERIODIC, BC_NONE = 0, 1, 2, 3, 4
CHEM_NONE, CHEM_IMEX, CHEM_TRAPEZOID, CHEM_TRAP_ADAPTIVE = 0, 1, 2, 3

ERRORS = {0: "OK", -1: "EU3D_ERR_ARG", -2: "EU3D_ERR_HIP", -3: "EU3D_ERR_MECH_MISMATCH",
          -4: "EU3D_ERR_NONFINITE", -5: "EU3D_ERR_NCHEM", -6: "EU3D_ERR_UNSUPPORTED"}

FIELDS = {"CS": 0, "T": 1, "E": 2, "U": 3, "V": 4, "W": 5, "P": 6, "D": 7, "H": 8, "GAMMA": 9,
          "C": 10, "MA": 11, "RGAS": 12, "CP": 13, "WAV": 14, "YI": 15, "DI": 16, "MC": 17,
          "PT": 18, "RHS": 19, "F": 20, "G": 21, "Q": 22, "CMS": 23, "MD": 24, "NCHEM": 25}

EXPORTED = [
    "eu3d_create", "eu3d_grid_sizes", "eu3d_destroy", "eu3d_last_error", "eu3d_sync", "eu3d_set_quirks",
    "eu3d_field_components", "eu3d_set_field", "eu3d_get_field", "eu3d_set_time", "eu3d_get_time",
    "eu3d_field_initial", "eu3d_cfl_dt", "eu3d_fill_ghosts", "eu3d_advection", "eu3d_advection_fused",
    "eu3d_update_after_adv", "eu3d_update_after_cs", "eu3d_get_partial_t", "eu3d_explicit",
    "eu3d_diagonalized", "eu3d_update_imex", "eu3d_trapezoid", "eu3d_trapezoid_prediction",
    "eu3d_trapezoid_adaptive", "eu3d_check_finite", "eu3d_step", "eu3d_get_chem_stats",
    "eu3d_chem_batch", "eu3d_chem_batch_noalloc", "eu3d_chem_set_schedule", "eu3d_chem_get_schedule", "eu3d_device_field",
    "eu3d_cells_with_ghosts", "eu3d_stream", "eu3d_time_kernel",
    "eu3d_mlp_create", "eu3d_mlp_destroy", "eu3d_mlp_param_count", "eu3d_mlp_get_params", "eu3d_mlp_set_params",
    "eu3d_mlp_get_grads", "eu3d_mlp_forward", "eu3d_mlp_forward_layers", "eu3d_mlp_train_step", "eu3d_gemm_bf16", "eu3d_mlp_stream",
    "eu3d_mlp_set_surrogate", "eu3d_surrogate_features", "eu3d_surrogate_targets", "eu3d_surrogate_batch",
    "eu3d_set_surrogate", "eu3d_get_surrogate_stats", "eu3d_surrogate_chem",
]
CHEM_SURROGATE = 4


class Eu3dError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class _Grid(C.Structure):
    _fields_ = [("ni", C.c_int), ("nj", C.c_int), ("nk", C.c_int), ("bc", C.c_int),
                ("xnode", C.POINTER(C.c_double)), ("ynode", C.POINTER(C.c_double)),
                ("znode", C.POINTER(C.c_double))]


class _Mech(C.Structure):
    _fields_ = [("ns", C.c_int), ("nr", C.c_int),
                ("stoi_f", C.POINTER(C.c_double)), ("stoi_b", C.POINTER(C.c_double)),
                ("af", C.POINTER(C.c_double)), ("bf", C.POINTER(C.c_double)), ("eaf", C.POINTER(C.c_double)),
                ("react_tb", C.POINTER(C.c_double)),
                ("coeff0", C.POINTER(C.c_double)), ("coeff1", C.POINTER(C.c_double)),
                ("mw", C.POINTER(C.c_double))]


class StepParams(C.Structure):
    _fields_ = [("cfl", C.c_double), ("final_time", C.c_double), ("bc_kinds", C.c_int * 6),
                ("diff", C.c_int), ("limiter", C.c_int), ("rk", C.c_int), ("chem", C.c_int),
                ("nchem_max", C.c_int), ("skip_derive", C.c_int)]


class ChemStats(C.Structure):
    _fields_ = [("cells", C.c_longlong), ("active_cells", C.c_longlong),
                ("substeps_total", C.c_longlong), ("nchem_max_seen", C.c_int)]


class ChemSchedule(C.Structure):
    _fields_ = [("block", C.c_int), ("blocks_per_cu", C.c_int), ("waves", C.c_int), ("tail", C.c_int)]


class SurrogateParams(C.Structure):
    _fields_ = [("feat_mean", C.c_float * 12), ("feat_std", C.c_float * 12), ("out_mean", C.c_float * 9),
                ("out_std", C.c_float * 9), ("hull_lo", C.c_float * 12), ("hull_hi", C.c_float * 12),
                ("dT_max", C.c_double), ("elem_tol", C.c_double), ("y_neg_tol", C.c_double),
                ("T_cold", C.c_double), ("nchem_threshold", C.c_int), ("out_eps", C.c_double),
                ("nchem_threshold_hi", C.c_int)]


class SurrogateStats(C.Structure):
    _fields_ = [("cells", C.c_longlong), ("fallback", C.c_longlong), ("hull", C.c_longlong),
                ("negative", C.c_longlong), ("element", C.c_longlong), ("dT", C.c_longlong),
                ("network", C.c_longlong)]


_lib = None


def load_library(path: str = LIB_PATH):
    """Load libeu3d.so. Raises if it has not been built: the product has no CPU path."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"libeu3d.so not found at {path}; build it with `python -m eu3d_amd.build`")
    lib = C.CDLL(path)
    vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)
    lib.eu3d_create.argtypes = [C.POINTER(vp), C.POINTER(_Grid), C.POINTER(_Mech), C.c_uint32]
    lib.eu3d_destroy.argtypes = [vp]
    lib.eu3d_chem_batch_noalloc.argtypes = [C.c_longlong, vp, vp, vp, C.c_double, C.c_int, C.c_uint32,
                                            C.POINTER(ChemSchedule), vp, vp, C.POINTER(ChemStats)]
    lib.eu3d_grid_sizes.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_longlong),
                                    C.POINTER(C.c_ulonglong)]
    lib.eu3d_last_error.argtypes = [vp]
    lib.eu3d_last_error.restype = C.c_char_p
    for name in ["eu3d_sync", "eu3d_field_initial", "eu3d_update_after_adv", "eu3d_update_after_cs",
                 "eu3d_get_partial_t", "eu3d_explicit", "eu3d_diagonalized", "eu3d_update_imex",
                 "eu3d_trapezoid", "eu3d_check_finite"]:
        getattr(lib, name).argtypes = [vp]
    lib.eu3d_set_quirks.argtypes = [vp, C.c_uint32]
    lib.eu3d_field_components.argtypes = [vp, C.c_int]
    lib.eu3d_set_field.argtypes = [vp, C.c_int, dp]
    lib.eu3d_get_field.argtypes = [vp, C.c_int, dp]
    lib.eu3d_set_time.argtypes = [vp, C.c_double, C.c_double]
    lib.eu3d_get_time.argtypes = [vp, dp, dp]
    lib.eu3d_cfl_dt.argtypes = [vp, C.c_double, C.c_double]
    lib.eu3d_fill_ghosts.argtypes = [vp, ip]
    lib.eu3d_advection.argtypes = [vp, C.c_int, C.c_int]
    lib.eu3d_advection_fused.argtypes = [vp, C.c_int, C.c_int]
    lib.eu3d_trapezoid_prediction.argtypes = [vp, C.c_int]
    lib.eu3d_trapezoid_adaptive.argtypes = [vp, C.c_int]
    lib.eu3d_step.argtypes = [vp, C.POINTER(StepParams), C.c_int]
    lib.eu3d_get_chem_stats.argtypes = [vp, C.POINTER(ChemStats)]
    lib.eu3d_chem_batch.argtypes = [vp, C.c_longlong, vp, vp, vp, C.c_double, C.c_int, vp,
                                    C.POINTER(ChemStats)]
    lib.eu3d_chem_set_schedule.argtypes = [vp, C.POINTER(ChemSchedule)]
    lib.eu3d_chem_get_schedule.argtypes = [vp, C.POINTER(ChemSchedule)]
    lib.eu3d_device_field.argtypes = [vp, C.c_int]
    lib.eu3d_device_field.restype = vp
    lib.eu3d_cells_with_ghosts.argtypes = [vp]
    lib.eu3d_cells_with_ghosts.restype = C.c_size_t
    lib.eu3d_stream.argtypes = [vp]
    lib.eu3d_stream.restype = vp
    lib.eu3d_time_kernel.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, dp]
    fp = C.POINTER(C.c_float)
    lib.eu3d_mlp_create.argtypes = [C.POINTER(vp), C.c_uint64]
    lib.eu3d_mlp_destroy.argtypes = [vp]
    lib.eu3d_mlp_param_count.argtypes = [vp]
    lib.eu3d_mlp_param_count.restype = C.c_size_t
    for n in ("eu3d_mlp_get_params", "eu3d_mlp_set_params", "eu3d_mlp_get_grads"):
        getattr(lib, n).argtypes = [vp, fp]
    lib.eu3d_mlp_forward.argtypes = [vp, C.c_int, vp, vp]
    lib.eu3d_mlp_forward_layers.argtypes = [vp, C.c_int, vp, vp]
    lib.eu3d_mlp_train_step.argtypes = [vp, C.c_int, vp, vp, C.c_float, C.c_float, fp]
    lib.eu3d_gemm_bf16.argtypes = [C.c_int, C.c_int, C.c_int, vp, vp, C.c_int, vp]
    lib.eu3d_mlp_stream.argtypes = [vp]
    lib.eu3d_mlp_stream.restype = vp
    lib.eu3d_mlp_set_surrogate.argtypes = [vp, C.POINTER(SurrogateParams)]
    lib.eu3d_surrogate_features.argtypes = [vp, C.c_longlong, vp, vp, C.c_double, vp]
    lib.eu3d_surrogate_targets.argtypes = [vp, C.c_longlong, vp, vp, vp]
    lib.eu3d_surrogate_batch.argtypes = [vp, vp, C.c_longlong, vp, vp, vp, C.c_double, C.c_int,
                                         C.POINTER(SurrogateStats)]
    lib.eu3d_set_surrogate.argtypes = [vp, vp]
    lib.eu3d_get_surrogate_stats.argtypes = [vp, C.POINTER(SurrogateStats)]
    _lib = lib
    return lib


def grid_sizes(ni: int, nj: int, nk: int, bc: int = 2):
    """(CFL partial count, device bytes) of a context for this grid; host arithmetic, no GPU."""
    lib = load_library()
    parts, nbytes = C.c_longlong(), C.c_ulonglong()
    rc = lib.eu3d_grid_sizes(ni, nj, nk, bc, C.byref(parts), C.byref(nbytes))
    if rc != 0:
        raise Eu3dError(rc, "eu3d_grid_sizes")
    return parts.value, nbytes.value


NOALLOC_MAX_CELLS = 1 << 22


def chem_batch_noalloc(n: int, rho_ptr: int, T_ptr: int, E_ptr: int, dtm: float, nchem_max: int = 0,
                       quirks: int = 0, schedule=None, stream: int | None = None, nchem_ptr: int | None = None,
                       stats: bool = True):
    """eu3d_chem_batch_noalloc: the batched chemistry with no context and no device allocation,
    on device pointers (rho_s [ns][n], T [n], E [n]). schedule: (block, blocks_per_cu, waves,
    tail) or None; stream: a hipStream_t handle (None: the null stream). With stats=False the
    call only enqueues and returns None."""
    lib = load_library()
    sp = None
    if schedule is not None:
        sp = C.byref(ChemSchedule(*schedule))
    st = ChemStats() if stats else None
    rc = lib.eu3d_chem_batch_noalloc(n, C.c_void_p(rho_ptr), C.c_void_p(T_ptr), C.c_void_p(E_ptr), dtm, nchem_max,
                                     quirks, sp, C.c_void_p(stream) if stream else None,
                                     C.c_void_p(nchem_ptr) if nchem_ptr else None, C.byref(st) if stats else None)
    if rc != 0:
        raise Eu3dError(rc, "eu3d_chem_batch_noalloc")
    if not stats:
        return None
    return {"cells": st.cells, "active_cells": st.active_cells, "substeps_total": st.substeps_total,
            "nchem_max_seen": st.nchem_max_seen}


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def mech_struct(m: Mechanism):
    """ctypes view of a mechanism; returns (struct, keepalive arrays)."""
    keep = [np.ascontiguousarray(x, dtype=np.float64) for x in
            (m.stoi_f, m.stoi_b, m.af, m.bf, m.eaf, m.react_tb, m.coeff0, m.coeff1, m.mw)]
    s = _Mech(m.ns, m.nr, *[_dptr(k) for k in keep])
    return s, keep


class Context:
    """One device-resident Flowfield + Reaction state (an eu3d_ctx)."""

    def __init__(self, xnode, ynode, znode, bc: int = 2, mech: Mechanism | None = None,
                 quirks: int = QUIRKS_FIXED):
        self.lib = load_library()
        self.mech = mech or load_default()
        self.ns = self.mech.ns
        self.bc = bc
        self.xnode = np.ascontiguousarray(xnode, dtype=np.float64)
        self.ynode = np.ascontiguousarray(ynode, dtype=np.float64)
        self.znode = np.ascontiguousarray(znode, dtype=np.float64)
        self.ni, self.nj, self.nk = (len(a) - 2 * bc for a in (self.xnode, self.ynode, self.znode))
        self.shape3 = (self.ni + 2 * bc, self.nj + 2 * bc, self.nk + 2 * bc)
        g = _Grid(self.ni, self.nj, self.nk, bc, _dptr(self.xnode), _dptr(self.ynode), _dptr(self.znode))
        ms, self._mkeep = mech_struct(self.mech)
        h = C.c_void_p()
        rc = self.lib.eu3d_create(C.byref(h), C.byref(g), C.byref(ms), quirks)
        if rc != 0:
            raise Eu3dError(rc, "eu3d_create failed")
        self.h = h
        self.quirks = quirks

    # -- lifetime
    def close(self):
        if getattr(self, "h", None):
            self.lib.eu3d_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc: int, what: str):
        if rc != 0:
            msg = self.lib.eu3d_last_error(self.h).decode(errors="replace")
            raise Eu3dError(rc, f"{what}: {msg}")

    # -- fields (reference layout)
    def field_shape(self, name: str):
        nc = self.lib.eu3d_field_components(self.h, FIELDS[name])
        return self.shape3 if nc == 1 else self.shape3 + (nc,)

    def set(self, name: str, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        if a.shape != self.field_shape(name):
            raise ValueError(f"{name}: shape {a.shape} != {self.field_shape(name)}")
        self._chk(self.lib.eu3d_set_field(self.h, FIELDS[name], _dptr(a)), f"set {name}")

    def get(self, name: str) -> np.ndarray:
        a = np.empty(self.field_shape(name), dtype=np.float64)
        self._chk(self.lib.eu3d_get_field(self.h, FIELDS[name], _dptr(a)), f"get {name}")
        return a

    def set_quirks(self, q: int):
        self._chk(self.lib.eu3d_set_quirks(self.h, q), "set_quirks")
        self.quirks = q

    def set_time(self, dt: float, time: float = 0.0):
        self._chk(self.lib.eu3d_set_time(self.h, dt, time), "set_time")

    def get_time(self):
        dt, t = C.c_double(), C.c_double()
        self._chk(self.lib.eu3d_get_time(self.h, C.byref(dt), C.byref(t)), "get_time")
        return dt.value, t.value

    # -- reference method bodies
    def field_initial(self):
        self._chk(self.lib.eu3d_field_initial(self.h), "field_initial")

    def cfl(self, cfl: float, final_time: float):
        self._chk(self.lib.eu3d_cfl_dt(self.h, cfl, final_time), "cfl")

    def fill_ghosts(self, kinds):
        k = (C.c_int * 6)(*kinds)
        self._chk(self.lib.eu3d_fill_ghosts(self.h, k), "fill_ghosts")

    def advection(self, diff: int = 1, limiter: int = 0):
        self._chk(self.lib.eu3d_advection(self.h, diff, limiter), "advection")

    def advection_fused(self, diff: int = 1, limiter: int = 0):
        self._chk(self.lib.eu3d_advection_fused(self.h, diff, limiter), "advection_fused")

    def update_after_adv(self):
        self._chk(self.lib.eu3d_update_after_adv(self.h), "update_after_adv")

    def update_after_cs(self):
        self._chk(self.lib.eu3d_update_after_cs(self.h), "update_after_cs")

    def partial_t(self):
        self._chk(self.lib.eu3d_get_partial_t(self.h), "partial_t")

    def explicit(self):
        self._chk(self.lib.eu3d_explicit(self.h), "explicit")

    def diagonalized(self):
        self._chk(self.lib.eu3d_diagonalized(self.h), "diagonalized")

    def update_imex(self):
        self._chk(self.lib.eu3d_update_imex(self.h), "update_imex")

    def trapezoid(self):
        self._chk(self.lib.eu3d_trapezoid(self.h), "trapezoid")

    def trapezoid_prediction(self, nchem_max: int = 0):
        self._chk(self.lib.eu3d_trapezoid_prediction(self.h, nchem_max), "trapezoid_prediction")

    def trapezoid_adaptive(self, nchem_max: int = 0):
        self._chk(self.lib.eu3d_trapezoid_adaptive(self.h, nchem_max), "trapezoid_adaptive")

    def check_finite(self) -> int:
        return self.lib.eu3d_check_finite(self.h)

    def step(self, nsteps: int = 1, cfl: float = 0.5, final_time: float = 1e30, bc_kinds=(1,) * 6,
             diff: int = 1, limiter: int = 0, rk: int = 1, chem: int = CHEM_IMEX, nchem_max: int = 0,
             skip_derive: bool = False):
        """nsteps steps of eu3d_step. skip_derive leaves the derived fields (Di, Mc, Wav, Rgas, Cp,
        Ma, Partion_T) unrefreshed, for callers stepping one call at a time."""
        p = StepParams(cfl, final_time, (C.c_int * 6)(*bc_kinds), diff, limiter, rk, chem, nchem_max,
                       1 if skip_derive else 0)
        self._chk(self.lib.eu3d_step(self.h, C.byref(p), nsteps), "step")

    def sync(self):
        self._chk(self.lib.eu3d_sync(self.h), "sync")

    def chem_stats(self) -> dict:
        s = ChemStats()
        self._chk(self.lib.eu3d_get_chem_stats(self.h, C.byref(s)), "chem_stats")
        return {"cells": s.cells, "active_cells": s.active_cells, "substeps_total": s.substeps_total,
                "nchem_max_seen": s.nchem_max_seen}

    def chem_batch(self, n: int, rho_ptr: int, T_ptr: int, E_ptr: int, dtm: float, nchem_max: int = 0,
                   nchem_ptr: int | None = None) -> dict:
        """Batched chemistry on device pointers (rho_s SoA [ns][n], T [n], E [n])."""
        _order_after_producer()
        s = ChemStats()
        rc = self.lib.eu3d_chem_batch(self.h, n, C.c_void_p(rho_ptr), C.c_void_p(T_ptr), C.c_void_p(E_ptr),
                                      dtm, nchem_max, C.c_void_p(nchem_ptr) if nchem_ptr else None, C.byref(s))
        self._chk(rc, "chem_batch")
        return {"cells": s.cells, "active_cells": s.active_cells, "substeps_total": s.substeps_total,
                "nchem_max_seen": s.nchem_max_seen}

    def set_schedule(self, block: int, blocks_per_cu: int, waves: int = 2, tail: int | None = None):
        """Persistent chemistry kernel schedule: threads per block, blocks per CU, the occupancy
        target (waves per SIMD: 1, 2 or 4) of its register allocation, and the sub-step count above
        which a cell runs one-per-wavefront (None keeps the current value, 0 disables)."""
        if tail is None:
            tail = self.get_schedule()["tail"]
        s = ChemSchedule(block, blocks_per_cu, waves, tail)
        self._chk(self.lib.eu3d_chem_set_schedule(self.h, C.byref(s)), "set_schedule")

    def get_schedule(self):
        s = ChemSchedule()
        self._chk(self.lib.eu3d_chem_get_schedule(self.h, C.byref(s)), "get_schedule")
        return {"block": s.block, "blocks_per_cu": s.blocks_per_cu, "waves": s.waves, "tail": s.tail}

    def set_surrogate(self, mlp):
        self._chk(self.lib.eu3d_set_surrogate(self.h, mlp.h if mlp is not None else None), "set_surrogate")

    def surrogate_stats(self) -> dict:
        s = SurrogateStats()
        self._chk(self.lib.eu3d_get_surrogate_stats(self.h, C.byref(s)), "surrogate_stats")
        return {k: getattr(s, k) for k, _ in SurrogateStats._fields_}

    def surrogate_batch(self, mlp, n, rho_ptr, T_ptr, E_ptr, dtm, nchem_max=0) -> dict:
        _order_after_producer()
        s = SurrogateStats()
        rc = self.lib.eu3d_surrogate_batch(self.h, mlp.h, n, C.c_void_p(rho_ptr), C.c_void_p(T_ptr),
                                           C.c_void_p(E_ptr), dtm, nchem_max, C.byref(s))
        self._chk(rc, "surrogate_batch")
        return {k: getattr(s, k) for k, _ in SurrogateStats._fields_}

    def device_field(self, name: str) -> int:
        return self.lib.eu3d_device_field(self.h, FIELDS[name])

    def stream(self) -> int:
        return self.lib.eu3d_stream(self.h)

    def time_kernel(self, kernel: str, reps: int = 10, diff: int = 1, limiter: int = 0) -> float:
        """Mean launch duration [ms] of one step kernel ("flux_z", "flux_xy", "imex_fused",
        "trap_first"), HIP events on the context stream."""
        ms = C.c_double()
        k = {"flux_z": 0, "flux_xy": 1, "imex_fused": 2, "trap_first": 3}[kernel]
        self._chk(self.lib.eu3d_time_kernel(self.h, k, reps, diff, limiter, C.byref(ms)), "time_kernel")
        return ms.value


def _vp(x):
    """Device pointer of a torch tensor (plumbing only)."""
    return C.c_void_p(x.data_ptr())


def _order_after_producer():
    """Wait for the work torch queued on its current stream (the producer of the device buffers
    handed to the library); the library's own streams are ordered after the null stream only."""
    import sys
    torch = sys.modules.get("torch")
    if torch is not None and torch.cuda.is_initialized():
        torch.cuda.current_stream().synchronize()


class MLP:
    """The chemistry surrogate network (an eu3d_mlp): 12 -> 4 x 256 GELU -> 9, BF16 MFMA."""

    NIN, NOUT = 32, 16

    def __init__(self, seed: int = 1):
        self.lib = load_library()
        h = C.c_void_p()
        rc = self.lib.eu3d_mlp_create(C.byref(h), seed)
        if rc != 0:
            raise Eu3dError(rc, "eu3d_mlp_create failed")
        self.h = h
        self.n_params = self.lib.eu3d_mlp_param_count(h)
        self.params_set = False

    def close(self):
        if getattr(self, "h", None):
            self.lib.eu3d_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc != 0:
            raise Eu3dError(rc, what)

    def get_params(self) -> np.ndarray:
        a = np.empty(self.n_params, dtype=np.float32)
        self._chk(self.lib.eu3d_mlp_get_params(self.h, a.ctypes.data_as(C.POINTER(C.c_float))), "get_params")
        return a

    def set_params(self, a):
        a = np.ascontiguousarray(a, dtype=np.float32)
        assert a.size == self.n_params
        self._chk(self.lib.eu3d_mlp_set_params(self.h, a.ctypes.data_as(C.POINTER(C.c_float))), "set_params")

    def get_grads(self) -> np.ndarray:
        a = np.empty(self.n_params, dtype=np.float32)
        self._chk(self.lib.eu3d_mlp_get_grads(self.h, a.ctypes.data_as(C.POINTER(C.c_float))), "get_grads")
        return a

    def forward(self, feats_bf16, out_f32):
        """feats: torch int16 [n, 32] holding BF16 bits; out: torch float32 [n, 16]."""
        _order_after_producer()
        n = feats_bf16.shape[0]
        self._chk(self.lib.eu3d_mlp_forward(self.h, n, _vp(feats_bf16), _vp(out_f32)), "mlp_forward")

    def forward_layers(self, feats_bf16, out_f32):
        """The same inference layer by layer (the training path's GEMMs)."""
        _order_after_producer()
        n = feats_bf16.shape[0]
        self._chk(self.lib.eu3d_mlp_forward_layers(self.h, n, _vp(feats_bf16), _vp(out_f32)), "mlp_forward_layers")

    def train_step(self, feats_bf16, targets_f32, lr: float, wd: float = 0.0) -> float:
        _order_after_producer()
        loss = C.c_float()
        n = feats_bf16.shape[0]
        self._chk(self.lib.eu3d_mlp_train_step(self.h, n, _vp(feats_bf16), _vp(targets_f32), lr, wd, C.byref(loss)),
                  "mlp_train_step")
        return loss.value

    def stream(self) -> int:
        return self.lib.eu3d_mlp_stream(self.h)

    def set_surrogate(self, p: "SurrogateParams"):
        self._chk(self.lib.eu3d_mlp_set_surrogate(self.h, C.byref(p)), "set_surrogate")
        self.params_set = True

    def features(self, rho, T, dt, out_bf16):
        _order_after_producer()
        n = T.shape[0]
        self._chk(self.lib.eu3d_surrogate_features(self.h, n, _vp(rho), _vp(T), dt, _vp(out_bf16)), "features")

    def targets(self, rho0, rho1, out_f32):
        _order_after_producer()
        n = rho0.shape[1]
        self._chk(self.lib.eu3d_surrogate_targets(self.h, n, _vp(rho0), _vp(rho1), _vp(out_f32)), "targets")


def gemm_bf16(A_bits, B_bits, b_kc: bool, C_out):
    """C = A * B on device through the surrogate's GEMM kernel (A [M x K] BF16 bits, B [N x K]
    when b_kc else [K x N], C FP32 [M x N])."""
    _order_after_producer()
    lib = load_library()
    M, K = A_bits.shape
    N = B_bits.shape[0] if b_kc else B_bits.shape[1]
    rc = lib.eu3d_gemm_bf16(M, N, K, _vp(A_bits), _vp(B_bits), 1 if b_kc else 0, _vp(C_out))
    if rc != 0:
        raise Eu3dError(rc, "gemm_bf16")
