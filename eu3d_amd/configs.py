"""Grids and seeded initial conditions for the workloads (shared by the GPU runs and the
parity checks so both sides receive identical inputs).

Grids follow Mesh::Initial/Boundary (src/Mesh.cpp:75-121): nodes xa + (i-bc) dx inside, ghost
nodes mirrored linearly. Initial conditions are primitive fields (T, P, U, V, W, Y) over every
cell including ghosts, turned into the conserved state by ``field_initial`` (the second loop of
Flowfield::FieldInitial, src/Flowfield.cpp:193-229).

Workloads:
  zn1d      1-D H2-air detonation: driver zone at 30 atm / 3000 K burnt, premix at 1 atm / 300 K.
  bubble3d  Mach-2 shock in air striking a spherical stoichiometric H2-air bubble.
  chem      batched cell-local chemistry states (T, p, mixture fraction, progress variable).
  det3d     planar 3-D H2-air detonation with a correlated density perturbation.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .mech import R_UNIV, Mechanism, load_default

ATM = 101325.0
BC = 2


def make_nodes(n: int, d: float, a: float = 0.0, bc: int = BC) -> np.ndarray:
    """Ghost-inclusive node coordinates of one axis (src/Mesh.cpp:80-81,112-120)."""
    x = np.zeros(n + 2 * bc)
    for i in range(bc, n + bc):
        x[i] = a + (i - bc) * d
    for i in range(bc):
        x[i] = 2 * x[bc] - x[2 * bc - i] if 2 * bc - i < n + bc else x[bc] - (bc - i) * d
        x[n + bc + i] = 2 * x[n + bc - 1] - x[n + bc - 1 - i - 1] if n + bc - 2 - i >= bc else x[n + bc - 1] + (i + 1) * d
    return x


def mass_fractions(mech: Mechanism, moles: dict) -> np.ndarray:
    X = np.zeros(mech.ns)
    for k, v in moles.items():
        X[mech.index(k)] = v
    Y = X * mech.mw
    return Y / Y.sum()


def premix(mech):   # 2 H2 + O2 + 3.76 N2
    return mass_fractions(mech, {"H2": 2.0, "O2": 1.0, "N2": 3.76})


def burnt(mech):    # 2 H2O + 3.76 N2
    return mass_fractions(mech, {"H2O": 2.0, "N2": 3.76})


def air(mech):      # O2 + 3.76 N2
    return mass_fractions(mech, {"O2": 1.0, "N2": 3.76})


@dataclass
class Workload:
    name: str
    xnode: np.ndarray
    ynode: np.ndarray
    znode: np.ndarray
    prims: dict           # T, P, U, V, W (3-D), YI (4-D), reference layout, ghosts included
    bc_kinds: tuple
    step: dict            # eu3d_step parameters
    nsteps: int

    @property
    def ncells(self):
        return (len(self.xnode) - 2 * BC) * (len(self.ynode) - 2 * BC) * (len(self.znode) - 2 * BC)


def _blank(shape, ns):
    return {"T": np.zeros(shape), "P": np.zeros(shape), "U": np.zeros(shape), "V": np.zeros(shape),
            "W": np.zeros(shape), "YI": np.zeros(shape + (ns,))}


def zn1d(ni: int = 2000, dx: float = 5e-5, n_driver: int = 40, nsteps: int = 2000, mech=None) -> Workload:
    """1-D detonation tube: wall at x-, outflow at x+, trapezoid chemistry, SSP-RK2, CFL 0.5."""
    mech = mech or load_default()
    x, y, z = make_nodes(ni, dx), make_nodes(1, dx), make_nodes(1, dx)
    sh = (ni + 2 * BC, 1 + 2 * BC, 1 + 2 * BC)
    p = _blank(sh, mech.ns)
    icell = np.arange(sh[0])[:, None, None] - BC          # interior index, ghosts negative / >= ni
    drv = np.broadcast_to(icell < n_driver, sh)
    p["T"][...] = np.where(drv, 3000.0, 300.0)
    p["P"][...] = np.where(drv, 30 * ATM, ATM)
    p["YI"][...] = np.where(drv[..., None], burnt(mech), premix(mech))
    return Workload("zn1d", x, y, z, p, (2, 1, 1, 1, 1, 1),
                    dict(cfl=0.5, rk=2, chem=3, diff=1, limiter=0, nchem_max=10000), nsteps)


def cfl_ramp(step: int, cfl: float = 0.5, start: float = 0.05, nsteps: int = 20) -> float:
    """CFL number of step `step` (from 0) of a detonation start: linear from `start` to `cfl` over
    the first `nsteps` steps. The first AUSM step across the 30:1 driver jump at CFL 0.5 leaves
    the finite range in the reference arithmetic too; the ramp is build-defined."""
    return start + (cfl - start) * min(1.0, step / nsteps)


def shock_post(M: float, T_pre: float, P_pre: float):
    """Post-shock state of the moving normal shock, as src/Flowfield.cpp:172-174 forms it."""
    U_post = M * (1 - (2 + 0.40 * M ** 2) / (2.40 * M ** 2)) * np.sqrt(1.40 * 287 * T_pre)
    P_post = P_pre * (1 + 2 * 1.4 / 2.4 * (M ** 2 - 1.0))
    T_post = T_pre * (2 + 0.40 * M ** 2) / (2.40 * M ** 2) * P_post / P_pre
    return U_post, P_post, T_post


def bubble3d(ni=384, nj=128, nk=128, dx=1e-4, radius_cells=25.0, mach=2.0, nsteps=500, seed=1234,
             shock_cells=None, center_cells=None, mech=None) -> Workload:
    """Mach-2 shock (x) into a spherical stoichiometric H2-air bubble, tanh interface over 2 cells,
    1 % seeded density perturbation on the bubble; outflow in x, reflective walls in y and z; IMEX."""
    mech = mech or load_default()
    x, y, z = make_nodes(ni, dx), make_nodes(nj, dx), make_nodes(nk, dx)
    sh = (ni + 2 * BC, nj + 2 * BC, nk + 2 * BC)
    p = _blank(sh, mech.ns)
    T0, P0 = 300.0, ATM
    Up, Pp, Tp = shock_post(mach, T0, P0)
    xs = (shock_cells if shock_cells is not None else max(4, ni // 16)) * dx
    cx = (center_cells[0] if center_cells else ni * 0.35) * dx
    cy = (center_cells[1] if center_cells else nj * 0.5) * dx
    cz = (center_cells[2] if center_cells else nk * 0.5) * dx
    X, Yc, Z = np.meshgrid(x, y, z, indexing="ij")
    r = np.sqrt((X - cx) ** 2 + (Yc - cy) ** 2 + (Z - cz) ** 2)
    phi = 0.5 * (1.0 - np.tanh((r - radius_cells * dx) / (2 * dx)))
    post = X < xs
    rng = np.random.default_rng(seed)
    xi = rng.uniform(-1.0, 1.0, size=sh)
    p["T"][...] = np.where(post, Tp, T0 / (1.0 + 0.01 * xi * phi))
    p["P"][...] = np.where(post, Pp, P0)
    p["U"][...] = np.where(post, Up, 0.0)
    Ya, Yb = air(mech), premix(mech)
    p["YI"][...] = (1 - phi)[..., None] * Ya + phi[..., None] * Yb
    return Workload("bubble3d", x, y, z, p, (1, 1, 2, 2, 2, 2),
                    dict(cfl=0.5, rk=1, chem=1, diff=1, limiter=0, nchem_max=0), nsteps)


def gaussian_field(shape, corr_cells: float, rms: float, seed: int) -> np.ndarray:
    """Periodic Gaussian random field with a Gaussian spectrum of correlation length corr_cells,
    normalised to the given RMS."""
    rng = np.random.default_rng(seed)
    w = rng.standard_normal(shape)
    k = [np.fft.fftfreq(n) for n in shape]
    K2 = sum(np.meshgrid(*[kk ** 2 for kk in k], indexing="ij", sparse=True))
    filt = np.exp(-0.5 * K2 * (2 * np.pi * corr_cells) ** 2)
    f = np.real(np.fft.ifftn(np.fft.fftn(w) * filt))
    f -= f.mean()
    return f * (rms / f.std())


def det3d(ni=512, nj=256, nk=128, dx=5e-5, n_driver=16, nsteps=1000, seed=4321, mech=None) -> Workload:
    """Planar detonation: driver slab in x-cells 0..15 (30 atm, 3000 K, burnt), stoichiometric premix
    at 1 atm / 300 K with a 2 % RMS density perturbation of 10-cell correlation length; wall at x-,
    outflow at x+, periodic in y and z."""
    mech = mech or load_default()
    x, y, z = make_nodes(ni, dx), make_nodes(nj, dx), make_nodes(nk, dx)
    sh = (ni + 2 * BC, nj + 2 * BC, nk + 2 * BC)
    p = _blank(sh, mech.ns)
    icell = np.arange(sh[0])[:, None, None] - BC
    drv = np.broadcast_to(icell < n_driver, sh)
    pert = np.zeros(sh)
    pert[BC:-BC, BC:-BC, BC:-BC] = gaussian_field((ni, nj, nk), 10.0, 0.02, seed)
    p["T"][...] = np.where(drv, 3000.0, 300.0 / (1.0 + pert))
    p["P"][...] = np.where(drv, 30 * ATM, ATM)
    p["YI"][...] = np.where(drv[..., None], burnt(mech), premix(mech))
    return Workload("det3d", x, y, z, p, (2, 1, 3, 3, 3, 3),
                    dict(cfl=0.5, rk=1, chem=3, diff=1, limiter=0, nchem_max=10000), nsteps)


# ----------------------------------------------------------------- batched chemistry states
def _nasa_e(mech: Mechanism, T: np.ndarray, Y: np.ndarray, c4: bool = False) -> np.ndarray:
    """Mixture internal energy [J/kg] = 1000 sum_s Y_s (h_s(T) - R_s T), h from the NASA-9 fit."""
    lo = (T < 1000)[:, None]
    c0, c1 = mech.coeff0, mech.coeff1
    a = [np.where(lo, c0[k][None, :], c1[k][None, :]) for k in range(7)]
    b1 = np.where(lo | c4, c0[7][None, :], c1[7][None, :])
    Tc = T[:, None]
    hR = (-a[0] / Tc + a[1] * np.log(Tc) + a[2] * Tc + a[3] * Tc ** 2 / 2 + a[4] * Tc ** 3 / 3
          + a[5] * Tc ** 4 / 4 + a[6] * Tc ** 5 / 5 + b1)
    ri = mech.ri[None, :]
    return 1000.0 * np.sum(Y * (ri * hR - ri * Tc), axis=1)


def chem_states(n: int, seed: int = 2024, mech=None):
    """Seeded chemistry states: T ~ U[900, 3000] K, ln p ~ U[ln 0.5, ln 40] atm, mixture fraction
    Z ~ U[0, 0.06] (fuel H2, oxidiser air), progress c ~ U[0, 1] between the unburnt mixture and
    complete combustion to H2O. Returns rho_s [n, ns] (kg/m^3), T [n], E [n] (J/kg), p [n]."""
    mech = mech or load_default()
    rng = np.random.default_rng(seed)
    T = rng.uniform(900.0, 3000.0, n)
    p = np.exp(rng.uniform(np.log(0.5), np.log(40.0), n)) * ATM
    Z = rng.uniform(0.0, 0.06, n)
    c = rng.uniform(0.0, 1.0, n)
    iH2, iO2, iN2, iH2O = (mech.index(s) for s in ("H2", "O2", "N2", "H2O"))
    ya = air(mech)
    Yu = np.zeros((n, mech.ns))
    Yu[:, iH2] = Z
    Yu[:, iO2] = (1 - Z) * ya[iO2]
    Yu[:, iN2] = (1 - Z) * ya[iN2]
    # complete combustion H2 + 1/2 O2 -> H2O, limited by the deficient reactant
    nu = 0.5 * mech.mw[iO2] / mech.mw[iH2]
    burn_h2 = np.minimum(Yu[:, iH2], Yu[:, iO2] / nu)
    Yb = Yu.copy()
    Yb[:, iH2] -= burn_h2
    Yb[:, iO2] -= burn_h2 * nu
    Yb[:, iH2O] += burn_h2 * (1 + nu)
    Y = (1 - c)[:, None] * Yu + c[:, None] * Yb
    Y = np.maximum(Y, 0.0)
    Y /= Y.sum(axis=1, keepdims=True)
    Rmix = R_UNIV * 1000.0 * np.sum(Y / mech.mw[None, :], axis=1)
    rho = p / (Rmix * T)
    E = _nasa_e(mech, T, Y)
    return rho[:, None] * Y, T, E, p
