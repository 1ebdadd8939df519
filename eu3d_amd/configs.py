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


def block_nodes(total: int, a: float, b: float, nblocks: int = 1, myid: int = 0, bc: int = BC) -> np.ndarray:
    """Ghost-inclusive nodes of block `myid` of an axis of `total` nodes over [a, b] split into
    `nblocks` blocks that share their end nodes (Mesh::Initial, src/Mesh.cpp:75-96: n = (total-1)
    / nblocks + 1 nodes, node i at a + (myid (n-1) + i - bc) d); ghosts as make_nodes."""
    d = (b - a) / (total - 1)
    n = (total - 1) // nblocks + 1
    x = make_nodes(n, d, a, bc)
    for i in range(bc, n + bc):
        x[i] = a + (myid * (n - 1) + i - bc) * d
    for i in range(bc):
        x[i] = 2 * x[bc] - x[2 * bc - i] if 2 * bc - i < n + bc else x[bc] - (bc - i) * d
        x[n + bc + i] = 2 * x[n + bc - 1] - x[n + bc - 1 - i - 1] if n + bc - 2 - i >= bc else x[n + bc - 1] + (i + 1) * d
    return x


def rsbi_reference(decks: dict, block=(0, 0, 0), mech=None) -> Workload:
    """The reference's own initial state of the 3dRSBI deck, for one block of its MPI
    decomposition (block = (myid_x, myid_y, myid_z); (0, 0, 0) with m_block 1 x 1 x 1 is the
    whole domain):
    - ReactionInitial (src/Reaction.cpp:340-397): a cylinder of pure H2 (axis along z through
      x = 0.038, y = 0) of radius 0.022 m in air (O2 + 3.76 N2 by moles), with a Gaussian
      transition out to 1.2 times that radius, Y_H2 = exp(-5 ((r - r_i) / (r_o - r_i))^2) and O2 and
      N2 sharing the rest in the air's mass ratio;
    - FieldInitial's first loop (src/Flowfield.cpp:166-190): at rest at P_pre, T_pre where
      x >= 0.01, the moving shock's post-shock state (shock_post) behind it; V = W = 0.
    `decks` holds the grid, initialization and ctrl deck values (text, as the decks give them)."""
    mech = mech or load_default()
    g, ini, c = decks["grid"], decks["initialization"], decks["ctrl"]
    nb = [int(c.get(f"m_block_{a}", 1)) for a in "xyz"]
    bc = int(g.get("bc", BC))
    nodes = [block_nodes(int(g[f"total_n{a}"]), float(g[f"{b}a"]), float(g[f"{b}b"]), nb[d], block[d], bc)
             for d, (a, b) in enumerate((("i", "x"), ("j", "y"), ("k", "z")))]
    x, y, z = nodes
    sh = (len(x), len(y), len(z))
    p = _blank(sh, mech.ns)
    P_pre, T_pre, M_wave = float(ini["P_pre"]), float(ini["T_pre"]), float(ini["M_wave"])
    U_post, P_post, T_post = shock_post(M_wave, T_pre, P_pre)
    pre = np.broadcast_to((x >= 0.01)[:, None, None], sh)
    p["P"][...] = np.where(pre, P_pre, P_post)
    p["T"][...] = np.where(pre, T_pre, T_post)
    p["U"][...] = np.where(pre, 0.0, U_post)
    cx, cy, r_in = 0.038, 0.0, 0.022
    r_out = 1.2 * r_in
    iO2, iH2, iN2 = mech.index("O2"), mech.index("H2"), mech.index("N2")
    r2 = (x[:, None] - cx) ** 2 + (y[None, :] - cy) ** 2
    mw = mech.mw
    Y = np.zeros((sh[0], sh[1], mech.ns))
    tot = 1.0 * mw[iO2] + 3.76 * mw[iN2]
    outer, inner = r2 > r_out ** 2, r2 < r_in ** 2
    Y[outer, iO2] = 1.0 * mw[iO2] / tot
    Y[outer, iN2] = 3.76 * mw[iN2] / tot
    Y[inner, iH2] = 1.0 * mw[iH2] / (1.0 * mw[iH2])
    mid = ~(outer | inner)
    r = np.sqrt(r2[mid])
    yh2 = np.exp(-5.0 * ((r - r_in) / (r_out - r_in)) ** 2)
    Y[mid, iH2] = yh2
    Y[mid, iO2] = (1 - yh2) * mw[iO2] / (mw[iO2] + 3.76 * mw[iN2])
    Y[mid, iN2] = (1 - yh2) * 3.76 * mw[iN2] / (mw[iO2] + 3.76 * mw[iN2])
    p["YI"][...] = Y[:, :, None, :]
    return Workload("rsbi_reference", x, y, z, p, (4, 4, 4, 4, 4, 4),
                    dict(cfl=float(c.get("cfl", 0.1)), rk=1, chem=1, diff=1, limiter=0, nchem_max=0), 0)


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
