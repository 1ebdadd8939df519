"""Batched chemistry workload: seeded states and their device layout.

States follow the batched-chemistry configuration:
- T uniform in [900, 3000] K;
- ln p uniform in [ln 0.5, ln 40] atm;
- mixture fraction Z in [0, 0.06] (fuel H2, oxidiser air);
- progress c in [0, 1] between the unburnt mixture and a burnt composition.

Two burnt compositions are available:
- "complete": the burnt composition is complete combustion to H2O (configs.chem_states).
- "equilibrium": the complete-combustion products at the sample's T and p are relaxed at constant
  density and internal energy with the solver's own FP64 chemistry (10 us in 20 steps, at most
  20000 sub-steps each). The burnt composition then carries the radical pool of the H2-air
  equilibrium line, and the cells have uneven sub-step counts at dt = 1e-8 s.

Device arrays are torch tensors used as plain memory: rho [ns][n], T [n], E [n].
"""
from __future__ import annotations

import numpy as np

from . import configs
from .mech import R_UNIV, load_default

ATM = 101325.0


def _torch():
    import torch
    return torch


def to_device(rho, T, E):
    torch = _torch()
    return (torch.tensor(np.ascontiguousarray(rho.T), device="cuda"), torch.tensor(T, device="cuda"),
            torch.tensor(E, device="cuda"))


def states(ctx, n: int, seed: int = 2024, burnt: str = "complete", relax_time: float = 1e-5, relax_steps: int = 20,
           relax_nchem_max: int = 20000, log=None):
    """Seeded batch states on the device: (rho [ns][n], T [n], E [n]) torch tensors, plus the
    host copies (rho [n][ns], T, E) they were made from."""
    mech = load_default()
    rho, T, E, p = configs.chem_states(n, seed=seed, mech=mech)
    if burnt == "complete":
        return to_device(rho, T, E), (rho, T, E)
    if burnt != "equilibrium":
        raise ValueError(f"burnt must be 'complete' or 'equilibrium', not {burnt!r}")
    # the same draws as configs.chem_states, to rebuild Y_u and Y_b per sample
    rng = np.random.default_rng(seed)
    T_s = rng.uniform(900.0, 3000.0, n)
    p_s = np.exp(rng.uniform(np.log(0.5), np.log(40.0), n)) * ATM
    Z = rng.uniform(0.0, 0.06, n)
    c = rng.uniform(0.0, 1.0, n)
    assert np.array_equal(T_s, T)
    iH2, iO2, iN2, iH2O = (mech.index(s) for s in ("H2", "O2", "N2", "H2O"))
    ya = configs.air(mech)
    Yu = np.zeros((n, mech.ns))
    Yu[:, iH2] = Z
    Yu[:, iO2] = (1 - Z) * ya[iO2]
    Yu[:, iN2] = (1 - Z) * ya[iN2]
    nu = 0.5 * mech.mw[iO2] / mech.mw[iH2]
    burn = np.minimum(Yu[:, iH2], Yu[:, iO2] / nu)
    Yb = Yu.copy()
    Yb[:, iH2] -= burn
    Yb[:, iO2] -= burn * nu
    Yb[:, iH2O] += burn * (1 + nu)
    Yb = np.maximum(Yb, 0.0)
    Yb /= Yb.sum(1, keepdims=True)
    # relax the products at the sample's T and p (constant rho and e)
    Rb = R_UNIV * 1000.0 * np.sum(Yb / mech.mw[None, :], axis=1)
    rho_b = (p_s / (Rb * T_s))[:, None] * Yb
    Eb = configs._nasa_e(mech, T_s, Yb)
    r, t, e = to_device(rho_b, T_s, Eb)
    dt = relax_time / relax_steps
    for k in range(relax_steps):
        st = ctx.chem_batch(n, r.data_ptr(), t.data_ptr(), e.data_ptr(), dt, relax_nchem_max)
        if log:
            log(f"  relax {k + 1}/{relax_steps}: active {st['active_cells']} max sub-steps {st['nchem_max_seen']}")
    Yeq = (r / r.sum(0)).T.cpu().numpy()
    bad = ~np.isfinite(Yeq).all(1)            # relaxations that left the finite range keep Y_b
    Yeq[bad] = Yb[bad]
    Yeq = np.maximum(Yeq, 0.0)
    Yeq /= Yeq.sum(1, keepdims=True)
    Y = (1 - c)[:, None] * Yu + c[:, None] * Yeq
    Y /= Y.sum(1, keepdims=True)
    Rm = R_UNIV * 1000.0 * np.sum(Y / mech.mw[None, :], axis=1)
    rho = (p_s / (Rm * T_s))[:, None] * Y
    E = configs._nasa_e(mech, T_s, Y)
    return to_device(rho, T_s, E), (rho, T_s, E)
