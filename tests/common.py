"""Shared helpers for the parity tests: build the same workload on the GPU context and on the
oracle, and compare fields over the interior."""
import sys

import numpy as np

from eu3d_amd import configs

TOL = 1e-10   # relative parity bound for FP64 fields


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def interior(x, bc=2):
    return x[bc:-bc, bc:-bc, bc:-bc]


def init_state(obj, wl):
    for k in ("T", "P", "U", "V", "W", "YI"):
        obj.set(k, wl.prims[k])
    obj.field_initial()


def make_pair(wl, quirks):
    from eu3d_amd.api import Context
    from oracle import OracleFlow
    g = Context(wl.xnode, wl.ynode, wl.znode, quirks=quirks)
    o = OracleFlow(wl.xnode, wl.ynode, wl.znode, quirks=quirks)
    init_state(g, wl)
    init_state(o, wl)
    return g, o


NS = 9
VEC = {"U", "V", "W"}


def _scale(name, b):
    """Normalisation of a field for the relative bound. Conserved state: species densities by the
    largest total density, momentum components by the largest momentum magnitude, energy by
    itself. Velocity components by the largest speed; mass fractions are absolute (scale 1).
    Everything else per component by its own largest magnitude."""
    if name == "CS":
        rho = np.max(np.abs(b[..., :NS].sum(axis=-1)))
        mom = np.max(np.sqrt((b[..., NS:NS + 3] ** 2).sum(axis=-1)))
        en = np.max(np.abs(b[..., NS + 3]))
        return np.array([rho] * NS + [mom] * 3 + [en])
    if name in ("YI",):
        return np.ones(b.shape[-1])
    if b.ndim == 4:
        return np.max(np.abs(b.reshape(-1, b.shape[-1])), axis=0)
    return np.array(np.max(np.abs(b)))


def rel_err(a, b, name="", vel_scale=None):
    """max |a-b| / scale, per component (see _scale)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    sc = _scale(name, b) if vel_scale is None else np.array(vel_scale)
    d = np.max(np.abs(a - b).reshape(-1, a.shape[-1]), axis=0) if a.ndim == 4 else np.array(np.max(np.abs(a - b)))
    sc = np.where(sc == 0, 1.0, sc)
    return float(np.max(d / sc))


def compare(g, o, names, tol=TOL, region=interior, report=None, allow_nan=False):
    """Normwise relative parity per field (per component). With allow_nan, the non-finite
    entries must coincide exactly and the finite ones are compared."""
    worst = {}
    vel = None
    if VEC & set(names):
        vel = np.max(np.sqrt(sum(region(o.get(v)) ** 2 for v in ("U", "V", "W"))))
    for n in names:
        a, b = region(g.get(n)), region(o.get(n))
        if allow_nan:
            fa, fb = np.isfinite(a), np.isfinite(b)
            assert (fa == fb).all(), f"{n}: non-finite pattern differs ({(fa != fb).sum()} entries)"
            a, b = np.where(fb, a, 0.0), np.where(fb, b, 0.0)
        else:
            assert np.isfinite(b).all(), f"oracle {n} has non-finite values"
        e = rel_err(a, b, n, vel_scale=vel if n in VEC else None)
        worst[n] = e
    if report is not None:
        report.update(worst)
    bad = {k: v for k, v in worst.items() if not v <= tol}
    assert not bad, f"parity failures (tol {tol}): {bad}"
    return worst


def small_bubble(seed=7):
    return configs.bubble3d(ni=24, nj=12, nk=10, dx=1e-4, radius_cells=4.0, shock_cells=5,
                            center_cells=(12, 6, 5), seed=seed)


def small_det(seed=9):
    return configs.det3d(ni=32, nj=12, nk=8, dx=5e-5, n_driver=6, seed=seed)


def small_zn(ni=120):
    return configs.zn1d(ni=ni, n_driver=10)


def oracle_batch(rho, T, E, dtm, quirks=0):
    """The oracle's Euler::Trapezoid(type 1) cell loop on states at rest, laid out along x."""
    from oracle import OracleFlow
    n = T.shape[0]
    x = configs.make_nodes(n, 1e-4)
    y = configs.make_nodes(1, 1e-4)
    o = OracleFlow(x, y, y, quirks=quirks)
    sh = o.shape3
    cs = np.zeros(sh + (13,))
    cs[..., 0:9] = rho[0]            # ghosts: any valid state
    cs[..., 12] = rho[0].sum() * E[0]
    cs[2:-2, 2, 2, :9] = rho
    cs[2:-2, 2, 2, 12] = rho.sum(1) * E
    To = np.full(sh, T[0])
    To[2:-2, 2, 2] = T
    o.set("CS", cs)
    o.set("T", To)
    o.update_after_cs()
    T_in = o.get("T")[2:-2, 2, 2].copy()
    E_in = o.get("E")[2:-2, 2, 2].copy()
    o.set_time(dtm)
    o.trapezoid_adaptive(10000)
    return T_in, E_in, o.get("CS")[2:-2, 2, 2, :9], o.get("T")[2:-2, 2, 2], o.get("NCHEM")[2:-2, 2, 2]
