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


def rel_err(a, b):
    """max |a-b| / max |b| (normwise, per component when a trailing component axis exists)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.ndim == 4:
        errs = []
        for c in range(a.shape[-1]):
            errs.append(rel_err(a[..., c], b[..., c]))
        return max(errs)
    scale = np.max(np.abs(b))
    d = np.max(np.abs(a - b))
    if scale == 0:
        return d
    return d / scale


def compare(g, o, names, tol=TOL, region=interior, report=None):
    worst = {}
    for n in names:
        a, b = region(g.get(n)), region(o.get(n))
        assert np.isfinite(b).all(), f"oracle {n} has non-finite values"
        e = rel_err(a, b)
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
