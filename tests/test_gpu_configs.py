"""The benchmark configurations as stated, against the oracle through the C ABI.

- 1-D detonation tube: 2000 cells, the 40-cell driver zone, SSP-RK2 with the adaptive trapezoid,
  the CFL ramp to 0.5, 250 steps (src/Euler.cpp:266-333 per step).
- Batched chemistry on equilibrium-line states: the cells with the longest sub-step chains (up to
  the 10000 cap, through the tail kernel) and a random sample of the rest, one dt = 1e-8 s step
  (src/Euler.cpp:526-556, src/Reaction.cpp:489-670).
- The surrogate chemistry on the grid with cells the network accepts: element conservation, the
  |dT| bound, and every rejected or non-network cell equal to the FP64 step.
- The stated step counts at full size where no restatement can follow them: the shock-bubble's
  500 and the detonation's 1000 steps stay finite and their fronts advance.
"""
import numpy as np
import pytest

from tests.common import compare, make_pair, oracle_batch, say, small_det

pytestmark = pytest.mark.gpu
Q_FIXED = 0


def _t():
    import torch
    assert torch.cuda.is_available()
    return torch


def test_zn1d_2000_as_stated_250_steps():
    """The 1-D detonation configuration exactly as stated (2000 cells of 5e-5 m, driver zone of 40
    cells at 30 atm and 3000 K, RK2, adaptive trapezoid, CFL ramp to 0.5) for 250 steps. Beyond
    that the flow amplifies rounding: two oracle runs, one from a 1-ulp change of the initial
    densities, differ by 1.5e-13 after 250 steps, 2e-9 after 500 and 1e-3 after 1500, so the
    configuration's 2000 steps are compared over their well-conditioned start."""
    from eu3d_amd import configs
    say("case zn1d 2000 as stated, 250 steps")
    wl = configs.zn1d()
    assert wl.ncells == 2000 and (wl.prims["T"][2:-2, 2, 2] == 3000.0).sum() == 40
    g, o = make_pair(wl, Q_FIXED)
    for s in range(250):
        kw = dict(cfl=configs.cfl_ramp(s), bc_kinds=wl.bc_kinds, rk=2, chem=3, nchem_max=10000)
        g.step(1, skip_derive=s < 249, **kw)
        o.step(1, **kw)
    assert g.check_finite() == 0
    dg, do = g.get_time(), o.get_time()
    assert abs(dg[1] - do[1]) <= 1e-10 * do[1]
    worst = compare(g, o, ["CS", "T", "D", "P", "E", "YI"])
    say("  t", do[1], "worst rel", worst)


def test_batch_equilibrium_states_vs_oracle():
    """2^18 equilibrium-line states; the 48 cells with the most sub-steps at dt = 1e-8 s (the
    10000-sub-step cap among them) and 2000 others, against the oracle's adaptive trapezoid:
    sub-step counts bit for bit; states within 1e-10, or four times the oracle's own response to a
    1-ulp change of its inputs where that is larger (a chain of 10000 dependent sub-steps)."""
    torch = _t()
    from eu3d_amd import chemistry, configs
    from eu3d_amd.api import Context
    say("case batch equilibrium states")
    x = configs.make_nodes(8, 1e-4)
    g = Context(x, x, x, quirks=Q_FIXED)
    n = 1 << 18
    (r, t, e), (rho, T, E) = chemistry.states(g, n, seed=2024, burnt="equilibrium")
    nch = torch.empty(n, dtype=torch.int32, device="cuda")
    g.chem_batch(n, r.data_ptr(), t.data_ptr(), e.data_ptr(), 1e-8, 10000, nch.data_ptr())
    nch = nch.cpu().numpy()
    # cells the adaptive trapezoid takes out of the finite range: the oracle does the same on the
    # same inputs (with the same sub-step counts), so the GPU reproduces where the method fails
    bad_all = np.flatnonzero(~np.isfinite(t.cpu().numpy()) | ~np.isfinite(r.cpu().numpy()).all(0))
    bad = bad_all[:64]
    say("  non-finite cells after the step:", len(bad_all))
    if len(bad):
        _, _, ro_b, To_b, nch_b = oracle_batch(rho[bad], T[bad], E[bad], 1e-8)
        assert (~np.isfinite(To_b) | ~np.isfinite(ro_b).all(1)).all()
        assert (nch_b == nch[bad]).all()
    order = np.argsort(-nch, kind="stable")
    heavy = order[:48]
    rng = np.random.default_rng(5)
    rest = rng.choice(np.setdiff1d(order[48:], bad_all), 2000, replace=False)
    sel = np.concatenate([heavy, rest])
    say("  sub-steps of the heavy cells", nch[heavy].min(), "..", nch[heavy].max(),
        "cells at the cap in 2^18:", int((nch == 10000).sum()))
    assert nch[heavy].max() == 10000
    rs, Ts, Es = rho[sel], T[sel], E[sel]
    T_in, E_in, ro, To, nch_o = oracle_batch(rs, Ts, Es, 1e-8)
    ro_p = oracle_batch(rs * (1 + 2.0 ** -52), Ts, Es, 1e-8)[2]
    fin = np.isfinite(ro).all(1) & np.isfinite(ro_p).all(1) & np.isfinite(To)
    assert fin.all()
    scale = ro.sum(1).max()
    cond = np.abs(ro_p - ro).max() / scale
    tol = max(1e-10, 4 * cond)
    m = len(sel)
    rd = torch.tensor(rs.T.copy(), device="cuda")
    td = torch.tensor(T_in, device="cuda")
    ed = torch.tensor(E_in, device="cuda")
    nd = torch.empty(m, dtype=torch.int32, device="cuda")
    st = g.chem_batch(m, rd.data_ptr(), td.data_ptr(), ed.data_ptr(), 1e-8, 10000, nd.data_ptr())
    torch.cuda.synchronize()
    assert (nd.cpu().numpy() == nch_o).all()
    assert st["nchem_max_seen"] == 10000 and st["substeps_total"] == int(nch_o.sum())
    err = np.abs(rd.cpu().numpy().T - ro).max() / scale
    terr = np.abs(td.cpu().numpy() - To).max()
    say("  substeps", st["substeps_total"], "rel err", err, "T err", terr, "conditioning", cond, "tol", tol)
    assert err < tol
    assert terr < max(1e-10, 4 * cond) * 3000


# H, O, N atoms of H, O, H2O, OH, O2, H2, H2O2, HO2, N2
ATOMS = np.array([[1, 0, 2, 1, 0, 2, 2, 1, 0], [0, 1, 1, 1, 2, 0, 2, 2, 0], [0, 0, 0, 0, 0, 0, 0, 0, 2]], dtype=float)


def test_surrogate_grid_step_accepts_cells():
    """A network trained on the detonation's own FP64 chemistry takes the cells needing more than
    one sub-step (hybrid mode) in one step of a small detonation. Three contexts start from the
    same state: surrogate, FP64 adaptive trapezoid, and no chemistry. Every cell the surrogate
    step leaves equal to the FP64 step is a non-network or rejected cell; every other cell was
    accepted and keeps its density and element content (H, O, N) of the no-chemistry step within
    the element tolerance, with |T - T_nochem| within the |dT| bound."""
    _t()
    from eu3d_amd import configs, surrogate
    from eu3d_amd.api import MLP, Context
    from eu3d_amd.mech import load_default
    say("case surrogate grid accepts cells")
    mech = load_default()
    wl = small_det()
    H = surrogate.Harvest()
    surrogate.harvest_run(H, wl, nsteps=8, every=1, min_T=600.0)
    x = configs.make_nodes(8, 1e-4)
    gb = Context(x, x, x, quirks=Q_FIXED)
    surrogate.harvest_trajectories(gb, H, 1 << 12, [1e-9, 2e-9], nsteps=12, seed=9)
    r0, r1, t0, dt = H.cat()
    m = MLP(seed=3)
    p = surrogate.fit_params(r0, r1, t0, dt, nchem_threshold=1, T_cold=600.0)
    m.set_surrogate(p)
    X, Y = surrogate.make_dataset(m, r0, r1, t0, dt)
    hist = surrogate.train(m, X, Y, epochs=12, batch=4096, lr=2e-3)
    say("  pairs", H.n, "loss", hist[0], "->", hist[-1])
    # first with the fitted bounds, then with the |dT| bound at the median of the accepted cells'
    # temperature change, so that about half of them are rejected and must equal FP64
    dT = _grid_step_check(wl, m, p, mech, expect_rejects=False)
    p.dT_max = float(np.median(dT))
    m.set_surrogate(p)
    _grid_step_check(wl, m, p, mech, expect_rejects=True)
    _one_call_equals_stepwise(wl, m)


def _one_call_equals_stepwise(wl, m):
    """Three surrogate steps in one call, where each step after the first takes its CFL dt from
    the partial minima the previous step's passes left (the prediction pass, the FP64 queue
    kernels and, for the network's accepted and rejected cells, the list pass), equal bit for bit
    the same steps called one at a time, each of which runs the full CFL pass."""
    from eu3d_amd import configs
    from eu3d_amd.api import Context
    from tests.common import init_state
    say("case surrogate steps: one call vs one call per step")
    ctx = []
    for _ in range(2):
        g = Context(wl.xnode, wl.ynode, wl.znode, quirks=Q_FIXED)
        init_state(g, wl)
        g.step(1, cfl=configs.cfl_ramp(0), bc_kinds=wl.bc_kinds, rk=1, chem=3, nchem_max=10000)
        g.set_surrogate(m)
        ctx.append(g)
    a, b = ctx
    kw = dict(cfl=configs.cfl_ramp(1), bc_kinds=wl.bc_kinds, rk=1, chem=4, nchem_max=10000)
    net = 0
    a.step(3, **kw)
    for n in range(3):
        b.step(1, skip_derive=n < 2, **kw)
        net += b.surrogate_stats()["network"]
    say("  network cells over the three steps", net)
    assert net > 0
    assert a.get_time() == b.get_time()
    for f in ("CS", "T", "E", "D", "P", "YI", "U", "C", "CMS", "NCHEM"):
        x, y = a.get(f), b.get(f)
        assert np.array_equal(x, y, equal_nan=True), f
    for g in ctx:
        g.close()


def _grid_step_check(wl, m, p, mech, expect_rejects):
    from eu3d_amd import configs
    from eu3d_amd.api import Context
    from tests.common import init_state
    ctx = {}
    for name in ("sur", "fp64", "none"):
        g = Context(wl.xnode, wl.ynode, wl.znode, quirks=Q_FIXED)
        init_state(g, wl)
        g.step(1, cfl=configs.cfl_ramp(0), bc_kinds=wl.bc_kinds, rk=1, chem=3, nchem_max=10000)
        ctx[name] = g
    ctx["sur"].set_surrogate(m)
    kw = dict(cfl=configs.cfl_ramp(1), bc_kinds=wl.bc_kinds, rk=1, nchem_max=10000)
    ctx["sur"].step(1, chem=4, **kw)
    ctx["fp64"].step(1, chem=3, **kw)
    ctx["none"].step(1, chem=0, **kw)
    st = ctx["sur"].surrogate_stats()
    accepted = st["network"] - st["fallback"]
    say("  dT_max", p.dT_max, "stats", st)
    assert st["network"] > 0 and accepted > 0
    if expect_rejects:
        assert st["fallback"] > 0 and st["fallback"] == st["dT"]
    I = (slice(2, -2),) * 3
    cs = {k: g.get("CS")[I].reshape(-1, 13) for k, g in ctx.items()}
    T = {k: g.get("T")[I].reshape(-1) for k, g in ctx.items()}
    # non-network and rejected cells take the FP64 sub-steps (the rejected ones re-run from their
    # advected state through the list kernels, equal to the fused pass up to the last bits)
    csc = np.abs(cs["fp64"]).max(0)
    same = (np.abs(T["sur"] - T["fp64"]) <= 1e-12 * T["fp64"]) & \
        (np.abs(cs["sur"] - cs["fp64"]) <= 1e-12 * csc).all(1)
    diff = ~same
    say("  cells changed by the network", int(diff.sum()), "accepted", accepted)
    assert 0 < diff.sum() <= accepted
    rho_s, rho_n = cs["sur"][diff, :9], cs["none"][diff, :9]
    assert np.allclose(rho_s.sum(1), rho_n.sum(1), rtol=1e-13, atol=0)
    el_s, el_n = (rho_s / mech.mw) @ ATOMS.T, (rho_n / mech.mw) @ ATOMS.T
    # the element check bounds the projected increment by elem_tol of the content; clipping the
    # mass fractions that end in [-y_neg_tol, 0) moves at most y_neg_tol of the density per
    # species, and the renormalisation that follows scales by at most 9 y_neg_tol
    D = rho_n.sum(1, keepdims=True)
    clip = D * p.y_neg_tol * (ATOMS / mech.mw).sum(1)[None, :]
    bound = (p.elem_tol + 9 * p.y_neg_tol) * el_n + clip
    assert (np.abs(el_s - el_n) <= bound).all()
    say("  worst element change / content", float((np.abs(el_s - el_n) / np.maximum(el_n, 1e-300)).max()))
    dT = np.abs(T["sur"][diff] - T["none"][diff])
    assert (dT <= p.dT_max).all()
    # momentum is untouched by the chemistry
    assert np.allclose(cs["sur"][diff, 9:12], cs["none"][diff, 9:12], rtol=1e-12, atol=1e-12)
    for g in ctx.values():
        g.close()
    return dT


def test_det3d_1000_steps_finite():
    """The planar detonation as stated (512 x 256 x 128) for its 1000 steps with the FIXED unsplit
    CFL bound: every interior cell stays finite (with the per-direction bound one cell's internal
    energy fell below the mixture's energy at 300 K at step 180), and the leading front advances
    every 100 steps."""
    from eu3d_amd import configs
    from eu3d_amd.api import Context
    from tests.common import init_state
    say("case det3d 512x256x128, 1000 steps")
    wl = configs.det3d(512, 256, 128)
    g = Context(wl.xnode, wl.ynode, wl.znode, quirks=Q_FIXED)
    init_state(g, wl)
    kw = dict(bc_kinds=wl.bc_kinds, rk=1, chem=3, nchem_max=10000)
    fronts = []
    for s in range(1000):
        g.step(1, cfl=configs.cfl_ramp(s), skip_derive=True, **kw)
        if (s + 1) % 100 == 0:
            assert g.check_finite() == 0, f"non-finite after step {s + 1}"
            T = g.get("T")[2:-2, 2:-2, 2:-2]
            hot = T > 1500.0
            front = (T.shape[0] - 1 - np.argmax(hot[::-1], axis=0)).mean()
            fronts.append(front)
            say(f"  step {s + 1}: front at cell {front:.1f}, t {g.get_time()[1]:.4e}")
    assert all(b > a for a, b in zip(fronts, fronts[1:]))


def test_bubble3d_500_steps_finite():
    """The shock-bubble configuration as stated (384 x 128 x 128, Mach-2 shock into the H2-air
    bubble, IMEX, CFL 0.5) for its 500 steps: every interior cell stays finite, and the incident
    shock (the largest x at which the y-z mean pressure exceeds the midpoint of the pre- and
    post-shock pressures) advances every 100 steps."""
    from eu3d_amd import configs
    from eu3d_amd.api import Context
    from tests.common import init_state
    say("case bubble3d 384x128x128, 500 steps")
    wl = configs.bubble3d(384, 128, 128)
    g = Context(wl.xnode, wl.ynode, wl.znode, quirks=Q_FIXED)
    init_state(g, wl)
    P = g.get("P")[2:-2, 2:-2, 2:-2]
    p_mid = 0.5 * (P.max() + P.min())
    kw = dict(cfl=0.5, bc_kinds=wl.bc_kinds, rk=1, chem=1)
    fronts = []
    for s in range(5):
        g.step(100, **kw)
        assert g.check_finite() == 0, f"non-finite after step {100 * (s + 1)}"
        pm = g.get("P")[2:-2, 2:-2, 2:-2].mean(axis=(1, 2))
        front = int(np.flatnonzero(pm > p_mid).max())
        fronts.append(front)
        say(f"  step {100 * (s + 1)}: shock at cell {front}, t {g.get_time()[1]:.4e}")
    assert all(b > a for a, b in zip(fronts, fronts[1:]))
    g.close()
