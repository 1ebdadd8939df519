"""GPU parity against the CPU oracle through the C ABI, per reference method and over steps."""
import numpy as np
import pytest

from tests.common import compare, make_pair, say, small_bubble, small_det, small_zn

pytestmark = pytest.mark.gpu

PRIMS = ["D", "T", "P", "U", "V", "W", "E", "H", "GAMMA", "C", "YI", "CS"]
Q_REF, Q_FIXED = 0x7FF, 0


@pytest.mark.parametrize("quirks", [Q_FIXED, Q_REF])
def test_field_initial(quirks):
    say("case field_initial quirks", quirks)
    g, o = make_pair(small_bubble(), quirks)
    compare(g, o, ["CS", "D", "E", "H", "GAMMA", "C", "MC", "DI", "PT"])


@pytest.mark.parametrize("quirks", [Q_FIXED, Q_REF])
def test_update_after_cs(quirks):
    say("case update_after_cs quirks", quirks)
    g, o = make_pair(small_bubble(), quirks)
    # perturb the conserved state so the Newton has work to do
    cs = o.get("CS")
    rng = np.random.default_rng(3)
    cs[..., -1] *= 1.0 + 0.05 * rng.uniform(-1, 1, cs.shape[:-1])
    for obj in (g, o):
        obj.set("CS", cs)
        obj.update_after_cs()
    compare(g, o, PRIMS + ["MA", "RGAS", "CP", "WAV", "MC", "DI"])


@pytest.mark.parametrize("quirks", [Q_FIXED, Q_REF])
@pytest.mark.parametrize("diff,limiter", [(1, 0), (0, 0), (1, 1), (1, 2), (1, 3), (1, 4)])
def test_advection(quirks, diff, limiter):
    say("case advection quirks", quirks, "diff", diff, "limiter", limiter)
    g, o = make_pair(small_bubble(), quirks)
    for obj in (g, o):
        obj.set_time(2e-8, 0.0)
        obj.fill_ghosts((0, 0, 0, 0, 0, 0) if quirks else (1, 1, 2, 2, 2, 2))
        obj.advection(diff, limiter)
    # Van Leer and Van Albada evaluate phi(1/r) at r = 0 as inf/inf = NaN in flat regions
    # (include/Differ.hpp:49,67); the GPU must reproduce exactly where that happens.
    compare(g, o, ["RHS"], allow_nan=limiter in (1, 2))


@pytest.mark.parametrize("quirks", [Q_FIXED, Q_REF])
@pytest.mark.parametrize("limiter", [0, 3])
def test_advection_fused(quirks, limiter):
    say("case advection_fused quirks", quirks, "limiter", limiter)
    g, o = make_pair(small_bubble(), quirks)
    for obj in (g, o):
        obj.set_time(2e-8, 0.0)
        obj.fill_ghosts((0, 0, 0, 0, 0, 0) if quirks else (1, 1, 2, 2, 2, 2))
        obj.advection_fused(1, limiter)
    compare(g, o, ["RHS"])


@pytest.mark.parametrize("kinds", [(0,) * 6, (1, 1, 2, 2, 2, 2), (2, 1, 3, 3, 3, 3), (2, 3, 0, 1, 3, 2),
                                   (4, 0, 0, 4, 1, 4), (3, 3, 0, 2, 0, 0)])
def test_fill_ghosts(kinds):
    """All ghost cells, corners included, against the oracle's three face passes (the GPU fills a
    grid like this one in one launch that composes them)."""
    say("case fill_ghosts", kinds)
    q = Q_REF if kinds[0] == 0 else Q_FIXED
    g, o = make_pair(small_bubble(), q)
    for obj in (g, o):
        obj.fill_ghosts(kinds)
    whole = lambda x: x
    compare(g, o, ["U", "V", "W", "P", "D", "T", "H", "GAMMA", "E", "YI", "MC", "RGAS", "CP"], region=whole)


def test_fill_ghosts_thin_grid():
    """A grid one cell thick in y and z (sources there can be ghosts): the three face passes."""
    say("case fill_ghosts thin")
    wl = small_zn()
    g, o = make_pair(wl, Q_FIXED)
    for obj in (g, o):
        obj.fill_ghosts((0, 2, 1, 1, 1, 1))
    compare(g, o, ["U", "V", "W", "P", "D", "T", "H", "GAMMA", "E", "YI"], region=lambda x: x)


@pytest.mark.parametrize("quirks", [Q_FIXED, Q_REF])
def test_cfl(quirks):
    say("case cfl quirks", quirks)
    g, o = make_pair(small_bubble(), quirks)
    for obj in (g, o):
        obj.cfl(0.5, 1.0)
    (dg, tg), (do, to) = g.get_time(), o.get_time()
    assert abs(dg - do) <= 1e-14 * do and abs(tg - to) <= 1e-14 * to


@pytest.mark.parametrize("quirks", [Q_FIXED, Q_REF])
def test_imex_kernels(quirks):
    say("case imex quirks", quirks)
    g, o = make_pair(small_det(), quirks)
    for obj in (g, o):
        obj.set_time(1e-8, 0.0)
        obj.fill_ghosts((2, 1, 3, 3, 3, 3))
        obj.advection(1, 0)
        obj.diagonalized()
    compare(g, o, ["CMS", "MD"])
    for obj in (g, o):
        obj.update_imex()
    compare(g, o, PRIMS + ["PT"])


def _plant_denormal_trace(g, o, wl):
    """H and HO2 at denormal densities in a block of air cells, as the shock-bubble's air ahead
    of the shock holds them at step 183 of the 384 x 128 x 128 case (the densities found there)."""
    cs = o.get("CS")
    cs[3:8, 3:6, 3:6, 0] = 5.305793907958817e-309
    cs[3:8, 3:6, 3:6, 7] = 1.7374740561301328e-307
    for obj in (g, o):
        obj.set("CS", cs)
        obj.update_after_cs()
        obj.set_time(2.794028552407666e-08, 0.0)


def test_imex_denormal_trace_species():
    """The diagonal Jacobian divides rate terms by a species' density (src/Reaction.cpp:800-807).
    For a density below about 5.6e-309 its reciprocal alone overflows, while the reference's
    quotient is finite (here -3.9e6 to -1.4e7 in MD of H): MD, CMS, the IMEX update and two fused
    steps equal the oracle's and stay finite."""
    say("case imex denormal trace species")
    wl = small_bubble()
    g, o = make_pair(wl, Q_FIXED)
    _plant_denormal_trace(g, o, wl)
    for obj in (g, o):
        obj.fill_ghosts(wl.bc_kinds)
        obj.advection(1, 0)
        obj.diagonalized()
    assert np.isfinite(g.get("MD")).all()
    compare(g, o, ["CMS", "MD"])
    for obj in (g, o):
        obj.update_imex()
    compare(g, o, PRIMS + ["PT"])
    g, o = make_pair(wl, Q_FIXED)
    _plant_denormal_trace(g, o, wl)
    kw = dict(nsteps=2, cfl=0.1, bc_kinds=wl.bc_kinds, rk=1, chem=1)
    g.step(**kw)
    o.step(**kw)
    assert g.check_finite() == 0
    compare(g, o, ["CS", "T", "D", "P", "E", "YI"])


@pytest.mark.parametrize("quirks", [Q_FIXED, Q_REF])
def test_trapezoid_full(quirks):
    say("case trapezoid quirks", quirks)
    g, o = make_pair(small_det(), quirks)
    for obj in (g, o):
        obj.set_time(5e-9, 0.0)
        obj.trapezoid()
    compare(g, o, ["DI", "CMS"])
    for obj in (g, o):
        obj.explicit()
    compare(g, o, ["CS", "D", "T", "E"] if quirks == Q_FIXED else ["CS", "D"])


@pytest.mark.parametrize("quirks", [Q_FIXED, Q_REF])
def test_trapezoid_adaptive(quirks):
    say("case trapezoid_adaptive quirks", quirks)
    g, o = make_pair(small_det(), quirks)
    for obj in (g, o):
        obj.set_time(2e-8, 0.0)
        obj.trapezoid_adaptive(200)
    nch_g, nch_o = g.get("NCHEM"), o.get("NCHEM")
    assert (nch_g[2:-2, 2:-2, 2:-2] == nch_o[2:-2, 2:-2, 2:-2]).all()
    if quirks == Q_REF:
        assert (nch_o[2:-2, 2:-2, 2:-2] == 1).all()
    compare(g, o, PRIMS + ["CMS"])


@pytest.mark.parametrize("chem", [1, 2, 3])
def test_steps_fixed(chem):
    say("case steps chem", chem)
    wl = small_det()
    g, o = make_pair(wl, Q_FIXED)
    # CFL 0.05: the 30:1 driver jump needs the start-up ramp (AUSM puts the first step's
    # pressure jump into momentum with no energy flux)
    kw = dict(nsteps=4, cfl=0.05, bc_kinds=wl.bc_kinds, rk=1, chem=chem, nchem_max=200)
    g.step(**kw)
    o.step(**kw)
    assert g.check_finite() == 0
    compare(g, o, ["CS", "T", "D", "P", "E", "YI"])


def test_steps_trap_adaptive_ref():
    """The step's adaptive-trapezoid path fuses Update_after_Adv into the prediction pass; under
    REF quirk F1 that update adds only the species components of RHS."""
    say("case steps trap adaptive REF")
    wl = small_det()
    g, o = make_pair(wl, Q_REF)
    kw = dict(nsteps=2, cfl=0.05, bc_kinds=wl.bc_kinds, rk=1, chem=3, nchem_max=200)
    g.step(**kw)
    o.step(**kw)
    compare(g, o, ["CS", "T", "D", "P", "E", "YI"], allow_nan=True)


def test_step_skip_derive_same_state():
    """Stepping one call at a time with skip_derive on all but the last call leaves every field,
    the derived ones included, as one call of all the steps does."""
    say("case skip_derive")
    from eu3d_amd.api import Context
    from tests.common import init_state
    wl = small_det()
    a = Context(wl.xnode, wl.ynode, wl.znode, quirks=Q_FIXED)
    b = Context(wl.xnode, wl.ynode, wl.znode, quirks=Q_FIXED)
    init_state(a, wl)
    init_state(b, wl)
    kw = dict(cfl=0.05, bc_kinds=wl.bc_kinds, rk=1, chem=3, nchem_max=200)
    a.step(3, **kw)
    b.step(1, skip_derive=True, **kw)
    b.step(1, skip_derive=True, **kw)
    b.step(1, **kw)
    for f in ("CS", "T", "E", "D", "P", "YI", "DI", "MC", "WAV", "RGAS", "CP", "MA", "PT", "CMS"):
        x, y = a.get(f), b.get(f)
        assert np.array_equal(x, y, equal_nan=True), f


def test_steps_rk2_zn1d():
    say("case rk2 zn1d")
    wl = small_zn()
    g, o = make_pair(wl, Q_FIXED)
    kw = dict(nsteps=5, cfl=0.05, bc_kinds=wl.bc_kinds, rk=2, chem=3, nchem_max=200)
    g.step(**kw)
    o.step(**kw)
    compare(g, o, ["CS", "T", "D", "P", "E", "YI"])


def test_steps_imex_bubble_ref():
    say("case imex bubble REF")
    wl = small_bubble()
    g, o = make_pair(wl, Q_REF)
    kw = dict(nsteps=2, cfl=0.1, bc_kinds=(0,) * 6, rk=1, chem=1)
    g.step(**kw)
    o.step(**kw)
    compare(g, o, ["CS", "T", "D", "P", "E", "YI"])


def test_steps_imex_ragged_grid():
    """x rows of 70 cells (one full and one partial 64-cell segment) and 7 y rows (a partial
    block of the 64 x 4 interior blocks): the lanes and rows past the interior compute nothing
    that is stored."""
    say("case imex ragged 70x7x5")
    from eu3d_amd import configs
    wl = configs.bubble3d(ni=70, nj=7, nk=5, dx=1e-4, radius_cells=3.0, shock_cells=5, center_cells=(30, 3, 2),
                          seed=11)
    g, o = make_pair(wl, Q_FIXED)
    kw = dict(nsteps=2, cfl=0.1, bc_kinds=(0,) * 6, rk=1, chem=1)
    g.step(**kw)
    o.step(**kw)
    assert g.check_finite() == 0
    compare(g, o, ["CS", "T", "D", "P", "E", "YI"])


@pytest.mark.parametrize("chem", [1, 3])
def test_steps_zn1d_2000(chem):
    """The reference's CPU-sized case: the 1-D ZND detonation tube of 2000 cells."""
    say("case zn1d 2000 chem", chem)
    wl = small_zn(2000)
    g, o = make_pair(wl, Q_FIXED)
    kw = dict(nsteps=3, cfl=0.05, bc_kinds=wl.bc_kinds, rk=1, chem=chem, nchem_max=200)
    g.step(**kw)
    o.step(**kw)
    assert g.check_finite() == 0
    compare(g, o, ["CS", "T", "D", "P", "E", "YI"])


def test_step_failed_prediction_stores_advected_state():
    """A step clamped to a final time already reached has dt = 0, so every cell's sub-step count
    is ceil(0/0): the reference's Nchem <= 0 abort (src/Reaction.cpp:664-668). The fused
    prediction pass leaves such cells to k_adv_nochem, which must store Update_after_Adv and
    Update_after_CS for every interior cell (src/Flowfield.cpp:1140-1157,1246-1284); the error
    surfaces on the chemistry statistics. With skip_derive nothing else writes Ma, so a poisoned
    Ma shows which cells were stored."""
    say("case failed prediction, dt = 0")
    from eu3d_amd.api import Context, Eu3dError
    from tests.common import init_state
    wl = small_det()
    g = Context(wl.xnode, wl.ynode, wl.znode, quirks=Q_FIXED)
    init_state(g, wl)
    ma = g.get("MA")
    ma[2:-2, 2:-2, 2:-2] = -7.0
    g.set("MA", ma)
    cs0 = g.get("CS")
    with pytest.raises(Eu3dError, match="Nchem"):
        g.step(1, cfl=0.05, final_time=0.0, bc_kinds=wl.bc_kinds, rk=1, chem=3, nchem_max=200, skip_derive=True)
    with pytest.raises(Eu3dError, match="Nchem"):
        g.chem_stats()
    inner = (slice(2, -2),) * 3
    ma = g.get("MA")[inner]
    assert (ma != -7.0).all()
    u, v, w, c = (g.get(f)[inner] for f in ("U", "V", "W", "C"))
    np.testing.assert_allclose(ma, np.sqrt(u * u + v * v + w * w) / c, rtol=1e-13, atol=0)
    # RHS = dt L = 0, so the conserved state is the one before the step
    assert np.array_equal(g.get("CS")[inner], cs0[inner])
    assert (g.get("NCHEM")[inner] == 0).all()


def test_nchem_error_does_not_stick():
    """A failed step reports Nchem <= 0 once: the next clean step, the next chemistry call and the
    statistics after it return OK; a multi-step call whose later steps fail reports it once."""
    say("case nchem error is per call")
    from eu3d_amd.api import Context, Eu3dError
    from tests.common import init_state
    wl = small_det()
    g = Context(wl.xnode, wl.ynode, wl.znode, quirks=Q_FIXED)
    init_state(g, wl)
    kw = dict(cfl=0.05, bc_kinds=wl.bc_kinds, rk=1, chem=3, nchem_max=200)
    with pytest.raises(Eu3dError, match="Nchem"):
        g.step(1, final_time=0.0, **kw)
    g.step(1, **kw)                      # clean: dt > 0
    g.chem_stats()
    g.set_time(1e-8, 0.0)
    g.trapezoid_adaptive(200)
    g.chem_stats()
    # a final time 1e-9 s ahead: the first step of the call reaches it, the later ones have dt = 0
    dt, t = g.get_time()
    with pytest.raises(Eu3dError, match="Nchem"):
        g.step(3, final_time=t + 1e-9, **kw)
    g.step(1, **kw)
    g.chem_stats()


@pytest.mark.parametrize("field,value,ijk,flagged", [
    ("D", np.nan, (5, 7, 2), True),      # first interior plane
    ("T", np.inf, (25, 13, 11), True),   # last interior plane: the partial group of z planes
    ("T", np.nan, (2, 2, 7), True),
    ("D", np.nan, (0, 7, 5), False),     # a ghost cell is not checked
])
def test_check_finite_flags_one_cell(field, value, ijk, flagged):
    """eu3d_check_finite (src/Euler.cpp:319-331) flags one non-finite D or T anywhere in the
    interior (24 x 12 x 10 cells, two ghost layers) and ignores the ghost cells."""
    say("case check_finite", field, value, ijk)
    from eu3d_amd.api import Context
    from tests.common import init_state
    wl = small_bubble()
    g = Context(wl.xnode, wl.ynode, wl.znode, quirks=Q_FIXED)
    init_state(g, wl)
    assert g.check_finite() == 0
    a = g.get(field)
    assert a.shape == (28, 16, 14)
    a[ijk] = value
    g.set(field, a)
    assert (g.check_finite() != 0) == flagged


@pytest.mark.parametrize("shape", [(1, 1, 200), (8, 64, 64)])
def test_steps_imex_thin_grid(shape):
    """Grids that write more CFL partials (one per 64x4 block per k layer) than n3/64 + 64: a
    1x1x200 z tube and an 8x64x64 slab. The partial buffer is sized by the block count, so the
    step's dt and state match the oracle."""
    say("case imex thin grid", shape)
    from eu3d_amd import configs
    from eu3d_amd.mech import load_default
    m = load_default()
    ni, nj, nk = shape
    dx = 1e-4
    x, y, z = configs.make_nodes(ni, dx), configs.make_nodes(nj, dx), configs.make_nodes(nk, dx)
    sh = (ni + 4, nj + 4, nk + 4)
    X, Y, Z = np.meshgrid(*(np.arange(n, dtype=float) for n in sh), indexing="ij")
    wave = np.sin(0.21 * X + 0.13 * Y + 0.17 * Z)
    prims = {"T": 1200.0 + 400.0 * wave, "P": 101325.0 * (2.0 + wave), "U": 80.0 * wave,
             "V": -40.0 * wave, "W": 60.0 * np.cos(0.17 * Z)}
    phi = 0.5 * (1 + wave)[..., None]
    prims["YI"] = phi * configs.premix(m) + (1 - phi) * configs.burnt(m)
    wl = configs.Workload("thin", x, y, z, prims, (1,) * 6, {}, 2)
    g, o = make_pair(wl, Q_FIXED)
    kw = dict(nsteps=2, cfl=0.3, bc_kinds=wl.bc_kinds, rk=1, chem=1)
    g.step(**kw)
    o.step(**kw)
    assert g.check_finite() == 0
    dg, do = g.get_time(), o.get_time()
    assert abs(dg[0] - do[0]) <= 1e-12 * do[0] and abs(dg[1] - do[1]) <= 1e-12 * do[1]
    compare(g, o, ["CS", "T", "D", "P", "E", "YI"])
