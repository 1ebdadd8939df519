"""Pins for the CPU oracle (the reference check of the HIP path).

The reference ships no tests and its Flowfield/Reaction translation units need <mpi.h>, which
this image lacks, so the oracle is pinned by: the reference's own Differ.hpp/TimeAdv.cpp/
Function.cpp built by oracle/Makefile.ref (bit-exact), the reference's shipped initial-state
output (every printed value of output_8/resultsIMEX0_*.dat, tests/golden/output8/, and the
sampled points of tests/golden/output8_initial.json), and closed forms / invariants chosen so a
dropped term, wrong sign or transposed index fails.
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

from eu3d_amd import configs
from eu3d_amd.mech import R_UNIV, load_default, parse_mechanism
from oracle import OracleFlow, load

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
REF = "/root/reference"
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libref.so")
M = load_default()
Q_REF, Q_FIXED = 0x7FF, 0
dp = C.POINTER(C.c_double)


def p(a):
    return np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(dp)


def flow(ni=6, nj=5, nk=4, quirks=Q_FIXED, dx=1e-4):
    x, y, z = configs.make_nodes(ni, dx), configs.make_nodes(nj, dx), configs.make_nodes(nk, dx)
    return OracleFlow(x, y, z, quirks=quirks)


def random_prims(o, seed=0, T=(800.0, 2600.0)):
    rng = np.random.default_rng(seed)
    sh = o.shape3
    Y = rng.uniform(0.01, 1.0, sh + (M.ns,))
    Y /= Y.sum(-1, keepdims=True)
    prims = {"T": rng.uniform(*T, sh), "P": rng.uniform(0.5, 5.0, sh) * 101325.0,
             "U": rng.uniform(-300, 300, sh), "V": rng.uniform(-300, 300, sh), "W": rng.uniform(-300, 300, sh),
             "YI": Y}
    for k, v in prims.items():
        o.set(k, v)
    o.field_initial()
    return prims


# ------------------------------------------------------------------ mechanism
@pytest.mark.skipif(not os.path.isdir(REF), reason="reference tree not mounted")
def test_mechanism_json_matches_reference_files():
    m = parse_mechanism(os.path.join(REF, "input", "H2_JA919.txt"), os.path.join(REF, "input", "thermo.dat"))
    for k in ("stoi_f", "stoi_b", "af", "bf", "eaf", "react_tb", "coeff0", "coeff1", "mw"):
        assert np.array_equal(getattr(m, k), getattr(M, k)), k


def test_mechanism_known_values():
    # species order and a few parsed values (input/H2_JA919.txt:5,49,71-76; input/thermo.dat:16-20)
    assert M.species == ["H", "O", "H2O", "OH", "O2", "H2", "H2O2", "HO2", "N2"]
    assert M.af[0] == 1.0e14 and M.eaf[0] == 5.603101560e4
    assert M.react_tb[2, 8] == 16.0 and M.react_tb[5, 6] == 2.0   # "16." parsed into int 16
    assert M.coeff0[2, 0] == 2.5 and M.mw[0] == 1.00794


def test_mechanism_element_balance():
    # every reaction conserves H and O atoms (catches a transposed or shifted stoichiometry row)
    H = np.array([1, 0, 2, 1, 0, 2, 2, 1, 0])
    O = np.array([0, 1, 1, 1, 2, 0, 2, 2, 0])
    d = M.stoi_b - M.stoi_f
    assert (H @ d == 0).all() and (O @ d == 0).all()
    assert np.allclose(M.mw @ d, 0.0, atol=1e-9)


# ------------------------------------------------------------------ thermo
def test_cp_h_closed_forms():
    o = flow()
    ri = M.ri
    for T in (300.0, 650.0, 999.0):
        assert o.getcpi(T, ri[0], 0) == pytest.approx(2.5 * ri[0], rel=1e-15)   # atomic H: c_p = 5/2 R
        assert o.gethi(T, ri[0], 0) == pytest.approx(ri[0] * (2.5 * T + M.coeff0[7, 0]), rel=1e-14)
    # c_p = dh/dT and c_p/T = ds/dT within each range (central differences)
    for sp in range(M.ns):
        for T in (400.0, 800.0, 1500.0, 2800.0):
            h = 1e-3
            dh = (o.gethi(T + h, R_UNIV, sp) - o.gethi(T - h, R_UNIV, sp)) / (2 * h)
            ds = (o.getsi(T + h, R_UNIV, sp) - o.getsi(T - h, R_UNIV, sp)) / (2 * h)
            cp = o.getcpi(T, R_UNIV, sp)
            assert dh == pytest.approx(cp, rel=1e-7), (sp, T)
            assert ds == pytest.approx(cp / T, rel=1e-7), (sp, T)


def test_c4_quirk_jump_at_1000K():
    """FIXED h is continuous at the 1000 K breakpoint to the fit's accuracy; REF (quirk C4) jumps
    by exactly R (b1_low - b1_high) on top of that."""
    fx, rf = flow(quirks=Q_FIXED), flow(quirks=Q_REF)
    for sp in range(M.ns):
        lo = fx.gethi(1000 - 1e-9, R_UNIV, sp)
        hi_fixed = fx.gethi(1000.0, R_UNIV, sp)
        hi_ref = rf.gethi(1000.0, R_UNIV, sp)
        scale = max(abs(lo), R_UNIV * 1000)
        assert abs(hi_fixed - lo) < 2e-3 * scale
        assert hi_ref - hi_fixed == pytest.approx(R_UNIV * (M.coeff0[7, sp] - M.coeff1[7, sp]), rel=1e-9, abs=1e-6)


def test_newton_pure_H_one_step():
    """Pure atomic H below 1000 K has constant c_p, so e(T) is linear and the Newton lands on T
    after one step; the second step confirms |dT| < 1e-6."""
    lib = load()
    o = flow()
    Y = np.zeros(M.ns)
    Y[0] = 1.0
    for T in (350.0, 700.0):
        E = 1000 * (1.5 * M.ri[0] * T + M.ri[0] * M.coeff0[7, 0])
        it = C.c_int()
        Tn = lib.orc_get_temp(o.om.ref(), 0, 900.0, E, p(Y), C.byref(it))
        assert Tn == pytest.approx(T, rel=1e-13) and it.value == 2


def test_newton_mixture_inverts_energy():
    lib = load()
    o = flow()
    rng = np.random.default_rng(1)
    for _ in range(50):
        Y = rng.uniform(0, 1, M.ns)
        Y /= Y.sum()
        T = rng.uniform(400, 2900)
        e = 1000 * sum(Y[s] * (o.gethi(T, M.ri[s], s) - M.ri[s] * T) for s in range(M.ns))
        Tn = lib.orc_get_temp(o.om.ref(), 0, T * 1.05, e, p(Y), None)
        assert Tn == pytest.approx(T, abs=1e-6)


def test_sum_order_matches_reference_build():
    ref = _ref()
    lib = load()
    rng = np.random.default_rng(2)
    a, b = rng.normal(size=17) * 1e3, rng.normal(size=17)
    for flag in range(4):
        assert lib.orc_sum(flag, p(a), p(b), 17) == ref.ref_sum(flag, p(a), p(b), 17)


# ------------------------------------------------------------------ reconstruction
def _ref():
    if not os.path.exists(REF_LIB):
        if os.path.isdir(REF):
            os.system(f"make -s -f {ROOT}/oracle/Makefile.ref")
    if not os.path.exists(REF_LIB):
        pytest.skip("oracle/_ref not built (needs the reference tree)")
    lib = C.CDLL(REF_LIB)
    lib.ref_muscl.argtypes = [C.c_int] * 6 + [dp, dp, C.c_int]
    lib.ref_ee.argtypes = [C.c_int, C.c_double] + [C.c_int] * 4 + [dp] * 7
    lib.ref_sum.argtypes = [C.c_int, dp, dp, C.c_int]
    lib.ref_sum.restype = C.c_double
    return lib


@pytest.mark.parametrize("limit", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("order", [1, 2])
def test_muscl_bit_exact_vs_reference_build(order, limit):
    ref = _ref()
    lib = load()
    o = flow(7, 5, 6)
    rng = np.random.default_rng(limit)
    fi = rng.normal(size=o.shape3)
    fi[2:4] = 1.5          # flat patches exercise r = 0 and 1/r = inf
    for d in (1, 2, 3):
        a = np.zeros(o.shape3 + (2,))
        b = np.zeros(o.shape3 + (2,))
        lib.orc_muscl(order, d, C.byref(o.f), a.ctypes.data_as(dp), p(fi), limit)
        ref.ref_muscl(order, d, o.ni, o.nj, o.nk, 2, p(fi), b.ctypes.data_as(dp), limit)
        assert np.array_equal(np.isnan(a), np.isnan(b))
        assert np.array_equal(np.nan_to_num(a), np.nan_to_num(b)), (order, limit, d)


def test_muscl_closed_forms():
    lib = load()
    o = flow(8, 3, 3)
    out = np.zeros(o.shape3 + (2,))
    # constant field -> both face states equal the constant
    lib.orc_muscl(2, 1, C.byref(o.f), out.ctypes.data_as(dp), p(np.full(o.shape3, 3.25)), 0)
    assert np.all(out[1:-2, 1:-2, 1:-2] == 3.25)
    # linear field -> r = A1/(A0 + 1e-12) just below 1, phi = r and f_L = f_i + r A0/2,
    # f_R = f_{i+1} - r A2/2
    lin = np.broadcast_to(np.arange(o.shape3[0], dtype=float)[:, None, None] * 0.5, o.shape3).copy()
    lib.orc_muscl(2, 1, C.byref(o.f), out.ctypes.data_as(dp), p(lin), 0)
    i = np.arange(1, o.shape3[0] - 2)
    r = 0.5 / (0.5 + 1e-12)
    assert np.allclose(out[1:-2, 2, 2, 0], lin[i, 2, 2] + 0.25 * r, rtol=0, atol=1e-15)
    assert np.allclose(out[1:-2, 2, 2, 1], lin[i + 1, 2, 2] - 0.25 * r, rtol=0, atol=1e-15)
    # TVD bound for minmod on rough data
    rng = np.random.default_rng(5)
    f = rng.normal(size=o.shape3)
    lib.orc_muscl(2, 1, C.byref(o.f), out.ctypes.data_as(dp), p(f), 0)
    for ii in range(1, o.shape3[0] - 2):
        lo = np.minimum(f[ii], f[ii + 1])[1:-2, 1:-2]
        hi = np.maximum(f[ii], f[ii + 1])[1:-2, 1:-2]
        for s in (0, 1):
            v = out[ii, 1:-2, 1:-2, s]
            assert np.all(v >= lo - 1e-12) and np.all(v <= hi + 1e-12)


# ------------------------------------------------------------------ flux and EE
def _uniform(o, T, P, U, V=0.0, W=0.0, Y=None):
    sh = o.shape3
    Y = configs.air(M) if Y is None else Y
    for k, v in (("T", T), ("P", P), ("U", U), ("V", V), ("W", W)):
        o.set(k, np.full(sh, float(v)))
    o.set("YI", np.broadcast_to(Y, sh + (M.ns,)))
    o.field_initial()


@pytest.mark.parametrize("U", [0.0, 120.0, -200.0, 900.0, -1200.0])
def test_ausm_uniform_state_is_physical_flux(U):
    """A uniform state gives the exact physical flux: supersonic by B = M, G = 1 or 0; subsonic
    because B1 + B2 = M and G1 + G2 = 1. Hence every RHS is zero."""
    o = flow(6, 4, 4)
    _uniform(o, 600.0, 2e5, U, 30.0, -40.0)
    o.set_time(1e-7)
    o.advection(1, 0)
    c = (3, 3, 3)
    rho, u, v, w = o.get("D")[c], U, 30.0, -40.0
    H = o.get("H")[c]
    pr = o.get("P")[c]
    F = o.get("F")[c]
    Y = o.get("YI")[c]
    ns = M.ns
    assert np.allclose(F[:ns], rho * u * Y, rtol=1e-12, atol=1e-12 * rho * 1000)
    assert F[ns] == pytest.approx(rho * u * u + pr, rel=1e-12)
    assert F[ns + 1] == pytest.approx(rho * u * v, rel=1e-12, abs=1e-9)
    assert F[ns + 3] == pytest.approx(u * (rho * H + 0.5 * rho * (u * u + v * v + w * w)), rel=1e-12, abs=1e-6)
    assert np.abs(o.get("RHS")[2:-2, 2:-2, 2:-2]).max() < 1e-9 * max(pr, 1.0) * 1e-7 / 1e-4


def test_ausm_mirror_antisymmetry():
    """Reflecting the flow (u -> -u, x -> -x) reverses the mass and energy fluxes and keeps the
    momentum flux."""
    o1, o2 = flow(8, 3, 3), flow(8, 3, 3)
    rng = np.random.default_rng(3)
    sh = o1.shape3
    T = rng.uniform(500, 1500, sh[0])
    P = rng.uniform(1e5, 3e5, sh[0])
    U = rng.uniform(-200, 200, sh[0])
    for o, flip in ((o1, 1), (o2, -1)):
        idx = np.arange(sh[0]) if flip == 1 else np.arange(sh[0])[::-1]
        o.set("T", np.broadcast_to(T[idx][:, None, None], sh))
        o.set("P", np.broadcast_to(P[idx][:, None, None], sh))
        o.set("U", np.broadcast_to(flip * U[idx][:, None, None], sh))
        o.set("V", np.zeros(sh))
        o.set("W", np.zeros(sh))
        o.set("YI", np.broadcast_to(configs.air(M), sh + (M.ns,)))
        o.field_initial()
        o.set_time(1e-8)
        o.advection(1, 0)
    F1, F2 = o1.get("F")[:, 2, 2], o2.get("F")[:, 2, 2]
    n = sh[0]
    ns = M.ns
    for i in range(1, n - 2):      # face i between i, i+1 maps to face n-2-i
        j = n - 2 - i
        assert F1[i, ns + 3] == pytest.approx(-F2[j, ns + 3], rel=1e-11, abs=1e-6)
        assert F1[i, ns] == pytest.approx(F2[j, ns], rel=1e-11)
        assert np.allclose(F1[i, :ns], -F2[j, :ns], rtol=1e-11, atol=1e-12)


def test_ee_bit_exact_vs_reference_build():
    ref = _ref()
    o = flow(5, 4, 6)
    random_prims(o, 4)
    o.set_time(3e-8)
    o.fill_ghosts((1,) * 6)
    o.advection(1, 0)
    out = np.zeros(o.shape3 + (M.ns + 4,))
    ref.ref_ee(M.ns, 3e-8, o.ni, o.nj, o.nk, 2, p(o.xnode), p(o.ynode), p(o.znode),
               p(o.get("F")), p(o.get("G")), p(o.get("Q")), out.ctypes.data_as(dp))
    assert np.array_equal(out[2:-2, 2:-2, 2:-2], o.get("RHS")[2:-2, 2:-2, 2:-2])


def test_periodic_domain_conserves_mass_and_energy():
    o = flow(6, 5, 4)
    random_prims(o, 5)
    o.set_time(1e-8)
    o.fill_ghosts((3,) * 6)
    o.advection(1, 0)
    R = o.get("RHS")[2:-2, 2:-2, 2:-2]
    cs = o.get("CS")[2:-2, 2:-2, 2:-2]
    for c in list(range(M.ns)) + [M.ns, M.ns + 1, M.ns + 2, M.ns + 3]:
        assert abs(R[..., c].sum()) < 1e-11 * np.abs(cs[..., c]).sum() + 1e-14


@pytest.mark.parametrize("shape", [(6, 5, 4), (6, 1, 4), (1, 1, 7)])
def test_cfl_brute_force(shape):
    """REF: cfl min(dx/(|u|+c), dy/(|v|+c)) (z ignored, src/Flowfield.cpp:254). FIXED: the unsplit
    three-direction bound cfl / max sum_d (|u_d|+c)/dx_d over the directions with more than one
    cell, never above cfl times the per-direction minimum."""
    for q in (Q_FIXED, Q_REF):
        o = flow(*shape, quirks=q)
        random_prims(o, 6)
        U, V, W, Cs = (o.get(k)[2:-2, 2:-2, 2:-2] for k in ("U", "V", "W", "C"))
        d = 1e-4
        if q == Q_FIXED:
            rate = sum((np.abs(X) + Cs) / d for X, n in zip((U, V, W), shape) if n > 1)
            b = 1.0 / rate
        else:
            b = np.minimum(d / (np.abs(U) + Cs), d / (np.abs(V) + Cs))
        o.cfl(0.5, 1.0)
        assert o.get_time()[0] == pytest.approx(0.5 * b.min(), rel=1e-12)
        if q == Q_FIXED:
            per_dir = min((d / (np.abs(X) + Cs)).min() for X, n in zip((U, V, W), shape) if n > 1)
            assert o.get_time()[0] <= 0.5 * per_dir * (1 + 1e-12)


# ------------------------------------------------------------------ kinetics
def test_wdot_conserves_mass_and_elements():
    H = np.array([1, 0, 2, 1, 0, 2, 2, 1, 0]) / M.mw
    O = np.array([0, 1, 1, 1, 2, 0, 2, 2, 0]) / M.mw
    for q in (Q_FIXED, Q_REF):
        o = flow(4, 3, 3, quirks=q)
        random_prims(o, 7, T=(1100.0, 2800.0))
        o.diagonalized()
        w = o.get("CMS")[2:-2, 2:-2, 2:-2]
        scale = np.abs(w).max(axis=-1)
        assert np.all(np.abs(w.sum(-1)) <= 1e-11 * scale)
        assert np.all(np.abs(w @ H) <= 1e-11 * scale) and np.all(np.abs(w @ O) <= 1e-11 * scale)


def test_jacobian_diagonal_matches_finite_difference():
    """With dT/drho_s = 0 the FIXED MD diagonal is d(CMS_s)/d(rho_s) at fixed T (rates depend on
    rho_s through the concentrations and the third-body sums)."""
    o = flow(1, 1, 1)
    random_prims(o, 8, T=(1500.0, 1600.0))
    base_di = o.get("DI").copy()
    o.set("PT", np.zeros_like(o.get("PT")))
    o.diagonalized()
    md = o.get("MD")[2, 2, 2]
    for s in range(M.ns):
        h = 1e-6 * base_di[2, 2, 2, s]
        vals = []
        for sgn in (1, -1):
            di = base_di.copy()
            di[2, 2, 2, s] += sgn * h
            o.set("DI", di)
            o.set("MC", di / M.mw * 1000)
            o.diagonalized()
            vals.append(o.get("CMS")[2, 2, 2, s])
        fd = (vals[0] - vals[1]) / (2 * h)
        assert fd == pytest.approx(md[s], rel=2e-5, abs=1e-6 * abs(md).max())


def test_ref_c2_gives_zero_rate_and_single_substep():
    o = flow(3, 3, 3, quirks=Q_REF)
    random_prims(o, 9, T=(1500.0, 2500.0))
    o.set_time(1e-7)
    o.trapezoid_prediction(0)
    assert (o.get("NCHEM")[2:-2, 2:-2, 2:-2] == 1).all()
    o.trapezoid_adaptive(0)
    assert np.abs(o.get("CMS")[2:-2, 2:-2, 2:-2]).max() == 0.0


def test_single_reaction_trapezoid_by_hand():
    """A one-reaction mechanism H2 + O2 <=> 2 OH: the point-implicit update computed by hand."""
    m = load_default()
    import copy
    one = copy.deepcopy(m)
    one.nr = 1
    iH2, iO2, iOH = m.index("H2"), m.index("O2"), m.index("OH")
    one.stoi_f = np.zeros((m.ns, 1)); one.stoi_f[iH2, 0] = 1; one.stoi_f[iO2, 0] = 1
    one.stoi_b = np.zeros((m.ns, 1)); one.stoi_b[iOH, 0] = 2
    one.af = np.array([1.7e13]); one.bf = np.array([0.0]); one.eaf = np.array([4.78e4])
    one.react_tb = np.zeros((m.ns, 1))
    x = configs.make_nodes(1, 1e-4)
    o = OracleFlow(x, x, x, mech=one, quirks=Q_FIXED)
    T = 1800.0
    Y = np.zeros(m.ns); Y[iH2] = 0.1; Y[iO2] = 0.6; Y[iOH] = 0.01; Y[m.index("N2")] = 0.29
    _uniform(o, T, 2 * 101325.0, 0.0, Y=Y)
    dt = 2e-7
    o.set_time(dt)
    di0 = o.get("DI")[2, 2, 2].copy()
    o.trapezoid()
    # by hand, reaction units (mol/cm^3, g/cm^3)
    Dr = di0 * 1e-3
    Mc = di0 / m.mw * 1000 * 1e-6
    g = np.array([o.gethi(T, R_UNIV, s) - o.getsi(T, R_UNIV, s) * T for s in range(m.ns)])
    kf = 1.7e13 * np.exp(-4.78e4 / (1.987 * T))
    dG = 2 * g[iOH] - g[iH2] - g[iO2]
    Kc = np.exp(-dG / (R_UNIV * T))     # dnu = 0
    kb = kf / Kc
    rf, rb = Mc[iH2] * Mc[iO2], Mc[iOH] ** 2
    Q = np.zeros(m.ns); P = np.zeros(m.ns)
    Q[iOH] = 2 * kf * rf; P[iOH] = 2 * kb * rb / Mc[iOH]
    Q[iH2] = kb * rb; P[iH2] = kf * rf / Mc[iH2]
    Q[iO2] = kb * rb; P[iO2] = kf * rf / Mc[iO2]
    Q *= m.mw
    expect = ((1 - dt / 2 * P) * Dr + dt * Q) / (1 + dt / 2 * P) * 1e3
    assert np.allclose(o.get("DI")[2, 2, 2], expect, rtol=1e-12, atol=1e-20)
    assert o.get("DI")[2, 2, 2, iOH] > di0[iOH]


def test_ignition_reaches_burnt_state():
    """FIXED adaptive trapezoid at constant volume ignites a 1500 K premix and settles near the
    constant-volume burnt temperature, with the water mass fraction of near-complete combustion."""
    x = configs.make_nodes(1, 1e-4)
    o = OracleFlow(x, x, x, quirks=Q_FIXED)
    _uniform(o, 1500.0, 101325.0, 0.0, Y=configs.premix(M))
    t, dt = 0.0, 1e-7
    for _ in range(1500):
        o.set_time(dt, t)
        o.trapezoid_adaptive(10000)
        t += dt
    T = o.get("T")[2, 2, 2]
    assert 2800 < T < 3300
    y = o.get("YI")[2, 2, 2]
    assert 0.18 < y[M.index("H2O")] < 0.26
    assert abs(y.sum() - 1) < 1e-6


def test_imex_small_dt_limit():
    """dt -> 0: the IMEX update reduces to CS + RHS (chemistry terms vanish with dt)."""
    o = flow(3, 3, 3)
    random_prims(o, 10, T=(1200, 2000))
    o.set_time(1e-14)
    o.fill_ghosts((1,) * 6)
    o.advection(1, 0)
    cs0, rhs = o.get("CS").copy(), o.get("RHS").copy()
    o.diagonalized()
    o.update_imex()
    d = (o.get("CS") - (cs0 + rhs))[2:-2, 2:-2, 2:-2]
    assert np.abs(d).max() <= 1e-12 * np.abs(cs0).max()


# ------------------------------------------------------------------ reference output fixture
def test_field_initial_matches_shipped_output():
    """Density of the reference's own initial state (output_8, 6 significant digits) from its
    T, P and Y through FieldInitial (src/Flowfield.cpp:201-203); N2 is the complement of the
    printed Y1..Y7 (HO2 is zero in that state). Also pins the shock relations used by the
    workload generator against the printed post-shock state."""
    fx = json.load(open(os.path.join(HERE, "golden", "output8_initial.json")))
    pts = fx["points"]
    n = len(pts)
    x = configs.make_nodes(n, 1e-3)
    y = configs.make_nodes(1, 1e-3)
    o = OracleFlow(x, y, y, quirks=Q_REF)
    sh = o.shape3
    Tf, Pf = np.full(sh, 1000.0), np.full(sh, 101325.0)
    Yf = np.zeros(sh + (M.ns,))
    Yf[..., M.index("N2")] = 1.0
    for i, q in enumerate(pts):
        Tf[i + 2], Pf[i + 2] = q["T"], q["P"]
        Yv = np.array([q[f"Y{k}"] for k in range(1, 8)] + [0.0, 0.0])
        Yv[8] = 1.0 - Yv.sum()
        Yf[i + 2] = Yv
    for k, v in (("T", Tf), ("P", Pf), ("U", np.zeros(sh)), ("V", np.zeros(sh)), ("W", np.zeros(sh)), ("YI", Yf)):
        o.set(k, v)
    o.field_initial()
    D = o.get("D")[2:-2, 2, 2]
    Dref = np.array([q["D"] for q in pts])
    assert np.allclose(D, Dref, rtol=6e-6)
    up, pp, tp = configs.shock_post(3.0, 1000.0, 101325.0)
    post = [q for q in pts if q["class"] == 0]
    assert post and all(abs(q["U"] - up) < 5e-6 * up and abs(q["P"] - pp) < 5e-6 * pp and
                        abs(q["T"] - tp) < 5e-6 * tp for q in post)


OUTPUT8 = os.path.join(HERE, "golden", "output8")
TECPLOT_VARS = "X,Y,Z,T,D,P,U,V,Ma,Y1,Y2,Y3,Y4,Y5,Y6,Y7,L,WaitTime".split(",")


def output8_values(rank):
    """The printed values of the reference's output_8/resultsIMEX0_<rank>.dat (18 blocks of
    61 x 41 x 2 values, i fastest), as the text lines."""
    import gzip
    with gzip.open(os.path.join(OUTPUT8, f"resultsIMEX0_{rank}.dat.gz"), "rt") as f:
        lines = f.read().split("\n")
    return lines[:3], [v for v in lines[3:] if v.strip()]


def test_reference_initial_state_whole_field():
    """The reference's own initial state, every printed value of its 8-rank run: for each rank's
    block (m_block 4 x 2 x 1, rank = bx + 4 by, src/Euler.cpp:213-215) the nodes of Mesh::Initial,
    ReactionInitial's H2 cylinder and FieldInitial's Mach-3 shock (configs.rsbi_reference), the
    oracle's FieldInitial (D from the mixture gas constant), Ma = 0 and Nchem = 1 as the reference
    leaves them at time 0, printed with 6 significant digits (std::ostream's default, C's %g),
    equal to output_8/resultsIMEX0_<rank>.dat string for string: 8 x 18 x 5002 values."""
    decks = json.load(open(os.path.join(OUTPUT8, "decks.json")))
    I = (slice(2, -2),) * 3
    for rank in range(8):
        wl = configs.rsbi_reference(decks, (rank % 4, rank // 4, 0))
        o = OracleFlow(wl.xnode, wl.ynode, wl.znode, quirks=Q_REF)
        for k in ("T", "P", "U", "V", "W", "YI"):
            o.set(k, wl.prims[k])
        o.field_initial()
        X, Y, Z = np.meshgrid(wl.xnode, wl.ynode, wl.znode, indexing="ij")
        Yi = o.get("YI")
        cols = ([X, Y, Z] + [o.get(f) for f in ("T", "D", "P", "U", "V", "MA")] + [Yi[..., s] for s in range(7)]
                + [o.get("NCHEM"), np.zeros_like(X)])
        head, ref = output8_values(rank)
        ni, nj, nk = (len(a) - 4 for a in (wl.xnode, wl.ynode, wl.znode))
        assert head[1].split() == ["ZONE", f"I={ni}", f"J={nj}", f"K={nk}"]
        n = ni * nj * nk
        assert len(ref) == 18 * n
        for v, (name, col) in enumerate(zip(TECPLOT_VARS, cols)):
            mine = ["%g" % x for x in col[I].transpose(2, 1, 0).reshape(-1)]
            bad = [i for i in range(n) if mine[i] != ref[v * n + i]]
            assert not bad, (rank, name, len(bad), [(mine[i], ref[v * n + i]) for i in bad[:3]])


# ------------------------------------------------------------------ ghost kinds and dT/drho_s
PRIM_FIELDS = ("U", "V", "W", "P", "D", "T", "H", "GAMMA", "E", "YI")


def test_ghost_kinds_closed_forms():
    """Outflow copies the first interior cell, periodic copies the cell one period away, reflect
    mirrors across the face with the normal velocity negated (x faces over the interior j, k)."""
    o = flow(ni=6, nj=5, nk=4)
    random_prims(o, seed=3)
    bc, ni = 2, 6
    for kind in (1, 2, 3):
        random_prims(o, seed=10 + kind)
        before = {f: o.get(f) for f in PRIM_FIELDS}
        o.fill_ghosts((kind, kind, 4, 4, 4, 4))
        after = {f: o.get(f) for f in PRIM_FIELDS}
        J, K = slice(bc, bc + 5), slice(bc, bc + 4)
        for f in PRIM_FIELDS:
            sign = -1.0 if (kind == 2 and f == "U") else 1.0
            for g in range(bc):
                lo, hi = bc - 1 - g, ni + bc + g              # ghosts at distance g + 1 from the faces
                if kind == 1:
                    slo, shi = bc, ni + bc - 1
                elif kind == 2:
                    slo, shi = bc + g, ni + bc - 1 - g
                else:
                    slo, shi = ni + bc - 1 - g, bc + g
                assert np.array_equal(after[f][lo, J, K], sign * before[f][slo, J, K]), (kind, f, g, "lo")
                assert np.array_equal(after[f][hi, J, K], sign * before[f][shi, J, K]), (kind, f, g, "hi")
        # interior untouched
        for f in PRIM_FIELDS:
            assert np.array_equal(after[f][bc:bc + ni, J, K], before[f][bc:bc + ni, J, K])


def _nasa_mix(T, Y):
    """Mixture R, c_p, h and e [J/kg(/K)] from the NASA-9 fits, computed here in numpy (range by
    T < 1000 as src/Reaction.cpp:893-926, FIXED: the high range keeps its own b1)."""
    ri = R_UNIV / M.mw                                     # kJ/kg/K
    lo = (T < 1000.0)[..., None, None]
    a = np.where(lo, M.coeff0.T, M.coeff1.T)               # (..., ns, 9)
    t = T[..., None]
    cp = a[..., 0] / t ** 2 + a[..., 1] / t + a[..., 2] + a[..., 3] * t + a[..., 4] * t ** 2 + \
        a[..., 5] * t ** 3 + a[..., 6] * t ** 4
    h = -a[..., 0] / t + a[..., 1] * np.log(t) + a[..., 2] * t + a[..., 3] * t ** 2 / 2 + \
        a[..., 4] * t ** 3 / 3 + a[..., 5] * t ** 4 / 4 + a[..., 6] * t ** 5 / 5 + a[..., 7]
    R = (Y * ri).sum(-1) * 1000
    return R, (Y * ri * cp).sum(-1) * 1000, (Y * ri * h).sum(-1) * 1000, (Y * ri * (h - t)).sum(-1) * 1000


def test_rsbi_ghosts_closed_forms():
    """FieldBoundary_3dRSBI (src/Flowfield.cpp:559-773) in closed form: every face mirrors
    (ghost at distance g + 1 takes the cell at distance g from the face inside; no sign change);
    x faces over the interior j, k, y faces over all i and the interior k, z faces over all i, j,
    in that order; the y-low face copies U, V, W, P, T, Y and Mc and recomputes R, rho = P / (R T),
    c_p, h, e and gamma = c_p / (c_p - R) from the NASA fits, the other faces copy rho, H, gamma
    and E too."""
    ni, nj, nk, bc = 6, 5, 4, 2
    o = flow(ni=ni, nj=nj, nk=nk)
    random_prims(o, seed=21)
    o.fill_ghosts((0,) * 6)
    f = {k: o.get(k) for k in PRIM_FIELDS + ("MC", "RGAS", "CP")}
    NX, NY, NZ = ni + 2 * bc, nj + 2 * bc, nk + 2 * bc
    I, J, K = slice(bc, ni + bc), slice(bc, nj + bc), slice(bc, nk + bc)
    copied = ("U", "V", "W", "P", "D", "T", "H", "GAMMA", "E", "YI")
    for g in range(bc):
        for k in copied:   # x faces, interior j and k
            assert np.array_equal(f[k][bc - 1 - g, J, K], f[k][bc + g, J, K]), ("x-", k, g)
            assert np.array_equal(f[k][ni + bc + g, J, K], f[k][ni + bc - 1 - g, J, K]), ("x+", k, g)
        for k in copied:   # y-high face, all i, interior k
            assert np.array_equal(f[k][:, nj + bc + g, K], f[k][:, nj + bc - 1 - g, K]), ("y+", k, g)
        for k in ("U", "V", "W", "P", "T", "YI", "MC"):   # y-low face copies
            assert np.array_equal(f[k][:, bc - 1 - g, K], f[k][:, bc + g, K]), ("y-", k, g)
        for k in copied:   # z faces, all i and j
            assert np.array_equal(f[k][:, :, bc - 1 - g], f[k][:, :, bc + g]), ("z-", k, g)
            assert np.array_equal(f[k][:, :, nk + bc + g], f[k][:, :, nk + bc - 1 - g]), ("z+", k, g)
    # y-low recompute from the copied T, P, Y
    sl = (slice(0, NX), slice(0, bc), K)
    T, P, Y = f["T"][sl], f["P"][sl], f["YI"][sl]
    R, cp, h, e = _nasa_mix(T, Y)
    for name, want, got in (("RGAS", R, f["RGAS"][sl]), ("D", P / (T * R), f["D"][sl]), ("CP", cp, f["CP"][sl]),
                            ("H", h, f["H"][sl]), ("E", e, f["E"][sl]),
                            ("GAMMA", cp / (cp - R), f["GAMMA"][sl])):
        assert np.allclose(got, want, rtol=1e-12, atol=0), (name, np.abs(got / want - 1).max())
    # the recompute is not a copy: the source cells hold other values
    assert not np.array_equal(f["D"][:, bc - 1, K], f["D"][:, bc, K])


def test_partial_t_matches_finite_difference():
    """GetPartial_T (FIXED units): dT/drho_s at constant rho u and rho E equals the central finite
    difference of the energy Newton, (1/2 |u|^2 - e_s) / (rho c_v)."""
    o = flow(ni=4, nj=3, nk=3)
    random_prims(o, seed=5, T=(900.0, 2400.0))
    o.update_after_cs()
    o.partial_t()
    pt = o.get("PT")
    cs0, T0 = o.get("CS"), o.get("T")
    cell = (3, 3, 3)
    for s in range(M.ns):
        d = 0.5 / abs(pt[cell + (s,)])          # a perturbation that moves T by about 0.5 K
        Ts = []
        for sgn in (1, -1):
            cs = cs0.copy()
            cs[cell + (s,)] += sgn * d
            o.set("CS", cs)
            o.set("T", T0)
            o.update_after_cs()
            Ts.append(o.get("T")[cell])
        fd = (Ts[0] - Ts[1]) / (2 * d)
        assert abs(fd - pt[cell + (s,)]) <= 2e-6 * abs(pt[cell + (s,)]), (s, fd, pt[cell + (s,)])
    o.set("CS", cs0)
    o.set("T", T0)


@pytest.mark.parametrize("d", [3.0, 5.0])
def test_constant_division_is_correctly_rounded(d):
    """The device kernels divide the NASA-9 integral terms by 3 and 5 as x * RN(1/d) refined by one
    FMA residual step. That equals IEEE division on 2e7 sampled operands over the exponent range
    and on the 1e7 mantissas next to 1.0 and 2.0; a plain product with RN(1/d) does not."""
    lib = load()
    assert lib.orc_check_div_const(d, 20_000_000, 12345) == 0
    # the check can fail: x * RN(1/3) alone differs from x / 3 on a sizeable fraction of operands
    x = np.random.default_rng(0).uniform(1.0, 2.0, 100_000)
    assert np.count_nonzero(x * (1.0 / d) != x / d) > 1000


def test_fast_division_equals_ieee():
    """The flux kernels' division (reciprocal estimate, two Newton steps, one residual correction)
    equals IEEE division on sampled operands even from a reciprocal estimate 1e-4 off; with no
    Newton step the same estimate misses on almost every pair (the check can fail)."""
    lib = load()
    assert lib.orc_check_div_fast(1e-4, 5_000_000, 777) == 0
    b = np.random.default_rng(1).uniform(1.0, 2.0, 100_000)
    a = np.random.default_rng(2).uniform(1.0, 2.0, 100_000)
    assert np.count_nonzero(a * ((1.0 / b) * (1 + 1e-4)) != a / b) > 90_000
