// Per-cell FP64 device math for the EU3D hot path on gfx950: NASA-9 thermo, the T-from-energy
// Newton, MUSCL reconstruction, the AUSM face flux and the H2-air kinetics (rates, trapezoid,
// diagonal Jacobian). All functions are inline and operate on registers; the kernels in
// eu3d_kernels.hip compose them.
//
// Arithmetic is restructured for the GPU, not transcribed: powers of T are shared across
// species and terms, the Newton works on per-cell mixture coefficients (one 8-term polynomial
// per iteration instead of 9 species x 2 fits), and the kinetics loops are unrolled over the
// compile-time mechanism so zero stoichiometry vanishes. The results agree with the reference
// arithmetic to rounding (a few ulp), far inside the 1e-10 parity bound.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "exp_fin.hpp"
#include "mech_gen.h"

namespace eu3d {
using namespace mech;

constexpr double R_UNIV = 8.31434;   // include/Flowfield.hpp:54
constexpr double RU_CAL = 1.987;     // include/Reaction.hpp:58
constexpr double P0_PA = 101325.0;   // include/Reaction.hpp:59
constexpr int NC = NS + 4;           // conserved components

enum : uint32_t {
    Q_F1 = 1u << 0, Q_F1B = 1u << 1, Q_F2 = 1u << 2, Q_F3 = 1u << 3, Q_F6 = 1u << 4,
    Q_C1 = 1u << 5, Q_C2 = 1u << 6, Q_C3 = 1u << 7, Q_C4 = 1u << 8, Q_C6 = 1u << 9, Q_C7 = 1u << 10
};

__host__ __device__ constexpr double RI(int s) { return R_UNIV / MW[s]; }   // src/Reaction.cpp:353-354
// 1 / Mw_s, folded at compile time: divisions by the molar masses become multiplications
__host__ __device__ constexpr double RMW(int s) { return 1.0 / MW[s]; }

// a / b for the face reconstruction and flux: the reciprocal instruction refined by two Newton
// steps, then one residual correction of the quotient (Markstein). It equals IEEE division on the
// oracle's check (2e7 operand pairs over 2^-60..2^60, from reciprocal estimates 1e-4 off) in eight
// instructions, where the general sequence adds the operand scaling that only extreme exponents
// need. For |b| outside about [2^-1000, 2^1000] (a zero, denormal or infinite
// divisor: a cell already out of the physical range) the result is NaN where IEEE division gives
// an infinity or a zero.
__device__ __forceinline__ double div_fast(double a, double b) {
    double r = __builtin_amdgcn_rcp(b);
    r = __builtin_fma(__builtin_fma(-b, r, 1.0), r, r);
    r = __builtin_fma(__builtin_fma(-b, r, 1.0), r, r);
    const double q = a * r;
    return __builtin_fma(__builtin_fma(-b, q, a), r, q);
}

// ------------------------------------------------------------------ NASA-9 thermo
struct TPow {
    double T, Tm1, Tm2, T2, T3, T4, T5, lnT;
};
__device__ __forceinline__ TPow tpow(double T) {
    TPow p;
    p.T = T;
    p.Tm1 = 1.0 / T;
    p.Tm2 = p.Tm1 * p.Tm1;
    p.T2 = T * T;
    p.T3 = p.T2 * T;
    p.T4 = p.T2 * p.T2;
    p.T5 = p.T4 * T;
    p.lnT = log(T);
    return p;
}

// Powers of T from those of a nearby temperature p.T (|T / p.T - 1| < 2^-12), without a log or a
// division: ln T = ln p.T + log1p(x), x = (T - p.T) / p.T, by its series to x^4 (truncation below
// 1e-17 relative), and 1/T by two Newton steps from 1/p.T. Each agrees with tpow(T) to about an
// ulp; farther temperatures take tpow.
__device__ __forceinline__ TPow tpow_near(const TPow& p, double T) {
    const double x = (T - p.T) * p.Tm1;
    if (!(fabs(x) < 0x1p-12)) return tpow(T);
    TPow q;
    q.T = T;
    double r = p.Tm1;
    r = __builtin_fma(r, __builtin_fma(-T, r, 1.0), r);
    r = __builtin_fma(r, __builtin_fma(-T, r, 1.0), r);
    q.Tm1 = r;
    q.Tm2 = r * r;
    q.T2 = T * T;
    q.T3 = q.T2 * T;
    q.T4 = q.T2 * q.T2;
    q.T5 = q.T4 * T;
    q.lnT = p.lnT + x * (1.0 + x * (-0.5 + x * (1.0 / 3 - 0.25 * x)));
    return q;
}

// c_p/R of species s (Reaction::GetCpi, src/Reaction.cpp:893-904); range by T < 1000.
__device__ __forceinline__ double cp_R(int s, bool lo, const TPow& p) {
    const double a0 = lo ? C0[0][s] : C1[0][s], a1 = lo ? C0[1][s] : C1[1][s];
    const double a2 = lo ? C0[2][s] : C1[2][s], a3 = lo ? C0[3][s] : C1[3][s];
    const double a4 = lo ? C0[4][s] : C1[4][s], a5 = lo ? C0[5][s] : C1[5][s];
    const double a6 = lo ? C0[6][s] : C1[6][s];
    return a0 * p.Tm2 + a1 * p.Tm1 + a2 + a3 * p.T + a4 * p.T2 + a5 * p.T3 + a6 * p.T4;
}

// x / D, correctly rounded, for the constant divisors D = 3 and 5 of the h/R and s/R integrals:
// the product with RN(1/D) refined by one FMA residual step (Markstein's correction). It equals
// IEEE division on every operand of the oracle's check (3e7 per divisor: random over 2^-300..2^300
// and the mantissas next to 1 and 2) at three instructions instead of a full division.
template <int D>
__device__ __forceinline__ double div_const(double x) {
    static_assert(D == 3 || D == 5, "checked for the divisors 3 and 5 only");
    constexpr double R = 1.0 / D;
    const double q = x * R;
    const double r = __builtin_fma(-q, (double)D, x);
    return __builtin_fma(r, R, q);
}

// h/R (Reaction::GetHi, src/Reaction.cpp:916-926); quirk C4 keeps the low-range b1.
__device__ __forceinline__ double h_R(int s, bool lo, bool c4, const TPow& p) {
    const double a0 = lo ? C0[0][s] : C1[0][s], a1 = lo ? C0[1][s] : C1[1][s];
    const double a2 = lo ? C0[2][s] : C1[2][s], a3 = lo ? C0[3][s] : C1[3][s];
    const double a4 = lo ? C0[4][s] : C1[4][s], a5 = lo ? C0[5][s] : C1[5][s];
    const double a6 = lo ? C0[6][s] : C1[6][s];
    const double b1 = (lo || c4) ? C0[7][s] : C1[7][s];
    return -a0 * p.Tm1 + a1 * p.lnT + a2 * p.T + a3 * p.T2 / 2 + div_const<3>(a4 * p.T3) + a5 * p.T4 / 4 +
           div_const<5>(a6 * p.T5) + b1;
}
// s/R (Reaction::GetSi, src/Reaction.cpp:938-948); quirk C4 keeps the low-range b2.
__device__ __forceinline__ double s_R(int s, bool lo, bool c4, const TPow& p) {
    const double a0 = lo ? C0[0][s] : C1[0][s], a1 = lo ? C0[1][s] : C1[1][s];
    const double a2 = lo ? C0[2][s] : C1[2][s], a3 = lo ? C0[3][s] : C1[3][s];
    const double a4 = lo ? C0[4][s] : C1[4][s], a5 = lo ? C0[5][s] : C1[5][s];
    const double a6 = lo ? C0[6][s] : C1[6][s];
    const double b2 = (lo || c4) ? C0[8][s] : C1[8][s];
    return -a0 * p.Tm2 / 2 - a1 * p.Tm1 + a2 * p.lnT + a3 * p.T + a4 * p.T2 / 2 + div_const<3>(a5 * p.T3) +
           a6 * p.T4 / 4 + b2;
}

// G_s / (R T) of every species at p.T, and h_s / R when H. The NASA range is taken per wavefront:
// each range runs only if some lane needs it, with that range's coefficients as compile-time
// constants, so a wavefront whose lanes agree (the usual case) selects no coefficients per lane.
template <bool H>
__device__ __forceinline__ void species_thermo(const TPow& p, bool c4, double (&gx)[NS], double (&hr)[NS]) {
    const bool lo = p.T < 1000;
#pragma unroll
    for (int pass = 0; pass < 2; pass++) {
        const bool range = pass == 0;
        if (__ballot(lo == range) != 0 && lo == range) {
#pragma unroll
            for (int s = 0; s < NS; s++) {
                const double h = h_R(s, range, c4, p);
                gx[s] = h * p.Tm1 - s_R(s, range, c4, p);
                if (H) hr[s] = h;
            }
        }
    }
}

// Mixture fit: sum_s w_s * (fit of s), for both ranges, w_s = Y_s * Ri_s [kJ/kg/K].
struct MixFit {
    double lo[8], hi[8];   // a1..a7, b1
    double rmix;           // sum_s Y_s Ri_s
};
__device__ __forceinline__ MixFit mix_fit(const double (&Y)[NS], bool c4) {
    MixFit m;
#pragma unroll
    for (int k = 0; k < 8; k++) { m.lo[k] = 0.0; m.hi[k] = 0.0; }
    m.rmix = 0.0;
#pragma unroll
    for (int s = 0; s < NS; s++) {
        const double w = Y[s] * RI(s);
        m.rmix += w;
#pragma unroll
        for (int k = 0; k < 7; k++) {
            m.lo[k] += w * C0[k][s];
            m.hi[k] += w * C1[k][s];
        }
        m.lo[7] += w * C0[7][s];
        m.hi[7] += w * (c4 ? C0[7][s] : C1[7][s]);
    }
    return m;
}

// One range of the mixture fit, in mix_fit's order: c = sum_s w_s (a1..a7, b1 of s).
__device__ __forceinline__ void mix_fit_range(const double (&Y)[NS], bool c4, bool lo, double (&c)[8]) {
#pragma unroll
    for (int k = 0; k < 8; k++) c[k] = 0.0;
#pragma unroll
    for (int s = 0; s < NS; s++) {
        const double w = Y[s] * RI(s);
#pragma unroll
        for (int k = 0; k < 7; k++) c[k] += w * (lo ? C0[k][s] : C1[k][s]);
        c[7] += w * ((lo || c4) ? C0[7][s] : C1[7][s]);
    }
}

// The mixture fit formed per range on demand for the whole wavefront: a range is summed only
// when some lane of the wavefront needs it (with the usual wavefront on one side of 1000 K, half
// of mix_fit's work). The values equal mix_fit's.
struct LazyMix {
    MixFit m;
    bool have_lo = false, have_hi = false;
    __device__ __forceinline__ void need(const double (&Y)[NS], bool c4, bool lo) {
        if (!have_lo && __ballot(lo) != 0) {
            mix_fit_range(Y, c4, true, m.lo);
            have_lo = true;
        }
        if (!have_hi && __ballot(!lo) != 0) {
            mix_fit_range(Y, c4, false, m.hi);
            have_hi = true;
        }
    }
};

// T from internal energy (Flowfield::Get_temp, src/Flowfield.cpp:1542-1566): Newton on
// sum_s Y_s e_s(T) = E, at most 10 iterations, stop when |dT| < 1e-6 K. E in J/kg. p holds the
// powers of the starting temperature (callers that already have them pass them in).
// The iterates' powers of T come from the previous iterate's by tpow_near (no log or division
// while the steps are small); *p_last receives the powers of the last iterate, from which the
// caller forms those of the result the same way.
__device__ __forceinline__ double newton_T_from(const MixFit& m, TPow p, double E, int* iters = nullptr,
                                                TPow* p_last = nullptr) {
    double Tn = p.T;
    const double Ek = E * 1e-3;
    int it = 0;
    for (int count = 0; count < 10; count++) {
        if (count > 0) p = tpow_near(p, Tn);
        const double T0 = p.T;
        const bool lo = T0 < 1000;
        const double* c = lo ? m.lo : m.hi;
        const double h = -c[0] * p.Tm1 + c[1] * p.lnT + c[2] * p.T + c[3] * p.T2 / 2 + div_const<3>(c[4] * p.T3) +
                         c[5] * p.T4 / 4 + div_const<5>(c[6] * p.T5) + c[7];
        const double cp = c[0] * p.Tm2 + c[1] * p.Tm1 + c[2] + c[3] * p.T + c[4] * p.T2 + c[5] * p.T3 +
                          c[6] * p.T4;
        const double f = (h - Ek) - m.rmix * T0;
        const double fp = cp - m.rmix;
        Tn = T0 - f / fp;
        it++;
        if (fabs(Tn - T0) < 1e-6) break;
    }
    if (iters) *iters = it;
    if (p_last) *p_last = p;
    return Tn;
}
__device__ __forceinline__ double newton_T(const MixFit& m, double T, double E, int* iters = nullptr) {
    return newton_T_from(m, tpow(T), E, iters);
}
// The same Newton on a LazyMix of the mass fractions Y.
__device__ __forceinline__ double newton_T_lazy(LazyMix& L, const double (&Y)[NS], bool c4, TPow p, double E,
                                                int* iters = nullptr, TPow* p_last = nullptr) {
    double Tn = p.T;
    const double Ek = E * 1e-3;
    int it = 0;
    for (int count = 0; count < 10; count++) {
        if (count > 0) p = tpow_near(p, Tn);
        const double T0 = p.T;
        const bool lo = T0 < 1000;
        L.need(Y, c4, lo);
        const double* c = lo ? L.m.lo : L.m.hi;
        const double h = -c[0] * p.Tm1 + c[1] * p.lnT + c[2] * p.T + c[3] * p.T2 / 2 + div_const<3>(c[4] * p.T3) +
                         c[5] * p.T4 / 4 + div_const<5>(c[6] * p.T5) + c[7];
        const double cp = c[0] * p.Tm2 + c[1] * p.Tm1 + c[2] + c[3] * p.T + c[4] * p.T2 + c[5] * p.T3 +
                          c[6] * p.T4;
        const double f = (h - Ek) - L.m.rmix * T0;
        const double fp = cp - L.m.rmix;
        Tn = T0 - f / fp;
        it++;
        if (fabs(Tn - T0) < 1e-6) break;
    }
    if (iters) *iters = it;
    if (p_last) *p_last = p;
    return Tn;
}

// c_p of the mixture [kJ/kg/K] from the mixture fit, at p.T.
__device__ __forceinline__ double mix_cp_p(const MixFit& m, const TPow& p) {
    const double* c = (p.T < 1000) ? m.lo : m.hi;
    return c[0] * p.Tm2 + c[1] * p.Tm1 + c[2] + c[3] * p.T + c[4] * p.T2 + c[5] * p.T3 + c[6] * p.T4;
}
__device__ __forceinline__ double mix_cp(const MixFit& m, double T) { return mix_cp_p(m, tpow(T)); }
// c_p of the mixture of mass fractions Y at p.T, summing only the ranges the wavefront needs.
__device__ __forceinline__ double mix_cp_Y(const double (&Y)[NS], bool c4, const TPow& p) {
    LazyMix L;
    L.need(Y, c4, p.T < 1000);
    return mix_cp_p(L.m, p);
}

// Primitive state of one cell, as Flowfield::Update_after_CS leaves it (src/Flowfield.cpp:1246-1284).
struct Prim {
    double D, U, V, W, E, T, P, H, Gamma, C, Rgas, Cp, Wav;
    double Y[NS];
};

// Update_after_CS for one cell: CS[NC], previous T (as its powers p_old) and E -> primitives.
// Quirk F3 solves T from the previous E; FIXED recomputes E from CS first. LAZY forms the
// mixture fit per range on demand (LazyMix): it pays where a wavefront's cells are neighbours in
// space and mostly on one side of 1000 K (the grid step), not where they are unrelated cells of
// a work queue (there both ranges are usually needed and the branches cost more than they save).
// pre(o) runs before the temperature Newton with the fields that do not depend on T (D, U, V, W,
// E, Y) set, so that a caller can store them then instead of holding them through the Newton.
// p_final, if given, receives the powers of the solved T (by series from the last iterate's).
struct NoPre {
    __device__ void operator()(const Prim&) const {}
};

template <bool LAZY = false, class Pre = NoPre>
__device__ __forceinline__ Prim prim_from_cs_p(const double (&cs)[NC], const TPow& p_old, double E_old, uint32_t q,
                                                int* iters = nullptr, const Pre& pre = Pre(),
                                                TPow* p_final = nullptr) {
    Prim o;
    double D = 0;
#pragma unroll
    for (int s = 0; s < NS; s++) D += cs[s];
    o.D = D;
    const double rD = 1.0 / D;
#pragma unroll
    for (int s = 0; s < NS; s++) o.Y[s] = cs[s] * rD;
    o.U = cs[NS] * rD;
    o.V = cs[NS + 1] * rD;
    o.W = cs[NS + 2] * rD;
    o.E = (cs[NS + 3] - 0.5 * D * (o.U * o.U + o.V * o.V + o.W * o.W)) * rD;
    pre(o);
    double cp;
    if constexpr (LAZY) {
        LazyMix L;
        L.m.rmix = 0.0;
#pragma unroll
        for (int s = 0; s < NS; s++) L.m.rmix += o.Y[s] * RI(s);
        TPow pl;
        o.T = newton_T_lazy(L, o.Y, q & Q_C4, p_old, (q & Q_F3) ? E_old : o.E, iters, &pl);
        L.need(o.Y, q & Q_C4, o.T < 1000);
        const TPow pf = tpow_near(pl, o.T);
        cp = mix_cp_p(L.m, pf);
        if (p_final) *p_final = pf;
    } else {
        const MixFit m = mix_fit(o.Y, q & Q_C4);
        TPow pl;
        o.T = newton_T_from(m, p_old, (q & Q_F3) ? E_old : o.E, iters, &pl);
        const TPow pf = tpow_near(pl, o.T);
        cp = mix_cp_p(m, pf);
        if (p_final) *p_final = pf;
    }
    double sy = 0;
#pragma unroll
    for (int s = 0; s < NS; s++) sy += o.Y[s] * RMW(s);
    o.Wav = 1.0 / sy;
    o.Rgas = R_UNIV * 1000 / o.Wav;
    o.P = o.T * D * o.Rgas;
    o.H = o.P * rD + o.E;
    o.Cp = cp * 1000;
    o.Gamma = o.Cp / (o.Cp - R_UNIV * sy * 1000);
    o.C = sqrt(o.Gamma * o.P * rD);
    return o;
}

__device__ __forceinline__ Prim prim_from_cs(const double (&cs)[NC], double T_old, double E_old, uint32_t q,
                                              int* iters = nullptr) {
    return prim_from_cs_p(cs, tpow(T_old), E_old, q, iters);
}

// dT/drho_s (Flowfield::GetPartial_T, src/Flowfield.cpp:1578-1595); quirk C7 keeps h_s in kJ/kg.
// hr: h_s / R at o.T of every species.
__device__ __forceinline__ void partial_T_h(const Prim& o, uint32_t q, const double (&hr)[NS], double (&pt)[NS]) {
    const double CV = o.Cp - o.Rgas;
    const double ke = 0.5 * (o.U * o.U + o.V * o.V + o.W * o.W);
    const double rden = 1.0 / (CV * o.D);
    const double hs = (q & Q_C7) ? 1.0 : 1e3;
#pragma unroll
    for (int s = 0; s < NS; s++) {
        const double hi = RI(s) * hr[s] * hs;
        pt[s] = (ke - hi + RI(s) * 1e3 * o.T) * rden;
    }
}
__device__ __forceinline__ void partial_T(const Prim& o, uint32_t q, double (&pt)[NS]) {
    const TPow p = tpow(o.T);
    double gx[NS], hr[NS];
    species_thermo<true>(p, q & Q_C4, gx, hr);
    partial_T_h(o, q, hr, pt);
}

// ------------------------------------------------------------------ reconstruction
// Limiters (include/Differ.hpp:33-82) and MUSCL_2 (include/Differ.hpp:149-238) for the face
// between f0 and f1 with neighbours fm (behind f0) and f2 (beyond f1).
__device__ __forceinline__ double lim(int limit, double R) {
    switch (limit) {
    case 1: return (R + fabs(R)) / (1.0 + fabs(R));
    case 2: return (R * R + R) / (1.0 + R * R);
    case 3: {
        double r = 0.0, b = (1.0 < 2 * R) ? 1.0 : 2 * R, c = (2.0 < R) ? 2.0 : R;
        if (r < b) r = b;
        if (r < c) r = c;
        return r;
    }
    case 4: {
        if (!(R > 0)) return 0.0;
        double r = 2 * R;
        if (1.0 < r) r = 1.0;
        const double c = (1 + R) / 2;
        if (c < r) r = c;
        return r;
    }
    default: return (R > 0) ? ((1.0 < R) ? 1.0 : R) : 0.0;
    }
}

// kappa = 1/3 weights as the reference forms them: (1 - 1/3) and (1 + 1/3).
constexpr double KM = 1.0 - 1.0 / 3.0;
constexpr double KP = 1 + 1.0 / 3.0;

// minmod branch of MUSCL_2 without the reciprocal divisions: for ratio r, the bracket
// (1-k) phi(r) + (1+k) phi(1/r) r is 0 for r <= 0, (1-k) r + (1+k) r for 0 < r <= 1 and
// (1-k) + (1+k) for r > 1 (the reference's (1/r)*r is 1 to rounding).
__device__ __forceinline__ double mm_bracket(double r) {
    if (!(r > 0)) return 0.0;
    return (1.0 < r) ? (KM + KP) : (KM * r + KP * r);
}

__device__ __forceinline__ void muscl(int order, int limit, double fm, double f0, double f1, double f2,
                                      double& fl, double& fr) {
    if (order == 0) {
        fl = f0;
        fr = f1;
        return;
    }
    const double A0 = f0 - fm, A1 = f1 - f0, A2 = f2 - f1;
    const double L = div_fast(A1, A0 + 1e-12);
    const double R = div_fast(A1, A2 + 1e-12);
    double bl, br;
    if (limit == 0) {
        bl = mm_bracket(L);
        br = mm_bracket(R);
    } else {
        bl = KM * lim(limit, L) + KP * lim(limit, 1.0 / L) * L;
        br = KM * lim(limit, R) + KP * lim(limit, 1.0 / R) * R;
    }
    fl = f0 + 0.25 * bl * A0;
    fr = f1 - 0.25 * br * A2;
}

// One face state (reconstructed).
struct Face {
    double P, D, U, V, W, H, G;
    double Y[NS];
};

// AUSM-type interface flux (Flowfield::AUSM face loop, src/Flowfield.cpp:1394-1500).
// dir: 0 = x, 1 = y, 2 = z. Writes NC flux components.
__device__ __forceinline__ void ausm_flux(int dir, const Face& L, const Face& R, bool zero_y, double (&F)[NC]) {
    const double CL = sqrt(L.G * L.P / L.D), CR = sqrt(R.G * R.P / R.D);
    const double CI = 0.5 * (CL + CR);
    const double unL = dir == 0 ? L.U : (dir == 1 ? L.V : L.W);
    const double unR = dir == 0 ? R.U : (dir == 1 ? R.V : R.W);
    const double MaL = unL / CI, MaR = unR / CI;
    double B1, G1, B2, G2;
    if (fabs(MaL) < 1.0) {
        const double a = (MaL + 1.0) * (MaL + 1.0), b = MaL * MaL - 1.0;
        B1 = 0.25 * a + 0.125 * (b * b);
        G1 = 0.25 * a * (2.0 - MaL) + 3.0 / 16.0 * MaL * (b * b);
    } else {
        B1 = 0.5 * (MaL + fabs(MaL));
        G1 = (MaL > 0) ? 1.0 : -0.0;    // B1 / MaL
    }
    if (fabs(MaR) < 1.0) {
        const double a = (MaR - 1.0) * (MaR - 1.0), b = MaR * MaR - 1.0;
        B2 = -0.25 * a - 0.125 * (b * b);
        G2 = 0.25 * a * (2.0 + MaR) - 3.0 / 16.0 * MaR * (b * b);
    } else {
        B2 = 0.5 * (MaR - fabs(MaR));
        G2 = (MaR > 0) ? 0.0 : 1.0;     // B2 / MaR
    }
    const double VI = (B1 + B2) * CI;
    const bool up = VI >= 0.0;
    const double Dx = up ? L.D : R.D, Ux = up ? L.U : R.U, Vx = up ? L.V : R.V, Wx = up ? L.W : R.W;
    const double Hx = up ? L.H : R.H;
    const double m = VI * Dx;
#pragma unroll
    for (int s = 0; s < NS; s++) F[s] = zero_y ? m * 0.0 : m * (up ? L.Y[s] : R.Y[s]);
    const double pterm = G1 * L.P;
    F[NS] = m * Ux;
    F[NS + 1] = m * Vx;
    F[NS + 2] = m * Wx;
    F[NS + dir] = F[NS + dir] + pterm + G2 * R.P;
    F[NS + 3] = VI * (Dx * Hx + 0.5 * Dx * (Ux * Ux + Vx * Vx + Wx * Wx));
}

// AUSM face flux from a 4-cell stencil with upwind-only species reconstruction. ld(v, m)
// returns primitive v (0 P, 1 rho, 2 u, 3 v, 4 w, 5 H, 6 gamma, 7.. Y_s) at stencil position
// m = 0..3 (cells c-1, c, c+1, c+2 of the face between c and c+1). The seven flow variables are
// reconstructed on both sides; a species' face state only enters the flux on the side the
// interface velocity selects (src/Flowfield.cpp:1440-1457), so each species is reconstructed
// once, on that side, by selecting the MUSCL inputs: the value equals the reference's.
__device__ __forceinline__ double muscl_side(int limit, bool left, double fm, double f0, double f1, double f2) {
    const double A1 = f1 - f0;
    const double base = left ? f0 : f1;
    const double d = left ? (f0 - fm) : (f2 - f1);
    const double r = div_fast(A1, d + 1e-12);
    const double b = (limit == 0) ? mm_bracket(r) : (KM * lim(limit, r) + KP * lim(limit, 1.0 / r) * r);
    const double t = 0.25 * b * d;
    return left ? base + t : base - t;
}

template <int dir, class Ld>
__device__ __forceinline__ void ausm_stencil(const Ld& ld, int order, int limit, bool zero_y, double (&F)[NC]) {
    double fl[7], fr[7];
#pragma unroll
    for (int v = 0; v < 7; v++) muscl(order, limit, ld(v, 0), ld(v, 1), ld(v, 2), ld(v, 3), fl[v], fr[v]);
    const double PL = fl[0], DL = fl[1], UL = fl[2], VL = fl[3], WL = fl[4], HL = fl[5], GL = fl[6];
    const double PR = fr[0], DR = fr[1], UR = fr[2], VR = fr[3], WR = fr[4], HR = fr[5], GR = fr[6];
    const double CL = sqrt(div_fast(GL * PL, DL)), CR = sqrt(div_fast(GR * PR, DR));
    const double CI = 0.5 * (CL + CR);
    const double unL = dir == 0 ? UL : (dir == 1 ? VL : WL);
    const double unR = dir == 0 ? UR : (dir == 1 ? VR : WR);
    const double MaL = div_fast(unL, CI), MaR = div_fast(unR, CI);
    double B1, G1, B2, G2;
    if (fabs(MaL) < 1.0) {
        const double a = (MaL + 1.0) * (MaL + 1.0), b = MaL * MaL - 1.0;
        B1 = 0.25 * a + 0.125 * (b * b);
        G1 = 0.25 * a * (2.0 - MaL) + 3.0 / 16.0 * MaL * (b * b);
    } else {
        B1 = 0.5 * (MaL + fabs(MaL));
        G1 = (MaL > 0) ? 1.0 : -0.0;
    }
    if (fabs(MaR) < 1.0) {
        const double a = (MaR - 1.0) * (MaR - 1.0), b = MaR * MaR - 1.0;
        B2 = -0.25 * a - 0.125 * (b * b);
        G2 = 0.25 * a * (2.0 + MaR) - 3.0 / 16.0 * MaR * (b * b);
    } else {
        B2 = 0.5 * (MaR - fabs(MaR));
        G2 = (MaR > 0) ? 0.0 : 1.0;
    }
    const double VI = (B1 + B2) * CI;
    const bool up = VI >= 0.0;
    const double Dx = up ? DL : DR, Ux = up ? UL : UR, Vx = up ? VL : VR, Wx = up ? WL : WR;
    const double Hx = up ? HL : HR;
    const double m = VI * Dx;
#pragma unroll
    for (int s = 0; s < NS; s++) {
        double y;
        if (order == 0) y = up ? ld(7 + s, 1) : ld(7 + s, 2);
        else y = muscl_side(limit, up, ld(7 + s, 0), ld(7 + s, 1), ld(7 + s, 2), ld(7 + s, 3));
        F[s] = zero_y ? m * 0.0 : m * y;
    }
    F[NS] = m * Ux;
    F[NS + 1] = m * Vx;
    F[NS + 2] = m * Wx;
    F[NS + dir] = F[NS + dir] + G1 * PL + G2 * PR;
    F[NS + 3] = VI * (Dx * Hx + 0.5 * Dx * (Ux * Ux + Vx * Vx + Wx * Wx));
}

// ------------------------------------------------------------------ kinetics
// One pass over the reactions of one cell (the common part of Reaction::Trapezoid,
// TrapezoidPrediction and Diagonalized; src/Reaction.cpp:420-482,499-588,744-809), accumulating
// per species as each reaction is evaluated, in reaction order as the reference sums:
//   w  = Mw sum_r (nu''-nu') R_TB RR                     [g/cm^3/s]   (all modes)
//   Q, P: production and destruction of the point-implicit trapezoid          (TRAP)
//   md: diagonal of the chemical Jacobian                                      (DIAG)
// Inputs in reaction units: Mc [mol/cm^3], Dr = Di [g/cm^3]. rtb: 0 keeps the summed
// third-body factor (quirk C1), 1 forces it to 0 (quirk C2), 2 uses 1 where the sum is 0.
// k_b is formed without the reference's two-step K_p, K_c division:
//   k_b = k_f / K_c = k_f exp(sum_s nu_s G_s / RT) (RT / (P0 1e-6))^dnu,
// with G_s / RT = h_s/(RT) - s_s/R from the NASA fits; the exponents of k_f (b ln T - E/RuT)
// are folded into one exp (exp_fin). Zero stoichiometry is skipped at compile time.
struct KinOut {
    double w[NS], P[NS], Q[NS], md[NS];
};

// gx: G_s / (R T) at p.T (species_thermo).
template <bool TRAP, bool DIAG, bool C3 = false>
__device__ __forceinline__ void kinetics_g(const TPow& p, const double (&gx)[NS], const double (&Mc)[NS],
                                           const double (&Dr)[NS], const double (&pts)[NS], int rtb, KinOut& o) {
    const double T = p.T;
#pragma unroll
    for (int s = 0; s < NS; s++) {
        o.w[s] = 0.0;
        if (TRAP) { o.P[s] = 0.0; o.Q[s] = 0.0; }
        if (DIAG) o.md[s] = 0.0;
    }
    // x / Dr as (x * rDr) * sDr: rDr = 1 / (Dr sDr), or 0 where Dr = 0 (those species contribute
    // no Jacobian term). sDr is 1, or 2^64 for |Dr| below 2^-1000, whose own reciprocal would
    // overflow (a trace species at a denormal density, where the reference's quotient is finite).
    // Scaling by powers of two is exact, so for every other Dr the value is x * (1 / Dr).
    double rDr[NS], sDr[NS];
    if (DIAG) {
#pragma unroll
        for (int s = 0; s < NS; s++) {
            sDr[s] = fabs(Dr[s]) < 0x1p-1000 ? 0x1p64 : 1.0;
            rDr[s] = (Dr[s] != 0) ? 1.0 / (Dr[s] * sDr[s]) : 0.0;
        }
    }
    const double RT = R_UNIV * T;
    constexpr double P0R = P0_PA * 1e-6;
    const double rt_p0 = RT / P0R, p0_rt = P0R / RT;
    const double rRuT = 1.0 / (RU_CAL * T);
    const double rRuT2 = rRuT * p.Tm1;
#pragma unroll
    for (int r = 0; r < NR; r++) {
        // k_f = A T^b exp(-E/(Ru T))
        double kf = AF[r];
        double ex = 0.0;
        bool use_exp = false;
        if (BF[r] == 1.0) kf *= T;
        else if (BF[r] == -1.0) kf *= p.Tm1;
        else if (BF[r] == -2.0) kf *= p.Tm2;
        else if (BF[r] != 0.0) { ex = BF[r] * p.lnT; use_exp = true; }
        if (EAF[r] != 0.0) { ex -= EAF[r] * rRuT; use_exp = true; }
        if (use_exp) kf *= exp_fin(ex);
        double xk = 0.0, rf = 1.0, rb = 1.0, tb = 0.0;
        int xr = 0;
#pragma unroll
        for (int s = 0; s < NS; s++) {
            const int nf = SF[s][r], nb = SB[s][r];
            if (nb - nf != 0) xk += (double)(nb - nf) * gx[s];
            xr += nb - nf;
            if (nf == 1) rf *= Mc[s];
            else if (nf == 2) rf *= Mc[s] * Mc[s];
            if (nb == 1) rb *= Mc[s];
            else if (nb == 2) rb *= Mc[s] * Mc[s];
            if (TB[s][r] != 0.0) tb += Mc[s] * TB[s][r];
        }
        double kb = kf * exp_fin(xk);
        if (xr == 1) kb *= rt_p0;
        else if (xr == -1) kb *= p0_rt;
        if (rtb == 1) tb = 0.0;
        else if (rtb == 2 && tb == 0.0) tb = 1.0;
        const double rr = kf * rf - kb * rb;
        const double dRdT = DIAG ? rr * (BF[r] * p.Tm1 + EAF[r] * rRuT2) : 0.0;
#pragma unroll
        for (int s = 0; s < NS; s++) {
            const int nf = SF[s][r], nb = SB[s][r], dn = nb - nf;
            if (dn != 0) o.w[s] += (double)dn * tb * rr;
            if (TRAP && (nf != 0 || nb != 0)) {
                const double qf = nb ? (double)nb * kf * rf : 0.0;
                const double qb = nf ? (double)nf * kb * rb : 0.0;
                const double pb = nb ? (double)nb * kb * rb : 0.0;
                const double pf = nf ? (double)nf * kf * rf : 0.0;
                if (C3) {
                    o.Q[s] += qf + qb * tb;
                    o.P[s] += pb + pf * tb;
                } else {
                    o.Q[s] += (qf + qb) * tb;
                    o.P[s] += (pb + pf) * tb;
                }
            }
            if (DIAG && dn != 0) {
                const double a = nf ? ((kf * (double)nf * rf) * rDr[s]) * sDr[s] : 0.0;
                const double b = nb ? ((kb * (double)nb * rb) * rDr[s]) * sDr[s] : 0.0;
                const double w1 = a - b;
                double term = (w1 + dRdT * pts[s]) * tb;
                if (TB[s][r] != 0.0) term += rr * (TB[s][r] / MW[s]);
                o.md[s] += (double)dn * term;
            }
        }
        // keep each reaction's temporaries local: without this the scheduler hoists every
        // reaction's exp() ahead and the accumulators spill
        __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int s = 0; s < NS; s++) {
        o.w[s] *= MW[s];
        if (TRAP) {
            o.Q[s] *= MW[s];
            o.P[s] = (Mc[s] == 0) ? 0.0 : o.P[s] / Mc[s];
        }
        if (DIAG) o.md[s] *= MW[s];
    }
}

template <bool TRAP, bool DIAG, bool C3 = false>
__device__ __forceinline__ void kinetics_p(const TPow& p, const double (&Mc)[NS], const double (&Dr)[NS],
                                           const double (&pts)[NS], int rtb, bool c4, KinOut& o) {
    double gx[NS], hr[NS];
    species_thermo<false>(p, c4, gx, hr);
    kinetics_g<TRAP, DIAG, C3>(p, gx, Mc, Dr, pts, rtb, o);
}
template <bool TRAP, bool DIAG, bool C3 = false>
__device__ __forceinline__ void kinetics(double T, const double (&Mc)[NS], const double (&Dr)[NS],
                                         const double (&pts)[NS], int rtb, bool c4, KinOut& o) {
    kinetics_p<TRAP, DIAG, C3>(tpow(T), Mc, Dr, pts, rtb, c4, o);
}

// Point-implicit trapezoid update (src/Reaction.cpp:481,587): Di' [kg/m^3] from Di [g/cm^3].
__device__ __forceinline__ void trap_update(const KinOut& k, const double (&Dr)[NS], double dt, double (&Dout)[NS]) {
#pragma unroll
    for (int s = 0; s < NS; s++)
        Dout[s] = ((1.0 - dt / 2.0 * k.P[s]) * Dr[s] + dt * k.Q[s]) / (1.0 + dt / 2.0 * k.P[s]) * 1e3;
}

}  // namespace eu3d
