/* CPU oracle: FP64 restatement of the EU3D reference hot path (TEST INFRASTRUCTURE ONLY).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this file's
 * library, and only to check the HIP path: the product path never calls it.
 *
 * Each function restates the reference line by line, keeping the reference's loop ranges,
 * evaluation order (left-to-right sums, pow() calls) and scratch-array semantics, because
 * parity is defined against that arithmetic. Built with -O2 -fno-fast-math
 * -ffp-contract=off so no FMA contraction or reassociation changes the rounding.
 * Never-written reference scratch (quirks F1b, I2) is taken as zero.
 *
 * Parity unpinned: what the reference leaves to uninitialised memory has no check to pin it:
 * MD(c,c) for c >= NS under quirk I2 (taken as 0), the F1b face states for i in [ni, ni+bc)
 * (taken as zero-filled) and F5's shallow copies (taken as deep copies). Every function has its
 * own pins in tests/test_oracle_golden.py.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define R_UNIV 8.31434  /* include/Flowfield.hpp:54, include/Reaction.hpp:57 */
#define RU_CAL 1.987    /* include/Reaction.hpp:58 */
#define P0_PA 101325.0  /* include/Reaction.hpp:59 */
#define NSMAX 16
#define NRMAX 24

#define NXg(f) ((f)->ni + 2 * (f)->bc)
#define NYg(f) ((f)->nj + 2 * (f)->bc)
#define NZg(f) ((f)->nk + 2 * (f)->bc)
#define I3(f, i, j, k) ((((size_t)(i)) * NYg(f) + (size_t)(j)) * NZg(f) + (size_t)(k))
#define I4(f, i, j, k, c, nc) (I3(f, i, j, k) * (size_t)(nc) + (size_t)(c))

static inline double dmin(double a, double b) { return (b < a) ? b : a; } /* std::min */
static inline double dmax(double a, double b) { return (a < b) ? b : a; } /* std::max */

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_num_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ---------------------------------------------------------------- NASA-9 thermo */
/* Reaction::GetCpi, src/Reaction.cpp:893-904 */
double orc_getcpi(double T, double R, int sp, const orc_mech *m, uint32_t q) {
    (void)q;
    const int ns = m->ns;
    const double *c = (T < 1000) ? m->coeff0 : m->coeff1;
#define CF(k) c[(k) * ns + sp]
    return R * (CF(0) * pow(T, -2) + CF(1) * pow(T, -1) + CF(2) + CF(3) * T + CF(4) * pow(T, 2) +
                CF(5) * pow(T, 3) + CF(6) * pow(T, 4));
#undef CF
}

/* Reaction::GetHi, src/Reaction.cpp:916-926 (quirk C4: high branch takes Coeff0(7)) */
double orc_gethi(double T, double R, int sp, const orc_mech *m, uint32_t q) {
    const int ns = m->ns;
    const double *c = (T < 1000) ? m->coeff0 : m->coeff1;
    const double b1 = (T < 1000 || (q & ORC_Q_C4)) ? m->coeff0[7 * ns + sp] : m->coeff1[7 * ns + sp];
#define CF(k) c[(k) * ns + sp]
    return R * (-CF(0) * pow(T, -1) + CF(1) * log(T) + CF(2) * T + CF(3) * pow(T, 2) / 2 +
                CF(4) * pow(T, 3) / 3 + CF(5) * pow(T, 4) / 4 + CF(6) * pow(T, 5) / 5 + b1);
#undef CF
}

/* Reaction::GetSi, src/Reaction.cpp:938-948 (quirk C4: high branch takes Coeff0(8)) */
double orc_getsi(double T, double R, int sp, const orc_mech *m, uint32_t q) {
    const int ns = m->ns;
    const double *c = (T < 1000) ? m->coeff0 : m->coeff1;
    const double b2 = (T < 1000 || (q & ORC_Q_C4)) ? m->coeff0[8 * ns + sp] : m->coeff1[8 * ns + sp];
#define CF(k) c[(k) * ns + sp]
    return R * (-CF(0) * pow(T, -2) / 2 - CF(1) * pow(T, -1) + CF(2) * log(T) + CF(3) * T +
                CF(4) * pow(T, 2) / 2 + CF(5) * pow(T, 3) / 3 + CF(6) * pow(T, 4) / 4 + b2);
#undef CF
}

/* Function::sum, src/Function.cpp:17-42 (left-to-right accumulation from 0) */
double orc_sum(int flag, const double *a, const double *b, int n) {
    double s = 0;
    switch (flag) {
    case 0: for (int i = 0; i < n; i++) s += (a[i] + b[i]); break;
    case 1: for (int i = 0; i < n; i++) s += (a[i] - b[i]); break;
    case 2: for (int i = 0; i < n; i++) s += (a[i] * b[i]); break;
    case 3: for (int i = 0; i < n; i++) s += (a[i] / b[i]); break;
    }
    return s;
}

/* Flowfield::Get_temp (the overload in use), src/Flowfield.cpp:1542-1566.
 * Newton on e(T) = E: at most 10 iterations, stop when |T_{n+1} - T_n| < 1e-6. */
double orc_get_temp(const orc_mech *m, uint32_t q, double T, double E, const double *Y, int *iters) {
    const int ns = m->ns;
    double T0 = T, T_temp = T, temp1, temp2;
    double Cpi0[NSMAX], Hi0[NSMAX];
    int count = 0, it = 0;
    while (count < 10) {
        T0 = T_temp;
        for (int s = 0; s < ns; s++) {
            Cpi0[s] = orc_getcpi(T0, m->ri[s], s, m, q);
            Hi0[s] = orc_gethi(T0, m->ri[s], s, m, q);
        }
        temp1 = (orc_sum(2, Y, Hi0, ns) - E * 1e-3) - orc_sum(2, Y, m->ri, ns) * T0;
        temp2 = orc_sum(2, Y, Cpi0, ns) - orc_sum(2, Y, m->ri, ns);
        T_temp = T0 - temp1 / temp2;
        it++;
        if (fabs(T_temp - T0) < 1e-6)
            break;
        else
            count = count + 1;
    }
    if (iters) *iters = it;
    return T_temp;
}

/* ------------------------------------------------------- primitive recovery */
/* Flowfield::Update_after_CS (per cell), src/Flowfield.cpp:1287-1319 (same arithmetic as the
 * full-field loop at :1246-1284). Quirk F3: T is solved from the E stored by the previous call;
 * FIXED recomputes E from the conserved energy first. */
void orc_update_after_cs_cell(orc_flow *f, const orc_mech *m, int i, int j, int k) {
    const int ns = m->ns, nc = ns + 4;
    const uint32_t q = f->quirks;
    const size_t c3 = I3(f, i, j, k);
    double Y[NSMAX], Cpi1[NSMAX];
    const double *cs = &f->CS[c3 * nc];
    double D = 0;
    for (int s = 0; s < ns; s++) {
        D += cs[s];
        f->Di[c3 * ns + s] = cs[s];
        f->Mc[c3 * ns + s] = f->Di[c3 * ns + s] / m->mw[s] * 1000;
    }
    f->D[c3] = D;
    for (int s = 0; s < ns; s++) {
        f->Yi[c3 * ns + s] = cs[s] / D;
        Y[s] = f->Yi[c3 * ns + s];
    }
    double U = cs[ns] / D, V = cs[ns + 1] / D, W = cs[ns + 2] / D;
    double Enew = (cs[ns + 3] - 0.5 * D * (pow(U, 2) + pow(V, 2) + pow(W, 2))) / D;
    double Eused = (q & ORC_Q_F3) ? f->E[c3] : Enew;
    double T = orc_get_temp(m, q, f->T[c3], Eused, Y, NULL);
    f->T[c3] = T;
    f->U[c3] = U;
    f->V[c3] = V;
    f->W[c3] = W;
    f->E[c3] = Enew;
    f->Wav[c3] = 1.0 / orc_sum(3, Y, m->mw, ns);
    f->Rgas[c3] = R_UNIV * 1000 / f->Wav[c3];
    f->P[c3] = T * D * f->Rgas[c3];
    f->H[c3] = f->P[c3] / D + Enew;
    for (int s = 0; s < ns; s++) Cpi1[s] = orc_getcpi(T, m->ri[s], s, m, q);
    f->Cp[c3] = orc_sum(2, Cpi1, Y, ns) * 1000;
    f->Gamma[c3] = f->Cp[c3] / (f->Cp[c3] - R_UNIV * orc_sum(3, Y, m->mw, ns) * 1000);
    f->C[c3] = sqrt(f->Gamma[c3] * f->P[c3] / D);
    f->Ma[c3] = sqrt(pow(U, 2) + pow(V, 2) + pow(W, 2)) / f->C[c3];
}

void orc_update_after_cs(orc_flow *f, const orc_mech *m) {
    const int bc = f->bc;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = bc; i < f->ni + bc; i++)
        for (int j = bc; j < f->nj + bc; j++)
            for (int k = bc; k < f->nk + bc; k++) orc_update_after_cs_cell(f, m, i, j, k);
}

/* Flowfield::GetPartial_T, src/Flowfield.cpp:1578-1595. Quirk C7: h_s in kJ/kg against J/kg
 * terms; FIXED scales h_s by 1e3. */
void orc_get_partial_t(orc_flow *f, const orc_mech *m) {
    const int ns = m->ns, nc = ns + 4, bc = f->bc;
    const uint32_t q = f->quirks;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = bc; i < f->ni + bc; i++)
        for (int j = bc; j < f->nj + bc; j++)
            for (int k = bc; k < f->nk + bc; k++) {
                const size_t c3 = I3(f, i, j, k);
                double CV = f->Cp[c3] - f->Rgas[c3];
                for (int s = 0; s < ns; s++) {
                    double Hi = orc_gethi(f->T[c3], m->ri[s], s, m, q);
                    if (!(q & ORC_Q_C7)) Hi = Hi * 1e3;
                    f->PT[c3 * nc + s] = (0.5 * (pow(f->U[c3], 2) + pow(f->V[c3], 2) + pow(f->W[c3], 2)) - Hi +
                                          m->ri[s] * 1e3 * f->T[c3]) /
                                         (CV * f->D[c3]);
                }
                f->PT[c3 * nc + ns + 1] = -f->U[c3] / (CV * f->D[c3]);
                f->PT[c3 * nc + ns + 2] = 1.0 / (CV * f->D[c3]);
            }
}

/* ------------------------------------------------------------ reconstruction */
/* Limiters, include/Differ.hpp:33-82 */
static inline double lim_minmod(double R) { return (R > 0) ? dmin(R, 1.0) : 0.0; }
static inline double lim_vanleer(double R) { return (R + fabs(R)) / (1.0 + fabs(R)); }
static inline double lim_superbee(double R) {
    double a = 0.0, b = dmin(2 * R, 1.0), c = dmin(R, 2.0);
    double r = a;              /* std::max({..}) keeps the first largest */
    if (r < b) r = b;
    if (r < c) r = c;
    return r;
}
static inline double lim_vanalbada(double R) { return (pow(R, 2) + R) / (1.0 + pow(R, 2)); }
static inline double lim_dminmod(double R) {
    if (R > 0) {
        double r = 2 * R;      /* std::min({..}) keeps the first smallest */
        if (1.0 < r) r = 1.0;
        double c = (1 + R) / 2;
        if (c < r) r = c;
        return r;
    }
    return 0.0;
}
static inline double limiter(int limit, double R) {
    switch (limit) {
    case 0: return lim_minmod(R);
    case 1: return lim_vanleer(R);
    case 2: return lim_vanalbada(R);
    case 3: return lim_superbee(R);
    case 4: return lim_dminmod(R);
    }
    return 0.0;
}

/* MUSCL_1 (include/Differ.hpp:110-136) and MUSCL_2 (:149-238); FLR(i,j,k,0/1) for each index in
 * [bc-1, n+bc) in all three dimensions. */
void orc_muscl(int order, int dir, const orc_flow *f, double *FLR, const double *fi, int limit) {
    const int bc = f->bc, ni = f->ni, nj = f->nj, nk = f->nk;
    const size_t si = (size_t)NYg(f) * NZg(f), sj = NZg(f), sk = 1;
    const size_t st = (dir == 1) ? si : (dir == 2) ? sj : sk;
    const double kk = 1.0 / 3.0;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = bc - 1; i < ni + bc; i++)
        for (int j = bc - 1; j < nj + bc; j++)
            for (int k = bc - 1; k < nk + bc; k++) {
                const size_t c = I3(f, i, j, k);
                if (order == 1) {
                    FLR[c * 2 + 0] = fi[c];
                    FLR[c * 2 + 1] = fi[c + st];
                    continue;
                }
                double A0 = fi[c] - fi[c - st];
                double A1 = fi[c + st] - fi[c];
                double A2 = fi[c + 2 * st] - fi[c + st];
                double L = A1 / (A0 + 1e-12);
                double R = A1 / (A2 + 1e-12);
                double L1 = limiter(limit, L), L2 = limiter(limit, 1.0 / L);
                double R1 = limiter(limit, R), R2 = limiter(limit, 1.0 / R);
                FLR[c * 2 + 0] = fi[c] + 0.25 * ((1.0 - kk) * L1 + (1 + kk) * L2 * L) * A0;
                FLR[c * 2 + 1] = fi[c + st] - 0.25 * ((1.0 - kk) * R1 + (1 + kk) * R2 * R) * A2;
            }
}

/* Flowfield::AUSM, src/Flowfield.cpp:1330-1502. Face (i,j,k) lies between cell (i,j,k) and its
 * +dir neighbour; faces are written for each index in [1, n+bc). Quirk F1b: the species face
 * states are copied only for i < ni, so faces with i >= ni read never-written (zero) Y. */
void orc_ausm(orc_flow *f, int dir, double *Fi, int order, int limit) {
    const int bc = f->bc, ni = f->ni, nj = f->nj, nk = f->nk;
    const size_t n3 = (size_t)NXg(f) * NYg(f) * NZg(f);
    const int S = f->ns;
    double *PLR = calloc(n3 * 2, sizeof(double)), *DLR = calloc(n3 * 2, sizeof(double));
    double *ULR = calloc(n3 * 2, sizeof(double)), *VLR = calloc(n3 * 2, sizeof(double));
    double *WLR = calloc(n3 * 2, sizeof(double)), *HLR = calloc(n3 * 2, sizeof(double));
    double *GLR = calloc(n3 * 2, sizeof(double)), *YLR = calloc(n3 * 2, sizeof(double));
    double *Yt = calloc(n3, sizeof(double));
    double *YLt = calloc(n3 * S, sizeof(double)), *YRt = calloc(n3 * S, sizeof(double));
    orc_muscl(order, dir, f, PLR, f->P, limit);
    orc_muscl(order, dir, f, DLR, f->D, limit);
    orc_muscl(order, dir, f, ULR, f->U, limit);
    orc_muscl(order, dir, f, VLR, f->V, limit);
    orc_muscl(order, dir, f, WLR, f->W, limit);
    orc_muscl(order, dir, f, HLR, f->H, limit);
    orc_muscl(order, dir, f, GLR, f->Gamma, limit);
    const int iy_end = (f->quirks & ORC_Q_F1B) ? ni : ni + bc;
    for (int s = 0; s < S; s++) {
        for (size_t c = 0; c < n3; c++) Yt[c] = f->Yi[c * S + s];
        orc_muscl(order, dir, f, YLR, Yt, limit);
        for (int i = 1; i < iy_end; i++)
            for (int j = 1; j < nj + bc; j++)
                for (int k = 1; k < nk + bc; k++) {
                    const size_t c = I3(f, i, j, k);
                    YLt[c * S + s] = YLR[c * 2 + 0];
                    YRt[c * S + s] = YLR[c * 2 + 1];
                }
    }
    const int nc = S + 4;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 1; i < ni + bc; i++)
        for (int j = 1; j < nj + bc; j++)
            for (int k = 1; k < nk + bc; k++) {
                const size_t c = I3(f, i, j, k);
                double PL = PLR[2 * c], PR = PLR[2 * c + 1], DL = DLR[2 * c], DR = DLR[2 * c + 1];
                double UL = ULR[2 * c], UR = ULR[2 * c + 1], VL = VLR[2 * c], VR = VLR[2 * c + 1];
                double WL = WLR[2 * c], WR = WLR[2 * c + 1], HL = HLR[2 * c], HR = HLR[2 * c + 1];
                double GL = GLR[2 * c], GR = GLR[2 * c + 1];
                const double *YL = &YLt[c * S], *YR = &YRt[c * S];
                double CL = sqrt(GL * PL / DL), CR = sqrt(GR * PR / DR);
                double CI = 0.5 * (CL + CR);
                double MaL, MaR, B1, B2, G1, G2;
                if (dir == 1) { MaL = UL / CI; MaR = UR / CI; }
                else if (dir == 2) { MaL = VL / CI; MaR = VR / CI; }
                else { MaL = WL / CI; MaR = WR / CI; }
                if (fabs(MaL) < 1.0) {
                    B1 = 0.25 * pow((MaL + 1.0), 2) + 0.125 * pow((pow(MaL, 2) - 1.0), 2);
                    G1 = 0.25 * pow((MaL + 1.0), 2) * (2.0 - MaL) + 3.0 / 16.0 * MaL * pow((pow(MaL, 2) - 1.0), 2);
                } else {
                    B1 = 0.5 * (MaL + fabs(MaL));
                    G1 = B1 / MaL;
                }
                if (fabs(MaR) < 1.0) {
                    B2 = -0.25 * pow((MaR - 1.0), 2) - 0.125 * pow((pow(MaR, 2) - 1.0), 2);
                    G2 = 0.25 * pow((MaR - 1.0), 2) * (2.0 + MaR) - 3.0 / 16.0 * MaR * pow((pow(MaR, 2) - 1.0), 2);
                } else {
                    B2 = 0.5 * (MaR - fabs(MaR));
                    G2 = B2 / MaR;
                }
                double MaI = B1 + B2;
                double VI = MaI * CI;
                double *Fo = &Fi[c * nc];
                const int up = (VI >= 0.0);
                double Dx = up ? DL : DR, Ux = up ? UL : UR, Vx = up ? VL : VR, Wx = up ? WL : WR;
                double Hx = up ? HL : HR;
                const double *Yx = up ? YL : YR;
                for (int s = 0; s < S; s++) Fo[s] = VI * Dx * Yx[s];
                Fo[S] = VI * Dx * Ux;
                Fo[S + 1] = VI * Dx * Vx;
                Fo[S + 2] = VI * Dx * Wx;
                double pterm = G1 * PL + G2 * PR;
                if (dir == 1) Fo[S] = VI * Dx * Ux + pterm;
                else if (dir == 2) Fo[S + 1] = VI * Dx * Vx + pterm;
                else Fo[S + 2] = VI * Dx * Wx + pterm;
                Fo[S + 3] = VI * (Dx * Hx + 0.5 * Dx * (pow(Ux, 2) + pow(Vx, 2) + pow(Wx, 2)));
            }
    free(PLR); free(DLR); free(ULR); free(VLR); free(WLR); free(HLR); free(GLR); free(YLR);
    free(Yt); free(YLt); free(YRt);
}

/* TimeAdv::EE, src/TimeAdv.cpp:29-43 */
void orc_ee(orc_flow *f) {
    const int S = f->ns, nc = S + 4, bc = f->bc;
    const size_t si = (size_t)NYg(f) * NZg(f), sj = NZg(f);
    const double dt = f->dt;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = bc; i < f->ni + bc; i++)
        for (int j = bc; j < f->nj + bc; j++)
            for (int k = bc; k < f->nk + bc; k++) {
                const size_t c = I3(f, i, j, k);
                for (int s = 0; s < nc; s++)
                    f->RHS[c * nc + s] =
                        -dt / (f->xnode[i] - f->xnode[i - 1]) * (f->F[c * nc + s] - f->F[(c - si) * nc + s]) -
                        dt / (f->ynode[j] - f->ynode[j - 1]) * (f->G[c * nc + s] - f->G[(c - sj) * nc + s]) -
                        dt / (f->znode[k] - f->znode[k - 1]) * (f->Q[c * nc + s] - f->Q[(c - 1) * nc + s]);
            }
}

/* Flowfield::Advection, src/Flowfield.cpp:1096-1128 (EE time advance) */
void orc_advection(orc_flow *f, int diff_order, int limit) {
    const size_t n = (size_t)NXg(f) * NYg(f) * NZg(f) * (f->ns + 4);
    memset(f->F, 0, n * sizeof(double));
    memset(f->G, 0, n * sizeof(double));
    memset(f->Q, 0, n * sizeof(double));
    orc_ausm(f, 1, f->F, diff_order, limit);
    orc_ausm(f, 2, f->G, diff_order, limit);
    orc_ausm(f, 3, f->Q, diff_order, limit);
    orc_ee(f);
}

/* Flowfield::CFLcondition, src/Flowfield.cpp:242-272 (single rank: the MPI min is the identity).
 * Quirk F6 (REF): dt = cfl min(dx/(|u|+c), dy/(|v|+c)), z ignored. FIXED: the bound of the unsplit
 * update in all three directions, dt = cfl / max sum_d (|u_d|+c)/dx_d, over the directions with
 * more than one cell (a one-cell direction has no flux difference). The per-direction minimum lets
 * the three-direction update run at an effective Courant number of up to 3 cfl, and the 3-D
 * detonation then drives a cell's internal energy below the mixture's lowest tabulated energy. */
void orc_cfl(orc_flow *f, double cfl, double final_time) {
    const int bc = f->bc;
    double dx = f->xnode[1] - f->xnode[0], dy = f->ynode[1] - f->ynode[0], dz = f->znode[1] - f->znode[0];
    const int fixed = !(f->quirks & ORC_Q_F6);
    if (fixed) {   /* a one-cell direction contributes (|u|+c)/inf = 0 */
        if (f->ni == 1) dx = INFINITY;
        if (f->nj == 1) dy = INFINITY;
        if (f->nk == 1) dz = INFINITY;
    }
    double dt_conv = 1e10;
    for (int i = bc; i < f->ni + bc; i++)
        for (int j = bc; j < f->nj + bc; j++)
            for (int k = bc; k < f->nk + bc; k++) {
                const size_t c = I3(f, i, j, k);
                double temp;
                if (fixed) {
                    const double rate = (fabs(f->U[c]) + f->C[c]) / dx + (fabs(f->V[c]) + f->C[c]) / dy +
                                        (fabs(f->W[c]) + f->C[c]) / dz;
                    temp = 1.0 / rate;
                    if (!(temp < 1e10)) temp = 1e10;
                } else {
                    temp = dmin(dx / (fabs(f->U[c]) + f->C[c]), dy / (fabs(f->V[c]) + f->C[c]));
                }
                if (temp < dt_conv) dt_conv = temp;
            }
    f->dt = cfl * dt_conv;
    if (f->time + f->dt > final_time) f->dt = final_time - f->time;
    f->time += f->dt;
}

/* ----------------------------------------------------------------- ghost fill */
#define NPRIM_COPY 10
static void copy_prims(orc_flow *f, size_t dst, size_t src, int S, int flip_comp) {
    f->U[dst] = f->U[src]; f->V[dst] = f->V[src]; f->W[dst] = f->W[src];
    f->P[dst] = f->P[src]; f->D[dst] = f->D[src]; f->T[dst] = f->T[src];
    f->H[dst] = f->H[src]; f->Gamma[dst] = f->Gamma[src]; f->E[dst] = f->E[src];
    for (int s = 0; s < S; s++) f->Yi[dst * S + s] = f->Yi[src * S + s];
    if (flip_comp == 1) f->U[dst] = -f->U[src];
    if (flip_comp == 2) f->V[dst] = -f->V[src];
    if (flip_comp == 3) f->W[dst] = -f->W[src];
}

/* Source index along one axis for ghost g of a face. lo: ghost index g in [0,bc) maps to the
 * interior; hi: ghost index n+bc+g. */
static int ghost_src(int kind, int lo, int g, int n, int bc) {
    if (lo) {
        switch (kind) {
        case ORC_BC_RSBI: case ORC_BC_REFLECT: return 2 * bc - 1 - g;   /* mirror */
        case ORC_BC_OUTFLOW: return bc;
        case ORC_BC_PERIODIC: return n + g;                               /* g - bc + n */
        }
    } else {
        switch (kind) {
        case ORC_BC_RSBI: case ORC_BC_REFLECT: return 2 * (n + bc) - 1 - (n + bc + g);
        case ORC_BC_OUTFLOW: return n + bc - 1;
        case ORC_BC_PERIODIC: return bc + g;
        }
    }
    return -1;
}

/* Ghost fill on the primitive fields. Face order and extents follow
 * Flowfield::FieldBoundary_3dRSBI (src/Flowfield.cpp:559-773): x faces over interior j,k; y faces
 * over all i and interior k; z faces over all i, j. ORC_BC_RSBI on the y- face recomputes R, rho,
 * cp, H, E and gamma from (p, T, Y) with the NASA fits and mirrors Mc (:616-647). */
void orc_fill_ghosts(orc_flow *f, const orc_mech *m, const int kinds[6]) {
    const int S = m->ns, bc = f->bc, ni = f->ni, nj = f->nj, nk = f->nk;
    const uint32_t q = f->quirks;
    for (int side = 0; side < 2; side++) {                 /* x */
        int kind = kinds[side];
        if (kind == ORC_BC_NONE) continue;
        for (int g = 0; g < bc; g++) {
            int gi = side == 0 ? g : ni + bc + g;
            int si = ghost_src(kind, side == 0, g, ni, bc);
            for (int j = bc; j < nj + bc; j++)
                for (int k = bc; k < nk + bc; k++)
                    copy_prims(f, I3(f, gi, j, k), I3(f, si, j, k), S, kind == ORC_BC_REFLECT ? 1 : 0);
        }
    }
    for (int side = 0; side < 2; side++) {                 /* y */
        int kind = kinds[2 + side];
        if (kind == ORC_BC_NONE) continue;
        for (int g = 0; g < bc; g++) {
            int gj = side == 0 ? g : nj + bc + g;
            int sj = ghost_src(kind, side == 0, g, nj, bc);
            for (int i = 0; i < ni + 2 * bc; i++)
                for (int k = bc; k < nk + bc; k++) {
                    size_t d = I3(f, i, gj, k), s = I3(f, i, sj, k);
                    if (kind == ORC_BC_RSBI && side == 0) {
                        double Y[NSMAX], Cpi[NSMAX], Hi[NSMAX], Ei[NSMAX];
                        f->U[d] = f->U[s]; f->V[d] = f->V[s]; f->W[d] = f->W[s];
                        f->P[d] = f->P[s]; f->T[d] = f->T[s];
                        for (int sp = 0; sp < S; sp++) {
                            f->Yi[d * S + sp] = f->Yi[s * S + sp];
                            Y[sp] = f->Yi[d * S + sp];
                            f->Mc[d * S + sp] = f->Mc[s * S + sp];
                            Cpi[sp] = orc_getcpi(f->T[d], m->ri[sp], sp, m, q);
                            Hi[sp] = orc_gethi(f->T[d], m->ri[sp], sp, m, q);
                            Ei[sp] = Hi[sp] - m->ri[sp] * f->T[d];
                        }
                        f->Rgas[d] = orc_sum(2, Y, m->ri, S) * 1000;
                        f->D[d] = f->P[d] / (f->T[d] * f->Rgas[d]);
                        f->Cp[d] = orc_sum(2, Cpi, Y, S) * 1000;
                        f->H[d] = orc_sum(2, Y, Hi, S) * 1000;
                        f->E[d] = orc_sum(2, Y, Ei, S) * 1000;
                        f->Gamma[d] = f->Cp[d] / (f->Cp[d] - f->Rgas[d]);
                    } else {
                        copy_prims(f, d, s, S, kind == ORC_BC_REFLECT ? 2 : 0);
                    }
                }
        }
    }
    for (int side = 0; side < 2; side++) {                 /* z */
        int kind = kinds[4 + side];
        if (kind == ORC_BC_NONE) continue;
        for (int g = 0; g < bc; g++) {
            int gk = side == 0 ? g : nk + bc + g;
            int sk = ghost_src(kind, side == 0, g, nk, bc);
            for (int i = 0; i < ni + 2 * bc; i++)
                for (int j = 0; j < nj + 2 * bc; j++)
                    copy_prims(f, I3(f, i, j, gk), I3(f, i, j, sk), S, kind == ORC_BC_REFLECT ? 3 : 0);
        }
    }
}

/* ------------------------------------------------------------- state updates */
/* Flowfield::Update_after_Adv, src/Flowfield.cpp:1140-1157. Quirk F1: species only. */
void orc_update_after_adv(orc_flow *f, const orc_mech *m) {
    const int S = m->ns, nc = S + 4, bc = f->bc;
    const int ncomp = (f->quirks & ORC_Q_F1) ? S : nc;
    for (int i = bc; i < f->ni + bc; i++)
        for (int j = bc; j < f->nj + bc; j++)
            for (int k = bc; k < f->nk + bc; k++) {
                const size_t c = I3(f, i, j, k);
                for (int s = 0; s < ncomp; s++) f->CS[c * nc + s] = f->CS[c * nc + s] + f->RHS[c * nc + s];
            }
    orc_update_after_cs(f, m);
}

/* Flowfield::Explicit(), src/Flowfield.cpp:1169-1186. Quirk F2: rhoE goes into the rho*w slot
 * and CS(NS+3) is left stale; FIXED follows the per-cell version (:1189-1203). */
void orc_explicit(orc_flow *f, const orc_mech *m) {
    const int S = m->ns, nc = S + 4, bc = f->bc;
    for (int i = bc; i < f->ni + bc; i++)
        for (int j = bc; j < f->nj + bc; j++)
            for (int k = bc; k < f->nk + bc; k++) {
                const size_t c = I3(f, i, j, k);
                double *cs = &f->CS[c * nc];
                f->D[c] = 0.0;
                for (int s = 0; s < S; s++) {
                    cs[s] = f->Di[c * S + s];
                    f->D[c] += cs[s];
                }
                double D = f->D[c];
                cs[S] = f->U[c] * D;
                cs[S + 1] = f->V[c] * D;
                double et = D * f->E[c] + 0.5 * (pow(f->U[c], 2) + pow(f->V[c], 2) + pow(f->W[c], 2)) * D;
                if (f->quirks & ORC_Q_F2) {
                    cs[S + 2] = et;
                } else {
                    cs[S + 2] = f->W[c] * D;
                    cs[S + 3] = et;
                }
            }
    orc_update_after_cs(f, m);
}

/* Flowfield::Explicit(step,i,j,k), src/Flowfield.cpp:1189-1203 */
void orc_explicit_cell(orc_flow *f, const orc_mech *m, int i, int j, int k) {
    const int S = m->ns, nc = S + 4;
    const size_t c = I3(f, i, j, k);
    double *cs = &f->CS[c * nc];
    f->D[c] = 0.0;
    for (int s = 0; s < S; s++) {
        cs[s] = f->Di[c * S + s];
        f->D[c] += cs[s];
    }
    double D = f->D[c];
    cs[S] = f->U[c] * D;
    cs[S + 1] = f->V[c] * D;
    cs[S + 2] = f->W[c] * D;
    cs[S + 3] = D * f->E[c] + 0.5 * (pow(f->U[c], 2) + pow(f->V[c], 2) + pow(f->W[c], 2)) * D;
    orc_update_after_cs_cell(f, m, i, j, k);
}

/* ---------------------------------------------------------------- kinetics */
/* Per-cell rate evaluation shared by Trapezoid (:412-485, :489-591), TrapezoidPrediction
 * (:593-670) and Diagonalized (:736-811). rtb_mode: 0 = R_TB kept as summed (C1),
 * 1 = R_TB forced to 0 (C2), 2 = R_TB==0 -> 1 (the correct default, :773-774). */
typedef struct {
    double Hi[NSMAX], Si[NSMAX], Gi[NSMAX], Di[NSMAX], Mc[NSMAX];
    double KF[NRMAX], KB[NRMAX], Kp[NRMAX], Kc[NRMAX], RR_F[NRMAX], RR_B[NRMAX], R_TB[NRMAX], RR[NRMAX];
    double Wi[NSMAX];
} kin_t;

static void kin_rates(const orc_mech *m, uint32_t q, double T_local, const double *Di_in, const double *Mc_in,
                      int rtb_mode, kin_t *ip) {
    const int S = m->ns, NR = m->nr;
    for (int s = 0; s < S; s++) {
        ip->Di[s] = Di_in[s] * 1e-3;
        ip->Mc[s] = Mc_in[s] * 1e-6;
        ip->Hi[s] = orc_gethi(T_local, R_UNIV, s, m, q);
        ip->Si[s] = orc_getsi(T_local, R_UNIV, s, m, q);
        ip->Gi[s] = ip->Hi[s] - ip->Si[s] * T_local;
    }
    for (int r = 0; r < NR; r++) { ip->RR_F[r] = 1.0; ip->RR_B[r] = 1.0; ip->R_TB[r] = 0.0; }
    for (int r = 0; r < NR; r++) {
        double Xr = 0.0, delta_g = 0.0;
        ip->KF[r] = m->af[r] * pow(T_local, m->bf[r]) * exp(-m->eaf[r] / (RU_CAL * T_local));
        for (int s = 0; s < S; s++) {
            const double sb = m->stoi_b[s * NR + r], sf = m->stoi_f[s * NR + r];
            Xr += (sb - sf);
            delta_g += (sb - sf) * ip->Gi[s];
            ip->RR_F[r] *= pow(ip->Mc[s], sf);
            ip->RR_B[r] *= pow(ip->Mc[s], sb);
            ip->R_TB[r] += ip->Mc[s] * m->react_tb[s * NR + r];
        }
        ip->Kp[r] = exp(-delta_g / (R_UNIV * T_local)) * pow((P0_PA * 1e-6), Xr);
        ip->Kc[r] = ip->Kp[r] * pow(R_UNIV * T_local, -Xr);
        ip->KB[r] = ip->KF[r] / ip->Kc[r];
        if (rtb_mode == 1) ip->R_TB[r] = 0.0;
        else if (rtb_mode == 2 && ip->R_TB[r] == 0.0) ip->R_TB[r] = 1.0;
        ip->RR[r] = ip->KF[r] * ip->RR_F[r] - ip->KB[r] * ip->RR_B[r];
    }
    for (int s = 0; s < S; s++) {
        double temp = 0.0;
        for (int r = 0; r < NR; r++) temp += (m->stoi_b[s * NR + r] - m->stoi_f[s * NR + r]) * ip->R_TB[r] * ip->RR[r];
        ip->Wi[s] = m->mw[s] * temp;
    }
}

/* Point-implicit trapezoid update of one cell's partial densities; returns Di' in kg/m^3.
 * Quirk C3: R_TB multiplies only the second term of Q and of P. */
static void kin_trap(const orc_mech *m, uint32_t q, const kin_t *ip, double dt, double *Di_out) {
    const int S = m->ns, NR = m->nr;
    for (int s = 0; s < S; s++) {
        double P = 0, Q = 0;
        for (int r = 0; r < NR; r++) {
            const double sb = m->stoi_b[s * NR + r], sf = m->stoi_f[s * NR + r];
            if (q & ORC_Q_C3) {
                Q += (sb * ip->KF[r] * ip->RR_F[r] + sf * ip->KB[r] * ip->RR_B[r] * ip->R_TB[r]);
                P += (sb * ip->KB[r] * ip->RR_B[r] + sf * ip->KF[r] * ip->RR_F[r] * ip->R_TB[r]);
            } else {
                Q += (sb * ip->KF[r] * ip->RR_F[r] + sf * ip->KB[r] * ip->RR_B[r]) * ip->R_TB[r];
                P += (sb * ip->KB[r] * ip->RR_B[r] + sf * ip->KF[r] * ip->RR_F[r]) * ip->R_TB[r];
            }
        }
        Q = m->mw[s] * Q;
        if (ip->Mc[s] == 0) P = 0.0;
        else P = P / ip->Mc[s];
        Di_out[s] = ((1.0 - dt / 2.0 * P) * ip->Di[s] + dt * Q) / (1.0 + dt / 2.0 * P) * 1e3;
    }
}

/* Reaction::Trapezoid full-field, src/Reaction.cpp:412-485 (quirk C1 for the R_TB default).
 * Reads Mc, Di (flow units), T; writes the new Di and CMS. The reference returns its member Di by
 * value (quirk F5); the oracle writes the values into f->Di, which is the assignment's intent. */
void orc_trapezoid_full(orc_flow *f, const orc_mech *m, double dt) {
    const int S = m->ns, bc = f->bc;
    const uint32_t q = f->quirks;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = bc; i < f->ni + bc; i++)
        for (int j = bc; j < f->nj + bc; j++)
            for (int k = bc; k < f->nk + bc; k++) {
                const size_t c = I3(f, i, j, k);
                kin_t ip;
                kin_rates(m, q, f->T[c], &f->Di[c * S], &f->Mc[c * S], (q & ORC_Q_C1) ? 0 : 2, &ip);
                for (int s = 0; s < S; s++) f->CMS[c * S + s] = ip.Wi[s] * 1e3;
                kin_trap(m, q, &ip, dt, &f->Di[c * S]);
            }
}

/* Reaction::Trapezoid per cell, src/Reaction.cpp:489-591 (quirk C2: R_TB assigned 0) */
void orc_trapezoid_cell(orc_flow *f, const orc_mech *m, double dtm, int i, int j, int k) {
    const int S = m->ns;
    const uint32_t q = f->quirks;
    const size_t c = I3(f, i, j, k);
    double dt = dtm / f->Nchem[c];
    kin_t ip;
    kin_rates(m, q, f->T[c], &f->Di[c * S], &f->Mc[c * S], (q & ORC_Q_C2) ? 1 : 2, &ip);
    for (int s = 0; s < S; s++) f->CMS[c * S + s] = ip.Wi[s] * 1e3;
    kin_trap(m, q, &ip, dt, &f->Di[c * S]);
}

/* Reaction::TrapezoidPrediction, src/Reaction.cpp:593-670. FIXED caps Nchem at nchem_max. */
void orc_trapezoid_prediction(orc_flow *f, const orc_mech *m, double dtm) {
    const int S = m->ns, bc = f->bc;
    const uint32_t q = f->quirks;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = bc; i < f->ni + bc; i++)
        for (int j = bc; j < f->nj + bc; j++)
            for (int k = bc; k < f->nk + bc; k++) {
                const size_t c = I3(f, i, j, k);
                double dt = dtm, dt_temp;
                kin_t ip;
                kin_rates(m, q, f->T[c], &f->Di[c * S], &f->Mc[c * S], (q & ORC_Q_C2) ? 1 : 2, &ip);
                for (int s = 0; s < S; s++) {
                    if (f->Yi[c * S + s] >= 1e-6) {
                        if (ip.Wi[s] != 0) {
                            dt_temp = fabs(-ip.Di[s] / ip.Wi[s]);
                            if (dt_temp < dt) dt = dt_temp;
                        }
                    }
                }
                double nd = ceil(dtm / dt);
                if (f->nchem_max > 0 && nd > f->nchem_max) nd = f->nchem_max;
                f->Nchem[c] = (int)nd;
                if (f->Nchem[c] <= 0) abort();   /* src/Reaction.cpp:664-668 */
            }
}

/* Euler::Trapezoid(type=1) cell loop without load balancing, src/Euler.cpp:383,526-556 */
void orc_trapezoid_adaptive(orc_flow *f, const orc_mech *m, double dtm) {
    const int bc = f->bc;
    orc_trapezoid_prediction(f, m, dtm);
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
    for (int i = bc; i < f->ni + bc; i++)
        for (int j = bc; j < f->nj + bc; j++)
            for (int k = bc; k < f->nk + bc; k++) {
                const size_t c = I3(f, i, j, k);
                for (int step = 0; step < f->Nchem[c]; step++) {
                    orc_trapezoid_cell(f, m, dtm, i, j, k);
                    orc_explicit_cell(f, m, i, j, k);
                }
            }
}

/* Reaction::Diagonalized, src/Reaction.cpp:736-811. Reads Mc, Di, T, PT; writes CMS and the MD
 * diagonal. Quirk C6: the temperature path mixes g/cm^3 rates with an SI dT/drho; FIXED scales
 * that term by 1e3. */
void orc_diagonalized(orc_flow *f, const orc_mech *m) {
    const int S = m->ns, NR = m->nr, nc = S + 4, bc = f->bc;
    const uint32_t q = f->quirks;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = bc; i < f->ni + bc; i++)
        for (int j = bc; j < f->nj + bc; j++)
            for (int k = bc; k < f->nk + bc; k++) {
                const size_t c = I3(f, i, j, k);
                const double T_local = f->T[c];
                kin_t ip;
                double WJH1[NSMAX][NRMAX], WJH2[NSMAX][NRMAX];
                kin_rates(m, q, T_local, &f->Di[c * S], &f->Mc[c * S], 2, &ip);
                for (int s = 0; s < S; s++) f->CMS[c * S + s] = ip.Wi[s] * 1e3;
                for (int r = 0; r < NR; r++)
                    for (int s = 0; s < S; s++) {
                        const double sb = m->stoi_b[s * NR + r], sf = m->stoi_f[s * NR + r];
                        if (ip.Di[s] == 0)
                            WJH1[s][r] = 0.0;
                        else
                            WJH1[s][r] = ip.KF[r] * sf * ip.RR_F[r] / ip.Di[s] - ip.KB[r] * sb * ip.RR_B[r] / ip.Di[s];
                        WJH2[s][r] = ip.RR[r] * (m->bf[r] / T_local + m->eaf[r] / (RU_CAL * pow(T_local, 2)));
                    }
                const double tscale = (q & ORC_Q_C6) ? 1.0 : 1e3;
                for (int s = 0; s < S; s++) {
                    double temp = 0.0;
                    const double pt = f->PT[c * nc + s] * tscale;
                    for (int r = 0; r < NR; r++)
                        temp += (m->stoi_b[s * NR + r] - m->stoi_f[s * NR + r]) *
                                ((WJH1[s][r] + WJH2[s][r] * pt) * ip.R_TB[r] + ip.RR[r] * m->react_tb[s * NR + r] / m->mw[s]);
                    f->MD[c * S + s] = m->mw[s] * temp;
                }
            }
}

/* Flowfield::Update_IMEX, src/Flowfield.cpp:1215-1234. MD(c,c) for c >= NS is never written in
 * the reference (quirk I2) and is taken as 0, so those components take RHS unchanged. */
void orc_update_imex(orc_flow *f, const orc_mech *m) {
    const int S = m->ns, nc = S + 4, bc = f->bc;
    const double dt = f->dt;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = bc; i < f->ni + bc; i++)
        for (int j = bc; j < f->nj + bc; j++)
            for (int k = bc; k < f->nk + bc; k++) {
                const size_t c = I3(f, i, j, k);
                for (int s = 0; s < S; s++) f->RHS[c * nc + s] = f->RHS[c * nc + s] + f->CMS[c * S + s] * dt;
                for (int s = 0; s < nc; s++) {
                    double md = (s < S) ? f->MD[c * S + s] : 0.0;
                    f->CS[c * nc + s] = f->CS[c * nc + s] + f->RHS[c * nc + s] / (1.0 - dt * md);
                }
            }
    orc_update_after_cs(f, m);
    orc_get_partial_t(f, m);
}

/* Conserved state from primitives over every cell including ghosts, as the second loop of
 * Flowfield::FieldInitial (src/Flowfield.cpp:193-226): reads T, P, U, V, W, Yi; writes Wav, Rgas,
 * D, Di, Mc, Cp, H, E, Gamma, C, CS. */
void orc_field_from_prims(orc_flow *f, const orc_mech *m) {
    const int S = m->ns, nc = S + 4;
    const uint32_t q = f->quirks;
    for (int i = 0; i < f->ni + 2 * f->bc; i++)
        for (int j = 0; j < f->nj + 2 * f->bc; j++)
            for (int k = 0; k < f->nk + 2 * f->bc; k++) {
                const size_t c = I3(f, i, j, k);
                double Y[NSMAX], Cpi[NSMAX], Hi[NSMAX], Ei[NSMAX];
                for (int s = 0; s < S; s++) Y[s] = f->Yi[c * S + s];
                f->Wav[c] = 1.0 / orc_sum(3, Y, m->mw, S);
                f->Rgas[c] = R_UNIV * 1000 / f->Wav[c];
                f->D[c] = f->P[c] / (f->Rgas[c] * f->T[c]);
                for (int s = 0; s < S; s++) {
                    f->Di[c * S + s] = Y[s] * f->D[c];
                    f->Mc[c * S + s] = f->Di[c * S + s] / m->mw[s] * 1000;
                    Cpi[s] = orc_getcpi(f->T[c], m->ri[s], s, m, q);
                    Hi[s] = orc_gethi(f->T[c], m->ri[s], s, m, q);
                    Ei[s] = Hi[s] - m->ri[s] * f->T[c];
                }
                f->Cp[c] = orc_sum(2, Cpi, Y, S) * 1000;
                f->H[c] = orc_sum(2, Y, Hi, S) * 1000;
                f->E[c] = orc_sum(2, Y, Ei, S) * 1000;
                f->Gamma[c] = f->Cp[c] / (f->Cp[c] - R_UNIV * orc_sum(3, Y, m->mw, S) * 1000);
                f->C[c] = sqrt(f->Gamma[c] * f->P[c] / f->D[c]);
                for (int s = 0; s < S; s++) f->CS[c * nc + s] = f->D[c] * Y[s];
                f->CS[c * nc + S] = f->D[c] * f->U[c];
                f->CS[c * nc + S + 1] = f->D[c] * f->V[c];
                f->CS[c * nc + S + 2] = f->D[c] * f->W[c];
                f->CS[c * nc + S + 3] = f->D[c] * f->E[c] +
                                        0.5 * (pow(f->U[c], 2) + pow(f->V[c], 2) + pow(f->W[c], 2)) * f->D[c];
            }
}

/* Check of the device kernels' division by the constant divisors 3 and 5 of the NASA-9 integrals
 * (x * RN(1/d) refined by one FMA residual step) against IEEE division: n operands drawn by a
 * xorshift generator over signs, mantissas and exponents 2^-300..2^300, then the n/4 mantissas
 * next to 1.0 and next to 2.0. Returns the number of operands where the two differ. */
long orc_check_div_const(double d, long n, unsigned long long seed) {
    uint64_t s = seed ? seed : 88172645463325252ull;
    const double R = 1.0 / d;
    long bad = 0;
    for (long i = 0; i < n + n / 2; i++) {
        uint64_t b;
        if (i < n) {
            s ^= s << 13; s ^= s >> 7; s ^= s << 17;
            const uint64_t r = s;
            s ^= s << 13; s ^= s >> 7; s ^= s << 17;
            const uint64_t e = 1023 - 300 + s % 600;
            b = (r & 0x800FFFFFFFFFFFFFull) | (e << 52);
        } else {
            const uint64_t m = (uint64_t)(i - n) / 2;
            b = (i - n) % 2 ? (1023ull << 52) | m : (1023ull << 52) | (0xFFFFFFFFFFFFFull - m);
        }
        double x;
        memcpy(&x, &b, sizeof x);
        const double q = x * R;
        const double r = fma(-q, d, x);
        if (fma(r, R, q) != x / d) bad++;
    }
    return bad;
}

/* Check of the device flux kernels' division a / b (a reciprocal estimate refined by two Newton
 * steps, then one residual correction of the quotient) against IEEE division: n operand pairs
 * over exponents 2^-60..2^60, with the reciprocal estimate perturbed by a relative error up to
 * rel_err (the hardware estimate is far better than the 1e-4 the test uses). Returns the number
 * of pairs where the two quotients differ. */
long orc_check_div_fast(double rel_err, long n, unsigned long long seed) {
    uint64_t s = seed ? seed : 88172645463325252ull;
    long bad = 0;
    for (long i = 0; i < n; i++) {
        double ab[2];
        for (int k = 0; k < 2; k++) {
            s ^= s << 13; s ^= s >> 7; s ^= s << 17;
            const uint64_t r = s;
            s ^= s << 13; s ^= s >> 7; s ^= s << 17;
            const uint64_t e = 1023 - 60 + s % 120;
            const uint64_t b = (r & 0x800FFFFFFFFFFFFFull) | (e << 52);
            memcpy(&ab[k], &b, sizeof(double));
        }
        s ^= s << 13; s ^= s >> 7; s ^= s << 17;
        const double eps = ((double)(s >> 11) / 9007199254740992.0 * 2 - 1) * rel_err;
        const double a = ab[0], b = ab[1];
        double r = (1.0 / b) * (1 + eps);
        r = fma(fma(-b, r, 1.0), r, r);
        r = fma(fma(-b, r, 1.0), r, r);
        const double q = a * r;
        if (fma(fma(-b, q, a), r, q) != a / b) bad++;
    }
    return bad;
}
