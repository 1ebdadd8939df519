/* CPU oracle: FP64 restatement of the EU3D reference hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This library checks the HIP path: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it. The product library
 * (eu3d_amd/libeu3d.so) never links, loads or calls it.
 *
 * Layout follows the reference Array (include/Array.hpp:33-46): row-major, last index fastest.
 * 3-D fields are X(i,j,k) with extents (ni+2bc, nj+2bc, nk+2bc); 4-D fields are X(i,j,k,c).
 * Every function cites the reference lines it restates. Quirk bits select the reference's
 * exact (REF) behaviour or the minimal physical correction (FIXED) per item.
 */
#ifndef EU3D_ORACLE_H
#define EU3D_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Quirk bits (1 = reference behaviour). */
#define ORC_Q_F1   (1u << 0)  /* Update_after_Adv adds RHS to species only (src/Flowfield.cpp:1146) */
#define ORC_Q_F1B  (1u << 1)  /* Y_L/Y_R copied only for i < ni (src/Flowfield.cpp:1353) */
#define ORC_Q_F2   (1u << 2)  /* Explicit() writes rhoE into the rho*w slot (src/Flowfield.cpp:1183) */
#define ORC_Q_F3   (1u << 3)  /* Newton uses the previous E (src/Flowfield.cpp:1268 vs :1272) */
#define ORC_Q_F6   (1u << 4)  /* CFL ignores z (src/Flowfield.cpp:254) */
#define ORC_Q_C1   (1u << 5)  /* full-field trapezoid: R_TB==0 default is a no-op (src/Reaction.cpp:449-450) */
#define ORC_Q_C2   (1u << 6)  /* per-cell trapezoid / prediction: R_TB assigned 0 (src/Reaction.cpp:528,634) */
#define ORC_Q_C3   (1u << 7)  /* R_TB multiplies one term of Q and of P (src/Reaction.cpp:469-470,575-576) */
#define ORC_Q_C4   (1u << 8)  /* high-T h, s use the low-range b1, b2 (src/Reaction.cpp:923,945) */
#define ORC_Q_C6   (1u << 9)  /* MD diagonal mixes g/cm^3 and SI units (src/Reaction.cpp:798-808) */
#define ORC_Q_C7   (1u << 10) /* dT/drho_s uses h_s in kJ/kg (src/Flowfield.cpp:1589-1590) */
#define ORC_Q_REF  0x7FFu
#define ORC_Q_FIXED 0u

/* Ghost-fill kinds per face (x-, x+, y-, y+, z-, z+). */
#define ORC_BC_RSBI     0   /* reference FieldBoundary_3dRSBI mirror (src/Flowfield.cpp:559-773) */
#define ORC_BC_OUTFLOW  1   /* zero-gradient copy of the first interior cell */
#define ORC_BC_REFLECT  2   /* mirror with the normal velocity negated */
#define ORC_BC_PERIODIC 3
#define ORC_BC_NONE     4   /* leave ghosts untouched */

typedef struct {
    int ns, nr;
    const double *stoi_f, *stoi_b;   /* [ns*nr], (s,r) -> s*nr + r */
    const double *af, *bf, *eaf;     /* [nr] */
    const double *react_tb;          /* [ns*nr] */
    const double *coeff0, *coeff1;   /* [9*ns], (k,s) -> k*ns + s */
    const double *mw, *ri;           /* [ns] */
} orc_mech;

typedef struct {
    int ni, nj, nk, bc, ns;
    const double *xnode, *ynode, *znode;           /* n+2bc each */
    /* 3-D fields */
    double *U, *V, *W, *P, *D, *T, *C, *Ma, *Wav, *Rgas, *Cp, *H, *E, *Gamma;
    /* 4-D fields */
    double *Mc, *Yi, *Di;                          /* ns */
    double *CS, *F, *G, *Q, *RHS, *PT;             /* ns+4 */
    double *CMS, *MD;                              /* ns (MD: diagonal entries only) */
    int *Nchem;                                    /* 3-D */
    double dt, time;
    uint32_t quirks;
    int nchem_max;                                 /* FIXED cap on adaptive sub-steps */
} orc_flow;

double orc_getcpi(double T, double R, int sp, const orc_mech *m, uint32_t q);
double orc_gethi(double T, double R, int sp, const orc_mech *m, uint32_t q);
double orc_getsi(double T, double R, int sp, const orc_mech *m, uint32_t q);
double orc_sum(int flag, const double *a, const double *b, int n);
double orc_get_temp(const orc_mech *m, uint32_t q, double T, double E, const double *Y, int *iters);

void orc_update_after_cs(orc_flow *f, const orc_mech *m);
void orc_update_after_cs_cell(orc_flow *f, const orc_mech *m, int i, int j, int k);
void orc_get_partial_t(orc_flow *f, const orc_mech *m);
void orc_muscl(int order, int dir, const orc_flow *f, double *FLR, const double *fi, int limit);
void orc_ausm(orc_flow *f, int dir, double *Fi, int order, int limit);
void orc_ee(orc_flow *f);
void orc_advection(orc_flow *f, int diff_order, int limit);
void orc_cfl(orc_flow *f, double cfl, double final_time);
void orc_fill_ghosts(orc_flow *f, const orc_mech *m, const int kinds[6]);
void orc_update_after_adv(orc_flow *f, const orc_mech *m);
void orc_explicit(orc_flow *f, const orc_mech *m);
void orc_explicit_cell(orc_flow *f, const orc_mech *m, int i, int j, int k);
void orc_diagonalized(orc_flow *f, const orc_mech *m);
void orc_update_imex(orc_flow *f, const orc_mech *m);
void orc_trapezoid_full(orc_flow *f, const orc_mech *m, double dt);
void orc_trapezoid_prediction(orc_flow *f, const orc_mech *m, double dtm);
void orc_trapezoid_cell(orc_flow *f, const orc_mech *m, double dtm, int i, int j, int k);
void orc_trapezoid_adaptive(orc_flow *f, const orc_mech *m, double dtm);
void orc_field_from_prims(orc_flow *f, const orc_mech *m);

int orc_num_threads(void);
long orc_check_div_const(double d, long n, unsigned long long seed);
long orc_check_div_fast(double rel_err, long n, unsigned long long seed);
void orc_set_num_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
