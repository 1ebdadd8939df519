/* eu3d.h - C ABI of the MI355X (gfx950) implementation of the EU3D per-step hot path.
 *
 * The reference exposes C++ classes whose methods operate on Array<> members
 * (include/Flowfield.hpp:142-211, include/Reaction.hpp:143-193, include/TimeAdv.hpp:20-23) and
 * a driver that reads their private state as a friend (include/Flowfield.hpp:139). This ABI
 * replaces those method bodies: one context owns the device state of one Flowfield + Reaction
 * pair; each entry point below is the body of the reference method it cites. Plain pointers and
 * sizes only; no exceptions cross the boundary.
 *
 * Layout at the boundary is the reference Array layout (include/Array.hpp:33-46): row-major,
 * last index fastest, X(i,j,k) or X(i,j,k,c) with ghost-inclusive extents
 * (ni+2bc, nj+2bc, nk+2bc). On the device every field is SoA with x fastest.
 *
 * Return value: 0 on success, a negative EU3D_ERR_* code on failure; eu3d_last_error() gives
 * the message. All work is ordered on the context's HIP stream. Entry points that return data
 * to the host synchronise the stream; the others return once the work is queued. The context
 * and network streams are blocking streams: they start after work the caller queued on the null
 * stream. A caller that produces the device buffers it passes in on another non-blocking
 * stream synchronises that stream first.
 */
#ifndef EU3D_H
#define EU3D_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EU3D_OK 0
#define EU3D_ERR_ARG (-1)
#define EU3D_ERR_HIP (-2)
#define EU3D_ERR_MECH_MISMATCH (-3)
#define EU3D_ERR_NONFINITE (-4)   /* NaN/inf in D or T: the reference aborts (src/Euler.cpp:319-331) */
#define EU3D_ERR_NCHEM (-5)       /* Nchem <= 0: the reference aborts (src/Reaction.cpp:664-668) */
#define EU3D_ERR_UNSUPPORTED (-6)

/* Quirk bits: 1 selects the reference's arithmetic, 0 the minimal physical correction. */
#define EU3D_Q_F1 (1u << 0)   /* src/Flowfield.cpp:1146 RHS added to species only */
#define EU3D_Q_F1B (1u << 1)  /* src/Flowfield.cpp:1353 species face states only for i < ni */
#define EU3D_Q_F2 (1u << 2)   /* src/Flowfield.cpp:1183 rhoE written into the rho*w slot */
#define EU3D_Q_F3 (1u << 3)   /* src/Flowfield.cpp:1268 Newton from the previous E */
#define EU3D_Q_F6 (1u << 4)   /* src/Flowfield.cpp:254 CFL ignores z */
#define EU3D_Q_C1 (1u << 5)   /* src/Reaction.cpp:449-450 R_TB default no-op (full-field trapezoid) */
#define EU3D_Q_C2 (1u << 6)   /* src/Reaction.cpp:528,634 R_TB assigned 0 (per-cell trapezoid) */
#define EU3D_Q_C3 (1u << 7)   /* src/Reaction.cpp:469-470 R_TB on one term of Q and of P */
#define EU3D_Q_C4 (1u << 8)   /* src/Reaction.cpp:923,945 high-T h,s use low-range b1,b2 */
#define EU3D_Q_C6 (1u << 9)   /* src/Reaction.cpp:798-808 mixed units in the MD diagonal */
#define EU3D_Q_C7 (1u << 10)  /* src/Flowfield.cpp:1589-1590 h_s in kJ/kg in dT/drho_s */
#define EU3D_QUIRKS_REF 0x7FFu
#define EU3D_QUIRKS_FIXED 0u

/* Field ids for eu3d_set_field / eu3d_get_field: name, component count, reference member. */
enum eu3d_field {
    EU3D_FIELD_CS = 0,     /* NS+4  Flowfield::CS conserved state rho_s, rho u, rho v, rho w, rho E */
    EU3D_FIELD_T = 1,      /* 1     Flowfield::T  */
    EU3D_FIELD_E = 2,      /* 1     Flowfield::E  internal energy per unit mass */
    EU3D_FIELD_U = 3, EU3D_FIELD_V = 4, EU3D_FIELD_W = 5,
    EU3D_FIELD_P = 6, EU3D_FIELD_D = 7, EU3D_FIELD_H = 8, EU3D_FIELD_GAMMA = 9,
    EU3D_FIELD_C = 10, EU3D_FIELD_MA = 11, EU3D_FIELD_RGAS = 12, EU3D_FIELD_CP = 13,
    EU3D_FIELD_WAV = 14,
    EU3D_FIELD_YI = 15,    /* NS    Flowfield::Yi */
    EU3D_FIELD_DI = 16,    /* NS    Flowfield::Di */
    EU3D_FIELD_MC = 17,    /* NS    Flowfield::Mc */
    EU3D_FIELD_PT = 18,    /* NS+4  Flowfield::Partion_T */
    EU3D_FIELD_RHS = 19,   /* NS+4  Flowfield::RHS */
    EU3D_FIELD_F = 20, EU3D_FIELD_G = 21, EU3D_FIELD_Q = 22, /* NS+4 face fluxes */
    EU3D_FIELD_CMS = 23,   /* NS    Reaction::CMS */
    EU3D_FIELD_MD = 24,    /* NS    diagonal of Reaction::MD */
    EU3D_FIELD_NCHEM = 25, /* 1     Reaction::Nchem, exchanged as double */
    EU3D_FIELD_COUNT = 26
};

/* Ghost-fill kinds per face, order x-, x+, y-, y+, z-, z+. */
#define EU3D_BC_RSBI 0      /* Flowfield::FieldBoundary_3dRSBI, src/Flowfield.cpp:559-773 */
#define EU3D_BC_OUTFLOW 1   /* zero-gradient copy of the first interior cell */
#define EU3D_BC_REFLECT 2   /* mirror with the normal velocity negated */
#define EU3D_BC_PERIODIC 3
#define EU3D_BC_NONE 4

/* Chemistry coupling for the fused step. */
#define EU3D_CHEM_NONE 0
#define EU3D_CHEM_IMEX 1          /* Euler::IMEX, src/Euler.cpp:748-760 */
#define EU3D_CHEM_TRAPEZOID 2     /* Euler::Trapezoid(0), src/Euler.cpp:369-379 */
#define EU3D_CHEM_TRAP_ADAPTIVE 3 /* Euler::Trapezoid(1) cell loop, src/Euler.cpp:383,526-556 */
#define EU3D_CHEM_SURROGATE 4     /* MLP surrogate with FP64 fallback (the reference's "DNN" slot,
                                     src/Euler.cpp:129-132,773-780, is a stub; this has no reference
                                     counterpart) */

typedef struct eu3d_ctx eu3d_ctx;

/* Grid: interior node counts, ghost width (must be 2) and ghost-inclusive node coordinates,
 * as built by Mesh::Initial/Boundary (src/Mesh.cpp:75-121). */
typedef struct {
    int ni, nj, nk, bc;
    const double *xnode, *ynode, *znode;   /* ni+2bc, nj+2bc, nk+2bc */
} eu3d_grid;

/* Mechanism as parsed by Reaction::ReactionRead (src/Reaction.cpp:150-326). */
typedef struct {
    int ns, nr;
    const double *stoi_f, *stoi_b;   /* [ns*nr] (s,r) row-major */
    const double *af, *bf, *eaf;     /* [nr] */
    const double *react_tb;          /* [ns*nr] */
    const double *coeff0, *coeff1;   /* [9*ns] (k,s) row-major */
    const double *mw;                /* [ns] */
} eu3d_mech;

/* Context lifetime: replaces Flowfield::Construction / Reaction::ReactionConstruction
 * (src/Flowfield.cpp:72-145, src/Reaction.cpp:49-99). Allocates every field on the current HIP
 * device. Fails with EU3D_ERR_MECH_MISMATCH unless the mechanism equals the compiled-in one. */
int eu3d_create(eu3d_ctx **out, const eu3d_grid *grid, const eu3d_mech *mech, uint32_t quirks);
/* Sizes a context for this grid takes, from host arithmetic alone (no GPU call): the number of
 * per-block CFL partial minima its kernels write (one per 64x4 block of the interior, per k
 * layer) and the device bytes eu3d_create allocates (fields, nodes, scalars, compaction list,
 * partials, RK2 copy; the lazily grown chemistry and surrogate lists come on top). Supports
 * planning the memory of Flowfield::Construction (src/Flowfield.cpp:72-145) on one device. */
int eu3d_grid_sizes(int ni, int nj, int nk, int bc, long long *cfl_partials, unsigned long long *device_bytes);
int eu3d_destroy(eu3d_ctx *ctx);
const char *eu3d_last_error(const eu3d_ctx *ctx);
int eu3d_sync(eu3d_ctx *ctx);
int eu3d_set_quirks(eu3d_ctx *ctx, uint32_t quirks);
int eu3d_field_components(const eu3d_ctx *ctx, int field);

/* Field exchange in reference layout (the friend access of src/Euler.cpp:319-326,833-903). */
int eu3d_set_field(eu3d_ctx *ctx, int field, const double *host);
int eu3d_get_field(eu3d_ctx *ctx, int field, double *host);
int eu3d_set_time(eu3d_ctx *ctx, double dt, double time);   /* Flowfield::dt, ::time */
int eu3d_get_time(eu3d_ctx *ctx, double *dt, double *time);

/* Flowfield::FieldInitial second loop (src/Flowfield.cpp:193-229): from the primitive fields
 * T, P, U, V, W, Yi (set over every cell, ghosts included) to the conserved state CS and the
 * derived fields D, Di, Mc, Wav, Rgas, Cp, H, E, Gamma, C; then GetPartial_T. */
int eu3d_field_initial(eu3d_ctx *ctx);

/* Flowfield::CFLcondition, src/Flowfield.cpp:242-272 (dt and time stay on the device). */
int eu3d_cfl_dt(eu3d_ctx *ctx, double cfl, double final_time);
/* Flowfield::FieldBoundary_3dRSBI (src/Flowfield.cpp:559-773) and the build's BC kinds. */
int eu3d_fill_ghosts(eu3d_ctx *ctx, const int kinds[6]);
/* Flowfield::Advection, src/Flowfield.cpp:1096-1128: MUSCL_1/MUSCL_2 (include/Differ.hpp) +
 * AUSM (src/Flowfield.cpp:1330-1502) per direction + TimeAdv::EE (src/TimeAdv.cpp:29-43).
 * diff: 0 = MUSCL_1, 1 = MUSCL_2; limiter 0..4 as include/Differ.hpp:188-220. */
int eu3d_advection(eu3d_ctx *ctx, int diff, int limiter);
/* The same RHS as eu3d_advection from one fused pass (no face-flux arrays): the form
 * eu3d_step uses. F, G and Q are left untouched. */
int eu3d_advection_fused(eu3d_ctx *ctx, int diff, int limiter);
/* Flowfield::Update_after_Adv, src/Flowfield.cpp:1140-1157 */
int eu3d_update_after_adv(eu3d_ctx *ctx);
/* Flowfield::Update_after_CS + Get_temp, src/Flowfield.cpp:1246-1284,1542-1566 */
int eu3d_update_after_cs(eu3d_ctx *ctx);
/* Flowfield::GetPartial_T, src/Flowfield.cpp:1578-1595 */
int eu3d_get_partial_t(eu3d_ctx *ctx);
/* Flowfield::Explicit(), src/Flowfield.cpp:1169-1186 (includes Update_after_CS) */
int eu3d_explicit(eu3d_ctx *ctx);
/* Reaction::Diagonalized, src/Reaction.cpp:736-811 */
int eu3d_diagonalized(eu3d_ctx *ctx);
/* Flowfield::Update_IMEX, src/Flowfield.cpp:1215-1234 (includes Update_after_CS, GetPartial_T) */
int eu3d_update_imex(eu3d_ctx *ctx);
/* Reaction::Trapezoid full field, src/Reaction.cpp:412-485, with dt = Flowfield::dt; writes Di
 * (the assignment at src/Euler.cpp:375) and CMS. */
int eu3d_trapezoid(eu3d_ctx *ctx);
/* Reaction::TrapezoidPrediction, src/Reaction.cpp:593-670 -> Nchem. nchem_max > 0 caps it. */
int eu3d_trapezoid_prediction(eu3d_ctx *ctx, int nchem_max);
/* Euler::Trapezoid(type 1) cell loop (src/Euler.cpp:383,526-556): prediction, then per cell
 * Nchem x {Reaction::Trapezoid(cell) (src/Reaction.cpp:489-591); Flowfield::Explicit(cell)
 * (src/Flowfield.cpp:1189-1203)}. On one GPU the MPI load balancer (src/LoadBalance.cpp) is
 * replaced by active-cell compaction and a persistent kernel. */
int eu3d_trapezoid_adaptive(eu3d_ctx *ctx, int nchem_max);
/* NaN checks of src/Euler.cpp:319-331: EU3D_ERR_NONFINITE if D or T is not finite. */
int eu3d_check_finite(eu3d_ctx *ctx);

/* One full time step of Euler::Computing (src/Euler.cpp:266-333) on the device, with no host
 * round trip: CFL dt (when cfl > 0; else the stored dt), ghost fill, advection, chemistry.
 * rk: 1 = explicit Euler (TimeAdv::EE), 2 = two-stage SSP Runge-Kutta built from EE.
 * The steps do not need the derived fields (Di, Mc, Wav, Rgas, Cp, Ma, Partion_T); eu3d_step
 * refreshes them once after its last step unless skip_derive is nonzero (for a caller stepping
 * one call at a time, on every call but the last). */
typedef struct {
    double cfl, final_time;
    int bc_kinds[6];
    int diff, limiter, rk, chem, nchem_max;
    int skip_derive;
} eu3d_step_params;
int eu3d_step(eu3d_ctx *ctx, const eu3d_step_params *p, int nsteps);

/* Statistics of the last adaptive chemistry call. */
typedef struct {
    long long cells, active_cells, substeps_total;
    int nchem_max_seen;
} eu3d_chem_stats;
int eu3d_get_chem_stats(eu3d_ctx *ctx, eu3d_chem_stats *out);

/* ---- Batched cell-local chemistry (no grid): the reaction step alone on n independent cells
 * at rest. Device pointers, SoA: rho_s [ns][n] (kg/m^3), T [n], E [n] (J/kg, internal energy).
 * Adaptive trapezoid exactly as eu3d_trapezoid_adaptive per cell; nchem_out [n] optional.
 * Work is scheduled by active-cell compaction into a persistent lane-level work queue; the
 * launch shape comes from eu3d_chem_tune or the built-in default. */
typedef struct {
    int block;          /* threads per block */
    int blocks_per_cu;  /* persistent grid = blocks_per_cu * CU count */
    int waves;          /* occupancy target of the persistent kernel's register allocation:
                           1, 2 or 4 waves per SIMD (other values select 2) */
    int tail;           /* cells needing more sub-steps than this run one cell per wavefront
                           (reactions across the lanes); 0 keeps every cell in the persistent
                           kernel */
} eu3d_chem_schedule;
int eu3d_chem_batch(eu3d_ctx *ctx, long long n, double *rho_s, double *T, double *E, double dtm,
                    int nchem_max, int *nchem_out, eu3d_chem_stats *stats);
int eu3d_chem_set_schedule(eu3d_ctx *ctx, const eu3d_chem_schedule *s);

/* eu3d_chem_batch with no context and no allocation. No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). Same cells, arguments and results as eu3d_chem_batch; its
 * scratch (counters, the compaction list, per-cell sub-step counts and the tail queue: 12 bytes a
 * cell) lives in one static device array per GPU sized for n <= EU3D_NOALLOC_MAX_CELLS (2^22,
 * 48 MiB); larger n returns EU3D_ERR_ARG, as do bad pointers, dtm <= 0 or a schedule that
 * eu3d_chem_set_schedule would refuse. quirks as eu3d_create (the compiled mechanism); sched NULL
 * takes the default schedule; stream is a hipStream_t (NULL: the null stream).
 * With stats NULL the call only enqueues kernels on the stream (no copy, no synchronisation);
 * with stats it waits, reads the counters back and reports Nchem <= 0 as EU3D_ERR_NCHEM.
 * The first call on a device looks up the array's address (hipGetSymbolAddress). Every call on
 * one GPU shares that GPU's array: calls that overlap in time on one GPU race, so serialise them
 * (one stream, or events between streams). */
#define EU3D_NOALLOC_MAX_CELLS (1ll << 22)
int eu3d_chem_batch_noalloc(long long n, double *rho_s, double *T, double *E, double dtm, int nchem_max,
                            uint32_t quirks, const eu3d_chem_schedule *sched, void *stream, int *nchem_out,
                            eu3d_chem_stats *stats);
int eu3d_chem_get_schedule(eu3d_ctx *ctx, eu3d_chem_schedule *s);

/* ---- MLP chemistry surrogate (no reference counterpart: the reference's "DNN" scheme only
 * advects, src/Euler.cpp:773-780). 12 features (T, ln p/atm, 9 ln max(Y_k, 1e-20), ln dt),
 * 4 hidden layers of 256 with GELU, 9 outputs (normalised dY_k or d ln(Y_k + eps)). BF16 weights and activations
 * on MFMA with FP32 accumulation; FP32 master weights and Adam. Device pointers throughout. */
typedef struct eu3d_mlp eu3d_mlp;
typedef struct {
    float feat_mean[12], feat_std[12];   /* feature normalisation */
    float out_mean[9], out_std[9];       /* output (dY_k) normalisation */
    float hull_lo[12], hull_hi[12];      /* training hull, in normalised features */
    double dT_max;                       /* largest accepted |dT| of one step [K] */
    double elem_tol;                     /* H/O/N conservation bound, relative to the cell's content */
    double y_neg_tol;                    /* most negative accepted mass fraction */
    double T_cold;                       /* cells below this temperature keep their state */
    int nchem_threshold;                 /* 0: every cell goes through the network; n > 0: only
                                            cells whose predicted sub-step count exceeds n, the
                                            others take the FP64 sub-steps */
    double out_eps;                      /* 0: the outputs are normalised dY_k; eps > 0: they are
                                            normalised d ln(max(Y_k, 0) + eps), so
                                            Y_k' = (max(Y_k, 0) + eps) exp(.) - eps
                                            stays above -eps, and the element projection is weighted
                                            by each species' abundance Y_k' + eps */
    int nchem_threshold_hi;              /* with nchem_threshold > 0: cells needing more than this
                                            many sub-steps stay on the FP64 path too (0: no bound) */
} eu3d_surrogate_params;
typedef struct {
    long long cells, fallback, hull, negative, element, dT;
    long long network;                   /* cells sent through the network */
} eu3d_surrogate_stats;
int eu3d_mlp_create(eu3d_mlp **out, uint64_t seed);
int eu3d_mlp_destroy(eu3d_mlp *m);
size_t eu3d_mlp_param_count(const eu3d_mlp *m);
int eu3d_mlp_get_params(eu3d_mlp *m, float *host);      /* per layer: W^T [dout x din] then bias */
int eu3d_mlp_set_params(eu3d_mlp *m, const float *host);
int eu3d_mlp_get_grads(eu3d_mlp *m, float *host);
/* Inference: BF16 features [n x 32] -> FP32 outputs [n x 16], the whole network in one fused
 * kernel (activations kept in LDS). eu3d_mlp_forward_layers runs the same network layer by layer
 * through the GEMM kernel that training uses. */
int eu3d_mlp_forward(eu3d_mlp *m, int n, const void *feats_bf16, float *out_f32);
int eu3d_mlp_forward_layers(eu3d_mlp *m, int n, const void *feats_bf16, float *out_f32);
int eu3d_mlp_train_step(eu3d_mlp *m, int n, const void *feats_bf16, const float *targets, float lr, float wd,
                        float *loss_out);
int eu3d_gemm_bf16(int M, int N, int K, const void *A, const void *B, int b_kc, float *C);
void *eu3d_mlp_stream(eu3d_mlp *m);
int eu3d_mlp_set_surrogate(eu3d_mlp *m, const eu3d_surrogate_params *p);
int eu3d_surrogate_features(eu3d_mlp *m, long long n, const double *rho, const double *T, double dt, void *feats);
int eu3d_surrogate_targets(eu3d_mlp *m, long long n, const double *rho0, const double *rho1, float *targets);
int eu3d_surrogate_batch(eu3d_ctx *ctx, eu3d_mlp *m, long long n, double *rho_s, double *T, double *E, double dtm,
                         int nchem_max, eu3d_surrogate_stats *stats);
/* Attach a trained surrogate to a context for eu3d_step with EU3D_CHEM_SURROGATE (NULL detaches). */
int eu3d_set_surrogate(eu3d_ctx *ctx, eu3d_mlp *m);
/* The chemistry half of a surrogate step on the grid's interior (after eu3d_update_after_adv),
 * over the context's dt, with the FP64 adaptive trapezoid on every rejected cell. Takes the
 * place of the reference's DNN chemistry slot (src/Euler.cpp:773-780). */
int eu3d_surrogate_chem(eu3d_ctx *ctx, int nchem_max, eu3d_surrogate_stats *stats);
int eu3d_get_surrogate_stats(eu3d_ctx *ctx, eu3d_surrogate_stats *stats);

/* Device memory of a field (SoA, x fastest; component c at offset c * n3) for zero-copy use. */
double *eu3d_device_field(eu3d_ctx *ctx, int field);
size_t eu3d_cells_with_ghosts(const eu3d_ctx *ctx);
void *eu3d_stream(eu3d_ctx *ctx);

/* Profiling hook (no reference counterpart): mean duration in ms of one kernel of the plain step,
 * timed with HIP events on the context stream around each of `reps` launches on identical
 * inputs. FLUX_Z / FLUX_XY are the two passes of eu3d_advection_fused; IMEX_FUSED is the fused
 * Diagonalized + Update_IMEX + Update_after_CS cell kernel of the IMEX step and TRAP_FIRST the
 * prediction pass of the adaptive-trapezoid step (for both, the CS, T, E inputs are restored
 * between launches, outside the timed region). The timed kernel's other outputs stay in the
 * context, so time a context only after its state is no longer needed. */
#define EU3D_KERNEL_FLUX_Z 0
#define EU3D_KERNEL_FLUX_XY 1
#define EU3D_KERNEL_IMEX_FUSED 2
#define EU3D_KERNEL_TRAP_FIRST 3    /* the adaptive-trapezoid step's prediction pass (Update_after_Adv,
                                       Update_after_CS, TrapezoidPrediction, first sub-step) */
int eu3d_time_kernel(eu3d_ctx *ctx, int kernel, int reps, int diff, int limiter, double *avg_ms);

#ifdef __cplusplus
}
#endif
#endif
