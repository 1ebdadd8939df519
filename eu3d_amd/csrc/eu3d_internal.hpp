// Context and launch declarations shared by the C-ABI layer and the kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/eu3d.h"
#include "eu3d_device.hpp"

namespace eu3d {

// Grid of one context. NX, NY, NZ are the ghost-inclusive extents (ni + 2 bc, ...). Device fields
// are SoA with x fastest: component c of cell (i, j, k) is field[c * n3 + IDX(g, i, j, k)]. A row
// of NX cells sits in a slot of PX doubles (NX rounded up to 16) at offset OX = 16 - bc, so the
// first interior cell of every row, and with it every 64-cell wavefront of a row, starts a
// 128-byte line: a wavefront's access touches 4 lines, not 5 (a 16-in / 13-out FP64 stream ran
// at 5.4 against 3.6 TB/s). A row's last ghost (slot offset OX + NX - 1 <= PX + OX - 1) ends
// before the next row's first ghost (offset PX + OX). n3, the per-component stride, is a multiple
// of 32 doubles so that every component starts on 256 bytes; ncell = NX NY NZ is the number of
// cells with ghosts (the reference Array's size).
struct G3 {
    int ni, nj, nk, bc, NX, NY, NZ, PX, OX;
    size_t n3, ncell;
};

struct Fields {
    double* f[EU3D_FIELD_COUNT];
};

// Interior kernels run 64x4 blocks over (i, j) and one grid layer per k. k_cfl and k_imex_fused
// write one CFL partial minimum per block of that grid, so the partial buffer holds exactly
// interior_blocks() doubles.
constexpr int BX = 64, BY = 4;
inline long long interior_blocks(int ni, int nj, int nk) {
    return (long long)((ni + BX - 1) / BX) * ((nj + BY - 1) / BY) * nk;
}

#define IDX(g, i, j, k) ((((size_t)(k)) * (size_t)(g).NY + (size_t)(j)) * (size_t)(g).PX + (size_t)(g).OX + (size_t)(i))
// The grid layout of a context with this interior and ghost width (PX, OX, n3, ncell above).
inline void grid_layout(G3& g) {
    g.NX = g.ni + 2 * g.bc;
    g.NY = g.nj + 2 * g.bc;
    g.NZ = g.nk + 2 * g.bc;
    g.PX = (g.NX + 15) / 16 * 16;
    g.OX = 16 - g.bc;
    g.ncell = (size_t)g.NX * g.NY * g.NZ;
    g.n3 = ((size_t)g.PX * g.NY * g.NZ + g.OX + 31) / 32 * 32;
}
#define FLD(F, id) ((F).f[id])

// Device-resident scalars of a context.
struct Scalars {
    double dt, time, dtm;            // dtm: chemistry interval of the current adaptive call
    int nonfinite;
    unsigned coop_count;             // cells queued for the one-cell-per-wavefront kernel
    unsigned long long chem_head;    // persistent-kernel work queue head
    unsigned long long chem_active;  // number of queued cells
    unsigned long long coop_head;    // work queue head of the tail kernel
    // sub-step sum ([slot][0]) and maximum ([slot][1]) of the chemistry passes, spread over
    // STAT_SLOTS slots of one 128-byte line each: one pair of atomics per wavefront on a single
    // address serialised in L2 (about 3 ms per step over 16.8 M cells)
    unsigned long long stat[32][16];
};
constexpr int STAT_SLOTS = 32;

}  // namespace eu3d

struct eu3d_ctx {
    eu3d::G3 g{};
    uint32_t quirks = 0;
    int device = 0;
    int num_cu = 256;
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;                                    // the tail kernel's stream (default priority)
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;               // context stream -> side -> context stream
    eu3d::Fields fl{};
    double *xnode = nullptr, *ynode = nullptr, *znode = nullptr;   // device, ghost-inclusive
    double dx0 = 0, dy0 = 0, dz0 = 0;                              // node(1) - node(0)
    eu3d::Scalars* sc = nullptr;                                  // device
    unsigned int* active = nullptr;                                // compaction list (interior cells)
    double* part = nullptr;                                        // per-block CFL partial minima
    bool part_ready = false;   // the last step's passes left the CFL partial minima of its state in part
    double* rk_save = nullptr;                                     // RK2 copy of CS, T, E
    // batched chemistry scratch
    unsigned int* batch_active = nullptr;
    long long batch_cap = 0;
    eu3d_chem_schedule sched{256, 4, 2, 512};
    unsigned int* coop_list = nullptr;                             // cells with the longest sub-step chains
    long long coop_cap = 0;
    int flux_variant = 0;                                          // tile shape / order of the x/y advection pass
    int zchunk = 0;                                                // z planes per column chunk of the z pass (0: by grid size)
    struct eu3d_mlp* mlp = nullptr;                                // surrogate used by EU3D_CHEM_SURROGATE
    unsigned* sur_list = nullptr;                                  // fallback list, work list, counters, keep
    long long sur_cap = 0;
    eu3d_surrogate_stats sur_stats{};
    eu3d_chem_stats stats{};
    std::string err;
};

namespace eu3d {
// Launchers (eu3d_kernels.hip). All enqueue on ctx->stream and return hipError_t.
hipError_t launch_aos_to_soa(eu3d_ctx* c, const double* aos, double* soa, int nc);
hipError_t launch_soa_to_aos(eu3d_ctx* c, const double* soa, double* aos, int nc);
hipError_t launch_update_after_cs(eu3d_ctx* c);
hipError_t launch_partial_t(eu3d_ctx* c);
hipError_t launch_field_initial(eu3d_ctx* c);
hipError_t launch_fill_ghosts(eu3d_ctx* c, const int kinds[6]);
hipError_t launch_advection(eu3d_ctx* c, int diff, int limiter);
hipError_t launch_cfl(eu3d_ctx* c, double cfl, double final_time);
hipError_t launch_cfl_finish(eu3d_ctx* c, double cfl, double final_time);
hipError_t launch_flux_fused(eu3d_ctx* c, int diff, int limiter);
hipError_t launch_flux_z(eu3d_ctx* c, int diff, int limiter);
hipError_t launch_flux_xy(eu3d_ctx* c, int diff, int limiter);
// last: the step is the last of its eu3d_step call (C, the sound speed, is stored then; within a
// call nothing reads it)
hipError_t launch_imex_fused(eu3d_ctx* c, bool last);
hipError_t launch_derive(eu3d_ctx* c);
hipError_t launch_clear_nchem_flag(eu3d_ctx* c);
hipError_t launch_update_after_adv(eu3d_ctx* c);
hipError_t launch_explicit(eu3d_ctx* c);
hipError_t launch_diagonalized(eu3d_ctx* c);
hipError_t launch_update_imex(eu3d_ctx* c);
hipError_t launch_trapezoid(eu3d_ctx* c);
// nthr > 0: cells above T_cold needing more than nthr sub-steps keep their state and are listed
// in slist / scount for the surrogate network instead of taking the FP64 sub-steps. adv: run the
// step's Update_after_Adv (CS += RHS) and Update_after_CS inside the prediction pass.
// step_src: store the chemical source and every cell's sub-step count (the last step of an
// eu3d_step call); step_cfl: the passes leave the next step's CFL partial minima in ctx->part.
hipError_t launch_trap_adaptive(eu3d_ctx* c, int nchem_max, bool predict_only, int nthr = 0, double T_cold = 0.0,
                                unsigned* slist = nullptr, unsigned* scount = nullptr, bool adv = false,
                                int nthr_hi = 0, bool step_src = true, bool step_cfl = false);
hipError_t launch_check_finite(eu3d_ctx* c);
hipError_t launch_chem_reset_step(eu3d_ctx* c);
hipError_t launch_trap_first_step(eu3d_ctx* c, int nchem_max);
hipError_t ensure_coop(eu3d_ctx* c, long long n);
hipError_t launch_set_scalars(eu3d_ctx* c, double dt, double time);
hipError_t launch_fill(eu3d_ctx* c, double* a, double v, size_t n);   // a[0..n) = v on the context stream
// first / last: the step is the first / the last of its eu3d_step call
hipError_t launch_step(eu3d_ctx* c, const eu3d_step_params* p, bool first, bool last);
// adv: the step's Update_after_Adv + Update_after_CS have not run yet (fused into the passes);
// last: the step is the last of its eu3d_step call (with adv: the chemical source is stored on
// that step only, and the steps before it leave the next step's CFL partial minima)
hipError_t launch_surrogate_grid(eu3d_ctx* c, int nchem_max, bool adv = false, bool last = true);
// The CFL bounds of the listed cells' final state (U, V, W, C in the fields; count on the device)
// folded into the block partial minima of ctx->part.
hipError_t launch_cfl_list(eu3d_ctx* c, const unsigned* list, const unsigned* count, unsigned max_count);
// Update_after_Adv + Update_after_CS of the listed cells (count on the device)
hipError_t launch_adv_list(eu3d_ctx* c, const unsigned* list, const unsigned* count, unsigned max_count);
hipError_t surrogate_grid_stats(eu3d_ctx* c);
hipError_t launch_trap_list(eu3d_ctx* c, const unsigned* list, const unsigned* count, unsigned max_count,
                            unsigned* work, int nchem_max);
hipError_t launch_batch_list(eu3d_ctx* c, long long n, double* rho, double* T, double* E, double dtm,
                             const unsigned* list, const unsigned* count, unsigned max_count, unsigned* work,
                             int* nchem_keep, int nchem_max);
hipError_t launch_chem_batch(eu3d_ctx* c, long long n, double* rho, double* T, double* E, double dtm, int nchem_max,
                             int* nchem_out, int nthr = 0, double T_cold = 0.0, unsigned* slist = nullptr,
                             unsigned* scount = nullptr, int nthr_hi = 0);
}  // namespace eu3d
