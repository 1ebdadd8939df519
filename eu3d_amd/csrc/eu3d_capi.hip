// C ABI of include/eu3d.h: context lifetime, field exchange and the entry points, each a thin
// wrapper that validates its arguments and enqueues the kernels of eu3d_kernels.hip.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "eu3d_internal.hpp"

using namespace eu3d;

namespace {

int ncomp(int field) {
    switch (field) {
    case EU3D_FIELD_CS: case EU3D_FIELD_PT: case EU3D_FIELD_RHS:
    case EU3D_FIELD_F: case EU3D_FIELD_G: case EU3D_FIELD_Q: return NC;
    case EU3D_FIELD_YI: case EU3D_FIELD_DI: case EU3D_FIELD_MC:
    case EU3D_FIELD_CMS: case EU3D_FIELD_MD: return NS;
    default: return (field >= 0 && field < EU3D_FIELD_COUNT) ? 1 : 0;
    }
}

int fail(eu3d_ctx* c, int code, const char* msg) {
    if (c) c->err = msg;
    return code;
}

int hip_fail(eu3d_ctx* c, hipError_t e, const char* where) {
    if (c) {
        c->err = std::string(where) + ": " + hipGetErrorString(e);
    }
    return EU3D_ERR_HIP;
}

#define HIPCHK(c, call, where)                          \
    do {                                                \
        hipError_t e_ = (call);                         \
        if (e_ != hipSuccess) return hip_fail(c, e_, where); \
    } while (0)

// ---- scratch of eu3d_chem_batch_noalloc: counters (one Scalars), the compaction list and the
// per-cell sub-step counts (launch_chem_batch keeps them side by side), and the tail queue.
// 12 bytes a cell for EU3D_NOALLOC_MAX_CELLS cells, 48 MiB and 4.3 KB per GPU, held from the
// library's load on that GPU until the process ends.
constexpr size_t NOALLOC_SCALARS = (sizeof(Scalars) + 255) & ~(size_t)255;
constexpr size_t NOALLOC_LISTS = (size_t)EU3D_NOALLOC_MAX_CELLS * (sizeof(unsigned) + sizeof(int));
constexpr size_t NOALLOC_BYTES = NOALLOC_SCALARS + NOALLOC_LISTS + (size_t)EU3D_NOALLOC_MAX_CELLS * sizeof(unsigned);
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

// The array's device address and the CU count, looked up once per device with that device
// current. In host code `scratch` names a host-side stand-in, never the device memory.
constexpr int NOALLOC_MAX_DEVICES = 64;
struct NoallocDevice {
    unsigned char* ptr = nullptr;
    int num_cu = 0;
};
NoallocDevice noalloc_devices[NOALLOC_MAX_DEVICES];
std::mutex noalloc_mutex;

hipError_t noalloc_device(unsigned char*& ptr, int& num_cu) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(noalloc_mutex);
    NoallocDevice& d = noalloc_devices[dev];
    if (!d.ptr) {
        void* p = nullptr;
        if ((e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch))) != hipSuccess) return e;
        int cu = 0;
        if ((e = hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return e;
        d.num_cu = cu > 0 ? cu : 256;
        d.ptr = static_cast<unsigned char*>(p);
    }
    ptr = d.ptr;
    num_cu = d.num_cu;
    return hipSuccess;
}

bool same(double a, double b) { return std::memcmp(&a, &b, sizeof(double)) == 0; }

// The kernels are compiled for one mechanism; accept only that one, bit for bit.
bool mech_matches(const eu3d_mech* m, std::string& why) {
    if (!m) { why = "mechanism is null"; return false; }
    if (m->ns != NS || m->nr != NR) { why = "species/reaction count differs from the compiled mechanism"; return false; }
    for (int s = 0; s < NS; s++) {
        if (!same(m->mw[s], MW[s])) { why = "Mw differs"; return false; }
        for (int r = 0; r < NR; r++) {
            if (!same(m->stoi_f[s * NR + r], (double)SF[s][r]) || !same(m->stoi_b[s * NR + r], (double)SB[s][r])) {
                why = "stoichiometry differs"; return false;
            }
            if (!same(m->react_tb[s * NR + r], TB[s][r])) { why = "third-body efficiencies differ"; return false; }
        }
        for (int k = 0; k < 9; k++)
            if (!same(m->coeff0[k * NS + s], C0[k][s]) || !same(m->coeff1[k * NS + s], C1[k][s])) {
                why = "NASA coefficients differ"; return false;
            }
    }
    for (int r = 0; r < NR; r++)
        if (!same(m->af[r], AF[r]) || !same(m->bf[r], BF[r]) || !same(m->eaf[r], EAF[r])) {
            why = "Arrhenius parameters differ"; return false;
        }
    return true;
}

}  // namespace

extern "C" {

int eu3d_create(eu3d_ctx** out, const eu3d_grid* grid, const eu3d_mech* mech, uint32_t quirks) {
    if (!out || !grid) return EU3D_ERR_ARG;
    *out = nullptr;
    if (grid->bc != 2 || grid->ni < 1 || grid->nj < 1 || grid->nk < 1 || !grid->xnode || !grid->ynode || !grid->znode)
        return EU3D_ERR_ARG;
    std::string why;
    if (!mech_matches(mech, why)) {
        fprintf(stderr, "eu3d_create: %s\n", why.c_str());
        return EU3D_ERR_MECH_MISMATCH;
    }
    eu3d_ctx* c = new eu3d_ctx();
    c->quirks = quirks;
    G3& g = c->g;
    g.ni = grid->ni; g.nj = grid->nj; g.nk = grid->nk; g.bc = grid->bc;
    grid_layout(g);
    if (g.n3 >= (1ull << 32)) { delete c; return EU3D_ERR_ARG; }
    if (const char* v = getenv("EU3D_FLUX_XY_TILE")) c->flux_variant = atoi(v);
    if (const char* v = getenv("EU3D_FLUX_Z_CHUNK")) c->zchunk = atoi(v);
    c->dx0 = grid->xnode[1] - grid->xnode[0];
    c->dy0 = grid->ynode[1] - grid->ynode[0];
    c->dz0 = grid->znode[1] - grid->znode[0];
    hipError_t e;
    if ((e = hipGetDevice(&c->device)) != hipSuccess) { delete c; return EU3D_ERR_HIP; }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, c->device) == hipSuccess && prop.multiProcessorCount > 0)
        c->num_cu = prop.multiProcessorCount;
    // a blocking stream: work the caller issued to the null stream (where it produced the device
    // buffers it hands over) completes before the context's kernels read them
    if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamDefault)) != hipSuccess) { delete c; return EU3D_ERR_HIP; }
    // the side stream of the chemistry tail kernel (ordered by events from the context stream)
    if (hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess) {
        eu3d_destroy(c);
        return EU3D_ERR_HIP;
    }
    for (int f = 0; f < EU3D_FIELD_COUNT; f++) {
        const size_t bytes = g.n3 * ncomp(f) * sizeof(double);
        if ((e = hipMalloc(&c->fl.f[f], bytes)) != hipSuccess) { eu3d_destroy(c); return EU3D_ERR_HIP; }
        if ((e = hipMemsetAsync(c->fl.f[f], 0, bytes, c->stream)) != hipSuccess) { eu3d_destroy(c); return EU3D_ERR_HIP; }
    }
    const size_t nx = g.NX, ny = g.NY, nz = g.NZ;
    if (hipMalloc(&c->xnode, nx * sizeof(double)) != hipSuccess || hipMalloc(&c->ynode, ny * sizeof(double)) != hipSuccess ||
        hipMalloc(&c->znode, nz * sizeof(double)) != hipSuccess || hipMalloc(&c->sc, sizeof(Scalars)) != hipSuccess ||
        hipMalloc(&c->active, g.n3 * sizeof(unsigned)) != hipSuccess ||
        hipMalloc(&c->part, (size_t)interior_blocks(g.ni, g.nj, g.nk) * sizeof(double)) != hipSuccess ||
        hipMalloc(&c->rk_save, g.n3 * NC * sizeof(double)) != hipSuccess) {
        eu3d_destroy(c);
        return EU3D_ERR_HIP;
    }
    (void)hipMemcpyAsync(c->xnode, grid->xnode, nx * sizeof(double), hipMemcpyHostToDevice, c->stream);
    (void)hipMemcpyAsync(c->ynode, grid->ynode, ny * sizeof(double), hipMemcpyHostToDevice, c->stream);
    (void)hipMemcpyAsync(c->znode, grid->znode, nz * sizeof(double), hipMemcpyHostToDevice, c->stream);
    (void)hipMemsetAsync(c->sc, 0, sizeof(Scalars), c->stream);
    (void)launch_set_scalars(c, 0.0, 0.0);
    // Nchem starts at 1 in every cell, as Reaction::ReactionConstruction fills it
    // (src/Reaction.cpp:95-96). The fused steps store a cell's count only where it differs from 1
    // (or on a call's last step), and the failed-prediction pass takes the cells whose stored
    // count is <= 0, so a count <= 0 is only ever one that a step wrote and reported.
    (void)launch_fill(c, c->fl.f[EU3D_FIELD_NCHEM], 1.0, g.n3);
    if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) { eu3d_destroy(c); return EU3D_ERR_HIP; }
    *out = c;
    return EU3D_OK;
}

int eu3d_grid_sizes(int ni, int nj, int nk, int bc, long long* cfl_partials, unsigned long long* device_bytes) {
    if (bc != 2 || ni < 1 || nj < 1 || nk < 1) return EU3D_ERR_ARG;
    G3 g{};
    g.ni = ni;
    g.nj = nj;
    g.nk = nk;
    g.bc = bc;
    grid_layout(g);
    const unsigned long long NX = g.NX, NY = g.NY, NZ = g.NZ, n3 = g.n3;
    const long long parts = interior_blocks(ni, nj, nk);
    if (cfl_partials) *cfl_partials = parts;
    if (device_bytes) {
        unsigned long long b = 0;
        for (int f = 0; f < EU3D_FIELD_COUNT; f++) b += n3 * ncomp(f) * sizeof(double);
        b += (NX + NY + NZ) * sizeof(double) + sizeof(Scalars) + n3 * sizeof(unsigned) +
             (unsigned long long)parts * sizeof(double) + n3 * NC * sizeof(double);
        *device_bytes = b;
    }
    return EU3D_OK;
}

int eu3d_destroy(eu3d_ctx* c) {
    if (!c) return EU3D_ERR_ARG;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (int f = 0; f < EU3D_FIELD_COUNT; f++)
        if (c->fl.f[f]) (void)hipFree(c->fl.f[f]);
    if (c->xnode) (void)hipFree(c->xnode);
    if (c->ynode) (void)hipFree(c->ynode);
    if (c->znode) (void)hipFree(c->znode);
    if (c->sc) (void)hipFree(c->sc);
    if (c->active) (void)hipFree(c->active);
    if (c->part) (void)hipFree(c->part);
    if (c->rk_save) (void)hipFree(c->rk_save);
    if (c->coop_list) (void)hipFree(c->coop_list);
    if (c->batch_active) (void)hipFree(c->batch_active);
    if (c->sur_list) (void)hipFree(c->sur_list);
    if (c->side) (void)hipStreamSynchronize(c->side);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    if (c->side) (void)hipStreamDestroy(c->side);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return EU3D_OK;
}

const char* eu3d_last_error(const eu3d_ctx* c) { return c ? c->err.c_str() : "null context"; }

int eu3d_sync(eu3d_ctx* c) {
    if (!c) return EU3D_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream), "eu3d_sync");
    return EU3D_OK;
}

int eu3d_set_quirks(eu3d_ctx* c, uint32_t q) {
    if (!c) return EU3D_ERR_ARG;
    c->quirks = q;
    return EU3D_OK;
}

int eu3d_field_components(const eu3d_ctx* c, int field) {
    (void)c;
    return ncomp(field);
}

int eu3d_set_field(eu3d_ctx* c, int field, const double* host) {
    if (!c || !host || ncomp(field) == 0) return fail(c, EU3D_ERR_ARG, "eu3d_set_field: bad field or pointer");
    const int nc = ncomp(field);
    const size_t bytes = c->g.ncell * nc * sizeof(double);
    // stage through the RHS-sized save buffer (n3 * NC doubles, the largest field)
    HIPCHK(c, hipMemcpyAsync(c->rk_save, host, bytes, hipMemcpyHostToDevice, c->stream), "set_field copy");
    HIPCHK(c, launch_aos_to_soa(c, c->rk_save, c->fl.f[field], nc), "set_field transpose");
    HIPCHK(c, hipStreamSynchronize(c->stream), "set_field sync");
    return EU3D_OK;
}

int eu3d_get_field(eu3d_ctx* c, int field, double* host) {
    if (!c || !host || ncomp(field) == 0) return fail(c, EU3D_ERR_ARG, "eu3d_get_field: bad field or pointer");
    const int nc = ncomp(field);
    const size_t bytes = c->g.ncell * nc * sizeof(double);
    HIPCHK(c, launch_soa_to_aos(c, c->fl.f[field], c->rk_save, nc), "get_field transpose");
    HIPCHK(c, hipMemcpyAsync(host, c->rk_save, bytes, hipMemcpyDeviceToHost, c->stream), "get_field copy");
    HIPCHK(c, hipStreamSynchronize(c->stream), "get_field sync");
    return EU3D_OK;
}

int eu3d_set_time(eu3d_ctx* c, double dt, double time) {
    if (!c) return EU3D_ERR_ARG;
    HIPCHK(c, launch_set_scalars(c, dt, time), "set_time");
    return EU3D_OK;
}

int eu3d_get_time(eu3d_ctx* c, double* dt, double* time) {
    if (!c) return EU3D_ERR_ARG;
    Scalars s;
    HIPCHK(c, hipMemcpyAsync(&s, c->sc, sizeof(Scalars), hipMemcpyDeviceToHost, c->stream), "get_time");
    HIPCHK(c, hipStreamSynchronize(c->stream), "get_time sync");
    if (dt) *dt = s.dt;
    if (time) *time = s.time;
    return EU3D_OK;
}

int eu3d_cfl_dt(eu3d_ctx* c, double cfl, double final_time) {
    if (!c || !(cfl > 0)) return fail(c, EU3D_ERR_ARG, "eu3d_cfl_dt: cfl must be > 0");
    HIPCHK(c, launch_cfl(c, cfl, final_time), "cfl");
    return EU3D_OK;
}

int eu3d_fill_ghosts(eu3d_ctx* c, const int kinds[6]) {
    if (!c || !kinds) return EU3D_ERR_ARG;
    for (int f = 0; f < 6; f++)
        if (kinds[f] < 0 || kinds[f] > EU3D_BC_NONE) return fail(c, EU3D_ERR_ARG, "bad boundary kind");
    HIPCHK(c, launch_fill_ghosts(c, kinds), "fill_ghosts");
    return EU3D_OK;
}

int eu3d_advection(eu3d_ctx* c, int diff, int limiter) {
    if (!c || diff < 0 || diff > 1 || limiter < 0 || limiter > 4) return fail(c, EU3D_ERR_ARG, "bad scheme");
    HIPCHK(c, launch_advection(c, diff, limiter), "advection");
    return EU3D_OK;
}

int eu3d_advection_fused(eu3d_ctx* c, int diff, int limiter) {
    if (!c || diff < 0 || diff > 1 || limiter < 0 || limiter > 4) return fail(c, EU3D_ERR_ARG, "bad scheme");
    HIPCHK(c, launch_flux_fused(c, diff, limiter), "advection_fused");
    return EU3D_OK;
}

int eu3d_update_after_adv(eu3d_ctx* c) {
    if (!c) return EU3D_ERR_ARG;
    HIPCHK(c, launch_update_after_adv(c), "update_after_adv");
    return EU3D_OK;
}

int eu3d_update_after_cs(eu3d_ctx* c) {
    if (!c) return EU3D_ERR_ARG;
    HIPCHK(c, launch_update_after_cs(c), "update_after_cs");
    return EU3D_OK;
}

int eu3d_field_initial(eu3d_ctx* c) {
    if (!c) return EU3D_ERR_ARG;
    HIPCHK(c, launch_field_initial(c), "field_initial");
    return EU3D_OK;
}

int eu3d_get_partial_t(eu3d_ctx* c) {
    if (!c) return EU3D_ERR_ARG;
    HIPCHK(c, launch_partial_t(c), "partial_t");
    return EU3D_OK;
}

int eu3d_explicit(eu3d_ctx* c) {
    if (!c) return EU3D_ERR_ARG;
    HIPCHK(c, launch_explicit(c), "explicit");
    return EU3D_OK;
}

int eu3d_diagonalized(eu3d_ctx* c) {
    if (!c) return EU3D_ERR_ARG;
    HIPCHK(c, launch_diagonalized(c), "diagonalized");
    return EU3D_OK;
}

int eu3d_update_imex(eu3d_ctx* c) {
    if (!c) return EU3D_ERR_ARG;
    HIPCHK(c, launch_update_imex(c), "update_imex");
    return EU3D_OK;
}

int eu3d_trapezoid(eu3d_ctx* c) {
    if (!c) return EU3D_ERR_ARG;
    HIPCHK(c, launch_trapezoid(c), "trapezoid");
    return EU3D_OK;
}

static int chem_status(eu3d_ctx* c) {
    Scalars s;
    HIPCHK(c, hipMemcpyAsync(&s, c->sc, sizeof(Scalars), hipMemcpyDeviceToHost, c->stream), "chem stats");
    HIPCHK(c, hipStreamSynchronize(c->stream), "chem sync");
    c->stats.active_cells = (long long)s.chem_active + (long long)s.coop_count;
    unsigned long long sub = 0, mx = 0;
    for (int k = 0; k < STAT_SLOTS; k++) {
        sub += s.stat[k][0];
        mx = s.stat[k][1] > mx ? s.stat[k][1] : mx;
    }
    c->stats.substeps_total = (long long)sub;
    c->stats.nchem_max_seen = (int)mx;
    if (s.nonfinite & 2) return fail(c, EU3D_ERR_NCHEM, "Nchem <= 0 (src/Reaction.cpp:664-668)");
    return EU3D_OK;
}

int eu3d_trapezoid_prediction(eu3d_ctx* c, int nchem_max) {
    if (!c) return EU3D_ERR_ARG;
    HIPCHK(c, launch_trap_adaptive(c, nchem_max, true), "prediction");
    c->stats.cells = (long long)c->g.ni * c->g.nj * c->g.nk;
    return chem_status(c);
}

int eu3d_trapezoid_adaptive(eu3d_ctx* c, int nchem_max) {
    if (!c) return EU3D_ERR_ARG;
    HIPCHK(c, launch_trap_adaptive(c, nchem_max, false), "trapezoid_adaptive");
    c->stats.cells = (long long)c->g.ni * c->g.nj * c->g.nk;
    return chem_status(c);
}

int eu3d_check_finite(eu3d_ctx* c) {
    if (!c) return EU3D_ERR_ARG;
    HIPCHK(c, launch_check_finite(c), "check_finite");
    Scalars s;
    HIPCHK(c, hipMemcpyAsync(&s, c->sc, sizeof(Scalars), hipMemcpyDeviceToHost, c->stream), "check_finite copy");
    HIPCHK(c, hipStreamSynchronize(c->stream), "check_finite sync");
    if (s.nonfinite & 1) return fail(c, EU3D_ERR_NONFINITE, "D or T is not finite (src/Euler.cpp:319-331)");
    return EU3D_OK;
}

int eu3d_step(eu3d_ctx* c, const eu3d_step_params* p, int nsteps) {
    if (!c || !p || nsteps < 0) return EU3D_ERR_ARG;
    if (p->rk != 1 && p->rk != 2) return fail(c, EU3D_ERR_ARG, "rk must be 1 or 2");
    if (p->rk == 2 && p->chem == EU3D_CHEM_IMEX) return fail(c, EU3D_ERR_UNSUPPORTED, "IMEX couples to the EE stage only");
    if (p->diff < 0 || p->diff > 1 || p->limiter < 0 || p->limiter > 4) return fail(c, EU3D_ERR_ARG, "bad scheme");
    if (p->chem == EU3D_CHEM_SURROGATE && !c->mlp) return fail(c, EU3D_ERR_ARG, "no surrogate attached");
    const bool adaptive = p->chem == EU3D_CHEM_TRAP_ADAPTIVE || p->chem == EU3D_CHEM_SURROGATE;
    if (adaptive && nsteps > 0) HIPCHK(c, launch_clear_nchem_flag(c), "step flag");
    for (int n = 0; n < nsteps; n++) HIPCHK(c, launch_step(c, p, n == 0, n == nsteps - 1), "step");
    if (p->chem == EU3D_CHEM_TRAP_ADAPTIVE) c->stats.cells = (long long)c->g.ni * c->g.nj * c->g.nk;
    if (p->chem == EU3D_CHEM_SURROGATE && nsteps > 0) HIPCHK(c, surrogate_grid_stats(c), "surrogate stats");
    if (nsteps > 0) {
        if (!p->skip_derive) HIPCHK(c, launch_derive(c), "step derive");
        HIPCHK(c, launch_check_finite(c), "step finite");
    }
    if (adaptive && nsteps > 0) {
        // The reference aborts on Nchem <= 0 (src/Reaction.cpp:664-668). The device keeps stepping
        // (such cells store the advected state); the error is reported once, after the steps.
        int flag = 0;
        HIPCHK(c, hipMemcpyAsync(&flag, &c->sc->nonfinite, sizeof(int), hipMemcpyDeviceToHost, c->stream), "step flag");
        HIPCHK(c, hipStreamSynchronize(c->stream), "step sync");
        if (flag & 6) return fail(c, EU3D_ERR_NCHEM, "Nchem <= 0 in a step (src/Reaction.cpp:664-668)");
    }
    return EU3D_OK;
}

int eu3d_get_chem_stats(eu3d_ctx* c, eu3d_chem_stats* out) {
    if (!c || !out) return EU3D_ERR_ARG;
    const int rc = chem_status(c);   // counters of the last chemistry launch, read back now
    *out = c->stats;
    return rc;
}

int eu3d_chem_batch(eu3d_ctx* c, long long n, double* rho_s, double* T, double* E, double dtm, int nchem_max,
                    int* nchem_out, eu3d_chem_stats* stats) {
    if (!c || n <= 0 || !rho_s || !T || !E || !(dtm > 0)) return fail(c, EU3D_ERR_ARG, "eu3d_chem_batch: bad arguments");
    if (n >= (1ll << 32)) return fail(c, EU3D_ERR_ARG, "eu3d_chem_batch: n too large");
    HIPCHK(c, launch_chem_batch(c, n, rho_s, T, E, dtm, nchem_max, nchem_out), "chem_batch");
    c->stats.cells = n;
    int rc = chem_status(c);
    if (stats) *stats = c->stats;
    return rc;
}

int eu3d_chem_set_schedule(eu3d_ctx* c, const eu3d_chem_schedule* s) {
    if (!c || !s) return EU3D_ERR_ARG;
    if (s->block != 64 && s->block != 128 && s->block != 256) return fail(c, EU3D_ERR_ARG, "block must be 64/128/256");
    if (s->blocks_per_cu < 1 || s->blocks_per_cu > 64) return fail(c, EU3D_ERR_ARG, "blocks_per_cu in [1,64]");
    c->sched = *s;
    if (c->sched.waves != 1 && c->sched.waves != 2 && c->sched.waves != 4) c->sched.waves = 2;
    if (c->sched.tail < 0) c->sched.tail = 0;
    return EU3D_OK;
}

int eu3d_chem_get_schedule(eu3d_ctx* c, eu3d_chem_schedule* s) {
    if (!c || !s) return EU3D_ERR_ARG;
    *s = c->sched;
    return EU3D_OK;
}

double* eu3d_device_field(eu3d_ctx* c, int field) {
    if (!c || field < 0 || field >= EU3D_FIELD_COUNT) return nullptr;
    return c->fl.f[field];
}

// Mean duration of one kernel of the step, timed with HIP events on the context stream around
// each of `reps` launches on identical inputs (the state the kernel writes is restored between
// launches, outside the timed region). A profiling hook with no reference counterpart.
int eu3d_time_kernel(eu3d_ctx* c, int kernel, int reps, int diff, int limiter, double* avg_ms) {
    if (!c || reps < 1 || !avg_ms || kernel < 0 || kernel > EU3D_KERNEL_TRAP_FIRST || diff < 0 || diff > 1 ||
        limiter < 0 || limiter > 4)
        return fail(c, EU3D_ERR_ARG, "eu3d_time_kernel: bad arguments");
    const size_t n3 = c->g.n3;
    double* save = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto cleanup = [&]() {
        if (save) (void)hipFree(save);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    };
    // k_imex_fused and k_trap_first overwrite their inputs CS, T, E: restored between launches
    const bool imex = kernel == EU3D_KERNEL_IMEX_FUSED || kernel == EU3D_KERNEL_TRAP_FIRST;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess ||
        (imex && hipMalloc(&save, n3 * (NC + 2) * sizeof(double)) != hipSuccess)) {
        cleanup();
        return fail(c, EU3D_ERR_HIP, "eu3d_time_kernel: allocation");
    }
    // k_imex_fused reads CS, T, E (and the RHS, which it leaves alone) and overwrites CS, T, E
    double* cs = c->fl.f[EU3D_FIELD_CS];
    double* T = c->fl.f[EU3D_FIELD_T];
    double* E = c->fl.f[EU3D_FIELD_E];
    const size_t b = n3 * sizeof(double);
    auto copy3 = [&](bool to_save) -> hipError_t {
        hipError_t e;
        if ((e = hipMemcpyAsync(to_save ? save : cs, to_save ? cs : save, NC * b, hipMemcpyDeviceToDevice, c->stream)) != hipSuccess) return e;
        if ((e = hipMemcpyAsync(to_save ? save + NC * n3 : T, to_save ? T : save + NC * n3, b, hipMemcpyDeviceToDevice, c->stream)) != hipSuccess) return e;
        return hipMemcpyAsync(to_save ? save + (NC + 1) * n3 : E, to_save ? E : save + (NC + 1) * n3, b, hipMemcpyDeviceToDevice, c->stream);
    };
    double total = 0;
    hipError_t e = imex ? copy3(true) : hipSuccess;
    for (int r = 0; r < reps && e == hipSuccess; r++) {
        if (kernel == EU3D_KERNEL_TRAP_FIRST && (e = launch_chem_reset_step(c)) != hipSuccess) break;
        if ((e = hipEventRecord(e0, c->stream)) != hipSuccess) break;
        switch (kernel) {
        case EU3D_KERNEL_FLUX_Z: e = launch_flux_z(c, diff, limiter); break;
        case EU3D_KERNEL_FLUX_XY: e = launch_flux_xy(c, diff, limiter); break;
        case EU3D_KERNEL_TRAP_FIRST: e = launch_trap_first_step(c, 10000); break;
        default: e = launch_imex_fused(c, false); break;   // as on a step that is not the last of its call
        }
        if (e != hipSuccess || (e = hipEventRecord(e1, c->stream)) != hipSuccess || (e = hipEventSynchronize(e1)) != hipSuccess)
            break;
        float ms = 0;
        if ((e = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess) break;
        total += ms;
        if (imex) e = copy3(false);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    cleanup();
    if (e != hipSuccess) return hip_fail(c, e, "eu3d_time_kernel");
    *avg_ms = total / reps;
    return EU3D_OK;
}

int eu3d_chem_batch_noalloc(long long n, double* rho_s, double* T, double* E, double dtm, int nchem_max,
                            uint32_t quirks, const eu3d_chem_schedule* sched, void* stream, int* nchem_out,
                            eu3d_chem_stats* stats) {
    if (n <= 0 || n > EU3D_NOALLOC_MAX_CELLS || !rho_s || !T || !E || !(dtm > 0)) return EU3D_ERR_ARG;
    if (sched && ((sched->block != 64 && sched->block != 128 && sched->block != 256) || sched->blocks_per_cu < 1 ||
                  sched->blocks_per_cu > 64))
        return EU3D_ERR_ARG;
    unsigned char* base = nullptr;
    int num_cu = 0;
    if (noalloc_device(base, num_cu) != hipSuccess) return EU3D_ERR_HIP;
    // A context on the stack that owns nothing: its scalars and lists point into this GPU's
    // scratch, with capacities that make launch_chem_batch's growth checks never allocate.
    eu3d_ctx c;
    c.quirks = quirks;
    c.num_cu = num_cu;
    c.stream = (hipStream_t)stream;
    c.sc = reinterpret_cast<Scalars*>(base);
    c.batch_active = reinterpret_cast<unsigned*>(base + NOALLOC_SCALARS);   // list, then keep
    c.batch_cap = EU3D_NOALLOC_MAX_CELLS;
    c.coop_list = reinterpret_cast<unsigned*>(base + NOALLOC_SCALARS + NOALLOC_LISTS);
    c.coop_cap = EU3D_NOALLOC_MAX_CELLS;
    if (sched) {
        c.sched = *sched;
        if (c.sched.waves != 1 && c.sched.waves != 2 && c.sched.waves != 4) c.sched.waves = 2;
        if (c.sched.tail < 0) c.sched.tail = 0;
    }
    // the previous call's counters and flags stay in the array: k_chem_reset clears the queue
    // heads and statistics, this clears the Nchem report
    if (launch_clear_nchem_flag(&c) != hipSuccess) return EU3D_ERR_HIP;
    if (launch_chem_batch(&c, n, rho_s, T, E, dtm, nchem_max, nchem_out) != hipSuccess) return EU3D_ERR_HIP;
    if (!stats) return EU3D_OK;
    c.stats = eu3d_chem_stats{};
    const int rc = chem_status(&c);
    c.stats.cells = n;
    *stats = c.stats;
    return rc;
}

size_t eu3d_cells_with_ghosts(const eu3d_ctx* c) { return c ? c->g.ncell : 0; }

void* eu3d_stream(eu3d_ctx* c) { return c ? (void*)c->stream : nullptr; }

}  // extern "C"
