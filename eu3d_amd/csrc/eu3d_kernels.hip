// HIP kernels for the EU3D hot path on gfx950 (FP64, SoA, x fastest) and their launchers.
//
// Each kernel is the body of one reference method over the interior cells, on registers:
// the per-cell math lives in eu3d_device.hpp. Launch shapes: interior kernels use 64x4 blocks
// over (i, j) and one grid layer per k, so a wavefront covers 64 consecutive x cells and every
// field access is a coalesced 512-byte row segment.
#include "eu3d_internal.hpp"

using namespace eu3d;


namespace {

inline dim3 interior_grid(const G3& g) { return dim3((g.ni + BX - 1) / BX, (g.nj + BY - 1) / BY, g.nk); }

__device__ __forceinline__ bool interior_ijk(const G3& g, int& i, int& j, int& k) {
    i = g.bc + (int)(blockIdx.x * BX + threadIdx.x);
    j = g.bc + (int)(blockIdx.y * BY + threadIdx.y);
    k = g.bc + (int)blockIdx.z;
    return i < g.ni + g.bc && j < g.nj + g.bc;
}

__device__ __forceinline__ void load_cs(const Fields& F, const G3& g, size_t c, double (&cs)[NC]) {
    const double* p = FLD(F, EU3D_FIELD_CS);
#pragma unroll
    for (int s = 0; s < NC; s++) cs[s] = p[s * g.n3 + c];
}

// Write everything Update_after_CS leaves behind for one cell.
// The primitives that do not depend on T (D, U, V, W, E, Y) and the densities derived from CS.
__device__ __forceinline__ void store_prims_pre(const Fields& F, const G3& g, size_t c, const double (&cs)[NC],
                                                const Prim& o) {
    FLD(F, EU3D_FIELD_D)[c] = o.D;
    FLD(F, EU3D_FIELD_U)[c] = o.U;
    FLD(F, EU3D_FIELD_V)[c] = o.V;
    FLD(F, EU3D_FIELD_W)[c] = o.W;
    FLD(F, EU3D_FIELD_E)[c] = o.E;
#pragma unroll
    for (int s = 0; s < NS; s++) {
        FLD(F, EU3D_FIELD_DI)[s * g.n3 + c] = cs[s];
        FLD(F, EU3D_FIELD_MC)[s * g.n3 + c] = cs[s] * RMW(s) * 1000;
        FLD(F, EU3D_FIELD_YI)[s * g.n3 + c] = o.Y[s];
    }
}

// The rest of the primitives.
__device__ __forceinline__ void store_prims_post(const Fields& F, size_t c, const Prim& o) {
    FLD(F, EU3D_FIELD_T)[c] = o.T;
    FLD(F, EU3D_FIELD_WAV)[c] = o.Wav;
    FLD(F, EU3D_FIELD_RGAS)[c] = o.Rgas;
    FLD(F, EU3D_FIELD_P)[c] = o.P;
    FLD(F, EU3D_FIELD_H)[c] = o.H;
    FLD(F, EU3D_FIELD_CP)[c] = o.Cp;
    FLD(F, EU3D_FIELD_GAMMA)[c] = o.Gamma;
    FLD(F, EU3D_FIELD_C)[c] = o.C;
    FLD(F, EU3D_FIELD_MA)[c] = sqrt(o.U * o.U + o.V * o.V + o.W * o.W) / o.C;
}

__device__ __forceinline__ void store_prims(const Fields& F, const G3& g, size_t c, const double (&cs)[NC],
                                            const Prim& o) {
    store_prims_pre(F, g, c, cs, o);
    store_prims_post(F, c, o);
}

// ---------------------------------------------------------------- layout conversion
// Reference Array layout (((i NY) + j) NZ + k) nc + c, ghosts included, to and from the device's
// SoA rows (IDX).
__global__ void k_aos_to_soa(G3 g, const double* __restrict__ aos, double* __restrict__ soa, int nc) {
    const size_t total = g.ncell * (size_t)nc;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const size_t c = t / g.ncell, cell = t - c * g.ncell;
        const size_t i = cell % g.NX, jk = cell / g.NX, j = jk % g.NY, k = jk / g.NY;
        soa[c * g.n3 + IDX(g, i, j, k)] = aos[((i * g.NY + j) * g.NZ + k) * nc + c];
    }
}

__global__ void k_soa_to_aos(G3 g, const double* __restrict__ soa, double* __restrict__ aos, int nc) {
    const size_t total = g.ncell * (size_t)nc;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const size_t c = t % nc, cell = t / nc;
        const size_t k = cell % g.NZ, ij = cell / g.NZ, j = ij % g.NY, i = ij / g.NY;
        aos[t] = soa[c * g.n3 + IDX(g, i, j, k)];
    }
}

// ---------------------------------------------------------------- primitive recovery
// Flowfield::Update_after_CS, src/Flowfield.cpp:1246-1284
__global__ void __launch_bounds__(256) k_update_after_cs(G3 g, Fields F, uint32_t q) {
    int i, j, k;
    if (!interior_ijk(g, i, j, k)) return;
    const size_t c = IDX(g, i, j, k);
    double cs[NC];
    load_cs(F, g, c, cs);
    const Prim o = prim_from_cs(cs, FLD(F, EU3D_FIELD_T)[c], FLD(F, EU3D_FIELD_E)[c], q);
    store_prims(F, g, c, cs, o);
}

// Flowfield::GetPartial_T, src/Flowfield.cpp:1578-1595
__global__ void __launch_bounds__(256) k_partial_t(G3 g, Fields F, uint32_t q) {
    int i, j, k;
    if (!interior_ijk(g, i, j, k)) return;
    const size_t c = IDX(g, i, j, k);
    Prim o;
    o.D = FLD(F, EU3D_FIELD_D)[c];
    o.U = FLD(F, EU3D_FIELD_U)[c];
    o.V = FLD(F, EU3D_FIELD_V)[c];
    o.W = FLD(F, EU3D_FIELD_W)[c];
    o.T = FLD(F, EU3D_FIELD_T)[c];
    o.Cp = FLD(F, EU3D_FIELD_CP)[c];
    o.Rgas = FLD(F, EU3D_FIELD_RGAS)[c];
    double pt[NS];
    partial_T(o, q, pt);
    double* PT = FLD(F, EU3D_FIELD_PT);
#pragma unroll
    for (int s = 0; s < NS; s++) PT[s * g.n3 + c] = pt[s];
    const double CV = o.Cp - o.Rgas;
    PT[(NS + 1) * g.n3 + c] = -o.U / (CV * o.D);
    PT[(NS + 2) * g.n3 + c] = 1.0 / (CV * o.D);
}

// Flowfield::FieldInitial second loop (src/Flowfield.cpp:193-226) over every cell including
// ghosts: from T, P, U, V, W, Yi to Wav, Rgas, D, Di, Mc, Cp, H, E, Gamma, C and CS.
__global__ void k_field_initial(G3 g, Fields F, uint32_t q) {
    const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (t >= g.ncell) return;
    const size_t row = t / g.NX;
    const size_t c = row * g.PX + g.OX + (t - row * g.NX);
    double Y[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) Y[s] = FLD(F, EU3D_FIELD_YI)[s * g.n3 + c];
    const double T = FLD(F, EU3D_FIELD_T)[c], P = FLD(F, EU3D_FIELD_P)[c];
    const double U = FLD(F, EU3D_FIELD_U)[c], V = FLD(F, EU3D_FIELD_V)[c], W = FLD(F, EU3D_FIELD_W)[c];
    const TPow p = tpow(T);
    const bool lo = T < 1000;
    double sy = 0, cp = 0, h = 0, e = 0;
#pragma unroll
    for (int s = 0; s < NS; s++) {
        sy += Y[s] * RMW(s);
        const double hi = RI(s) * h_R(s, lo, q & Q_C4, p);
        cp += RI(s) * cp_R(s, lo, p) * Y[s];
        h += Y[s] * hi;
        e += Y[s] * (hi - RI(s) * T);
    }
    const double Wav = 1.0 / sy, Rg = R_UNIV * 1000 / Wav, D = P / (Rg * T);
    const double Cp = cp * 1000, E = e * 1000;
    const double Gm = Cp / (Cp - R_UNIV * sy * 1000);
    FLD(F, EU3D_FIELD_WAV)[c] = Wav;
    FLD(F, EU3D_FIELD_RGAS)[c] = Rg;
    FLD(F, EU3D_FIELD_D)[c] = D;
    FLD(F, EU3D_FIELD_CP)[c] = Cp;
    FLD(F, EU3D_FIELD_H)[c] = h * 1000;
    FLD(F, EU3D_FIELD_E)[c] = E;
    FLD(F, EU3D_FIELD_GAMMA)[c] = Gm;
    FLD(F, EU3D_FIELD_C)[c] = sqrt(Gm * P / D);
    double* CS = FLD(F, EU3D_FIELD_CS);
#pragma unroll
    for (int s = 0; s < NS; s++) {
        const double di = Y[s] * D;
        FLD(F, EU3D_FIELD_DI)[s * g.n3 + c] = di;
        FLD(F, EU3D_FIELD_MC)[s * g.n3 + c] = di * RMW(s) * 1000;
        CS[s * g.n3 + c] = D * Y[s];
    }
    CS[NS * g.n3 + c] = D * U;
    CS[(NS + 1) * g.n3 + c] = D * V;
    CS[(NS + 2) * g.n3 + c] = D * W;
    CS[(NS + 3) * g.n3 + c] = D * E + 0.5 * (U * U + V * V + W * W) * D;
}

// ---------------------------------------------------------------- ghost fill
__device__ __forceinline__ int ghost_src(int kind, bool lo, int gi, int n, int bc) {
    if (lo) {
        if (kind == EU3D_BC_OUTFLOW) return bc;
        if (kind == EU3D_BC_PERIODIC) return n + gi;
        return 2 * bc - 1 - gi;                       // RSBI, REFLECT
    }
    if (kind == EU3D_BC_OUTFLOW) return n + bc - 1;
    if (kind == EU3D_BC_PERIODIC) return bc + gi;
    return n + bc - 1 - gi;
}

__device__ __forceinline__ void copy_prims(const Fields& F, const G3& g, size_t d, size_t s, int flip) {
    const int ids[9] = {EU3D_FIELD_U, EU3D_FIELD_V, EU3D_FIELD_W, EU3D_FIELD_P, EU3D_FIELD_D,
                        EU3D_FIELD_T, EU3D_FIELD_H, EU3D_FIELD_GAMMA, EU3D_FIELD_E};
#pragma unroll
    for (int t = 0; t < 9; t++) FLD(F, ids[t])[d] = FLD(F, ids[t])[s];
    if (flip == 1) FLD(F, EU3D_FIELD_U)[d] = -FLD(F, EU3D_FIELD_U)[s];
    if (flip == 2) FLD(F, EU3D_FIELD_V)[d] = -FLD(F, EU3D_FIELD_V)[s];
    if (flip == 3) FLD(F, EU3D_FIELD_W)[d] = -FLD(F, EU3D_FIELD_W)[s];
    double* Y = FLD(F, EU3D_FIELD_YI);
#pragma unroll
    for (int sp = 0; sp < NS; sp++) Y[sp * g.n3 + d] = Y[sp * g.n3 + s];
}

// x faces over interior (j, k); src/Flowfield.cpp:564-610
__global__ void k_fill_x(G3 g, Fields F, int klo, int khi) {
    const int j = g.bc + blockIdx.x * blockDim.x + threadIdx.x, k = g.bc + blockIdx.y;
    const int side = blockIdx.z / g.bc, gi = blockIdx.z % g.bc;
    if (j >= g.nj + g.bc) return;
    const int kind = side == 0 ? klo : khi;
    if (kind == EU3D_BC_NONE) return;
    const int di = side == 0 ? gi : g.ni + g.bc + gi;
    const int si = ghost_src(kind, side == 0, gi, g.ni, g.bc);
    copy_prims(F, g, IDX(g, di, j, k), IDX(g, si, j, k), kind == EU3D_BC_REFLECT ? 1 : 0);
}

// y faces over all i and interior k; src/Flowfield.cpp:616-673. RSBI on y- recomputes the
// thermodynamic state from (p, T, Y) with the NASA fits (:616-647).
__global__ void k_fill_y(G3 g, Fields F, int klo, int khi, uint32_t q) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, k = g.bc + blockIdx.y;
    const int side = blockIdx.z / g.bc, gi = blockIdx.z % g.bc;
    if (i >= g.NX) return;
    const int kind = side == 0 ? klo : khi;
    if (kind == EU3D_BC_NONE) return;
    const int dj = side == 0 ? gi : g.nj + g.bc + gi;
    const int sj = ghost_src(kind, side == 0, gi, g.nj, g.bc);
    const size_t d = IDX(g, i, dj, k), s = IDX(g, i, sj, k);
    if (kind == EU3D_BC_RSBI && side == 0) {
        const int ids[5] = {EU3D_FIELD_U, EU3D_FIELD_V, EU3D_FIELD_W, EU3D_FIELD_P, EU3D_FIELD_T};
#pragma unroll
        for (int t = 0; t < 5; t++) FLD(F, ids[t])[d] = FLD(F, ids[t])[s];
        const double T = FLD(F, EU3D_FIELD_T)[d];
        const TPow p = tpow(T);
        const bool lo = T < 1000;
        double rg = 0, cp = 0, h = 0, e = 0;
#pragma unroll
        for (int sp = 0; sp < NS; sp++) {
            const double y = FLD(F, EU3D_FIELD_YI)[sp * g.n3 + s];
            FLD(F, EU3D_FIELD_YI)[sp * g.n3 + d] = y;
            FLD(F, EU3D_FIELD_MC)[sp * g.n3 + d] = FLD(F, EU3D_FIELD_MC)[sp * g.n3 + s];
            const double hi = RI(sp) * h_R(sp, lo, q & Q_C4, p);
            rg += y * RI(sp);
            cp += RI(sp) * cp_R(sp, lo, p) * y;
            h += y * hi;
            e += y * (hi - RI(sp) * T);
        }
        const double Rg = rg * 1000;
        FLD(F, EU3D_FIELD_RGAS)[d] = Rg;
        FLD(F, EU3D_FIELD_D)[d] = FLD(F, EU3D_FIELD_P)[d] / (T * Rg);
        FLD(F, EU3D_FIELD_CP)[d] = cp * 1000;
        FLD(F, EU3D_FIELD_H)[d] = h * 1000;
        FLD(F, EU3D_FIELD_E)[d] = e * 1000;
        FLD(F, EU3D_FIELD_GAMMA)[d] = (cp * 1000) / (cp * 1000 - Rg);
    } else {
        copy_prims(F, g, d, s, kind == EU3D_BC_REFLECT ? 2 : 0);
    }
}

// z faces over all (i, j); src/Flowfield.cpp:679-769
__global__ void k_fill_z(G3 g, Fields F, int klo, int khi) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    const int side = blockIdx.z / g.bc, gi = blockIdx.z % g.bc;
    if (i >= g.NX) return;
    const int kind = side == 0 ? klo : khi;
    if (kind == EU3D_BC_NONE) return;
    const int dk = side == 0 ? gi : g.nk + g.bc + gi;
    const int sk = ghost_src(kind, side == 0, gi, g.nk, g.bc);
    copy_prims(F, g, IDX(g, i, j, dk), IDX(g, i, j, sk), kind == EU3D_BC_REFLECT ? 3 : 0);
}

// ---------------------------------------------------------------- ghost fill, one launch
// The three face passes above composed per ghost cell, for grids whose interior extents are all
// at least bc (then every source of every face is an interior cell, and no cell this launch
// writes is read by it). A ghost cell's final value is that of one interior cell, carried through
// the passes that write it in the reference's order x, y, z: x copies over the interior (j, k), y
// over all i and the interior k (RSBI on y- copies U, V, W, P, T, Y, Mc and recomputes the
// thermodynamic state), z over all (i, j). Regions: z-ghost planes (all i, j), y-ghost rows
// (interior k, all i), x-ghost columns (interior j, k).
struct GhostCell {
    double U, V, W, P, D, T, H, G, E, Y[NS];
};
__device__ __forceinline__ GhostCell gc_load(const Fields& F, const G3& g, size_t c) {
    GhostCell r;
    r.U = FLD(F, EU3D_FIELD_U)[c]; r.V = FLD(F, EU3D_FIELD_V)[c]; r.W = FLD(F, EU3D_FIELD_W)[c];
    r.P = FLD(F, EU3D_FIELD_P)[c]; r.D = FLD(F, EU3D_FIELD_D)[c]; r.T = FLD(F, EU3D_FIELD_T)[c];
    r.H = FLD(F, EU3D_FIELD_H)[c]; r.G = FLD(F, EU3D_FIELD_GAMMA)[c]; r.E = FLD(F, EU3D_FIELD_E)[c];
#pragma unroll
    for (int s = 0; s < NS; s++) r.Y[s] = FLD(F, EU3D_FIELD_YI)[s * g.n3 + c];
    return r;
}
// the copy set of the face passes (copy_prims)
__device__ __forceinline__ void gc_store(const Fields& F, const G3& g, size_t c, const GhostCell& r) {
    FLD(F, EU3D_FIELD_U)[c] = r.U; FLD(F, EU3D_FIELD_V)[c] = r.V; FLD(F, EU3D_FIELD_W)[c] = r.W;
    FLD(F, EU3D_FIELD_P)[c] = r.P; FLD(F, EU3D_FIELD_D)[c] = r.D; FLD(F, EU3D_FIELD_T)[c] = r.T;
    FLD(F, EU3D_FIELD_H)[c] = r.H; FLD(F, EU3D_FIELD_GAMMA)[c] = r.G; FLD(F, EU3D_FIELD_E)[c] = r.E;
#pragma unroll
    for (int s = 0; s < NS; s++) FLD(F, EU3D_FIELD_YI)[s * g.n3 + c] = r.Y[s];
}
// RSBI y-: R, rho, c_p, h, e, gamma from (p, T, Y), in k_fill_y's arithmetic
__device__ __forceinline__ void rsbi_state(GhostCell& r, uint32_t q, double& Rg, double& Cp) {
    const double T = r.T;
    const TPow p = tpow(T);
    const bool lo = T < 1000;
    double rg = 0, cp = 0, h = 0, e = 0;
#pragma unroll
    for (int sp = 0; sp < NS; sp++) {
        const double y = r.Y[sp];
        const double hi = RI(sp) * h_R(sp, lo, q & Q_C4, p);
        rg += y * RI(sp);
        cp += RI(sp) * cp_R(sp, lo, p) * y;
        h += y * hi;
        e += y * (hi - RI(sp) * T);
    }
    Rg = rg * 1000;
    Cp = cp * 1000;
    r.D = r.P / (T * Rg);
    r.H = h * 1000;
    r.E = e * 1000;
    r.G = Cp / (Cp - Rg);
}
// the source index along one dimension of ghost index t, or -1 where that face is not filled
__device__ __forceinline__ int gsrc(int t, int n, int bc, int klo, int khi, int& kind) {
    if (t < bc) {
        kind = klo;
        return klo == EU3D_BC_NONE ? -1 : ghost_src(klo, true, t, n, bc);
    }
    kind = khi;
    return khi == EU3D_BC_NONE ? -1 : ghost_src(khi, false, t - (n + bc), n, bc);
}
struct GhostKinds {
    int k[6];
};
__device__ __forceinline__ bool ghost_dim(int t, int n, int bc) { return t < bc || t >= n + bc; }

// (i, j, k) after the x pass, j and k interior
__device__ __forceinline__ GhostCell gc_after_x(const Fields& F, const G3& g, const GhostKinds& K, int i, int j, int k) {
    if (ghost_dim(i, g.ni, g.bc)) {
        int kind;
        const int si = gsrc(i, g.ni, g.bc, K.k[0], K.k[1], kind);
        if (si >= 0) {
            GhostCell r = gc_load(F, g, IDX(g, si, j, k));
            if (kind == EU3D_BC_REFLECT) r.U = -r.U;
            return r;
        }
    }
    return gc_load(F, g, IDX(g, i, j, k));
}
// (i, j, k) after the x and y passes, k interior
__device__ __forceinline__ GhostCell gc_after_xy(const Fields& F, const G3& g, const GhostKinds& K, uint32_t q, int i,
                                                 int j, int k) {
    if (ghost_dim(j, g.nj, g.bc)) {
        int kind;
        const int sj = gsrc(j, g.nj, g.bc, K.k[2], K.k[3], kind);
        if (sj >= 0) {
            GhostCell r = gc_after_x(F, g, K, i, sj, k);
            if (kind == EU3D_BC_RSBI && j < g.bc) {
                double Rg, Cp;
                rsbi_state(r, q, Rg, Cp);
            } else if (kind == EU3D_BC_REFLECT) {
                r.V = -r.V;
            }
            return r;
        }
        return gc_load(F, g, IDX(g, i, j, k));
    }
    return gc_after_x(F, g, K, i, j, k);
}

// The z-ghost planes and y-ghost rows are enumerated PX + 16 threads a row, thread q of a row
// holding i = q - OX at slot offset q (a row's cells span slot offsets OX .. OX + NX - 1 < PX + 16),
// so that 16 consecutive threads are one aligned 128-byte line; threads outside the row return.
__device__ __forceinline__ long long fill_row_threads(const G3& g) { return (long long)g.PX + 16; }
__global__ void __launch_bounds__(256) k_fill_all(G3 g, Fields F, GhostKinds K, uint32_t q) {
    const long long RT = fill_row_threads(g);
    const long long nR1 = RT * g.NY * (2 * g.bc), nR2 = RT * (2 * g.bc) * g.nk;
    const long long nR3 = (long long)(2 * g.bc) * g.nj * g.nk;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nR1) {   // z-ghost planes, all (i, j)
        const int i = (int)(t % RT) - g.OX, j = (int)((t / RT) % g.NY), kg = (int)(t / (RT * g.NY));
        if (i < 0 || i >= g.NX) return;
        const int k = kg < g.bc ? kg : g.nk + kg;
        int kind;
        const int sk = gsrc(k, g.nk, g.bc, K.k[4], K.k[5], kind);
        if (sk < 0) return;
        GhostCell r = gc_after_xy(F, g, K, q, i, j, sk);
        if (kind == EU3D_BC_REFLECT) r.W = -r.W;
        gc_store(F, g, IDX(g, i, j, k), r);
        return;
    }
    t -= nR1;
    if (t < nR2) {   // y-ghost rows, interior k, all i
        const int i = (int)(t % RT) - g.OX, jg = (int)((t / RT) % (2 * g.bc));
        const int k = g.bc + (int)(t / (RT * 2 * g.bc));
        if (i < 0 || i >= g.NX) return;
        const int j = jg < g.bc ? jg : g.nj + jg;
        int kind;
        const int sj = gsrc(j, g.nj, g.bc, K.k[2], K.k[3], kind);
        if (sj < 0) return;
        GhostCell r = gc_after_x(F, g, K, i, sj, k);
        const size_t d = IDX(g, i, j, k);
        if (kind == EU3D_BC_RSBI && j < g.bc) {
            double Rg, Cp;
            rsbi_state(r, q, Rg, Cp);
            gc_store(F, g, d, r);
            FLD(F, EU3D_FIELD_RGAS)[d] = Rg;
            FLD(F, EU3D_FIELD_CP)[d] = Cp;
            const size_t s = IDX(g, i, sj, k);   // Mc is copied from the mirror cell as it stands
#pragma unroll
            for (int sp = 0; sp < NS; sp++) FLD(F, EU3D_FIELD_MC)[sp * g.n3 + d] = FLD(F, EU3D_FIELD_MC)[sp * g.n3 + s];
        } else {
            if (kind == EU3D_BC_REFLECT) r.V = -r.V;
            gc_store(F, g, d, r);
        }
        return;
    }
    t -= nR2;
    if (t < nR3) {   // x-ghost columns, interior (j, k)
        const int ig = (int)(t % (2 * g.bc)), j = g.bc + (int)((t / (2 * g.bc)) % g.nj);
        const int k = g.bc + (int)(t / ((long long)2 * g.bc * g.nj));
        const int i = ig < g.bc ? ig : g.ni + ig;
        int kind;
        if (gsrc(i, g.ni, g.bc, K.k[0], K.k[1], kind) < 0) return;
        gc_store(F, g, IDX(g, i, j, k), gc_after_x(F, g, K, i, j, k));
    }
}

// ---------------------------------------------------------------- advection
// One face per thread: MUSCL on the 16 primitives from the 4-cell stencil, then the AUSM flux
// (Flowfield::AUSM, src/Flowfield.cpp:1330-1502). Faces along dir cover [bc-1, n+bc); the other
// two indices cover the interior, which is every face TimeAdv::EE reads.
template <int dir>
__global__ void __launch_bounds__(256) k_face_flux(G3 g, Fields F, int order, int limiter, uint32_t q,
                                                   double* __restrict__ Fo) {
    const int i = (dir == 0 ? g.bc - 1 : g.bc) + (int)(blockIdx.x * BX + threadIdx.x);
    const int j = (dir == 1 ? g.bc - 1 : g.bc) + (int)(blockIdx.y * BY + threadIdx.y);
    const int k = (dir == 2 ? g.bc - 1 : g.bc) + (int)blockIdx.z;
    if (i >= g.ni + g.bc || j >= g.nj + g.bc || k >= g.nk + g.bc) return;
    const size_t c = IDX(g, i, j, k);
    const size_t st = dir == 0 ? 1 : (dir == 1 ? (size_t)g.PX : (size_t)g.PX * g.NY);
    Face L, R;
    auto rec = [&](const double* f, double& l, double& r) {
        muscl(order, limiter, f[c - st], f[c], f[c + st], f[c + 2 * st], l, r);
    };
    rec(FLD(F, EU3D_FIELD_P), L.P, R.P);
    rec(FLD(F, EU3D_FIELD_D), L.D, R.D);
    rec(FLD(F, EU3D_FIELD_U), L.U, R.U);
    rec(FLD(F, EU3D_FIELD_V), L.V, R.V);
    rec(FLD(F, EU3D_FIELD_W), L.W, R.W);
    rec(FLD(F, EU3D_FIELD_H), L.H, R.H);
    rec(FLD(F, EU3D_FIELD_GAMMA), L.G, R.G);
    const double* Y = FLD(F, EU3D_FIELD_YI);
#pragma unroll
    for (int s = 0; s < NS; s++) rec(Y + s * g.n3, L.Y[s], R.Y[s]);
    double fl[NC];
    ausm_flux(dir, L, R, (q & Q_F1B) && i >= g.ni, fl);
#pragma unroll
    for (int s = 0; s < NC; s++) Fo[s * g.n3 + c] = fl[s];
}

// TimeAdv::EE, src/TimeAdv.cpp:29-43
__global__ void __launch_bounds__(256) k_ee(G3 g, Fields F, const double* __restrict__ xn, const double* __restrict__ yn,
                                            const double* __restrict__ zn, const Scalars* __restrict__ sc) {
    int i, j, k;
    if (!interior_ijk(g, i, j, k)) return;
    const size_t c = IDX(g, i, j, k);
    const double dt = sc->dt;
    const double cx = -dt / (xn[i] - xn[i - 1]);
    const double cy = dt / (yn[j] - yn[j - 1]);
    const double cz = dt / (zn[k] - zn[k - 1]);
    const size_t sy = g.PX, sz = (size_t)g.PX * g.NY;
    const double *Fx = FLD(F, EU3D_FIELD_F), *Fy = FLD(F, EU3D_FIELD_G), *Fz = FLD(F, EU3D_FIELD_Q);
    double* R = FLD(F, EU3D_FIELD_RHS);
#pragma unroll
    for (int s = 0; s < NC; s++) {
        const size_t o = s * g.n3 + c;
        R[o] = cx * (Fx[o] - Fx[o - 1]) - cy * (Fy[o] - Fy[o - sy]) - cz * (Fz[o] - Fz[o - sz]);
    }
}

// ---------------------------------------------------------------- CFL
// Flowfield::CFLcondition, src/Flowfield.cpp:242-272: each block reduces its cells' bound and
// writes one partial; k_cfl_finish reduces the partials in one block (no contended atomics).
// The per-cell time-step bound of Flowfield::CFLcondition (src/Flowfield.cpp:242-272). REF (quirk
// F6): min(dx/(|u|+c), dy/(|v|+c)), z ignored. FIXED: the bound of the unsplit three-direction
// update, 1 / sum_d (|u_d|+c)/dx_d, where the launcher passes dx_d = inf for a one-cell direction.
__device__ __forceinline__ double cfl_bound(double dx, double dy, double dz, double U, double V, double W, double C,
                                            uint32_t q) {
    double v;
    if (q & Q_F6) {
        const double a = dx / (fabs(U) + C), b = dy / (fabs(V) + C);
        v = (b < a) ? b : a;
    } else {
        v = 1.0 / ((fabs(U) + C) / dx + (fabs(V) + C) / dy + (fabs(W) + C) / dz);
    }
    return (v < 1e10) ? v : 1e10;
}

// block-wide min of v (blockDim.x * blockDim.y threads, multiple of 64); thread 0 writes it
__device__ __forceinline__ void block_min_store(double v, double* out) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(v, off);
        v = (o < v) ? o : v;
    }
    __shared__ double red[16];
    const int t = threadIdx.y * blockDim.x + threadIdx.x;
    const int nw = (blockDim.x * blockDim.y) >> 6;
    if ((t & 63) == 0) red[t >> 6] = v;
    __syncthreads();
    if (t == 0) {
        double m = red[0];
        for (int w = 1; w < nw; w++) m = (red[w] < m) ? red[w] : m;
        out[(blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] = m;
    }
}

__global__ void __launch_bounds__(256) k_cfl(G3 g, Fields F, double dx, double dy, double dz, uint32_t q,
                                             double* part) {
    int i, j, k;
    double v = 1e10;
    if (interior_ijk(g, i, j, k)) {
        const size_t c = IDX(g, i, j, k);
        v = cfl_bound(dx, dy, dz, FLD(F, EU3D_FIELD_U)[c], FLD(F, EU3D_FIELD_V)[c], FLD(F, EU3D_FIELD_W)[c],
                      FLD(F, EU3D_FIELD_C)[c], q);
    }
    block_min_store(v, part);
}

__global__ void __launch_bounds__(1024) k_cfl_finish(Scalars* sc, const double* __restrict__ part, int n, double cfl,
                                                     double final_time) {
    // eight independent loads in flight per thread (one block reads every partial)
    double v = 1e10, u[8];
#pragma unroll
    for (int m = 0; m < 8; m++) u[m] = 1e10;
    const int bd = blockDim.x;
    int t = threadIdx.x;
    for (; t + 7 * bd < n; t += 8 * bd) {
#pragma unroll
        for (int m = 0; m < 8; m++) {
            const double a = part[t + m * bd];
            u[m] = (a < u[m]) ? a : u[m];
        }
    }
    for (; t < n; t += bd) v = (part[t] < v) ? part[t] : v;
#pragma unroll
    for (int m = 0; m < 8; m++) v = (u[m] < v) ? u[m] : v;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(v, off);
        v = (o < v) ? o : v;
    }
    __shared__ double red[16];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = red[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) m = (red[w] < m) ? red[w] : m;
        double dt = cfl * m;
        if (sc->time + dt > final_time) dt = final_time - sc->time;
        sc->dt = dt;
        sc->time += dt;
    }
}

__global__ void k_set_scalars(Scalars* sc, double dt, double time) {
    sc->dt = dt;
    sc->time = time;
    sc->nonfinite = 0;
}

__global__ void k_advance_time(Scalars* sc) { sc->time += sc->dt; }

__global__ void k_clear_nchem_flag(Scalars* sc) { sc->nonfinite &= ~6; }

// ---------------------------------------------------------------- state updates
// Flowfield::Update_after_Adv, src/Flowfield.cpp:1140-1157 (quirk F1: species only)
__global__ void __launch_bounds__(256) k_add_rhs(G3 g, Fields F, int ncomp) {
    int i, j, k;
    if (!interior_ijk(g, i, j, k)) return;
    const size_t c = IDX(g, i, j, k);
    double* CS = FLD(F, EU3D_FIELD_CS);
    const double* R = FLD(F, EU3D_FIELD_RHS);
    for (int s = 0; s < ncomp; s++) CS[s * g.n3 + c] = CS[s * g.n3 + c] + R[s * g.n3 + c];
}

// Flowfield::Explicit(), src/Flowfield.cpp:1169-1186 (quirk F2: rhoE into the rho*w slot)
__global__ void __launch_bounds__(256) k_explicit(G3 g, Fields F, uint32_t q) {
    int i, j, k;
    if (!interior_ijk(g, i, j, k)) return;
    const size_t c = IDX(g, i, j, k);
    double* CS = FLD(F, EU3D_FIELD_CS);
    const double* Di = FLD(F, EU3D_FIELD_DI);
    double D = 0.0;
#pragma unroll
    for (int s = 0; s < NS; s++) {
        const double v = Di[s * g.n3 + c];
        CS[s * g.n3 + c] = v;
        D += v;
    }
    FLD(F, EU3D_FIELD_D)[c] = D;
    const double U = FLD(F, EU3D_FIELD_U)[c], V = FLD(F, EU3D_FIELD_V)[c], W = FLD(F, EU3D_FIELD_W)[c];
    const double et = D * FLD(F, EU3D_FIELD_E)[c] + 0.5 * (U * U + V * V + W * W) * D;
    CS[NS * g.n3 + c] = U * D;
    CS[(NS + 1) * g.n3 + c] = V * D;
    if (q & Q_F2) {
        CS[(NS + 2) * g.n3 + c] = et;
    } else {
        CS[(NS + 2) * g.n3 + c] = W * D;
        CS[(NS + 3) * g.n3 + c] = et;
    }
}

// ---------------------------------------------------------------- advection, two passes
// The RHS of Flowfield::Advection + TimeAdv::EE in two streaming passes with no face arrays:
//   k_flux_z (first): one thread per (x, y) column marching up a chunk of z planes with a
//     4-plane register window (each primitive is read from HBM once, the next plane prefetched);
//     it stores the z term dt/dz dQ into RHS without reading it.
//   k_flux_xy (second): one z plane per block. An X x Y thread block (32 x 8 by default) stages
//     the primitives of its tile plus the two-cell halo ((X+3) x (Y+3) cells x 16 variables,
//     48 KB at 32 x 8, three blocks per CU) in LDS with coalesced loads, computes the x face left
//     of and the y face below every cell from LDS, exchanges the x and y face fluxes together
//     through the same LDS (sized for both: 53 KB at 16 x 16, three blocks per CU) and writes
//     RHS = (-dt/dx dF - dt/dy dG) - dt/dz dQ for (X-1) x (Y-1) cells, EE's order of operations.
constexpr int NV = 16;
constexpr int ZX = 64, ZY = 4;   // k_flux_z block: 64 x 4 columns; the z chunk is a template parameter

__device__ __forceinline__ const double* prim_ptr(const Fields& F, const G3& g, int v) {
    switch (v) {
    case 0: return FLD(F, EU3D_FIELD_P);
    case 1: return FLD(F, EU3D_FIELD_D);
    case 2: return FLD(F, EU3D_FIELD_U);
    case 3: return FLD(F, EU3D_FIELD_V);
    case 4: return FLD(F, EU3D_FIELD_W);
    case 5: return FLD(F, EU3D_FIELD_H);
    case 6: return FLD(F, EU3D_FIELD_GAMMA);
    default: return FLD(F, EU3D_FIELD_YI) + (size_t)(v - 7) * g.n3;
    }
}

// XCD: workgroups are dispatched to the 8 XCDs round-robin by linear id, and each XCD has its own
// L2. The remap hands every XCD a contiguous range of tiles (x fastest, then y, then z), so the
// halo columns and rows a tile shares with its neighbours are fetched into one L2, not eight.
// MM: the MUSCL_2 + minmod configuration compiled in (order and limiter arguments ignored);
// otherwise both are taken at run time.
template <int XY_X, int XY_Y, bool XCD, bool MM>
__global__ void __launch_bounds__(XY_X * XY_Y) k_flux_xy(G3 g, Fields F, int order_a, int limiter_a, uint32_t q,
                                                 const double* __restrict__ xn, const double* __restrict__ yn,
                                                 const Scalars* __restrict__ sc) {
    constexpr int XY_TX = XY_X + 3, XY_TY = XY_Y + 3, NT = XY_X * XY_Y;
    const int order = MM ? 1 : order_a, limiter = MM ? 0 : limiter_a;
    // the staged tile, then (reused) the x and y face fluxes side by side
    constexpr int NSTAGE = NV * XY_TY * XY_TX, NEX = 2 * NC * NT;
    __shared__ double lds[NSTAGE > NEX ? NSTAGE : NEX];
    auto sp = reinterpret_cast<double (*)[XY_TY][XY_TX]>(lds);
    int bx = blockIdx.x, by = blockIdx.y, bz = blockIdx.z;
    if (XCD) {
        const unsigned nbx = gridDim.x, nby = gridDim.y, total = nbx * nby * gridDim.z;
        const unsigned L = bx + nbx * (by + nby * bz), per = total / 8;
        if (L < per * 8) {
            const unsigned T = (L % 8) * per + L / 8;
            bx = T % nbx;
            by = (T / nbx) % nby;
            bz = T / (nbx * nby);
        }
    }
    const int tx = threadIdx.x, ty = threadIdx.y, t = ty * XY_X + tx;
    const int X0 = g.bc + bx * (XY_X - 1), Y0 = g.bc + by * (XY_Y - 1);
    const int z = g.bc + bz;
    const int x = X0 + tx, y = Y0 + ty;
    const int xe = g.ni + g.bc, ye = g.nj + g.bc;
    const bool out = tx < XY_X - 1 && ty < XY_Y - 1 && x < xe && y < ye;
    double* Rh = FLD(F, EU3D_FIELD_RHS);
    const size_t cc = out ? IDX(g, x, y, z) : 0;
    double zt[NC];   // z term written by k_flux_z, fetched while the tile is staged
    if (out) {
#pragma unroll
        for (int c = 0; c < NC; c++) zt[c] = Rh[c * g.n3 + cc];
    }
    // stage the tile: column a <-> x = X0 - 2 + a, row b <-> y = Y0 - 2 + b
    for (int idx = t; idx < XY_TX * XY_TY; idx += XY_X * XY_Y) {
        const int a = idx % XY_TX, b = idx / XY_TX;
        const int gx = X0 - 2 + a, gy = Y0 - 2 + b;
        if (gx < g.NX && gy < g.NY) {
            const size_t c = IDX(g, gx, gy, z);
#pragma unroll
            for (int v = 0; v < NV; v++) sp[v][b][a] = prim_ptr(F, g, v)[c];
        }
    }
    __syncthreads();
    const bool xface = ty < XY_Y - 1 && x <= xe && y < ye;
    const bool yface = tx < XY_X - 1 && x < xe && y <= ye;
    const bool f1b = q & Q_F1B;
    double fx[NC], fy[NC];
    if (xface) {   // face between x-1 and x: columns x-2 .. x+1 = a tx .. tx+3, row ty+2
        auto ld = [&](int v, int m) { return sp[v][ty + 2][tx + m]; };
        ausm_stencil<0>(ld, order, limiter, f1b && x - 1 >= g.ni, fx);
    }
    if (yface) {   // face between y-1 and y: rows ty .. ty+3, column tx+2
        auto ld = [&](int v, int m) { return sp[v][ty + m][tx + 2]; };
        ausm_stencil<1>(ld, order, limiter, f1b && x >= g.ni, fy);
    }
    __syncthreads();
    double* sfx = lds;             // the staging area, reused for the face-flux exchange
    double* sfy = lds + NC * NT;
    if (xface) {
#pragma unroll
        for (int c = 0; c < NC; c++) sfx[c * NT + t] = fx[c];
    }
    if (yface) {
#pragma unroll
        for (int c = 0; c < NC; c++) sfy[c * NT + t] = fy[c];
    }
    __syncthreads();
    if (out) {
        const double dt = sc->dt;
        const double cx = -dt / (xn[x] - xn[x - 1]), cy = dt / (yn[y] - yn[y - 1]);
#pragma unroll
        for (int c = 0; c < NC; c++) {
            double rhs = cx * (sfx[c * NT + t + 1] - sfx[c * NT + t]);
            rhs -= cy * (sfy[c * NT + t + XY_X] - sfy[c * NT + t]);
            Rh[c * g.n3 + cc] = rhs - zt[c];
        }
    }
}

// One wave per SIMD: the 4-plane window and the prefetched plane (80 values) stay in registers
// with no spill (two waves per SIMD spilled and ran 17 % slower).
template <int ZCH, bool MM>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1))) k_flux_z(G3 g, Fields F, int order_a, int limiter_a, uint32_t q,
                                                const double* __restrict__ zn, const Scalars* __restrict__ sc) {
    const int order = MM ? 1 : order_a, limiter = MM ? 0 : limiter_a;
    const int x = g.bc + blockIdx.x * ZX + threadIdx.x, y = g.bc + blockIdx.y * ZY + threadIdx.y;
    if (x >= g.ni + g.bc || y >= g.nj + g.bc) return;
    const int z0 = g.bc + blockIdx.z * ZCH, z1 = min(z0 + ZCH, g.nk + g.bc);
    const size_t base = IDX(g, x, y, 0), sz = (size_t)g.PX * g.NY;
    const bool zy = (q & Q_F1B) && x >= g.ni;
    double w[NV][4], nxt[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) {
        const double* f = prim_ptr(F, g, v) + base;
#pragma unroll
        for (int m = 0; m < 4; m++) w[v][m] = f[(size_t)(z0 - 2 + m) * sz];
        nxt[v] = f[(size_t)(z0 + 2) * sz];
    }
    auto ld = [&](int v, int m) { return w[v][m]; };
    double fb[NC], ft[NC];
    ausm_stencil<2>(ld, order, limiter, zy, fb);
    const double dt = sc->dt;
    double* Rh = FLD(F, EU3D_FIELD_RHS);
    for (int z = z0; z < z1; z++) {
#pragma unroll
        for (int v = 0; v < NV; v++) {
            w[v][0] = w[v][1];
            w[v][1] = w[v][2];
            w[v][2] = w[v][3];
            w[v][3] = nxt[v];
        }
        if (z + 1 < z1) {
#pragma unroll
            for (int v = 0; v < NV; v++) nxt[v] = prim_ptr(F, g, v)[base + (size_t)(z + 3) * sz];
        }
        ausm_stencil<2>(ld, order, limiter, zy, ft);
        const double cz = dt / (zn[z] - zn[z - 1]);
        const size_t cc = base + (size_t)z * sz;
#pragma unroll
        for (int c = 0; c < NC; c++) {
            Rh[c * g.n3 + cc] = cz * (ft[c] - fb[c]);
            fb[c] = ft[c];
        }
    }
}

// ---------------------------------------------------------------- fused IMEX
// Reaction::Diagonalized + Flowfield::Update_IMEX + Update_after_CS + GetPartial_T
// (src/Reaction.cpp:736-811, src/Flowfield.cpp:1215-1234,1246-1284,1578-1595) for one cell in
// registers. The inputs of Diagonalized (Mc, Di, dT/drho_s) are recomputed from the state at
// the start of the step (CS, T) instead of being read back; the outputs are the new CS, T, E and
// the primitives the next advection and CFL bound read. Each block writes its CFL partial.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) k_imex_fused(G3 g, Fields F, uint32_t q, const Scalars* __restrict__ sc,
                                                    double dx, double dy, double dz, double* part, int store_c) {
    int i, j, k;
    double v = 1e10;
    if (interior_ijk(g, i, j, k)) {
        const size_t c = IDX(g, i, j, k);
        // the old E enters the Newton only under REF quirk F3
        const double T0 = FLD(F, EU3D_FIELD_T)[c], E0 = (q & Q_F3) ? FLD(F, EU3D_FIELD_E)[c] : 0.0;
        const TPow p = tpow(T0);
        KinOut ko;
        // the conserved state stays in registers through the kinetics (224 VGPRs, still two waves
        // per SIMD): a second read after the kinetics missed L2 often enough to cost 2 %
        double cs[NC];
        load_cs(F, g, c, cs);
        {
            double rho[NS], D = 0;
#pragma unroll
            for (int s = 0; s < NS; s++) {
                rho[s] = cs[s];
                D += rho[s];
            }
            Prim o;
            o.D = D;
            const double rD = 1.0 / D;
            o.U = cs[NS] * rD;
            o.V = cs[NS + 1] * rD;
            o.W = cs[NS + 2] * rD;
            o.T = T0;
            double Y[NS], Mc[NS], Dr[NS], sy = 0;
#pragma unroll
            for (int s = 0; s < NS; s++) {
                Y[s] = rho[s] * rD;
                sy += Y[s] * RMW(s);
                Mc[s] = (rho[s] * RMW(s) * 1000) * 1e-6;
                Dr[s] = rho[s] * 1e-3;
            }
            o.Rgas = R_UNIV * 1000 * sy;
            o.Cp = mix_cp_Y(Y, q & Q_C4, p) * 1000;
            double gx[NS], hr[NS], pt[NS];
            species_thermo<true>(p, q & Q_C4, gx, hr);
            partial_T_h(o, q, hr, pt);
            const double tsc = (q & Q_C6) ? 1.0 : 1e3;
#pragma unroll
            for (int s = 0; s < NS; s++) pt[s] *= tsc;
            kinetics_g<false, true>(p, gx, Mc, Dr, pt, 2, ko);
        }
        __builtin_amdgcn_sched_barrier(0);
        const double dt = sc->dt;
        const double* R = FLD(F, EU3D_FIELD_RHS);
#pragma unroll
        for (int s = 0; s < NS; s++) cs[s] = cs[s] + (R[s * g.n3 + c] + ko.w[s] * 1e3 * dt) / (1.0 - dt * ko.md[s]);
#pragma unroll
        for (int s = NS; s < NC; s++) cs[s] = cs[s] + R[s * g.n3 + c];
        const Prim n = prim_from_cs_p<true>(cs, p, E0, q);
        double* CS = FLD(F, EU3D_FIELD_CS);
#pragma unroll
        for (int s = 0; s < NC; s++) CS[s * g.n3 + c] = cs[s];
        FLD(F, EU3D_FIELD_T)[c] = n.T;
        FLD(F, EU3D_FIELD_E)[c] = n.E;
        FLD(F, EU3D_FIELD_D)[c] = n.D;
        FLD(F, EU3D_FIELD_P)[c] = n.P;
        FLD(F, EU3D_FIELD_U)[c] = n.U;
        FLD(F, EU3D_FIELD_V)[c] = n.V;
        FLD(F, EU3D_FIELD_W)[c] = n.W;
        FLD(F, EU3D_FIELD_H)[c] = n.H;
        FLD(F, EU3D_FIELD_GAMMA)[c] = n.Gamma;
        if (store_c) FLD(F, EU3D_FIELD_C)[c] = n.C;
#pragma unroll
        for (int s = 0; s < NS; s++) FLD(F, EU3D_FIELD_YI)[s * g.n3 + c] = n.Y[s];
        v = cfl_bound(dx, dy, dz, n.U, n.V, n.W, n.C, q);
    }
    block_min_store(v, part);
}

// Derived fields the fused step does not write (Di, Mc, Wav, Rgas, Cp, Ma, Partion_T), refreshed
// from CS and T once at the end of eu3d_step so every field is consistent for the caller.
__global__ void __launch_bounds__(256) k_derive(G3 g, Fields F, uint32_t q) {
    int i, j, k;
    if (!interior_ijk(g, i, j, k)) return;
    const size_t c = IDX(g, i, j, k);
    double cs[NC], Y[NS], sy = 0;
    load_cs(F, g, c, cs);
    double D = 0;
#pragma unroll
    for (int s = 0; s < NS; s++) D += cs[s];
#pragma unroll
    for (int s = 0; s < NS; s++) {
        Y[s] = cs[s] / D;
        sy += Y[s] * RMW(s);
        FLD(F, EU3D_FIELD_DI)[s * g.n3 + c] = cs[s];
        FLD(F, EU3D_FIELD_MC)[s * g.n3 + c] = cs[s] * RMW(s) * 1000;
    }
    Prim o;
    o.D = D;
    o.U = FLD(F, EU3D_FIELD_U)[c];
    o.V = FLD(F, EU3D_FIELD_V)[c];
    o.W = FLD(F, EU3D_FIELD_W)[c];
    o.T = FLD(F, EU3D_FIELD_T)[c];
    o.Wav = 1.0 / sy;
    o.Rgas = R_UNIV * 1000 / o.Wav;
    o.Cp = mix_cp_Y(Y, q & Q_C4, tpow(o.T)) * 1000;
    FLD(F, EU3D_FIELD_WAV)[c] = o.Wav;
    FLD(F, EU3D_FIELD_RGAS)[c] = o.Rgas;
    FLD(F, EU3D_FIELD_CP)[c] = o.Cp;
    FLD(F, EU3D_FIELD_MA)[c] = sqrt(o.U * o.U + o.V * o.V + o.W * o.W) / FLD(F, EU3D_FIELD_C)[c];
    double pt[NS];
    partial_T(o, q, pt);
    double* PT = FLD(F, EU3D_FIELD_PT);
#pragma unroll
    for (int s = 0; s < NS; s++) PT[s * g.n3 + c] = pt[s];
    const double CV = o.Cp - o.Rgas;
    PT[(NS + 1) * g.n3 + c] = -o.U / (CV * o.D);
    PT[(NS + 2) * g.n3 + c] = 1.0 / (CV * o.D);
}

// ---------------------------------------------------------------- chemistry
__device__ __forceinline__ void react_units(const Fields& F, const G3& g, size_t c, double (&Mc)[NS],
                                            double (&Dr)[NS]) {
#pragma unroll
    for (int s = 0; s < NS; s++) {
        Mc[s] = FLD(F, EU3D_FIELD_MC)[s * g.n3 + c] * 1e-6;
        Dr[s] = FLD(F, EU3D_FIELD_DI)[s * g.n3 + c] * 1e-3;
    }
}

// Reaction::Diagonalized, src/Reaction.cpp:736-811
__global__ void __launch_bounds__(256) k_diagonalized(G3 g, Fields F, uint32_t q) {
    int i, j, k;
    if (!interior_ijk(g, i, j, k)) return;
    const size_t c = IDX(g, i, j, k);
    double Mc[NS], Dr[NS], pt[NS];
    react_units(F, g, c, Mc, Dr);
    const double tsc = (q & Q_C6) ? 1.0 : 1e3;
#pragma unroll
    for (int s = 0; s < NS; s++) pt[s] = FLD(F, EU3D_FIELD_PT)[s * g.n3 + c] * tsc;
    KinOut ko;
    kinetics<false, true>(FLD(F, EU3D_FIELD_T)[c], Mc, Dr, pt, 2, q & Q_C4, ko);
#pragma unroll
    for (int s = 0; s < NS; s++) {
        FLD(F, EU3D_FIELD_CMS)[s * g.n3 + c] = ko.w[s] * 1e3;
        FLD(F, EU3D_FIELD_MD)[s * g.n3 + c] = ko.md[s];
    }
}

// Flowfield::Update_IMEX conserved update, src/Flowfield.cpp:1219-1228. MD(c,c) for c >= NS is
// never written by the reference (quirk I2) and is 0, so those components add RHS unchanged.
__global__ void __launch_bounds__(256) k_update_imex(G3 g, Fields F, const Scalars* __restrict__ sc) {
    int i, j, k;
    if (!interior_ijk(g, i, j, k)) return;
    const size_t c = IDX(g, i, j, k);
    const double dt = sc->dt;
    double* CS = FLD(F, EU3D_FIELD_CS);
    double* R = FLD(F, EU3D_FIELD_RHS);
#pragma unroll
    for (int s = 0; s < NS; s++) {
        const size_t o = s * g.n3 + c;
        const double r = R[o] + FLD(F, EU3D_FIELD_CMS)[o] * dt;
        R[o] = r;
        CS[o] = CS[o] + r / (1.0 - dt * FLD(F, EU3D_FIELD_MD)[o]);
    }
#pragma unroll
    for (int s = NS; s < NC; s++) CS[s * g.n3 + c] = CS[s * g.n3 + c] + R[s * g.n3 + c];
}

// Reaction::Trapezoid full field, src/Reaction.cpp:412-485 (quirk C1 for the R_TB default)
template <bool C3>
__global__ void __launch_bounds__(256) k_trapezoid(G3 g, Fields F, uint32_t q, const Scalars* __restrict__ sc) {
    int i, j, k;
    if (!interior_ijk(g, i, j, k)) return;
    const size_t c = IDX(g, i, j, k);
    double Mc[NS], Dr[NS], Dn[NS];
    react_units(F, g, c, Mc, Dr);
    KinOut ko;
    kinetics<true, false, C3>(FLD(F, EU3D_FIELD_T)[c], Mc, Dr, Dr, (q & Q_C1) ? 0 : 2, q & Q_C4, ko);
    trap_update(ko, Dr, sc->dt, Dn);
#pragma unroll
    for (int s = 0; s < NS; s++) {
        FLD(F, EU3D_FIELD_CMS)[s * g.n3 + c] = ko.w[s] * 1e3;
        FLD(F, EU3D_FIELD_DI)[s * g.n3 + c] = Dn[s];
    }
}

__global__ void __launch_bounds__(256) k_check_finite(G3 g, Fields F, Scalars* sc) {
    int i, j, k;
    if (!interior_ijk(g, i, j, k)) return;
    const size_t c = IDX(g, i, j, k);
    const double d = FLD(F, EU3D_FIELD_D)[c], t = FLD(F, EU3D_FIELD_T)[c];
    if (!isfinite(d) || !isfinite(t)) atomicOr(&sc->nonfinite, 1);
}

// SSP-RK2 second stage: U^{n+1} = 1/2 U^n + 1/2 (U* + dt L(U*)), with U* in CS and
// dt L(U*) in RHS; T and E of U* seed the Newton.
__global__ void __launch_bounds__(256) k_rk2_combine(G3 g, Fields F, const double* __restrict__ save) {
    int i, j, k;
    if (!interior_ijk(g, i, j, k)) return;
    const size_t c = IDX(g, i, j, k);
    double* CS = FLD(F, EU3D_FIELD_CS);
    const double* R = FLD(F, EU3D_FIELD_RHS);
#pragma unroll
    for (int s = 0; s < NC; s++) {
        const size_t o = s * g.n3 + c;
        CS[o] = 0.5 * save[o] + 0.5 * (CS[o] + R[o]);
    }
}

__global__ void k_copy(const double* __restrict__ a, double* __restrict__ b, size_t n) {
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
        b[t] = a[t];
}

__global__ void k_fill(double* __restrict__ a, double v, size_t n) {
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
        a[t] = v;
}

// ---------------------------------------------------------------- adaptive trapezoid
// Cell state carried through the sub-steps of Euler::Trapezoid(type 1): partial densities
// [kg/m^3], velocity, internal energy, temperature.
struct ChemCell {
    double Di[NS];
    double U, V, W, E, T;
};

// Reaction-unit inputs for one sub-step, as Update_after_CS leaves Mc and the reaction code
// rescales it (src/Flowfield.cpp:1261, src/Reaction.cpp:502-503).
__device__ __forceinline__ void chem_inputs(const ChemCell& c, double (&Mc)[NS], double (&Dr)[NS]) {
#pragma unroll
    for (int s = 0; s < NS; s++) {
        Mc[s] = (c.Di[s] * RMW(s) * 1000) * 1e-6;
        Dr[s] = c.Di[s] * 1e-3;
    }
}

// Sub-step tail: the trapezoid update from the kinetics k, then Flowfield::Explicit(cell) and
// Update_after_CS(cell) (src/Flowfield.cpp:1189-1203,1287-1319). Leaves the conserved state in cs.
// tp: powers of c.T (the sub-step's starting temperature, shared with the kinetics).
// pre: passed to prim_from_cs_p (runs before the temperature Newton).
template <class Pre = NoPre>
__device__ __forceinline__ Prim chem_apply(ChemCell& c, const KinOut& k, const double (&Dr)[NS], double dt,
                                           uint32_t q, double (&cs)[NC], const TPow& tp, const Pre& pre = Pre()) {
    double Dn[NS];
    trap_update(k, Dr, dt, Dn);
    double D = 0.0;
#pragma unroll
    for (int s = 0; s < NS; s++) {
        cs[s] = Dn[s];
        D += cs[s];
    }
    cs[NS] = c.U * D;
    cs[NS + 1] = c.V * D;
    cs[NS + 2] = c.W * D;
    cs[NS + 3] = D * c.E + 0.5 * (c.U * c.U + c.V * c.V + c.W * c.W) * D;
    const Prim o = prim_from_cs_p<false>(cs, tp, c.E, q, nullptr, pre);
#pragma unroll
    for (int s = 0; s < NS; s++) c.Di[s] = cs[s];
    c.U = o.U;
    c.V = o.V;
    c.W = o.W;
    c.E = o.E;
    c.T = o.T;
    return o;
}

// Reaction::TrapezoidPrediction sub-step count (src/Reaction.cpp:650-668) from the same rates
// the first sub-step uses. nmax caps the count.
__device__ __forceinline__ int chem_nchem(const KinOut& k, const double (&Dr)[NS], const double (&Y)[NS],
                                          double dtm, int nmax) {
    double dt = dtm;
#pragma unroll
    for (int s = 0; s < NS; s++) {
        if (Y[s] >= 1e-6 && k.w[s] != 0) {
            const double t = fabs(-Dr[s] / k.w[s]);
            if (t < dt) dt = t;
        }
    }
    double nd = ceil(dtm / dt);
    // dtm = 0 gives 0/0: a count that is NaN or below 1 is the reference's Nchem <= 0, returned
    // as 0 rather than through an undefined NaN-to-int conversion
    if (!(nd >= 1.0)) return 0;
    if (nd > (double)nmax) nd = (double)nmax;
    return (int)nd;
}

// Grid accessor for the persistent kernel.
struct GridCells {
    G3 g;
    Fields F;
    __device__ void load(unsigned id, ChemCell& c, int& n) const {
        const size_t cc = id;
#pragma unroll
        for (int s = 0; s < NS; s++) c.Di[s] = FLD(F, EU3D_FIELD_DI)[s * g.n3 + cc];
        c.U = FLD(F, EU3D_FIELD_U)[cc];
        c.V = FLD(F, EU3D_FIELD_V)[cc];
        c.W = FLD(F, EU3D_FIELD_W)[cc];
        c.E = FLD(F, EU3D_FIELD_E)[cc];
        c.T = FLD(F, EU3D_FIELD_T)[cc];
        n = (int)FLD(F, EU3D_FIELD_NCHEM)[cc];
    }
    // the chemical source, stored as soon as the last sub-step's rates are known so that they are
    // not held through its energy Newton
    __device__ void store_src(unsigned id, const KinOut& k) const {
#pragma unroll
        for (int s = 0; s < NS; s++) FLD(F, EU3D_FIELD_CMS)[s * g.n3 + (size_t)id] = k.w[s] * 1e3;
    }
    // store_pre: what does not depend on T (from the hook before the Newton); store_post: the rest
    __device__ void store_pre(unsigned id, const double (&cs)[NC], const Prim& o) const {
        const size_t cc = id;
        double* CS = FLD(F, EU3D_FIELD_CS);
#pragma unroll
        for (int s = 0; s < NC; s++) CS[s * g.n3 + cc] = cs[s];
        store_prims_pre(F, g, cc, cs, o);
    }
    __device__ void store_post(unsigned id, const Prim& o) const { store_prims_post(F, id, o); }
    __device__ void store_state(unsigned id, const double (&cs)[NC], const Prim& o) const {
        store_pre(id, cs, o);
        store_post(id, o);
    }
    __device__ void store(unsigned id, const double (&cs)[NC], const Prim& o, const KinOut& k) const {
        store_state(id, cs, o);
        store_src(id, k);
    }
};

// Grid accessor of eu3d_step's adaptive trapezoid: stores what the step reads next (the conserved
// state, the primitives of the advection and the CFL bound, Y, the chemical source) and not the
// derived fields (Di, Mc, Wav, Rgas, Cp, Ma), which eu3d_step refreshes after its last step;
// Di is read from the conserved state, which holds the same values.
struct GridCellsStep {
    G3 g;
    Fields F;
    bool src = true;           // store the chemical source (CMS): on the last step of an eu3d_step call
    // store C (the sound speed): on the last step of a call. Within a call nothing reads it: the
    // chemistry passes leave the next step's CFL partials themselves.
    bool store_c = true;
    double* cfl = nullptr;     // fold each stored cell's CFL bound into the next step's block partials
    double cdx = 0, cdy = 0, cdz = 0;
    uint32_t q = 0;
    __device__ void load(unsigned id, ChemCell& c, int& n) const {
        const size_t cc = id;
#pragma unroll
        for (int s = 0; s < NS; s++) c.Di[s] = FLD(F, EU3D_FIELD_CS)[s * g.n3 + cc];
        c.U = FLD(F, EU3D_FIELD_U)[cc];
        c.V = FLD(F, EU3D_FIELD_V)[cc];
        c.W = FLD(F, EU3D_FIELD_W)[cc];
        c.E = FLD(F, EU3D_FIELD_E)[cc];
        c.T = FLD(F, EU3D_FIELD_T)[cc];
        n = (int)FLD(F, EU3D_FIELD_NCHEM)[cc];
    }
    __device__ void store_src(unsigned id, const KinOut& k) const {
        if (!src) return;
#pragma unroll
        for (int s = 0; s < NS; s++) FLD(F, EU3D_FIELD_CMS)[s * g.n3 + (size_t)id] = k.w[s] * 1e3;
    }
    __device__ void store(unsigned id, const double (&cs)[NC], const Prim& o, const KinOut& k) const {
        store_state(id, cs, o);
        store_src(id, k);
    }
    // the cell's CFL bound into the partial of the interior block that owns it (the layout
    // block_min_store writes: 64 x 4 blocks over (i, j), one grid layer per k); the bounds are
    // positive, so their bit patterns order as unsigned integers
    __device__ void cfl_min(unsigned id, const Prim& o) const {
        const size_t row = (id - (size_t)g.OX) / g.PX;   // the inverse of IDX
        const int k = (int)(row / g.NY), j = (int)(row % g.NY), i = (int)(id - g.OX - row * g.PX);
        const int gx = (g.ni + BX - 1) / BX, gy = (g.nj + BY - 1) / BY;
        const size_t b = ((size_t)(k - g.bc) * gy + (j - g.bc) / BY) * gx + (i - g.bc) / BX;
        const double v = cfl_bound(cdx, cdy, cdz, o.U, o.V, o.W, o.C, q);
        atomicMin(reinterpret_cast<unsigned long long*>(cfl + b), (unsigned long long)__double_as_longlong(v));
    }
    __device__ void store_pre(unsigned id, const double (&cs)[NC], const Prim& o) const {
        const size_t cc = id;
        double* CS = FLD(F, EU3D_FIELD_CS);
#pragma unroll
        for (int s = 0; s < NC; s++) CS[s * g.n3 + cc] = cs[s];
        FLD(F, EU3D_FIELD_D)[cc] = o.D;
        FLD(F, EU3D_FIELD_U)[cc] = o.U;
        FLD(F, EU3D_FIELD_V)[cc] = o.V;
        FLD(F, EU3D_FIELD_W)[cc] = o.W;
        FLD(F, EU3D_FIELD_E)[cc] = o.E;
#pragma unroll
        for (int s = 0; s < NS; s++) FLD(F, EU3D_FIELD_YI)[s * g.n3 + cc] = o.Y[s];
    }
    __device__ void store_post(unsigned id, const Prim& o) const {
        const size_t cc = id;
        FLD(F, EU3D_FIELD_T)[cc] = o.T;
        FLD(F, EU3D_FIELD_P)[cc] = o.P;
        FLD(F, EU3D_FIELD_H)[cc] = o.H;
        FLD(F, EU3D_FIELD_GAMMA)[cc] = o.Gamma;
        if (store_c) FLD(F, EU3D_FIELD_C)[cc] = o.C;
        if (cfl) cfl_min(id, o);
    }
    __device__ void store_state(unsigned id, const double (&cs)[NC], const Prim& o) const {
        store_pre(id, cs, o);
        store_post(id, o);
    }
};

// Flat batch accessor: rho[s * n + id], T[id], E[id]; the cells are at rest.
struct BatchCells {
    long long n;
    double *rho, *T, *E;
    const int* nchem;
    __device__ void load(unsigned id, ChemCell& c, int& nn) const {
#pragma unroll
        for (int s = 0; s < NS; s++) c.Di[s] = rho[(size_t)s * n + id];
        c.U = c.V = c.W = 0.0;
        c.E = E[id];
        c.T = T[id];
        nn = nchem[id];
    }
    __device__ void store_src(unsigned, const KinOut&) const {}
    __device__ void store_pre(unsigned id, const double (&cs)[NC], const Prim& o) const {
#pragma unroll
        for (int s = 0; s < NS; s++) rho[(size_t)s * n + id] = cs[s];
        E[id] = o.E;
    }
    __device__ void store_post(unsigned id, const Prim& o) const { T[id] = o.T; }
    __device__ void store_state(unsigned id, const double (&cs)[NC], const Prim& o) const {
        store_pre(id, cs, o);
        store_post(id, o);
    }
    __device__ void store(unsigned id, const double (&cs)[NC], const Prim& o, const KinOut&) const {
        store_state(id, cs, o);
    }
};

__device__ __forceinline__ void chem_count(Scalars* sc, int nch, bool active, unsigned* list, unsigned id) {
    // wave-aggregated compaction of the active cells and the sub-step statistics
    const unsigned long long am = __ballot(active);
    const int lane = __lane_id();
    const int leader = __ffsll((unsigned long long)__ballot(1)) - 1;
    unsigned long long base = 0;
    int tot = nch, mx = nch;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        tot += __shfl_xor(tot, off);
        const int o = __shfl_xor(mx, off);
        mx = o > mx ? o : mx;
    }
    if (lane == leader) {
        if (am) base = atomicAdd(&sc->chem_active, (unsigned long long)__popcll(am));
        const unsigned wave = ((((blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x) * blockDim.z +
                                threadIdx.z) * blockDim.y + threadIdx.y) * blockDim.x + threadIdx.x;
        unsigned long long* st = sc->stat[(wave >> 6) % STAT_SLOTS];
        atomicAdd(&st[0], (unsigned long long)tot);
        atomicMax(&st[1], (unsigned long long)mx);
    }
    base = __shfl(base, leader);
    if (active) list[base + __popcll(am & ((1ull << lane) - 1ull))] = id;
}

// What the step's chemistry passes store beyond the state the next step reads. src: the chemical
// source and the sub-step count of every cell (the last step of an eu3d_step call; on the other
// steps only the cells that continue, or whose prediction failed, keep their count). part: the
// next step's per-block CFL partial minima, from each cell's final state (null: none).
struct StepOut {
    int src = 1;
    double* part = nullptr;
    double dx = 0, dy = 0, dz = 0;
};

// Where a cell with more than one sub-step continues: the persistent kernel's list, or, above the
// tail threshold, the list of the one-cell-per-wavefront kernel.
struct TailRoute {
    unsigned* list;
    int tail;         // 0: no routing
    unsigned few = 0; // list passes: at most this many listed cells send every multi-sub-step cell
                      // to the tail kernel (as many cells as its wavefronts hold at once)
    // the threshold for a list pass of cnt cells: a short list (the rejects of the surrogate
    // network) would leave its long chains to single lanes of a nearly idle persistent kernel,
    // several times slower per sub-step than the tail kernel's 32 lanes a cell
    __device__ __forceinline__ TailRoute for_list(unsigned cnt) const {
        TailRoute r = *this;
        if (tail > 0 && cnt <= few) r.tail = 1;
        return r;
    }
};

// Wave-aggregated append of the flagged lanes' ids to a device list (one atomic per wave).
__device__ __forceinline__ void wave_append(unsigned* list, unsigned* count, bool flag, unsigned id) {
    const unsigned long long m = __ballot(flag);
    if (!m) return;   // the same for every lane of the wave
    const int lane = __lane_id();
    const int leader = __ffsll(m) - 1;
    unsigned base = 0;
    if (lane == leader) base = atomicAdd(count, (unsigned)__popcll(m));
    base = __shfl(base, leader);
    if (flag) list[base + __popcll(m & ((1ull << lane) - 1ull))] = id;
}

// chem_count with the tail split: active cells needing more than rt.tail sub-steps go to rt.list.
__device__ __forceinline__ void chem_count_tail(Scalars* sc, int nch, bool active, unsigned* list, unsigned id,
                                                const TailRoute& rt) {
    const bool tail = active && rt.tail > 0 && nch > rt.tail;
    chem_count(sc, nch, active && !tail, list, id);
    if (rt.tail > 0) wave_append(rt.list, &sc->coop_count, tail, id);
}

// Prediction + first sub-step for every interior cell; cells needing more sub-steps are
// compacted into the work list. With nthr > 0 (surrogate chemistry), a cell at or above T_cold
// needing more than nthr sub-steps (and at most nthr_hi, when nthr_hi > 0) keeps its state and
// goes to the network list instead.
// ADV: the step's Update_after_Adv (CS += RHS, src/Flowfield.cpp:1140-1157) and Update_after_CS
// run first in registers, and the inputs of the prediction come from that state instead of the
// fields it would have written. A cell whose prediction fails (Nchem <= 0, the reference's abort)
// is left to k_adv_nochem, which stores that state; holding it here through the kinetics would
// cost registers on every cell.
template <bool C3, bool ADV>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(ADV ? 2 : 1, ADV ? 2 : 1)))
k_trap_first(G3 g, Fields F, uint32_t q, const Scalars* __restrict__ scc, Scalars* sc, unsigned* list, int nmax,
             int predict_only, int nthr, int nthr_hi, double T_cold, unsigned* slist, unsigned* scount,
             TailRoute rt, StepOut so) {
    int i, j, k;
    const bool in = interior_ijk(g, i, j, k);
    const size_t c = in ? IDX(g, i, j, k) : 0;
    int nch = 0;
    bool active = false, to_net = false;
    double cflv = 1e10;   // this cell's CFL bound when its state is final here
    if (in) {
        ChemCell cell;
        double Mc[NS];
        TPow tp;   // powers of cell.T: with ADV from the Newton that solved it
        if (ADV) {
            double cs0[NC];
            load_cs(F, g, c, cs0);
            const double* R = FLD(F, EU3D_FIELD_RHS);
            const int nadd = (q & Q_F1) ? NS : NC;
#pragma unroll
            for (int s = 0; s < NC; s++)
                if (s < nadd) cs0[s] = cs0[s] + R[s * g.n3 + c];
            // the old E enters the Newton only under REF quirk F3
            const double E_old = (q & Q_F3) ? FLD(F, EU3D_FIELD_E)[c] : 0.0;
            const Prim o0 = prim_from_cs_p(cs0, tpow(FLD(F, EU3D_FIELD_T)[c]), E_old, q, nullptr, NoPre(), &tp);
#pragma unroll
            for (int s = 0; s < NS; s++) {
                cell.Di[s] = cs0[s];
                Mc[s] = (cs0[s] * RMW(s) * 1000) * 1e-6;
            }
            cell.U = o0.U;
            cell.V = o0.V;
            cell.W = o0.W;
            cell.E = o0.E;
            cell.T = o0.T;
        } else {
#pragma unroll
            for (int s = 0; s < NS; s++) {
                cell.Di[s] = FLD(F, EU3D_FIELD_DI)[s * g.n3 + c];
                Mc[s] = FLD(F, EU3D_FIELD_MC)[s * g.n3 + c] * 1e-6;
            }
            cell.U = FLD(F, EU3D_FIELD_U)[c];
            cell.V = FLD(F, EU3D_FIELD_V)[c];
            cell.W = FLD(F, EU3D_FIELD_W)[c];
            cell.E = FLD(F, EU3D_FIELD_E)[c];
            cell.T = FLD(F, EU3D_FIELD_T)[c];
        }
        KinOut ko;
        if (!ADV) tp = tpow(cell.T);
        // the trapezoid's rates do not read Dr; it and Y are formed after the kinetics so that they
        // are not held through them (Y as Update_after_CS forms it: Di / sum Di)
        kinetics_p<true, false, C3>(tp, Mc, Mc, Mc, (q & Q_C2) ? 1 : 2, q & Q_C4, ko);
        double Dr[NS], Y[NS];
        if (ADV) {
            double D = 0;
#pragma unroll
            for (int s = 0; s < NS; s++) D += cell.Di[s];
            const double rD = 1.0 / D;
#pragma unroll
            for (int s = 0; s < NS; s++) Y[s] = cell.Di[s] * rD;
        } else {
#pragma unroll
            for (int s = 0; s < NS; s++) Y[s] = FLD(F, EU3D_FIELD_YI)[s * g.n3 + c];
        }
#pragma unroll
        for (int s = 0; s < NS; s++) Dr[s] = cell.Di[s] * 1e-3;
        const double dtm = scc->dt;
        nch = chem_nchem(ko, Dr, Y, dtm, nmax);
        if (so.src || nch != 1) FLD(F, EU3D_FIELD_NCHEM)[c] = (double)nch;
        if (nch <= 0) atomicOr(&sc->nonfinite, 2);
        to_net = nthr > 0 && nch > nthr && (nthr_hi <= 0 || nch <= nthr_hi) && cell.T >= T_cold;
        if (!predict_only && nch > 0 && !to_net) {
            using Acc = typename std::conditional<ADV, GridCellsStep, GridCells>::type;
            Acc acc{g, F};
            if constexpr (ADV) acc.src = acc.store_c = so.src != 0;
            acc.store_src((unsigned)c, ko);
            double cs[NC];
            const Prim o = chem_apply(cell, ko, Dr, dtm / nch, q, cs, tp,
                                      [&](const Prim& o1) { acc.store_pre((unsigned)c, cs, o1); });
            acc.store_post((unsigned)c, o);
            active = nch > 1;
            if (!active) cflv = cfl_bound(so.dx, so.dy, so.dz, o.U, o.V, o.W, o.C, q);
        }
    }
    chem_count_tail(sc, in && !to_net ? nch : 0, active, list, (unsigned)c, rt);
    if (nthr > 0) wave_append(slist, scount, to_net, (unsigned)c);
    // the next step's CFL partial of this block from its finished cells; the cells that continue
    // (and those whose prediction failed) fold theirs in where their final state is stored
    if (so.part) block_min_store(cflv, so.part);
}

// Update_after_Adv (CS += RHS; REF quirk F1 adds the species only) and Update_after_CS of one
// cell, stored: for the cells the fused prediction pass leaves unstored.
__device__ __forceinline__ void adv_update_cell(const G3& g, const Fields& F, uint32_t q, size_t c) {
    double cs[NC];
    load_cs(F, g, c, cs);
    const double* R = FLD(F, EU3D_FIELD_RHS);
    const int nadd = (q & Q_F1) ? NS : NC;
    for (int s = 0; s < nadd; s++) cs[s] = cs[s] + R[s * g.n3 + c];
    const Prim o = prim_from_cs(cs, FLD(F, EU3D_FIELD_T)[c], FLD(F, EU3D_FIELD_E)[c], q);
    double* CS = FLD(F, EU3D_FIELD_CS);
#pragma unroll
    for (int s = 0; s < NC; s++) CS[s * g.n3 + c] = cs[s];
    store_prims(F, g, c, cs, o);
}

// After the fused prediction pass: the cells whose prediction failed (Nchem <= 0). The pass sets
// bit 2 of sc->nonfinite for any such cell; without it there is nothing to do, so the kernel is
// launched on a small grid that scans the interior only when the bit is set.
__global__ void __launch_bounds__(256) k_adv_nochem(G3 g, Fields F, uint32_t q, const Scalars* __restrict__ sc,
                                                    StepOut so) {
    if (!(sc->nonfinite & 2)) return;
    const size_t n = (size_t)g.ni * g.nj * g.nk;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
        const size_t r = t / g.ni;
        const int i = (int)(t - r * g.ni) + g.bc, j = (int)(r % g.nj) + g.bc, k = (int)(r / g.nj) + g.bc;
        const size_t c = IDX(g, i, j, k);
        if (FLD(F, EU3D_FIELD_NCHEM)[c] > 0) continue;
        adv_update_cell(g, F, q, c);
        if (so.part) {
            GridCellsStep acc{g, F};
            acc.cfl = so.part;
            acc.cdx = so.dx;
            acc.cdy = so.dy;
            acc.cdz = so.dz;
            acc.q = q;
            Prim o;
            o.U = FLD(F, EU3D_FIELD_U)[c];
            o.V = FLD(F, EU3D_FIELD_V)[c];
            o.W = FLD(F, EU3D_FIELD_W)[c];
            o.C = FLD(F, EU3D_FIELD_C)[c];
            acc.cfl_min((unsigned)c, o);
        }
    }
}

// After a surrogate step: the CFL bounds of the cells that went to the network (accepted, or
// redone on FP64 after a rejection), whose final state neither the prediction pass nor the FP64
// queue kernels saw, into the next step's block partials. Grid-stride over the list.
__global__ void __launch_bounds__(256) k_cfl_list(G3 g, Fields F, uint32_t q, StepOut so,
                                                  const unsigned* __restrict__ list, const unsigned* __restrict__ count) {
    const unsigned n = *count;
    GridCellsStep acc{g, F};
    acc.cfl = so.part;
    acc.cdx = so.dx;
    acc.cdy = so.dy;
    acc.cdz = so.dz;
    acc.q = q;
    for (unsigned t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        const size_t c = list[t];
        Prim o;
        o.U = FLD(F, EU3D_FIELD_U)[c];
        o.V = FLD(F, EU3D_FIELD_V)[c];
        o.W = FLD(F, EU3D_FIELD_W)[c];
        o.C = FLD(F, EU3D_FIELD_C)[c];
        acc.cfl_min((unsigned)c, o);
    }
}

// After the fused prediction pass with surrogate routing: the cells listed for the network
// (count on the device), grid-stride.
__global__ void __launch_bounds__(256) k_adv_list(G3 g, Fields F, uint32_t q, const unsigned* __restrict__ list,
                                                  const unsigned* __restrict__ count) {
    const unsigned n = *count;
    for (unsigned t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x)
        adv_update_cell(g, F, q, list[t]);
}

// Persistent chemistry kernel: every lane pulls cells from the compacted list through a
// wave-aggregated atomic queue and runs that cell's remaining sub-steps, taking a new cell as
// soon as one finishes. All lanes execute the same sub-step code on different cells, so cells of
// very different sub-step counts do not serialise a wavefront.
// W is the occupancy target (waves per SIMD) the register allocation is built for: the schedule
// chooses among 1 (up to 512 registers, no spills), 2 and 4.
template <class Acc, bool C3, int W>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(W, W)))
k_chem_persist(Acc acc, const unsigned* __restrict__ list, Scalars* sc, uint32_t q, const double* __restrict__ dtm_p) {
    const unsigned long long n_active = sc->chem_active;
    // nothing queued (the usual case of a grid step): leave before every wave's atomic on the
    // queue head, which cost about 50 us over a whole-GPU grid
    if (n_active == 0) return;
    const double dtm = *dtm_p;
    const int lane = __lane_id();
    ChemCell cell;
    int rem = 0, nch = 1;
    unsigned id = 0;
    bool have = false, drained = false;
    while (true) {
        const bool need = !have && !drained;
        const unsigned long long nm = __ballot(need);
        if (nm) {
            const int leader = __ffsll(nm) - 1;
            unsigned long long base = 0;
            if (lane == leader) base = atomicAdd(&sc->chem_head, (unsigned long long)__popcll(nm));
            base = __shfl(base, leader);
            if (need) {
                const unsigned long long slot = base + __popcll(nm & ((1ull << lane) - 1ull));
                if (slot < n_active) {
                    id = list[slot];
                    acc.load(id, cell, nch);
                    rem = nch - 1;   // the first sub-step ran in the compaction pass
                    have = rem > 0;
                } else {
                    drained = true;
                }
            }
        }
        if (__ballot(have) == 0) {
            if (__ballot(!drained) == 0) break;
            continue;
        }
        if (have) {
            double Mc[NS], Dr[NS], cs[NC];
            chem_inputs(cell, Mc, Dr);
            KinOut ko;
            const TPow tp = tpow(cell.T);
            kinetics_p<true, false, C3>(tp, Mc, Dr, Dr, (q & Q_C2) ? 1 : 2, q & Q_C4, ko);
            const bool last = --rem == 0;
            if (last) acc.store_src(id, ko);
            const Prim o = chem_apply(cell, ko, Dr, dtm / nch, q, cs, tp, [&](const Prim& o1) {
                if (last) acc.store_pre(id, cs, o1);
            });
            if (last) {
                acc.store_post(id, o);
                have = false;
            }
        }
    }
}

// ---------------------------------------------------------------- tail kernel
// Cells with the longest sub-step chains run two to a wavefront, 32 lanes per cell. A sub-step's
// independent pieces go to lanes: lane s < NS evaluates species s (reaction units, G/RT and, after
// the reduction, its trapezoid update), lane r < NR reaction r, lane q < 3 NS one reaction sum,
// lane k < 17 one coefficient of the mixture fit. Values cross lanes through LDS. Each piece keeps
// the arithmetic of the per-lane path (src/Reaction.cpp:548-585, src/Flowfield.cpp:1246-1284),
// and every sum runs in the same order, so the results agree with it to rounding (FMA
// contraction differs). A sub-step then issues a few hundred wavefront instructions instead of
// several thousand, which is what bounds a chain of thousands of dependent sub-steps.

// Sparse stoichiometry of each reaction: the species with a non-zero forward or backward
// coefficient, in species order, padded with species NS (a sink row that is never summed).
constexpr int RX_TERMS = 4;
struct RxnSparse {
    int sp[NR][RX_TERMS], nf[NR][RX_TERMS], nb[NR][RX_TERMS], xr[NR];
};
constexpr int rxn_max_terms() {
    int m = 0;
    for (int r = 0; r < NR; r++) {
        int c = 0;
        for (int s = 0; s < NS; s++) c += (SF[s][r] != 0 || SB[s][r] != 0) ? 1 : 0;
        m = c > m ? c : m;
    }
    return m;
}
static_assert(rxn_max_terms() <= RX_TERMS, "a reaction involves more species than the tail kernel holds");
constexpr RxnSparse make_rxn_sparse() {
    RxnSparse t{};
    for (int r = 0; r < NR; r++) {
        int j = 0, x = 0;
        for (int s = 0; s < NS; s++) {
            x += SB[s][r] - SF[s][r];
            if (SF[s][r] != 0 || SB[s][r] != 0) {
                t.sp[r][j] = s;
                t.nf[r][j] = SF[s][r];
                t.nb[r][j] = SB[s][r];
                j++;
            }
        }
        for (; j < RX_TERMS; j++) {
            t.sp[r][j] = NS;
            t.nf[r][j] = 0;
            t.nb[r][j] = 0;
        }
        t.xr[r] = x;
    }
    return t;
}
__constant__ RxnSparse RXS = make_rxn_sparse();

// LDS of one cell (one half wavefront).
struct TailLds {
    double mc[NS + 1];               // reaction units [mol/cm^3]; slot NS pads the sparse lists
    double gx[NS + 1];               // G_s / (R T)
    double pw[NS + 1][NR];           // per-reaction terms of w; row NS is the sink
    double2 pqp[NS + 1][NR];         // per-reaction terms of Q and P, side by side
    double fin[NS];                  // w of the last sub-step, scaled
    double cs[NS];                   // updated partial densities
    double mix[2][8];                // mixture fit, low and high range: a1..a7, b1
    double rmix;                     // sum_s Y_s R_s
};

// Newton on the mixture energy (Flowfield::Get_temp, src/Flowfield.cpp:1542-1566), newton_T's
// arithmetic with the mixture fit in LDS and the powers of the starting T given; the iterates'
// powers come from the start's by tpow_near (anchored to the start, so nothing accumulates
// over a chain of sub-steps).
__device__ __forceinline__ double tail_newton(const TailLds& S, const TPow& p0, double E) {
    const double Ek = E * 1e-3;
    TPow p = p0;
    double Tn = p.T;
    for (int count = 0; count < 10; count++) {
        const double T0 = p.T;
        const int r = T0 < 1000 ? 0 : 1;
        const double* c = S.mix[r];
        const double h = -c[0] * p.Tm1 + c[1] * p.lnT + c[2] * p.T + c[3] * p.T2 / 2 + div_const<3>(c[4] * p.T3) +
                         c[5] * p.T4 / 4 + div_const<5>(c[6] * p.T5) + c[7];
        const double cp = c[0] * p.Tm2 + c[1] * p.Tm1 + c[2] + c[3] * p.T + c[4] * p.T2 + c[5] * p.T3 + c[6] * p.T4;
        const double f = (h - Ek) - S.rmix * T0;
        const double fp = cp - S.rmix;
        Tn = T0 - f / fp;
        if (fabs(Tn - T0) < 1e-6) break;
        p = tpow_near(p0, Tn);
    }
    return Tn;
}

template <class Acc, bool C3>
__global__ void __launch_bounds__(64) k_chem_tail(Acc acc, const unsigned* __restrict__ list, Scalars* sc, uint32_t q,
                                                  const double* __restrict__ dtm_p) {
    if (sc->coop_count == 0) return;   // nothing queued: no set-up, no queue atomics
    __shared__ TailLds L[2];
    const int lane = threadIdx.x, hl = lane & 31;
    TailLds& S = L[lane >> 5];
    const bool c4 = q & Q_C4;
    const int rtb = (q & Q_C2) ? 1 : 2;
    // lane constants: species sS, reaction rR, sum row, mixture coefficient
    const int sS = hl < NS ? hl : 0, rR = hl < NR ? hl : 0, kM = hl < 17 ? hl : 0;
    double clo[9], chi[9];   // a1..a7 b1 b2 of species sS, both ranges (C4: low b1, b2 throughout)
#pragma unroll
    for (int k = 0; k < 9; k++) {
        clo[k] = C0[k][sS];
        chi[k] = (k >= 7 && c4) ? C0[k][sS] : C1[k][sS];
    }
    // reaction rR: k_f = A T^b exp(-E/(Ru T)) with b in {1, -1, -2} as a product (src/Reaction.cpp:522-530)
    const double af = AF[rR], bf = BF[rR], eaf = EAF[rR];
    const int tmode = bf == 1.0 ? 1 : bf == -1.0 ? 2 : bf == -2.0 ? 3 : 0;
    const bool b_exp = tmode == 0 && bf != 0.0;
    const bool use_exp = b_exp || eaf != 0.0;
    double tbr[NS], cf[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) {
        tbr[s] = TB[s][rR];
        cf[s] = kM < 8 ? C0[kM][s] : kM < 16 ? ((kM == 15 && c4) ? C0[7][s] : C1[kM - 8][s]) : 1.0;
    }
    int sp[RX_TERMS], nfj[RX_TERMS], nbj[RX_TERMS];
#pragma unroll
    for (int j = 0; j < RX_TERMS; j++) {
        sp[j] = RXS.sp[rR][j];
        nfj[j] = RXS.nf[rR][j];
        nbj[j] = RXS.nb[rR][j];
    }
    const int xr = RXS.xr[rR];
    const double mws = MW[sS];
    const double rmw = RMW(sS);
    const int kd = kM & 7;
    for (int i = hl; i < (NS + 1) * NR; i += 32) {
        (&S.pw[0][0])[i] = 0.0;
        (&S.pqp[0][0])[i] = make_double2(0.0, 0.0);
    }
    if (hl == 0) {
        S.mc[NS] = 1.0;
        S.gx[NS] = 0.0;
    }
    __syncthreads();

    const unsigned cnt = sc->coop_count;
    const double dtm = *dtm_p;
    bool have = false, drained = false;
    int rem = 0, nch = 1;
    unsigned id = 0;
    double T = 0, E = 0, U = 0, V = 0, W = 0, di = 0, dt = 0;
    while (true) {
        const bool need = !have && !drained;   // the same on the 32 lanes of a cell
        unsigned long long slot = 0;
        if (need && hl == 0) slot = atomicAdd(&sc->coop_head, 1ull);
        slot = __shfl(slot, lane & 32);
        if (need) {
            if (slot < cnt) {
                id = list[slot];
                ChemCell c;
                acc.load(id, c, nch);
#pragma unroll
                for (int s = 0; s < NS; s++)
                    if (s == sS) di = c.Di[s];
                T = c.T;
                E = c.E;
                U = c.U;
                V = c.V;
                W = c.W;
                dt = dtm / nch;
                rem = nch - 1;   // the first sub-step ran in the compaction pass
                have = rem > 0;
            } else {
                drained = true;
            }
        }
        if (__ballot(have) == 0) {
            if (__ballot(!drained) == 0) break;
            continue;
        }
        if (!have) continue;
        const TPow p = tpow(T);
        const bool lo = T < 1000;
        // species lanes: reaction units and G/RT (Reaction::GetHi/GetSi, src/Reaction.cpp:906-948);
        // reaction lanes: k_f. Every lane evaluates both (lanes past NS or NR on species or reaction
        // 0, stored nowhere) so that the two independent chains share one basic block.
        const double Dr = di * 1e-3;
        const double mcs = (di * rmw * 1000) * 1e-6;
        double gxs;
        {
            double a[9];
#pragma unroll
            for (int k = 0; k < 9; k++) a[k] = lo ? clo[k] : chi[k];
            const double h = -a[0] * p.Tm1 + a[1] * p.lnT + a[2] * p.T + a[3] * p.T2 / 2 +
                             div_const<3>(a[4] * p.T3) + a[5] * p.T4 / 4 + div_const<5>(a[6] * p.T5) + a[7];
            const double sr = -a[0] * p.Tm2 / 2 - a[1] * p.Tm1 + a[2] * p.lnT + a[3] * p.T + a[4] * p.T2 / 2 +
                              div_const<3>(a[5] * p.T3) + a[6] * p.T4 / 4 + a[8];
            gxs = h * p.Tm1 - sr;
        }
        const double RT = R_UNIV * T;
        constexpr double P0R = P0_PA * 1e-6;
        double kf;
        {
            const double rRuT = 1.0 / (RU_CAL * T);
            const double tfac = tmode == 1 ? T : tmode == 2 ? p.Tm1 : tmode == 3 ? p.Tm2 : 1.0;
            const double ex = (b_exp ? bf * p.lnT : 0.0) - eaf * rRuT;
            const double ekf = exp_fin(ex);
            kf = af * tfac;
            if (use_exp) kf *= ekf;
        }
        if (hl < NS) {
            S.mc[hl] = mcs;
            S.gx[hl] = gxs;
        }
        __syncthreads();
        // reaction lanes: rates and the per-species terms (src/Reaction.cpp:548-585)
        if (hl < NR) {
            double mc[NS];
#pragma unroll
            for (int s = 0; s < NS; s++) mc[s] = S.mc[s];
            double tb = 0.0;
#pragma unroll
            for (int s = 0; s < NS; s++) tb += mc[s] * tbr[s];
            double xk = 0.0, rf = 1.0, rb = 1.0;
#pragma unroll
            for (int j = 0; j < RX_TERMS; j++) {
                const double g = S.gx[sp[j]], m = S.mc[sp[j]];
                const int dn = nbj[j] - nfj[j];
                xk += dn != 0 ? (double)dn * g : 0.0;
                rf *= nfj[j] == 0 ? 1.0 : nfj[j] == 1 ? m : m * m;
                rb *= nbj[j] == 0 ? 1.0 : nbj[j] == 1 ? m : m * m;
            }
            double kb = kf * exp_fin(xk);
            kb *= xr == 1 ? RT / P0R : xr == -1 ? P0R / RT : 1.0;
            if (rtb == 1) tb = 0.0;
            else if (rtb == 2 && tb == 0.0) tb = 1.0;
            const double rr = kf * rf - kb * rb;
#pragma unroll
            for (int j = 0; j < RX_TERMS; j++) {
                const int nf = nfj[j], nb = nbj[j], dn = nb - nf, s = sp[j];
                const double qf = nb ? (double)nb * kf * rf : 0.0;
                const double qb = nf ? (double)nf * kb * rb : 0.0;
                const double pb = nb ? (double)nb * kb * rb : 0.0;
                const double pf = nf ? (double)nf * kf * rf : 0.0;
                S.pw[s][hl] = dn != 0 ? (double)dn * tb * rr : 0.0;
                if (C3) S.pqp[s][hl] = make_double2(qf + qb * tb, pb + pf * tb);
                else S.pqp[s][hl] = make_double2((qf + qb) * tb, (pb + pf) * tb);
            }
        }
        __syncthreads();
        // species lanes: Q and P of the species summed over the reactions in reaction order (two
        // chains side by side), then the point-implicit trapezoid update (src/Reaction.cpp:481,587);
        // w only on the last sub-step, for the store
        if (hl < NS) {
            double Qs = 0.0, Ps = 0.0;
#pragma unroll
            for (int r = 0; r < NR; r++) {
                const double2 v = S.pqp[hl][r];
                Qs += v.x;
                Ps += v.y;
            }
            Qs *= mws;
            Ps = (mcs == 0) ? 0.0 : Ps / mcs;
            di = ((1.0 - dt / 2.0 * Ps) * Dr + dt * Qs) / (1.0 + dt / 2.0 * Ps) * 1e3;
            S.cs[hl] = di;
            if (rem == 1) {
                double w = 0.0;
#pragma unroll
                for (int r = 0; r < NR; r++) w += S.pw[hl][r];
                S.fin[hl] = w * mws;
            }
        }
        __syncthreads();
        // every lane: the conserved state (Flowfield::Explicit for the cell)
        double cs[NC];
        double D = 0.0;
#pragma unroll
        for (int s = 0; s < NS; s++) {
            cs[s] = S.cs[s];
            D += cs[s];
        }
        cs[NS] = U * D;
        cs[NS + 1] = V * D;
        cs[NS + 2] = W * D;
        cs[NS + 3] = D * E + 0.5 * (U * U + V * V + W * W) * D;
        if (--rem == 0) {
            // last sub-step: the full primitive state, stored as the persistent kernel stores it
            const Prim o = prim_from_cs(cs, T, E, q);
            if (hl == 0) {
                KinOut ko{};
#pragma unroll
                for (int s = 0; s < NS; s++) ko.w[s] = S.fin[s];
                acc.store(id, cs, o, ko);
            }
            have = false;
            continue;
        }
        // Update_after_CS with the mixture fit summed one coefficient per lane
        const double rD = 1.0 / D;
        double Y[NS];
#pragma unroll
        for (int s = 0; s < NS; s++) Y[s] = cs[s] * rD;
        const double Un = cs[NS] * rD, Vn = cs[NS + 1] * rD, Wn = cs[NS + 2] * rD;
        const double En = (cs[NS + 3] - 0.5 * D * (Un * Un + Vn * Vn + Wn * Wn)) * rD;
        if (hl < 17) {
            double a = 0.0;
#pragma unroll
            for (int s = 0; s < NS; s++) a += (Y[s] * RI(s)) * cf[s];
            if (hl < 16) S.mix[hl >> 3][kd] = a;
            else S.rmix = a;
        }
        __syncthreads();
        T = tail_newton(S, p, (q & Q_F3) ? E : En);
        E = En;
        U = Un;
        V = Vn;
        W = Wn;
    }
}

// Batch: prediction + first sub-step on flat cells at rest.
template <bool C3>
__global__ void __launch_bounds__(256) k_batch_first(BatchCells acc, int* nchem_out, int* nchem_keep, uint32_t q,
                                                     double dtm, int nmax, Scalars* sc, unsigned* list, int nthr, int nthr_hi,
                                                     double T_cold, unsigned* slist, unsigned* scount, TailRoute rt) {
    const long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const bool in = id < acc.n;
    int nch = 0;
    bool active = false, to_net = false;
    if (in) {
        ChemCell cell;
#pragma unroll
        for (int s = 0; s < NS; s++) cell.Di[s] = acc.rho[(size_t)s * acc.n + id];
        cell.U = cell.V = cell.W = 0.0;
        cell.E = acc.E[id];
        cell.T = acc.T[id];
        double Mc[NS], Dr[NS], Y[NS], D = 0;
        chem_inputs(cell, Mc, Dr);
#pragma unroll
        for (int s = 0; s < NS; s++) D += cell.Di[s];
#pragma unroll
        for (int s = 0; s < NS; s++) Y[s] = cell.Di[s] / D;
        KinOut ko;
        const TPow tp = tpow(cell.T);
        kinetics_p<true, false, C3>(tp, Mc, Dr, Dr, (q & Q_C2) ? 1 : 2, q & Q_C4, ko);
        nch = chem_nchem(ko, Dr, Y, dtm, nmax);
        nchem_keep[id] = nch;
        if (nchem_out) nchem_out[id] = nch;
        if (nch <= 0) atomicOr(&sc->nonfinite, 2);
        to_net = nthr > 0 && nch > nthr && (nthr_hi <= 0 || nch <= nthr_hi) && cell.T >= T_cold;
        if (nch > 0 && !to_net) {
            double cs[NC];
            const Prim o = chem_apply(cell, ko, Dr, dtm / nch, q, cs, tp);
            acc.store((unsigned)id, cs, o, ko);
            active = nch > 1;
        }
    }
    chem_count_tail(sc, in && !to_net ? nch : 0, active, list, (unsigned)id, rt);
    if (nthr > 0) wave_append(slist, scount, to_net, (unsigned)id);
}

// Prediction + first sub-step for the cells of a list (the surrogate's fallback cells), grid
// and batch forms; cells needing more sub-steps are compacted into out_list as above.
template <bool C3>
__global__ void __launch_bounds__(256) k_trap_first_list(G3 g, Fields F, uint32_t q, const unsigned* __restrict__ in_list,
                                                         const unsigned* __restrict__ in_count, Scalars* sc,
                                                         unsigned* out_list, int nmax, TailRoute rt) {
    // grid-stride over the listed cells (the count is known only on the device)
    const unsigned cnt = *in_count;
    const TailRoute rl = rt.for_list(cnt);
    for (unsigned b0 = blockIdx.x * blockDim.x; b0 < cnt; b0 += gridDim.x * blockDim.x) {
        const unsigned t = b0 + threadIdx.x;
        const bool in = t < cnt;
        const size_t c = in ? in_list[t] : 0;
        int nch = 0;
        bool active = false;
        if (in) {
            ChemCell cell;
    #pragma unroll
            for (int s = 0; s < NS; s++) cell.Di[s] = FLD(F, EU3D_FIELD_DI)[s * g.n3 + c];
            cell.U = FLD(F, EU3D_FIELD_U)[c];
            cell.V = FLD(F, EU3D_FIELD_V)[c];
            cell.W = FLD(F, EU3D_FIELD_W)[c];
            cell.E = FLD(F, EU3D_FIELD_E)[c];
            cell.T = FLD(F, EU3D_FIELD_T)[c];
            double Mc[NS], Dr[NS], Y[NS];
            react_units(F, g, c, Mc, Dr);
    #pragma unroll
            for (int s = 0; s < NS; s++) Y[s] = FLD(F, EU3D_FIELD_YI)[s * g.n3 + c];
            KinOut ko;
            const TPow tp = tpow(cell.T);
            kinetics_p<true, false, C3>(tp, Mc, Dr, Dr, (q & Q_C2) ? 1 : 2, q & Q_C4, ko);
            const double dtm = sc->dtm;
            nch = chem_nchem(ko, Dr, Y, dtm, nmax);
            FLD(F, EU3D_FIELD_NCHEM)[c] = (double)nch;
            if (nch <= 0) atomicOr(&sc->nonfinite, 2);
            if (nch > 0) {
                double cs[NC];
                const Prim o = chem_apply(cell, ko, Dr, dtm / nch, q, cs, tp);
                GridCells acc{g, F};
                acc.store((unsigned)c, cs, o, ko);
                active = nch > 1;
            }
        }
        chem_count_tail(sc, in ? nch : 0, active, out_list, (unsigned)c, rl);
    }
}

template <bool C3>
__global__ void __launch_bounds__(256) k_batch_first_list(BatchCells acc, int* nchem_keep, uint32_t q, double dtm, int nmax,
                                                          const unsigned* __restrict__ in_list,
                                                          const unsigned* __restrict__ in_count, Scalars* sc,
                                                          unsigned* out_list, TailRoute rt) {
    // grid-stride over the listed cells (the count is known only on the device)
    const unsigned cnt = *in_count;
    const TailRoute rl = rt.for_list(cnt);
    for (unsigned b0 = blockIdx.x * blockDim.x; b0 < cnt; b0 += gridDim.x * blockDim.x) {
        const unsigned t = b0 + threadIdx.x;
        const bool in = t < cnt;
        const long long id = in ? in_list[t] : 0;
        int nch = 0;
        bool active = false;
        if (in) {
            ChemCell cell;
    #pragma unroll
            for (int s = 0; s < NS; s++) cell.Di[s] = acc.rho[(size_t)s * acc.n + id];
            cell.U = cell.V = cell.W = 0.0;
            cell.E = acc.E[id];
            cell.T = acc.T[id];
            double Mc[NS], Dr[NS], Y[NS], D = 0;
            chem_inputs(cell, Mc, Dr);
    #pragma unroll
            for (int s = 0; s < NS; s++) D += cell.Di[s];
    #pragma unroll
            for (int s = 0; s < NS; s++) Y[s] = cell.Di[s] / D;
            KinOut ko;
            const TPow tp = tpow(cell.T);
            kinetics_p<true, false, C3>(tp, Mc, Dr, Dr, (q & Q_C2) ? 1 : 2, q & Q_C4, ko);
            nch = chem_nchem(ko, Dr, Y, dtm, nmax);
            nchem_keep[id] = nch;
            if (nch <= 0) atomicOr(&sc->nonfinite, 2);
            if (nch > 0) {
                double cs[NC];
                const Prim o = chem_apply(cell, ko, Dr, dtm / nch, q, cs, tp);
                acc.store((unsigned)id, cs, o, ko);
                active = nch > 1;
            }
        }
        chem_count_tail(sc, in ? nch : 0, active, out_list, (unsigned)id, rl);
    }
}

// Bit 2 of sc->nonfinite reports a cell with Nchem <= 0 in the current chemistry call (the passes
// of that call read it); each call starts clear and keeps an earlier call's report in bit 4,
// which eu3d_step reads back once after its steps.
__global__ void k_chem_reset(Scalars* sc, double* dtm_slot, double dtm, int use_dt) {
    if (sc->nonfinite & 2) sc->nonfinite = (sc->nonfinite & ~2) | 4;
    sc->coop_count = 0;
    sc->coop_head = 0;
    sc->chem_head = 0;
    sc->chem_active = 0;
    for (int k = 0; k < STAT_SLOTS; k++) {
        sc->stat[k][0] = 0;
        sc->stat[k][1] = 0;
    }
    *dtm_slot = use_dt ? sc->dt : dtm;
}

// Persistent-kernel launch with the context's schedule (block, blocks per CU, occupancy target).
template <class Acc>
void launch_persist(eu3d_ctx* c, const Acc& acc, const unsigned* list, const double* dtm_slot) {
    const int nb = c->sched.blocks_per_cu * c->num_cu, blk = c->sched.block;
    const bool c3 = c->quirks & Q_C3;
    hipStream_t s = c->stream;
    switch (c->sched.waves) {
    case 1:
        if (c3) k_chem_persist<Acc, true, 1><<<nb, blk, 0, s>>>(acc, list, c->sc, c->quirks, dtm_slot);
        else k_chem_persist<Acc, false, 1><<<nb, blk, 0, s>>>(acc, list, c->sc, c->quirks, dtm_slot);
        break;
    case 4:
        if (c3) k_chem_persist<Acc, true, 4><<<nb, blk, 0, s>>>(acc, list, c->sc, c->quirks, dtm_slot);
        else k_chem_persist<Acc, false, 4><<<nb, blk, 0, s>>>(acc, list, c->sc, c->quirks, dtm_slot);
        break;
    default:
        if (c3) k_chem_persist<Acc, true, 2><<<nb, blk, 0, s>>>(acc, list, c->sc, c->quirks, dtm_slot);
        else k_chem_persist<Acc, false, 2><<<nb, blk, 0, s>>>(acc, list, c->sc, c->quirks, dtm_slot);
    }
}

inline G3 G(const eu3d_ctx* c) { return c->g; }

}  // namespace

namespace eu3d {

static int grid1d(size_t n, int blk) {
    size_t b = (n + blk - 1) / blk;
    if (b > 65536 * 8) b = 65536 * 8;
    return (int)(b ? b : 1);
}

hipError_t launch_aos_to_soa(eu3d_ctx* c, const double* aos, double* soa, int nc) {
    k_aos_to_soa<<<grid1d(c->g.ncell * nc, 256), 256, 0, c->stream>>>(G(c), aos, soa, nc);
    return hipGetLastError();
}

hipError_t launch_soa_to_aos(eu3d_ctx* c, const double* soa, double* aos, int nc) {
    k_soa_to_aos<<<grid1d(c->g.ncell * nc, 256), 256, 0, c->stream>>>(G(c), soa, aos, nc);
    return hipGetLastError();
}

hipError_t launch_update_after_cs(eu3d_ctx* c) {
    k_update_after_cs<<<interior_grid(c->g), dim3(BX, BY), 0, c->stream>>>(G(c), c->fl, c->quirks);
    return hipGetLastError();
}

hipError_t launch_field_initial(eu3d_ctx* c) {
    k_field_initial<<<(unsigned)((c->g.ncell + 255) / 256), 256, 0, c->stream>>>(G(c), c->fl, c->quirks);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_partial_t(c);
}

hipError_t launch_partial_t(eu3d_ctx* c) {
    k_partial_t<<<interior_grid(c->g), dim3(BX, BY), 0, c->stream>>>(G(c), c->fl, c->quirks);
    return hipGetLastError();
}

hipError_t launch_fill_ghosts(eu3d_ctx* c, const int kinds[6]) {
    const G3& g = c->g;
    if (g.ni >= g.bc && g.nj >= g.bc && g.nk >= g.bc) {   // else the three passes (a source may be a ghost)
        GhostKinds K;
        for (int f = 0; f < 6; f++) K.k[f] = kinds[f];
        // the thread count of k_fill_all's three ranges (z planes and y rows PX + 16 threads a row)
        const long long rt = (long long)g.PX + 16;
        const long long n = rt * g.NY * (2 * g.bc) + rt * (2 * g.bc) * g.nk + (long long)(2 * g.bc) * g.nj * g.nk;
        k_fill_all<<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>(g, c->fl, K, c->quirks);
        return hipGetLastError();
    }
    k_fill_x<<<dim3((g.nj + 63) / 64, g.nk, 2 * g.bc), 64, 0, c->stream>>>(g, c->fl, kinds[0], kinds[1]);
    k_fill_y<<<dim3((g.NX + 63) / 64, g.nk, 2 * g.bc), 64, 0, c->stream>>>(g, c->fl, kinds[2], kinds[3], c->quirks);
    k_fill_z<<<dim3((g.NX + 63) / 64, g.NY, 2 * g.bc), 64, 0, c->stream>>>(g, c->fl, kinds[4], kinds[5]);
    return hipGetLastError();
}

hipError_t launch_advection(eu3d_ctx* c, int diff, int limiter) {
    const G3& g = c->g;
    const int fid[3] = {EU3D_FIELD_F, EU3D_FIELD_G, EU3D_FIELD_Q};
    for (int d = 0; d < 3; d++) {
        const int nx = g.ni + (d == 0), ny = g.nj + (d == 1), nz = g.nk + (d == 2);
        dim3 grid((nx + BX - 1) / BX, (ny + BY - 1) / BY, nz);
        if (d == 0) k_face_flux<0><<<grid, dim3(BX, BY), 0, c->stream>>>(g, c->fl, diff, limiter, c->quirks, c->fl.f[fid[d]]);
        if (d == 1) k_face_flux<1><<<grid, dim3(BX, BY), 0, c->stream>>>(g, c->fl, diff, limiter, c->quirks, c->fl.f[fid[d]]);
        if (d == 2) k_face_flux<2><<<grid, dim3(BX, BY), 0, c->stream>>>(g, c->fl, diff, limiter, c->quirks, c->fl.f[fid[d]]);
    }
    k_ee<<<interior_grid(g), dim3(BX, BY), 0, c->stream>>>(g, c->fl, c->xnode, c->ynode, c->znode, c->sc);
    return hipGetLastError();
}

static int n_interior_blocks(const G3& g) { return (int)interior_blocks(g.ni, g.nj, g.nk); }

// The tail kernel after the persistent kernel: one wavefront per queued cell, grid-stride over the
// device count.
template <class Acc>
void launch_coop(eu3d_ctx* c, const Acc& acc, const double* dtm_slot) {
    if (c->sched.tail <= 0) return;
    const unsigned nb = (unsigned)c->num_cu * 16;
    if (c->quirks & Q_C3) k_chem_tail<Acc, true><<<nb, 64, 0, c->stream>>>(acc, c->coop_list, c->sc, c->quirks, dtm_slot);
    else k_chem_tail<Acc, false><<<nb, 64, 0, c->stream>>>(acc, c->coop_list, c->sc, c->quirks, dtm_slot);
}

// The rest of a chemistry call after its first pass: the persistent kernel on the context stream
// and, when cells were routed to it, the tail kernel. The two take disjoint cells, so with a side
// stream the tail kernel (the longest sub-step chains, a few dozen wavefronts) runs beside the
// persistent kernel instead of after it, and the context stream waits for both.
template <class Acc>
void launch_rest(eu3d_ctx* c, const Acc& acc, const unsigned* list, const double* dtm_slot) {
    if (c->side && c->sched.tail > 0 && hipEventRecord(c->ev_fork, c->stream) == hipSuccess &&
        hipStreamWaitEvent(c->side, c->ev_fork, 0) == hipSuccess) {
        const unsigned nb = (unsigned)c->num_cu * 16;
        if (c->quirks & Q_C3) k_chem_tail<Acc, true><<<nb, 64, 0, c->side>>>(acc, c->coop_list, c->sc, c->quirks, dtm_slot);
        else k_chem_tail<Acc, false><<<nb, 64, 0, c->side>>>(acc, c->coop_list, c->sc, c->quirks, dtm_slot);
        launch_persist(c, acc, list, dtm_slot);
        if (hipEventRecord(c->ev_join, c->side) == hipSuccess) (void)hipStreamWaitEvent(c->stream, c->ev_join, 0);
        return;
    }
    launch_persist(c, acc, list, dtm_slot);
    launch_coop(c, acc, dtm_slot);
}

hipError_t ensure_coop(eu3d_ctx* c, long long n) {
    if (c->coop_cap >= n) return hipSuccess;
    if (c->coop_list) (void)hipFree(c->coop_list);
    c->coop_list = nullptr;
    c->coop_cap = 0;
    hipError_t e = hipMalloc(&c->coop_list, (size_t)n * sizeof(unsigned));
    if (e == hipSuccess) c->coop_cap = n;
    return e;
}

// Node spacings for cfl_bound: FIXED takes a one-cell direction out of the sum (spacing inf).
static void cfl_spacing(const eu3d_ctx* c, double& dx, double& dy, double& dz) {
    dx = c->dx0;
    dy = c->dy0;
    dz = c->dz0;
    if (!(c->quirks & Q_F6)) {
        if (c->g.ni == 1) dx = INFINITY;
        if (c->g.nj == 1) dy = INFINITY;
        if (c->g.nk == 1) dz = INFINITY;
    }
}

hipError_t launch_cfl(eu3d_ctx* c, double cfl, double final_time) {
    double dx, dy, dz;
    cfl_spacing(c, dx, dy, dz);
    k_cfl<<<interior_grid(c->g), dim3(BX, BY), 0, c->stream>>>(G(c), c->fl, dx, dy, dz, c->quirks, c->part);
    k_cfl_finish<<<1, 1024, 0, c->stream>>>(c->sc, c->part, n_interior_blocks(c->g), cfl, final_time);
    return hipGetLastError();
}

template <int X, int Y, bool XCD>
static void launch_xy(eu3d_ctx* c, int diff, int limiter) {
    const G3& g = c->g;
    dim3 gxy((g.ni + X - 2) / (X - 1), (g.nj + Y - 2) / (Y - 1), g.nk);
    if (diff == 1 && limiter == 0)
        k_flux_xy<X, Y, XCD, true><<<gxy, dim3(X, Y), 0, c->stream>>>(g, c->fl, diff, limiter, c->quirks, c->xnode,
                                                                      c->ynode, c->sc);
    else
        k_flux_xy<X, Y, XCD, false><<<gxy, dim3(X, Y), 0, c->stream>>>(g, c->fl, diff, limiter, c->quirks, c->xnode,
                                                                       c->ynode, c->sc);
}

template <int ZCH>
static void launch_z(eu3d_ctx* c, int diff, int limiter) {
    const G3& g = c->g;
    dim3 gz((g.ni + ZX - 1) / ZX, (g.nj + ZY - 1) / ZY, (g.nk + ZCH - 1) / ZCH);
    if (diff == 1 && limiter == 0)
        k_flux_z<ZCH, true><<<gz, dim3(ZX, ZY), 0, c->stream>>>(g, c->fl, diff, limiter, c->quirks, c->znode, c->sc);
    else
        k_flux_z<ZCH, false><<<gz, dim3(ZX, ZY), 0, c->stream>>>(g, c->fl, diff, limiter, c->quirks, c->znode, c->sc);
}

hipError_t launch_flux_z(eu3d_ctx* c, int diff, int limiter) {
    // z planes per column chunk: shorter chunks give more waves to hide latency with, at the cost
    // of re-reading 3 planes of stencil per chunk. With the 128-byte-aligned rows (MUSCL_2 +
    // minmod, one wave per SIMD): detonation 512x256x128 8 planes 0.873 ms, 32 planes 0.806 ms,
    // 64 planes 0.782 ms; bubble3d 384x128x128 0.327, 0.303, 0.326 ms. (Unaligned rows had
    // favoured 8 planes: bubble3d 0.391 against 0.403 ms for 32.) With zchunk 0 the chunk is 64
    // planes when that still gives every CU four blocks (one runs at a time, at one wave per
    // SIMD), else 32: the detonation's 1,024 blocks of 64 planes fill four rounds exactly, the
    // bubble's 384 would leave half the GPU idle in their second round.
    int zc = c->zchunk;
    if (zc == 0) {
        const long long cols = (long long)((c->g.ni + ZX - 1) / ZX) * ((c->g.nj + ZY - 1) / ZY);
        zc = cols * ((c->g.nk + 63) / 64) >= 4LL * c->num_cu ? 64 : 32;
    }
    switch (zc) {
    case 8: launch_z<8>(c, diff, limiter); break;
    case 16: launch_z<16>(c, diff, limiter); break;
    case 64: launch_z<64>(c, diff, limiter); break;
    default: launch_z<32>(c, diff, limiter);
    }
    return hipGetLastError();
}

hipError_t launch_flux_xy(eu3d_ctx* c, int diff, int limiter) {
    // tile shape of the x/y pass (threads in x, y) and the XCD-aware tile order. With the
    // 128-byte-aligned rows: 32x8 1.490 ms against 1.514 ms for 16x16 on the detonation
    // 512x256x128, 0.550 against 0.558 ms on bubble3d 384x128x128. (With unaligned rows 16x16 had
    // been faster on bubble3d, 0.671 against 0.697 ms; 64x8 1.185 ms; 64x6, 64x4, 32x4 slower.)
    switch (c->flux_variant) {
    case 1: launch_xy<16, 16, true>(c, diff, limiter); break;
    case 2: launch_xy<64, 8, true>(c, diff, limiter); break;
    case 3: launch_xy<32, 8, false>(c, diff, limiter); break;
    default: launch_xy<32, 8, true>(c, diff, limiter);
    }
    return hipGetLastError();
}

hipError_t launch_flux_fused(eu3d_ctx* c, int diff, int limiter) {
    hipError_t e = launch_flux_z(c, diff, limiter);
    return e != hipSuccess ? e : launch_flux_xy(c, diff, limiter);
}

hipError_t launch_imex_fused(eu3d_ctx* c, bool last) {
    double dx, dy, dz;
    cfl_spacing(c, dx, dy, dz);
    k_imex_fused<<<interior_grid(c->g), dim3(BX, BY), 0, c->stream>>>(G(c), c->fl, c->quirks, c->sc, dx, dy, dz,
                                                                     c->part, last ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_clear_nchem_flag(eu3d_ctx* c) {
    k_clear_nchem_flag<<<1, 1, 0, c->stream>>>(c->sc);
    return hipGetLastError();
}

hipError_t launch_cfl_finish(eu3d_ctx* c, double cfl, double final_time) {
    k_cfl_finish<<<1, 1024, 0, c->stream>>>(c->sc, c->part, n_interior_blocks(c->g), cfl, final_time);
    return hipGetLastError();
}

hipError_t launch_derive(eu3d_ctx* c) {
    k_derive<<<interior_grid(c->g), dim3(BX, BY), 0, c->stream>>>(G(c), c->fl, c->quirks);
    return hipGetLastError();
}

hipError_t launch_set_scalars(eu3d_ctx* c, double dt, double time) {
    k_set_scalars<<<1, 1, 0, c->stream>>>(c->sc, dt, time);
    return hipGetLastError();
}

hipError_t launch_fill(eu3d_ctx* c, double* a, double v, size_t n) {
    if (n) k_fill<<<grid1d(n, 256), 256, 0, c->stream>>>(a, v, n);
    return hipGetLastError();
}

hipError_t launch_adv_list(eu3d_ctx* c, const unsigned* list, const unsigned* count, unsigned max_count) {
    const unsigned nb = std::min<unsigned>((max_count + 255) / 256, (unsigned)c->num_cu * 8);
    if (nb) k_adv_list<<<nb, 256, 0, c->stream>>>(G(c), c->fl, c->quirks, list, count);
    return hipGetLastError();
}

hipError_t launch_cfl_list(eu3d_ctx* c, const unsigned* list, const unsigned* count, unsigned max_count) {
    StepOut so;
    so.part = c->part;
    cfl_spacing(c, so.dx, so.dy, so.dz);
    const unsigned nb = std::min<unsigned>((max_count + 255) / 256, (unsigned)c->num_cu * 8);
    if (nb) k_cfl_list<<<nb, 256, 0, c->stream>>>(G(c), c->fl, c->quirks, so, list, count);
    return hipGetLastError();
}

hipError_t launch_update_after_adv(eu3d_ctx* c) {
    k_add_rhs<<<interior_grid(c->g), dim3(BX, BY), 0, c->stream>>>(G(c), c->fl, (c->quirks & Q_F1) ? NS : NC);
    return launch_update_after_cs(c);
}

hipError_t launch_explicit(eu3d_ctx* c) {
    k_explicit<<<interior_grid(c->g), dim3(BX, BY), 0, c->stream>>>(G(c), c->fl, c->quirks);
    return launch_update_after_cs(c);
}

hipError_t launch_diagonalized(eu3d_ctx* c) {
    k_diagonalized<<<interior_grid(c->g), dim3(BX, BY), 0, c->stream>>>(G(c), c->fl, c->quirks);
    return hipGetLastError();
}

hipError_t launch_update_imex(eu3d_ctx* c) {
    k_update_imex<<<interior_grid(c->g), dim3(BX, BY), 0, c->stream>>>(G(c), c->fl, c->sc);
    hipError_t e = launch_update_after_cs(c);
    if (e != hipSuccess) return e;
    return launch_partial_t(c);
}

hipError_t launch_trapezoid(eu3d_ctx* c) {
    if (c->quirks & Q_C3)
        k_trapezoid<true><<<interior_grid(c->g), dim3(BX, BY), 0, c->stream>>>(G(c), c->fl, c->quirks, c->sc);
    else
        k_trapezoid<false><<<interior_grid(c->g), dim3(BX, BY), 0, c->stream>>>(G(c), c->fl, c->quirks, c->sc);
    return hipGetLastError();
}

hipError_t launch_check_finite(eu3d_ctx* c) {
    k_check_finite<<<interior_grid(c->g), dim3(BX, BY), 0, c->stream>>>(G(c), c->fl, c->sc);
    return hipGetLastError();
}

hipError_t launch_trap_adaptive(eu3d_ctx* c, int nchem_max, bool predict_only, int nthr, double T_cold,
                                unsigned* slist, unsigned* scount, bool adv, int nthr_hi, bool step_src,
                                bool step_cfl) {
    const int nmax = nchem_max > 0 ? nchem_max : 10000;
    StepOut so;
    so.src = step_src ? 1 : 0;
    if (step_cfl) {
        so.part = c->part;
        cfl_spacing(c, so.dx, so.dy, so.dz);
    }
    double* slot = &c->sc->dtm;
    hipError_t e = ensure_coop(c, (long long)c->g.ni * c->g.nj * c->g.nk);
    if (e != hipSuccess) return e;
    const TailRoute rt{c->coop_list, predict_only ? 0 : c->sched.tail};
    k_chem_reset<<<1, 1, 0, c->stream>>>(c->sc, slot, 0.0, 1);
    const bool c3 = c->quirks & Q_C3;
    const dim3 gr = interior_grid(c->g), bl(BX, BY);
    const int po = predict_only ? 1 : 0;
    if (c3 && adv)
        k_trap_first<true, true><<<gr, bl, 0, c->stream>>>(G(c), c->fl, c->quirks, c->sc, c->sc, c->active, nmax, po,
                                                           nthr, nthr_hi, T_cold, slist, scount, rt, so);
    else if (c3)
        k_trap_first<true, false><<<gr, bl, 0, c->stream>>>(G(c), c->fl, c->quirks, c->sc, c->sc, c->active, nmax, po,
                                                            nthr, nthr_hi, T_cold, slist, scount, rt, so);
    else if (adv)
        k_trap_first<false, true><<<gr, bl, 0, c->stream>>>(G(c), c->fl, c->quirks, c->sc, c->sc, c->active, nmax, po,
                                                            nthr, nthr_hi, T_cold, slist, scount, rt, so);
    else
        k_trap_first<false, false><<<gr, bl, 0, c->stream>>>(G(c), c->fl, c->quirks, c->sc, c->sc, c->active, nmax, po,
                                                             nthr, nthr_hi, T_cold, slist, scount, rt, so);
    if (adv) k_adv_nochem<<<4 * c->num_cu, 256, 0, c->stream>>>(G(c), c->fl, c->quirks, c->sc, so);
    if (!predict_only && adv) {   // the step: the derived fields are refreshed after its last step
        GridCellsStep acc{c->g, c->fl};
        acc.src = step_src;
        acc.store_c = step_src;
        acc.cfl = so.part;
        acc.cdx = so.dx;
        acc.cdy = so.dy;
        acc.cdz = so.dz;
        acc.q = c->quirks;
        launch_rest(c, acc, c->active, slot);
    } else if (!predict_only) {
        GridCells acc{c->g, c->fl};
        launch_rest(c, acc, c->active, slot);
    }
    return hipGetLastError();
}

// The step's prediction pass alone (k_trap_first with the fused Update_after_Adv and
// Update_after_CS, the prediction and the first sub-step), for timing it: launch_chem_reset_step
// is the per-call reset that precedes it in the step.
hipError_t launch_chem_reset_step(eu3d_ctx* c) {
    k_chem_reset<<<1, 1, 0, c->stream>>>(c->sc, &c->sc->dtm, 0.0, 1);
    return hipGetLastError();
}

hipError_t launch_trap_first_step(eu3d_ctx* c, int nchem_max) {
    hipError_t e = ensure_coop(c, (long long)c->g.ni * c->g.nj * c->g.nk);
    if (e != hipSuccess) return e;
    const TailRoute rt{c->coop_list, c->sched.tail};
    const int nmax = nchem_max > 0 ? nchem_max : 10000;
    const dim3 gr = interior_grid(c->g), bl(BX, BY);
    StepOut so;   // as on a step that is not the last of its call
    so.src = 0;
    so.part = c->part;
    cfl_spacing(c, so.dx, so.dy, so.dz);
    if (c->quirks & Q_C3)
        k_trap_first<true, true><<<gr, bl, 0, c->stream>>>(G(c), c->fl, c->quirks, c->sc, c->sc, c->active, nmax, 0, 0,
                                                           0, 0.0, nullptr, nullptr, rt, so);
    else
        k_trap_first<false, true><<<gr, bl, 0, c->stream>>>(G(c), c->fl, c->quirks, c->sc, c->sc, c->active, nmax, 0, 0,
                                                            0, 0.0, nullptr, nullptr, rt, so);
    return hipGetLastError();
}

hipError_t launch_chem_batch(eu3d_ctx* c, long long n, double* rho, double* T, double* E, double dtm, int nchem_max,
                             int* nchem_out, int nthr, double T_cold, unsigned* slist, unsigned* scount, int nthr_hi) {
    const int nmax = nchem_max > 0 ? nchem_max : 10000;
    if (c->batch_cap < n) {
        if (c->batch_active) (void)hipFree(c->batch_active);
        c->batch_active = nullptr;
        c->batch_cap = 0;
        hipError_t e = hipMalloc(&c->batch_active, (size_t)n * (sizeof(unsigned) + sizeof(int)));
        if (e != hipSuccess) return e;
        c->batch_cap = n;
    }
    unsigned* list = c->batch_active;
    int* keep = reinterpret_cast<int*>(list + c->batch_cap);
    double* slot = &c->sc->dtm;
    hipError_t e = ensure_coop(c, n);
    if (e != hipSuccess) return e;
    const TailRoute rt{c->coop_list, c->sched.tail};
    k_chem_reset<<<1, 1, 0, c->stream>>>(c->sc, slot, dtm, 0);
    BatchCells acc{n, rho, T, E, keep};
    const unsigned gb = (unsigned)((n + 255) / 256);
    if (c->quirks & Q_C3) {
        k_batch_first<true><<<gb, 256, 0, c->stream>>>(acc, nchem_out, keep, c->quirks, dtm, nmax, c->sc, list, nthr, nthr_hi,
                                                       T_cold, slist, scount, rt);
    } else {
        k_batch_first<false><<<gb, 256, 0, c->stream>>>(acc, nchem_out, keep, c->quirks, dtm, nmax, c->sc, list, nthr, nthr_hi,
                                                        T_cold, slist, scount, rt);
    }
    launch_rest(c, acc, list, slot);
    return hipGetLastError();
}

// FP64 adaptive trapezoid on the cells of a device list (count on the device), grid form;
// the chemistry interval is the context's dt. max_count bounds the launch.
hipError_t launch_trap_list(eu3d_ctx* c, const unsigned* list, const unsigned* count, unsigned max_count,
                            unsigned* work, int nchem_max) {
    const int nmax = nchem_max > 0 ? nchem_max : 10000;
    hipError_t e = ensure_coop(c, max_count);
    if (e != hipSuccess) return e;
    // a short list sends its multi-sub-step cells to the tail kernel, which holds two cells per
    // wavefront at one wavefront per SIMD
    const TailRoute rt{c->coop_list, c->sched.tail, 8u * (unsigned)c->num_cu};
    k_chem_reset<<<1, 1, 0, c->stream>>>(c->sc, &c->sc->dtm, 0.0, 1);
    const unsigned nb = std::min<unsigned>((max_count + 255) / 256, (unsigned)c->num_cu * 8);
    const bool c3 = c->quirks & Q_C3;
    if (nb) {
        if (c3)
            k_trap_first_list<true><<<nb, 256, 0, c->stream>>>(G(c), c->fl, c->quirks, list, count, c->sc, work, nmax, rt);
        else
            k_trap_first_list<false><<<nb, 256, 0, c->stream>>>(G(c), c->fl, c->quirks, list, count, c->sc, work, nmax, rt);
    }
    GridCells acc{c->g, c->fl};
    launch_rest(c, acc, work, &c->sc->dtm);
    return hipGetLastError();
}

// Batch form of launch_trap_list: flat cells at rest (rho [ns][n], T, E).
hipError_t launch_batch_list(eu3d_ctx* c, long long n, double* rho, double* T, double* E, double dtm,
                             const unsigned* list, const unsigned* count, unsigned max_count, unsigned* work,
                             int* nchem_keep, int nchem_max) {
    const int nmax = nchem_max > 0 ? nchem_max : 10000;
    hipError_t e = ensure_coop(c, max_count);
    if (e != hipSuccess) return e;
    // a short list sends its multi-sub-step cells to the tail kernel, which holds two cells per
    // wavefront at one wavefront per SIMD
    const TailRoute rt{c->coop_list, c->sched.tail, 8u * (unsigned)c->num_cu};
    k_chem_reset<<<1, 1, 0, c->stream>>>(c->sc, &c->sc->dtm, dtm, 0);
    BatchCells acc{n, rho, T, E, nchem_keep};
    const unsigned nb = std::min<unsigned>((max_count + 255) / 256, (unsigned)c->num_cu * 8);
    const bool c3 = c->quirks & Q_C3;
    if (nb) {
        if (c3)
            k_batch_first_list<true><<<nb, 256, 0, c->stream>>>(acc, nchem_keep, c->quirks, dtm, nmax, list, count, c->sc,
                                                                 work, rt);
        else
            k_batch_first_list<false><<<nb, 256, 0, c->stream>>>(acc, nchem_keep, c->quirks, dtm, nmax, list, count, c->sc,
                                                                  work, rt);
    }
    launch_rest(c, acc, work, &c->sc->dtm);
    return hipGetLastError();
}

hipError_t launch_step(eu3d_ctx* c, const eu3d_step_params* p, bool first, bool last) {
    hipError_t e;
    if (p->cfl > 0) {
        // the fused IMEX epilogue and the step passes of the adaptive trapezoid and of the
        // surrogate leave the CFL partials of the new state behind for the next step of the call
        e = (!first && c->part_ready) ? launch_cfl_finish(c, p->cfl, p->final_time)
                                      : launch_cfl(c, p->cfl, p->final_time);
        if (e != hipSuccess) return e;
    } else {
        k_advance_time<<<1, 1, 0, c->stream>>>(c->sc);
    }
    c->part_ready = false;
    if ((e = launch_fill_ghosts(c, p->bc_kinds)) != hipSuccess) return e;
    if (p->rk == 2) {
        const size_t n = c->g.n3 * NC;
        k_copy<<<grid1d(n, 256), 256, 0, c->stream>>>(c->fl.f[EU3D_FIELD_CS], c->rk_save, n);
        if ((e = launch_flux_fused(c, p->diff, p->limiter)) != hipSuccess) return e;
        if ((e = launch_update_after_adv(c)) != hipSuccess) return e;
        if ((e = launch_fill_ghosts(c, p->bc_kinds)) != hipSuccess) return e;
        if ((e = launch_flux_fused(c, p->diff, p->limiter)) != hipSuccess) return e;
        k_rk2_combine<<<interior_grid(c->g), dim3(BX, BY), 0, c->stream>>>(G(c), c->fl, c->rk_save);
        if ((e = launch_update_after_cs(c)) != hipSuccess) return e;
        switch (p->chem) {
        case EU3D_CHEM_TRAPEZOID:
            if ((e = launch_trapezoid(c)) != hipSuccess) return e;
            e = launch_explicit(c);
            break;
        case EU3D_CHEM_TRAP_ADAPTIVE: e = launch_trap_adaptive(c, p->nchem_max, false); break;
        case EU3D_CHEM_SURROGATE: e = launch_surrogate_grid(c, p->nchem_max); break;
        default: e = hipSuccess;
        }
        return e;
    }
    if ((e = launch_flux_fused(c, p->diff, p->limiter)) != hipSuccess) return e;
    switch (p->chem) {
    case EU3D_CHEM_IMEX:
        e = launch_imex_fused(c, last);
        c->part_ready = true;
        break;
    case EU3D_CHEM_TRAPEZOID:
        if ((e = launch_update_after_adv(c)) != hipSuccess) return e;
        if ((e = launch_trapezoid(c)) != hipSuccess) return e;
        e = launch_explicit(c);
        break;
    case EU3D_CHEM_TRAP_ADAPTIVE:   // Update_after_Adv fused into the prediction pass
        e = launch_trap_adaptive(c, p->nchem_max, false, 0, 0.0, nullptr, nullptr, true, 0, last, !last);
        c->part_ready = !last;
        break;
    case EU3D_CHEM_SURROGATE: e = launch_surrogate_grid(c, p->nchem_max, true, last); break;
    default: e = launch_update_after_adv(c);
    }
    return e;
}

}  // namespace eu3d
