// MLP chemistry surrogate on gfx950: BF16 MFMA GEMMs with FP32 accumulation, fused epilogues,
// backward pass and Adam; plus the features, apply and fallback kernels that couple it to the
// FP64 state. This part has no counterpart in the reference (its "DNN" scheme is a stub that
// only advects, src/Euler.cpp:773-780).
//
// Network: x (12 features, padded to 32) -> 4 x [Linear 256 + GELU] -> Linear 9 (padded to 16).
// Storage: activations row-major [rows x features] in BF16; weights kept as FP32 masters in
// "out x in" layout (Wt) with BF16 copies in both layouts (Wt and W) so every GEMM reads both
// operands K-contiguous.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "eu3d_internal.hpp"

namespace {

typedef uint16_t bf16_t;
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even (NaN stays NaN): the gfx950 conversion instruction.
__device__ __forceinline__ bf16_t f2bf(float f) { return __builtin_bit_cast(bf16_t, static_cast<__bf16>(f)); }
__host__ __device__ __forceinline__ float bf2f(bf16_t h) {
    uint32_t u = (uint32_t)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

// GELU in its tanh form, 0.5 z (1 + tanh(sqrt(2/pi) (z + 0.044715 z^3))), evaluated as the equal
// z * sigmoid(s) with s = 2 sqrt(2/pi) (z + 0.044715 z^3): one exp2 and one reciprocal per
// element, so the epilogue of the fused network kernel stays short next to its MFMA work.
// Training, inference and the CPU oracle (oracle/mlp.py) all use this form.
constexpr float GELU_C1 = 1.5957691216057308f;                 // 2 sqrt(2/pi)
constexpr float GELU_C2 = GELU_C1 * 0.044715f;
constexpr float LOG2E = 1.4426950408889634f;
__device__ __forceinline__ float gelu_sig(float z) {          // sigmoid(s)
    const float w = z * fmaf(-GELU_C2 * LOG2E, z * z, -GELU_C1 * LOG2E);   // -s log2(e)
    return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(w));
}
__device__ __forceinline__ float gelu(float z) { return z * gelu_sig(z); }
__device__ __forceinline__ float dgelu(float z) {
    const float g = gelu_sig(z);
    return g + z * g * (1.0f - g) * fmaf(3.0f * GELU_C2, z * z, GELU_C1);
}

// ------------------------------------------------------------------ GEMM
// C[M x N] = A[M x K] * B[K x N]. AKC: A stored K-contiguous (A[m * lda + k]) else
// M-contiguous (A[k * lda + m]); BKC: B stored K-contiguous (B[n * ldb + k]) else (B[k * ldb + n]).
// Block tile 128 x 128 x 32, four waves in 2 x 2, each wave 64 x 64 = 4 x 4 MFMA 16x16x32 tiles.
// Tiles are staged in LDS as [row][k] (K-contiguous, padded pitch) so a lane's MFMA fragment,
// 8 consecutive k of one row, is one 16-byte LDS read. Split-K over gridDim.z writes partials.
constexpr int BM = 128, BN = 128, BK = 32, LDK = BK + 8;

enum Epi { EPI_F32 = 0, EPI_BIAS_F32 = 1, EPI_BIAS_GELU = 2, EPI_DGELU = 3 };

struct GemmArgs {
    int M, N, K, kchunk;
    const bf16_t* A;
    int lda;
    const bf16_t* B;
    int ldb;
    float* Cf;            // EPI_F32 / EPI_BIAS_F32 output (split s at Cf + s * M * ldc)
    bf16_t* Cb;           // EPI_BIAS_GELU activation / EPI_DGELU gradient output
    bf16_t* Z;            // EPI_BIAS_GELU: pre-activation store (may be null); EPI_DGELU: read
    const float* bias;
    int ldc;
};

// One operand tile (ROWS x BK) moves in two phases, so that the loads of the next K step are in
// flight while the MFMAs of the current one run: stage_load takes the thread's two 16-byte
// pieces into registers, stage_store writes them to S[row][k] (the M- or N-contiguous case
// transposing through 8 two-byte stores).
template <int P>
struct StageRegsN {
    uint4 v[P];
};
template <bool KC, int ROWS, int NTH = 256, int BKX = BK, int P = ROWS * BKX / 8 / NTH>
__device__ __forceinline__ StageRegsN<P> stage_load(const bf16_t* G, int ld, int r0, int nrows, int k0, int kend,
                                                    int tid) {
    static_assert(ROWS * BKX / 8 == P * NTH, "whole pieces per thread");
    StageRegsN<P> R;
#pragma unroll
    for (int p = 0; p < P; p++) {
        const int c = tid + NTH * p;
        uint4 v = make_uint4(0, 0, 0, 0);
        if (KC) {   // G[r * ld + k]; 8-element pieces along k
            const int r = c / (BKX / 8), kc = (c % (BKX / 8)) * 8;
            const int gr = r0 + r, gk = k0 + kc;
            if (gr < nrows) {
                const bf16_t* src = G + (size_t)gr * ld + gk;
                if (gk + 8 <= kend && (((uintptr_t)src) & 15) == 0) {
                    v = *reinterpret_cast<const uint4*>(src);
                } else {
                    bf16_t tmp[8];
                    for (int e = 0; e < 8; e++) tmp[e] = (gk + e < kend) ? src[e] : (bf16_t)0;
                    memcpy(&v, tmp, 16);
                }
            }
        } else {    // G[k * ld + r]; 8-element pieces along r, k fastest over the lanes
            const int k = c % BKX, rc = (c / BKX) * 8;
            const int gk = k0 + k, gr = r0 + rc;
            if (gk < kend) {
                const bf16_t* src = G + (size_t)gk * ld + gr;
                if (gr + 8 <= nrows && (((uintptr_t)src) & 15) == 0) {
                    v = *reinterpret_cast<const uint4*>(src);
                } else {
                    bf16_t tmp[8];
                    for (int e = 0; e < 8; e++) tmp[e] = (gr + e < nrows) ? src[e] : (bf16_t)0;
                    memcpy(&v, tmp, 16);
                }
            }
        }
        R.v[p] = v;
    }
    return R;
}
template <bool KC, int ROWS, int NTH = 256, int BKX = BK, int LD = LDK, int P = ROWS * BKX / 8 / NTH>
__device__ __forceinline__ void stage_store(bf16_t (*S)[LD], const StageRegsN<P>& R, int tid) {
#pragma unroll
    for (int p = 0; p < P; p++) {
        const int c = tid + NTH * p;
        if (KC) {
            const int r = c / (BKX / 8), kc = (c % (BKX / 8)) * 8;
            *reinterpret_cast<uint4*>(&S[r][kc]) = R.v[p];
        } else {
            // k fastest over the lanes: each two-byte store of a wavefront covers 32 consecutive k
            // of two rows (with r fastest, the rows 8 apart fell in one LDS bank)
            const int k = c % BKX, rc = (c / BKX) * 8;
            bf16_t tmp[8];
            memcpy(tmp, &R.v[p], 16);
            for (int e = 0; e < 8; e++) S[rc + e][k] = tmp[e];
        }
    }
}

template <bool AKC, bool BKC, int EPI>
__global__ void __launch_bounds__(256) k_gemm(GemmArgs a) {
    __shared__ bf16_t As[BM][LDK];
    __shared__ bf16_t Bs[BN][LDK];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;   // rows on x: no 65535 limit on M
    const int kbeg = blockIdx.z * a.kchunk, kend = min(a.K, kbeg + a.kchunk);
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
    const int fr = lane & 15, fk = (lane >> 4) * 8;
    f32x4_t acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = (f32x4_t){0.f, 0.f, 0.f, 0.f};
    auto ra = stage_load<AKC, BM>(a.A, a.lda, m0, a.M, kbeg, kend, tid);
    auto rb = stage_load<BKC, BN>(a.B, a.ldb, n0, a.N, kbeg, kend, tid);
    for (int k0 = kbeg; k0 < kend; k0 += BK) {
        stage_store<AKC, BM>(As, ra, tid);
        stage_store<BKC, BN>(Bs, rb, tid);
        __syncthreads();
        if (k0 + BK < kend) {   // the next K step's loads, in flight during this step's MFMAs
            ra = stage_load<AKC, BM>(a.A, a.lda, m0, a.M, k0 + BK, kend, tid);
            rb = stage_load<BKC, BN>(a.B, a.ldb, n0, a.N, k0 + BK, kend, tid);
        }
        bf16x8_t af[4], bfr[4];
#pragma unroll
        for (int i = 0; i < 4; i++) af[i] = *reinterpret_cast<const bf16x8_t*>(&As[wm + 16 * i + fr][fk]);
#pragma unroll
        for (int j = 0; j < 4; j++) bfr[j] = *reinterpret_cast<const bf16x8_t*>(&Bs[wn + 16 * j + fr][fk]);
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }
    // D fragment: lane holds rows 4 * (lane / 16) + r, column lane % 16 of each 16 x 16 tile
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = m0 + wm + 16 * i + 4 * (lane >> 4) + r;
                const int col = n0 + wn + 16 * j + (lane & 15);
                if (row >= a.M || col >= a.N) continue;
                const float v = acc[i][j][r];
                const size_t o = (size_t)row * a.ldc + col;
                if (EPI == EPI_F32) {
                    a.Cf[(size_t)blockIdx.z * a.M * a.ldc + o] = v;
                } else if (EPI == EPI_BIAS_F32) {
                    a.Cf[o] = v + a.bias[col];
                } else if (EPI == EPI_BIAS_GELU) {
                    const float z = v + a.bias[col];
                    const bf16_t zb = f2bf(z);
                    if (a.Z) a.Z[o] = zb;
                    a.Cb[o] = f2bf(gelu(bf2f(zb)));
                } else {   // EPI_DGELU: gradient wrt the pre-activation of the previous layer
                    a.Cb[o] = f2bf(v * dgelu(bf2f(a.Z[o])));
                }
            }
}

// The weight-gradient GEMM of a 256 x 256 layer, C = A B with A and B stored k-major (the rows
// of dZ and of the activations): one block computes all 256 x 256 outputs of its K chunk, so each
// operand is read once (two 128 x 128 tiles read each twice). Eight waves in 4 x 2, each 64 x 128
// = 4 x 8 MFMA tiles; the next K step is loaded during the current step's MFMAs, as in k_gemm
// (steps of 64 needed 11 more registers than two waves per SIMD allow and were slower). Split-K
// partials out, as k_gemm's EPI_F32.
constexpr int WB = 256, WBK = BK, WLD = WBK + 8;
__global__ void __launch_bounds__(512) k_gemm_wgrad(GemmArgs a) {
    __shared__ bf16_t As[WB][WLD];
    __shared__ bf16_t Bs[WB][WLD];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int kbeg = blockIdx.z * a.kchunk, kend = min(a.K, kbeg + a.kchunk);
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 128;
    const int fr = lane & 15, fk = (lane >> 4) * 8;
    f32x4_t acc[4][8];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 8; j++) acc[i][j] = (f32x4_t){0.f, 0.f, 0.f, 0.f};
    auto ra = stage_load<false, WB, 512, WBK>(a.A, a.lda, 0, a.M, kbeg, kend, tid);
    auto rb = stage_load<false, WB, 512, WBK>(a.B, a.ldb, 0, a.N, kbeg, kend, tid);
    for (int k0 = kbeg; k0 < kend; k0 += WBK) {
        stage_store<false, WB, 512, WBK, WLD>(As, ra, tid);
        stage_store<false, WB, 512, WBK, WLD>(Bs, rb, tid);
        __syncthreads();
        if (k0 + WBK < kend) {   // the next K step's loads, in flight during this step's MFMAs
            ra = stage_load<false, WB, 512, WBK>(a.A, a.lda, 0, a.M, k0 + WBK, kend, tid);
            rb = stage_load<false, WB, 512, WBK>(a.B, a.ldb, 0, a.N, k0 + WBK, kend, tid);
        }
#pragma unroll
        for (int kk = 0; kk < WBK; kk += 32) {
            bf16x8_t af[4];
#pragma unroll
            for (int i = 0; i < 4; i++) af[i] = *reinterpret_cast<const bf16x8_t*>(&As[wm + 16 * i + fr][kk + fk]);
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const bf16x8_t b = *reinterpret_cast<const bf16x8_t*>(&Bs[wn + 16 * j + fr][kk + fk]);
#pragma unroll
                for (int i = 0; i < 4; i++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], b, acc[i][j], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    float* C = a.Cf + (size_t)blockIdx.z * a.M * a.ldc;
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 8; j++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = wm + 16 * i + 4 * (lane >> 4) + r, col = wn + 16 * j + (lane & 15);
                if (row < a.M && col < a.N) C[(size_t)row * a.ldc + col] = acc[i][j][r];
            }
}

// Sum split-K partials: out[i] = sum_s part[s * n + i]
__global__ void k_reduce_splits(const float* __restrict__ part, float* __restrict__ out, int n, int splits) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.f;
    for (int k = 0; k < splits; k++) s += part[(size_t)k * n + i];
    out[i] = s;
}

// First level of a two-level split-K sum: out[g * n + i] = sum of partials g*per .. g*per+per-1.
__global__ void k_reduce_splits_grouped(const float* __restrict__ part, float* __restrict__ out, int n, int splits,
                                        int per) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, g = blockIdx.y;
    if (i >= n) return;
    const int k1 = min(splits, (g + 1) * per);
    float s = 0.f;
    for (int k = g * per; k < k1; k++) s += part[(size_t)k * n + i];
    out[(size_t)g * n + i] = s;
}

// Column sums of a bf16 [rows x ncol] matrix (bias gradients; ncol a multiple of 8, at most 256),
// deterministic two-level: block b sums rows [b * rows_per_block, ...) per column into
// part[b * ncol + col]. A row is read by ncol / 8 threads with 16-byte loads; the block's
// 2048 / ncol row groups stride over its rows, and their sums are combined in LDS in group order.
__global__ void __launch_bounds__(256) k_colsum_bf16(const bf16_t* __restrict__ X, int rows, int ncol,
                                                     int rows_per_block, float* __restrict__ part) {
    __shared__ float red[2048];
    const int tpr = ncol / 8, groups = 256 / tpr;
    const int t = threadIdx.x, cg = t % tpr, rg = t / tpr;
    const int r0 = blockIdx.x * rows_per_block, r1 = min(rows, r0 + rows_per_block);
    float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int r = r0 + rg; r < r1; r += groups) {
        const uint4 v = *reinterpret_cast<const uint4*>(X + (size_t)r * ncol + cg * 8);
        const unsigned w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int j = 0; j < 4; j++) {
            s[2 * j] += __uint_as_float(w[j] << 16);
            s[2 * j + 1] += __uint_as_float(w[j] & 0xffff0000u);
        }
    }
#pragma unroll
    for (int j = 0; j < 8; j++) red[rg * ncol + cg * 8 + j] = s[j];
    __syncthreads();
    if (t < ncol) {
        float a = 0.f;
        for (int g = 0; g < groups; g++) a += red[g * ncol + t];
        part[(size_t)blockIdx.x * ncol + t] = a;
    }
}

// Loss gradient: dY = 2 (Y - T) / (n * nout) on the nout real outputs, zero on the padding;
// per-block partial sums of the squared error.
__global__ void k_mse_grad(const float* __restrict__ Y, const float* __restrict__ T, int n, int ld, int nout,
                           float scale, bf16_t* __restrict__ dY, float* __restrict__ part) {
    __shared__ float red[256];
    float s = 0.f;
    const size_t total = (size_t)n * ld;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const int col = (int)(t % ld);
        float g = 0.f;
        if (col < nout) {
            const float d = Y[t] - T[t];
            s += d * d;
            g = scale * d;
        }
        dY[t] = f2bf(g);
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Adam on FP32 masters (Wt layout [out x in] and bias [out]); refreshes the BF16 copies Wt and
// W ([in x out]). Padded rows/columns stay zero.
__global__ void k_adam(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v, const float* __restrict__ g,
                       int n_out, int n_in, int real_out, int real_in, float lr, float b1, float b2, float eps,
                       float bc1, float bc2, float wd, bf16_t* __restrict__ wt_b, bf16_t* __restrict__ w_b) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_out * n_in) return;
    const int o = i / n_in, k = i % n_in;
    if (o >= real_out || k >= real_in) {
        w[i] = 0.f;
        if (wt_b) wt_b[i] = 0;
        if (w_b) w_b[(size_t)k * n_out + o] = 0;
        return;
    }
    const float gi = g[i] + wd * w[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
    w[i] = wi;
    const bf16_t h = f2bf(wi);
    if (wt_b) wt_b[i] = h;
    if (w_b) w_b[(size_t)k * n_out + o] = h;
}

__global__ void k_cast_weights(const float* __restrict__ w, int n_out, int n_in, bf16_t* __restrict__ wt_b,
                               bf16_t* __restrict__ w_b) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_out * n_in) return;
    const int o = i / n_in, k = i % n_in;
    const bf16_t h = f2bf(w[i]);
    wt_b[i] = h;
    w_b[(size_t)k * n_out + o] = h;
}

__global__ void k_f32_to_bf16(const float* __restrict__ x, bf16_t* __restrict__ y, size_t n) {
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
        y[t] = f2bf(x[t]);
}

// ------------------------------------------------------------------ fused inference
// The whole network for a tile of FM rows in one workgroup: the activations stay in LDS from
// the input features to the last layer, so HBM sees only the features in (64 B per row) and the
// outputs out (64 B per row). Each layer is [FM x K] x [K x N] on MFMA 16x16x32 BF16. Eight
// waves (two per SIMD, so one wave's epilogue overlaps the other's MFMAs) each own 16 rows and
// every output column (N/16 accumulator tiles). K goes in chunks of 32 whose weights (W^T rows,
// K-contiguous) pass through registers into a double-buffered LDS stage shared by the eight
// waves; each chunk is loaded two chunks ahead of its use. The epilogue (bias, BF16 rounding of
// the pre-activation, GELU, BF16) writes each wave's own rows back in place, which no other
// wave reads.
constexpr int FM = 128, FT = 512, FPITCH = 256 + 8, WPITCH = 32 + 8;

struct MlpView {
    const bf16_t* wt[5];
    const float* b[5];
};

// chunk c of W^T (N rows x 32 k): thread tid holds 16-byte pieces tid and tid + FT
struct WPiece {
    uint4 a, b;
};
// KSRC: the row length of wt when it is shorter than the chunk (16 < 32: zeros beyond it)
template <int N, int K, int KSRC = K>
__device__ __forceinline__ WPiece load_w(const bf16_t* __restrict__ wt, int c, int tid) {
    WPiece r{make_uint4(0, 0, 0, 0), make_uint4(0, 0, 0, 0)};
    const int p0 = tid, p1 = tid + FT;
    if (p0 < N * 4 && c * 32 + (p0 & 3) * 8 < KSRC)
        r.a = *reinterpret_cast<const uint4*>(wt + (size_t)(p0 >> 2) * KSRC + c * 32 + (p0 & 3) * 8);
    if (p1 < N * 4 && c * 32 + (p1 & 3) * 8 < KSRC)
        r.b = *reinterpret_cast<const uint4*>(wt + (size_t)(p1 >> 2) * KSRC + c * 32 + (p1 & 3) * 8);
    return r;
}
template <int N>
__device__ __forceinline__ void store_w(bf16_t (*ws)[WPITCH], const WPiece& r, int tid) {
    const int p0 = tid, p1 = tid + FT;
    if (p0 < N * 4) *reinterpret_cast<uint4*>(&ws[p0 >> 2][(p0 & 3) * 8]) = r.a;
    if (p1 < N * 4) *reinterpret_cast<uint4*>(&ws[p1 >> 2][(p1 & 3) * 8]) = r.b;
}

// Training outputs of one hidden layer: the pre-activation Z and the activation GELU(Z), BF16
// [n x 256] each (the backward pass reads both).
struct TrainOut {
    bf16_t* z;
    bf16_t* a;
};

// How the eight wavefronts split a layer's FM x N output: RG row groups x CG column groups,
// each wavefront RT x CT tiles of 16 x 16. For N = 256, 4 x 2 groups of 2 x 8 tiles: a
// wavefront reads 2 + 8 fragments per 32-deep K chunk for 16 MFMAs (16 rows x all columns read
// 1 + 16), which keeps the LDS reads under what the MFMAs consume.
template <int N>
struct WaveTile {
    static constexpr int CT = N >= 128 ? 8 : N / 16, CG = N / (16 * CT), RG = 8 / CG, RT = FM / (16 * RG);
    static_assert(CG * RG == 8 && RT * 16 * RG == FM, "eight wavefronts over FM rows");
    int wr, wc;   // first row and column of the wavefront's block
    __device__ __forceinline__ explicit WaveTile(int tid) {
        const int w = tid >> 6;
        wr = (w / CG) * RT * 16;
        wc = (w % CG) * CT * 16;
    }
};

// acc = act[the wavefront's rows][0:K] x W^T[its columns][0:K]^T on MFMA, the weights streamed
// in 32-column chunks through the double-buffered stage ws. With CG > 1 two wavefronts read the
// same rows of act, so the function ends with a barrier before anyone writes them.
template <int K, int N, int KSRC = K>
__device__ __forceinline__ void fused_mfma(const bf16_t (*act)[FPITCH], bf16_t (*ws)[256][WPITCH],
                                           const bf16_t* __restrict__ wt, int tid, int lane, const WaveTile<N>& wt_,
                                           f32x4_t (&acc)[WaveTile<N>::RT][WaveTile<N>::CT]) {
    constexpr int RT = WaveTile<N>::RT, CT = WaveTile<N>::CT, KS = K / 32;
    const int fr = lane & 15, fk = (lane >> 4) * 8;
#pragma unroll
    for (int i = 0; i < RT; i++)
#pragma unroll
        for (int j = 0; j < CT; j++) acc[i][j] = (f32x4_t){0.f, 0.f, 0.f, 0.f};
    // chunk c waits in register set c % 2 from its load (two chunks ahead) to its LDS store
    WPiece r0 = load_w<N, K, KSRC>(wt, 0, tid), r1{};
    if (KS > 1) r1 = load_w<N, K, KSRC>(wt, 1, tid);
    __syncthreads();   // the previous layer is done with ws and act
    store_w<N>(ws[0], r0, tid);
    if (KS > 2) r0 = load_w<N, K, KSRC>(wt, 2, tid);
#pragma unroll
    for (int ks = 0; ks < KS; ks++) {   // unrolled: the register set of each chunk is known
        __syncthreads();   // chunk ks staged; the other buffer's readers are done
        if (ks + 1 < KS) store_w<N>(ws[(ks + 1) & 1], ((ks + 1) & 1) ? r1 : r0, tid);
        if (ks + 3 < KS) {
            if ((ks + 1) & 1) r1 = load_w<N, K, KSRC>(wt, ks + 3, tid);
            else r0 = load_w<N, K, KSRC>(wt, ks + 3, tid);
        }
        bf16x8_t a[RT];
#pragma unroll
        for (int i = 0; i < RT; i++) a[i] = *reinterpret_cast<const bf16x8_t*>(&act[wt_.wr + 16 * i + fr][ks * 32 + fk]);
#pragma unroll
        for (int j = 0; j < CT; j++) {
            const bf16x8_t b = *reinterpret_cast<const bf16x8_t*>(&ws[ks & 1][wt_.wc + 16 * j + fr][fk]);
#pragma unroll
            for (int i = 0; i < RT; i++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b, acc[i][j], 0, 0, 0);
        }
    }
    if (WaveTile<N>::CG > 1) __syncthreads();
}

template <int K, int N, bool LAST, bool TRAIN = false>
__device__ __forceinline__ void fused_layer(bf16_t (*act)[FPITCH], bf16_t (*ws)[256][WPITCH], const bf16_t* __restrict__ wt,
                                            const float* __restrict__ bias, int tid, int lane, long long row0,
                                            int n, float* __restrict__ Y, TrainOut tr = {}) {
    using TL = WaveTile<N>;
    constexpr int RT = TL::RT, CT = TL::CT;
    const TL tl(tid);
    const int fr = lane & 15;
    f32x4_t acc[RT][CT];
    fused_mfma<K, N>(act, ws, wt, tid, lane, tl, acc);
    // D fragment: lane holds rows 4 * (lane / 16) + r, column lane % 16 of each 16 x 16 tile; the
    // wavefront writes only its own block of act
    if constexpr (!LAST && !TRAIN) {
#pragma unroll
        for (int j = 0; j < CT; j++) {
            const int col = tl.wc + 16 * j + fr;
            const float bv = bias[col];
#pragma unroll
            for (int i = 0; i < RT; i++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const bf16_t zb = f2bf(acc[i][j][r] + bv);
                    act[tl.wr + 16 * i + 4 * (lane >> 4) + r][col] = f2bf(gelu(bf2f(zb)));
                }
        }
    } else if constexpr (!LAST) {
        // Z into the wavefront's block of act; then, in 16-byte pieces (row-contiguous, so the
        // global stores coalesce), Z out to HBM, GELU(Z) back into act and out to HBM
#pragma unroll
        for (int j = 0; j < CT; j++) {
            const int col = tl.wc + 16 * j + fr;
            const float bv = bias[col];
#pragma unroll
            for (int i = 0; i < RT; i++)
#pragma unroll
                for (int r = 0; r < 4; r++) act[tl.wr + 16 * i + 4 * (lane >> 4) + r][col] = f2bf(acc[i][j][r] + bv);
        }
        constexpr int PR = CT * 2;   // 16-byte pieces per row of the block
        static_assert(RT * 16 * PR == 8 * 64, "8 pieces per lane");
#pragma unroll
        for (int t = 0; t < 8; t++) {
            const int c = lane + 64 * t, r = c / PR, cc = tl.wc + (c % PR) * 8;
            uint4 zv = *reinterpret_cast<const uint4*>(&act[tl.wr + r][cc]);
            const unsigned zw[4] = {zv.x, zv.y, zv.z, zv.w};
            unsigned gw[4];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const bf16_t lo = f2bf(gelu(bf2f((bf16_t)(zw[q] & 0xffffu))));
                const bf16_t hi = f2bf(gelu(bf2f((bf16_t)(zw[q] >> 16))));
                gw[q] = (unsigned)lo | ((unsigned)hi << 16);
            }
            const uint4 gv = make_uint4(gw[0], gw[1], gw[2], gw[3]);
            *reinterpret_cast<uint4*>(&act[tl.wr + r][cc]) = gv;
            const long long row = row0 + tl.wr + r;
            if (row < n) {
                *reinterpret_cast<uint4*>(tr.z + (size_t)row * N + cc) = zv;
                *reinterpret_cast<uint4*>(tr.a + (size_t)row * N + cc) = gv;
            }
        }
    } else {
        static_assert(RT == 1 && CT == 1, "the output layer: one tile per wavefront");
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const long long row = row0 + tl.wr + 4 * (lane >> 4) + r;
            if (row < n) Y[row * N + fr] = acc[0][0][r] + bias[fr];
        }
    }
}

// count (device, may be null): rows beyond *count are skipped, whole tiles at once. TRAIN: the
// training forward pass, which also writes each hidden layer's Z and GELU(Z) (tr[l], l = 0..3).
template <bool TRAIN>
__global__ void __launch_bounds__(FT) k_mlp_fused(MlpView v, int n_max, const unsigned* __restrict__ count,
                                                  const bf16_t* __restrict__ X, float* __restrict__ Y,
                                                  TrainOut tr0 = {}, TrainOut tr1 = {}, TrainOut tr2 = {},
                                                  TrainOut tr3 = {}) {
    __shared__ bf16_t act[FM][FPITCH];
    __shared__ bf16_t ws[2][256][WPITCH];
    const int tid = threadIdx.x, lane = tid & 63;
    const int n = count ? min(n_max, (int)*count) : n_max;
    // persistent over the row tiles: the grid is sized for n_max, the work by the device count
    for (long long row0 = (long long)blockIdx.x * FM; row0 < n; row0 += (long long)gridDim.x * FM) {
        __syncthreads();   // the previous tile's last layer is done with act
        for (int c = tid; c < FM * 4; c += FT) {   // input features: FM rows x 32 BF16
            const int r = c >> 2, kc = (c & 3) * 8;
            uint4 val = make_uint4(0, 0, 0, 0);
            if (row0 + r < n) val = *reinterpret_cast<const uint4*>(X + (size_t)(row0 + r) * 32 + kc);
            *reinterpret_cast<uint4*>(&act[r][kc]) = val;
        }
        fused_layer<32, 256, false, TRAIN>(act, ws, v.wt[0], v.b[0], tid, lane, row0, n, Y, tr0);
        fused_layer<256, 256, false, TRAIN>(act, ws, v.wt[1], v.b[1], tid, lane, row0, n, Y, tr1);
        fused_layer<256, 256, false, TRAIN>(act, ws, v.wt[2], v.b[2], tid, lane, row0, n, Y, tr2);
        fused_layer<256, 256, false, TRAIN>(act, ws, v.wt[3], v.b[3], tid, lane, row0, n, Y, tr3);
        fused_layer<256, 16, true>(act, ws, v.wt[4], v.b[4], tid, lane, row0, n, Y);
    }
}

// One wavefront's 16 rows x 256 columns between HBM and LDS in 16-byte pieces (8 per lane),
// rows at or beyond n skipped (loads give zeros).
// NCOL 16: rows of 16 columns, padded with zeros to the 32 of one K chunk.
template <int NCOL = 256>
__device__ __forceinline__ void rows_in(bf16_t (*act)[FPITCH], const bf16_t* __restrict__ g, long long row0, int wr,
                                        int lane, int n) {
    static_assert(NCOL == 256 || NCOL == 16, "row widths of the network");
    if constexpr (NCOL == 256) {
#pragma unroll
        for (int t = 0; t < 8; t++) {
            const int c = lane + 64 * t, r = c >> 5, cc = (c & 31) * 8;
            const long long row = row0 + wr + r;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (row < n) v = *reinterpret_cast<const uint4*>(g + (size_t)row * 256 + cc);
            *reinterpret_cast<uint4*>(&act[wr + r][cc]) = v;
        }
    } else {
        const int r = lane >> 2, cc = (lane & 3) * 8;
        const long long row = row0 + wr + r;
        uint4 v = make_uint4(0, 0, 0, 0);
        if (row < n && cc < 16) v = *reinterpret_cast<const uint4*>(g + (size_t)row * 16 + cc);
        *reinterpret_cast<uint4*>(&act[wr + r][cc]) = v;
    }
}
__device__ __forceinline__ void rows_out(const bf16_t (*act)[FPITCH], bf16_t* __restrict__ g, long long row0, int wr,
                                         int lane, int n) {
#pragma unroll
    for (int t = 0; t < 8; t++) {
        const int c = lane + 64 * t, r = c >> 5, cc = (c & 31) * 8;
        const long long row = row0 + wr + r;
        if (row < n) *reinterpret_cast<uint4*>(g + (size_t)row * 256 + cc) = *reinterpret_cast<const uint4*>(&act[wr + r][cc]);
    }
}

// Backward through one hidden layer: dZ_{l-1} = (dZ_l W_l) .* GELU'(Z_{l-1}) for a tile of FM rows,
// all [n x 256] BF16. Each wavefront stages its own 16 rows of dZ_l in LDS, the weights (W_l as
// [in x out], K-contiguous rows) stream through the forward kernel's double-buffered stage, then
// the wavefront's rows of Z_{l-1} replace its dZ_l rows in LDS for the GELU' epilogue, whose
// result leaves with 16-byte stores. Same arithmetic as the GEMM's DGELU epilogue.
// KIN: the width of dZ_l, 256 for the hidden layers, 16 for the output layer (its W_l rows are 16
// long; the MFMA chunk of 32 is padded with zeros).
template <int KIN>
__global__ void __launch_bounds__(FT) k_mlp_dx(const bf16_t* __restrict__ dz, const bf16_t* __restrict__ wn,
                                               const bf16_t* __restrict__ Z, bf16_t* __restrict__ out, int n) {
    __shared__ bf16_t act[FM][FPITCH];
    __shared__ bf16_t ws[2][256][WPITCH];
    // the row slices (16 rows per wavefront) move rows between HBM and LDS; the MFMA blocks
    // (WaveTile) span two slices, so the passes between them are separated by barriers
    const int tid = threadIdx.x, lane = tid & 63, wr = (tid >> 6) * 16, fr = lane & 15;
    using TL = WaveTile<256>;
    const TL tl(tid);
    for (long long row0 = (long long)blockIdx.x * FM; row0 < n; row0 += (long long)gridDim.x * FM) {
        rows_in<KIN>(act, dz, row0, wr, lane, n);
        // the wavefront's slice of Z, loaded now and held in registers through the MFMAs
        uint4 zr[8];
#pragma unroll
        for (int t = 0; t < 8; t++) {
            const int c = lane + 64 * t, r = c >> 5, cc = (c & 31) * 8;
            const long long row = row0 + wr + r;
            zr[t] = row < n ? *reinterpret_cast<const uint4*>(Z + (size_t)row * 256 + cc) : make_uint4(0, 0, 0, 0);
        }
        f32x4_t acc[TL::RT][TL::CT];
        fused_mfma<(KIN < 32 ? 32 : KIN), 256, KIN>(act, ws, wn, tid, lane, tl, acc);   // ends with a barrier
#pragma unroll
        for (int t = 0; t < 8; t++) {
            const int c = lane + 64 * t, r = c >> 5, cc = (c & 31) * 8;
            *reinterpret_cast<uint4*>(&act[wr + r][cc]) = zr[t];
        }
        __syncthreads();   // every slice of Z staged
#pragma unroll
        for (int j = 0; j < TL::CT; j++) {
            const int col = tl.wc + 16 * j + fr;
#pragma unroll
            for (int i = 0; i < TL::RT; i++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    bf16_t& e = act[tl.wr + 16 * i + 4 * (lane >> 4) + r][col];
                    e = f2bf(acc[i][j][r] * dgelu(bf2f(e)));
                }
        }
        __syncthreads();   // every block of the result written
        rows_out(act, out, row0, wr, lane, n);
        __syncthreads();   // the next tile's rows_in overwrites act
    }
}

}  // namespace

// ------------------------------------------------------------------ MLP object
struct eu3d_mlp {
    int nl = 5;                   // linear layers
    int din[5], dout[5];          // padded
    int rin[5], rout[5];          // real
    float *w[5], *b[5];           // FP32 masters: w[l] is [dout x din] (Wt), b[l] [dout]
    float *mw[5], *vw[5], *mb[5], *vb[5];
    float *gw[5], *gb[5];         // gradients
    bf16_t *wt[5], *wn[5];        // BF16 copies: Wt [dout x din], W [din x dout]
    int cap = 0;                  // workspace rows
    bf16_t* act[5] = {};          // act[0] = input [cap x 32], act[l] = output of layer l (l=1..4)
    bf16_t* zs[5] = {};           // pre-activations of layers 1..4
    bf16_t *dz_a = nullptr, *dz_b = nullptr;
    float* yout = nullptr;        // [cap x 16]
    int num_cu = 256;             // compute units of the device (grid sizes of the list kernels)
    int icap = 0;                 // inference rows (features and outputs only)
    bf16_t* ifeat = nullptr;      // [icap x 32]
    float* iout = nullptr;        // [icap x 16]
    float* part = nullptr;        // split-K / reduction partials
    size_t part_cap = 0;
    int adam_t = 0;
    hipStream_t stream = nullptr;
    eu3d_surrogate_params sp{};
    bool sp_set = false;
    std::string err;
};

namespace {

constexpr int NIN_REAL = 12, NIN = 32, NH = 256, NOUT_REAL = 9, NOUT = 16;

template <bool AKC, bool BKC, int EPI>
hipError_t gemm(hipStream_t s, const GemmArgs& a0, int splits = 1) {
    GemmArgs a = a0;
    a.kchunk = ((a.K + splits - 1) / splits + BK - 1) / BK * BK;
    const int sp = (a.K + a.kchunk - 1) / a.kchunk;
    dim3 grid((a.M + BM - 1) / BM, (a.N + BN - 1) / BN, sp);
    k_gemm<AKC, BKC, EPI><<<grid, 256, 0, s>>>(a);
    return hipGetLastError();
}

// split-K count of the weight-gradient GEMMs (M = N = 256, K = rows): enough K chunks for about
// eight 128 x 128 tiles per CU at 2^20 rows
int splits_for(int K) { return std::max(1, std::min(512, K / 2048)); }

hipError_t ensure_ws(eu3d_mlp* m, int n) {
    if (n <= m->cap) return hipSuccess;
    for (int l = 0; l < 5; l++) {
        if (m->act[l]) (void)hipFree(m->act[l]);
        if (m->zs[l]) (void)hipFree(m->zs[l]);
        m->act[l] = m->zs[l] = nullptr;
    }
    if (m->dz_a) (void)hipFree(m->dz_a);
    if (m->dz_b) (void)hipFree(m->dz_b);
    if (m->yout) (void)hipFree(m->yout);
    hipError_t e;
    if ((e = hipMalloc(&m->act[0], (size_t)n * NIN * 2)) != hipSuccess) return e;
    for (int l = 1; l < 5; l++) {
        if ((e = hipMalloc(&m->act[l], (size_t)n * NH * 2)) != hipSuccess) return e;
        if ((e = hipMalloc(&m->zs[l], (size_t)n * NH * 2)) != hipSuccess) return e;
    }
    if ((e = hipMalloc(&m->dz_a, (size_t)n * NH * 2)) != hipSuccess) return e;
    if ((e = hipMalloc(&m->dz_b, (size_t)n * NH * 2)) != hipSuccess) return e;
    if ((e = hipMalloc(&m->yout, (size_t)n * NOUT * 4)) != hipSuccess) return e;
    m->cap = n;
    return hipSuccess;
}

// Inference needs only the features and the outputs of each row (the fused kernel keeps the
// activations on chip): 128 bytes per row.
hipError_t ensure_infer(eu3d_mlp* m, int n) {
    if (n <= m->icap) return hipSuccess;
    if (m->ifeat) (void)hipFree(m->ifeat);
    if (m->iout) (void)hipFree(m->iout);
    m->ifeat = nullptr;
    m->iout = nullptr;
    m->icap = 0;
    hipError_t e;
    if ((e = hipMalloc(&m->ifeat, (size_t)n * NIN * 2)) != hipSuccess) return e;
    if ((e = hipMalloc(&m->iout, (size_t)n * NOUT * 4)) != hipSuccess) return e;
    m->icap = n;
    return hipSuccess;
}

hipError_t ensure_part(eu3d_mlp* m, size_t nfloats) {
    if (nfloats <= m->part_cap) return hipSuccess;
    if (m->part) (void)hipFree(m->part);
    m->part = nullptr;
    hipError_t e = hipMalloc(&m->part, nfloats * 4);
    if (e == hipSuccess) m->part_cap = nfloats;
    return e;
}

// Forward through all layers from act[0] (n rows); output FP32 [n x 16] in m->yout.
hipError_t forward(eu3d_mlp* m, int n, bool save_z, hipStream_t st = nullptr) {
    hipError_t e;
    hipStream_t s0 = st ? st : m->stream;
    for (int l = 0; l < 4; l++) {
        GemmArgs a{};
        a.M = n; a.N = m->dout[l]; a.K = m->din[l];
        a.A = m->act[l]; a.lda = m->din[l];
        a.B = m->wt[l]; a.ldb = m->din[l];
        a.Cb = m->act[l + 1]; a.Z = save_z ? m->zs[l + 1] : nullptr; a.bias = m->b[l]; a.ldc = m->dout[l];
        if ((e = gemm<true, true, EPI_BIAS_GELU>(s0, a)) != hipSuccess) return e;
    }
    GemmArgs a{};
    a.M = n; a.N = m->dout[4]; a.K = m->din[4];
    a.A = m->act[4]; a.lda = m->din[4];
    a.B = m->wt[4]; a.ldb = m->din[4];
    a.Cf = m->yout; a.bias = m->b[4]; a.ldc = m->dout[4];
    return gemm<true, true, EPI_BIAS_F32>(s0, a);
}

// Inference through the fused kernel: features (BF16 [n x 32]) -> outputs (FP32 [n x 16]).
hipError_t forward_fused(eu3d_mlp* m, int n, const bf16_t* X, float* Y, hipStream_t st, const unsigned* count = nullptr) {
    MlpView v;
    for (int l = 0; l < 5; l++) {
        v.wt[l] = m->wt[l];
        v.b[l] = m->b[l];
    }
    // one 8-wave block per CU fits its LDS; a device count means a grid-stride over the tiles
    const unsigned tiles = (unsigned)((n + FM - 1) / FM);
    const unsigned nb = count ? std::min<unsigned>(tiles, (unsigned)m->num_cu) : tiles;
    k_mlp_fused<false><<<nb, FT, 0, st>>>(v, n, count, X, Y);
    return hipGetLastError();
}

// The training forward pass through the fused kernel: outputs into yout, and each hidden layer's
// Z and GELU(Z) into zs[l + 1] / act[l + 1] for the backward pass.
hipError_t forward_fused_train(eu3d_mlp* m, int n) {
    MlpView v;
    for (int l = 0; l < 5; l++) {
        v.wt[l] = m->wt[l];
        v.b[l] = m->b[l];
    }
    TrainOut t[4];
    for (int l = 0; l < 4; l++) t[l] = TrainOut{m->zs[l + 1], m->act[l + 1]};
    const unsigned tiles = (unsigned)((n + FM - 1) / FM);
    k_mlp_fused<true><<<tiles, FT, 0, m->stream>>>(v, n, nullptr, m->act[0], m->yout, t[0], t[1], t[2], t[3]);
    return hipGetLastError();
}

}  // namespace

namespace {
using namespace eu3d;

constexpr int NEL = 3;   // H, O, N atoms per species (mechanism order H, O, H2O, OH, O2, H2, H2O2, HO2, N2)
__device__ __forceinline__ int atoms(int e, int s) {
    const int H[NS] = {1, 0, 2, 1, 0, 2, 2, 1, 0}, O[NS] = {0, 1, 1, 1, 2, 0, 2, 2, 0}, N[NS] = {0, 0, 0, 0, 0, 0, 0, 0, 2};
    return e == 0 ? H[s] : (e == 1 ? O[s] : N[s]);
}

// Projection of a mass-fraction increment onto the increments that conserve the H, O and N
// atoms, in the metric weighted by w: dY -= W A^T (A W A^T)^-1 A dY with A[e][s] = atoms(e, s) /
// Mw_s and W = diag(w) (w = 1: the orthogonal projection). The network's raw increment conserves
// the elements only to its training error; the projected one conserves them to rounding. With
// w_s the species' abundance, each species is corrected in proportion to its own amount, so a
// trace species is not driven negative by the correction of a major one.
__device__ __forceinline__ void project_elements(double (&dY)[NS], const double (&w)[NS]) {
    double M[NEL][NEL], b[NEL];
#pragma unroll
    for (int e = 0; e < NEL; e++) {
        b[e] = 0.0;
#pragma unroll
        for (int s = 0; s < NS; s++) b[e] += atoms(e, s) * RMW(s) * dY[s];
#pragma unroll
        for (int f = 0; f < NEL; f++) {
            double m = 0.0;
#pragma unroll
            for (int s = 0; s < NS; s++) m += (atoms(e, s) * RMW(s)) * w[s] * (atoms(f, s) * RMW(s));
            M[e][f] = m;
        }
    }
    // 3 x 3 symmetric solve M lam = b by cofactors
    const double c00 = M[1][1] * M[2][2] - M[1][2] * M[2][1], c01 = M[1][2] * M[2][0] - M[1][0] * M[2][2],
                 c02 = M[1][0] * M[2][1] - M[1][1] * M[2][0];
    const double det = M[0][0] * c00 + M[0][1] * c01 + M[0][2] * c02;
    const double c10 = M[0][2] * M[2][1] - M[0][1] * M[2][2], c11 = M[0][0] * M[2][2] - M[0][2] * M[2][0],
                 c12 = M[0][1] * M[2][0] - M[0][0] * M[2][1];
    const double c20 = M[0][1] * M[1][2] - M[0][2] * M[1][1], c21 = M[0][2] * M[1][0] - M[0][0] * M[1][2],
                 c22 = M[0][0] * M[1][1] - M[0][1] * M[1][0];
    const double rd = 1.0 / det;
    const double lam[NEL] = {(c00 * b[0] + c10 * b[1] + c20 * b[2]) * rd, (c01 * b[0] + c11 * b[1] + c21 * b[2]) * rd,
                             (c02 * b[0] + c12 * b[1] + c22 * b[2]) * rd};
#pragma unroll
    for (int s = 0; s < NS; s++) {
        double corr = 0.0;
#pragma unroll
        for (int e = 0; e < NEL; e++) corr += atoms(e, s) * RMW(s) * lam[e];
        dY[s] -= w[s] * corr;
    }
}

// Surrogate features of one cell: T, ln(p / 1 atm), ln(max(Y_k, 1e-20)), ln(dt), normalised.
__device__ __forceinline__ bool sur_features(const eu3d_surrogate_params& sp, const double (&Y)[NS], double T, double D,
                                             double dt, float (&f)[NIN_REAL]) {
    double sy = 0;
#pragma unroll
    for (int s = 0; s < NS; s++) sy += Y[s] / MW[s];
    const double p = T * D * (R_UNIV * 1000 * sy);
    double raw[NIN_REAL];
    raw[0] = T;
    raw[1] = log(p / 101325.0);
#pragma unroll
    for (int s = 0; s < NS; s++) raw[2 + s] = log(fmax(Y[s], 1e-20));
    raw[11] = log(dt);
    bool in = true;
#pragma unroll
    for (int i = 0; i < NIN_REAL; i++) {
        f[i] = (float)((raw[i] - sp.feat_mean[i]) / sp.feat_std[i]);
        in = in && f[i] >= sp.hull_lo[i] && f[i] <= sp.hull_hi[i];
    }
    return in;
}

__device__ __forceinline__ void store_feats(bf16_t* out, size_t r, const float (&f)[NIN_REAL]) {
    bf16_t* o = out + r * NIN;
#pragma unroll
    for (int i = 0; i < NIN; i++) o[i] = i < NIN_REAL ? f2bf(f[i]) : (bf16_t)0;
}

// Batch cells at rest: features of (rho_s, T) with interval dt. raw != null also receives the
// un-normalised features in FP32 (for the normalisation statistics).
__global__ void k_sur_feat_batch(eu3d_surrogate_params sp, long long n, const double* __restrict__ rho,
                                 const double* __restrict__ T, double dt, bf16_t* __restrict__ feats) {
    const long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (id >= n) return;
    double Y[NS], D = 0;
#pragma unroll
    for (int s = 0; s < NS; s++) D += rho[(size_t)s * n + id];
#pragma unroll
    for (int s = 0; s < NS; s++) Y[s] = rho[(size_t)s * n + id] / D;
    float f[NIN_REAL];
    sur_features(sp, Y, T[id], D, dt, f);
    store_feats(feats, (size_t)id, f);
}

// Training targets: normalised mass-fraction increments between two batch states.
__global__ void k_sur_targets(eu3d_surrogate_params sp, long long n, const double* __restrict__ rho0,
                              const double* __restrict__ rho1, float* __restrict__ tgt) {
    const long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (id >= n) return;
    double D0 = 0, D1 = 0;
#pragma unroll
    for (int s = 0; s < NS; s++) {
        D0 += rho0[(size_t)s * n + id];
        D1 += rho1[(size_t)s * n + id];
    }
#pragma unroll
    for (int s = 0; s < NOUT; s++) {
        float v = 0.f;
        if (s < NS) {
            const double y0 = rho0[(size_t)s * n + id] / D0, y1 = rho1[(size_t)s * n + id] / D1;
            // (the FP64 chemistry leaves some mass fractions a rounding error below zero)
            const double dy = sp.out_eps > 0 ? log(fmax(y1, 0.0) + sp.out_eps) - log(fmax(y0, 0.0) + sp.out_eps)
                                             : y1 - y0;
            v = (float)((dy - sp.out_mean[s]) / sp.out_std[s]);
        }
        tgt[(size_t)id * NOUT + s] = v;
    }
}

// Apply the network's increment to one cell and run the acceptance checks: training hull,
// mass fractions not below -y_neg_tol, H/O/N element conservation within elem_tol of the
// cell's element content, and |dT| <= dT_max after the energy Newton at constant internal
// energy. On acceptance returns the new primitive state and the new partial densities in cs.
struct SurCell {
    double Di[NS];
    double U, V, W, E, T;
};

__device__ __forceinline__ int sur_apply(const eu3d_surrogate_params& sp, const SurCell& c, const float* out, double dt,
                                         uint32_t q, double (&cs)[NC], Prim& o) {
    double D = 0, Y[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) D += c.Di[s];
#pragma unroll
    for (int s = 0; s < NS; s++) Y[s] = c.Di[s] / D;
    float f[NIN_REAL];
    if (!sur_features(sp, Y, c.T, D, dt, f)) return 1;
    double dY[NS], Yn[NS], w[NS], ymin = 1.0;
    if (sp.out_eps > 0) {
        // log increments: the unprojected new mass fractions are above -eps by construction
#pragma unroll
        for (int s = 0; s < NS; s++) {
            const double yn = (fmax(Y[s], 0.0) + sp.out_eps) * exp((double)out[s] * sp.out_std[s] + sp.out_mean[s]);
            dY[s] = (yn - sp.out_eps) - Y[s];
            w[s] = yn;
        }
    } else {
#pragma unroll
        for (int s = 0; s < NS; s++) {
            dY[s] = (double)out[s] * sp.out_std[s] + sp.out_mean[s];
            w[s] = 1.0;
        }
    }
    project_elements(dY, w);
#pragma unroll
    for (int s = 0; s < NS; s++) {
        Yn[s] = Y[s] + dY[s];
        ymin = fmin(ymin, Yn[s]);
    }
    if (!(ymin >= -sp.y_neg_tol)) return 2;
#pragma unroll
    for (int e = 0; e < NEL; e++) {
        double have = 0, change = 0;
#pragma unroll
        for (int s = 0; s < NS; s++) {
            const int a = atoms(e, s);
            if (a) {
                have += a * Y[s] / MW[s];
                change += a * dY[s] / MW[s];
            }
        }
        if (have > 0 && !(fabs(change) <= sp.elem_tol * have)) return 3;
    }
    double sum = 0;
#pragma unroll
    for (int s = 0; s < NS; s++) {
        Yn[s] = fmax(Yn[s], 0.0);
        sum += Yn[s];
    }
#pragma unroll
    for (int s = 0; s < NS; s++) cs[s] = D * (Yn[s] / sum);
    double Dn = 0;
#pragma unroll
    for (int s = 0; s < NS; s++) Dn += cs[s];
    cs[NS] = c.U * Dn;
    cs[NS + 1] = c.V * Dn;
    cs[NS + 2] = c.W * Dn;
    cs[NS + 3] = Dn * c.E + 0.5 * (c.U * c.U + c.V * c.V + c.W * c.W) * Dn;
    o = prim_from_cs(cs, c.T, c.E, q);
    if (!(fabs(o.T - c.T) <= sp.dT_max)) return 4;
    return 0;
}

__device__ __forceinline__ void sur_count(unsigned* cnt, unsigned* list, unsigned id, int why) {
    const bool fb = why != 0;
    const unsigned long long m = __ballot(fb);
    const int lane = __lane_id();
    const int leader = __ffsll((unsigned long long)__ballot(1)) - 1;
    unsigned base = 0;
    if (lane == leader && m) base = atomicAdd(&cnt[0], (unsigned)__popcll(m));
    base = __shfl(base, leader);
    if (fb) {
        list[base + __popcll(m & ((1ull << lane) - 1ull))] = id;
        atomicAdd(&cnt[why], 1u);
    }
}

// Batch apply: cells at rest with internal energy E; cells colder than T_cold keep their state.
__global__ void __launch_bounds__(256) k_sur_apply_batch(eu3d_surrogate_params sp, long long n, double* rho, double* T,
                                                         double* E, const float* __restrict__ out, double dt, uint32_t q,
                                                         unsigned* cnt, unsigned* list) {
    const long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    int why = 0;
    if (id < n && T[id] >= sp.T_cold) {
        SurCell c;
#pragma unroll
        for (int s = 0; s < NS; s++) c.Di[s] = rho[(size_t)s * n + id];
        c.U = c.V = c.W = 0.0;
        c.E = E[id];
        c.T = T[id];
        double cs[NC];
        Prim o;
        why = sur_apply(sp, c, out + (size_t)id * NOUT, dt, q, cs, o);
        if (why == 0) {
#pragma unroll
            for (int s = 0; s < NS; s++) rho[(size_t)s * n + id] = cs[s];
            T[id] = o.T;
            E[id] = o.E;
        }
    }
    sur_count(cnt, list, (unsigned)id, why);
}

// Grid: interior cell (i, j, k) is feature row ((k - bc) * nj + (j - bc)) * ni + (i - bc).
__global__ void __launch_bounds__(256) k_sur_feat_grid(eu3d_surrogate_params sp, G3 g, Fields F, const Scalars* sc,
                                                       bf16_t* __restrict__ feats) {
    const int i = g.bc + blockIdx.x * 64 + threadIdx.x, j = g.bc + blockIdx.y * 4 + threadIdx.y, k = g.bc + blockIdx.z;
    if (i >= g.ni + g.bc || j >= g.nj + g.bc) return;
    const size_t c = IDX(g, i, j, k);
    const size_t r = ((size_t)(k - g.bc) * g.nj + (j - g.bc)) * g.ni + (i - g.bc);
    double Y[NS], D = 0;
#pragma unroll
    for (int s = 0; s < NS; s++) D += FLD(F, EU3D_FIELD_CS)[s * g.n3 + c];
#pragma unroll
    for (int s = 0; s < NS; s++) Y[s] = FLD(F, EU3D_FIELD_CS)[s * g.n3 + c] / D;
    float f[NIN_REAL];
    sur_features(sp, Y, FLD(F, EU3D_FIELD_T)[c], D, sc->dt, f);
    store_feats(feats, r, f);
}

// Apply the network's output row to grid cell c with the checks; on acceptance write the cell's
// conserved state and everything Update_after_CS leaves. Returns the rejection code (0 accepted).
__device__ __forceinline__ int sur_apply_grid_cell(const eu3d_surrogate_params& sp, const G3& g, const Fields& F, double dt,
                                                   uint32_t q, size_t c, const float* out) {
    SurCell cl;
#pragma unroll
    for (int s = 0; s < NS; s++) cl.Di[s] = FLD(F, EU3D_FIELD_CS)[s * g.n3 + c];
    cl.U = FLD(F, EU3D_FIELD_U)[c];
    cl.V = FLD(F, EU3D_FIELD_V)[c];
    cl.W = FLD(F, EU3D_FIELD_W)[c];
    cl.E = FLD(F, EU3D_FIELD_E)[c];
    cl.T = FLD(F, EU3D_FIELD_T)[c];
    double cs[NC];
    Prim o;
    const int why = sur_apply(sp, cl, out, dt, q, cs, o);
    if (why == 0) {
        double* CS = FLD(F, EU3D_FIELD_CS);
#pragma unroll
        for (int s = 0; s < NC; s++) CS[s * g.n3 + c] = cs[s];
        FLD(F, EU3D_FIELD_D)[c] = o.D;
        FLD(F, EU3D_FIELD_T)[c] = o.T;
        FLD(F, EU3D_FIELD_U)[c] = o.U;
        FLD(F, EU3D_FIELD_V)[c] = o.V;
        FLD(F, EU3D_FIELD_W)[c] = o.W;
        FLD(F, EU3D_FIELD_E)[c] = o.E;
        FLD(F, EU3D_FIELD_P)[c] = o.P;
        FLD(F, EU3D_FIELD_H)[c] = o.H;
        FLD(F, EU3D_FIELD_GAMMA)[c] = o.Gamma;
        FLD(F, EU3D_FIELD_C)[c] = o.C;
#pragma unroll
        for (int s = 0; s < NS; s++) {
            FLD(F, EU3D_FIELD_DI)[s * g.n3 + c] = cs[s];
            FLD(F, EU3D_FIELD_MC)[s * g.n3 + c] = cs[s] * RMW(s) * 1000;
            FLD(F, EU3D_FIELD_YI)[s * g.n3 + c] = o.Y[s];
        }
    }
    return why;
}

__global__ void __launch_bounds__(256) k_sur_apply_grid(eu3d_surrogate_params sp, G3 g, Fields F, const Scalars* sc,
                                                        const float* __restrict__ out, uint32_t q, unsigned* cnt,
                                                        unsigned* list) {
    const int i = g.bc + blockIdx.x * 64 + threadIdx.x, j = g.bc + blockIdx.y * 4 + threadIdx.y, k = g.bc + blockIdx.z;
    const bool in = i < g.ni + g.bc && j < g.nj + g.bc;
    const size_t c = in ? IDX(g, i, j, k) : 0;
    int why = 0;
    if (in && FLD(F, EU3D_FIELD_T)[c] >= sp.T_cold) {
        const size_t r = ((size_t)(k - g.bc) * g.nj + (j - g.bc)) * g.ni + (i - g.bc);
        why = sur_apply_grid_cell(sp, g, F, sc->dt, q, c, out + r * NOUT);
    }
    sur_count(cnt, list, (unsigned)c, why);
}

// ---- network lists (surrogate chemistry on the cells whose sub-step count exceeds the
// threshold): row r of the features / outputs is cell list[r], r < *count.
__global__ void __launch_bounds__(256) k_sur_feat_batch_list(eu3d_surrogate_params sp, long long n, const double* __restrict__ rho,
                                                             const double* __restrict__ T, double dt,
                                                             const unsigned* __restrict__ list, const unsigned* __restrict__ count,
                                                             bf16_t* __restrict__ feats) {
    const unsigned cnt = *count;
    for (unsigned r = blockIdx.x * blockDim.x + threadIdx.x; r < cnt; r += gridDim.x * blockDim.x) {
        const size_t id = list[r];
        double Y[NS], D = 0;
#pragma unroll
        for (int s = 0; s < NS; s++) D += rho[(size_t)s * n + id];
#pragma unroll
        for (int s = 0; s < NS; s++) Y[s] = rho[(size_t)s * n + id] / D;
        float f[NIN_REAL];
        sur_features(sp, Y, T[id], D, dt, f);
        store_feats(feats, r, f);
    }
}

__global__ void __launch_bounds__(256) k_sur_apply_batch_list(eu3d_surrogate_params sp, long long n, double* rho, double* T,
                                                              double* E, const float* __restrict__ out, double dt, uint32_t q,
                                                              const unsigned* __restrict__ list, const unsigned* __restrict__ count,
                                                              unsigned* cnt_out, unsigned* rej) {
    const unsigned cnt = *count;
    for (unsigned b0 = blockIdx.x * blockDim.x; b0 < cnt; b0 += gridDim.x * blockDim.x) {
        const unsigned r = b0 + threadIdx.x;
        const bool in = r < cnt;
        const unsigned id = in ? list[r] : 0;
        int why = 0;
        if (in) {
            SurCell c;
#pragma unroll
            for (int s = 0; s < NS; s++) c.Di[s] = rho[(size_t)s * n + id];
            c.U = c.V = c.W = 0.0;
            c.E = E[id];
            c.T = T[id];
            double cs[NC];
            Prim o;
            why = sur_apply(sp, c, out + (size_t)r * NOUT, dt, q, cs, o);
            if (why == 0) {
#pragma unroll
                for (int s = 0; s < NS; s++) rho[(size_t)s * n + id] = cs[s];
                T[id] = o.T;
                E[id] = o.E;
            }
        }
        sur_count(cnt_out, rej, id, why);
    }
}

__global__ void __launch_bounds__(256) k_sur_feat_grid_list(eu3d_surrogate_params sp, G3 g, Fields F, const Scalars* sc,
                                                            const unsigned* __restrict__ list, const unsigned* __restrict__ count,
                                                            bf16_t* __restrict__ feats) {
    const unsigned cnt = *count;
    for (unsigned r = blockIdx.x * blockDim.x + threadIdx.x; r < cnt; r += gridDim.x * blockDim.x) {
        const size_t c = list[r];
        double Y[NS], D = 0;
#pragma unroll
        for (int s = 0; s < NS; s++) D += FLD(F, EU3D_FIELD_CS)[s * g.n3 + c];
#pragma unroll
        for (int s = 0; s < NS; s++) Y[s] = FLD(F, EU3D_FIELD_CS)[s * g.n3 + c] / D;
        float f[NIN_REAL];
        sur_features(sp, Y, FLD(F, EU3D_FIELD_T)[c], D, sc->dt, f);
        store_feats(feats, r, f);
    }
}

__global__ void __launch_bounds__(256) k_sur_apply_grid_list(eu3d_surrogate_params sp, G3 g, Fields F, const Scalars* sc,
                                                             const float* __restrict__ out, uint32_t q,
                                                             const unsigned* __restrict__ list, const unsigned* __restrict__ count,
                                                             unsigned* cnt_out, unsigned* rej) {
    const unsigned cnt = *count;
    for (unsigned b0 = blockIdx.x * blockDim.x; b0 < cnt; b0 += gridDim.x * blockDim.x) {
        const unsigned r = b0 + threadIdx.x;
        const bool in = r < cnt;
        const size_t c = in ? list[r] : 0;
        const int why = in ? sur_apply_grid_cell(sp, g, F, sc->dt, q, c, out + (size_t)r * NOUT) : 0;
        sur_count(cnt_out, rej, (unsigned)c, why);
    }
}

__global__ void k_zero_u32(unsigned* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

}  // namespace

extern "C" {

int eu3d_mlp_create(eu3d_mlp** out, uint64_t seed) {
    if (!out) return EU3D_ERR_ARG;
    eu3d_mlp* m = new eu3d_mlp();
    const int din[5] = {NIN, NH, NH, NH, NH}, dout[5] = {NH, NH, NH, NH, NOUT};
    const int rin[5] = {NIN_REAL, NH, NH, NH, NH}, rout[5] = {NH, NH, NH, NH, NOUT_REAL};
    // blocking stream: ordered after the caller's null-stream work on the buffers it passes in
    if (hipStreamCreateWithFlags(&m->stream, hipStreamDefault) != hipSuccess) { delete m; return EU3D_ERR_HIP; }
    {
        int dev = 0, ncu = 0;
        if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
            ncu > 0)
            m->num_cu = ncu;
    }
    uint64_t st = seed * 6364136223846793005ull + 1442695040888963407ull;
    auto rnd = [&]() {   // uniform (-1, 1) from a 64-bit LCG
        st = st * 6364136223846793005ull + 1442695040888963407ull;
        return ((double)(st >> 11) / (double)(1ull << 53)) * 2.0 - 1.0;
    };
    for (int l = 0; l < 5; l++) {
        m->din[l] = din[l]; m->dout[l] = dout[l]; m->rin[l] = rin[l]; m->rout[l] = rout[l];
        const size_t nw = (size_t)dout[l] * din[l];
        std::vector<float> hw(nw, 0.f), hb(dout[l], 0.f);
        const double lim = std::sqrt(6.0 / (rin[l] + rout[l]));   // Glorot uniform
        for (int o = 0; o < rout[l]; o++)
            for (int k = 0; k < rin[l]; k++) hw[(size_t)o * din[l] + k] = (float)(lim * rnd());
        float** bufs[] = {&m->w[l], &m->mw[l], &m->vw[l], &m->gw[l]};
        for (float** p : bufs) {
            if (hipMalloc(p, nw * 4) != hipSuccess) return EU3D_ERR_HIP;
            (void)hipMemset(*p, 0, nw * 4);
        }
        float** bb[] = {&m->b[l], &m->mb[l], &m->vb[l], &m->gb[l]};
        for (float** p : bb) {
            if (hipMalloc(p, dout[l] * 4) != hipSuccess) return EU3D_ERR_HIP;
            (void)hipMemset(*p, 0, dout[l] * 4);
        }
        if (hipMalloc(&m->wt[l], nw * 2) != hipSuccess || hipMalloc(&m->wn[l], nw * 2) != hipSuccess) return EU3D_ERR_HIP;
        (void)hipMemcpy(m->w[l], hw.data(), nw * 4, hipMemcpyHostToDevice);
        k_cast_weights<<<(unsigned)((nw + 255) / 256), 256, 0, m->stream>>>(m->w[l], dout[l], din[l], m->wt[l], m->wn[l]);
    }
    if (hipStreamSynchronize(m->stream) != hipSuccess) return EU3D_ERR_HIP;
    *out = m;
    return EU3D_OK;
}

int eu3d_mlp_destroy(eu3d_mlp* m) {
    if (!m) return EU3D_ERR_ARG;
    (void)hipStreamSynchronize(m->stream);
    for (int l = 0; l < 5; l++) {
        for (float* p : {m->w[l], m->mw[l], m->vw[l], m->gw[l], m->b[l], m->mb[l], m->vb[l], m->gb[l]})
            if (p) (void)hipFree(p);
        if (m->wt[l]) (void)hipFree(m->wt[l]);
        if (m->wn[l]) (void)hipFree(m->wn[l]);
        if (m->act[l]) (void)hipFree(m->act[l]);
        if (m->zs[l]) (void)hipFree(m->zs[l]);
    }
    for (void* p : {(void*)m->dz_a, (void*)m->dz_b, (void*)m->yout, (void*)m->part, (void*)m->ifeat, (void*)m->iout})
        if (p) (void)hipFree(p);
    (void)hipStreamDestroy(m->stream);
    delete m;
    return EU3D_OK;
}

// Parameters in host FP32, layer by layer: W^T [dout x din] (padded) then bias [dout].
size_t eu3d_mlp_param_count(const eu3d_mlp* m) {
    size_t n = 0;
    for (int l = 0; l < 5; l++) n += (size_t)m->dout[l] * m->din[l] + m->dout[l];
    return n;
}

static int xfer_params(eu3d_mlp* m, float* host, bool to_host, bool grads) {
    size_t off = 0;
    for (int l = 0; l < 5; l++) {
        const size_t nw = (size_t)m->dout[l] * m->din[l];
        float* dw = grads ? m->gw[l] : m->w[l];
        float* db = grads ? m->gb[l] : m->b[l];
        if (to_host) {
            if (hipMemcpy(host + off, dw, nw * 4, hipMemcpyDeviceToHost) != hipSuccess) return EU3D_ERR_HIP;
            if (hipMemcpy(host + off + nw, db, m->dout[l] * 4, hipMemcpyDeviceToHost) != hipSuccess) return EU3D_ERR_HIP;
        } else {
            if (hipMemcpy(dw, host + off, nw * 4, hipMemcpyHostToDevice) != hipSuccess) return EU3D_ERR_HIP;
            if (hipMemcpy(db, host + off + nw, m->dout[l] * 4, hipMemcpyHostToDevice) != hipSuccess) return EU3D_ERR_HIP;
            k_cast_weights<<<(unsigned)((nw + 255) / 256), 256, 0, m->stream>>>(m->w[l], m->dout[l], m->din[l], m->wt[l],
                                                                           m->wn[l]);
        }
        off += nw + m->dout[l];
    }
    return hipStreamSynchronize(m->stream) == hipSuccess ? EU3D_OK : EU3D_ERR_HIP;
}

int eu3d_mlp_get_params(eu3d_mlp* m, float* host) { return m && host ? xfer_params(m, host, true, false) : EU3D_ERR_ARG; }
int eu3d_mlp_set_params(eu3d_mlp* m, const float* host) {
    return m && host ? xfer_params(m, const_cast<float*>(host), false, false) : EU3D_ERR_ARG;
}
int eu3d_mlp_get_grads(eu3d_mlp* m, float* host) { return m && host ? xfer_params(m, host, true, true) : EU3D_ERR_ARG; }

// Inference on n rows of BF16 features (device, [n x 32], zero padded) -> FP32 outputs
// (device, [n x 16]; columns 9..15 are padding).
int eu3d_mlp_forward(eu3d_mlp* m, int n, const void* feats_bf16, float* out_f32) {
    if (!m || n <= 0 || !feats_bf16 || !out_f32) return EU3D_ERR_ARG;
    if (forward_fused(m, n, static_cast<const bf16_t*>(feats_bf16), out_f32, m->stream) != hipSuccess)
        return EU3D_ERR_HIP;
    return hipStreamSynchronize(m->stream) == hipSuccess ? EU3D_OK : EU3D_ERR_HIP;
}

// The same inference layer by layer through the GEMM kernel (the path training uses), with the
// activations in HBM.
int eu3d_mlp_forward_layers(eu3d_mlp* m, int n, const void* feats_bf16, float* out_f32) {
    if (!m || n <= 0 || !feats_bf16 || !out_f32) return EU3D_ERR_ARG;
    if (ensure_ws(m, n) != hipSuccess) return EU3D_ERR_HIP;
    if (hipMemcpyAsync(m->act[0], feats_bf16, (size_t)n * NIN * 2, hipMemcpyDeviceToDevice, m->stream) != hipSuccess)
        return EU3D_ERR_HIP;
    if (forward(m, n, false) != hipSuccess) return EU3D_ERR_HIP;
    if (hipMemcpyAsync(out_f32, m->yout, (size_t)n * NOUT * 4, hipMemcpyDeviceToDevice, m->stream) != hipSuccess)
        return EU3D_ERR_HIP;
    return hipStreamSynchronize(m->stream) == hipSuccess ? EU3D_OK : EU3D_ERR_HIP;
}

// One optimisation step on n rows: BF16 features [n x 32], FP32 targets [n x 16] (device).
// Computes the MSE loss over the 9 real outputs, the gradients of every parameter (kept, see
// eu3d_mlp_get_grads) and, when lr > 0, applies Adam (beta 0.9, 0.999, eps 1e-8) with weight
// decay wd. loss_out receives the mean squared error of the forward pass.
int eu3d_mlp_train_step(eu3d_mlp* m, int n, const void* feats_bf16, const float* targets, float lr, float wd,
                        float* loss_out) {
    if (!m || n <= 0 || !feats_bf16 || !targets) return EU3D_ERR_ARG;
    hipStream_t s = m->stream;
    if (ensure_ws(m, n) != hipSuccess) return EU3D_ERR_HIP;
    // partials: split-K dW [sp x 256 x 256], the bias column sums [nb x 256], then the first-level
    // sums in groups of 32, of the split-K partials [ceil(sp / 32) x 256 x 256] or (later, same
    // space) of the column sums [ceil(nb / 32) x 256]
    const int sp = splits_for(n), rpb = 512, nb = (n + rpb - 1) / rpb, kgrp = 32;
    const size_t ng = (size_t)std::max((sp + kgrp - 1) / kgrp * NH * NH, (nb + kgrp - 1) / kgrp * NH);
    if (ensure_part(m, (size_t)sp * NH * NH + (size_t)nb * NH + ng) != hipSuccess) return EU3D_ERR_HIP;
    if (hipMemcpyAsync(m->act[0], feats_bf16, (size_t)n * NIN * 2, hipMemcpyDeviceToDevice, s) != hipSuccess)
        return EU3D_ERR_HIP;
    if (forward_fused_train(m, n) != hipSuccess) return EU3D_ERR_HIP;
    // loss and dY (bf16, [n x 16]) into dz_a
    const int lb = 1024;
    k_mse_grad<<<lb, 256, 0, s>>>(m->yout, targets, n, NOUT, NOUT_REAL, 2.0f / ((float)n * NOUT_REAL), m->dz_a, m->part);
    std::vector<float> hp(lb);
    if (hipMemcpyAsync(hp.data(), m->part, lb * 4, hipMemcpyDeviceToHost, s) != hipSuccess) return EU3D_ERR_HIP;
    if (hipStreamSynchronize(s) != hipSuccess) return EU3D_ERR_HIP;
    double loss = 0;
    for (float v : hp) loss += v;
    if (loss_out) *loss_out = (float)(loss / ((double)n * NOUT_REAL));
    // backward: dz holds dL/dZ of layer l (bf16 [n x dout[l]])
    bf16_t* dz = m->dz_a;
    bf16_t* dz_next = m->dz_b;
    for (int l = 4; l >= 0; l--) {
        // dW^T [dout x din] = dZ^T [dout x n] * A_l [n x din], split over n
        GemmArgs g{};
        g.M = m->dout[l]; g.N = m->din[l]; g.K = n;
        g.A = dz; g.lda = m->dout[l];          // element (m = o, k = b) at dz[b * dout + o]
        g.B = m->act[l]; g.ldb = m->din[l];    // element (n = i, k = b) at act[b * din + i]
        g.Cf = m->part; g.ldc = m->din[l];
        const bool wide = g.M == WB && g.N == WB;
        // K chunk of the split (a multiple of the kernel's K step), and the number of partials
        const int kstep = wide ? WBK : BK;
        const int kch = ((n + sp - 1) / sp + kstep - 1) / kstep * kstep, nsp = (n + kch - 1) / kch;
        if (wide) {   // the 256 x 256 layers: one block per K chunk
            g.kchunk = kch;
            k_gemm_wgrad<<<dim3(1, 1, (n + g.kchunk - 1) / g.kchunk), 512, 0, s>>>(g);
            if (hipGetLastError() != hipSuccess) return EU3D_ERR_HIP;
        } else if (gemm<false, false, EPI_F32>(s, g, sp) != hipSuccess) {
            return EU3D_ERR_HIP;
        }
        const int nw = m->dout[l] * m->din[l];
        if (nsp > kgrp) {   // two levels: many split-K partials
            float* grp = m->part + (size_t)sp * NH * NH + (size_t)nb * NH;
            const int ngr = (nsp + kgrp - 1) / kgrp;
            k_reduce_splits_grouped<<<dim3((nw + 255) / 256, ngr), 256, 0, s>>>(m->part, grp, nw, nsp, kgrp);
            k_reduce_splits<<<(nw + 255) / 256, 256, 0, s>>>(grp, m->gw[l], nw, ngr);
        } else {
            k_reduce_splits<<<(nw + 255) / 256, 256, 0, s>>>(m->part, m->gw[l], nw, nsp);
        }
        // db = column sums of dZ
        float* cpart = m->part + (size_t)sp * NH * NH;
        k_colsum_bf16<<<nb, 256, 0, s>>>(dz, n, m->dout[l], rpb, cpart);
        if (nb > kgrp) {   // two levels, as for the weights: one thread per column cannot sum nb rows fast
            float* grp = m->part + (size_t)sp * NH * NH + (size_t)nb * NH;
            const int ngr = (nb + kgrp - 1) / kgrp;
            k_reduce_splits_grouped<<<dim3((m->dout[l] + 255) / 256, ngr), 256, 0, s>>>(cpart, grp, m->dout[l], nb, kgrp);
            k_reduce_splits<<<(m->dout[l] + 255) / 256, 256, 0, s>>>(grp, m->gb[l], m->dout[l], ngr);
        } else {
            k_reduce_splits<<<(m->dout[l] + 255) / 256, 256, 0, s>>>(cpart, m->gb[l], m->dout[l], nb);
        }
        if (l > 0) {
            // dZ_{l-1} = (dZ_l [n x dout] * W_l^T) .* GELU'(Z_{l-1}); B operand element
            // (n = i, k = o) = W[i][o] = wn[i * dout + o]
            if (m->din[l] == NH && (m->dout[l] == NH || m->dout[l] == NOUT)) {   // the fused row-tile kernel
                if (m->dout[l] == NH) k_mlp_dx<NH><<<(n + FM - 1) / FM, FT, 0, s>>>(dz, m->wn[l], m->zs[l], dz_next, n);
                else k_mlp_dx<NOUT><<<(n + FM - 1) / FM, FT, 0, s>>>(dz, m->wn[l], m->zs[l], dz_next, n);
                if (hipGetLastError() != hipSuccess) return EU3D_ERR_HIP;
            } else {
                GemmArgs d{};
                d.M = n; d.N = m->din[l]; d.K = m->dout[l];
                d.A = dz; d.lda = m->dout[l];
                d.B = m->wn[l]; d.ldb = m->dout[l];
                d.Cb = dz_next; d.Z = m->zs[l]; d.ldc = m->din[l];
                if (gemm<true, true, EPI_DGELU>(s, d) != hipSuccess) return EU3D_ERR_HIP;
            }
            std::swap(dz, dz_next);
        }
    }
    if (lr > 0) {
        m->adam_t++;
        const float b1 = 0.9f, b2 = 0.999f;
        const float bc1 = 1.f - std::pow(b1, (float)m->adam_t), bc2 = 1.f - std::pow(b2, (float)m->adam_t);
        for (int l = 0; l < 5; l++) {
            const int nw = m->dout[l] * m->din[l];
            k_adam<<<(nw + 255) / 256, 256, 0, s>>>(m->w[l], m->mw[l], m->vw[l], m->gw[l], m->dout[l], m->din[l],
                                                    m->rout[l], m->rin[l], lr, b1, b2, 1e-8f, bc1, bc2, wd, m->wt[l],
                                                    m->wn[l]);
            k_adam<<<(m->dout[l] + 255) / 256, 256, 0, s>>>(m->b[l], m->mb[l], m->vb[l], m->gb[l], m->dout[l], 1,
                                                            m->rout[l], 1, lr, b1, b2, 1e-8f, bc1, bc2, 0.f, nullptr,
                                                            nullptr);
        }
    }
    return hipStreamSynchronize(s) == hipSuccess ? EU3D_OK : EU3D_ERR_HIP;
}

// General BF16 GEMM through the same kernel (exposed for its tests): C[M x N] (FP32,
// row-major) = A[M x K] (K-contiguous) * B where B is given as [N x K] (K-contiguous) when
// b_kc != 0, else as [K x N].
int eu3d_gemm_bf16(int M, int N, int K, const void* A, const void* B, int b_kc, float* C) {
    if (M <= 0 || N <= 0 || K <= 0 || !A || !B || !C) return EU3D_ERR_ARG;
    GemmArgs a{};
    a.M = M; a.N = N; a.K = K;
    a.A = static_cast<const bf16_t*>(A); a.lda = K;
    a.B = static_cast<const bf16_t*>(B); a.ldb = b_kc ? K : N;
    a.Cf = C; a.ldc = N;
    hipError_t e = b_kc ? gemm<true, true, EPI_F32>(nullptr, a) : gemm<true, false, EPI_F32>(nullptr, a);
    if (e != hipSuccess) return EU3D_ERR_HIP;
    return hipDeviceSynchronize() == hipSuccess ? EU3D_OK : EU3D_ERR_HIP;
}

void* eu3d_mlp_stream(eu3d_mlp* m) { return m ? (void*)m->stream : nullptr; }

int eu3d_mlp_set_surrogate(eu3d_mlp* m, const eu3d_surrogate_params* p) {
    if (!m || !p) return EU3D_ERR_ARG;
    for (int i = 0; i < NIN_REAL; i++)
        if (!(p->feat_std[i] > 0)) return EU3D_ERR_ARG;
    for (int i = 0; i < NOUT_REAL; i++)
        if (!(p->out_std[i] > 0)) return EU3D_ERR_ARG;
    if (!(p->out_eps >= 0)) return EU3D_ERR_ARG;
    m->sp = *p;
    m->sp_set = true;
    return EU3D_OK;
}

// Normalised BF16 features [n x 32] of batch cells (rho_s [ns][n], T [n]) for an interval dt.
int eu3d_surrogate_features(eu3d_mlp* m, long long n, const double* rho, const double* T, double dt, void* feats) {
    if (!m || !m->sp_set || n <= 0 || !rho || !T || !feats) return EU3D_ERR_ARG;
    k_sur_feat_batch<<<(unsigned)((n + 255) / 256), 256, 0, m->stream>>>(m->sp, n, rho, T, dt, (bf16_t*)feats);
    return hipStreamSynchronize(m->stream) == hipSuccess ? EU3D_OK : EU3D_ERR_HIP;
}

// Normalised training targets [n x 16] (FP32) from the states before and after the FP64 step.
int eu3d_surrogate_targets(eu3d_mlp* m, long long n, const double* rho0, const double* rho1, float* targets) {
    if (!m || !m->sp_set || n <= 0 || !rho0 || !rho1 || !targets) return EU3D_ERR_ARG;
    k_sur_targets<<<(unsigned)((n + 255) / 256), 256, 0, m->stream>>>(m->sp, n, rho0, rho1, targets);
    return hipStreamSynchronize(m->stream) == hipSuccess ? EU3D_OK : EU3D_ERR_HIP;
}

// Surrogate work buffers of a context: the rejected-cell list, the FP64 work list and the network
// list (n entries each), 64 counters (0 rejected cells, 1..4 rejection reasons, 8 network cells)
// and the per-cell sub-step counts.
struct SurBufs {
    unsigned *rej, *work, *net, *cnt;
    int* keep;
};
static SurBufs sur_bufs(eu3d_ctx* c) {
    unsigned* b = c->sur_list;
    const size_t n = (size_t)c->sur_cap;
    return {b, b + n, b + 2 * n, b + 3 * n, reinterpret_cast<int*>(b + 3 * n + 64)};
}
constexpr int CNT_NET = 8;

static int sur_ensure(eu3d_ctx* c, long long n) {
    if (c->sur_cap >= n) return EU3D_OK;
    if (c->sur_list) (void)hipFree(c->sur_list);
    c->sur_list = nullptr;
    c->sur_cap = 0;
    if (hipMalloc(&c->sur_list, (size_t)n * 3 * sizeof(unsigned) + 64 * sizeof(unsigned) + (size_t)n * sizeof(int)) != hipSuccess)
        return EU3D_ERR_HIP;
    c->sur_cap = n;
    return EU3D_OK;
}

static int sur_stats(eu3d_ctx* c, eu3d_surrogate_stats* st, long long n) {
    unsigned h[16];
    if (hipMemcpyAsync(h, sur_bufs(c).cnt, sizeof(h), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return EU3D_ERR_HIP;
    if (st) {
        st->cells = n;
        st->fallback = h[0];
        st->hull = h[1];
        st->negative = h[2];
        st->element = h[3];
        st->dT = h[4];
        st->network = h[CNT_NET];
    }
    return EU3D_OK;
}

// Surrogate chemistry on batch cells at rest, with the FP64 adaptive trapezoid on every cell the
// checks reject. Device pointers as eu3d_chem_batch.
int eu3d_surrogate_batch(eu3d_ctx* c, eu3d_mlp* m, long long n, double* rho, double* T, double* E, double dtm,
                         int nchem_max, eu3d_surrogate_stats* st) {
    if (!c || !m || !m->sp_set || n <= 0 || n >= (1ll << 31) || !rho || !T || !E || !(dtm > 0)) return EU3D_ERR_ARG;
    if (ensure_infer(m, (int)n) != hipSuccess || sur_ensure(c, n) != EU3D_OK) return EU3D_ERR_HIP;
    const SurBufs b = sur_bufs(c);
    const eu3d_surrogate_params& sp = m->sp;
    const unsigned nb = (unsigned)((n + 255) / 256);
    k_zero_u32<<<1, 64, 0, c->stream>>>(b.cnt, 64);
    if (sp.nchem_threshold <= 0) {   // every cell through the network
        k_sur_feat_batch<<<nb, 256, 0, c->stream>>>(sp, n, rho, T, dtm, m->ifeat);
        if (forward_fused(m, (int)n, m->ifeat, m->iout, c->stream) != hipSuccess) return EU3D_ERR_HIP;
        k_sur_apply_batch<<<nb, 256, 0, c->stream>>>(sp, n, rho, T, E, m->iout, dtm, c->quirks, b.cnt, b.rej);
    } else {   // stiff cells through the network, the rest through the FP64 sub-steps
        if (eu3d::launch_chem_batch(c, n, rho, T, E, dtm, nchem_max, nullptr, sp.nchem_threshold, sp.T_cold, b.net,
                                    b.cnt + CNT_NET, sp.nchem_threshold_hi) != hipSuccess)
            return EU3D_ERR_HIP;
        const unsigned nl = std::min<unsigned>(nb, (unsigned)c->num_cu * 8);
        k_sur_feat_batch_list<<<nl, 256, 0, c->stream>>>(sp, n, rho, T, dtm, b.net, b.cnt + CNT_NET, m->ifeat);
        if (forward_fused(m, (int)n, m->ifeat, m->iout, c->stream, b.cnt + CNT_NET) != hipSuccess) return EU3D_ERR_HIP;
        k_sur_apply_batch_list<<<nl, 256, 0, c->stream>>>(sp, n, rho, T, E, m->iout, dtm, c->quirks, b.net,
                                                          b.cnt + CNT_NET, b.cnt, b.rej);
    }
    if (eu3d::launch_batch_list(c, n, rho, T, E, dtm, b.rej, b.cnt, (unsigned)n, b.work, b.keep, nchem_max) != hipSuccess)
        return EU3D_ERR_HIP;
    return sur_stats(c, st, n);
}

// Surrogate chemistry over the grid's interior with the FP64 fallback, on the context's dt (the
// chemistry half of a step with EU3D_CHEM_SURROGATE, after eu3d_update_after_adv).
int eu3d_surrogate_chem(eu3d_ctx* c, int nchem_max, eu3d_surrogate_stats* st) {
    if (!c || !c->mlp) return EU3D_ERR_ARG;
    if (eu3d::launch_surrogate_grid(c, nchem_max) != hipSuccess) return EU3D_ERR_HIP;
    const long long n = (long long)c->g.ni * c->g.nj * c->g.nk;
    const int rc = sur_stats(c, &c->sur_stats, n);
    if (st) *st = c->sur_stats;
    return rc;
}

int eu3d_set_surrogate(eu3d_ctx* c, eu3d_mlp* m) {
    if (!c) return EU3D_ERR_ARG;
    if (m && !m->sp_set) return EU3D_ERR_ARG;
    c->mlp = m;
    return EU3D_OK;
}

int eu3d_get_surrogate_stats(eu3d_ctx* c, eu3d_surrogate_stats* st) {
    if (!c || !st) return EU3D_ERR_ARG;
    *st = c->sur_stats;
    return EU3D_OK;
}

}  // extern "C"

namespace eu3d {
// Surrogate chemistry over the interior of a grid context (the chemistry half of a step with
// EU3D_CHEM_SURROGATE): features, network, apply with checks, FP64 fallback on the rejects.
hipError_t launch_surrogate_grid(eu3d_ctx* c, int nchem_max, bool adv, bool last) {
    eu3d_mlp* m = c->mlp;
    const long long n = (long long)c->g.ni * c->g.nj * c->g.nk;
    if (!m || ensure_infer(m, (int)n) != hipSuccess || sur_ensure(c, n) != EU3D_OK) return hipErrorInvalidValue;
    const SurBufs b = sur_bufs(c);
    const eu3d_surrogate_params& sp = m->sp;
    k_zero_u32<<<1, 64, 0, c->stream>>>(b.cnt, 64);
    if (sp.nchem_threshold <= 0) {   // every cell through the network
        if (adv) {
            hipError_t e = launch_update_after_adv(c);
            if (e != hipSuccess) return e;
        }
        const dim3 grid((c->g.ni + 63) / 64, (c->g.nj + 3) / 4, c->g.nk), blk(64, 4);
        k_sur_feat_grid<<<grid, blk, 0, c->stream>>>(sp, c->g, c->fl, c->sc, m->ifeat);
        hipError_t e = forward_fused(m, (int)n, m->ifeat, m->iout, c->stream);
        if (e != hipSuccess) return e;
        k_sur_apply_grid<<<grid, blk, 0, c->stream>>>(sp, c->g, c->fl, c->sc, m->iout, c->quirks, b.cnt, b.rej);
    } else {   // stiff cells through the network, the rest through the FP64 sub-steps
        // within an eu3d_step call (adv) the source is stored on the call's last step, and the
        // steps before it leave the next step's CFL partials: the prediction pass and the FP64
        // queue kernels for the cells they finish, k_cfl_list for the network's cells
        const bool part = adv && !last;
        hipError_t e = launch_trap_adaptive(c, nchem_max, false, sp.nchem_threshold, sp.T_cold, b.net, b.cnt + CNT_NET,
                                            adv, sp.nchem_threshold_hi, !part, part);
        if (e != hipSuccess) return e;
        // the network cells' state after Update_after_CS, which the fused pass does not store
        if (adv && (e = launch_adv_list(c, b.net, b.cnt + CNT_NET, (unsigned)n)) != hipSuccess) return e;
        const unsigned nb = std::min<unsigned>((unsigned)((n + 255) / 256), (unsigned)c->num_cu * 8);
        k_sur_feat_grid_list<<<nb, 256, 0, c->stream>>>(sp, c->g, c->fl, c->sc, b.net, b.cnt + CNT_NET, m->ifeat);
        if ((e = forward_fused(m, (int)n, m->ifeat, m->iout, c->stream, b.cnt + CNT_NET)) != hipSuccess) return e;
        k_sur_apply_grid_list<<<nb, 256, 0, c->stream>>>(sp, c->g, c->fl, c->sc, m->iout, c->quirks, b.net,
                                                         b.cnt + CNT_NET, b.cnt, b.rej);
        if ((e = launch_trap_list(c, b.rej, b.cnt, (unsigned)n, b.work, nchem_max)) != hipSuccess) return e;
        if (part) {
            if ((e = launch_cfl_list(c, b.net, b.cnt + CNT_NET, (unsigned)n)) != hipSuccess) return e;
            c->part_ready = true;
        }
        return hipSuccess;
    }
    return launch_trap_list(c, b.rej, b.cnt, (unsigned)n, b.work, nchem_max);
}

hipError_t surrogate_grid_stats(eu3d_ctx* c) {
    const long long n = (long long)c->g.ni * c->g.nj * c->g.nk;
    return sur_stats(c, &c->sur_stats, n) == EU3D_OK ? hipSuccess : hipErrorUnknown;
}
}  // namespace eu3d

extern "C" {

}  // extern "C"
