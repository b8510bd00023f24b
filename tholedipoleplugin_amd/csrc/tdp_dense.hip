// tdp_dense.hip -- dense direct solve of the induced-dipole equations of small systems (gfx950).
//
// For each system of a batch context the scaled matrix  A^ = I - D^1/2 T D^1/2  (D = diag(alpha),
// T the Thole-damped dipole field tensor of calculateInducedDipolePairIxn,
// ReferenceTholeDipoleForce.cpp:766-819, with the same iScale table) is assembled in FP16 or BF16,
// factored as L L^T by a blocked left-looking Cholesky whose tile products run on MFMA with FP32
// accumulation, and used as the preconditioner of an FP64 iterative refinement whose residuals are
// formed matrix-free by the FP64 T*mu kernels.
//
// Layout of one system with nt particle tiles: unknown r = c * 64 nt + s (component c of local
// slot s), so block R = c nt + I of 64 unknowns is component c of particle tile I, and a 64x64
// block of A^ couples two particle tiles in one component pair.  m = 192 nt.  The factor is kept
// in place: block (i, j), i > j, holds L_ij and block (j, i) its transpose; the inverse of each
// diagonal block L_jj is kept in FP32.  Padding slots have alpha = 0 and identity rows.
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include "tdp_device.hpp"
#include "tdp_kernels.h"

namespace tdp {
namespace {

typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float16_t __attribute__((ext_vector_type(16)));

constexpr int kB = 64;   // block size of the factorization (one particle tile)

// Storage type of the factor: FP16 or BF16 bit patterns in 16-bit words.
template <bool BF16>
__device__ __forceinline__ unsigned short to_lp(float x) {
    if (BF16) {
        // round to nearest even on the upper 16 bits
        unsigned u = __float_as_uint(x);
        unsigned r = u + 0x7fffu + ((u >> 16) & 1u);
        return (unsigned short)(r >> 16);
    }
    // keep x an FP32 value: otherwise the FP64 -> FP32 -> FP16 conversions of the callers fold into
    // one FP64 -> FP16 conversion, which gfx950 has no instruction for (a long integer sequence)
    asm volatile("" : "+v"(x));
    _Float16 h = (_Float16)x;
    return __builtin_bit_cast(unsigned short, h);
}
template <bool BF16>
__device__ __forceinline__ float from_lp(unsigned short b) {
    if (BF16) return __uint_as_float(((unsigned)b) << 16);
    return (float)__builtin_bit_cast(_Float16, b);
}

// 32x32x8 MFMA with FP32 accumulation.  Operand lane layout (both A and B): lane l holds k =
// 4 (l / 32) + t, t = 0..3, of row (A) / column (B) l % 32; accumulator element v of lane l is row
// 8 (v / 4) + 4 (l / 32) + v % 4, column l % 32.
template <bool BF16>
__device__ __forceinline__ float16_t mfma32(uint2 a, uint2 b, float16_t c) {
    if (BF16)
        return __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(__builtin_bit_cast(short4_t, a), __builtin_bit_cast(short4_t, b),
                                                       c, 0, 0, 0);
    return __builtin_amdgcn_mfma_f32_32x32x8f16(__builtin_bit_cast(half4_t, a), __builtin_bit_cast(half4_t, b), c, 0,
                                                0, 0);
}

__device__ __forceinline__ size_t blk_off(const DenseSys& d, int R, int C) {
    const size_t m = (size_t)192 * d.nt;
    return d.moff + (size_t)R * kB * m + (size_t)C * kB;
}

// ------------------------------------------------------------------ assembly
// One workgroup per (system, particle tile pair I, J): the 3x3 tensors of its 64 x 64 particle
// pairs, written into the 9 component blocks (c I, c' J).  A^_(ci)(c'j) = delta - s_i s_j T^cc'_ij
// with s = sqrt(alpha) and T^cc' = iScale (rr5 D_c D_c' - rr3 delta_cc') (calculateInducedDipolePairIxn).
template <bool BF16>
__global__ __launch_bounds__(256) void k_dense_assemble(const DenseSys* __restrict__ sys, const int* __restrict__ tcnt,
                                                        const double4* __restrict__ posa,
                                                        const double* __restrict__ alpha,
                                                        const double* __restrict__ inva,
                                                        const int* __restrict__ exc_start,
                                                        const int* __restrict__ exc_J,
                                                        const unsigned long long* __restrict__ exc_mask,
                                                        unsigned short* __restrict__ mat) {
    const DenseSys d = sys[blockIdx.z];
    const int I = blockIdx.x, J = blockIdx.y;
    if (I >= d.nt || J >= d.nt) return;
    __shared__ unsigned long long s_mi[kB];
    __shared__ double4 s_pj[kB];
    __shared__ double s_aj[kB], s_iaj[kB];
    const int t = threadIdx.x;
    const int TI = d.t0 + I, TJ = d.t0 + J;
    const int cntI = tcnt[TI], cntJ = tcnt[TJ];
    if (t < kB) {
        unsigned long long m = 0ull;
        for (int e = exc_start[TI]; e < exc_start[TI + 1]; e++)
            if (exc_J[e] == TJ) m = exc_mask[((size_t)e * 3 + 2) * kTile + t];
        s_mi[t] = m;
        s_pj[t] = posa[TJ * kTile + t];
        s_aj[t] = alpha[TJ * kTile + t];
        s_iaj[t] = inva[TJ * kTile + t];
    }
    __syncthreads();
    const int j = t & (kB - 1);
    const size_t m = (size_t)192 * d.nt;
    for (int i = t >> 6; i < kB; i += 4) {
        const double4 pi = posa[TI * kTile + i];
        const double ai = alpha[TI * kTile + i];
        const double4 pj = s_pj[j];
        const bool valid = i < cntI && j < cntJ && !((s_mi[i] >> j) & 1ull);
        double dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
        double r2 = valid ? dx * dx + dy * dy + dz * dz : 1.0;
        double rr3, rr5;
        damped_rr(r2, pi.w, pj.w, inva[TI * kTile + i], s_iaj[j], rr3, rr5);
        const double sc = valid ? sqrt(ai * s_aj[j]) : 0.0;
        const double D[3] = {dx, dy, dz};
        const bool diag = TI == TJ && i == j;
        for (int ci = 0; ci < 3; ci++)
            for (int cj = 0; cj < 3; cj++) {
                double tv = rr5 * D[ci] * D[cj] - (ci == cj ? rr3 : 0.0);
                double v = (diag && ci == cj ? 1.0 : 0.0) - sc * tv;
                const size_t row = (size_t)(ci * d.nt + I) * kB + i, col = (size_t)(cj * d.nt + J) * kB + j;
                mat[d.moff + row * m + col] = to_lp<BF16>((float)v);
            }
    }
}

// ------------------------------------------------------------------ factorization
// S = A^_ij - sum_{k<j} L_ik L_jk^T for one 64x64 block, on MFMA: four waves, each a 32x32 quarter.
// Operand tiles are staged in LDS as 64 rows of 16 8-byte words (4 values each), row stride 17
// words to spread the banks.
constexpr int kLd2 = kB / 4 + 1;
static_assert(2 * kB * (kB + 1) * sizeof(float) <= 4 * kB * kLd2 * sizeof(uint2), "LDS overlay");

template <bool BF16>
__device__ __forceinline__ void block_update(const DenseSys& d, int i, int j, uint2* s_a, uint2* s_b,
                                             const unsigned short* __restrict__ mat, float16_t& acc) {
    // s_a, s_b: two buffers each of 64 rows x kLd2 8-byte words.  The tiles of the next kDepth steps
    // are in flight in a register ring while step k's MFMAs run on one buffer and step k + 1's tiles
    // are written to the other; the k loop is unrolled by kDepth so that the ring is indexed statically.
    constexpr int kDepth = 4;
    const int t = threadIdx.x, w = t >> 6, l = t & 63;
    const int wr = w >> 1, wc = w & 1;
    const size_t m = (size_t)192 * d.nt;
    if (j == 0) return;
    // 64 rows x 16 words per tile: thread t loads words (t, t + 256, t + 512, t + 768)
    uint2 ra[kDepth][4], rb[kDepth][4];
    auto fetch = [&](int k, uint2 (&xa)[4], uint2 (&xb)[4]) {
        const uint2* gi = reinterpret_cast<const uint2*>(mat + blk_off(d, i, k));
        const uint2* gj = reinterpret_cast<const uint2*>(mat + blk_off(d, j, k));
        const size_t mw = m / 4;
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int e = t + 256 * u, r = e >> 4, c = e & 15;
            xa[u] = gi[(size_t)r * mw + c];
            xb[u] = gj[(size_t)r * mw + c];
        }
    };
    auto stash = [&](int buf, const uint2 (&xa)[4], const uint2 (&xb)[4]) {
        uint2* pa = s_a + buf * kB * kLd2;
        uint2* pb = s_b + buf * kB * kLd2;
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int e = t + 256 * u, r = e >> 4, c = e & 15;
            pa[r * kLd2 + c] = xa[u];
            pb[r * kLd2 + c] = xb[u];
        }
    };
#pragma unroll
    for (int q = 0; q < kDepth; q++)
        if (q < j) fetch(q, ra[q], rb[q]);
    stash(0, ra[0], rb[0]);
    __syncthreads();
    for (int k0 = 0; k0 < j; k0 += kDepth) {
#pragma unroll
        for (int q = 0; q < kDepth; q++) {
            const int k = k0 + q;
            if (k < j) {
                const int buf = k & 1;
                if (k + kDepth < j) fetch(k + kDepth, ra[q], rb[q]);   // slot q held tile k, already staged
                const uint2* pa = s_a + buf * kB * kLd2 + (wr * 32 + (l & 31)) * kLd2 + (l >> 5);
                const uint2* pb = s_b + buf * kB * kLd2 + (wc * 32 + (l & 31)) * kLd2 + (l >> 5);
#pragma unroll
                for (int kk = 0; kk < kB / 4; kk += 2) acc = mfma32<BF16>(pa[kk], pb[kk], acc);
                if (k + 1 < j) stash(buf ^ 1, ra[(q + 1) % kDepth], rb[(q + 1) % kDepth]);
                __syncthreads();
            }
        }
    }
}

// Diagonal block j of every system with more than j blocks: S = A^_jj - sum_{k<j} L_jk L_jk^T, its
// Cholesky factor in FP32 and the inverse X of that factor (kept in FP32 for the off-diagonal blocks
// and the triangular solves; the factor itself is not needed).  The sum over k < j arrives
// precomputed: k_chol_col adds L_jk L_jk^T of every finished column k to the FP32 running block
// dacc_j, so the serial chain of diagonal factorizations carries no block products.
// Right-looking in registers: wave w holds rows 16 w .. 16 w + 15 of S and X, lane c column c.  At
// pivot p the owner of row p scales it (by symmetry of the trailing S, row p past the diagonal is
// column p of L) and scales row p of X; both go through LDS (double-buffered by the parity of p, one
// barrier per pivot), and every wave updates S -= L_rp L_cp and, for its rows r > p, X_r -= L_rp X_p.
template <bool BF16>
__global__ __launch_bounds__(256) void k_chol_diag(int j, const DenseSys* __restrict__ sys,
                                                   const unsigned short* __restrict__ mat,
                                                   const float* __restrict__ dacc, float* __restrict__ dinv,
                                                   int* __restrict__ status) {
    const DenseSys d = sys[blockIdx.x];
    const int nb = 3 * d.nt;
    if (j >= nb) return;
    __shared__ __align__(16) float s_col[2][kB];
    __shared__ __align__(16) float s_xr[2][kB];
    const int t = threadIdx.x, w = t >> 6, c = t & 63;
    const size_t m = (size_t)192 * d.nt;
    const unsigned short* g = mat + blk_off(d, j, j);
    const float* ac = dacc + d.doff + (size_t)j * kB * kB;
    float S[16], X[16];
#pragma unroll
    for (int u = 0; u < 16; u++) {
        const int r = 16 * w + u;
        S[u] = from_lp<BF16>(g[(size_t)r * m + c]) - ac[r * kB + c];
        X[u] = r == c ? 1.f : 0.f;
    }
    bool bad = false;
    for (int wp = 0; wp < 4; wp++) {
#pragma unroll
        for (int up = 0; up < 16; up++) {
            const int p = 16 * wp + up, b = p & 1;
            if (w == wp) {
                float dv = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(S[up]), p));   // S_pp
                if (!(dv > 0.f)) {
                    bad = true;
                    dv = 1.f;
                }
                const float piv = sqrtf(dv), ip = 1.f / piv;
                s_col[b][c] = c == p ? piv : S[up] * ip;   // L_cp for c > p
                X[up] *= ip;
                s_xr[b][c] = X[up];
            }
            __syncthreads();
            const float lc = s_col[b][c], xp = s_xr[b][c];
#pragma unroll
            for (int u = 0; u < 16; u++) {
                const int r = 16 * w + u;
                const float lr = s_col[b][r];
                S[u] -= lr * lc;              // rows or columns <= p are finished and never read again
                if (r > p) X[u] -= lr * xp;   // X stays lower triangular
            }
        }
    }
    if (bad && c == 0) atomicOr(status, 1);
    float* out = dinv + d.doff + (size_t)j * kB * kB;
#pragma unroll
    for (int u = 0; u < 16; u++) out[(16 * w + u) * kB + c] = X[u];
}

// Off-diagonal blocks (i, j), i > j: L_ij = (A^_ij - sum_{k<j} L_ik L_jk^T) L_jj^-T, stored in the
// 16-bit format at (i, j) and transposed at (j, i).
// Each block also adds L_ij L_ij^T (from the stored 16-bit L_ij, on MFMA) to the FP32 running sum
// dacc_i of its diagonal block; one workgroup per (i, j), and the columns j are successive launches.
template <bool BF16>
__global__ __launch_bounds__(256) void k_chol_col(int j, const DenseSys* __restrict__ sys,
                                                  unsigned short* __restrict__ mat, const float* __restrict__ dinv,
                                                  float* __restrict__ dacc) {
    const DenseSys d = sys[blockIdx.y];
    const int nb = 3 * d.nt;
    const int i = j + 1 + blockIdx.x;
    if (i >= nb) return;
    // the staging buffers of the update and the FP32 blocks used after it share one LDS area
    __shared__ uint2 s_raw[4 * kB * kLd2];
    uint2* s_a = s_raw;
    uint2* s_b = s_raw + 2 * kB * kLd2;
    float (*s_s)[kB + 1] = reinterpret_cast<float (*)[kB + 1]>(s_raw);
    float (*s_x)[kB + 1] = s_s + kB;
    const int t = threadIdx.x, w = t >> 6, l = t & 63;
    const int wr = w >> 1, wc = w & 1;
    float16_t acc = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    block_update<BF16>(d, i, j, s_a, s_b, mat, acc);
    const size_t m = (size_t)192 * d.nt;
    unsigned short* g = mat + blk_off(d, i, j);
    for (int v = 0; v < 16; v++) {
        const int r = wr * 32 + 8 * (v >> 2) + 4 * (l >> 5) + (v & 3), c = wc * 32 + (l & 31);
        s_s[r][c] = from_lp<BF16>(g[(size_t)r * m + c]) - acc[v];
    }
    const float* inv = dinv + d.doff + (size_t)j * kB * kB;
    for (int e = t; e < kB * kB; e += 256) s_x[e >> 6][e & 63] = inv[e];
    __syncthreads();
    // L_ij[r][c] = sum_k S[r][k] Linv[c][k]  (Linv lower: k <= c); thread: rows 4 rb.., cols 4 cb..
    unsigned short* gt = mat + blk_off(d, j, i);
    {
        const int rb = (t >> 4) * 4, cb = (t & 15) * 4;
        float o[4][4] = {};
        for (int k = 0; k <= cb + 3; k++) {
            float sv[4], xv[4];
            for (int u = 0; u < 4; u++) {
                sv[u] = s_s[rb + u][k];
                xv[u] = s_x[cb + u][k];
            }
            for (int u = 0; u < 4; u++)
                for (int v = 0; v < 4; v++) o[u][v] += sv[u] * xv[v];
        }
        unsigned short lb[4][4];
        for (int u = 0; u < 4; u++)
            for (int v = 0; v < 4; v++) lb[u][v] = to_lp<BF16>(o[u][v]);
        // four 16-bit values per 8-byte store: rows of L_ij and rows of its transpose
        for (int u = 0; u < 4; u++)
            *reinterpret_cast<uint2*>(g + (size_t)(rb + u) * m + cb) =
                make_uint2((unsigned)lb[u][0] | ((unsigned)lb[u][1] << 16), (unsigned)lb[u][2] | ((unsigned)lb[u][3] << 16));
        for (int v = 0; v < 4; v++)
            *reinterpret_cast<uint2*>(gt + (size_t)(cb + v) * m + rb) =
                make_uint2((unsigned)lb[0][v] | ((unsigned)lb[1][v] << 16), (unsigned)lb[2][v] | ((unsigned)lb[3][v] << 16));
        __syncthreads();   // s_s and s_x are read no more: their LDS now stages L_ij
        for (int u = 0; u < 4; u++)
            s_a[(rb + u) * kLd2 + (cb >> 2)] = make_uint2((unsigned)lb[u][0] | ((unsigned)lb[u][1] << 16),
                                                          (unsigned)lb[u][2] | ((unsigned)lb[u][3] << 16));
    }
    __syncthreads();
    {
        float16_t acc2 = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        const uint2* pa = s_a + (wr * 32 + (l & 31)) * kLd2 + (l >> 5);
        const uint2* pb = s_a + (wc * 32 + (l & 31)) * kLd2 + (l >> 5);
#pragma unroll
        for (int kk = 0; kk < kB / 4; kk += 2) acc2 = mfma32<BF16>(pa[kk], pb[kk], acc2);
        float* ac = dacc + d.doff + (size_t)i * kB * kB;
        for (int v = 0; v < 16; v++) {
            const int r = wr * 32 + 8 * (v >> 2) + 4 * (l >> 5) + (v & 3), c = wc * 32 + (l & 31);
            ac[r * kB + c] += acc2[v];
        }
    }
}

// ------------------------------------------------------------------ triangular solves
// One workgroup of 16 waves per system solves L L^T y = rhs for the 4 right-hand sides, left-looking
// by blocks.  Step R forms b_R - (row block R of L or L^T, off the diagonal) y on MFMA
// (v_mfma_f32_16x16x32 f16/bf16, FP32 accumulation): the 16-bit factor rows are the A operand,
// read once from memory, and y the B operand, split into a high and a low 16-bit part per field
// (columns 0-3 and 4-7; columns 8-15 are zero), so the products carry y to about 2^-22 (FP16) or
// 2^-16 (BF16) relative, below the rounding of the factor itself.  Row R*64 + r of L left of the
// diagonal block sits in the lower blocks and row R*64 + r of L^T right of it in the upper
// (transposed) blocks, both contiguous.  The 32-column chunks of the row block are dealt to the
// waves in turn; each wave covers all 64 rows (four 16-row tiles) of its chunks and its partial
// sums are added in a fixed order, then the step applies the FP32 inverse of the diagonal block.
// y stays in LDS (FP32, and its 16-bit parts).  Each field is scaled by a power of two that brings
// its largest right-hand side entry into [1, 2) (exact; undone on output), so that the 16-bit parts
// of the small corrections of later refinement steps stay within the FP16 exponent range.  rhs and
// y are FP64 arrays of m x 4 values (unknown, then right-hand side) at d.voff.
constexpr int kMaxDenseTiles = 16;   // up to 1,024 particles per system: m <= 3,072
constexpr int kMaxM = 192 * kMaxDenseTiles;
constexpr int kSolveWaves = 16;
constexpr int kSolveThreads = kSolveWaves * 64;
constexpr int kSolveRound = 2;       // chunks per wave whose factor loads are in flight together

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float float4_t __attribute__((ext_vector_type(4)));

// 16x16x32 MFMA with FP32 accumulation.  Lane l holds row l % 16 of A and column l % 16 of B, with
// the same 8 values of k (group l / 16) in both; accumulator element v of lane l is row
// 4 (l / 16) + v, column l % 16.
template <bool BF16>
__device__ __forceinline__ float4_t mfma16(uint4 a, uint4 b, float4_t c) {
    if (BF16)
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, a), __builtin_bit_cast(bf16x8_t, b),
                                                       c, 0, 0, 0);
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(half8_t, a), __builtin_bit_cast(half8_t, b), c, 0,
                                                  0, 0);
}

template <bool BF16>
__global__ __launch_bounds__(kSolveThreads) void k_dense_solve(const DenseSys* __restrict__ sys,
                                                               const unsigned short* __restrict__ mat,
                                                               const float* __restrict__ dinv,
                                                               const double* __restrict__ rhs, double* __restrict__ y,
                                                               const int* __restrict__ active) {
    const DenseSys d = sys[blockIdx.x];
    if (active && !active[blockIdx.x]) return;
    const int nb = 3 * d.nt;
    const int m = 192 * d.nt;
    __shared__ float4 s_y[kMaxM];                               // 48 KB: y, then x, 4 fields
    __shared__ __align__(16) unsigned short s_h[8][kMaxM];      // 48 KB: high parts (fields 0-3), low parts
    __shared__ float s_part[kSolveWaves][8][kB + 1];            // 32.5 KB: per-wave partial products
    __shared__ float s_inv[kB][kB + 1];
    __shared__ float s_b[4][kB];
    __shared__ float s_max[kSolveWaves][4];
    const int t = threadIdx.x, w = t >> 6, l = t & 63;
    const int n = l & 15, g = l >> 4;
    float4 mx = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int e = t; e < m; e += kSolveThreads) {
        const double* q = rhs + d.voff + 4 * (size_t)e;
        const float4 v = make_float4((float)q[0], (float)q[1], (float)q[2], (float)q[3]);
        s_y[e] = v;
        mx = make_float4(fmaxf(mx.x, fabsf(v.x)), fmaxf(mx.y, fabsf(v.y)), fmaxf(mx.z, fabsf(v.z)),
                         fmaxf(mx.w, fabsf(v.w)));
    }
    for (int o = 32; o > 0; o >>= 1) {
        mx.x = fmaxf(mx.x, __shfl_xor(mx.x, o));
        mx.y = fmaxf(mx.y, __shfl_xor(mx.y, o));
        mx.z = fmaxf(mx.z, __shfl_xor(mx.z, o));
        mx.w = fmaxf(mx.w, __shfl_xor(mx.w, o));
    }
    if (l == 0) {
        s_max[w][0] = mx.x; s_max[w][1] = mx.y; s_max[w][2] = mx.z; s_max[w][3] = mx.w;
    }
    __syncthreads();
    // per field: 2^-e with the largest |rhs| in [2^e, 2^(e+1)) (1 for a zero field)
    float sc[4], isc[4];
#pragma unroll
    for (int f = 0; f < 4; f++) {
        float v = 0.f;
        for (int u = 0; u < kSolveWaves; u++) v = fmaxf(v, s_max[u][f]);
        const int e = v > 0.f ? ilogbf(v) : 0;
        sc[f] = ldexpf(1.f, -e);
        isc[f] = ldexpf(1.f, e);
    }
    for (int e = t; e < m; e += kSolveThreads) {
        const float4 v = s_y[e];
        s_y[e] = make_float4(v.x * sc[0], v.y * sc[1], v.z * sc[2], v.w * sc[3]);
    }
    __syncthreads();
    constexpr int kInvPer = kB * kB / kSolveThreads;   // diagonal-inverse words per thread
    // Steps 0..nb-1: L y = rhs, blocks R = 0..nb-1; steps nb..2nb-1: L^T x = y, blocks R = nb-1..0.
    // Step q reads the columns [c0, c0 + 32 nch) of row block R: left of the diagonal (forward) or
    // right of it (backward).  The factor does not depend on y, so the loads of a step's first round
    // of chunks and of its diagonal inverse are issued at the end of the step before, ahead of its
    // barriers.  Lane (n, g) loads row 16 rt + n of the row block, columns 32 ch + 8 g .. + 7.
    auto geom = [&](int q, int& R, int& c0, int& nch) {
        R = q < nb ? q : 2 * nb - 1 - q;
        c0 = q < nb ? 0 : (R + 1) * kB;
        nch = (q < nb ? R * kB : m - c0) >> 5;
    };
    uint4 a[kSolveRound][4];
    float iv[kInvPer];
    auto issue = [&](int R, int c0, int nch, int ch0) {
        const unsigned short* arow = mat + d.moff + (size_t)(R * kB + n) * m + c0 + 8 * g;
#pragma unroll
        for (int u = 0; u < kSolveRound; u++) {
            const int ch = ch0 + kSolveWaves * u;
            if (ch < nch) {
#pragma unroll
                for (int rt = 0; rt < 4; rt++)
                    a[u][rt] = *reinterpret_cast<const uint4*>(arow + (size_t)(16 * rt) * m + 32 * ch);
            }
        }
    };
    auto issue_inv = [&](int R) {
        const float* inv = dinv + d.doff + (size_t)R * kB * kB;
#pragma unroll
        for (int u = 0; u < kInvPer; u++) iv[u] = inv[t + kSolveThreads * u];
    };
    {
        int R, c0, nch;
        geom(0, R, c0, nch);
        issue(R, c0, nch, w);
        issue_inv(R);
    }
    for (int q = 0; q < 2 * nb; q++) {
        const int pass = q < nb ? 0 : 1;
        int R, c0, nch;
        geom(q, R, c0, nch);
        const unsigned short* brow = &s_h[n & 7][c0 + 8 * g];
        float4_t acc[4];
#pragma unroll
        for (int rt = 0; rt < 4; rt++) acc[rt] = float4_t{0.f, 0.f, 0.f, 0.f};
        for (int ch0 = w;;) {
#pragma unroll
            for (int u = 0; u < kSolveRound; u++) {
                const int ch = ch0 + kSolveWaves * u;   // uniform over the wave
                if (ch < nch) {
                    uint4 bv = *reinterpret_cast<const uint4*>(brow + 32 * ch);
                    if (n >= 8) bv = make_uint4(0u, 0u, 0u, 0u);
#pragma unroll
                    for (int rt = 0; rt < 4; rt++) acc[rt] = mfma16<BF16>(a[u][rt], bv, acc[rt]);
                }
            }
            ch0 += kSolveWaves * kSolveRound;
            if (ch0 >= nch) break;
            issue(R, c0, nch, ch0);
        }
        if (n < 8) {
#pragma unroll
            for (int rt = 0; rt < 4; rt++)
#pragma unroll
                for (int v = 0; v < 4; v++) s_part[w][n][16 * rt + 4 * g + v] = acc[rt][v];
        }
#pragma unroll
        for (int u = 0; u < kInvPer; u++) {
            const int e = t + kSolveThreads * u;
            s_inv[e >> 6][e & 63] = iv[u];
        }
        if (q + 1 < 2 * nb) {
            int R1, c01, nch1;
            geom(q + 1, R1, c01, nch1);
            issue(R1, c01, nch1, w);
            issue_inv(R1);
        }
        __syncthreads();
        if (t < 4 * kB) {
            const int r = t & 63, f = t >> 6;
            float sum = 0.f;
            for (int u = 0; u < kSolveWaves; u++) sum += s_part[u][f][r] + s_part[u][f + 4][r];
            s_b[f][r] = reinterpret_cast<const float*>(&s_y[R * kB + r])[f] - sum;
        }
        __syncthreads();
        // diagonal block: thread (r, f) applies Linv (forward) or Linv^T (backward) to field f
        if (t < 4 * kB) {
            const int r = t & 63, f = t >> 6;
            // the inverse is lower triangular with explicit zeros above the diagonal, so the sums run
            // over all 64 columns (a fixed count the compiler unrolls, with the LDS reads batched)
            float o = 0.f;
            if (pass == 0) {
#pragma unroll
                for (int k = 0; k < kB; k++) o += s_inv[r][k] * s_b[f][k];
            } else {
#pragma unroll
                for (int k = 0; k < kB; k++) o += s_inv[k][r] * s_b[f][k];
            }
            reinterpret_cast<float*>(&s_y[R * kB + r])[f] = o;
            const unsigned short hi = to_lp<BF16>(o);
            s_h[f][R * kB + r] = hi;
            s_h[4 + f][R * kB + r] = to_lp<BF16>(o - from_lp<BF16>(hi));
        }
        __syncthreads();
    }
    for (int e = t; e < m; e += kSolveThreads) {
        const float4 v = s_y[e];
        double* q = y + d.voff + 4 * (size_t)e;
        q[0] = (double)v.x * isc[0];
        q[1] = (double)v.y * isc[1];
        q[2] = (double)v.z * isc[2];
        q[3] = (double)v.w * isc[3];
    }
}

// ------------------------------------------------------------------ refinement vectors
// Slot i of tile T (system g, local tile I, lane s) is unknown c * 64 nt + 64 I + s.
__device__ __forceinline__ size_t vec_index(const DenseSys& d, int T, int s, int c) {
    return d.voff + 4 * ((size_t)c * kB * d.nt + (size_t)(T - d.t0) * kB + s);
}

// rhs^ = sqrt(alpha) [E0, e_x, e_y, e_z]; x^ = 0; r^ = rhs^
__global__ __launch_bounds__(64) void k_dense_rhs(const int* __restrict__ tseg, const DenseSys* __restrict__ sys,
                                                  const int* __restrict__ tcnt, const double4* __restrict__ E0,
                                                  const double* __restrict__ alpha, double* __restrict__ b,
                                                  double* __restrict__ x, double* __restrict__ res) {
    const int T = blockIdx.x, s = threadIdx.x;
    const DenseSys d = sys[tseg[T]];
    const int i = T * kTile + s;
    const bool valid = s < tcnt[T];
    const double sa = valid ? sqrt(alpha[i]) : 0.0;
    const double4 e = E0[i];
    const double e0[3] = {e.x, e.y, e.z};
    for (int c = 0; c < 3; c++) {
        const size_t k = vec_index(d, T, s, c);
        const double v[4] = {sa * e0[c], c == 0 ? sa : 0.0, c == 1 ? sa : 0.0, c == 2 ? sa : 0.0};
        for (int q = 0; q < 4; q++) {
            b[k + q] = v[q];
            x[k + q] = 0.0;
            res[k + q] = v[q];
        }
    }
}

// x^ += y (systems still refining); mu_q = sqrt(alpha) x^_q for right-hand side q (slot order)
__global__ __launch_bounds__(64) void k_dense_update(const int* __restrict__ tseg, const DenseSys* __restrict__ sys,
                                                     const double* __restrict__ alpha, const double* __restrict__ y,
                                                     double* __restrict__ x, double4* __restrict__ mu4,
                                                     size_t mu_stride, const int* __restrict__ active) {
    const int T = blockIdx.x, s = threadIdx.x;
    const int g = tseg[T];
    const DenseSys d = sys[g];
    const int i = T * kTile + s;
    const double sa = sqrt(alpha[i]);
    double m[4][3];
    for (int c = 0; c < 3; c++) {
        const size_t k = vec_index(d, T, s, c);
        for (int q = 0; q < 4; q++) {
            double v = x[k + q];
            if (active[g]) {
                v += y[k + q];
                x[k + q] = v;
            }
            m[q][c] = sa * v;
        }
    }
    for (int q = 0; q < 4; q++) mu4[q * mu_stride + i] = make_double4(m[q][0], m[q][1], m[q][2], 0.0);
}

// r^_q = rhs^_q - x^_q + sqrt(alpha) (T mu_q) for one right-hand side q; per-tile partial sums of
// |r^_q|^2 and |rhs^_q|^2.
__global__ __launch_bounds__(64) void k_dense_residual(int q, const int* __restrict__ tseg,
                                                       const DenseSys* __restrict__ sys, const double* __restrict__ alpha,
                                                       const double4* __restrict__ Tmu, const double* __restrict__ b,
                                                       const double* __restrict__ x, double* __restrict__ res,
                                                       double* __restrict__ red, int ntile) {
    const int T = blockIdx.x, s = threadIdx.x;
    const DenseSys d = sys[tseg[T]];
    const int i = T * kTile + s;
    const double sa = sqrt(alpha[i]);
    const double4 f = Tmu[i];
    const double fv[3] = {f.x, f.y, f.z};
    double rr = 0.0, bb = 0.0;
    for (int c = 0; c < 3; c++) {
        const size_t k = vec_index(d, T, s, c) + q;
        const double r = b[k] - x[k] + sa * fv[c];
        res[k] = r;
        rr += r * r;
        bb += b[k] * b[k];
    }
    rr = wave_sum(rr);
    bb = wave_sum(bb);
    if (s == 0) {
        red[T] = rr;
        red[ntile + T] = bb;
    }
}

// relative residual of system g for right-hand side q: out[4 g + q]
__global__ __launch_bounds__(64) void k_dense_norms(int q, const int4* __restrict__ seg, const double* __restrict__ red,
                                                    int ntile, double* __restrict__ out) {
    const int4 sg = seg[blockIdx.x];
    double a = 0.0, b = 0.0;
    for (int T = sg.x + (int)threadIdx.x; T < sg.y; T += kTile) {
        a += red[T];
        b += red[ntile + T];
    }
    a = wave_sum(a);
    b = wave_sum(b);
    if (threadIdx.x == 0) out[4 * blockIdx.x + q] = b > 0.0 ? sqrt(a / b) : sqrt(a);
}

}  // namespace

void launch_dense_assemble(hipStream_t s, bool bf16, int nsys, int max_nt, const DenseSys* sys, const int* tcnt,
                           const double4* posa, const double* alpha, const double* inva, const int* exc_start,
                           const int* exc_J, const unsigned long long* exc_mask, unsigned short* mat) {
    dim3 g(max_nt, max_nt, nsys);
    if (bf16)
        hipLaunchKernelGGL(k_dense_assemble<true>, g, dim3(256), 0, s, sys, tcnt, posa, alpha, inva, exc_start, exc_J,
                           exc_mask, mat);
    else
        hipLaunchKernelGGL(k_dense_assemble<false>, g, dim3(256), 0, s, sys, tcnt, posa, alpha, inva, exc_start, exc_J,
                           exc_mask, mat);
}

void launch_dense_factor(hipStream_t s, bool bf16, int nsys, int max_nt, const DenseSys* sys, unsigned short* mat,
                         float* dinv, float* dacc, int* status) {
    const int nb = 3 * max_nt;
    for (int j = 0; j < nb; j++) {
        if (bf16) {
            hipLaunchKernelGGL(k_chol_diag<true>, dim3(nsys), dim3(256), 0, s, j, sys, mat, dacc, dinv, status);
            if (j + 1 < nb)
                hipLaunchKernelGGL(k_chol_col<true>, dim3(nb - j - 1, nsys), dim3(256), 0, s, j, sys, mat, dinv, dacc);
        } else {
            hipLaunchKernelGGL(k_chol_diag<false>, dim3(nsys), dim3(256), 0, s, j, sys, mat, dacc, dinv, status);
            if (j + 1 < nb)
                hipLaunchKernelGGL(k_chol_col<false>, dim3(nb - j - 1, nsys), dim3(256), 0, s, j, sys, mat, dinv,
                                   dacc);
        }
    }
}

void launch_dense_solve(hipStream_t s, bool bf16, int nsys, const DenseSys* sys, const unsigned short* mat,
                        const float* dinv, const double* rhs, double* y, const int* active) {
    if (bf16)
        hipLaunchKernelGGL(k_dense_solve<true>, dim3(nsys), dim3(kSolveThreads), 0, s, sys, mat, dinv, rhs, y, active);
    else
        hipLaunchKernelGGL(k_dense_solve<false>, dim3(nsys), dim3(kSolveThreads), 0, s, sys, mat, dinv, rhs, y, active);
}

void launch_dense_rhs(hipStream_t s, const SegLaunch& G, const DenseSys* sys, const int* tcnt, const double4* E0,
                      const double* alpha, double* b, double* x, double* res) {
    hipLaunchKernelGGL(k_dense_rhs, dim3(G.ntile), dim3(kTile), 0, s, G.tseg, sys, tcnt, E0, alpha, b, x, res);
}

void launch_dense_update(hipStream_t s, const SegLaunch& G, const DenseSys* sys, const double* alpha, const double* y,
                         double* x, double4* mu4, size_t mu_stride, const int* active) {
    hipLaunchKernelGGL(k_dense_update, dim3(G.ntile), dim3(kTile), 0, s, G.tseg, sys, alpha, y, x, mu4, mu_stride,
                       active);
}

void launch_dense_residual(hipStream_t s, const SegLaunch& G, int q, const DenseSys* sys, const double* alpha,
                           const double4* Tmu, const double* b, const double* x, double* res, double* red,
                           double* rel) {
    hipLaunchKernelGGL(k_dense_residual, dim3(G.ntile), dim3(kTile), 0, s, q, G.tseg, sys, alpha, Tmu, b, x, res, red,
                       G.ntile);
    hipLaunchKernelGGL(k_dense_norms, dim3(G.nseg), dim3(kTile), 0, s, q, G.seg, (const double*)red, G.ntile, rel);
}

int dense_max_tiles() { return kMaxDenseTiles; }

// E0 = unit field along axis at every particle slot (padding slots 0)
__global__ void k_uniform_field(int npad, const int* __restrict__ perm, int axis, double4* __restrict__ E0) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npad) return;
    const double v = perm[i] >= 0 ? 1.0 : 0.0;
    E0[i] = make_double4(axis == 0 ? v : 0.0, axis == 1 ? v : 0.0, axis == 2 ? v : 0.0, 0.0);
}

void launch_uniform_field(hipStream_t s, int npad, const int* perm, int axis, double4* E0) {
    hipLaunchKernelGGL(k_uniform_field, dim3((npad + 255) / 256), dim3(256), 0, s, npad, perm, axis, E0);
}

}  // namespace tdp
