// tdp_ewald.hip -- the Ewald option of the periodic method on gfx950: charges and induced dipoles
// with Thole damping summed over all periodic images.
//
// The reference's PME method is the minimum image over all pairs
// (ReferenceTholeDipoleForce.cpp:1375-1423); this option keeps its pair terms (fixed field :678-733,
// induced field :766-818, Thole damping :646-676, scale maps :609-644) and sums them by Ewald
// splitting instead:
//   real space     pairs whose minimum-image distance is below the cutoff, over neighbour tiles:
//                  every tile row I owns its particles' sums (full neighbour list, no j-side
//                  partials), four waves share a row and add their sums in a fixed order;
//   reciprocal     smooth PME of order 5: charges and dipoles spread into a fixed-point int64 grid
//                  (scale 2^40, order-independent and therefore deterministic), a 3D FFT by
//                  Stockham passes in LDS, the influence function folded into the x pass, and the
//                  potential gathered with B-spline derivatives in the epilogue of the real-space
//                  kernel;
//   corrections    the Ewald self-field and self-energy, and the pairs with a scale below 1 (masked
//                  out of the neighbour pass) evaluated per slot in FP64 from their own list with
//                  (scale * lambda - 1) times the bare kernel added to the screened one.
// Screened radial functions B_0 = erfc(a r)/r, B_n = ((2n-1) B_{n-1} + (2a^2)^n/(a sqrt(pi))
// e^{-a^2 r^2}) / r^2; damped kernels C_n = B_n + (lambda_{2n+1} - 1) b_n with b = 1/r, 1/r^3, 3/r^5,
// 15/r^7 and the reference's Thole factors lambda3 = 1 - e^s, lambda5 = 1 - (1-s) e^s plus
// lambda7 = 1 - (1 - s + 3/5 s^2) e^s for the gradient of the dipole field.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <type_traits>

#include "tdp_device.hpp"
#include "tdp_kernels.h"

namespace tdp {

namespace {

constexpr double kFix = 1099511627776.0;     // 2^40: fixed-point scale of the spreading grid
constexpr double kInvFix = 1.0 / 1099511627776.0;

// ------------------------------------------------------------------ exact FP64 pair geometry
// The neighbour criterion is evaluated with correctly rounded operations in a fixed order and
// no contraction, so that the host can reproduce it bit for bit: d = x_i - x_j, then per axis
// d -= L * floor(d / L + 0.5), r2 = (dx^2 + dy^2) + dz^2, inside = r2 < rc^2.
__device__ __forceinline__ double ew_image(double d, double L) {
    return __dsub_rn(d, __dmul_rn(L, floor(__dadd_rn(__ddiv_rn(d, L), 0.5))));
}
__device__ __forceinline__ double ew_r2(double dx, double dy, double dz) {
    return __dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz));
}

// B-spline weights of order 5 for grid coordinate u: th[j] = M_5(w + 4 - j) at grid point
// floor(u) - 4 + j, first and second derivatives with respect to u.
template <typename T>
__device__ __forceinline__ void ew_bspline(T w, T th[kEwOrder], T d1[kEwOrder], T d2[kEwOrder]) {
    T m1[kEwOrder], m2[kEwOrder], m3[kEwOrder];
    // levels 2..5 of M_n(x) = (x M_{n-1}(x) + (n - x) M_{n-1}(x-1)) / (n-1), index j <-> M_n(w + n-1-j)
    m1[0] = T(1) - w;   // n = 2
    m1[1] = w;
    // n = 3
    m2[0] = T(0.5) * (T(1) - w) * m1[0];
    m2[1] = T(0.5) * ((w + T(1)) * m1[0] + (T(2) - w) * m1[1]);
    m2[2] = T(0.5) * w * m1[1];
    // n = 4
    const T t3 = T(1) / T(3);
    m3[0] = t3 * (T(1) - w) * m2[0];
    m3[1] = t3 * ((w + T(2)) * m2[0] + (T(2) - w) * m2[1]);
    m3[2] = t3 * ((w + T(1)) * m2[1] + (T(3) - w) * m2[2]);
    m3[3] = t3 * w * m2[2];
    // n = 5
    th[0] = T(0.25) * (T(1) - w) * m3[0];
    th[1] = T(0.25) * ((w + T(3)) * m3[0] + (T(2) - w) * m3[1]);
    th[2] = T(0.25) * ((w + T(2)) * m3[1] + (T(3) - w) * m3[2]);
    th[3] = T(0.25) * ((w + T(1)) * m3[2] + (T(4) - w) * m3[3]);
    th[4] = T(0.25) * w * m3[3];
    // M_5' = M_4(x) - M_4(x-1): d1[j] = m3[j-1] - m3[j]
    d1[0] = -m3[0];
    d1[1] = m3[0] - m3[1];
    d1[2] = m3[1] - m3[2];
    d1[3] = m3[2] - m3[3];
    d1[4] = m3[3];
    // M_5'' = M_3(x) - 2 M_3(x-1) + M_3(x-2): d2[j] = m2[j-2] - 2 m2[j-1] + m2[j]
    d2[0] = m2[0];
    d2[1] = m2[1] - T(2) * m2[0];
    d2[2] = m2[2] - T(2) * m2[1] + m2[0];
    d2[3] = -T(2) * m2[2] + m2[1];
    d2[4] = m2[2];
}

// Grid coordinate of one axis: u = K * frac(x / L) in [0, K), base grid index floor(u) - 4.
__device__ __forceinline__ void ew_grid_coord(double x, double L, int K, int& base, double& w) {
    double f = x / L;
    f -= floor(f);
    const double u = f * K;
    int iu = (int)floor(u);
    w = u - iu;
    if (iu >= K) iu -= K;
    if (!(iu >= 0 && iu < K)) {   // a coordinate that is not finite: keep the indices in the grid
        iu = 0;
        w = 0.0;
    }
    base = iu - (kEwOrder - 1);
}

__device__ __forceinline__ int ew_wrap(int g, int K) { return g < 0 ? g + K : (g >= K ? g - K : g); }

template <typename T> struct Cplx;
template <> struct Cplx<double> { using type = double2; };
template <> struct Cplx<float> { using type = float2; };

template <typename C>
__device__ __forceinline__ C cmul(C a, C b) {
    C r;
    r.x = a.x * b.x - a.y * b.y;
    r.y = a.x * b.y + a.y * b.x;
    return r;
}

template <typename T>
__device__ __forceinline__ typename Cplx<T>::type twiddle(const EwAxis& A, int q, bool inverse);
template <>
__device__ __forceinline__ double2 twiddle<double>(const EwAxis& A, int q, bool inverse) {
    double2 w = A.tw64[q];
    if (inverse) w.y = -w.y;
    return w;
}
template <>
__device__ __forceinline__ float2 twiddle<float>(const EwAxis& A, int q, bool inverse) {
    float2 w = A.tw32[q];
    if (inverse) w.y = -w.y;
    return w;
}

}  // namespace

// ------------------------------------------------------------------ neighbour tiles
// Tile row I lists every tile J whose bounding box (minimum image of the centres, FP32 with a
// margin) comes within the cutoff of I's: a superset of the tiles holding a pair within the cutoff;
// beside each, the tile pair's entry in the exclusion tiles (or -1), so that the pair passes need no
// search.
__global__ __launch_bounds__(64) void k_ew_neighbors(EwLaunch E) {
    const int I = blockIdx.x, lane = threadIdx.x;
    const float4 cI = E.tb_c[I], hI = E.tb_h[I];
    const float bx = (float)E.L[0], by = (float)E.L[1], bz = (float)E.L[2];
    int count = 0;
    for (int J0 = 0; J0 < E.nblk; J0 += kTile) {
        const int J = J0 + lane;
        bool ok = false;
        if (J < E.nblk && E.tcnt[J] > 0) {
            const float4 cJ = E.tb_c[J], hJ = E.tb_h[J];
            float dx = cJ.x - cI.x, dy = cJ.y - cI.y, dz = cJ.z - cI.z;
            dx -= bx * rintf(dx / bx);
            dy -= by * rintf(dy / by);
            dz -= bz * rintf(dz / bz);
            dx = fmaxf(0.f, fabsf(dx) - hI.x - hJ.x);
            dy = fmaxf(0.f, fabsf(dy) - hI.y - hJ.y);
            dz = fmaxf(0.f, fabsf(dz) - hI.z - hJ.z);
            ok = dx * dx + dy * dy + dz * dz < E.cut2_tiles;
        }
        const unsigned long long m = __builtin_amdgcn_ballot_w64(ok);
        const int before = __popcll(m & ((1ull << lane) - 1ull));
        if (ok) {
            int ex = -1;   // the tile pair's entry in the exclusion tiles of row I, if any
            for (int e = E.exc_start[I]; e < E.exc_start[I + 1]; e++)
                if (E.exc_J[e] == J) ex = e;
            E.nbr[(size_t)I * E.nblk + count + before] = J;
            E.nbr_exc[(size_t)I * E.nblk + count + before] = ex;
        }
        count += __popcll(m);
    }
    if (lane == 0) E.nbr_cnt[I] = count;
}

// ------------------------------------------------------------------ spreading
// One workgroup per tile: each slot's contribution q theta_x theta_y theta_z + sum_d v_d (K_d/L_d)
// theta'_d ... is added in fixed point (an associative integer sum, so the grid does not depend on
// the order of the additions) into an LDS box of kSpB^3 grid points around the tile, and the box's
// nonzero points are then added to the global grid (a compact tile touches about 13^3 points, so
// the global atomics fall from 125 per slot to a fifth of that).  A tile whose points do not fit in
// the box adds its contributions to the global grid directly.
constexpr int kSpB = 16;

__global__ __launch_bounds__(64) void k_ew_spread(EwLaunch E, const double4* __restrict__ posa,
                                                  const double* __restrict__ charge, const double4* __restrict__ vec,
                                                  const double* __restrict__ scalars) {
    if (scalars && scalars[SC_DONE] != 0.0) return;
    __shared__ unsigned long long box[kSpB * kSpB * kSpB];
    const int T = blockIdx.x, lane = threadIdx.x;
    const int i = T * kTile + lane;
    const double q = charge ? charge[i] : 0.0;
    const double4 v = vec ? vec[i] : make_double4(0.0, 0.0, 0.0, 0.0);
    const bool live = lane < E.tcnt[T] && !(q == 0.0 && v.x == 0.0 && v.y == 0.0 && v.z == 0.0);
    const double4 p = posa[i];
    const double x[3] = {p.x, p.y, p.z};
    int base[3], rel[3], ref[3], lo[3], ext[3];
    double th[3][kEwOrder], d1[3][kEwOrder], d2[kEwOrder];
    bool fits = true;
    for (int k = 0; k < 3; k++) {
        double w;
        ew_grid_coord(x[k], E.L[k], E.K[k], base[k], w);
        ew_bspline<double>(w, th[k], d1[k], d2);
        // offsets from the first slot's base index, taken across the periodic wrap
        ref[k] = __shfl(base[k], 0, 64);
        int r = base[k] - ref[k];
        if (r > E.K[k] / 2) r -= E.K[k];
        if (r < -E.K[k] / 2) r += E.K[k];
        rel[k] = r;
        int mn = live ? r : INT_MAX, mx = live ? r : INT_MIN;
        for (int off = 32; off > 0; off >>= 1) {
            mn = min(mn, __shfl_xor(mn, off, 64));
            mx = max(mx, __shfl_xor(mx, off, 64));
        }
        lo[k] = mn;
        ext[k] = mx - mn + kEwOrder;
        if (mn == INT_MAX || ext[k] > kSpB || ext[k] > E.K[k]) fits = false;
    }
    const double vx = v.x * E.K[0] / E.L[0], vy = v.y * E.K[1] / E.L[1], vz = v.z * E.K[2] / E.L[2];
    if (!fits) {   // (an empty tile also lands here and adds nothing)
        if (!live) return;
        for (int a = 0; a < kEwOrder; a++) {
            const int gx = ew_wrap(base[0] + a, E.K[0]);
            for (int b = 0; b < kEwOrder; b++) {
                const int gy = ew_wrap(base[1] + b, E.K[1]);
                const double txy = th[0][a] * th[1][b];
                const double cq = q * txy + vx * d1[0][a] * th[1][b] + vy * th[0][a] * d1[1][b];
                const double cz = vz * txy;
                unsigned long long* row = E.qfix + ((size_t)gx * E.K[1] + gy) * E.K[2];
                for (int c = 0; c < kEwOrder; c++) {
                    const int gz = ew_wrap(base[2] + c, E.K[2]);
                    atomicAdd(row + gz, (unsigned long long)__double2ll_rn((cq * th[2][c] + cz * d1[2][c]) * kFix));
                }
            }
        }
        return;
    }
    const int ncell = ext[0] * ext[1] * ext[2];
    for (int c = lane; c < ncell; c += kTile) box[c] = 0ull;
    __syncthreads();
    if (live) {
        const int o0 = rel[0] - lo[0], o1 = rel[1] - lo[1], o2 = rel[2] - lo[2];
        for (int a = 0; a < kEwOrder; a++)
            for (int b = 0; b < kEwOrder; b++) {
                const double txy = th[0][a] * th[1][b];
                const double cq = q * txy + vx * d1[0][a] * th[1][b] + vy * th[0][a] * d1[1][b];
                const double cz = vz * txy;
                unsigned long long* row = box + ((o0 + a) * ext[1] + (o1 + b)) * ext[2] + o2;
                for (int c = 0; c < kEwOrder; c++)
                    atomicAdd(row + c, (unsigned long long)__double2ll_rn((cq * th[2][c] + cz * d1[2][c]) * kFix));
            }
    }
    __syncthreads();
    for (int c = lane; c < ncell; c += kTile) {
        const unsigned long long val = box[c];
        if (!val) continue;
        const int cz = c % ext[2], cy = (c / ext[2]) % ext[1], cx = c / (ext[2] * ext[1]);
        int g[3] = {ref[0] + lo[0] + cx, ref[1] + lo[1] + cy, ref[2] + lo[2] + cz};
        for (int k = 0; k < 3; k++) g[k] = ((g[k] % E.K[k]) + E.K[k]) % E.K[k];
        atomicAdd(E.qfix + ((size_t)g[0] * E.K[1] + g[1]) * E.K[2] + g[2], val);
    }
}

// ------------------------------------------------------------------ FFT passes
// Batched 1D FFTs along one axis of the K0 x K1 x K2 grid (z fastest), LB lines per workgroup in
// LDS, Stockham radix stages (autosorting, ping-pong buffers).  Stage with radix R after the
// stages so far of total span Ns: butterfly j takes x[j + r n/R] times w^(r (j mod Ns) n/(Ns R)),
// forms the R-point DFT and writes it to (j div Ns) Ns R + (j mod Ns) + k Ns.
// Pass kinds: 0 z forward from the fixed-point grid (which it clears), 1 y forward, 2 x forward,
// influence function, x inverse, 3 y inverse, 4 z inverse to the real potential grid.  The grids in
// real space are real, so the complex grid holds only kz = 0 .. K2/2 (K2h values, [K0][K1][K2h]):
// kind 0 keeps that half of each z transform, kinds 1-3 run on K2h columns, and kind 4 completes
// each z line by X(K2 - kz) = conj(X(kz)) (the spectrum stays Hermitian: the influence function is
// even) before its inverse transform.
// One radix-R stage of the lines in LDS (R known at compile time, so the butterfly stays in registers).
template <typename T, int R>
__device__ __forceinline__ void ew_fft_stage(const typename Cplx<T>::type* __restrict__ src,
                                             typename Cplx<T>::type* __restrict__ dst, int LB, int n, int Ns,
                                             const EwAxis& A, bool inverse) {
    using C = typename Cplx<T>::type;
    const int nb = n / R;
    const int span = n / (Ns * R);
    for (int idx = threadIdx.x; idx < LB * nb; idx += blockDim.x) {
        const int line = idx / nb, j = idx - line * nb;
        const C* x = src + line * n;
        C* y = dst + line * n;
        const int k = j % Ns;
        const int tstep = k * span;
        C v[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            C t = x[j + r * nb];
            if (r > 0 && tstep) t = cmul(t, twiddle<T>(A, (r * tstep) % n, inverse));
            v[r] = t;
        }
        const int base = (j / Ns) * Ns * R + k;
#pragma unroll
        for (int kk = 0; kk < R; kk++) {
            C acc = v[0];
#pragma unroll
            for (int r = 1; r < R; r++) {
                const C t = cmul(v[r], twiddle<T>(A, ((r * kk) % R) * nb, inverse));
                acc.x += t.x;
                acc.y += t.y;
            }
            y[base + kk * Ns] = acc;
        }
    }
}

template <typename T>
__device__ void ew_fft_lines(typename Cplx<T>::type* a, typename Cplx<T>::type* b, int LB, const EwAxis& A,
                             bool inverse, typename Cplx<T>::type*& out) {
    using C = typename Cplx<T>::type;
    const int n = A.n;
    int Ns = 1;
    C* src = a;
    C* dst = b;
    for (int st = 0; st < A.nst; st++) {
        const int R = A.r[st];
        switch (R) {
            case 2: ew_fft_stage<T, 2>(src, dst, LB, n, Ns, A, inverse); break;
            case 3: ew_fft_stage<T, 3>(src, dst, LB, n, Ns, A, inverse); break;
            case 4: ew_fft_stage<T, 4>(src, dst, LB, n, Ns, A, inverse); break;
            case 5: ew_fft_stage<T, 5>(src, dst, LB, n, Ns, A, inverse); break;
            default: ew_fft_stage<T, 7>(src, dst, LB, n, Ns, A, inverse); break;
        }
        __syncthreads();
        C* t = src;
        src = dst;
        dst = t;
        Ns *= R;
    }
    out = src;
}

template <typename T, int KIND>
__global__ __launch_bounds__(256) void k_ew_fft(EwLaunch E, int LB, void* phi_out, const double* __restrict__ scalars) {
    using C = typename Cplx<T>::type;
    if (scalars && scalars[SC_DONE] != 0.0) return;
    extern __shared__ unsigned char smem[];
    const int axis = (KIND == 0 || KIND == 4) ? 2 : ((KIND == 1 || KIND == 3) ? 1 : 0);
    const EwAxis& A = E.ax[axis];
    const int n = A.n;
    const int K1 = E.K[1], K2 = E.K[2], K2h = K2 / 2 + 1;
    const int nlines = axis == 2 ? E.K[0] * K1 : (axis == 1 ? E.K[0] * K2h : K1 * K2h);
    const size_t estride = axis == 2 ? 1 : (axis == 1 ? (size_t)K2h : (size_t)K1 * K2h);
    C* bufA = reinterpret_cast<C*>(smem);
    C* bufB = bufA + LB * n;
    C* grid = reinterpret_cast<C*>(E.cgrid);
    const int L0 = blockIdx.x * LB;
    const int lb = min(LB, nlines - L0);
    auto line_base = [&](int Lidx) -> size_t {   // in the complex grid (z lines: their first K2h values)
        if (axis == 2) return (size_t)Lidx * K2h;
        if (axis == 1) return (size_t)(Lidx / K2h) * K1 * K2h + (Lidx % K2h);
        return (size_t)Lidx;
    };
    // load (z lines: elements of a line are contiguous; y and x lines: consecutive lines are)
    for (int idx = threadIdx.x; idx < lb * n; idx += blockDim.x) {
        int b, e;
        if (axis == 2) { b = idx / n; e = idx - b * n; }
        else { e = idx / lb; b = idx - e * lb; }
        C v;
        if (KIND == 0) {
            const size_t g = (size_t)(L0 + b) * K2 + e;   // the real grid
            const long long f = (long long)E.qfix[g];
            v.x = (T)((double)f * kInvFix);
            v.y = T(0);
            E.qfix[g] = 0ull;
        } else if (KIND == 4) {   // the z line from its kz <= K2/2 half
            const size_t lb0 = line_base(L0 + b);
            if (e < K2h) {
                v = grid[lb0 + e];
            } else {
                v = grid[lb0 + (K2 - e)];
                v.y = -v.y;
            }
        } else {
            v = grid[line_base(L0 + b) + (size_t)e * estride];
        }
        bufA[b * n + e] = v;
    }
    __syncthreads();
    C* res;
    ew_fft_lines<T>(bufA, bufB, lb, A, KIND >= 3, res);
    if (KIND == 2) {   // influence function ([K0][K1][K2]), then the inverse transform along x
        for (int idx = threadIdx.x; idx < lb * n; idx += blockDim.x) {
            const int e = idx / lb, b = idx - e * lb;
            const int Lb = L0 + b;   // = ky K2h + kz
            const size_t g = ((size_t)e * K1 + Lb / K2h) * K2 + Lb % K2h;
            const T gf = E.mixed ? (T)E.G32[g] : (T)E.G64[g];
            C v = res[b * n + e];
            v.x *= gf;
            v.y *= gf;
            res[b * n + e] = v;
        }
        __syncthreads();
        C* other = res == bufA ? bufB : bufA;
        ew_fft_lines<T>(res, other, lb, A, true, res);
    }
    for (int idx = threadIdx.x; idx < lb * n; idx += blockDim.x) {
        int b, e;
        if (axis == 2) { b = idx / n; e = idx - b * n; }
        else { e = idx / lb; b = idx - e * lb; }
        if (KIND == 4) reinterpret_cast<T*>(phi_out)[(size_t)(L0 + b) * K2 + e] = res[b * n + e].x;
        else if (KIND == 0) {
            if (e < K2h) grid[line_base(L0 + b) + e] = res[b * n + e];
        } else {
            grid[line_base(L0 + b) + (size_t)e * estride] = res[b * n + e];
        }
    }
}

// ------------------------------------------------------------------ reciprocal gather
// Potential derivatives at slot x from a potential grid: g1 = grad_u phi (3), g2 = second
// derivatives in u (xx, yy, zz, xy, xz, yz), phi itself.  NEED: 0 phi and g1, 1 also g2.
template <typename T, int NEED>
__device__ __forceinline__ void ew_gather(const EwLaunch& E, const T* __restrict__ phi, double4 p, double& ph,
                                          double g1[3], double g2[6], int a0 = 0, int a1 = kEwOrder) {
    int base[3];
    double th[3][kEwOrder], d1[3][kEwOrder], d2[3][kEwOrder];
    const double x[3] = {p.x, p.y, p.z};
    for (int k = 0; k < 3; k++) {
        double w;
        ew_grid_coord(x[k], E.L[k], E.K[k], base[k], w);
        ew_bspline<double>(w, th[k], d1[k], d2[k]);
    }
    ph = 0.0;
    for (int k = 0; k < 3; k++) g1[k] = 0.0;
    for (int k = 0; k < 6; k++) g2[k] = 0.0;
#pragma unroll 1
    for (int a = a0; a < a1; a++) {   // x planes a0 .. a1-1 of the spline's support
        const int gx = ew_wrap(base[0] + a, E.K[0]);
#pragma unroll 1
        for (int b = 0; b < kEwOrder; b++) {
            const int gy = ew_wrap(base[1] + b, E.K[1]);
            const T* row = phi + ((size_t)gx * E.K[1] + gy) * E.K[2];
            double s0 = 0.0, s1 = 0.0, s2 = 0.0;   // sum_c v th_c, v d1_c, v d2_c
            for (int c = 0; c < kEwOrder; c++) {
                const double v = (double)row[ew_wrap(base[2] + c, E.K[2])];
                s0 += v * th[2][c];
                s1 += v * d1[2][c];
                if (NEED) s2 += v * d2[2][c];
            }
            const double txy = th[0][a] * th[1][b];
            ph += txy * s0;
            g1[0] += d1[0][a] * th[1][b] * s0;
            g1[1] += th[0][a] * d1[1][b] * s0;
            g1[2] += txy * s1;
            if (NEED) {
                g2[0] += d2[0][a] * th[1][b] * s0;
                g2[1] += th[0][a] * d2[1][b] * s0;
                g2[2] += txy * s2;
                g2[3] += d1[0][a] * d1[1][b] * s0;
                g2[4] += d1[0][a] * th[1][b] * s1;
                g2[5] += th[0][a] * d1[1][b] * s1;
            }
        }
    }
}

// ------------------------------------------------------------------ real space
// Minimum image with the bit-exact result of d - L floor(d / L + 0.5): floor(d (1/L) + 0.5) is the
// same integer unless d/L + 0.5 lies within 1e-12 of one, where the correctly rounded division
// decides.
__device__ __forceinline__ double ew_image_fast(double d, double L, double iL) {
    const double t = d * iL + 0.5;
    double n = floor(t);
    if (fabs(t - rint(t)) < 1e-12) n = floor(__dadd_rn(__ddiv_rn(d, L), 0.5));
    return __dsub_rn(d, __dmul_rn(L, n));
}

// Screened and damped radial coefficients of one pair (B_0..B_3, bare b_1..b_3, Thole factors).
template <typename T>
struct EwR {
    T B0, B1, B2, B3, b1, b2, b3, l3, l5, l7;
};

// erfcx(x) = e^{x^2} erfc(x) on [0, 6] as P(x) / Q(x), degree 9, coefficients in ascending powers
// from fit_erfcx.py (relative error of the fit 2.3e-18; below 1e-15 evaluated in FP64).
constexpr double kInvSqrtPi = 0.56418958354775628;   // 1 / sqrt(pi)
constexpr double kErfcxP[10] = {1, 1.8246330203534158, 1.6678312021636672, 0.95726832494803793,
                                0.37200178985934518, 0.099564333415988568, 0.017898050519465095,
                                0.0019808686982774721, 0.00010407600456896987, -3.7341872226394312e-13};
constexpr double kErfcxQ[10] = {1, 2.9530121874489286, 3.9999486346601589, 3.2699676243660476,
                                1.7832281212683818, 0.67512166134826335, 0.17822904124994046,
                                0.031815678241935146, 0.0035109996444155928, 0.00018446987284089452};

// erfc(x) for x >= 0 from e = e^{-x^2}, which the radial functions need anyway: e P(x) / Q(x) up to
// x = 6 (every pair inside the cutoff: a r_c is about 3, below 6 for any tolerance above 1e-15).
// Beyond, the asymptotic series e / (x sqrt(pi)) (1 - 1/(2x^2) + 3/(4x^4) - 15/(8x^6)), relative
// error below 105 / (16 x^8) < 4e-6 where erfc(x) < 2.2e-17: the kernel's terms there are below 1e-22
// of the bare ones.  (The math library's erfc in that branch would hold its registers in every caller
// and cost them a wave per SIMD.)
__device__ __forceinline__ double ew_erfc(double x, double e) {
    if (x <= 6.0) {
        double p = kErfcxP[9], q = kErfcxQ[9];
#pragma unroll
        for (int k = 8; k >= 0; k--) {
            p = fma(p, x, kErfcxP[k]);
            q = fma(q, x, kErfcxQ[k]);
        }
        return e * p / q;
    }
    const double t = 1.0 / (x * x);
    return e * (kInvSqrtPi / x) * (1.0 - t * (0.5 - t * (0.75 - 1.875 * t)));
}

// FP64: erfc through ew_erfc, the exponentials of the math library; THOLE32: the Thole exponential
// in FP32, for pairs whose exponent s is at most -4 (the mixed lists' FP32 pairs beyond r64): e^s
// <= e^-4 enters C_n as (lambda - 1) b_n = -e^s b_n, so the FP32 rounding of s and e^s (relative
// error below 1.5e-7 |s| + 1.2e-7) moves C_1 by less than 2e-8 b_1.
template <bool THOLE32 = false>
__device__ __forceinline__ void ew_radial(const EwLaunch& E, double r2, double ai, double aj, double iai, double iaj,
                                          bool inside, bool need_l7, EwR<double>& R) {
    const double ri = rsqrt(r2);
    const double r = r2 * ri;
    const double ri2 = ri * ri;
    R.b1 = ri * ri2;
    R.b2 = 3.0 * R.b1 * ri2;
    R.b3 = 5.0 * R.b2 * ri2;
    if (inside) {
        const double ex = exp(-E.alpha * E.alpha * r2);
        R.B0 = ew_erfc(E.alpha * r, ex) * ri;
        R.B1 = (R.B0 + E.c1 * ex) * ri2;
        R.B2 = (3.0 * R.B1 + E.c2 * ex) * ri2;
        R.B3 = (5.0 * R.B2 + E.c3 * ex) * ri2;
    } else {
        R.B0 = R.B1 = R.B2 = R.B3 = 0.0;
    }
    R.l3 = R.l5 = R.l7 = 1.0;
    const double g = ai < aj ? ai : aj;
    const double u = r * (iai * iaj);
    const double s = -g * (u * u * u);
    if (ai * aj != 0.0 && s > -50.0) {
        const double es = THOLE32 ? (double)__expf((float)s) : exp(s);
        R.l3 = 1.0 - es;
        R.l5 = 1.0 - (1.0 - s) * es;
        if (need_l7) R.l7 = 1.0 - (1.0 - s + 0.6 * s * s) * es;
    }
}

// FP32 (pairs inside the cutoff): erfc(x) by Abramowitz & Stegun 7.1.26 (Hastings),
// t = 1/(1 + p x), erfc = t (a1 + t (a2 + t (a3 + t (a4 + t a5)))) e^{-x^2}, absolute error below
// 1.5e-7, sharing e^{-x^2} with B_1..B_3; the Thole exponential only where some lane is damped.
__device__ __forceinline__ void ew_radial32(float alpha, float c1, float c2, float c3, float r2, float ai, float aj,
                                            float iai, float iaj, bool need_l7, EwR<float>& R) {
    const float ri = __builtin_amdgcn_rsqf(r2);
    const float r = r2 * ri;
    const float ri2 = ri * ri;
    R.b1 = ri * ri2;
    R.b2 = 3.f * R.b1 * ri2;
    R.b3 = 5.f * R.b2 * ri2;
    const float x = alpha * r;
    const float ex = expf(-x * x);
    const float t = __builtin_amdgcn_rcpf(1.f + 0.3275911f * x);
    const float poly = t * (0.254829592f + t * (-0.284496736f + t * (1.421413741f + t * (-1.453152027f + t * 1.061405429f))));
    R.B0 = poly * ex * ri;
    R.B1 = (R.B0 + c1 * ex) * ri2;
    R.B2 = (3.f * R.B1 + c2 * ex) * ri2;
    R.B3 = (5.f * R.B2 + c3 * ex) * ri2;
    R.l3 = R.l5 = R.l7 = 1.f;
    const float g = ai < aj ? ai : aj;
    const float u = r * (iai * iaj);
    const float s = -g * (u * u * u);
    const bool damped = ai * aj != 0.f && s > -50.f;
    if (__builtin_amdgcn_ballot_w64(damped)) {
        const float es = damped ? __expf(s) : 0.f;
        R.l3 = 1.f - es;
        R.l5 = 1.f - (1.f - s) * es;
        if (need_l7) R.l7 = 1.f - (1.f - s + 0.6f * s * s) * es;
    }
}

// The real-space passes give wave w of a tile row the slots 8w + m (m = 0..7) of every neighbour
// tile (contiguous, so that a wave's slots of a tile are one 128-byte line of a float4 array).
__device__ __forceinline__ int ew_wave_slot(int w, int m) { return 8 * w + m; }
// Bits 8w..8w+7 of x (the slots of wave w in a tile's 64-bit mask) as bits 0-7.
__device__ __forceinline__ unsigned long long ew_slots_of_wave(unsigned long long x, int w) {
    return (x >> (8 * w)) & 0xffull;
}

// The i-side terms of one unscaled pair inside the cutoff (d = x_i - x_j): the field of charge qj
// (E0), the field of dipole vj (T v), or the force on i and half the pair's charge-charge energy.
template <typename T, int MODE, bool MUTUAL>
__device__ __forceinline__ void ew_pair_terms(const EwR<T>& R, T dx, T dy, T dz, T qi, T vix, T viy, T viz, T qj,
                                              T vjx, T vjy, T vjz, T& sx, T& sy, T& sz, T& se) {
    const T C1 = R.B1 + (R.l3 - T(1)) * R.b1;
    const T C2 = R.B2 + (R.l5 - T(1)) * R.b2;
    if (MODE == kEwE0) {
        const T f = qj * C1;
        sx += f * dx;
        sy += f * dy;
        sz += f * dz;
    } else if (MODE == kEwTmu) {
        const T vd = vjx * dx + vjy * dy + vjz * dz;
        sx += C2 * vd * dx - C1 * vjx;
        sy += C2 * vd * dy - C1 * vjy;
        sz += C2 * vd * dz - C1 * vjz;
    } else {
        // g = grad_d U; the force on i is -g
        const T qq = qi * qj * R.B1;
        T gx = -qq * dx, gy = -qq * dy, gz = -qq * dz;
        const T ux = qi * vjx - qj * vix, uy = qi * vjy - qj * viy, uz = qi * vjz - qj * viz;
        const T ud = ux * dx + uy * dy + uz * dz;
        gx += C1 * ux - C2 * ud * dx;
        gy += C1 * uy - C2 * ud * dy;
        gz += C1 * uz - C2 * ud * dz;
        if (MUTUAL) {
            const T C3 = R.B3 + (R.l7 - T(1)) * R.b3;
            const T mid = vix * dx + viy * dy + viz * dz;
            const T mjd = vjx * dx + vjy * dy + vjz * dz;
            const T mm = vix * vjx + viy * vjy + viz * vjz;
            const T cd = -C2 * mm + C3 * mid * mjd;
            gx += cd * dx - C2 * (vix * mjd + vjx * mid);
            gy += cd * dy - C2 * (viy * mjd + vjy * mid);
            gz += cd * dz - C2 * (viz * mjd + vjz * mid);
        }
        sx -= gx;
        sy -= gy;
        sz -= gz;
        se += T(0.5) * qi * qj * R.B0;
    }
}

// Row-owner real-space pass over the neighbour tiles of tile row I: NP workgroups of kEwWaves waves
// per row, part h taking the neighbour tiles h, h + NP, ... (the T v and force passes use two parts:
// twice as many workgroups as rows fill the CUs more evenly, one 8-wave workgroup per CU at these
// register counts; the E0 pass, which builds the stored lists, one, for longer lists).  Every wave
// visits every tile of its part and takes its 8 slots k = 8w + m, eight tiles at a time
// staged in the wave's own LDS slice, so that each wave sees an even spatial sample of the row's
// neighbourhood and the waves' stored lists have nearly equal lengths per slot (their padding to
// the longest is small).  Only the slots within the cutoff (plus a margin) of row I's bounding box
// are visited, and a step whose 64 pairs all lie beyond the cutoff is skipped.  MIXED: pairs beyond
// E.r64sq are evaluated in FP32 (group sums in FP32, added to the FP64 row sums); the steps that
// hold a closer pair are marked and evaluated afterwards in FP64 for those pairs.  Output: the
// part's real-space sums for the slot (field or force, and the charge-charge energy) in rs.
template <int MODE, bool MUTUAL, bool MIXED>
__global__ __launch_bounds__(kEwWaves * 64) void k_ew_pairs(EwLaunch E, const double4* __restrict__ posa,
                                                            const double* __restrict__ inva,
                                                            const double* __restrict__ charge,
                                                            const double4* __restrict__ vec,
                                                            const double* __restrict__ scalars,
                                                            double4* __restrict__ rs) {
    if (scalars && scalars[SC_DONE] != 0.0) return;
    constexpr int NP = ew_parts(MODE);
    const int I = blockIdx.x / NP, h = blockIdx.x % NP;
    const int lane = threadIdx.x & (kTile - 1), w = threadIdx.x / kTile;
    static_assert(kEwWaves * 8 == kTile, "each wave takes 8 slots of a tile");
    __shared__ double4 s_p[kEwWaves][kTile];
    __shared__ double4 s_v[kEwWaves][kTile];   // charge in w
    __shared__ double s_ia[kEwWaves][kTile];
    const int i = I * kTile + lane;
    const bool ivalid = lane < E.tcnt[I];
    const double4 pi = posa[i];
    const double iai = inva[i];
    const double qi = charge[i];
    const double4 vi = (MODE == kEwE0) ? make_double4(0.0, 0.0, 0.0, 0.0) : vec[i];
    const float4 cI = E.tb_c[I], hI = E.tb_h[I];
    const float bxf = (float)E.L[0], byf = (float)E.L[1], bzf = (float)E.L[2];
    const double iL0 = 1.0 / E.L[0], iL1 = 1.0 / E.L[1], iL2 = 1.0 / E.L[2];
    double ax = 0.0, ay = 0.0, az = 0.0, ae = 0.0;
    // this part's neighbour tiles: t = NP l + h of the row's list, l < cnt
    const int cnt = (E.nbr_cnt[I] - h + NP - 1) / NP;
    const int* nb = E.nbr + (size_t)I * E.nblk + h;
    const int* nbx = E.nbr_exc + (size_t)I * E.nblk + h;
    for (int t0 = 0; t0 < cnt; t0 += 8) {
        // a group of up to 8 neighbour tiles: lane 8g + m holds slot k = 8w + m of tile
        // t0 + g and tests it against row I's bounding box (cutoff plus margin)
        const int g8 = lane >> 3, m8 = lane & 7;
        int Jl = 0, exl = -1;
        double4 pl = make_double4(0.0, 0.0, 0.0, 0.0), vl = make_double4(0.0, 0.0, 0.0, 0.0);
        double ial = 0.0;
        bool near_box = false;
        if (t0 + g8 < cnt) {
            Jl = nb[NP * (t0 + g8)];
            exl = nbx[NP * (t0 + g8)];
            const int ks = ew_wave_slot(w, m8);
            if (ks < E.tcnt[Jl]) {
                const int j = Jl * kTile + ks;
                pl = posa[j];
                if (MODE != kEwE0) vl = vec[j];
                vl.w = charge[j];
                ial = inva[j];
                float dx = (float)pl.x - cI.x, dy = (float)pl.y - cI.y, dz = (float)pl.z - cI.z;
                dx -= bxf * rintf(dx / bxf);
                dy -= byf * rintf(dy / byf);
                dz -= bzf * rintf(dz / bzf);
                dx = fmaxf(0.f, fabsf(dx) - hI.x);
                dy = fmaxf(0.f, fabsf(dy) - hI.y);
                dz = fmaxf(0.f, fabsf(dz) - hI.z);
                near_box = dx * dx + dy * dy + dz * dz < E.cut2_tiles;
            }
        }
        const unsigned long long visit = __builtin_amdgcn_ballot_w64(near_box);
        if (!visit) continue;
        s_p[w][lane] = pl;
        s_v[w][lane] = vl;
        s_ia[w][lane] = ial;
        // the slice is this wave's own: its LDS operations complete in order
        __builtin_amdgcn_wave_barrier();
        // scaled pairs of these tile pairs are masked (the epilogue evaluates them).  Bit 8g + m of
        // the group's mask is slot 8w + m of tile t0 + g.
        unsigned long long mask = 0ull;
        for (int g = 0; g < 8; g++) {
            if (!((visit >> (8 * g)) & 0xffull)) continue;
            const int J = __builtin_amdgcn_readlane(Jl, 8 * g);
            const int e = __builtin_amdgcn_readlane(exl, 8 * g);
            if (e >= 0) {
                const unsigned long long* mm = E.exc_mask + (size_t)e * 3 * kTile;
                unsigned long long mk;
                if (MODE == kEwE0) {
                    mk = mm[lane] | mm[kTile + lane];
                } else if (MODE == kEwTmu) {
                    mk = mm[2 * kTile + lane];
                } else {
                    mk = mm[lane] | mm[kTile + lane] | mm[2 * kTile + lane];
                }
                mask |= ew_slots_of_wave(mk, w) << (8 * g);
            }
            if (J == I && (lane >> 3) == w) mask |= 1ull << (8 * g + (lane & 7));
        }
        if (!ivalid) mask = ~0ull;
        // Phase A (one step per visited slot, all lanes): the FP64 pair criterion, giving each lane
        // the bits of the pairs it evaluates.  E0: mB = the field's pairs; mixed T v / force:
        // mA = pairs beyond r64 (FP32), mB = closer pairs (FP64); FP64: mB.
        unsigned long long mA = 0ull, mB = 0ull;
        {
            unsigned long long v = visit;
            while (v) {
                const int q = __builtin_ctzll(v);
                v &= v - 1ull;
                const double4 pj = s_p[w][q];
                const double dx = ew_image_fast(__dsub_rn(pi.x, pj.x), E.L[0], iL0);
                const double dy = ew_image_fast(__dsub_rn(pi.y, pj.y), E.L[1], iL1);
                const double dz = ew_image_fast(__dsub_rn(pi.z, pj.z), E.L[2], iL2);
                const double r2x = ew_r2(dx, dy, dz);
                const unsigned long long bit = 1ull << q;
                if (r2x < E.rc2 && !(mask & bit)) {
                    if (MODE != kEwE0 && MIXED && r2x >= E.r64sq) mA |= bit;
                    else mB |= bit;
                }
            }
        }
        // Phase B: each lane runs through its own pairs (ascending q, the same order for every
        // pass), so that the radial functions are evaluated for pairs inside the cutoff only.
        if (MIXED) {
            float fx = 0.f, fy = 0.f, fz = 0.f, fe = 0.f;
            const float alpha32 = (float)E.alpha, c1f = (float)E.c1, c2f = (float)E.c2, c3f = (float)E.c3;
            unsigned long long v = mA;
            while (v) {
                const int q = __builtin_ctzll(v);
                v &= v - 1ull;
                const double4 pj = s_p[w][q];
                const double dx = ew_image_fast(__dsub_rn(pi.x, pj.x), E.L[0], iL0);
                const double dy = ew_image_fast(__dsub_rn(pi.y, pj.y), E.L[1], iL1);
                const double dz = ew_image_fast(__dsub_rn(pi.z, pj.z), E.L[2], iL2);
                const double4 vj = s_v[w][q];
                const float gx = (float)dx, gy = (float)dy, gz = (float)dz;
                EwR<float> R;
                ew_radial32(alpha32, c1f, c2f, c3f, gx * gx + gy * gy + gz * gz, (float)pi.w, (float)pj.w, (float)iai,
                            (float)s_ia[w][q], MODE == kEwForce && MUTUAL, R);
                ew_pair_terms<float, MODE, MUTUAL>(R, gx, gy, gz, (float)qi, (float)vi.x, (float)vi.y, (float)vi.z,
                                                   (float)vj.w, (float)vj.x, (float)vj.y, (float)vj.z, fx, fy, fz, fe);
            }
            ax += (double)fx;
            ay += (double)fy;
            az += (double)fz;
            ae += (double)fe;
        }
        double sx = 0.0, sy = 0.0, sz = 0.0, se = 0.0;
        unsigned long long v = mB;
        while (v) {
            const int q = __builtin_ctzll(v);
            v &= v - 1ull;
            const double4 pj = s_p[w][q];
            const double dx = ew_image_fast(__dsub_rn(pi.x, pj.x), E.L[0], iL0);
            const double dy = ew_image_fast(__dsub_rn(pi.y, pj.y), E.L[1], iL1);
            const double dz = ew_image_fast(__dsub_rn(pi.z, pj.z), E.L[2], iL2);
            const double r2x = ew_r2(dx, dy, dz);
            const double4 vj = s_v[w][q];
            EwR<double> R;
            ew_radial(E, r2x, pi.w, pj.w, iai, s_ia[w][q], true, MODE == kEwForce && MUTUAL, R);
            ew_pair_terms<double, MODE, MUTUAL>(R, dx, dy, dz, qi, vi.x, vi.y, vi.z, vj.w, vj.x, vj.y, vj.z, sx, sy, sz,
                                                se);
        }
        ax += sx;
        ay += sy;
        az += sz;
        ae += se;
        __builtin_amdgcn_wave_barrier();   // every lane has read the slice before the next group is staged
    }
    // the row's sums over the waves, in wave order (the staging slices are reused)
    __syncthreads();
    double* red = reinterpret_cast<double*>(&s_p[0][0]);   // [kEwWaves][4][kTile] doubles: 16 KB of the 32 KB
    red[(w * 4 + 0) * kTile + lane] = ax;
    red[(w * 4 + 1) * kTile + lane] = ay;
    red[(w * 4 + 2) * kTile + lane] = az;
    red[(w * 4 + 3) * kTile + lane] = ae;
    __syncthreads();
    if (w != 0) return;
    for (int u = 1; u < kEwWaves; u++) {
        ax += red[(u * 4 + 0) * kTile + lane];
        ay += red[(u * 4 + 1) * kTile + lane];
        az += red[(u * 4 + 2) * kTile + lane];
        ae += red[(u * 4 + 3) * kTile + lane];
    }
    rs[(size_t)h * E.npad + i] = make_double4(ax, ay, az, ae);
}

// The E0 pass when the blocks of T are stored, in two phases per wave (list (I, w)).
// Phase A visits, as k_ew_pairs does, the wave's slots of the row's neighbour tiles that lie near
// row I's box and records for each lane the pairs inside the cutoff as candidates 8 t + m (slot
// 8w + m of neighbour tile t) with two flags (in the field sum, stored as a block of T),
// in two streams: the pairs whose blocks are FP32 (beyond r64 in mixed precision) and the others
// (FP64).  A lane's candidates go to its own 32-byte blocks of the stream (8 per block, so that the
// lane's stores fill whole sectors).  Phase B runs through each lane's streams densely, every lane
// busy: the FP64 radial functions are evaluated for pairs inside the cutoff only (a step of
// k_ew_pairs keeps the lanes of its pairs beyond the cutoff idle), and the block of candidate n of
// a stream goes to row n of that stream's list, so that the 64 lanes store one row together (zero
// blocks where candidate n is not stored or lies past the lane's end).  The candidates and their
// counts stay for the force pass (k_ew_force_cand).
constexpr unsigned kCandSel = 1u << 16, kCandStore = 1u << 17;
constexpr int kCandTabShift = 18;   // FP32 candidates: bits 18-31 their slot's place in the list's slot table

// x, y, z wrapped into [0, L) in FP64, then rounded to FP32.
__device__ __forceinline__ float4 ew_wrap32(const double4& p, const double* L, double iL0, double iL1, double iL2) {
    return make_float4((float)(p.x - L[0] * floor(p.x * iL0)), (float)(p.y - L[1] * floor(p.y * iL1)),
                       (float)(p.z - L[2] * floor(p.z * iL2)), 0.f);
}

// Minimum image of the separation of two tile origins, FP64 (iL = 1/L; the same function serves the
// E0 pass's image test and the T v pass's offsets, so that both take the same image).
__device__ __forceinline__ double ew_tile_shift(double d, double L, double iL) { return d - L * floor(d * iL + 0.5); }

// True when every pair of tiles I and J within the cutoff takes the image of the tiles' origins:
// per axis |D| + hI + hJ < L - rc with D = o_J - o_I under its minimum image.  A pair's separation is
// D + u_j - u_i up to whole boxes with |u_j - u_i| <= hI + hJ; its minimum image is below rc and
// every other image of it lies beyond L - rc.
__device__ __forceinline__ bool ew_tile_image_ok(const EwLaunch& E, const double4& oI, const float4& hI,
                                                 const double4& oJ, const float4& hJ, double rc) {
    const double Dx = ew_tile_shift(oJ.x - oI.x, E.L[0], E.iL[0]);
    const double Dy = ew_tile_shift(oJ.y - oI.y, E.L[1], E.iL[1]);
    const double Dz = ew_tile_shift(oJ.z - oI.z, E.L[2], E.iL[2]);
    return fabs(Dx) + (double)hI.x + (double)hJ.x < E.L[0] - rc &&
           fabs(Dy) + (double)hI.y + (double)hJ.y < E.L[1] - rc &&
           fabs(Dz) + (double)hI.z + (double)hJ.z < E.L[2] - rc;
}

// Slot j of tile J as the FP32 T v pass of row I stages it: (v_j, y_x) and (y_y, y_z), y = u_j + D_IJ
// formed in FP64 and rounded once (the slot's offset from tile I's origin under the tile pair's image).
__device__ __forceinline__ void ew_slot32(const EwLaunch& E, const double4& oI, int J, int j,
                                          const float4* __restrict__ v32, float4& a, float2& b) {
    const double4 oJ = E.tb_o[J];
    const float4 u = E.posl[j];
    const float4 v = v32[j];
    a = make_float4(v.x, v.y, v.z, (float)((double)u.x + ew_tile_shift(oJ.x - oI.x, E.L[0], E.iL[0])));
    b = make_float2((float)((double)u.y + ew_tile_shift(oJ.y - oI.y, E.L[1], E.iL[1])),
                    (float)((double)u.z + ew_tile_shift(oJ.z - oI.z, E.L[2], E.iL[2])));
}

// The stored lists are jagged: row n of a list (entry n of every lane, or entry pair n of the FP32
// lists, whose entries 2n and 2n+1 of a lane are adjacent: a 4-byte index pair and a 16-byte
// coefficient pair) holds only the lanes whose list reaches it, packed in lane order, and the rows
// follow one another, each rounded up to whole 128-byte lines of the coefficient array (8 FP32 entry
// pairs, 4 FP64 entries).  The row's lanes are the set bits of ballot(n < rows of the lane); a lane's
// place is the number of set bits below it, the row's start the sum of the earlier rows' rounded
// counts.  A pass that walks the rows in order keeps that sum as it goes, so that no line is streamed
// for lanes past their list's end.  A list's rows lie in its region of the arrays, (cap32 / 2) 64
// pairs or cap64 64 entries from the list's base.
__device__ __forceinline__ size_t ew_base32(int list, int cap32) { return (size_t)list * (cap32 >> 1) * kTile; }
__device__ __forceinline__ size_t ew_base64(int list, int cap64) { return (size_t)list * cap64 * kTile; }
__device__ __forceinline__ int ew_row_slots32(unsigned long long m) { return ((int)__popcll(m) + 7) & ~7; }
__device__ __forceinline__ int ew_row_slots64(unsigned long long m) { return ((int)__popcll(m) + 3) & ~3; }
__device__ __forceinline__ int ew_lane_rank(unsigned long long m) {
    return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

__device__ __forceinline__ unsigned* ew_cand_base(unsigned* c, int list, int cap, int lane) {
    return c + (size_t)list * cap * kTile + lane * 8;   // candidate n at [(n / 8) 512 + n % 8]
}
// Candidate n of a lane's stream of nl entries; past the end its last entry, or entry 0 (which phase A
// writes for a lane without candidates).
__device__ __forceinline__ unsigned ew_cand_at(const unsigned* cb, int n, int nl) {
    const int k = min(n, max(nl, 1) - 1);
    return cb[(k >> 3) * 512 + (k & 7)];
}
// Its neighbour tile (index into the row's list).
__device__ __forceinline__ int ew_cand_tile(unsigned c) { return (int)((c & 0xffffu) >> 3); }
__device__ __forceinline__ int ew_cand_slot(const int* nb, unsigned c, int w) {
    const int loc = (int)(c & 0xffffu);
    return nb[loc >> 3] * kTile + ew_wave_slot(w, loc & 7);
}

template <bool MIXED>
__global__ __launch_bounds__(kEwE0Waves * 64) void k_ew_e0_build(EwLaunch E, const double4* __restrict__ posa,
                                                                 const double* __restrict__ inva,
                                                                 const double* __restrict__ charge,
                                                                 double4* __restrict__ rs) {
    // workgroup h of row I: waves w = kEwE0Waves h + wl (kEwWaves / kEwE0Waves workgroups per row, so
    // that several fit on a CU at this register count)
    const int I = blockIdx.x / kEwE0Parts, h = blockIdx.x % kEwE0Parts;
    const int lane = threadIdx.x & (kTile - 1), wl = threadIdx.x / kTile, w = kEwE0Waves * h + wl;
    __shared__ double4 s_p[kEwE0Waves][kTile];
    __shared__ float4 s_f[kEwE0Waves][kTile];   // the slots wrapped into the box, FP32
    __shared__ double red[kEwE0Waves * 3 * kTile];
    const int i = I * kTile + lane;
    const bool ivalid = lane < E.tcnt[I];
    const double4 pi = posa[i];
    const double iai = inva[i];
    const float4 cI = E.tb_c[I], hI = E.tb_h[I];
    const float bxf = (float)E.L[0], byf = (float)E.L[1], bzf = (float)E.L[2];
    const double iL0 = 1.0 / E.L[0], iL1 = 1.0 / E.L[1], iL2 = 1.0 / E.L[2];
    const int cnt = E.nbr_cnt[I];
    const int* nb = E.nbr + (size_t)I * E.nblk;
    const int* nbx = E.nbr_exc + (size_t)I * E.nblk;
    const int list = I * kEwWaves + w;
    // FP32 screen of the pair criterion on coordinates wrapped into the box.  Each separation
    // component is off by at most 4 u (u = max L 2^-23: the wraps' and the box length's roundings),
    // so |r2_FP32 - r2_FP64| < 2 sqrt(3) r 4 u + r^2 2^-21; within the band 4e-5 r_t^2 + 32 r_t u
    // around a threshold r_t (r_c, r64), which holds that bound several times over, the FP64 rule
    // decides
    const float Lx = (float)E.L[0], Ly = (float)E.L[1], Lz = (float)E.L[2];
    const float iLx = 1.f / Lx, iLy = 1.f / Ly, iLz = 1.f / Lz;
    const float rc2f = (float)E.rc2, r64f = (float)E.r64sq;
    const float upos = fmaxf(Lx, fmaxf(Ly, Lz)) * 0x1p-23f;
    const float band_c2 = 4e-5f * rc2f + 32.f * sqrtf(rc2f) * upos;
    const float band_64 = 4e-5f * r64f + 32.f * sqrtf(r64f) * upos;
    const float4 pif = ew_wrap32(pi, E.L, iL0, iL1, iL2);
    unsigned* cf = ew_cand_base(E.cand32, list, E.cap32, lane);   // FP32 stream (mixed precision)
    unsigned* cn = ew_cand_base(E.cand64, list, E.cap64, lane);   // FP64 stream
    const double4 oI = E.tb_o[I];
    const double rc = sqrt(E.rc2);
    int nf = 0, nn = 0;
    int ntab = 0;   // mixed: entries of the list's slot table so far
    const float4 ui = E.posl[i];   // offset from tile I's origin (FP32)
    // ---- phase A
    // Group data of lane 8g + m: slot 8w + m of neighbour tile t0 + g (read unconditionally at a
    // valid tile and slot; a lane past the row's count or the tile's particles is not visited).  The
    // next group's neighbour tile and data are requested while the current group is screened.
    struct Grp {
        int J, ex, tc;
        double4 oJ, pl;
        float4 hJ, u;
    };
    const int g8 = lane >> 3, m8 = lane & 7, ks = ew_wave_slot(w, m8);
    auto load_grp = [&](int J, int ex, Grp& G) {
        G.J = J;
        G.ex = ex;
        G.oJ = E.tb_o[J];
        G.hJ = E.tb_h[J];
        G.tc = E.tcnt[J];
        G.pl = posa[J * kTile + ks];
        G.u = E.posl[J * kTile + ks];
    };
    Grp cur;
    load_grp(nb[min(g8, cnt - 1)], nbx[min(g8, cnt - 1)], cur);
    int Jnext = nb[min(8 + g8, cnt - 1)], Xnext = nbx[min(8 + g8, cnt - 1)];
    for (int t0 = 0; t0 < cnt; t0 += 8) {
        Grp nxt;
        load_grp(Jnext, Xnext, nxt);
        Jnext = nb[min(t0 + 16 + g8, cnt - 1)];
        Xnext = nbx[min(t0 + 16 + g8, cnt - 1)];
        const int Jl = cur.J, exl = cur.ex;
        const double4 pl = cur.pl, oJ = cur.oJ;
        float4 fl = make_float4(0.f, 0.f, 0.f, 0.f);
        bool near_box = false, img_ok = false;
        if (t0 + g8 < cnt) {
            // the pairs of tile pairs that take the image of the tile origins are screened on FP32
            // offsets from tile I's origin (and only they may have FP32 blocks); the others on
            // coordinates wrapped into the box
            img_ok = ew_tile_image_ok(E, oI, hI, oJ, cur.hJ, rc);
            if (ks < cur.tc) {
                float dx = (float)pl.x - cI.x, dy = (float)pl.y - cI.y, dz = (float)pl.z - cI.z;
                dx -= bxf * rintf(dx / bxf);
                dy -= byf * rintf(dy / byf);
                dz -= bzf * rintf(dz / bzf);
                dx = fmaxf(0.f, fabsf(dx) - hI.x);
                dy = fmaxf(0.f, fabsf(dy) - hI.y);
                dz = fmaxf(0.f, fabsf(dz) - hI.z);
                near_box = dx * dx + dy * dy + dz * dz < E.cut2_tiles;
                if (img_ok) {
                    const float4 u = cur.u;
                    fl = make_float4((float)((double)u.x + ew_tile_shift(oJ.x - oI.x, E.L[0], E.iL[0])),
                                     (float)((double)u.y + ew_tile_shift(oJ.y - oI.y, E.L[1], E.iL[1])),
                                     (float)((double)u.z + ew_tile_shift(oJ.z - oI.z, E.L[2], E.iL[2])), 0.f);
                } else {
                    fl = ew_wrap32(pl, E.L, iL0, iL1, iL2);
                }
            }
        }
        cur = nxt;
        const unsigned long long visit = __builtin_amdgcn_ballot_w64(near_box);
        if (!visit) continue;
        const unsigned long long frame_ok = __builtin_amdgcn_ballot_w64(img_ok);   // bit 8 g + m: tile t0 + g
        s_p[wl][lane] = pl;
        s_f[wl][lane] = fl;
        __builtin_amdgcn_wave_barrier();   // the slice is this wave's own: its LDS operations complete in order
        // masks of the group (bit 8 g + m): pairs scaled in the field (m0, mhalf: the epilogue
        // evaluates them) and pairs without a block of T (i0)
        unsigned long long mask = 0ull, maskT = 0ull;
        for (int g = 0; g < 8; g++) {
            if (!((visit >> (8 * g)) & 0xffull)) continue;
            const int J = __builtin_amdgcn_readlane(Jl, 8 * g);
            const int e = __builtin_amdgcn_readlane(exl, 8 * g);
            if (e >= 0) {
                const unsigned long long* mm = E.exc_mask + (size_t)e * 3 * kTile;
                mask |= ew_slots_of_wave(mm[lane] | mm[kTile + lane], w) << (8 * g);
                maskT |= ew_slots_of_wave(mm[2 * kTile + lane], w) << (8 * g);
            }
            if (J == I && (lane >> 3) == w) {
                const unsigned long long self = 1ull << (8 * g + (lane & 7));
                mask |= self;
                maskT |= self;
            }
        }
        if (!ivalid) mask = maskT = ~0ull;
        // the pair criterion for every visited slot, as per-lane bit masks of the group (inside the
        // cutoff, beyond r64), one 32-slot half at a time
        unsigned long long inm = 0ull, farm = 0ull;
#pragma unroll
        for (int half = 0; half < 2; half++) {
            unsigned v = (unsigned)(visit >> (32 * half));
            unsigned in32 = 0u, far32 = 0u;
            while (v) {
                const int b = __builtin_ctz(v);
                v &= v - 1u;
                const int q = 32 * half + b;
                const float4 pjf = s_f[wl][q];
                float fx, fy, fz;
                if ((frame_ok >> q) & 1ull) {   // uniform
                    fx = ui.x - pjf.x;
                    fy = ui.y - pjf.y;
                    fz = ui.z - pjf.z;
                } else {
                    fx = pif.x - pjf.x;
                    fy = pif.y - pjf.y;
                    fz = pif.z - pjf.z;
                    fx -= Lx * rintf(fx * iLx);
                    fy -= Ly * rintf(fy * iLy);
                    fz -= Lz * rintf(fz * iLz);
                }
                const float r2f = fx * fx + fy * fy + fz * fz;
                bool in = r2f < rc2f;
                bool far = r2f >= r64f;
                if (fabsf(r2f - rc2f) <= band_c2 || (MIXED && fabsf(r2f - r64f) <= band_64)) {   // the FP64 rule (rare)
                    const double4 pj = s_p[wl][q];
                    const double dx = ew_image_fast(__dsub_rn(pi.x, pj.x), E.L[0], iL0);
                    const double dy = ew_image_fast(__dsub_rn(pi.y, pj.y), E.L[1], iL1);
                    const double dz = ew_image_fast(__dsub_rn(pi.z, pj.z), E.L[2], iL2);
                    const double r2x = ew_r2(dx, dy, dz);
                    in = r2x < E.rc2;
                    far = r2x >= E.r64sq;
                }
                in32 |= in ? 1u << b : 0u;
                far32 |= far ? 1u << b : 0u;
            }
            inm |= (unsigned long long)in32 << (32 * half);
            farm |= (unsigned long long)far32 << (32 * half);
        }
        __builtin_amdgcn_wave_barrier();   // every lane has read the slice before the next group is staged
        // the group's candidates (inside, in the field sum or stored) into the lane's streams in
        // ascending slot order: FP32 blocks for pairs beyond r64 of tile pairs with the origins'
        // image (mixed precision), FP64 blocks for the others
        const unsigned long long cand = inm & ~(mask & maskT);
        const unsigned long long c32 = MIXED ? cand & farm & frame_ok : 0ull;
        // mixed: the group's slots that some lane's FP32 blocks name (stored candidates) join the
        // list's slot table in ascending order; lane 8 g + m enters its own slot
        unsigned long long used = 0ull;
        if (MIXED) {
            unsigned lo = (unsigned)(c32 & ~maskT), hi = (unsigned)((c32 & ~maskT) >> 32);
            for (int off = 32; off > 0; off >>= 1) {
                lo |= __shfl_xor(lo, off, 64);
                hi |= __shfl_xor(hi, off, 64);
            }
            used = ((unsigned long long)hi << 32) | lo;
            if ((used >> lane) & 1ull) {
                const int k = ntab + __popcll(used & ((1ull << lane) - 1ull));
                if (k < E.tcap) E.slot_tab[(size_t)list * E.tcap + k] = Jl * kTile + ks;
            }
        }
        auto emit = [&](unsigned long long m, unsigned* cb, int& n, int cap, bool tab) {
            while (__builtin_amdgcn_ballot_w64(m != 0ull)) {
                if (m) {
                    const int q = __builtin_ctzll(m);
                    m &= m - 1ull;
                    const bool st = !((maskT >> q) & 1ull);
                    unsigned c = (unsigned)(8 * t0 + q) | (((mask >> q) & 1ull) ? 0u : kCandSel) | (st ? kCandStore : 0u);
                    if (tab && st)   // (a table that overflows keeps every index inside it; the host regrows it)
                        c |= (unsigned)min(ntab + __popcll(used & ((1ull << q) - 1ull)), min(E.tcap - 1, kEwMaxSlots))
                             << kCandTabShift;
                    if (n < cap) cb[(n >> 3) * 512 + (n & 7)] = c;
                    n++;
                }
            }
        };
        if (MIXED) emit(c32, cf, nf, E.cap32, true);
        emit(cand & ~c32, cn, nn, E.cap64, false);
        ntab += __popcll(used);
    }
    E.ccnt[(size_t)list * kTile + lane] = make_int2(min(nf, E.cap32), min(nn, E.cap64));
    // a lane without candidates gets the entry 0 = slot 8w of the first neighbour tile, no flags: every
    // read of a stream (ew_cand_at) is then of a written entry
    if (MIXED && nf == 0) cf[0] = 0u;
    if (nn == 0) cn[0] = 0u;
    // ---- phase B
    int mf = nf, mn = nn;
    for (int off = 32; off > 0; off >>= 1) {
        mf = max(mf, __shfl_xor(mf, off, 64));
        mn = max(mn, __shfl_xor(mn, off, 64));
    }
    double sx = 0.0, sy = 0.0, sz = 0.0;
    int s32 = 0, s64 = 0;
    // one stream: FP32 blocks (FP32 = true) or FP64 blocks.  Software pipeline: while pair n is
    // evaluated, the slot data of candidate n+1, the neighbour tile of candidate n+2 and candidate n+3
    // itself are in flight, so that every wait is for a load issued an iteration earlier
    auto stream = [&](auto fp32_tag, const unsigned* cb, int nl, int wd, int& slots) {
        constexpr bool FP32 = decltype(fp32_tag)::value;
        const size_t lbase = FP32 ? ew_base32(list, E.cap32) : ew_base64(list, E.cap64);
        const int rows = FP32 ? (nl + 1) >> 1 : nl;   // the lane's rows of the jagged list
        int rs = 0, rank = 0;                          // start of the current row, the lane's place in it
        unsigned long long rm = 0ull;                  // lanes of the current row
        unsigned c0 = ew_cand_at(cb, 0, nl), c1 = ew_cand_at(cb, 1, nl), c2 = ew_cand_at(cb, 2, nl);
        int J1 = nb[ew_cand_tile(c1)];
        int j0 = ew_cand_slot(nb, c0, w);
        double4 p0 = posa[j0];
        double q0 = charge[j0], a0 = inva[j0];
        for (int n = 0; n < wd; n++) {
            const int row = FP32 ? n >> 1 : n;
            if (!FP32 || (n & 1) == 0) {
                rs += FP32 ? ew_row_slots32(rm) : ew_row_slots64(rm);
                rm = __builtin_amdgcn_ballot_w64(row < rows);
                rank = ew_lane_rank(rm);
            }
            const unsigned c3 = ew_cand_at(cb, n + 3, nl);
            const int J2 = nb[ew_cand_tile(c2)];
            const int j1 = J1 * kTile + ew_wave_slot(w, c1 & 7);
            const double4 p1 = posa[j1];
            const double q1 = charge[j1], a1 = inva[j1];
            const unsigned c = n < nl ? c0 : 0u;
            double C1 = 0.0, C2 = 0.0, dx = 0.0, dy = 0.0, dz = 0.0;
            if (n < nl) {
                dx = ew_image_fast(__dsub_rn(pi.x, p0.x), E.L[0], iL0);
                dy = ew_image_fast(__dsub_rn(pi.y, p0.y), E.L[1], iL1);
                dz = ew_image_fast(__dsub_rn(pi.z, p0.z), E.L[2], iL2);
                EwR<double> R;
                ew_radial<FP32>(E, ew_r2(dx, dy, dz), pi.w, p0.w, iai, a0, true, false, R);
                C1 = R.B1 + (R.l3 - 1.0) * R.b1;
                C2 = R.B2 + (R.l5 - 1.0) * R.b2;
                if (c & kCandSel) {
                    const double f = q0 * C1;
                    sx += f * dx;
                    sy += f * dy;
                    sz += f * dz;
                }
            }
            // the block T_ij = C2 d d^T - C1: FP32 as (C1, C2) (the T v pass forms d), FP64 as
            // (-C1, sqrt|C2| d) with the sign of C2 in the index
            const bool st = (c & kCandStore) != 0u;
            if (row >= rows) {
                // past the lane's list: nothing stored
            } else if (FP32) {   // entries in pairs per lane; the second of an odd list's last pair is zero
                const size_t at = (lbase + rs + rank) * 2 + (n & 1);
                E.idx16[at] = st ? (unsigned short)(c >> kCandTabShift) : (unsigned short)0;
                E.t32[at] = st ? make_float2((float)C1, (float)C2) : make_float2(0.f, 0.f);
                s32 += st ? 1 : 0;
            } else {
                const double g = st ? sqrt(fabs(C2)) : 0.0;
                const size_t at = lbase + rs + rank;
                // j itself in mixed precision (its FP64 pairs gather v_j from memory), its place among
                // the wave's slots in FP64 mode (the T v pass stages them in LDS)
                const int jx = MIXED ? j0 : (int)(c & 0xffffu);
                E.idx64[at] = st ? (jx | (C2 < 0.0 ? (int)0x80000000u : 0)) : (MIXED ? i : 0);
                E.t64[at] = st ? make_double4(-C1, g * dx, g * dy, g * dz) : make_double4(0.0, 0.0, 0.0, 0.0);
                s64 += st ? 1 : 0;
            }
            c0 = c1;
            c1 = c2;
            c2 = c3;
            J1 = J2;
            j0 = j1;
            p0 = p1;
            q0 = q1;
            a0 = a1;
        }
        // slots of the list's rows (entries, zero halves of odd FP32 lists, candidates in the field sum
        // only, and the rows' rounding to whole lines), FP32 entries counted singly
        slots = FP32 ? 2 * (rs + ew_row_slots32(rm)) : rs + ew_row_slots64(rm);
    };
    const int wd32 = MIXED ? min((mf + 1) & ~1, E.cap32) : 0;   // whole entry pairs
    const int wd64 = min(mn, E.cap64);
    int k32 = 0, k64 = 0;
    if (MIXED) stream(std::true_type(), cf, min(nf, E.cap32), wd32, k32);
    stream(std::false_type(), cn, min(nn, E.cap64), wd64, k64);
    int s32w = s32, s64w = s64;
    for (int off = 32; off > 0; off >>= 1) {
        s32w += __shfl_xor(s32w, off, 64);
        s64w += __shfl_xor(s64w, off, 64);
    }
    if (lane == 0) {
        E.w32[list] = wd32;
        E.w64[list] = wd64;
        if (MIXED) {
            E.slot_cnt[list] = min(ntab, E.tcap);
            atomicMax(&E.need[7], ntab);
        }
        atomicMax(&E.need[0], MIXED ? mf : 0);
        atomicMax(&E.need[1], mn);
        atomicAdd(&E.need[2], s32w);
        atomicAdd(&E.need[3], s64w);
        atomicAdd(&E.need[4], k32);
        atomicAdd(&E.need[5], k64);
        if (w == 0) atomicMax(&E.need[6], cnt);
    }
    // the part's field over its waves, in wave order (the epilogue adds the parts in order)
    red[(wl * 3 + 0) * kTile + lane] = sx;
    red[(wl * 3 + 1) * kTile + lane] = sy;
    red[(wl * 3 + 2) * kTile + lane] = sz;
    __syncthreads();
    if (wl != 0) return;
    for (int u = 1; u < kEwE0Waves; u++) {
        sx += red[(u * 3 + 0) * kTile + lane];
        sy += red[(u * 3 + 1) * kTile + lane];
        sz += red[(u * 3 + 2) * kTile + lane];
    }
    rs[(size_t)h * E.npad + i] = make_double4(sx, sy, sz, 0.0);
}

// The force pass from the E0 pass's candidates (same positions): each lane runs through its unscaled
// pairs (in the field sum and stored, i.e. no scale factor: the epilogue evaluates the scaled ones),
// FP32 stream first (FP32 terms in mixed precision, summed in FP32 over at most 64 candidates, then
// added to the FP64 sums), then the FP64 stream.  Workgroups as in k_ew_e0_build.
template <bool MUTUAL, bool MIXED>
__global__ __launch_bounds__(kEwE0Waves * 64) void k_ew_force_cand(EwLaunch E, const double4* __restrict__ posa,
                                                                   const double* __restrict__ inva,
                                                                   const double* __restrict__ charge,
                                                                   const double4* __restrict__ vec,
                                                                   double4* __restrict__ rs) {
    const int I = blockIdx.x / kEwE0Parts, h = blockIdx.x % kEwE0Parts;
    const int lane = threadIdx.x & (kTile - 1), wl = threadIdx.x / kTile, w = kEwE0Waves * h + wl;
    __shared__ double red[kEwE0Waves * 4 * kTile];
    const int i = I * kTile + lane;
    const double4 pi = posa[i];
    const double iai = inva[i];
    const double qi = charge[i];
    const double4 vi = vec[i];
    const double iL0 = 1.0 / E.L[0], iL1 = 1.0 / E.L[1], iL2 = 1.0 / E.L[2];
    const int* nb = E.nbr + (size_t)I * E.nblk;
    const int list = I * kEwWaves + w;
    const int2 nc = E.ccnt[(size_t)list * kTile + lane];
    int mf = nc.x, mn = nc.y;
    for (int off = 32; off > 0; off >>= 1) {
        mf = max(mf, __shfl_xor(mf, off, 64));
        mn = max(mn, __shfl_xor(mn, off, 64));
    }
    double ax = 0.0, ay = 0.0, az = 0.0, ae = 0.0;
    constexpr unsigned kUnscaled = kCandSel | kCandStore;
    // each stream: while pair n is evaluated, the slot data of candidate n+1, the neighbour tile of
    // candidate n+2 and candidate n+3 itself are in flight (every wait is for a load issued an
    // iteration earlier)
    if (MIXED) {
        // FP32 candidates: their tile pairs take the image of the tile origins (k_ew_e0_build), so d is
        // u_i - (u_j + D_t) with D_t the image of the origins' separation (FP64 in LDS for the row's
        // neighbour tiles when they fit, else from the origins) and the slot data are three FP32
        // records (offset, dipole, parameters)
        extern __shared__ double4 s_D[];   // [E.lds_tiles] (shared by the workgroup's waves: one row)
        const int cnt = E.nbr_cnt[I];
        const double4 oI = E.tb_o[I];
        // a row with more neighbour tiles than the launch's LDS holds (the host sizes it from the
        // previous list build) forms its shifts from the tile origins instead
        const bool lds = E.lds_tiles > 0 && cnt <= E.lds_tiles;
        if (lds) {
            for (int t = threadIdx.x; t < cnt; t += blockDim.x) {
                const double4 oJ = E.tb_o[nb[t]];
                s_D[t] = make_double4(ew_tile_shift(oJ.x - oI.x, E.L[0], E.iL[0]),
                                      ew_tile_shift(oJ.y - oI.y, E.L[1], E.iL[1]),
                                      ew_tile_shift(oJ.z - oI.z, E.L[2], E.iL[2]), 0.0);
            }
            __syncthreads();
        }
        const float4 ui = E.posl[i];
        const unsigned* cb = ew_cand_base(E.cand32, list, E.cap32, lane);
        const float alpha32 = (float)E.alpha, c1f = (float)E.c1, c2f = (float)E.c2, c3f = (float)E.c3;
        const float qif = (float)qi, aif = (float)pi.w, iaif = (float)iai;
        const float vixf = (float)vi.x, viyf = (float)vi.y, vizf = (float)vi.z;
        const int nl = nc.x;
        unsigned c0 = ew_cand_at(cb, 0, nl), c1 = ew_cand_at(cb, 1, nl), c2 = ew_cand_at(cb, 2, nl);
        int J1 = nb[ew_cand_tile(c1)];
        const int j0 = ew_cand_slot(nb, c0, w);
        float4 u0 = E.posl[j0], m0 = E.mu32[j0], r0 = E.slot32[j0];
        float fx = 0.f, fy = 0.f, fz = 0.f, fe = 0.f;
        for (int n = 0; n < mf; n++) {
            const unsigned c3 = ew_cand_at(cb, n + 3, nl);
            const int J2 = nb[ew_cand_tile(c2)];
            const int j1 = J1 * kTile + ew_wave_slot(w, c1 & 7);
            const float4 u1 = E.posl[j1], m1 = E.mu32[j1], r1 = E.slot32[j1];
            if (n < nl && (c0 & kUnscaled) == kUnscaled) {
                double4 D;
                if (lds) {
                    D = s_D[ew_cand_tile(c0)];
                } else {
                    const double4 oJ = E.tb_o[nb[ew_cand_tile(c0)]];
                    D = make_double4(ew_tile_shift(oJ.x - oI.x, E.L[0], E.iL[0]),
                                     ew_tile_shift(oJ.y - oI.y, E.L[1], E.iL[1]),
                                     ew_tile_shift(oJ.z - oI.z, E.L[2], E.iL[2]), 0.0);
                }
                const float gx = ui.x - (float)((double)u0.x + D.x);
                const float gy = ui.y - (float)((double)u0.y + D.y);
                const float gz = ui.z - (float)((double)u0.z + D.z);
                EwR<float> R;
                ew_radial32(alpha32, c1f, c2f, c3f, gx * gx + gy * gy + gz * gz, aif, r0.y, iaif, r0.z, MUTUAL, R);
                ew_pair_terms<float, kEwForce, MUTUAL>(R, gx, gy, gz, qif, vixf, viyf, vizf, r0.x, m0.x, m0.y, m0.z, fx,
                                                       fy, fz, fe);
            }
            if ((n & 63) == 63) {   // FP32 sums over at most 64 candidates
                ax += (double)fx;
                ay += (double)fy;
                az += (double)fz;
                ae += (double)fe;
                fx = fy = fz = fe = 0.f;
            }
            c0 = c1;
            c1 = c2;
            c2 = c3;
            J1 = J2;
            u0 = u1;
            m0 = m1;
            r0 = r1;
        }
        ax += (double)fx;
        ay += (double)fy;
        az += (double)fz;
        ae += (double)fe;
    }
    {
        const unsigned* cb = ew_cand_base(E.cand64, list, E.cap64, lane);
        const int nl = nc.y;
        unsigned c0 = ew_cand_at(cb, 0, nl), c1 = ew_cand_at(cb, 1, nl), c2 = ew_cand_at(cb, 2, nl);
        int J1 = nb[ew_cand_tile(c1)];
        const int j0 = ew_cand_slot(nb, c0, w);
        double4 p0 = posa[j0], v0 = vec[j0];
        double q0 = charge[j0], a0 = inva[j0];
        for (int n = 0; n < mn; n++) {
            const unsigned c3 = ew_cand_at(cb, n + 3, nl);
            const int J2 = nb[ew_cand_tile(c2)];
            const int j1 = J1 * kTile + ew_wave_slot(w, c1 & 7);
            const double4 p1 = posa[j1], v1 = vec[j1];
            const double q1 = charge[j1], a1 = inva[j1];
            if (n < nl && (c0 & kUnscaled) == kUnscaled) {
                const double dx = ew_image_fast(__dsub_rn(pi.x, p0.x), E.L[0], iL0);
                const double dy = ew_image_fast(__dsub_rn(pi.y, p0.y), E.L[1], iL1);
                const double dz = ew_image_fast(__dsub_rn(pi.z, p0.z), E.L[2], iL2);
                EwR<double> R;
                ew_radial(E, ew_r2(dx, dy, dz), pi.w, p0.w, iai, a0, true, MUTUAL, R);
                ew_pair_terms<double, kEwForce, MUTUAL>(R, dx, dy, dz, qi, vi.x, vi.y, vi.z, q0, v0.x, v0.y, v0.z, ax, ay,
                                                        az, ae);
            }
            c0 = c1;
            c1 = c2;
            c2 = c3;
            J1 = J2;
            p0 = p1;
            v0 = v1;
            q0 = q1;
            a0 = a1;
        }
    }
    red[(wl * 4 + 0) * kTile + lane] = ax;
    red[(wl * 4 + 1) * kTile + lane] = ay;
    red[(wl * 4 + 2) * kTile + lane] = az;
    red[(wl * 4 + 3) * kTile + lane] = ae;
    __syncthreads();
    if (wl != 0) return;
    for (int u = 1; u < kEwE0Waves; u++) {
        ax += red[(u * 4 + 0) * kTile + lane];
        ay += red[(u * 4 + 1) * kTile + lane];
        az += red[(u * 4 + 2) * kTile + lane];
        ae += red[(u * 4 + 3) * kTile + lane];
    }
    rs[(size_t)h * E.npad + i] = make_double4(ax, ay, az, ae);
}

// FP32 records of every slot for the force pass's FP32 pairs: the dipole, and (charge, Thole a, 1/a).
__global__ __launch_bounds__(256) void k_ew_force_records(int npad, const double4* __restrict__ posa,
                                                          const double* __restrict__ inva,
                                                          const double* __restrict__ charge,
                                                          const double4* __restrict__ vec, float4* __restrict__ mu32,
                                                          float4* __restrict__ slot32) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npad) return;
    const double4 v = vec[i];
    mu32[i] = make_float4((float)v.x, (float)v.y, (float)v.z, 0.f);
    slot32[i] = make_float4((float)charge[i], (float)posa[i].w, (float)inva[i], 0.f);
}

typedef float ew_f4 __attribute__((ext_vector_type(4)));

// FP32 entries of one list: the sum of T_ij v_j = C2 d (d . v_j) - C1 v_j with d = u_i - y_j and the
// slot data (v_j, y_j) from the list's LDS copy (LDS) or formed from the neighbour-tile list.
// Software pipeline over steps of kEwSpmvU entry pairs, three register sets by step mod 3: while step
// s is summed, the indices and coefficients of steps s+1 and s+2 are in flight (the memory counter
// completes in order, so waiting for step s never waits for a later step); the first two steps are
// issued by the caller before the slot data is staged.  The rows are jagged (ew_lane_rank): steps
// that every lane's list covers are summed without weights, the others weight each row by whether
// the lane's list reaches it.
struct EwSpmvPipe {
    unsigned j[3][kEwSpmvU];
    ew_f4 t[3][kEwSpmvU];
    size_t base;   // the list's first entry pair
    int rs;        // start of the next row to be requested (rows are requested in order)
    int rows;      // the lane's rows (entry pairs)
};

// Requests the rows of step `step` into register set `set`; a lane past its list reads the list's
// first pair (cached) and the sum gives it weight zero.
__device__ __forceinline__ void ew_spmv_load(const EwLaunch& E, int lane, EwSpmvPipe& P, int set, int step) {
    const unsigned* ip = reinterpret_cast<const unsigned*>(E.idx16);
    const ew_f4* tp = reinterpret_cast<const ew_f4*>(E.t32);
#pragma unroll
    for (int u = 0; u < kEwSpmvU; u++) {
        const unsigned long long m = __builtin_amdgcn_ballot_w64(step * kEwSpmvU + u < P.rows);
        const size_t at = P.base + (((m >> lane) & 1ull) ? (size_t)(P.rs + ew_lane_rank(m)) : 0);
        P.rs += ew_row_slots32(m);
        P.j[set][u] = ip[at];
        P.t[set][u] = __builtin_nontemporal_load(tp + at);
    }
}

template <bool LDS>
__device__ __forceinline__ void ew_spmv_fp32(const EwLaunch& E, int w, int lane, int n32, int minrows, EwSpmvPipe& P,
                                             const float4* __restrict__ s_a, const float2* __restrict__ s_b,
                                             const float4* __restrict__ v32, const int* __restrict__ tab,
                                             const double4& oI, const float4& ui, float& fx, float& fy,
                                             float& fz) {
    const int ns = ((n32 + 1) / 2 + kEwSpmvU - 1) / kEwSpmvU;
    auto term = [&](int loc, float c1, float c2) {
        float4 a;
        float2 b;
        if (LDS) {
            a = s_a[loc];
            b = s_b[loc];
        } else {
            const int j = tab[loc];
            ew_slot32(E, oI, j / kTile, j, v32, a, b);
        }
        const float dx = ui.x - a.w, dy = ui.y - b.x, dz = ui.z - b.y;
        const float p = c2 * fmaf(dz, a.z, fmaf(dy, a.y, dx * a.x));
        fx = fmaf(-c1, a.x, fmaf(p, dx, fx));
        fy = fmaf(-c1, a.y, fmaf(p, dy, fy));
        fz = fmaf(-c1, a.z, fmaf(p, dz, fz));
    };
    auto sum = [&](int set, int step) {
        if ((step + 1) * kEwSpmvU <= minrows) {   // every lane's list holds the whole step (uniform)
#pragma unroll
            for (int u = 0; u < kEwSpmvU; u++) {
                const ew_f4 c = P.t[set][u];
                term((int)(P.j[set][u] & 0xffffu), c.x, c.y);
                term((int)(P.j[set][u] >> 16), c.z, c.w);
            }
        } else {
#pragma unroll
            for (int u = 0; u < kEwSpmvU; u++) {
                const float on = step * kEwSpmvU + u < P.rows ? 1.f : 0.f;
                const ew_f4 c = P.t[set][u];
                term((int)(P.j[set][u] & 0xffffu), on * c.x, on * c.y);
                term((int)(P.j[set][u] >> 16), on * c.z, on * c.w);
            }
        }
    };
    for (int s0 = 0; s0 < ns; s0 += 3) {
        ew_spmv_load(E, lane, P, 2, s0 + 2);
        __builtin_amdgcn_sched_barrier(0);
        sum(0, s0);
        __builtin_amdgcn_sched_barrier(0);
        ew_spmv_load(E, lane, P, 0, s0 + 3);
        __builtin_amdgcn_sched_barrier(0);
        if (s0 + 1 < ns) sum(1, s0 + 1);
        __builtin_amdgcn_sched_barrier(0);
        ew_spmv_load(E, lane, P, 1, s0 + 4);
        __builtin_amdgcn_sched_barrier(0);
        if (s0 + 2 < ns) sum(2, s0 + 2);
        __builtin_amdgcn_sched_barrier(0);
    }
}

// T v of the real-space pairs from the stored blocks: one wave per list (tile row I, wave w), so that
// the nblk * kEwWaves waves spread evenly over the CUs.  The FP32 lists read their slot data from an
// LDS copy of the slots their entries name (the list's slot table, up to E.lds_slots entries of
// kEwSlotLds bytes: v_j and y_j, staged once, then every gather is an LDS read); their first entries
// are requested before the copy is staged.  The FP64 lists gather v_j from HBM (L2), four entries at a
// time (their index and block loads, then the four gathers).  Each wave writes its partial sums; the
// epilogue adds the kEwWaves partials of a slot in list order.  10 bytes per FP32 entry and 36 per
// FP64 entry stream from HBM once per pass.
template <bool MIXED>
__global__ __launch_bounds__(64) void k_ew_spmv(EwLaunch E, const double4* __restrict__ vec,
                                                const float4* __restrict__ v32, const double* __restrict__ scalars,
                                                double4* __restrict__ part) {
    // mixed: [E.lds_slots] (v, y_x), then float2 [E.lds_slots] (y_y, y_z); FP64 mode: see below
    extern __shared__ float4 s_a[];
    if (scalars && scalars[SC_DONE] != 0.0) return;
    const int row = blockIdx.x, lane = threadIdx.x;   // row = I kEwWaves + w
    const int I = row / kEwWaves, w = row % kEwWaves;
    const int* nb = E.nbr + (size_t)I * E.nblk;
    const int n32 = MIXED ? E.w32[row] : 0;
    const int ntab = MIXED ? E.slot_cnt[row] : 0;
    const int* tab = MIXED ? E.slot_tab + (size_t)row * E.tcap : nullptr;
    // a list whose slot table is longer than the launch's LDS holds (sized from the previous list
    // build) gathers its slot data from HBM
    // (a list whose FP32 blocks are all zero, of pairs without a T block, names no slot: skipped)
    const bool has32 = MIXED && n32 > 0 && ntab > 0;
    const bool lds = has32 && E.lds_slots > 0 && ntab <= E.lds_slots;
    float2* s_b = reinterpret_cast<float2*>(s_a + (size_t)E.lds_slots);
    const int2 rows = E.ccnt[(size_t)row * kTile + lane];   // the lane's FP32 entries and FP64 rows
    EwSpmvPipe P;
    P.base = ew_base32(row, E.cap32);
    P.rs = 0;
    P.rows = (rows.x + 1) >> 1;
    int minrows = P.rows;
    double4 oI = make_double4(0.0, 0.0, 0.0, 0.0);
    float4 ui = make_float4(0.f, 0.f, 0.f, 0.f);
    if (has32) {
        ew_spmv_load(E, lane, P, 0, 0);
        ew_spmv_load(E, lane, P, 1, 1);
        oI = E.tb_o[I];
        ui = E.posl[I * kTile + lane];
        for (int off = 32; off > 0; off >>= 1) minrows = min(minrows, __shfl_xor(minrows, off, 64));
    }
    if (lds) {
        // entry k: slot tab[k] (the wave's own LDS: its operations complete in order); four slots per
        // lane at a time, their loads issued together, and the table entries of the next four
        // requested before them
        int jn[4];
#pragma unroll
        for (int u = 0; u < 4; u++) jn[u] = tab[min(lane + u * kTile, ntab - 1)];
        for (int e0 = lane; e0 < ntab; e0 += 4 * kTile) {
            int j[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                j[u] = jn[u];
                jn[u] = tab[min(e0 + (4 + u) * kTile, ntab - 1)];
            }
            float4 a[4];
            float2 b[4];
#pragma unroll
            for (int u = 0; u < 4; u++) ew_slot32(E, oI, j[u] / kTile, j[u], v32, a[u], b[u]);
#pragma unroll
            for (int u = 0; u < 4; u++)
                if (e0 + u * kTile < ntab) {
                    s_a[e0 + u * kTile] = a[u];
                    s_b[e0 + u * kTile] = b[u];
                }
        }
        __builtin_amdgcn_wave_barrier();
    }
    // FP64 mode: the FP64 lists index the wave's slots (8 t + m) and their v_j (FP64, 24 B per slot)
    // are staged in LDS like the FP32 lists' slot data; mixed: j itself, v_j from memory
    const bool lds64 = !MIXED && E.lds_tiles > 0 && E.nbr_cnt[I] <= E.lds_tiles;
    double2* s_vxy = reinterpret_cast<double2*>(s_a);
    double* s_vz = reinterpret_cast<double*>(s_vxy + (size_t)E.lds_tiles * 8);
    if (lds64) {
        const int ne = 8 * E.nbr_cnt[I];
        for (int e0 = lane; e0 < ne; e0 += 4 * kTile) {
            double4 v[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int e = min(e0 + u * kTile, ne - 1);
                v[u] = vec[nb[e >> 3] * kTile + ew_wave_slot(w, e & 7)];
            }
#pragma unroll
            for (int u = 0; u < 4; u++)
                if (e0 + u * kTile < ne) {
                    s_vxy[e0 + u * kTile] = make_double2(v[u].x, v[u].y);
                    s_vz[e0 + u * kTile] = v[u].z;
                }
        }
        __builtin_amdgcn_wave_barrier();
    }
    double ax = 0.0, ay = 0.0, az = 0.0;
    {
        const int n64 = E.w64[row];
        const size_t base = ew_base64(row, E.cap64);
        int rs = 0;   // start of the next row; a lane past its list reads the list's first entry
        for (int n0 = 0; n0 < n64; n0 += 4) {
            int j[4];
            double4 t[4], v[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const unsigned long long m = __builtin_amdgcn_ballot_w64(n0 + u < rows.y);
                const size_t at = base + (((m >> lane) & 1ull) ? (size_t)(rs + ew_lane_rank(m)) : 0);
                rs += ew_row_slots64(m);
                j[u] = E.idx64[at];
                t[u] = E.t64[at];
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int loc = j[u] & 0x7fffffff;
                if (lds64) {
                    const double2 xy = s_vxy[loc];
                    v[u] = make_double4(xy.x, xy.y, s_vz[loc], 0.0);
                } else if (!MIXED) {
                    v[u] = vec[nb[loc >> 3] * kTile + ew_wave_slot(w, loc & 7)];
                } else {
                    v[u] = vec[loc];
                }
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {   // rows past the lane's list with weight zero
                const double on = n0 + u < rows.y ? 1.0 : 0.0;
                double p = t[u].y * v[u].x + t[u].z * v[u].y + t[u].w * v[u].z;
                p = j[u] < 0 ? -on * p : on * p;
                const double c = on * t[u].x;
                ax += c * v[u].x + p * t[u].y;
                ay += c * v[u].y + p * t[u].z;
                az += c * v[u].z + p * t[u].w;
            }
        }
    }
    if (has32) {
        float fx = 0.f, fy = 0.f, fz = 0.f;
        if (lds) ew_spmv_fp32<true>(E, w, lane, n32, minrows, P, s_a, s_b, v32, tab, oI, ui, fx, fy, fz);
        else ew_spmv_fp32<false>(E, w, lane, n32, minrows, P, s_a, s_b, v32, tab, oI, ui, fx, fy, fz);
        ax += fx;
        ay += fy;
        az += fz;
    }
    part[(size_t)row * kTile + lane] = make_double4(ax, ay, az, 0.0);
}

// Epilogue of a pass for the 64 slots of tile I, on four waves so that their dependent loads
// overlap: waves 0-2 gather the reciprocal field (E = -(K/L) grad_u phi) or force (q E + grad E . v)
// from x planes 0-1, 2-3 and 4 of the spline support of the potential grids; wave 3 adds the row
// parts' real-space sums (in part order), the scaled pairs of each slot in FP64 with their scales
// ([inside] B_n + (s lambda - 1) b_n) and the self field; wave 0 adds the four parts in wave order
// and forms the outputs of the mode.
template <int MODE, bool MUTUAL>
__global__ __launch_bounds__(256) void k_ew_epilogue(EwLaunch E, const double4* __restrict__ posa,
                                                     const double* __restrict__ inva, const double* __restrict__ charge,
                                                     const double4* __restrict__ vec, const double4* __restrict__ E0,
                                                     const double* __restrict__ scalars,
                                                     const double4* __restrict__ rs, double4* __restrict__ out4, PcgQ pq,
                                                     double* __restrict__ epart, double* __restrict__ frc,
                                                     double* __restrict__ eout) {
    if (scalars && scalars[SC_DONE] != 0.0) return;
    __shared__ double s_part[3][5][kTile];   // waves 1-3: field (3), energy, potential
    const int I = blockIdx.x, w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i = I * kTile + lane;
    const bool valid = lane < E.tcnt[I];
    const double4 pi = posa[i];
    const double qi = charge[i];
    const double4 vi = (MODE == kEwE0) ? make_double4(0.0, 0.0, 0.0, 0.0) : vec[i];
    double ax = 0.0, ay = 0.0, az = 0.0, ae = 0.0, ph = 0.0;
    if (w == 3) {
        {   // the row parts' sums, in part order
            const double4 r0 = rs[i];
            ax = r0.x;
            ay = r0.y;
            az = r0.z;
            ae = r0.w;
            for (int h = 1; h < (MODE == kEwE0 ? E.e0_parts : ew_parts(MODE)); h++) {
                const double4 r = rs[(size_t)h * E.npad + i];
                ax += r.x;
                ay += r.y;
                az += r.z;
                ae += r.w;
            }
        }
        if (MODE == kEwTmu) {
            ax += E.self_field * vi.x;
            ay += E.self_field * vi.y;
            az += E.self_field * vi.z;
        }
        const double iai = inva[i];
        const double iL0 = 1.0 / E.L[0], iL1 = 1.0 / E.L[1], iL2 = 1.0 / E.L[2];
        if (valid) {
            for (int q = E.sc_off[i]; q < E.sc_off[i + 1]; q++) {
                const int j = E.sc_j[q];
                const double2 sc = E.sc_s[q];
                const double m = sc.x, is = sc.y;
                if (MODE == kEwE0 && m == 1.0) continue;
                if (MODE == kEwTmu && is == 1.0) continue;
                const double4 pj = posa[j];
                const double dx = ew_image_fast(__dsub_rn(pi.x, pj.x), E.L[0], iL0);
                const double dy = ew_image_fast(__dsub_rn(pi.y, pj.y), E.L[1], iL1);
                const double dz = ew_image_fast(__dsub_rn(pi.z, pj.z), E.L[2], iL2);
                const double r2 = ew_r2(dx, dy, dz);
                const bool inside = r2 < E.rc2;
                EwR<double> R;
                ew_radial(E, r2, pi.w, pj.w, iai, inva[j], inside, true, R);
                const bool cm = inside || m != 1.0, ci = inside || is != 1.0;
                const double Cqm1 = R.B1 + (cm ? (m * R.l3 - 1.0) * R.b1 : 0.0);
                const double Cqm2 = R.B2 + (cm ? (m * R.l5 - 1.0) * R.b2 : 0.0);
                const double Cmm1 = R.B1 + (ci ? (is * R.l3 - 1.0) * R.b1 : 0.0);
                const double Cmm2 = R.B2 + (ci ? (is * R.l5 - 1.0) * R.b2 : 0.0);
                const double qj = charge[j];
                if (MODE == kEwE0) {
                    ax += qj * Cqm1 * dx;
                    ay += qj * Cqm1 * dy;
                    az += qj * Cqm1 * dz;
                } else if (MODE == kEwTmu) {
                    const double4 vj = vec[j];
                    const double vd = vj.x * dx + vj.y * dy + vj.z * dz;
                    ax += Cmm2 * vd * dx - Cmm1 * vj.x;
                    ay += Cmm2 * vd * dy - Cmm1 * vj.y;
                    az += Cmm2 * vd * dz - Cmm1 * vj.z;
                } else {
                    const double4 vj = vec[j];
                    const double b0 = R.b1 * r2;   // 1/r
                    const double Cqq0 = R.B0 + (cm ? (m - 1.0) * b0 : 0.0);
                    const double Cqq1 = R.B1 + (cm ? (m - 1.0) * R.b1 : 0.0);
                    double gx = -qi * qj * Cqq1 * dx, gy = -qi * qj * Cqq1 * dy, gz = -qi * qj * Cqq1 * dz;
                    const double ux = qi * vj.x - qj * vi.x, uy = qi * vj.y - qj * vi.y, uz = qi * vj.z - qj * vi.z;
                    const double ud = ux * dx + uy * dy + uz * dz;
                    gx += Cqm1 * ux - Cqm2 * ud * dx;
                    gy += Cqm1 * uy - Cqm2 * ud * dy;
                    gz += Cqm1 * uz - Cqm2 * ud * dz;
                    if (MUTUAL) {
                        const double Cmm3 = R.B3 + (ci ? (is * R.l7 - 1.0) * R.b3 : 0.0);
                        const double mid = vi.x * dx + vi.y * dy + vi.z * dz;
                        const double mjd = vj.x * dx + vj.y * dy + vj.z * dz;
                        const double mm = vi.x * vj.x + vi.y * vj.y + vi.z * vj.z;
                        const double cd = -Cmm2 * mm + Cmm3 * mid * mjd;
                        gx += cd * dx - Cmm2 * (vi.x * mjd + vj.x * mid);
                        gy += cd * dy - Cmm2 * (vi.y * mjd + vj.y * mid);
                        gz += cd * dz - Cmm2 * (vi.z * mjd + vj.z * mid);
                    }
                    ax -= gx;
                    ay -= gy;
                    az -= gz;
                    ae += 0.5 * qi * qj * Cqq0;
                }
            }
        }
    } else {
        const int a0 = 2 * w, a1 = w == 2 ? kEwOrder : 2 * w + 2;   // x planes of the support
        const double sxk = E.K[0] / E.L[0], syk = E.K[1] / E.L[1], szk = E.K[2] / E.L[2];
        double g1[3], g2[6];
        if (MODE == kEwE0 || MODE == kEwTmu) {
            ew_gather<double, 0>(E, reinterpret_cast<const double*>(MODE == kEwE0 ? E.phi_q : E.phi_v), pi, ph, g1,
                                 g2, a0, a1);
            ax = -g1[0] * sxk;
            ay = -g1[1] * syk;
            az = -g1[2] * szk;
        } else {
            double gv[3], hv[6], phv;
            ew_gather<double, 1>(E, reinterpret_cast<const double*>(E.phi_q), pi, ph, g1, g2, a0, a1);
            ew_gather<double, 1>(E, reinterpret_cast<const double*>(E.phi_v), pi, phv, gv, hv, a0, a1);
            const double Ex = -(g1[0] + gv[0]) * sxk, Ey = -(g1[1] + gv[1]) * syk, Ez = -(g1[2] + gv[2]) * szk;
            double h[6];
            for (int k = 0; k < 6; k++) h[k] = g2[k] + (MUTUAL ? hv[k] : 0.0);
            const double hxx = -h[0] * sxk * sxk, hyy = -h[1] * syk * syk, hzz = -h[2] * szk * szk;
            const double hxy = -h[3] * sxk * syk, hxz = -h[4] * sxk * szk, hyz = -h[5] * syk * szk;
            ax = qi * Ex + hxx * vi.x + hxy * vi.y + hxz * vi.z;
            ay = qi * Ey + hxy * vi.x + hyy * vi.y + hyz * vi.z;
            az = qi * Ez + hxz * vi.x + hyz * vi.y + hzz * vi.z;
        }
    }
    if (w > 0) {
        s_part[w - 1][0][lane] = ax;
        s_part[w - 1][1][lane] = ay;
        s_part[w - 1][2][lane] = az;
        s_part[w - 1][3][lane] = ae;
        s_part[w - 1][4][lane] = ph;
    }
    __syncthreads();
    if (w != 0) return;
    for (int u = 0; u < 3; u++) {
        ax += s_part[u][0][lane];
        ay += s_part[u][1][lane];
        az += s_part[u][2][lane];
        ae += s_part[u][3][lane];
        ph += s_part[u][4][lane];
    }
    if (MODE == kEwE0) {
        if (valid) {
            out4[i] = make_double4(ax, ay, az, 0.0);
            epart[i] = 0.5 * qi * ph + E.self_energy * qi * qi;
        } else {
            out4[i] = make_double4(0.0, 0.0, 0.0, 0.0);
            epart[i] = 0.0;
        }
    } else if (MODE == kEwTmu) {
        if (!valid) ax = ay = az = 0.0;
        if (pq.q) {   // PCG: q = p/alpha - T p and the tile sum of p.q
            const double4 p = pq.p[i];
            const double a = pq.alpha[i];
            const double ia = a > 0.0 ? 1.0 / a : 0.0;
            const double4 qv = make_double4(p.x * ia - ax, p.y * ia - ay, p.z * ia - az, 0.0);
            pq.q[i] = qv;
            const double d = wave_sum(p.x * qv.x + p.y * qv.y + p.z * qv.z);
            if (lane == 0) pq.red[I] = d;
        } else {
            out4[i] = make_double4(ax, ay, az, 0.0);
        }
    } else {
        if (valid) {
            const double4 e0 = E0[i];
            frc[3 * i] = ax;
            frc[3 * i + 1] = ay;
            frc[3 * i + 2] = az;
            eout[i] = ae + epart[i] - 0.5 * (vi.x * e0.x + vi.y * e0.y + vi.z * e0.z);
        } else {
            frc[3 * i] = frc[3 * i + 1] = frc[3 * i + 2] = 0.0;
            eout[i] = 0.0;
        }
    }
}

// Energy: slots summed in a fixed order by one workgroup of 1024 threads (each thread its slots,
// eight loads in flight at a time, then the waves' sums in wave order), plus the background term.
__global__ __launch_bounds__(1024) void k_ew_energy(int npad, const double* __restrict__ eout, double background,
                                                    double* __restrict__ energy) {
    __shared__ double s[16];
    double acc = 0.0;
    for (int i0 = threadIdx.x; i0 < npad; i0 += 8 * blockDim.x) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int i = i0 + u * blockDim.x;
            v[u] = i < npad ? eout[i] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 8; u++) acc += v[u];
    }
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x / 64] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double e = 0.0;
        for (int k = 0; k < (int)(blockDim.x / 64); k++) e += s[k];
        energy[0] = e + background;
    }
}

// Every pair within the cutoff, from the side of its lower caller index.
__global__ __launch_bounds__(64) void k_ew_dump_pairs(EwLaunch E, const double4* __restrict__ posa,
                                                      const int* __restrict__ perm, int* __restrict__ pairs,
                                                      long long cap, unsigned long long* __restrict__ count) {
    const int I = blockIdx.x, lane = threadIdx.x;
    const int i = I * kTile + lane;
    const bool ivalid = lane < E.tcnt[I];
    const double4 pi = posa[i];
    const int oi = perm[i];
    const int cnt = E.nbr_cnt[I];
    for (int t = 0; t < cnt; t++) {
        const int J = E.nbr[(size_t)I * E.nblk + t];
        for (int k = 0; k < E.tcnt[J]; k++) {
            const int j = J * kTile + k;
            const int oj = perm[j];
            if (!ivalid || oj <= oi) continue;
            const double4 pj = posa[j];
            const double dx = ew_image(__dsub_rn(pi.x, pj.x), E.L[0]);
            const double dy = ew_image(__dsub_rn(pi.y, pj.y), E.L[1]);
            const double dz = ew_image(__dsub_rn(pi.z, pj.z), E.L[2]);
            if (!(ew_r2(dx, dy, dz) < E.rc2)) continue;
            const unsigned long long at = atomicAdd(count, 1ull);
            if ((long long)at < cap) {
                pairs[2 * at] = oi;
                pairs[2 * at + 1] = oj;
            }
        }
    }
}

// Epilogue of the stored-block T v pass for the 64 slots of tile I, on four waves so that its
// dependent loads overlap: waves 0 and 1 gather the reciprocal field from x planes 0-1 and 2-3 of
// the spline support, wave 2 from plane 4 and adds the kEwWaves list partials (in list order),
// wave 3 the scaled pairs (FP64, [inside] B_n + (s lambda - 1) b_n) and the self field; wave 0 adds
// the four parts in wave order and forms the outputs (q = p/alpha - T p and the tile sum of p.q for
// PCG, else T p).
__global__ __launch_bounds__(256) void k_ew_tmu_epilogue(EwLaunch E, const double4* __restrict__ posa,
                                                         const double* __restrict__ inva,
                                                         const double4* __restrict__ vec,
                                                         const double* __restrict__ scalars,
                                                         const double4* __restrict__ rs, double4* __restrict__ out4,
                                                         PcgQ pq) {
    if (scalars && scalars[SC_DONE] != 0.0) return;
    __shared__ double s_part[3][3][kTile];   // waves 1-3: their parts of the field
    const int I = blockIdx.x, w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i = I * kTile + lane;
    const bool valid = lane < E.tcnt[I];
    const double4 pi = posa[i];
    double ax = 0.0, ay = 0.0, az = 0.0;
    if (w <= 2) {
        if (w == 2) {   // the list partials, in list order
            const double4 r0 = rs[(size_t)(I * kEwWaves) * kTile + lane];
            ax = r0.x;
            ay = r0.y;
            az = r0.z;
            for (int u = 1; u < kEwWaves; u++) {
                const double4 r = rs[(size_t)(I * kEwWaves + u) * kTile + lane];
                ax += r.x;
                ay += r.y;
                az += r.z;
            }
        }
        double ph, g1[3], g2[6];
        const int a0 = 2 * w, a1 = w == 2 ? kEwOrder : 2 * w + 2;
        ew_gather<double, 0>(E, reinterpret_cast<const double*>(E.phi_v), pi, ph, g1, g2, a0, a1);
        ax -= g1[0] * (E.K[0] / E.L[0]);
        ay -= g1[1] * (E.K[1] / E.L[1]);
        az -= g1[2] * (E.K[2] / E.L[2]);
    } else {
        const double4 vi = vec[i];
        ax = E.self_field * vi.x;
        ay = E.self_field * vi.y;
        az = E.self_field * vi.z;
        if (valid) {
            const double iai = inva[i];
            const double iL0 = 1.0 / E.L[0], iL1 = 1.0 / E.L[1], iL2 = 1.0 / E.L[2];
            for (int q = E.sc_off[i]; q < E.sc_off[i + 1]; q++) {
                const double is = E.sc_s[q].y;
                if (is == 1.0) continue;
                const int j = E.sc_j[q];
                const double4 pj = posa[j];
                const double dx = ew_image_fast(__dsub_rn(pi.x, pj.x), E.L[0], iL0);
                const double dy = ew_image_fast(__dsub_rn(pi.y, pj.y), E.L[1], iL1);
                const double dz = ew_image_fast(__dsub_rn(pi.z, pj.z), E.L[2], iL2);
                const double r2 = ew_r2(dx, dy, dz);
                const bool inside = r2 < E.rc2;
                EwR<double> R;
                ew_radial(E, r2, pi.w, pj.w, iai, inva[j], inside, true, R);
                const bool ci = inside || is != 1.0;
                const double Cmm1 = R.B1 + (ci ? (is * R.l3 - 1.0) * R.b1 : 0.0);
                const double Cmm2 = R.B2 + (ci ? (is * R.l5 - 1.0) * R.b2 : 0.0);
                const double4 vj = vec[j];
                const double vd = vj.x * dx + vj.y * dy + vj.z * dz;
                ax += Cmm2 * vd * dx - Cmm1 * vj.x;
                ay += Cmm2 * vd * dy - Cmm1 * vj.y;
                az += Cmm2 * vd * dz - Cmm1 * vj.z;
            }
        }
    }
    if (w > 0) {
        s_part[w - 1][0][lane] = ax;
        s_part[w - 1][1][lane] = ay;
        s_part[w - 1][2][lane] = az;
    }
    __syncthreads();
    if (w != 0) return;
    for (int u = 0; u < 3; u++) {
        ax += s_part[u][0][lane];
        ay += s_part[u][1][lane];
        az += s_part[u][2][lane];
    }
    if (!valid) ax = ay = az = 0.0;
    if (pq.q) {   // PCG: q = p/alpha - T p and the tile sum of p.q
        const double4 p = pq.p[i];
        const double a = pq.alpha[i];
        const double ia = a > 0.0 ? 1.0 / a : 0.0;
        const double4 qv = make_double4(p.x * ia - ax, p.y * ia - ay, p.z * ia - az, 0.0);
        pq.q[i] = qv;
        const double d = wave_sum(p.x * qv.x + p.y * qv.y + p.z * qv.z);
        if (lane == 0) pq.red[I] = d;
    } else {
        out4[i] = make_double4(ax, ay, az, 0.0);
    }
}

// ------------------------------------------------------------------ launchers
void launch_ew_neighbors(hipStream_t s, const EwLaunch& E) {
    hipLaunchKernelGGL(k_ew_neighbors, dim3(E.nblk), dim3(64), 0, s, E);
}

namespace {
template <typename T>
void fft_pass(hipStream_t s, const EwLaunch& E, int kind, void* phi, const double* scalars, bool small) {
    using C = typename Cplx<T>::type;
    const int axis = (kind == 0 || kind == 4) ? 2 : ((kind == 1 || kind == 3) ? 1 : 0);
    const int n = E.K[axis];
    const int K2h = E.K[2] / 2 + 1;   // the complex grid's z extent (k_ew_fft)
    const int nlines = axis == 2 ? E.K[0] * E.K[1] : (axis == 1 ? E.K[0] * K2h : E.K[1] * K2h);
    // small: two lines per one-wave workgroup (4 KB of LDS for 64-point FP64 lines).  The T v pass's
    // workgroups leave a few KB of LDS and a wave slot per CU free, and small workgroups start there
    // as soon as they are launched (512 workgroups of eight lines, 16 KB, waited for the pass's
    // workgroups to retire: 50-70 us for a pass that takes 8 us alone).  Beside the force pass,
    // which is bound by its VALU work, the chain does better waiting (it still ends first).
    int LB, threads;
    if (small) {
        LB = std::max(1, std::min(2, (int)(4096 / sizeof(C)) / n));
        threads = 64;
    } else {   // at most 24 KB per LDS buffer, and enough workgroups (512) to spread over the CUs
        LB = (int)(24576 / sizeof(C)) / n;
        LB = std::min(LB, std::max(1, nlines / 512));
        LB = std::max(1, std::min(LB, 32));
        threads = 256;
    }
    const size_t shm = 2 * (size_t)LB * n * sizeof(C);
    const dim3 grid((nlines + LB - 1) / LB), block(threads);
    switch (kind) {
        case 0: hipLaunchKernelGGL((k_ew_fft<T, 0>), grid, block, shm, s, E, LB, phi, scalars); break;
        case 1: hipLaunchKernelGGL((k_ew_fft<T, 1>), grid, block, shm, s, E, LB, phi, scalars); break;
        case 2: hipLaunchKernelGGL((k_ew_fft<T, 2>), grid, block, shm, s, E, LB, phi, scalars); break;
        case 3: hipLaunchKernelGGL((k_ew_fft<T, 3>), grid, block, shm, s, E, LB, phi, scalars); break;
        default: hipLaunchKernelGGL((k_ew_fft<T, 4>), grid, block, shm, s, E, LB, phi, scalars); break;
    }
}
}  // namespace

void launch_ew_recip(hipStream_t s, const EwLaunch& E, const double4* posa, const double* charge, const double4* vec,
                     void* phi, const double* scalars, bool small_fft) {
    hipLaunchKernelGGL(k_ew_spread, dim3(E.nblk), dim3(64), 0, s, E, posa, charge, vec, scalars);
    for (int kind = 0; kind < 5; kind++) {
        if (E.mixed) fft_pass<float>(s, E, kind, phi, scalars, small_fft);
        else fft_pass<double>(s, E, kind, phi, scalars, small_fft);
    }
}

void launch_ew_pairs(hipStream_t s, const EwLaunch& E, int mode, bool mutual, const double4* posa, const double* inva,
                     const double* charge, const double4* vec, const double4* E0, const double* scalars, double4* out4,
                     const PcgQ& pq, double* epart, double* frc, double* eout) {
    const dim3 grid(E.nblk * (mode == kEwE0 ? ew_parts(kEwE0) : kEwParts)), block(kEwWaves * 64), rows(E.nblk);
#define TDP_EW_PAIRS(M, MU, MX) \
    hipLaunchKernelGGL((k_ew_pairs<M, MU, MX>), grid, block, 0, s, E, posa, inva, charge, vec, scalars, E.rs)
#define TDP_EW_EPI(M, MU)                                                                                      \
    hipLaunchKernelGGL((k_ew_epilogue<M, MU>), rows, dim3(256), 0, s, E, posa, inva, charge, vec, E0, scalars, E.rs, \
                       out4, pq, epart, frc, eout)
    // the permanent field stays in FP64 in both precisions (cancellation of the molecules' charge fields)
    if (mode == kEwE0) {
        TDP_EW_PAIRS(kEwE0, false, false);
        TDP_EW_EPI(kEwE0, false);
    } else if (mode == kEwTmu) {
        if (E.pairs32) TDP_EW_PAIRS(kEwTmu, false, true);
        else TDP_EW_PAIRS(kEwTmu, false, false);
        TDP_EW_EPI(kEwTmu, false);
    } else if (mutual) {
        if (E.pairs32) TDP_EW_PAIRS(kEwForce, true, true);
        else TDP_EW_PAIRS(kEwForce, true, false);
        TDP_EW_EPI(kEwForce, true);
    } else {
        if (E.pairs32) TDP_EW_PAIRS(kEwForce, false, true);
        else TDP_EW_PAIRS(kEwForce, false, false);
        TDP_EW_EPI(kEwForce, false);
    }
#undef TDP_EW_PAIRS
#undef TDP_EW_EPI
}

void launch_ew_e0_stored(hipStream_t s, const EwLaunch& E, const double4* posa, const double* inva,
                         const double* charge, double4* out4, double* epart, hipEvent_t join) {
    EwLaunch Ep = E;
    Ep.e0_parts = kEwE0Parts;
    const dim3 grid(E.nblk * kEwE0Parts), block(kEwE0Waves * 64);
    if (E.pairs32)
        hipLaunchKernelGGL((k_ew_e0_build<true>), grid, block, 0, s, Ep, posa, inva, charge, E.rs);
    else
        hipLaunchKernelGGL((k_ew_e0_build<false>), grid, block, 0, s, Ep, posa, inva, charge, E.rs);
    if (join) hipStreamWaitEvent(s, join, 0);
    hipLaunchKernelGGL((k_ew_epilogue<kEwE0, false>), dim3(E.nblk), dim3(256), 0, s, Ep, posa, inva, charge,
                       (const double4*)nullptr, (const double4*)nullptr, (const double*)nullptr, E.rs, out4, PcgQ(),
                       epart, (double*)nullptr, (double*)nullptr);
}

void launch_ew_force_stored(hipStream_t s, const EwLaunch& E, bool mutual, const double4* posa, const double* inva,
                            const double* charge, const double4* vec, const double4* E0, double* epart, double* frc,
                            double* eout, hipEvent_t join) {
    static_assert(kEwE0Parts == ew_parts(kEwForce), "the force epilogue adds the parts of the force pass");
    const dim3 grid(E.nblk * kEwE0Parts), block(kEwE0Waves * 64), rows(E.nblk);
    const size_t shm = E.pairs32 ? (size_t)E.lds_tiles * sizeof(double4) : 0;
    if (E.pairs32)
        hipLaunchKernelGGL(k_ew_force_records, dim3((E.npad + 255) / 256), dim3(256), 0, s, E.npad, posa, inva, charge, vec,
                           E.mu32, E.slot32);
    if (mutual) {
        if (E.pairs32) hipLaunchKernelGGL((k_ew_force_cand<true, true>), grid, block, shm, s, E, posa, inva, charge, vec, E.rs);
        else hipLaunchKernelGGL((k_ew_force_cand<true, false>), grid, block, 0, s, E, posa, inva, charge, vec, E.rs);
        if (join) hipStreamWaitEvent(s, join, 0);
        hipLaunchKernelGGL((k_ew_epilogue<kEwForce, true>), rows, dim3(256), 0, s, E, posa, inva, charge, vec, E0,
                           (const double*)nullptr, E.rs, (double4*)nullptr, PcgQ(), epart, frc, eout);
    } else {
        if (E.pairs32) hipLaunchKernelGGL((k_ew_force_cand<false, true>), grid, block, shm, s, E, posa, inva, charge, vec, E.rs);
        else hipLaunchKernelGGL((k_ew_force_cand<false, false>), grid, block, 0, s, E, posa, inva, charge, vec, E.rs);
        if (join) hipStreamWaitEvent(s, join, 0);
        hipLaunchKernelGGL((k_ew_epilogue<kEwForce, false>), rows, dim3(256), 0, s, E, posa, inva, charge, vec, E0,
                           (const double*)nullptr, E.rs, (double4*)nullptr, PcgQ(), epart, frc, eout);
    }
}

constexpr size_t kEwSpmvMinLds = 22 * 1024 + 512;   // 7 x 23,040 B <= 160 KB < 8 x 23,040 B

void launch_ew_spmv(hipStream_t s, const EwLaunch& E, const double4* vec, const float4* v32, const double* scalars) {
    const dim3 lists(E.nblk * kEwWaves);
    if (E.pairs32)
        // at least 22.5 KB of LDS per workgroup: at most 7 of the pass's one-wave workgroups per CU (160 KB),
        // which leaves a SIMD with room for the reciprocal chain's FFT workgroups beside it (with 8, its
        // register limit, the pass took 115 us instead of 121 but the chain waited and the evaluation
        // took 4.02 instead of 3.93 ms at 36,864 sites; with 6, 26 KB each, 3.93 instead of 3.85 ms)
        hipLaunchKernelGGL((k_ew_spmv<true>), lists, dim3(64), std::max((size_t)E.lds_slots * kEwSlotLds, kEwSpmvMinLds), s,
                           E, vec, v32,
                           scalars, E.rsw);
    else
        hipLaunchKernelGGL((k_ew_spmv<false>), lists, dim3(64), (size_t)E.lds_tiles * 8 * kEwSlotLds, s, E, vec, v32,
                           scalars, E.rsw);
}

void launch_ew_tmu_epilogue(hipStream_t s, const EwLaunch& E, const double4* posa, const double* inva,
                            const double* charge, const double4* vec, const double* scalars, double4* out4,
                            const PcgQ& pq) {
    (void)charge;
    hipLaunchKernelGGL(k_ew_tmu_epilogue, dim3(E.nblk), dim3(256), 0, s, E, posa, inva, vec, scalars, E.rsw, out4, pq);
}

void launch_ew_energy(hipStream_t s, int npad, const double* eout, double background, double* energy) {
    hipLaunchKernelGGL(k_ew_energy, dim3(1), dim3(1024), 0, s, npad, eout, background, energy);
}

void launch_ew_dump_pairs(hipStream_t s, const EwLaunch& E, const double4* posa, const int* perm, int* pairs,
                          long long cap, unsigned long long* count) {
    hipLaunchKernelGGL(k_ew_dump_pairs, dim3(E.nblk), dim3(64), 0, s, E, posa, perm, pairs, cap, count);
}

}  // namespace tdp
