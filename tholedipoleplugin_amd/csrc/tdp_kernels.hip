// tdp_kernels.hip -- gfx950 kernels of the TholeDipoleForce force + induction path.
//
// Pair kernels visit every pair once: a workgroup is one 64-lane wavefront that owns a tile of
// 64 particles i (one per lane) and sweeps a short range of j tiles J >= I (a work item); each j
// tile is staged once into LDS with coalesced loads.  Lane l keeps the sums of its own particle in
// registers and pairs with slot (l + k) mod 64 at step k, while the j-side sums rotate one lane per
// step (DPP).  Work items write their i sums and each tile pair's j sums to separate partial
// buffers that k_col_reduce adds in a fixed order, so every result is deterministic.
//
// Exclusions/scalings come from per-tile bit masks built on the host (tdp_context.hip):
// for a tile pair (I,J) that holds at least one scaled pair, three 64-bit masks per lane mark
// mScale == 0, mScale == 0.5 and iScale == 0 (self pairs set all of them).  Tiles without masks
// take a branch-free fast path.
#include "tdp_kernels.h"
#include "tdp_device.hpp"

using namespace tdp;

namespace {

struct BoxT64 { double bx, by, bz, ibx, iby, ibz; };

// Scalars shared between the solver kernels (device memory, see SolverScalars in tdp_kernels.h).
__device__ __forceinline__ bool solver_done(const double* sc) { return sc[SC_DONE] != 0.0; }

__device__ __forceinline__ double inv_or_zero(double a) { return a > 0.0 ? 1.0 / a : 0.0; }

// ------------------------------------------------------------------ frame rotation (a3)
// checkChiralCenterAtParticle + applyRotationMatrixToParticle
// (ReferenceTholeDipoleForce.cpp:432-452, :477-546, called for Z>=0 && Z!=i at :556).
__global__ void k_rotate_dipoles(int n, const double4* __restrict__ posa, const double* __restrict__ dmol,
                                 const double* __restrict__ charge, const int4* __restrict__ axis,
                                 double4* __restrict__ qd) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int4 ax = axis[i];  // (type, Z, X, Y)
    double d0 = dmol[3 * i], d1 = dmol[3 * i + 1], d2 = dmol[3 * i + 2];
    double4 pi = posa[i];
    if (ax.w > -1 && ax.x == 0) {  // checkChiral: ZThenX with a Y atom (:437)
        double4 pz = posa[ax.y], px = posa[ax.z], py = posa[ax.w];
        double adx = pi.x - py.x, ady = pi.y - py.y, adz = pi.z - py.z;
        double bdx = pz.x - py.x, bdy = pz.y - py.y, bdz = pz.z - py.z;
        double cdx = px.x - py.x, cdy = px.y - py.y, cdz = px.z - py.z;
        double cx = bdy * cdz - bdz * cdy, cy = bdz * cdx - bdx * cdz, cz = bdx * cdy - bdy * cdx;
        double vol = cx * adx + cy * ady + cz * adz;
        if (vol < 0.0) d1 = -d1;
    }
    double lx = d0, ly = d1, lz = d2;
    if (ax.y >= 0 && ax.y != i) {
        double4 pz = posa[ax.y];
        double zx = pz.x - pi.x, zy = pz.y - pi.y, zz = pz.z - pi.z;
        double nz = sqrt(zx * zx + zy * zy + zz * zz);
        if (nz > 0.0) { double s = 1.0 / nz; zx *= s; zy *= s; zz *= s; }
        double xx, xy, xz;
        if (ax.x == 4) {  // ZOnly
            if (fabs(zx) < 0.866) { xx = 1.0; xy = 0.0; xz = 0.0; } else { xx = 0.0; xy = 1.0; xz = 0.0; }
        } else {
            double4 px = posa[ax.z];
            xx = px.x - pi.x; xy = px.y - pi.y; xz = px.z - pi.z;
            if (ax.x == 1 || ax.x == 2 || ax.x == 3) {
                double nx = sqrt(xx * xx + xy * xy + xz * xz);
                if (nx > 0.0) { double s = 1.0 / nx; xx *= s; xy *= s; xz *= s; }
            }
            if (ax.x == 1) {  // Bisector
                zx += xx; zy += xy; zz += xz;
                double n2 = sqrt(zx * zx + zy * zy + zz * zz);
                if (n2 > 0.0) { double s = 1.0 / n2; zx *= s; zy *= s; zz *= s; }
            } else if (ax.x == 2) {  // ZBisect
                double4 py = posa[ax.w];
                double yx = py.x - pi.x, yy = py.y - pi.y, yz = py.z - pi.z;
                double ny = sqrt(yx * yx + yy * yy + yz * yz);
                if (ny > 0.0) { double s = 1.0 / ny; yx *= s; yy *= s; yz *= s; }
                xx += yx; xy += yy; xz += yz;
                double n2 = sqrt(xx * xx + xy * xy + xz * xz);
                if (n2 > 0.0) { double s = 1.0 / n2; xx *= s; xy *= s; xz *= s; }
            } else if (ax.x == 3) {  // ThreeFold
                double4 py = posa[ax.w];
                double yx = py.x - pi.x, yy = py.y - pi.y, yz = py.z - pi.z;
                double ny = sqrt(yx * yx + yy * yy + yz * yz);
                if (ny > 0.0) { double s = 1.0 / ny; yx *= s; yy *= s; yz *= s; }
                zx += xx + yx; zy += xy + yy; zz += xz + yz;
                double n2 = sqrt(zx * zx + zy * zy + zz * zz);
                if (n2 > 0.0) { double s = 1.0 / n2; zx *= s; zy *= s; zz *= s; }
            }
        }
        double dt = zx * xx + zy * xy + zz * xz;
        xx -= zx * dt; xy -= zy * dt; xz -= zz * dt;
        double nx = sqrt(xx * xx + xy * xy + xz * xz);
        if (nx > 0.0) { double s = 1.0 / nx; xx *= s; xy *= s; xz *= s; }
        double yx = zy * xz - zz * xy, yy = zz * xx - zx * xz, yz = zx * xy - zy * xx;
        lx = d0 * xx + d1 * yx + d2 * zx;
        ly = d0 * xy + d1 * yy + d2 * zy;
        lz = d0 * xz + d1 * yz + d2 * zz;
    }
    qd[i] = make_double4(lx, ly, lz, charge[i]);
}

// ------------------------------------------------------------------ positions
// Gathers the caller's positions (original order) into the internal spatial order:
// slot i holds particle perm[i].
__global__ void k_load_positions(int npad, const double* __restrict__ pos3, const int* __restrict__ perm,
                                 const double* __restrict__ thole, double4* __restrict__ posa) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npad) return;
    double4 p = make_double4(0.0, 0.0, 0.0, 0.0);
    const int o = perm[i];
    if (o >= 0) p = make_double4(pos3[3 * o], pos3[3 * o + 1], pos3[3 * o + 2], thole[i]);
    posa[i] = p;
}

// Vectors between the caller's order (3 doubles per particle) and the internal order (double4 per
// slot, with an optional FP32 copy); slots with perm < 0 are padding (zero).
__global__ void k_gather_vec3(int npad, const int* __restrict__ perm, const double* __restrict__ src3,
                              double4* __restrict__ dst4, float4* __restrict__ dst32) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npad) return;
    const int o = perm[i];
    const double3 v = o >= 0 ? make_double3(src3[3 * o], src3[3 * o + 1], src3[3 * o + 2]) : make_double3(0.0, 0.0, 0.0);
    dst4[i] = make_double4(v.x, v.y, v.z, 0.0);
    if (dst32) dst32[i] = make_float4((float)v.x, (float)v.y, (float)v.z, 0.f);
}

__global__ void k_scatter_vec3(int npad, const int* __restrict__ perm, const double4* __restrict__ src4,
                               double* __restrict__ dst3) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npad) return;
    const int o = perm[i];
    if (o < 0) return;
    const double4 v = src4[i];
    dst3[3 * o] = v.x;
    dst3[3 * o + 1] = v.y;
    dst3[3 * o + 2] = v.z;
}

// 32-bit key of a particle's raw (y, z): equal coordinates (-0 and +0 alike) give equal keys, so
// different keys prove that the positions differ in y or z.  Mixing by the SplitMix64 finaliser.
__device__ __forceinline__ unsigned yz_key(double y, double z) {
    unsigned long long a = (unsigned long long)__double_as_longlong(y + 0.0);
    const unsigned long long b = (unsigned long long)__double_as_longlong(z + 0.0);
    a ^= b * 0x9E3779B97F4A7C15ull;
    a ^= a >> 30;
    a *= 0xBF58476D1CE4E5B9ull;
    a ^= a >> 27;
    a *= 0x94D049BB133111EBull;
    a ^= a >> 31;
    return (unsigned)a;
}

// ------------------------------------------------------------------ tile classification
// Per tile of 64 particles (internal order), one wavefront: centre and half extent of its bounding
// box (PBC: offsets from the tile's first particle wrapped to the minimum image, so a tile that
// straddles the box edge stays compact) and the largest/smallest Thole parameter.  Also the tile
// frame of the FP32 far-field paths: origin o = first particle + box centre (FP64) and each
// particle's offset u from it in FP32, so that x = o + u + n L with n the integer wrap of the
// particle's offset from the first particle (|u| <= half extent per axis); w carries the bits of
// yz_key of the particle's raw position.
__global__ __launch_bounds__(64) void k_tile_bounds(const int* __restrict__ tcnt, const double4* __restrict__ posa,
                                                    double bx, double by, double bz, int pbc,
                                                    float4* __restrict__ tb_c, float4* __restrict__ tb_h,
                                                    double4* __restrict__ tb_o, float4* __restrict__ posl,
                                                    double4* __restrict__ posl64) {
    const int T = blockIdx.x, lane = threadIdx.x;
    const int i = T * kTile + lane;
    const bool valid = lane < tcnt[T];
    const double4 p0 = posa[T * kTile];
    const double4 p = posa[i];
    double d[3] = {p.x - p0.x, p.y - p0.y, p.z - p0.z};
    if (pbc) {
        const double L[3] = {bx, by, bz};
        for (int k = 0; k < 3; k++)
            if (L[k] > 0) d[k] -= L[k] * floor(d[k] / L[k] + 0.5);
    }
    double lo[3], hi[3];
    for (int k = 0; k < 3; k++) lo[k] = hi[k] = valid ? d[k] : 0.0;   // the first particle has d = 0
    double amax = valid ? p.w : p0.w, amin = amax;
    for (int m = 1; m < kTile; m <<= 1) {
        for (int k = 0; k < 3; k++) {
            lo[k] = fmin(lo[k], __shfl_xor(lo[k], m));
            hi[k] = fmax(hi[k], __shfl_xor(hi[k], m));
        }
        amax = fmax(amax, __shfl_xor(amax, m));
        amin = fmin(amin, __shfl_xor(amin, m));
    }
    const double c[3] = {0.5 * (lo[0] + hi[0]), 0.5 * (lo[1] + hi[1]), 0.5 * (lo[2] + hi[2])};
    // padding slots sit 1e6 nm from the tile origin: their FP32 pairs stay finite (and carry zero
    // charge, dipole and Thole parameters), so the FP32 paths need no masks on partial tiles
    posl[i] = valid ? make_float4((float)(d[0] - c[0]), (float)(d[1] - c[1]), (float)(d[2] - c[2]),
                                  __uint_as_float(yz_key(p.y, p.z)))
                    : make_float4(1e6f, 1e6f, 1e6f, 0.f);
    if (posl64) posl64[i] = valid ? make_double4(d[0] - c[0], d[1] - c[1], d[2] - c[2], 0.0) : make_double4(1e6, 1e6, 1e6, 0.0);
    if (lane == 0) {
        tb_c[T] = make_float4((float)(p0.x + c[0]), (float)(p0.y + c[1]), (float)(p0.z + c[2]), (float)amax);
        // half extents rounded up so the box still contains every particle after the float conversion
        tb_h[T] = make_float4((float)(0.5 * (hi[0] - lo[0])) * 1.0001f + 1e-6f,
                              (float)(0.5 * (hi[1] - lo[1])) * 1.0001f + 1e-6f,
                              (float)(0.5 * (hi[2] - lo[2])) * 1.0001f + 1e-6f, (float)amin);
        tb_o[T] = make_double4(p0.x + c[0], p0.y + c[1], p0.z + c[2], 0.0);
    }
}

// True when every pair of tiles I and J lies beyond the Thole range, i.e. the reference's
// exponent s = -gamma (r/(a_i a_j))^3 is below -50 for all of them and no damping applies
// (getAndScaleInverseRs, ReferenceTholeDipoleForce.cpp:660-675; the same bound gives a u^3 >= 50
// in calculateElectrostaticPairIxn, where every Thole factor is then exactly 1 in FP64).  Uses a
// lower bound of the minimum-image distance between the two bounding boxes, with a 2% margin.
__device__ __forceinline__ bool tiles_beyond_damping(float4 cI, float4 hI, float4 cJ, float4 hJ, bool pbc, float bx,
                                                     float by, float bz) {
    float dx = cJ.x - cI.x, dy = cJ.y - cI.y, dz = cJ.z - cI.z;
    if (pbc) {
        dx -= bx * rintf(dx / bx);
        dy -= by * rintf(dy / by);
        dz -= bz * rintf(dz / bz);
    }
    dx = fmaxf(0.f, fabsf(dx) - hI.x - hJ.x);
    dy = fmaxf(0.f, fabsf(dy) - hI.y - hJ.y);
    dz = fmaxf(0.f, fabsf(dz) - hI.z - hJ.z);
    float rmin2 = dx * dx + dy * dy + dz * dz;
    float dmax = cI.w * cJ.w;           // largest a_i a_j
    float gmin = fminf(hI.w, hJ.w);      // smallest gamma
    if (dmax == 0.f) return true;        // no pair is damped
    if (!(gmin > 0.f)) return false;
    float rd = dmax * cbrtf(50.f / gmin) * 1.02f;
    return rmin2 > rd * rd;
}

__device__ __forceinline__ void load_site(ElecSite& s, double4 p, double4 q, double4 u, double ia) {
    s.x = p.x; s.y = p.y; s.z = p.z; s.a = p.w; s.ia = ia;
    s.dx = q.x; s.dy = q.y; s.dz = q.z; s.q = q.w;
    s.ux = u.x; s.uy = u.y; s.uz = u.z;
}

// ------------------------------------------------------------------ symmetric pair sweep
// calculateInducedDipoleFields + calculateInducedDipolePairIxn (ReferenceTholeDipoleForce.cpp:766-819),
// calculateFixedDipoleField (:715-733) and calculateElectrostatic (:368-393) visit every pair once
// (i < j).  The kernels below do the same over 64 x 64 tiles of the internal order.  Work item w
// covers the tiles (I, J) for J in [J0, J1), J >= I: lane l owns particle i = 64 I + l (its sums
// stay in registers), and at step k evaluates the pair with slot (l + k) mod 64 of the staged j
// tile.  The j-side sums travel with their slot: after every step they rotate one lane down the
// wavefront (DPP wave_rol:1), so lane l always holds the sums of slot (l + k) mod 64.  On the
// diagonal tile steps k = 1..32 cover each pair once (lanes >= 32 sit out k = 32).  Each tile's
// j sums go to the column-major partial slab Pj (component c of slot s at Pj[c S + s]), each
// item's i sums to Pi[w]; k_col_reduce adds them in a fixed order, so results are deterministic.

// Last rotation step of a diagonal tile holding cnt sites.  Steps 1..32 cover every pair (a, b), a < b,
// of a full tile: at step b - a from lane a when b - a <= 32, at step 64 - (b - a) from lane b otherwise.
// With fewer sites the largest step needed is min(32, cnt - 1); the steps left out would evaluate no
// pair, so the sums are the same.
__device__ __forceinline__ int diag_last_step(int cnt) { return cnt > kTile / 2 ? kTile / 2 : cnt - 1; }

// Tile (I, J), I <= J, of one system in column-major upper-triangle order: the tiles of column J
// are contiguous, starting at col[J] with the system's first tile.
__device__ __forceinline__ size_t slab_tile(const TileMap& tm, int I, int J) {
    return tm.col[J] + (size_t)(I - tm.first[J]);
}

__device__ __forceinline__ void store3(double* __restrict__ P, size_t S, size_t slot, double x, double y, double z) {
    P[slot] = x;
    P[S + slot] = y;
    P[2 * S + slot] = z;
}

// value of lane (l + 1) mod 64 (all lanes must be active)
__device__ __forceinline__ float rot1(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x134, 0xF, 0xF, true));
}
__device__ __forceinline__ double rot1(double v) {
    long long u = __double_as_longlong(v);
    int lo = __builtin_amdgcn_mov_dpp((int)(u & 0xffffffffll), 0x134, 0xF, 0xF, false);
    int hi = __builtin_amdgcn_mov_dpp((int)(u >> 32), 0x134, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// damped_rr in FP64 for a full wavefront: the exponential is evaluated only when some lane of
// the wavefront has a damped pair (s > -50), which in tiles near the Thole range is a minority
// of the steps.  Same results as damped_rr<double>.  All lanes must be active.
__device__ __forceinline__ void damped_rr_wave(double r2, double ai, double aj, double iai, double iaj, double& rr3,
                                               double& rr5) {
    double rinv = rsqrt(r2);
    double r = r2 * rinv;
    double rinv2 = rinv * rinv;
    rr3 = rinv * rinv2;
    rr5 = 3.0 * rr3 * rinv2;
    double g = ai < aj ? ai : aj;
    double ratio = r * (iai * iaj);
    double s = -g * (ratio * ratio * ratio);
    const bool damped = (ai * aj != 0.0) && (s > -50.0);
    if (__builtin_amdgcn_ballot_w64(damped)) {
        double e = exp(damped ? s : -50.0);
        if (damped) {
            rr3 *= 1.0 - e;
            rr5 *= 1.0 - (1.0 - s) * e;
        }
    }
}

// ------------------------------------------------------------------ T * mu, FP64 (a7)
template <bool PBC>
__global__ __launch_bounds__(64) void k_sym_matvec(TileMap tm, int nblk, const int4* __restrict__ work,
                                                   const double4* __restrict__ posa, const double4* __restrict__ mu,
                                                   const double* __restrict__ inva, const int* __restrict__ exc_start,
                                                   const int* __restrict__ exc_J,
                                                   const unsigned long long* __restrict__ exc_mask,
                                                   const float4* __restrict__ tb_c, const float4* __restrict__ tb_h, const TilePair* __restrict__ tp,
                                                   const double4* __restrict__ posl64, BoxT64 box,
                                                   const double* __restrict__ scalars,
                                                   double4* __restrict__ Pi, double* __restrict__ Pj, size_t S) {
    __shared__ double4 s_p[kTile];
    __shared__ double4 s_u[kTile];
    __shared__ double4 s_o[kTile];   // offsets of the j tile's particles from its origin
    __shared__ double s_ia[kTile];
    const int lane = threadIdx.x, w = blockIdx.x;
    const int4 wk = work[w];
    const int I = wk.x, J0 = wk.y, J1 = wk.z;
    if (scalars && solver_done(scalars + SC_COUNT * tm.seg[I])) return;
    const int i = I * kTile + lane;
    const int cntI = tm.cnt[I];
    const bool ivalid = lane < cntI;
    const double4 pi = posa[i];
    const double4 ui = mu[i];
    const double iai = inva[i];
    const double4 oi = posl64[i];
    const float4 cI = tb_c[I], hI = tb_h[I];
    const bool ifull = cntI == kTile;
    double ax = 0.0, ay = 0.0, az = 0.0;
    int e = exc_start[I];
    const int eEnd = exc_start[I + 1];
    while (e < eEnd && exc_J[e] < J0) e++;
    for (int J = J0; J < J1; J++) {
        const TilePair tpr = tp[slab_tile(tm, I, J)];
        const int jb = J * kTile;
        const int cntJ = tm.cnt[J];
        __syncthreads();
        s_p[lane] = posa[jb + lane];
        s_u[lane] = mu[jb + lane];
        s_o[lane] = posl64[jb + lane];
        s_ia[lane] = inva[jb + lane];
        __syncthreads();
        const bool diag = J == I;
        const bool masked = (e < eEnd && exc_J[e] == J);
        unsigned long long mi = 0ull;
        if (masked) {
            mi = exc_mask[((size_t)e * 3 + 2) * kTile + lane];
            e++;
        }
        double jx = 0.0, jy = 0.0, jz = 0.0;   // j side of slot (lane + k) mod 64
        int kend;
        if (!masked && !diag && ifull && cntJ == kTile &&
            (tpr.flags & kTpFar)) {
            // a tile pair on a constant frame (every pair takes the image of the tile origins'
            // separation D, none near half a box): d = o_j - (o_i - D) from the offsets, as in the
            // FP32 far path, one subtraction per component; else the minimum image of each pair
            const bool frame = (tpr.flags & (kTpStraddle | kTpGeneral)) == 0u;
            const double six = oi.x - tpr.shx, siy = oi.y - tpr.shy, siz = oi.z - tpr.shz;
#pragma unroll 4
            for (int k = 0; k < kTile; k++) {
                const int jj = (lane + k) & (kTile - 1);
                const double4 uj = s_u[jj];
                double dx, dy, dz;
                if (frame) {
                    const double4 oj = s_o[jj];
                    dx = oj.x - six;
                    dy = oj.y - siy;
                    dz = oj.z - siz;
                } else {
                    const double4 pj = s_p[jj];
                    dx = pj.x - pi.x;
                    dy = pj.y - pi.y;
                    dz = pj.z - pi.z;
                    if (PBC) min_image(dx, dy, dz, box.bx, box.by, box.bz, box.ibx, box.iby, box.ibz);
                }
                double r2 = dx * dx + dy * dy + dz * dz;
                double rinv = rsqrt(r2);
                double rinv2 = rinv * rinv;
                double rr3 = rinv * rinv2;
                double rr5 = 3.0 * rr3 * rinv2;
                double udj = rr5 * (uj.x * dx + uj.y * dy + uj.z * dz);
                double udi = rr5 * (ui.x * dx + ui.y * dy + ui.z * dz);
                ax += udj * dx - rr3 * uj.x;
                ay += udj * dy - rr3 * uj.y;
                az += udj * dz - rr3 * uj.z;
                jx += udi * dx - rr3 * ui.x;
                jy += udi * dy - rr3 * ui.y;
                jz += udi * dz - rr3 * ui.z;
                jx = rot1(jx);
                jy = rot1(jy);
                jz = rot1(jz);
            }
            kend = kTile;
        } else {
            const int k0 = diag ? 1 : 0, k1 = diag ? diag_last_step(cntI) : kTile - 1;
            for (int k = k0; k <= k1; k++) {
                const int jj = (lane + k) & (kTile - 1);
                const bool ok = ivalid && (jj < cntJ) && !((mi >> jj) & 1ull) &&
                                !(diag && k == kTile / 2 && lane >= kTile / 2);
                const double4 pj = s_p[jj];
                const double4 uj = s_u[jj];
                double dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
                if (PBC) min_image(dx, dy, dz, box.bx, box.by, box.bz, box.ibx, box.iby, box.ibz);
                double r2 = ok ? dx * dx + dy * dy + dz * dz : 1.0;
                double rr3, rr5;
                damped_rr_wave(r2, pi.w, pj.w, iai, s_ia[jj], rr3, rr5);
                if (!ok) {
                    rr3 = 0.0;
                    rr5 = 0.0;
                }
                double udj = rr5 * (uj.x * dx + uj.y * dy + uj.z * dz);
                double udi = rr5 * (ui.x * dx + ui.y * dy + ui.z * dz);
                ax += udj * dx - rr3 * uj.x;
                ay += udj * dy - rr3 * uj.y;
                az += udj * dz - rr3 * uj.z;
                jx += udi * dx - rr3 * ui.x;
                jy += udi * dy - rr3 * ui.y;
                jz += udi * dz - rr3 * ui.z;
                jx = rot1(jx);
                jy = rot1(jy);
                jz = rot1(jz);
            }
            kend = k1 + 1;
        }
        // after the last rotation this lane holds the sums of slot (lane + kend) mod 64
        store3(Pj, S, slab_tile(tm, I, J) * kTile + ((lane + kend) & (kTile - 1)), jx, jy, jz);
    }
    Pi[(size_t)w * kTile + lane] = make_double4(ax, ay, az, 0.0);
}

// ------------------------------------------------------------------ FP32 far-field frames
// For tiles I and J with origins oI, oJ (k_tile_bounds) let D = oJ - oI, reduced by the minimum-image
// rule.  The separation of a pair (i, j) is then D + u_j - u_i up to whole boxes, with |u| bounded by
// the tile half extents hI, hJ.  Per axis:
//  - constant: |D| + hI + hJ < L/2.  Every pair takes the image of D, exactly as the FP64 rule
//    d - L floor(d/L + 1/2) applied to the raw separation would pick it (the margin between the pair
//    and the half box exceeds the rounding of the rule), and the imaged separation is u_j - s_i with
//    the per-lane constant s_i = u_i - D (one FP32 subtraction per pair).
//  - straddling: the pairs may fall on either side of the half box.  With h = copysign(L/2, D) and
//    s_i = u_i - (D - h), y = u_j - s_i is the signed distance of the pair past the half box, and
//    the imaged separation is y - copysign(L/2, y).  Pairs with |y| inside a band that exceeds the
//    FP32 rounding of y cannot be decided in FP32: the tile pair is then redone from FP64
//    positions (rare: a pair within ~1e-6 nm of half a box).
//  - general (both boxes together span half a box or more): FP64 separations for the tile pair.
// Non-periodic systems have constant frames only.
struct FarFrame {
    double shx, shy, shz;   // per-lane constant: s_i = u_i - sh
    float hx, hy, hz;       // L/2 of the straddling axes
    float bx, by, bz;       // undecidable band of the straddling axes
    int smask;              // straddling axes (bit 0 x, 1 y, 2 z); -1: general
};

__device__ __forceinline__ double frame_axis(double D, float hI, float hJ, double L, double iL, bool pbc, int bit,
                                             FarFrame& f, float& h, float& band) {
    h = 0.f;
    band = 0.f;
    if (!pbc) return D;
    D -= L * floor(D * iL + 0.5);
    const double w = (double)hI + (double)hJ;
    if (w >= 0.5 * L) {
        f.smask = -1;
        return D;
    }
    if (fabs(D) + w < 0.5 * L) return D;
    if (f.smask >= 0) f.smask |= bit;
    h = (float)(0.5 * L);
    band = (float)((2.0 * hI + hJ + 1.0) * 0x1p-20);
    return D - copysign(0.5 * L, D);
}

__device__ __forceinline__ FarFrame far_frame(double4 oI, double4 oJ, float4 hI, float4 hJ, bool pbc, const BoxT64& box) {
    FarFrame f;
    f.smask = 0;
    f.shx = frame_axis(oJ.x - oI.x, hI.x, hJ.x, box.bx, box.ibx, pbc, 1, f, f.hx, f.bx);
    f.shy = frame_axis(oJ.y - oI.y, hI.y, hJ.y, box.by, box.iby, pbc, 2, f, f.hy, f.by);
    f.shz = frame_axis(oJ.z - oI.z, hI.z, hJ.z, box.bz, box.ibz, pbc, 4, f, f.hz, f.bz);
    return f;
}

// The frame of a tile pair from its record (the same values far_frame gives).
__device__ __forceinline__ FarFrame frame_of(const TilePair& t, const BoxT64& box) {
    FarFrame f;
    f.shx = t.shx;
    f.shy = t.shy;
    f.shz = t.shz;
    f.hx = (t.flags & 1u) ? (float)(0.5 * box.bx) : 0.f;
    f.hy = (t.flags & 2u) ? (float)(0.5 * box.by) : 0.f;
    f.hz = (t.flags & 4u) ? (float)(0.5 * box.bz) : 0.f;
    f.bx = t.bx;
    f.by = t.by;
    f.bz = t.bz;
    f.smask = (t.flags & kTpGeneral) ? -1 : (int)(t.flags & kTpStraddle);
    return f;
}

// imaged FP32 separation component on a constant (S = false) or straddling (S = true) axis
template <bool S>
__device__ __forceinline__ float far_sep(float uj, float si, float h, float band, bool& amb) {
    float d = uj - si;
    if (S) {
        amb = amb || (fabsf(d) < band);
        d -= copysignf(h, d);
    }
    return d;
}

// Far tile pair of T * mu in FP32 on frame SM (straddling-axis mask): s_l, s_u hold the j tile's
// offsets and FP32 dipoles twice over (slot (lane + k) mod 64 is entry lane + k).  i side as two
// sums per component (sum udj*d and sum rr3*u_j), j side lagging one step.
template <int SM>
__device__ __forceinline__ void far_matvec_frame(const float4* __restrict__ s_l, const float4* __restrict__ s_u,
                                                 int lane, float si_x, float si_y, float si_z, const FarFrame& f,
                                                 float4 uf, float3 uf3, float& tx, float& ty, float& tz, float& sx3,
                                                 float& sy3, float& sz3, float& fx, float& fy, float& fz, bool& amb) {
    const float hx = f.hx, hy = f.hy, hz = f.hz, bx = f.bx, by = f.by, bz = f.bz;
#pragma unroll 4
    for (int k = 0; k < kTile; k++) {
        const int jj = lane + k;
        const float4 pj = s_l[jj];
        const float4 uj = s_u[jj];
        const float dx = far_sep<(SM & 1) != 0>(pj.x, si_x, hx, bx, amb);
        const float dy = far_sep<(SM & 2) != 0>(pj.y, si_y, hy, by, amb);
        const float dz = far_sep<(SM & 4) != 0>(pj.z, si_z, hz, bz, amb);
        float r2 = dx * dx + dy * dy + dz * dz;
        float rinv = rsqrt_t(r2);
        float rinv2 = rinv * rinv;
        float rr3 = rinv * rinv2;
        float rr5 = rr3 * rinv2;   // 1/r^5: the factor 3 sits in uf3 and in the flush of tx..tz
        float udj = rr5 * (uj.x * dx + uj.y * dy + uj.z * dz);
        float udi = rr5 * (uf3.x * dx + uf3.y * dy + uf3.z * dz);
        tx += udj * dx;
        ty += udj * dy;
        tz += udj * dz;
        sx3 += rr3 * uj.x;
        sy3 += rr3 * uj.y;
        sz3 += rr3 * uj.z;
        // before step k, lane l holds slot (l + k - 1); the rotation and the add form one DPP instruction
        fx = rot1(fx) + (udi * dx - rr3 * uf.x);
        fy = rot1(fy) + (udi * dy - rr3 * uf.y);
        fz = rot1(fz) + (udi * dz - rr3 * uf.z);
    }
}

// Squared separation beyond which no pair of tiles I and J is damped by more than 3e-12 relative
// (s = -gamma (r/(a_i a_j))^3 <= -30, so 1 - (1 - s) e^s differs from 1 by less than 3e-12): the
// damping factors are 1 to FP32 precision there.  Bound from the tiles' largest and smallest a
// with a 2% margin; 0 when no pair is damped, infinite when a zero a could meet a nonzero one.
__device__ __forceinline__ float undamped_r2_fp32(float4 cI, float4 hI, float4 cJ, float4 hJ) {
    const float dmax = cI.w * cJ.w, gmin = fminf(hI.w, hJ.w);
    if (dmax == 0.f) return 0.f;
    if (!(gmin > 0.f)) return __builtin_inff();
    const float rd = dmax * cbrtf(30.f / gmin) * 1.02f;
    return rd * rd;
}

// Squared separation beyond which the Thole factors of a pair may be evaluated in FP32 (s <= -4,
// r beyond 2.2 damping lengths for gamma = 0.39): the closest pairs, whose large tensor terms
// would carry the FP32 rounding of their separations, stay in FP64.  Same bound as
// undamped_r2_fp32 with exponent 4; infinite when a zero a could meet a nonzero one.
__device__ __forceinline__ float damped32_r2(float4 cI, float4 hI, float4 cJ, float4 hJ) {
    const float dmax = cI.w * cJ.w, gmin = fminf(hI.w, hJ.w);
    if (dmax == 0.f) return 0.f;
    if (!(gmin > 0.f)) return __builtin_inff();
    const float rd = dmax * cbrtf(4.f / gmin) * 1.02f;
    return rd * rd;
}

// Classification of the tile pairs of the work items (one thread per work item, over its tiles J).
// Positions are fixed during an evaluation, so the pair kernels read these records instead of
// recomputing the bounds, frames and radii on every pass.
template <bool PBC>
__global__ __launch_bounds__(64) void k_tile_pairs(TileMap tm, const int4* __restrict__ work, int nwork,
                                                   const float4* __restrict__ tb_c, const float4* __restrict__ tb_h,
                                                   const double4* __restrict__ tb_o, BoxT64 box,
                                                   TilePair* __restrict__ out) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nwork) return;
    const int4 wk = work[w];
    const int I = wk.x;
    const float4 cI = tb_c[I], hI = tb_h[I];
    const double4 oI = tb_o[I];
    for (int J = wk.y; J < wk.z; J++) {
        const float4 cJ = tb_c[J], hJ = tb_h[J];
        const FarFrame fr = far_frame(oI, tb_o[J], hI, hJ, PBC, box);
        TilePair t;
        t.shx = fr.shx;
        t.shy = fr.shy;
        t.shz = fr.shz;
        t.bx = fr.bx;
        t.by = fr.by;
        t.bz = fr.bz;
        unsigned fl = fr.smask < 0 ? kTpGeneral : (unsigned)fr.smask;
        if (tiles_beyond_damping(cI, hI, cJ, hJ, PBC, (float)box.bx, (float)box.by, (float)box.bz)) fl |= kTpFar;
        t.flags = fl;
        t.r2min = undamped_r2_fp32(cI, hI, cJ, hJ);
        t.r2d = damped32_r2(cI, hI, cJ, hJ);
        out[slab_tile(tm, I, J)] = t;
    }
}

// Near tile pair of T * mu in mixed precision (off-diagonal tiles on frame SM, full or partial).
// Pass 1 (near_matvec_mixed32) runs the 64 steps in FP32 like a far tile pair, with the Thole
// factors in FP32 on steps where some pair lies within r2min (every a is then nonzero: r2min is
// infinite otherwise), except steps in which some pair lies within r2d (or a straddling pair
// cannot be decided in FP32): those are returned in a step mask and only rotate the FP32 j sums.  Pass 2 (near_matvec_mixed64) evaluates
// the marked steps in FP64 as k_sym_matvec does; in a step every lane serves a different j slot,
// so its j terms are added to the slot's LDS accumulator s_jn (3 x 64 doubles) without conflicts
// or rotation.
// MASKED: the tile pair holds excluded pairs (bit s of mskip: the pair with slot s has iScale 0);
// a step with such a pair goes to the FP64 pass, which skips it.
template <int SM, bool MASKED>
__device__ __forceinline__ unsigned long long near_matvec_mixed32(const float4* __restrict__ s_l,
                                                                  const float4* __restrict__ s_uf, int lane,
                                                                  float si_x, float si_y, float si_z,
                                                                  const FarFrame& f, float4 uf, float3 uf3,
                                                                  float r2min, float r2d, float2 nai,
                                                                  unsigned long long mskip,
                                                                  double& ax, double& ay, double& az,
                                                                  double& jx, double& jy, double& jz) {
    float tx = 0.f, ty = 0.f, tz = 0.f, sx3 = 0.f, sy3 = 0.f, sz3 = 0.f, fx = 0.f, fy = 0.f, fz = 0.f;
    unsigned long long steps64 = 0ull;
    for (int k = 0; k < kTile; k++) {
        const int jj = lane + k;   // FP32 arrays are staged twice over
        const float4 pl = s_l[jj];
        bool amb = false;
        const float dx = far_sep<(SM & 1) != 0>(pl.x, si_x, f.hx, f.bx, amb);
        const float dy = far_sep<(SM & 2) != 0>(pl.y, si_y, f.hy, f.by, amb);
        const float dz = far_sep<(SM & 4) != 0>(pl.z, si_z, f.hz, f.bz, amb);
        const float r2 = dx * dx + dy * dy + dz * dz;
        if (MASKED) amb = amb || ((mskip >> (jj & (kTile - 1))) & 1ull);
        if (__builtin_amdgcn_ballot_w64(amb || !(r2 > r2d))) {
            steps64 |= 1ull << k;
            fx = rot1(fx);
            fy = rot1(fy);
            fz = rot1(fz);
            continue;
        }
        const float4 uj = s_uf[jj];
        float rinv = rsqrt_t(r2);
        float rinv2 = rinv * rinv;
        float rr3 = rinv * rinv2;
        float rr5 = rr3 * rinv2;   // 1/r^5: the factor 3 sits in uf3 and in the flush of tx..tz
        if (__builtin_amdgcn_ballot_w64(!(r2 > r2min))) {
            // getAndScaleInverseRs in FP32: s = -min(a_i, a_j) (r / (a_i a_j))^3, rr3 *= 1 - e^s,
            // rr5 *= 1 - (1 - s) e^s; for s <= -30 both factors round to 1
            // staged: 1/a_j in the offsets' w, a_j in the dipoles' w
            const float rat = r2 * rinv * (nai.y * pl.w);
            const float sv = -fminf(nai.x, uj.w) * (rat * rat * rat);
            const float ev = __expf(sv);
            const float l3 = 1.f - ev;
            rr3 *= l3;
            rr5 *= fmaf(sv, ev, l3);
        }
        float udj = rr5 * (uj.x * dx + uj.y * dy + uj.z * dz);
        float udi = rr5 * (uf3.x * dx + uf3.y * dy + uf3.z * dz);
        tx += udj * dx;
        ty += udj * dy;
        tz += udj * dz;
        sx3 += rr3 * uj.x;
        sy3 += rr3 * uj.y;
        sz3 += rr3 * uj.z;
        fx = rot1(fx) + (udi * dx - rr3 * uf.x);
        fy = rot1(fy) + (udi * dy - rr3 * uf.y);
        fz = rot1(fz) + (udi * dz - rr3 * uf.z);
    }
    ax += 3.0 * (double)tx - (double)sx3;
    ay += 3.0 * (double)ty - (double)sy3;
    az += 3.0 * (double)tz - (double)sz3;
    // the FP32 j sums lag one step: after 64 steps lane l holds slot l + 63, one more rotation gives slot l
    jx = (double)rot1(fx);
    jy = (double)rot1(fy);
    jz = (double)rot1(fz);
    return steps64;
}

template <bool PBC>
__device__ __forceinline__ void near_matvec_mixed64(unsigned long long steps64, const double4* __restrict__ s_p,
                                                    const double4* __restrict__ s_u, const double* __restrict__ s_ia,
                                                    double* __restrict__ s_jn, int lane, bool ivalid, int cntJ,
                                                    unsigned long long mskip,
                                                    const double4& pi, const double4& ui, double iai,
                                                    const BoxT64& box, double& ax, double& ay, double& az,
                                                    double& jx, double& jy, double& jz) {
    s_jn[lane] = 0.0;
    s_jn[kTile + lane] = 0.0;
    s_jn[2 * kTile + lane] = 0.0;
    __syncthreads();
    while (steps64) {
        const int k = __builtin_ctzll(steps64);
        steps64 &= steps64 - 1ull;
        const int jd = (lane + k) & (kTile - 1);
        const double4 pj = s_p[jd];
        const double4 uj = s_u[jd];
        double ex = pj.x - pi.x, ey = pj.y - pi.y, ez = pj.z - pi.z;
        if (PBC) min_image(ex, ey, ez, box.bx, box.by, box.bz, box.ibx, box.iby, box.ibz);
        // padding slots of partial tiles sit at the origin; excluded pairs (iScale 0) are skipped
        const bool ok = ivalid && jd < cntJ && !((mskip >> jd) & 1ull);
        double rr3, rr5;
        damped_rr_wave(ok ? ex * ex + ey * ey + ez * ez : 1.0, pi.w, pj.w, iai, s_ia[jd], rr3, rr5);
        if (!ok) {
            rr3 = 0.0;
            rr5 = 0.0;
        }
        double udj = rr5 * (uj.x * ex + uj.y * ey + uj.z * ez);
        double udi = rr5 * (ui.x * ex + ui.y * ey + ui.z * ez);
        ax += udj * ex - rr3 * uj.x;
        ay += udj * ey - rr3 * uj.y;
        az += udj * ez - rr3 * uj.z;
        s_jn[jd] += udi * ex - rr3 * ui.x;
        s_jn[kTile + jd] += udi * ey - rr3 * ui.y;
        s_jn[2 * kTile + jd] += udi * ez - rr3 * ui.z;
    }
    __syncthreads();
    jx += s_jn[lane];
    jy += s_jn[kTile + lane];
    jz += s_jn[2 * kTile + lane];
}

// ------------------------------------------------------------------ T * mu, mixed precision (a7)
// Tile pairs beyond the Thole range, most of the pairs of a large system, carry only the undamped
// 1/r^3 and 3/r^5 terms and run in FP32 with FP32 dipoles, on the tile pair's FP32 frame (which
// picks the image of the FP64 rule for every pair, or sends the tile pair to FP64 separations).
// On the other off-diagonal tile pairs, a step whose pairs all lie beyond a u^3 >= 4 also runs in
// FP32, with FP32 Thole factors; the steps with a closer pair or an excluded or scaled pair are
// evaluated afterwards in FP64 (near_matvec_mixed64), as are the diagonal tiles.  The closest pairs
// carry the largest terms, so keeping them in FP64 holds the operator's error near that of the FP32
// far field.  Tile and item partials are stored in FP32, the reductions run in FP64.
template <bool PBC>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void k_sym_matvec_mixed(TileMap tm, int nblk, const int4* __restrict__ work,
                                                         const double4* __restrict__ posa,
                                                         const double4* __restrict__ mu,
                                                         const float4* __restrict__ mu32,
                                                         const double* __restrict__ inva,
                                                         const int* __restrict__ exc_start,
                                                         const int* __restrict__ exc_J,
                                                         const unsigned long long* __restrict__ exc_mask,
                                                         const float4* __restrict__ tb_c, const float4* __restrict__ tb_h, const TilePair* __restrict__ tp,
                                                         const double4* __restrict__ tb_o,
                                                         const float4* __restrict__ posl, BoxT64 box, const double* __restrict__ scalars,
                                                         float* __restrict__ Pjf, size_t S) {
    // One LDS area holds either the near tile (FP64 dipoles and 1/a) or the far tile (FP32
    // dipoles).  The far tile is staged twice over, so that slot (lane + k) mod 64 is entry
    // lane + k: the step's address is the lane's base plus a constant offset.
    // FP32 frames: offsets s_l (float4) at the start, dipoles s_uf behind the FP64 positions s_pf of
    // the general far path, so a fallback from a frame to FP64 positions restages only s_pf.
    // Near tiles in mixed steps add FP32 offsets and dipoles (twice over) behind the FP64 arrays.
    constexpr int kBytesNear = kTile * (2 * sizeof(double4) + sizeof(double));
    constexpr int kBytesNearMixed = kBytesNear + 2 * kTile * 2 * sizeof(float4) + 3 * kTile * sizeof(double);
    constexpr int kBytesFar = 2 * kTile * (sizeof(double4) + sizeof(float4));
    __shared__ __align__(16) unsigned char s_raw[kBytesNearMixed > kBytesFar ? kBytesNearMixed : kBytesFar];
    double4* s_p = reinterpret_cast<double4*>(s_raw);
    double4* s_u = s_p + kTile;
    double* s_ia = reinterpret_cast<double*>(s_u + kTile);
    float4* s_nl = reinterpret_cast<float4*>(s_raw + kBytesNear);
    float4* s_nu = s_nl + 2 * kTile;
    double* s_jn = reinterpret_cast<double*>(s_nu + 2 * kTile);
    double4* s_pf = reinterpret_cast<double4*>(s_raw);
    float4* s_uf = reinterpret_cast<float4*>(s_pf + 2 * kTile);
    float4* s_l = reinterpret_cast<float4*>(s_raw);
    const int lane = threadIdx.x, w = blockIdx.x;
    const int4 wk = work[w];
    const int I = wk.x, J0 = wk.y, J1 = wk.z;
    if (scalars && solver_done(scalars + SC_COUNT * tm.seg[I])) return;
    const int i = I * kTile + lane;
    const int cntI = tm.cnt[I];
    const bool ivalid = lane < cntI;
    const double4 pi = posa[i];
    const double4 ui = mu[i];
    const double iai = inva[i];
    const float4 uf = mu32[i];
    const float3 uf3 = make_float3(3.f * uf.x, 3.f * uf.y, 3.f * uf.z);
    const float4 cI = tb_c[I], hI = tb_h[I];
    const double4 oI = tb_o[I];
    const float4 li = posl[i];
    const float2 nai = make_float2((float)pi.w, (float)iai);
    const bool ifull = cntI == kTile;
    double ax = 0.0, ay = 0.0, az = 0.0;
    int e = exc_start[I];
    const int eEnd = exc_start[I + 1];
    while (e < eEnd && exc_J[e] < J0) e++;
    for (int J = J0; J < J1; J++) {
        const TilePair tpr = tp[slab_tile(tm, I, J)];
        const int jb = J * kTile;
        const int cntJ = tm.cnt[J];
        const bool diag = J == I;
        const bool masked = (e < eEnd && exc_J[e] == J);
        unsigned long long mi = 0ull;
        if (masked) {
            mi = exc_mask[((size_t)e * 3 + 2) * kTile + lane];
            e++;
        }
        const bool far = !masked && !diag && ifull && cntJ == kTile &&
                         (tpr.flags & kTpFar);
        double jx, jy, jz;
        int kend;
        FarFrame fr_n;
        __syncthreads();
        if (far) {
            const FarFrame fr = frame_of(tpr, box);
            {
                const float4 u = mu32[jb + lane];
                s_uf[lane] = u;
                s_uf[lane + kTile] = u;
            }
            float tx = 0.f, ty = 0.f, tz = 0.f, sx3 = 0.f, sy3 = 0.f, sz3 = 0.f, fx = 0.f, fy = 0.f, fz = 0.f;
            bool general = fr.smask < 0;
            if (!general) {
                const float4 l = posl[jb + lane];
                s_l[lane] = l;
                s_l[lane + kTile] = l;
                __syncthreads();
                const float si_x = (float)((double)li.x - fr.shx), si_y = (float)((double)li.y - fr.shy),
                            si_z = (float)((double)li.z - fr.shz);
                bool amb = false;
#define TDP_FAR_CASE(M)                                                                                        \
    case M:                                                                                                    \
        far_matvec_frame<M>(s_l, s_uf, lane, si_x, si_y, si_z, fr, uf, uf3, tx, ty, tz, sx3, sy3, sz3, fx, fy, \
                            fz, amb);                                                                          \
        break;
                switch (fr.smask) {
                    TDP_FAR_CASE(0)
                    TDP_FAR_CASE(1)
                    TDP_FAR_CASE(2)
                    TDP_FAR_CASE(3)
                    TDP_FAR_CASE(4)
                    TDP_FAR_CASE(5)
                    TDP_FAR_CASE(6)
                    TDP_FAR_CASE(7)
                }
#undef TDP_FAR_CASE
                if (__builtin_amdgcn_ballot_w64(amb)) {
                    general = true;
                    tx = ty = tz = sx3 = sy3 = sz3 = fx = fy = fz = 0.f;
                    __syncthreads();
                }
            }
            if (general) {
                const double4 p = posa[jb + lane];
                s_pf[lane] = p;
                s_pf[lane + kTile] = p;
                __syncthreads();
#pragma unroll 4
                for (int k = 0; k < kTile; k++) {
                    const int jj = lane + k;
                    const double4 pj = s_pf[jj];
                    const float4 uj = s_uf[jj];
                    double ex = pj.x - pi.x, ey = pj.y - pi.y, ez = pj.z - pi.z;
                    if (PBC) min_image(ex, ey, ez, box.bx, box.by, box.bz, box.ibx, box.iby, box.ibz);
                    const float dx = (float)ex, dy = (float)ey, dz = (float)ez;
                    float r2 = dx * dx + dy * dy + dz * dz;
                    float rinv = rsqrt_t(r2);
                    float rinv2 = rinv * rinv;
                    float rr3 = rinv * rinv2;
                    float rr5 = rr3 * rinv2;
                    float udj = rr5 * (uj.x * dx + uj.y * dy + uj.z * dz);
                    float udi = rr5 * (uf3.x * dx + uf3.y * dy + uf3.z * dz);
                    tx += udj * dx;
                    ty += udj * dy;
                    tz += udj * dz;
                    sx3 += rr3 * uj.x;
                    sy3 += rr3 * uj.y;
                    sz3 += rr3 * uj.z;
                    fx = rot1(fx) + (udi * dx - rr3 * uf.x);
                    fy = rot1(fy) + (udi * dy - rr3 * uf.y);
                    fz = rot1(fz) + (udi * dz - rr3 * uf.z);
                }
            }
            ax += 3.0 * (double)tx - (double)sx3;
            ay += 3.0 * (double)ty - (double)sy3;
            az += 3.0 * (double)tz - (double)sz3;
            jx = (double)fx;
            jy = (double)fy;
            jz = (double)fz;
            kend = kTile - 1;   // the lagging sums end on slot (lane + 63)
        } else if (!diag && (fr_n = frame_of(tpr, box)).smask >= 0) {
            // near tile pair, or a far one with a partial tile or excluded pairs: padding slots need no
            // masks in FP32 (their staged offsets lie 1e6 nm away) and are masked in the FP64 pass, as
            // are excluded pairs, whose steps all go to the FP64 pass
            const double4 pj = posa[jb + lane];
            const double iaj = inva[jb + lane];
            s_p[lane] = pj;
            s_u[lane] = mu[jb + lane];
            s_ia[lane] = iaj;
            {
                float4 l = posl[jb + lane];
                float4 u = mu32[jb + lane];
                l.w = (float)iaj;   // the Thole parameters of the FP32 damped steps ride in the unused w
                u.w = (float)pj.w;
                s_nl[lane] = l;
                s_nl[lane + kTile] = l;
                s_nu[lane] = u;
                s_nu[lane + kTile] = u;
            }
            __syncthreads();
            const float r2min = tpr.r2min;
            const float r2d = tpr.r2d;
            const float si_x = (float)((double)li.x - fr_n.shx), si_y = (float)((double)li.y - fr_n.shy),
                        si_z = (float)((double)li.z - fr_n.shz);
            unsigned long long steps64 = 0ull;
#define TDP_NEAR_CASE(M)                                                                                       \
    case M:                                                                                                    \
        steps64 = masked ? near_matvec_mixed32<M, true>(s_nl, s_nu, lane, si_x, si_y, si_z, fr_n, uf, uf3, r2min,   \
                                                        r2d, nai, mi, ax, ay, az, jx, jy, jz)                   \
                         : near_matvec_mixed32<M, false>(s_nl, s_nu, lane, si_x, si_y, si_z, fr_n, uf, uf3, r2min,  \
                                                         r2d, nai, 0ull, ax, ay, az, jx, jy, jz);               \
        break;
            switch (fr_n.smask) {
                TDP_NEAR_CASE(0)
                TDP_NEAR_CASE(1)
                TDP_NEAR_CASE(2)
                TDP_NEAR_CASE(3)
                TDP_NEAR_CASE(4)
                TDP_NEAR_CASE(5)
                TDP_NEAR_CASE(6)
                TDP_NEAR_CASE(7)
            }
#undef TDP_NEAR_CASE
            if (steps64)
                near_matvec_mixed64<PBC>(steps64, s_p, s_u, s_ia, s_jn, lane, ivalid, cntJ, mi, pi, ui, iai, box, ax, ay,
                                         az, jx, jy, jz);
            kend = kTile;
        } else {
            s_p[lane] = posa[jb + lane];
            s_u[lane] = mu[jb + lane];
            s_ia[lane] = inva[jb + lane];
            __syncthreads();
            jx = jy = jz = 0.0;
            const int k0 = diag ? 1 : 0, k1 = diag ? diag_last_step(cntI) : kTile - 1;
            for (int k = k0; k <= k1; k++) {
                const int jj = (lane + k) & (kTile - 1);
                const bool ok = ivalid && (jj < cntJ) && !((mi >> jj) & 1ull) &&
                                !(diag && k == kTile / 2 && lane >= kTile / 2);
                const double4 pj = s_p[jj];
                const double4 uj = s_u[jj];
                double dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
                if (PBC) min_image(dx, dy, dz, box.bx, box.by, box.bz, box.ibx, box.iby, box.ibz);
                double r2 = ok ? dx * dx + dy * dy + dz * dz : 1.0;
                double rr3, rr5;
                damped_rr_wave(r2, pi.w, pj.w, iai, s_ia[jj], rr3, rr5);
                if (!ok) {
                    rr3 = 0.0;
                    rr5 = 0.0;
                }
                double udj = rr5 * (uj.x * dx + uj.y * dy + uj.z * dz);
                double udi = rr5 * (ui.x * dx + ui.y * dy + ui.z * dz);
                ax += udj * dx - rr3 * uj.x;
                ay += udj * dy - rr3 * uj.y;
                az += udj * dz - rr3 * uj.z;
                jx += udi * dx - rr3 * ui.x;
                jy += udi * dy - rr3 * ui.y;
                jz += udi * dz - rr3 * ui.z;
                jx = rot1(jx);
                jy = rot1(jy);
                jz = rot1(jz);
            }
            kend = k1 + 1;
        }
        // the j sums of a tile pair go to the FP32 slab: exact for far tile pairs (FP32 sums), one
        // rounding of the FP64 sums of the other tile pairs
        const size_t slot = slab_tile(tm, I, J) * kTile + ((lane + kend) & (kTile - 1));
        Pjf[slot] = (float)jx;
        Pjf[S + slot] = (float)jy;
        Pjf[2 * S + slot] = (float)jz;
    }
    // item partials in FP32 behind the slab: component c of item w at Pjf[3 S + c NW + 64 w + lane]
    const size_t NW = (size_t)gridDim.x * kTile, iw = 3 * S + (size_t)w * kTile + lane;
    Pjf[iw] = (float)ax;
    Pjf[iw + NW] = (float)ay;
    Pjf[iw + 2 * NW] = (float)az;
}


// ------------------------------------------------------------------ symmetric permanent field (a6)
// calculateFixedDipoleField (ReferenceTholeDipoleForce.cpp:715-733) with each pair once, in the
// sweep of k_sym_matvec: Delta = r_j - r_i,
//   E_i -= m [Delta (rr3 q_j - rr5 d_j.Delta) + rr3 d_j],  E_j += m [Delta (rr3 q_i + rr5 d_i.Delta) - rr3 d_i].
template <bool PBC, bool PERM>
__global__ __launch_bounds__(64) void k_sym_field(TileMap tm, int nblk, const int4* __restrict__ work,
                                                  const double4* __restrict__ posa, const double4* __restrict__ qd,
                                                  const double* __restrict__ inva, const int* __restrict__ exc_start,
                                                  const int* __restrict__ exc_J,
                                                  const unsigned long long* __restrict__ exc_mask,
                                                  const float4* __restrict__ tb_c, const float4* __restrict__ tb_h, const TilePair* __restrict__ tp,
                                                  BoxT64 box, double4* __restrict__ Pi, double* __restrict__ Pj,
                                                  size_t S) {
    __shared__ double4 s_p[kTile];
    __shared__ double4 s_q[kTile];
    __shared__ double s_ia[kTile];
    const int lane = threadIdx.x, w = blockIdx.x;
    const int4 wk = work[w];
    const int I = wk.x, J0 = wk.y, J1 = wk.z;
    const int i = I * kTile + lane;
    const int cntI = tm.cnt[I];
    const bool ivalid = lane < cntI;
    const double4 pi = posa[i];
    const double4 qi = qd[i];
    const double iai = inva[i];
    const float4 cI = tb_c[I], hI = tb_h[I];
    const bool ifull = cntI == kTile;
    double ax = 0.0, ay = 0.0, az = 0.0;
    int e = exc_start[I];
    const int eEnd = exc_start[I + 1];
    while (e < eEnd && exc_J[e] < J0) e++;
    for (int J = J0; J < J1; J++) {
        const TilePair tpr = tp[slab_tile(tm, I, J)];
        const int jb = J * kTile;
        const int cntJ = tm.cnt[J];
        __syncthreads();
        s_p[lane] = posa[jb + lane];
        s_q[lane] = qd[jb + lane];
        s_ia[lane] = inva[jb + lane];
        __syncthreads();
        const bool diag = J == I;
        const bool masked = (e < eEnd && exc_J[e] == J);
        unsigned long long m0 = 0ull, mh = 0ull;
        if (masked) {
            m0 = exc_mask[((size_t)e * 3 + 0) * kTile + lane];
            mh = exc_mask[((size_t)e * 3 + 1) * kTile + lane];
            e++;
        }
        double jx = 0.0, jy = 0.0, jz = 0.0;
        int kend;
        if (!masked && !diag && ifull && cntJ == kTile &&
            (tpr.flags & kTpFar)) {
#pragma unroll 2
            for (int k = 0; k < kTile; k++) {
                const int jj = (lane + k) & (kTile - 1);
                const double4 pj = s_p[jj];
                const double4 qj = s_q[jj];
                double dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
                if (PBC) min_image(dx, dy, dz, box.bx, box.by, box.bz, box.ibx, box.iby, box.ibz);
                double r2 = dx * dx + dy * dy + dz * dz;
                double rinv = rsqrt(r2);
                double rinv2 = rinv * rinv;
                double rr3 = rinv * rinv2;
                if (PERM) {
                    double rr5 = 3.0 * rr3 * rinv2;
                    double fj = rr3 * qj.w - rr5 * (qj.x * dx + qj.y * dy + qj.z * dz);
                    double fi = rr3 * qi.w + rr5 * (qi.x * dx + qi.y * dy + qi.z * dz);
                    ax -= dx * fj + qj.x * rr3;
                    ay -= dy * fj + qj.y * rr3;
                    az -= dz * fj + qj.z * rr3;
                    jx += dx * fi - qi.x * rr3;
                    jy += dy * fi - qi.y * rr3;
                    jz += dz * fi - qi.z * rr3;
                } else {   // charges only
                    const double fj = rr3 * qj.w, fi = rr3 * qi.w;
                    ax -= dx * fj;
                    ay -= dy * fj;
                    az -= dz * fj;
                    jx += dx * fi;
                    jy += dy * fi;
                    jz += dz * fi;
                }
                jx = rot1(jx);
                jy = rot1(jy);
                jz = rot1(jz);
            }
            kend = kTile;
        } else {
            const int k0 = diag ? 1 : 0, k1 = diag ? diag_last_step(cntI) : kTile - 1;
            for (int k = k0; k <= k1; k++) {
                const int jj = (lane + k) & (kTile - 1);
                const bool ok = ivalid && (jj < cntJ) && !((m0 >> jj) & 1ull) &&
                                !(diag && k == kTile / 2 && lane >= kTile / 2);
                const double scale = ((mh >> jj) & 1ull) ? 0.5 : 1.0;
                const double4 pj = s_p[jj];
                const double4 qj = s_q[jj];
                double dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
                if (PBC) min_image(dx, dy, dz, box.bx, box.by, box.bz, box.ibx, box.iby, box.ibz);
                double r2 = ok ? dx * dx + dy * dy + dz * dz : 1.0;
                double rr3, rr5;
                damped_rr_wave(r2, pi.w, pj.w, iai, s_ia[jj], rr3, rr5);
                const double m = ok ? scale : 0.0;
                rr3 = ok ? rr3 : 0.0;
                rr5 = ok ? rr5 : 0.0;
                if (PERM) {
                    double fj = rr3 * qj.w - rr5 * (qj.x * dx + qj.y * dy + qj.z * dz);
                    double fi = rr3 * qi.w + rr5 * (qi.x * dx + qi.y * dy + qi.z * dz);
                    ax -= (dx * fj + qj.x * rr3) * m;
                    ay -= (dy * fj + qj.y * rr3) * m;
                    az -= (dz * fj + qj.z * rr3) * m;
                    jx += (dx * fi - qi.x * rr3) * m;
                    jy += (dy * fi - qi.y * rr3) * m;
                    jz += (dz * fi - qi.z * rr3) * m;
                } else {   // charges only
                    const double fj = rr3 * qj.w, fi = rr3 * qi.w;
                    ax -= (dx * fj) * m;
                    ay -= (dy * fj) * m;
                    az -= (dz * fj) * m;
                    jx += (dx * fi) * m;
                    jy += (dy * fi) * m;
                    jz += (dz * fi) * m;
                }
                jx = rot1(jx);
                jy = rot1(jy);
                jz = rot1(jz);
            }
            kend = k1 + 1;
        }
        store3(Pj, S, slab_tile(tm, I, J) * kTile + ((lane + kend) & (kTile - 1)), jx, jy, jz);
    }
    Pi[(size_t)w * kTile + lane] = make_double4(ax, ay, az, 0.0);
}

// ------------------------------------------------------------------ symmetric energy/force/torque (a9)
// calculateElectrostatic (ReferenceTholeDipoleForce.cpp:368-393) with each pair once.  The pair is
// evaluated in the reference's orientation (I = lower caller index, orig[]); the force on I is -f,
// on K +f; torques tI, tK; the energy is added by the lane.
struct ElecAcc {
    double fx, fy, fz, tx, ty, tz;
};

template <bool MUTUAL, bool DAMP, bool PERM>
__device__ __forceinline__ void elec_sym_step(const ElecSite& me, int my_orig, const ElecSite& other, int other_orig,
                                              bool ok, double mS, double iS, bool pbc, const BoxT64& box, ElecAcc& a,
                                              ElecAcc& b, double& en) {
    const bool lo = my_orig < other_orig;
    const ElecSite& A = lo ? me : other;
    const ElecSite& B = lo ? other : me;
    double dx = B.x - A.x, dy = B.y - A.y, dz = B.z - A.z;
    if (pbc) min_image(dx, dy, dz, box.bx, box.by, box.bz, box.ibx, box.iby, box.ibz);
    if (!ok) { dx = 1.0; dy = 0.0; dz = 0.0; mS = 0.0; iS = 0.0; }
    ElecOut o;
    elec_pair<MUTUAL, DAMP, PERM>(A, B, dx, dy, dz, mS, iS, o);
    if (!ok) { o.e = o.fx = o.fy = o.fz = o.tix = o.tiy = o.tiz = o.tkx = o.tky = o.tkz = 0.0; }
    const double s = lo ? -1.0 : 1.0;   // force on this lane's particle = s f
    a.fx += s * o.fx; a.fy += s * o.fy; a.fz += s * o.fz;
    b.fx -= s * o.fx; b.fy -= s * o.fy; b.fz -= s * o.fz;
    if (PERM) {
        a.tx += lo ? o.tix : o.tkx; a.ty += lo ? o.tiy : o.tky; a.tz += lo ? o.tiz : o.tkz;
        b.tx += lo ? o.tkx : o.tix; b.ty += lo ? o.tky : o.tiy; b.tz += lo ? o.tkz : o.tiz;
    }
    en += o.e;
}

// Rotation of the j-side sums of the energy pass by one lane (torques only with permanent dipoles).
template <bool PERM>
__device__ __forceinline__ void rot_acc(ElecAcc& b) {
    b.fx = rot1(b.fx); b.fy = rot1(b.fy); b.fz = rot1(b.fz);
    if (PERM) { b.tx = rot1(b.tx); b.ty = rot1(b.ty); b.tz = rot1(b.tz); }
}

// FP32 far tile pair of the energy pass: tile frame (offsets s_l) when fr.smask >= 0, else FP64
// separations from s_p.  Pair orientation, the frame's axis choice (raw y and z compared in FP64)
// and the force/torque signs follow elec_sym_step.  Returns false when a straddling pair could not
// be decided in FP32 (the caller then repeats the tile pair with FP64 separations).
template <bool PBC, bool MUTUAL, bool PERM, bool FRAME>
__device__ __forceinline__ bool elec_far32(const double4* __restrict__ s_p, const float4* __restrict__ s_l,
                                           const float4* __restrict__ s_qf, const float4* __restrict__ s_uf,
                                           const int* __restrict__ s_o, int lane, const ElecSite& me, int my_orig,
                                           float si_x, float si_y, float si_z, const FarFrame& fr, const BoxT64& box,
                                           ElecAcc& acc, ElecAcc& jacc, double& en) {
    const ElecSiteT<float> m32 = {(float)me.q, (float)me.dx, (float)me.dy, (float)me.dz, (float)me.ux,
                                  (float)me.uy, (float)me.uz, 0.f, 0.f};
    float ifx = 0.f, ify = 0.f, ifz = 0.f, itx = 0.f, ity = 0.f, itz = 0.f, e32 = 0.f;
    float jfx = 0.f, jfy = 0.f, jfz = 0.f, jtx = 0.f, jty = 0.f, jtz = 0.f;
    bool amb = false;
    for (int k = 0; k < kTile; k++) {
        const int jj = (lane + k) & (kTile - 1);
        const double4 pj = s_p[jj];
        const bool lo = my_orig < s_o[jj];
        float dx, dy, dz;   // x_j - x_i, minimum image
        if (FRAME) {
            const float4 lj = s_l[jj];
            dx = (fr.smask & 1) ? far_sep<true>(lj.x, si_x, fr.hx, fr.bx, amb) : lj.x - si_x;
            dy = (fr.smask & 2) ? far_sep<true>(lj.y, si_y, fr.hy, fr.by, amb) : lj.y - si_y;
            dz = (fr.smask & 4) ? far_sep<true>(lj.z, si_z, fr.hz, fr.bz, amb) : lj.z - si_z;
        } else {
            double ex = pj.x - me.x, ey = pj.y - me.y, ez = pj.z - me.z;
            if (PBC) min_image(ex, ey, ez, box.bx, box.by, box.bz, box.ibx, box.iby, box.ibz);
            dx = (float)ex;
            dy = (float)ey;
            dz = (float)ez;
        }
        const float4 qj = s_qf[jj], uj = s_uf[jj];
        const ElecSiteT<float> o32 = {qj.w, qj.x, qj.y, qj.z, uj.x, uj.y, uj.z, 0.f, 0.f};
        const float sg = lo ? -1.f : 1.f;   // force on this lane's particle = sg f; separation A -> B = -sg d
        ElecOutT<float> o;
        elec_pair_t<float, MUTUAL, false, PERM>(lo ? m32 : o32, lo ? o32 : m32, -sg * dx, -sg * dy, -sg * dz,
                                                me.y != pj.y || me.z != pj.z, 1.f, 1.f, o);
        ifx += sg * o.fx; ify += sg * o.fy; ifz += sg * o.fz;
        e32 += o.e;
        // j side lagging one step (rotation fused into the add)
        jfx = rot1(jfx) - sg * o.fx;
        jfy = rot1(jfy) - sg * o.fy;
        jfz = rot1(jfz) - sg * o.fz;
        if (PERM) {
            itx += lo ? o.tix : o.tkx; ity += lo ? o.tiy : o.tky; itz += lo ? o.tiz : o.tkz;
            jtx = rot1(jtx) + (lo ? o.tkx : o.tix);
            jty = rot1(jty) + (lo ? o.tky : o.tiy);
            jtz = rot1(jtz) + (lo ? o.tkz : o.tiz);
        }
    }
    if (FRAME && __builtin_amdgcn_ballot_w64(amb)) return false;
    acc.fx += ifx; acc.fy += ify; acc.fz += ifz;
    acc.tx += itx; acc.ty += ity; acc.tz += itz;
    en += e32;
    jacc.fx = jfx; jacc.fy = jfy; jacc.fz = jfz;
    jacc.tx = jtx; jacc.ty = jty; jacc.tz = jtz;
    return true;
}

// FP32 far tile pair of the energy pass without permanent dipoles (charges and induced dipoles, no
// damping, unit scales) in the lab frame.  For such a pair calculateElectrostaticPairIxn
// (ReferenceTholeDipoleForce.cpp:129-366) reduces to a closed form in the frame's X axis alone:
// with d = x_B - x_A (A the lower caller index), Z = d/r and X = (e_x - Z_x Z)/|e_x - Z_x Z| (the
// frame for raw positions that differ in y or z), a = mu_A, b = mu_B, w = q_A b - q_B a,
//   energy = q_A q_B / r - (X.w) / (2 r^2),
//   F      = w / r^3 + X (q_A q_B / r^2 - 3 (X.w) / r^3)
//            [+ 6 ((X.b) a + (X.a) b) / r^4 + X (6 a.b - 30 (X.a)(X.b)) / r^4 for Mutual],
// and the force on B is (F_y, F_z, F_x): the reference rotates the force back with the index
// pattern at :150-159, which permutes the components (Y(Y.v) + Z(Z.v) = v - X(X.v) removes Y and
// Z).  X = (rho^2, -d_x d_y, -d_x d_z) / (r rho), rho^2 = d_y^2 + d_z^2, is formed from products
// only, with no cancellation when d is nearly along x.  The lane's own site is m, the other o, s = +1 when the
// lane's particle is A; d, X and (X.m)(X.o) do not depend on the orientation, w and the a/b terms
// change sign with it.  Returns false, with nothing accumulated, when a straddling pair cannot be
// decided in FP32 (SM >= 0), when some pair's raw y and z may be equal (keys match; the reference
// then builds the frame from e_y) or when a separation lies along x (rho = 0: no frame); the
// caller then uses elec_far32.  SM: straddling-axis mask of the tile frame; -1: FP64 separations.
template <bool PBC, bool MUTUAL, int SM>
__device__ __forceinline__ bool elec_far32_lab(const double4* __restrict__ s_p, const float4* __restrict__ s_l,
                                               const float4* __restrict__ s_uf, const int* __restrict__ s_o,
                                               int lane, const ElecSite& me, int my_orig, unsigned key,
                                               float si_x, float si_y, float si_z, const FarFrame& fr,
                                               const BoxT64& box, ElecAcc& acc, ElecAcc& jacc, double& en) {
    const float Qm = (float)me.q, mx = (float)me.ux, my = (float)me.uy, mz = (float)me.uz;
    float ifx = 0.f, ify = 0.f, ifz = 0.f, e32 = 0.f, jfx = 0.f, jfy = 0.f, jfz = 0.f;
    bool amb = false, bad = false;
    for (int k = 0; k < kTile; k++) {
        const int jj = (lane + k) & (kTile - 1);
        float dx, dy, dz;   // x_other - x_me, minimum image
        if (SM >= 0) {
            const float4 lj = s_l[jj];
            dx = far_sep<(SM & 1) != 0>(lj.x, si_x, fr.hx, fr.bx, amb);
            dy = far_sep<(SM & 2) != 0>(lj.y, si_y, fr.hy, fr.by, amb);
            dz = far_sep<(SM & 4) != 0>(lj.z, si_z, fr.hz, fr.bz, amb);
            bad = bad || __float_as_uint(lj.w) == key;
        } else {
            const double4 pj = s_p[jj];
            double ex = pj.x - me.x, ey = pj.y - me.y, ez = pj.z - me.z;
            if (PBC) min_image(ex, ey, ez, box.bx, box.by, box.bz, box.ibx, box.iby, box.ibz);
            dx = (float)ex;
            dy = (float)ey;
            dz = (float)ez;
            bad = bad || !(me.y != pj.y || me.z != pj.z);
        }
        const float4 oq = s_uf[jj];   // induced dipole, charge
        const float s = my_orig < s_o[jj] ? 1.f : -1.f;
        const float rho2 = dy * dy + dz * dz;
        bad = bad || !(rho2 > 0.f);
        const float r2 = dx * dx + rho2;
        const float rinv = rsqrt_t(r2);
        const float rI2 = rinv * rinv, rI3 = rI2 * rinv;
        const float kx = rinv * rsqrt_t(rho2);
        // X = (rho^2, -d_x d_y, -d_x d_z) / (r rho): products only, no cancellation; X.v as a dot product
        const float ex = -dx * kx;
        const float Xx = rho2 * kx, Xy = ex * dy, Xz = ex * dz;
        const float Xm = Xx * mx + Xy * my + Xz * mz;
        const float Xo = Xx * oq.x + Xy * oq.y + Xz * oq.z;
        const float QQ = Qm * oq.w;
        const float Xw = Qm * Xo - oq.w * Xm;
        float am, ao, inner;
        if (MUTUAL) {
            const float rI4 = rI2 * rI2;
            const float A = s * (6.f * rI4);
            am = A * Xo - rI3 * oq.w;
            ao = A * Xm + rI3 * Qm;
            inner = rI2 * QQ + rI4 * (6.f * (mx * oq.x + my * oq.y + mz * oq.z) - 30.f * Xm * Xo);
        } else {
            am = -rI3 * oq.w;
            ao = rI3 * Qm;
            inner = rI2 * QQ;
        }
        const float lam = s * inner - 3.f * rI3 * Xw;
        const float Gx = mx * am + oq.x * ao + Xx * lam;
        const float Gy = my * am + oq.y * ao + Xy * lam;
        const float Gz = mz * am + oq.z * ao + Xz * lam;
        e32 += rinv * QQ - 0.5f * s * rI2 * Xw;
        // force on the lane's particle -(G_y, G_z, G_x), on the other +(G_y, G_z, G_x); j side lagging one step
        ifx -= Gy;
        ify -= Gz;
        ifz -= Gx;
        jfx = rot1(jfx) + Gy;
        jfy = rot1(jfy) + Gz;
        jfz = rot1(jfz) + Gx;
    }
    if (__builtin_amdgcn_ballot_w64(amb || bad)) return false;
    acc.fx += ifx; acc.fy += ify; acc.fz += ifz;
    en += e32;
    jacc.fx = jfx; jacc.fy = jfy; jacc.fz = jfz;
    jacc.tx = jacc.ty = jacc.tz = 0.0;
    return true;
}

// FP64 far tile pair of the energy pass without permanent dipoles: the closed form of elec_far32_lab
// in FP64, on separations by the minimum image of each pair.  Returns false, with nothing
// accumulated, when some pair's raw y and z are equal (the reference then builds the frame from
// e_y) or a separation lies along x (rho = 0); the caller then evaluates the explicit frame.
template <bool PBC, bool MUTUAL>
__device__ __forceinline__ bool elec_far64_lab(const double4* __restrict__ s_p, const double4* __restrict__ s_q,
                                               const double4* __restrict__ s_u, const int* __restrict__ s_o, int lane,
                                               const ElecSite& me, int my_orig, const BoxT64& box, ElecAcc& acc,
                                               ElecAcc& jacc, double& en) {
    const double Qm = me.q, mx = me.ux, my = me.uy, mz = me.uz;
    double ifx = 0.0, ify = 0.0, ifz = 0.0, e64 = 0.0, jfx = 0.0, jfy = 0.0, jfz = 0.0;
    bool bad = false;
    for (int k = 0; k < kTile; k++) {
        const int jj = (lane + k) & (kTile - 1);
        const double4 pj = s_p[jj];
        double dx = pj.x - me.x, dy = pj.y - me.y, dz = pj.z - me.z;   // x_other - x_me, minimum image
        bad = bad || !(me.y != pj.y || me.z != pj.z);
        if (PBC) min_image(dx, dy, dz, box.bx, box.by, box.bz, box.ibx, box.iby, box.ibz);
        const double4 uo = s_u[jj];
        const double qo = s_q[jj].w;
        const double s = my_orig < s_o[jj] ? 1.0 : -1.0;
        const double rho2 = dy * dy + dz * dz;
        bad = bad || !(rho2 > 0.0);
        const double r2 = dx * dx + rho2;
        const double rinv = rsqrt(r2);
        const double rI2 = rinv * rinv, rI3 = rI2 * rinv;
        const double kx = rinv * rsqrt(rho2 > 0.0 ? rho2 : 1.0);
        const double ex = -dx * kx;
        const double Xx = rho2 * kx, Xy = ex * dy, Xz = ex * dz;
        const double Xm = Xx * mx + Xy * my + Xz * mz;
        const double Xo = Xx * uo.x + Xy * uo.y + Xz * uo.z;
        const double QQ = Qm * qo;
        const double Xw = Qm * Xo - qo * Xm;
        double am, ao, inner;
        if (MUTUAL) {
            const double rI4 = rI2 * rI2;
            const double A = s * (6.0 * rI4);
            am = A * Xo - rI3 * qo;
            ao = A * Xm + rI3 * Qm;
            inner = rI2 * QQ + rI4 * (6.0 * (mx * uo.x + my * uo.y + mz * uo.z) - 30.0 * Xm * Xo);
        } else {
            am = -rI3 * qo;
            ao = rI3 * Qm;
            inner = rI2 * QQ;
        }
        const double lam = s * inner - 3.0 * rI3 * Xw;
        const double Gx = mx * am + uo.x * ao + Xx * lam;
        const double Gy = my * am + uo.y * ao + Xy * lam;
        const double Gz = mz * am + uo.z * ao + Xz * lam;
        e64 += rinv * QQ - 0.5 * s * rI2 * Xw;
        // force on the lane's particle -(G_y, G_z, G_x), on the other +(G_y, G_z, G_x); j side lagging one step
        ifx -= Gy;
        ify -= Gz;
        ifz -= Gx;
        jfx = rot1(jfx) + Gy;
        jfy = rot1(jfy) + Gz;
        jfz = rot1(jfz) + Gx;
    }
    if (__builtin_amdgcn_ballot_w64(bad)) return false;
    acc.fx += ifx;
    acc.fy += ify;
    acc.fz += ifz;
    en += e64;
    jacc.fx = jfx;
    jacc.fy = jfy;
    jacc.fz = jfz;
    jacc.tx = jacc.ty = jacc.tz = 0.0;
    return true;
}

// Near tile pair of the energy pass without permanent dipoles in mixed precision (off-diagonal
// tiles on frame SM, full or partial), pass 1: steps whose pairs all lie beyond r2d (Thole exponent
// a u^3 >= 4) in FP32 by the closed form of elec_far32_lab with the Thole factors of
// calculateElectrostaticPairIxn (:205-220): with e = exp(-a u^3), thole_c = 1 - e, thole_d0 =
// 1 - e (1 + 1.5 a u^3), dthole_d0 = 1 - e (1 + a u^3 + 1.5 (a u^3)^2), dthole_d1 = 1 - e (1 + a u^3),
//   energy = q_A q_B / r - thole_c (X.w) / (2 r^2),
//   F      = thole_c w / r^3 + X (q_A q_B / r^2 - (2 thole_d0 + thole_c) (X.w) / r^3)
//            [+ c6 ((X.b) a + (X.a) b) / r^4 + X (6 dthole_d1 a.b - (12 dthole_d0 + 6 dthole_d1 + 2 c6)
//               (X.a)(X.b)) / r^4 for Mutual, c6 = 4 thole_d0 + 2 thole_c].
// Steps with a closer pair, an undecidable straddling pair, possibly equal raw y and z or a
// separation along x are returned in a step mask for the FP64 pass (their j sums only rotate).
// MASKED: the tile pair holds scaled pairs (bit s of mscaled: the pair with slot s has mScale or
// iScale below 1); a step with such a pair goes to the FP64 pass, which applies the scales.
template <bool PBC, bool MUTUAL, int SM, bool MASKED>
__device__ __forceinline__ unsigned long long elec_near32_lab(const float4* __restrict__ s_l,
                                                              const float4* __restrict__ s_uf,
                                                              const float2* __restrict__ s_na,
                                                              const int* __restrict__ s_o, int lane,
                                                              const ElecSite& me, int my_orig, unsigned key,
                                                              float si_x, float si_y, float si_z, const FarFrame& fr,
                                                              float r2d, unsigned long long mscaled, ElecAcc& acc,
                                                              ElecAcc& jacc, double& en) {
    const float Qm = (float)me.q, mx = (float)me.ux, my = (float)me.uy, mz = (float)me.uz;
    const float am_ = (float)me.a, iam = (float)me.ia;
    float ifx = 0.f, ify = 0.f, ifz = 0.f, e32 = 0.f, jfx = 0.f, jfy = 0.f, jfz = 0.f;
    unsigned long long steps64 = 0ull;
    for (int k = 0; k < kTile; k++) {
        const int jj = (lane + k) & (kTile - 1);
        const float4 lj = s_l[jj];
        bool amb = false;
        const float dx = far_sep<(SM & 1) != 0>(lj.x, si_x, fr.hx, fr.bx, amb);
        const float dy = far_sep<(SM & 2) != 0>(lj.y, si_y, fr.hy, fr.by, amb);
        const float dz = far_sep<(SM & 4) != 0>(lj.z, si_z, fr.hz, fr.bz, amb);
        const float rho2 = dy * dy + dz * dz;
        const float r2 = dx * dx + rho2;
        if (MASKED) amb = amb || ((mscaled >> jj) & 1ull);
        if (__builtin_amdgcn_ballot_w64(amb || __float_as_uint(lj.w) == key || !(rho2 > 0.f) || !(r2 > r2d))) {
            steps64 |= 1ull << k;
            jfx = rot1(jfx);
            jfy = rot1(jfy);
            jfz = rot1(jfz);
            continue;
        }
        const float4 oq = s_uf[jj];
        const float2 na = s_na[jj];
        const float s = my_orig < s_o[jj] ? 1.f : -1.f;
        const float rinv = rsqrt_t(r2);
        const float rI2 = rinv * rinv, rI3 = rI2 * rinv;
        const float uu = r2 * rinv * (iam * na.y);
        const float au3 = fminf(fminf(am_, na.x) * (uu * uu * uu), 100.f);
        const float ev = __expf(-au3);
        const float thc = 1.f - ev;
        const float th0 = 1.f - ev * fmaf(1.5f, au3, 1.f);
        const float kx = rinv * rsqrt_t(rho2);
        // X = (rho^2, -d_x d_y, -d_x d_z) / (r rho): products only, no cancellation; X.v as a dot product
        const float ex = -dx * kx;
        const float Xx = rho2 * kx, Xy = ex * dy, Xz = ex * dz;
        const float Xm = Xx * mx + Xy * my + Xz * mz;
        const float Xo = Xx * oq.x + Xy * oq.y + Xz * oq.z;
        const float QQ = Qm * oq.w;
        const float Xw = Qm * Xo - oq.w * Xm;
        const float t3 = rI3 * thc;
        float amc, aoc, inner;
        if (MUTUAL) {
            const float rI4 = rI2 * rI2;
            const float dd1 = 1.f - ev * (1.f + au3);
            const float dd0 = dd1 - ev * (1.5f * au3 * au3);
            const float c6 = 4.f * th0 + 2.f * thc;
            const float A = s * c6 * rI4;
            amc = A * Xo - t3 * oq.w;
            aoc = A * Xm + t3 * Qm;
            inner = rI2 * QQ + rI4 * (6.f * dd1 * (mx * oq.x + my * oq.y + mz * oq.z) -
                                      (12.f * dd0 + 6.f * dd1 + 2.f * c6) * Xm * Xo);
        } else {
            amc = -t3 * oq.w;
            aoc = t3 * Qm;
            inner = rI2 * QQ;
        }
        const float lam = s * inner - rI3 * (2.f * th0 + thc) * Xw;
        const float Gx = mx * amc + oq.x * aoc + Xx * lam;
        const float Gy = my * amc + oq.y * aoc + Xy * lam;
        const float Gz = mz * amc + oq.z * aoc + Xz * lam;
        e32 += rinv * QQ - 0.5f * s * rI2 * thc * Xw;
        ifx -= Gy;
        ify -= Gz;
        ifz -= Gx;
        jfx = rot1(jfx) + Gy;
        jfy = rot1(jfy) + Gz;
        jfz = rot1(jfz) + Gx;
    }
    acc.fx += ifx; acc.fy += ify; acc.fz += ifz;
    en += e32;
    // the j sums lag one step: after 64 steps lane l holds slot l + 63, one more rotation gives slot l
    jacc.fx = (double)rot1(jfx);
    jacc.fy = (double)rot1(jfy);
    jacc.fz = (double)rot1(jfz);
    jacc.tx = jacc.ty = jacc.tz = 0.0;
    return steps64;
}

// MIXED: tile pairs beyond the Thole range evaluate the pair terms in FP32 (separations from the tile
// frames, or FP64 separations where a frame cannot decide an image; the frame's axis choice compares
// FP64 positions; sums per tile flushed to FP64); every other tile stays FP64.  PERM = false: the
// system has no permanent dipoles, so no torques are formed or written.
template <bool PBC, bool MUTUAL, bool MIXED, bool PERM>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(PERM ? 2 : 3))) void k_sym_elec(TileMap tm, int nblk, const int4* __restrict__ work,
                                                 const double4* __restrict__ posa, const double4* __restrict__ qd,
                                                 const double4* __restrict__ mu, const double* __restrict__ inva,
                                                 const int* __restrict__ orig, const int* __restrict__ exc_start,
                                                 const int* __restrict__ exc_J,
                                                 const unsigned long long* __restrict__ exc_mask,
                                                 const float4* __restrict__ tb_c, const float4* __restrict__ tb_h, const TilePair* __restrict__ tp,
                                                 const double4* __restrict__ tb_o, const float4* __restrict__ posl,
                                                 BoxT64 box, double4* __restrict__ PiF, double* __restrict__ PjF,
                                                 double4* __restrict__ PiT, double* __restrict__ PjT, size_t S,
                                                 double* __restrict__ part_e, float* __restrict__ Pf32) {
    // MIXED without permanent dipoles: the force partials go to Pf32 in FP32, in the layout of the
    // mixed T*mu pass (tile slab, then the item partials), one rounding per tile pair or work item
    constexpr bool kF32Out = MIXED && !PERM;
    __shared__ double4 s_p[kTile];
    __shared__ double4 s_q[kTile];
    __shared__ double4 s_u[kTile];
    __shared__ double s_ia[kTile];
    __shared__ int s_o[kTile];
    __shared__ float4 s_l[MIXED ? kTile : 1];
    __shared__ float4 s_qf[MIXED ? kTile : 1];
    __shared__ float4 s_uf[MIXED ? kTile : 1];
    __shared__ float2 s_na[MIXED && !PERM ? kTile : 1];      // (a, 1/a) of the j tile, near mixed steps
    __shared__ double s_jn[MIXED && !PERM ? 3 * kTile : 1];  // per-slot j sums of the FP64 steps
    const int lane = threadIdx.x, w = blockIdx.x;
    const int4 wk = work[w];
    const int I = wk.x, J0 = wk.y, J1 = wk.z;
    const int i = I * kTile + lane;
    const int cntI = tm.cnt[I];
    const bool ivalid = lane < cntI;
    ElecSite me;
    load_site(me, posa[i], qd[i], mu[i], inva[i]);
    const int my_orig = orig[i];
    const float4 cI = tb_c[I], hI = tb_h[I];
    const double4 oI = tb_o[I];
    const float4 li = posl[i];
    const bool ifull = cntI == kTile;
    ElecAcc acc = {0, 0, 0, 0, 0, 0};
    double en = 0.0;
    int e = exc_start[I];
    const int eEnd = exc_start[I + 1];
    while (e < eEnd && exc_J[e] < J0) e++;
    for (int J = J0; J < J1; J++) {
        const TilePair tpr = tp[slab_tile(tm, I, J)];
        const int jb = J * kTile;
        const int cntJ = tm.cnt[J];
        __syncthreads();
        s_p[lane] = posa[jb + lane];
        s_q[lane] = qd[jb + lane];
        s_u[lane] = mu[jb + lane];
        s_ia[lane] = inva[jb + lane];
        s_o[lane] = orig[jb + lane];
        __syncthreads();
        const bool diag = J == I;
        const bool masked = (e < eEnd && exc_J[e] == J);
        unsigned long long m0 = 0ull, mh = 0ull, mi = 0ull;
        if (masked) {
            m0 = exc_mask[((size_t)e * 3 + 0) * kTile + lane];
            mh = exc_mask[((size_t)e * 3 + 1) * kTile + lane];
            mi = exc_mask[((size_t)e * 3 + 2) * kTile + lane];
            e++;
        }
        ElecAcc jacc = {0, 0, 0, 0, 0, 0};
        int kend;
        FarFrame fr_n;
        if (!masked && !diag && ifull && cntJ == kTile &&
            (tpr.flags & kTpFar)) {
            if (MIXED) {
                const FarFrame fr = frame_of(tpr, box);
                {
                    const double4 q = s_q[lane], u = s_u[lane];
                    s_qf[lane] = make_float4((float)q.x, (float)q.y, (float)q.z, (float)q.w);
                    s_uf[lane] = make_float4((float)u.x, (float)u.y, (float)u.z, (float)q.w);
                    s_l[lane] = posl[jb + lane];
                }
                __syncthreads();
                const float si_x = (float)((double)li.x - fr.shx), si_y = (float)((double)li.y - fr.shy),
                            si_z = (float)((double)li.z - fr.shz);
                bool done = false;
                if (!PERM) {   // closed form in the lab frame; elec_far32 where it cannot decide
                    const unsigned key = __float_as_uint(li.w);
#define TDP_LAB_CASE(M)                                                                                        \
    case M:                                                                                                    \
        done = elec_far32_lab<PBC, MUTUAL, M>(s_p, s_l, s_uf, s_o, lane, me, my_orig, key, si_x, si_y, si_z, fr, \
                                              box, acc, jacc, en);                                             \
        break;
                    switch (fr.smask) {
                        TDP_LAB_CASE(0)
                        TDP_LAB_CASE(1)
                        TDP_LAB_CASE(2)
                        TDP_LAB_CASE(3)
                        TDP_LAB_CASE(4)
                        TDP_LAB_CASE(5)
                        TDP_LAB_CASE(6)
                        TDP_LAB_CASE(7)
                        default:
                            break;
                    }
#undef TDP_LAB_CASE
                    if (!done)
                        done = elec_far32_lab<PBC, MUTUAL, -1>(s_p, s_l, s_uf, s_o, lane, me, my_orig, key, si_x, si_y,
                                                               si_z, fr, box, acc, jacc, en);
                } else {
                    done = fr.smask >= 0 &&
                           elec_far32<PBC, MUTUAL, PERM, true>(s_p, s_l, s_qf, s_uf, s_o, lane, me, my_orig, si_x,
                                                               si_y, si_z, fr, box, acc, jacc, en);
                }
                if (!done)
                    elec_far32<PBC, MUTUAL, PERM, false>(s_p, s_l, s_qf, s_uf, s_o, lane, me, my_orig, si_x, si_y,
                                                         si_z, fr, box, acc, jacc, en);
                kend = kTile - 1;   // the lagging sums end on slot (lane + 63)
            } else if (!PERM && elec_far64_lab<PBC, MUTUAL>(s_p, s_q, s_u, s_o, lane, me, my_orig, box, acc, jacc, en)) {
                kend = kTile - 1;   // closed form in the lab frame: the lagging sums end on slot (lane + 63)
            } else {
                for (int k = 0; k < kTile; k++) {
                    const int jj = (lane + k) & (kTile - 1);
                    ElecSite other;
                    load_site(other, s_p[jj], s_q[jj], s_u[jj], s_ia[jj]);
                    elec_sym_step<MUTUAL, false, PERM>(me, my_orig, other, s_o[jj], true, 1.0, 1.0, PBC, box, acc, jacc,
                                                       en);
                    rot_acc<PERM>(jacc);
                }
                kend = kTile;
            }
        } else if (MIXED && !PERM && !diag && (fr_n = frame_of(tpr, box)).smask >= 0) {
            // full or partial tiles: padding carries no charge or dipole and its staged offsets lie
            // 1e6 nm away (finite FP32 terms); the FP64 steps mask it.  Steps with scaled pairs go to
            // the FP64 steps, which apply the scales.
            // near tile pair, mixed: FP32 steps by the closed form, the rest in FP64 (pass 2)
            {
                const double4 q = s_q[lane], u = s_u[lane];
                s_uf[lane] = make_float4((float)u.x, (float)u.y, (float)u.z, (float)q.w);
                s_l[lane] = posl[jb + lane];
                s_na[lane] = make_float2((float)s_p[lane].w, (float)s_ia[lane]);
                s_jn[lane] = 0.0;
                s_jn[kTile + lane] = 0.0;
                s_jn[2 * kTile + lane] = 0.0;
            }
            __syncthreads();
            const float r2d = tpr.r2d;
            const float si_x = (float)((double)li.x - fr_n.shx), si_y = (float)((double)li.y - fr_n.shy),
                        si_z = (float)((double)li.z - fr_n.shz);
            const unsigned key = __float_as_uint(li.w);
            unsigned long long steps64 = 0ull;
#define TDP_ENEAR_CASE(M)                                                                                        \
    case M:                                                                                                      \
        steps64 = masked ? elec_near32_lab<PBC, MUTUAL, M, true>(s_l, s_uf, s_na, s_o, lane, me, my_orig, key, si_x, \
                                                                 si_y, si_z, fr_n, r2d, m0 | mh | mi, acc, jacc, en) \
                         : elec_near32_lab<PBC, MUTUAL, M, false>(s_l, s_uf, s_na, s_o, lane, me, my_orig, key,     \
                                                                  si_x, si_y, si_z, fr_n, r2d, 0ull, acc, jacc, en); \
        break;
            switch (fr_n.smask) {
                TDP_ENEAR_CASE(0)
                TDP_ENEAR_CASE(1)
                TDP_ENEAR_CASE(2)
                TDP_ENEAR_CASE(3)
                TDP_ENEAR_CASE(4)
                TDP_ENEAR_CASE(5)
                TDP_ENEAR_CASE(6)
                TDP_ENEAR_CASE(7)
            }
#undef TDP_ENEAR_CASE
            // pass 2: the marked steps in FP64; in a step every lane serves a different j slot, whose
            // sums collect in s_jn without conflicts
            while (steps64) {
                const int k = __builtin_ctzll(steps64);
                steps64 &= steps64 - 1ull;
                const int jd = (lane + k) & (kTile - 1);
                ElecSite other;
                load_site(other, s_p[jd], s_q[jd], s_u[jd], s_ia[jd]);
                ElecAcc jb1 = {0, 0, 0, 0, 0, 0};
                // the scales of the pair as in the FP64 path below (a pair with both scales 0 is skipped)
                const bool z0 = (m0 >> jd) & 1ull, zh = (mh >> jd) & 1ull, zi = (mi >> jd) & 1ull;
                elec_sym_step<MUTUAL, true, PERM>(me, my_orig, other, s_o[jd], ivalid && jd < cntJ && !(z0 && zi),
                                                  z0 ? 0.0 : (zh ? 0.5 : 1.0), zi ? 0.0 : 1.0, PBC, box, acc, jb1,
                                                  en);
                s_jn[jd] += jb1.fx;
                s_jn[kTile + jd] += jb1.fy;
                s_jn[2 * kTile + jd] += jb1.fz;
            }
            __syncthreads();
            jacc.fx += s_jn[lane];
            jacc.fy += s_jn[kTile + lane];
            jacc.fz += s_jn[2 * kTile + lane];
            kend = kTile;
        } else {
            const int k0 = diag ? 1 : 0, k1 = diag ? diag_last_step(cntI) : kTile - 1;
            for (int k = k0; k <= k1; k++) {
                const int jj = (lane + k) & (kTile - 1);
                const bool z0 = (m0 >> jj) & 1ull, h = (mh >> jj) & 1ull, zi = (mi >> jj) & 1ull;
                // a pair with both scales 0 contributes nothing (this also covers the self pair)
                const bool ok = ivalid && (jj < cntJ) && !(z0 && zi) && !(diag && k == kTile / 2 && lane >= kTile / 2);
                const double mS = z0 ? 0.0 : (h ? 0.5 : 1.0);
                const double iS = zi ? 0.0 : 1.0;
                ElecSite other;
                load_site(other, s_p[jj], s_q[jj], s_u[jj], s_ia[jj]);
                elec_sym_step<MUTUAL, true, PERM>(me, my_orig, other, s_o[jj], ok, mS, iS, PBC, box, acc, jacc, en);
                rot_acc<PERM>(jacc);
            }
            kend = k1 + 1;
        }
        const size_t slot = slab_tile(tm, I, J) * kTile + ((lane + kend) & (kTile - 1));
        if (kF32Out) {
            Pf32[slot] = (float)jacc.fx;
            Pf32[S + slot] = (float)jacc.fy;
            Pf32[2 * S + slot] = (float)jacc.fz;
        } else {
            store3(PjF, S, slot, jacc.fx, jacc.fy, jacc.fz);
        }
        if (PERM) store3(PjT, S, slot, jacc.tx, jacc.ty, jacc.tz);
    }
    if (kF32Out) {
        const size_t NW = (size_t)gridDim.x * kTile, iw = 3 * S + (size_t)w * kTile + lane;
        Pf32[iw] = (float)acc.fx;
        Pf32[iw + NW] = (float)acc.fy;
        Pf32[iw + 2 * NW] = (float)acc.fz;
    } else {
        PiF[(size_t)w * kTile + lane] = make_double4(acc.fx, acc.fy, acc.fz, 0.0);
    }
    if (PERM) PiT[(size_t)w * kTile + lane] = make_double4(acc.tx, acc.ty, acc.tz, 0.0);
    en = wave_sum(en);
    if (lane == 0) part_e[w] = en;
}

// ------------------------------------------------------------------ fixed-order reduction
// Sum of the symmetric partials of tile column B (one workgroup of kRedWaves wavefronts per
// column): wave v adds the i partials of work items row_start[B] + v, + v + kRedWaves, ... and
// the j partials of tiles (I', B) for I' = v, v + kRedWaves, ...; the waves' sums are then added
// in wave order.  The order is fixed, so the result is deterministic.  out4 (double4 per
// particle) or out3 (three doubles per particle) receives the sums.
constexpr int kRedWaves = 8;
// F32: the partials are FP32 (the slab and item partials of the mixed T*mu pass in Pjf), else FP64
// (Pj and Pi).
template <bool F32>
__global__ __launch_bounds__(kRedWaves * 64) void k_col_reduce(TileMap tm, const int* __restrict__ row_start,
                                                               const double4* __restrict__ Pi,
                                                               const double* __restrict__ Pj,
                                                               const float* __restrict__ Pjf, size_t S,
                                                               const double* __restrict__ scalars,
                                                               double4* __restrict__ out4, double* __restrict__ out3,
                                                               PcgQ pq) {
    const int B = blockIdx.x, l = threadIdx.x & (kTile - 1), v = threadIdx.x / kTile;
    if (scalars && solver_done(scalars + SC_COUNT * tm.seg[B])) return;
    __shared__ double red[3][kRedWaves][kTile];
    double sx = 0.0, sy = 0.0, sz = 0.0;
    if (F32) {   // FP32 item partials behind the slab (k_sym_matvec_mixed)
        const size_t NW = (size_t)row_start[gridDim.x] * kTile;   // all items (the grid is one block per tile)
        const float* P = Pjf + 3 * S;
        for (int w = row_start[B] + v; w < row_start[B + 1]; w += kRedWaves) {
            const size_t k = (size_t)w * kTile + l;
            sx += (double)P[k];
            sy += (double)P[NW + k];
            sz += (double)P[2 * NW + k];
        }
    } else {
        for (int w = row_start[B] + v; w < row_start[B + 1]; w += kRedWaves) {
            const double4 p = Pi[(size_t)w * kTile + l];
            sx += p.x;
            sy += p.y;
            sz += p.z;
        }
    }
    const size_t col = tm.col[B];
    const int first = tm.first[B];
#pragma unroll 4
    for (int Ip = first + v; Ip <= B; Ip += kRedWaves) {
        const size_t k = (col + (size_t)(Ip - first)) * kTile + l;
        if (F32) {
            sx += (double)Pjf[k];
            sy += (double)Pjf[S + k];
            sz += (double)Pjf[2 * S + k];
        } else {
            sx += Pj[k];
            sy += Pj[S + k];
            sz += Pj[2 * S + k];
        }
    }
    red[0][v][l] = sx;
    red[1][v][l] = sy;
    red[2][v][l] = sz;
    __syncthreads();
    const int x = B * kTile + l;
    if (v == 0) {
        for (int u = 1; u < kRedWaves; u++) {
            sx += red[0][u][l];
            sy += red[1][u][l];
            sz += red[2][u][l];
        }
        if (pq.q) {   // PCG: q = p/alpha - T p and the tile sum of p.q (wave 0 holds the whole tile)
            const double4 p = pq.p[x];
            const double ia = inv_or_zero(pq.alpha[x]);
            const double4 q = make_double4(p.x * ia - sx, p.y * ia - sy, p.z * ia - sz, 0.0);
            pq.q[x] = q;
            const double d = wave_sum(p.x * q.x + p.y * q.y + p.z * q.z);
            if (l == 0) pq.red[B] = d;
        } else if (out3) {
            out3[3 * x] = sx;
            out3[3 * x + 1] = sy;
            out3[3 * x + 2] = sz;
        } else {
            out4[x] = make_double4(sx, sy, sz, 0.0);
        }
    }
}

// ------------------------------------------------------------------ solver vector kernels
// The vector kernels run one 64-thread block per tile; each tile belongs to one system (segment)
// of the context, whose solver scalars sit at scalars + SC_COUNT * s.  Padding slots carry
// alpha = 0 and stay zero.  Norms are summed per tile (wave_sum, fixed order) into red[T] and
// red[ntile + T] (the PCG p.q sums into red[2 ntile + T]), then per system over its tiles in a
// fixed assignment and order, by one 64-thread block or by every block of the next kernel alike
// (segment_sum), so every result is deterministic.  A system whose solve stops adds 1 to done[0].

// E0 = reduced field; mu = alpha*E0 (calculateInducedDipoles, ReferenceTholeDipoleForce.cpp:743-746).
__global__ void k_field_init(int npad, const double4* __restrict__ part, const double* __restrict__ alpha,
                             double4* __restrict__ E0, double4* __restrict__ mu, float4* __restrict__ mu32) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npad) return;
    double4 f = part[i];
    f.w = 0.0;
    double a = alpha[i];
    double4 m = make_double4(f.x * a, f.y * a, f.z * a, 0.0);
    E0[i] = f;
    mu[i] = m;
    if (mu32) mu32[i] = make_float4((float)m.x, (float)m.y, (float)m.z, 0.f);
}

__device__ __forceinline__ void tile_partials(double a, double b, double* __restrict__ red, int ntile) {
    a = wave_sum(a);
    b = wave_sum(b);
    if (threadIdx.x == 0) {
        red[blockIdx.x] = a;
        red[ntile + blockIdx.x] = b;
    }
}

// Sums of the tile partials of system sg over its tiles [sg.x, sg.y), in a fixed order: every
// workgroup that forms them gets the same values.
__device__ __forceinline__ double segment_sum(const int4 sg, const double* __restrict__ red) {
    double a = 0.0;
    for (int T = sg.x + (int)threadIdx.x; T < sg.y; T += kTile) a += red[T];
    return wave_sum(a);
}

__device__ __forceinline__ void segment_sums(const int4 sg, const double* __restrict__ red, int ntile, double& a,
                                             double& b) {
    a = 0.0;
    b = 0.0;
    for (int T = sg.x + (int)threadIdx.x; T < sg.y; T += kTile) {
        a += red[T];
        b += red[ntile + T];
    }
    a = wave_sum(a);
    b = wave_sum(b);
}

__device__ __forceinline__ void mark_done(double* __restrict__ sc, int* __restrict__ done) {
    sc[SC_DONE] = 1.0;
    atomicAdd(done, 1);
}

// One Jacobi step of convergeInducedDipolesByDIIS (ReferenceTholeDipoleForce.cpp:835-846):
// mu_new = alpha*(E0 + T mu); partial sums of |mu_new - mu|^2.
__global__ __launch_bounds__(64) void k_jacobi_update(const int* __restrict__ tseg, const double4* __restrict__ part,
                                                      const double4* __restrict__ E0, const double* __restrict__ alpha,
                                                      double4* __restrict__ mu, float4* __restrict__ mu32,
                                                      const double* __restrict__ scalars, double* __restrict__ red,
                                                      int ntile) {
    if (solver_done(scalars + SC_COUNT * tseg[blockIdx.x])) return;
    const int i = blockIdx.x * kTile + threadIdx.x;
    double4 f = part[i];
    double4 e0 = E0[i];
    double a = alpha[i];
    double nx = (e0.x + f.x) * a, ny = (e0.y + f.y) * a, nz = (e0.z + f.z) * a;
    double4 o = mu[i];
    double dx = nx - o.x, dy = ny - o.y, dz = nz - o.z;
    mu[i] = make_double4(nx, ny, nz, 0.0);
    if (mu32) mu32[i] = make_float4((float)nx, (float)ny, (float)nz, 0.f);
    tile_partials(dx * dx + dy * dy + dz * dz, 0.0, red, ntile);
}

// Jacobi bookkeeping (ReferenceTholeDipoleForce.cpp:829-850): eps = sqrt(sum/n), iteration count,
// stop when iterations == max or !(eps > target).
__global__ __launch_bounds__(64) void k_jacobi_finalize(const int4* __restrict__ seg, const double* __restrict__ red,
                                                        int ntile, double* __restrict__ scalars, int* __restrict__ done) {
    double* sc = scalars + SC_COUNT * blockIdx.x;
    if (solver_done(sc)) return;
    const int4 sg = seg[blockIdx.x];
    double sum, unused;
    segment_sums(sg, red, ntile, sum, unused);
    if (threadIdx.x != 0) return;
    double eps = sqrt(sum / sg.z);
    double it = sc[SC_ITER] + 1.0;
    sc[SC_EPS] = eps;
    sc[SC_ITER] = it;
    if (it >= sc[SC_MAXIT] || !(eps > sc[SC_TARGET])) mark_done(sc, done);
}

// ---- Jacobi-preconditioned CG on (alpha^-1 - T) mu = E0, stopped on the same RMS(alpha r) the
// Jacobi iteration uses (alpha r is exactly the Jacobi update).  A particle with alpha = 0 keeps
// mu = 0 (its z and p stay 0), as in the reference's update mu = alpha (E0 + T mu).
// init: x = mu (= alpha E0), r = E0 - x/alpha + T x, z = alpha r, p = z.
__global__ __launch_bounds__(64) void k_pcg_init(const double4* __restrict__ part, const double4* __restrict__ E0,
                                                 const double* __restrict__ alpha, const double4* __restrict__ x,
                                                 double4* __restrict__ r, double4* __restrict__ p,
                                                 float4* __restrict__ p32, double* __restrict__ red, int ntile) {
    const int i = blockIdx.x * kTile + threadIdx.x;
    double4 f = part[i];
    double4 e0 = E0[i], xv = x[i];
    double a = alpha[i];
    double ia = inv_or_zero(a);
    double4 rv = make_double4(e0.x - xv.x * ia + f.x, e0.y - xv.y * ia + f.y, e0.z - xv.z * ia + f.z, 0.0);
    double4 zv = make_double4(rv.x * a, rv.y * a, rv.z * a, 0.0);
    r[i] = rv;
    p[i] = zv;
    if (p32) p32[i] = make_float4((float)zv.x, (float)zv.y, (float)zv.z, 0.f);
    tile_partials(rv.x * zv.x + rv.y * zv.y + rv.z * zv.z, zv.x * zv.x + zv.y * zv.y + zv.z * zv.z, red, ntile);
}

__global__ __launch_bounds__(64) void k_pcg_init_finalize(const int4* __restrict__ seg, const double* __restrict__ red,
                                                          int ntile, double* __restrict__ scalars,
                                                          int* __restrict__ done) {
    double* sc = scalars + SC_COUNT * blockIdx.x;
    const int4 sg = seg[blockIdx.x];
    double rz, zz;
    segment_sums(sg, red, ntile, rz, zz);
    if (threadIdx.x != 0) return;
    double eps = sqrt(zz / sg.z);
    sc[SC_RZ] = rz;
    sc[SC_EPS] = eps;
    sc[SC_ITER] = 0.0;
    sc[SC_DONE] = 0.0;
    if (sc[SC_MAXIT] <= 0.0 || !(eps > sc[SC_TARGET])) mark_done(sc, done);
}

// One PCG iteration after the reduction of T p, which left q = p/alpha - T p and the tile sums of
// p.q in red_pq.  Every workgroup of a system sums the system's tile sums itself (segment_sum, the
// same order in every workgroup, so all of them get the same value) instead of waiting for a
// separate launch; the scalars that these sums replace are written by the system's first tile
// only, into slots that no workgroup of the same launch reads.
// x += gamma p; r -= gamma q; z = alpha r (gamma = r.z / p.q); partial r.z and z.z.
__global__ __launch_bounds__(64) void k_pcg_update(const int* __restrict__ tseg, const int4* __restrict__ seg,
                                                   const double* __restrict__ alpha, const double4* __restrict__ p,
                                                   const double4* __restrict__ q, double4* __restrict__ x,
                                                   double4* __restrict__ r, double4* __restrict__ z,
                                                   double* __restrict__ scalars, const double* __restrict__ red_pq,
                                                   double* __restrict__ red, int ntile) {
    // the site's vectors are loaded first: their latency overlaps that of the scalars and the sums
    const int i = blockIdx.x * kTile + threadIdx.x;
    double4 pv = p[i], qv = q[i], xv = x[i], rv = r[i];
    const double a = alpha[i];
    const int g = tseg[blockIdx.x];
    double* sc = scalars + SC_COUNT * g;
    if (solver_done(sc)) return;
    const int4 sg = seg[g];
    const double rz = sc[SC_RZ];
    const double gam = rz / segment_sum(sg, red_pq);
    if ((int)blockIdx.x == sg.x && threadIdx.x == 0) sc[SC_RZOLD] = rz;   // for k_pcg_direction
    xv.x += gam * pv.x; xv.y += gam * pv.y; xv.z += gam * pv.z;
    rv.x -= gam * qv.x; rv.y -= gam * qv.y; rv.z -= gam * qv.z;
    double4 zv = make_double4(rv.x * a, rv.y * a, rv.z * a, 0.0);
    x[i] = xv;
    r[i] = rv;
    z[i] = zv;
    tile_partials(rv.x * zv.x + rv.y * zv.y + rv.z * zv.z, zv.x * zv.x + zv.y * zv.y + zv.z * zv.z, red, ntile);
}

// p = z + beta p (beta = new r.z / old r.z).  The system's first tile also records r.z, the RMS of
// alpha r (the Jacobi update, eps) and the iteration count, and stops the solve when iterations ==
// max or !(eps > target).  Workgroups of the same launch that then see the stop flag skip the p
// update: p is not used after the last iteration.
__global__ __launch_bounds__(64) void k_pcg_direction(const int* __restrict__ tseg, const int4* __restrict__ seg,
                                                      const double4* __restrict__ z, double4* __restrict__ p,
                                                      float4* __restrict__ p32, double* __restrict__ scalars,
                                                      const double* __restrict__ red, int ntile,
                                                      int* __restrict__ done) {
    const int i = blockIdx.x * kTile + threadIdx.x;   // (loaded first, as in k_pcg_update)
    const double4 zv = z[i], pv = p[i];
    const int g = tseg[blockIdx.x];
    double* sc = scalars + SC_COUNT * g;
    if (solver_done(sc)) return;
    const int4 sg = seg[g];
    double rz, zz;
    segment_sums(sg, red, ntile, rz, zz);
    const double b = rz / sc[SC_RZOLD];
    double4 np = make_double4(zv.x + b * pv.x, zv.y + b * pv.y, zv.z + b * pv.z, 0.0);
    p[i] = np;
    if (p32) p32[i] = make_float4((float)np.x, (float)np.y, (float)np.z, 0.f);
    if ((int)blockIdx.x == sg.x && threadIdx.x == 0) {
        const double eps = sqrt(zz / sg.z);
        const double it = sc[SC_ITER] + 1.0;
        sc[SC_RZ] = rz;
        sc[SC_EPS] = eps;
        sc[SC_ITER] = it;
        if (it >= sc[SC_MAXIT] || !(eps > sc[SC_TARGET])) mark_done(sc, done);
    }
}

// Extrapolated polarization (convergeInducedDipolesByExtrapolation, ReferenceTholeDipoleForce.cpp:876-887):
// mu = alpha*(E0 + T mu); hist[o] = mu.
__global__ void k_extrap_step(int npad, const double4* __restrict__ part, const double4* __restrict__ E0,
                              const double* __restrict__ alpha, double4* __restrict__ mu, float4* __restrict__ mu32,
                              double4* __restrict__ hist_o) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npad) return;
    double4 f = part[i];
    double4 e0 = E0[i];
    double a = alpha[i];
    double4 m = make_double4((e0.x + f.x) * a, (e0.y + f.y) * a, (e0.z + f.z) * a, 0.0);
    mu[i] = m;
    if (mu32) mu32[i] = make_float4((float)m.x, (float)m.y, (float)m.z, 0.f);
    hist_o[i] = m;
}

// mu = sum_o c_o hist[o]  (:890-895)
__global__ void k_extrap_combine(int npad, int order, const double4* __restrict__ hist, const double* __restrict__ coef,
                                 double4* __restrict__ mu, float4* __restrict__ mu32) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npad) return;
    double x = 0, y = 0, z = 0;
    for (int o = 0; o < order; o++) {
        double4 h = hist[(size_t)o * npad + i];
        double c = coef[o];
        x += h.x * c; y += h.y * c; z += h.z * c;
    }
    mu[i] = make_double4(x, y, z, 0.0);
    if (mu32) mu32[i] = make_float4((float)x, (float)y, (float)z, 0.f);
}

// final error of the extrapolated dipoles (:897-906): partial sums of |alpha(E0+T mu) - mu|^2
__global__ __launch_bounds__(64) void k_extrap_error(const double4* __restrict__ part, const double4* __restrict__ E0,
                                                     const double* __restrict__ alpha, const double4* __restrict__ mu,
                                                     double* __restrict__ red, int ntile) {
    const int i = blockIdx.x * kTile + threadIdx.x;
    double4 f = part[i];
    double4 e0 = E0[i], m = mu[i];
    double a = alpha[i];
    double dx = (e0.x + f.x) * a - m.x, dy = (e0.y + f.y) * a - m.y, dz = (e0.z + f.z) * a - m.z;
    tile_partials(dx * dx + dy * dy + dz * dz, 0.0, red, ntile);
}

__global__ __launch_bounds__(64) void k_sum_eps(const int4* __restrict__ seg, const double* __restrict__ red,
                                                int ntile, double* __restrict__ scalars) {
    const int4 sg = seg[blockIdx.x];
    double sum, unused;
    segment_sums(sg, red, ntile, sum, unused);
    if (threadIdx.x == 0) scalars[SC_COUNT * blockIdx.x + SC_EPS] = sqrt(sum / sg.z);
}

// Energy of system s: the pair-kernel partials of its work items [w0, w1) in a fixed order
// (strided per thread, then a tree over the 256 threads).
constexpr int kEnergyThreads = 1024;
__global__ __launch_bounds__(kEnergyThreads) void k_reduce_energy(const int2* __restrict__ segw,
                                                                  const double* __restrict__ part_e,
                                                                  double* __restrict__ out) {
    __shared__ double sh[kEnergyThreads];
    const int2 w = segw[blockIdx.x];
    const int t = threadIdx.x;
    constexpr int T = kEnergyThreads;
    // four independent loads in flight per thread, then a fixed-order tree
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int k = w.x + t;
    for (; k + 3 * T < w.y; k += 4 * T) {
        a0 += part_e[k];
        a1 += part_e[k + T];
        a2 += part_e[k + 2 * T];
        a3 += part_e[k + 3 * T];
    }
    for (; k < w.y; k += T) a0 += part_e[k];
    sh[t] = (a0 + a1) + (a2 + a3);
    __syncthreads();
    for (int h = T / 2; h > 0; h >>= 1) {
        if (t < h) sh[t] += sh[t + h];
        __syncthreads();
    }
    if (t == 0) out[blockIdx.x] = sh[0];
}

// ------------------------------------------------------------------ torque -> force (a10)
// mapTorqueToForceForParticle (ReferenceTholeDipoleForce.cpp:915-1090), one thread per particle
// with an axis; writes the contributions to I, U(=Z), V(=X), W(=Y) into contrib[4*i + role].
__device__ __forceinline__ double nrm3(double& x, double& y, double& z) {
    double n = sqrt(x * x + y * y + z * z);
    if (n > 0.0) { double s = 1.0 / n; x *= s; y *= s; z *= s; }
    return n;
}
__device__ __forceinline__ void crs(double ax, double ay, double az, double bx, double by, double bz, double& cx,
                                    double& cy, double& cz) {
    cx = ay * bz - az * by;
    cy = az * bx - ax * bz;
    cz = ax * by - ay * bx;
}

__global__ void k_map_torque(int n, const double4* __restrict__ posa, const int4* __restrict__ axis,
                             const double4* __restrict__ torque, double4* __restrict__ contrib) {
    int I = blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= n) return;
    int4 ax = axis[I];
    double4 cI = make_double4(0, 0, 0, 0), cU = cI, cV = cI, cW = cI;
    const int type = ax.x;
    if (type != 5) {
        const double4 pI = posa[I], pU = posa[ax.y], pV = posa[ax.z];
        const int W = ax.w > -1 ? ax.w : -1;
        double ux = pU.x - pI.x, uy = pU.y - pI.y, uz = pU.z - pI.z;
        double vx = pV.x - pI.x, vy = pV.y - pI.y, vz = pV.z - pI.z;
        double normU = nrm3(ux, uy, uz);
        double normV = nrm3(vx, vy, vz);
        double wx, wy, wz;
        if (W >= 0 && (type == 2 || type == 3)) {
            double4 pW = posa[W];
            wx = pW.x - pI.x; wy = pW.y - pI.y; wz = pW.z - pI.z;
        } else {
            crs(ux, uy, uz, vx, vy, vz, wx, wy, wz);
        }
        double normW = nrm3(wx, wy, wz);
        double uvx, uvy, uvz, uwx, uwy, uwz, vwx, vwy, vwz;
        crs(vx, vy, vz, ux, uy, uz, uvx, uvy, uvz);
        crs(wx, wy, wz, ux, uy, uz, uwx, uwy, uwz);
        crs(wx, wy, wz, vx, vy, vz, vwx, vwy, vwz);
        nrm3(uvx, uvy, uvz);
        nrm3(uwx, uwy, uwz);
        nrm3(vwx, vwy, vwz);
        double cosUV = ux * vx + uy * vy + uz * vz, sinUV = sqrt(1.0 - cosUV * cosUV);
        double cosUW = ux * wx + uy * wy + uz * wz, sinUW = sqrt(1.0 - cosUW * cosUW);
        double cosVW = vx * wx + vy * wy + vz * wz, sinVW = sqrt(1.0 - cosVW * cosVW);
        double4 t = torque[I];
        double dphi0 = -(ux * t.x + uy * t.y + uz * t.z);
        double dphi1 = -(vx * t.x + vy * t.y + vz * t.z);
        double dphi2 = -(wx * t.x + wy * t.y + wz * t.z);
        if (type == 0 || type == 1) {
            double f1 = dphi1 / (normU * sinUV);
            double f2 = dphi2 / normU;
            double f3 = -dphi0 / (normV * sinUV);
            double f4 = 0.0;
            if (type == 1) {
                f2 *= 0.5;
                f4 = 0.5 * dphi2 / normV;
            }
            double FUx = uvx * f1 + f2 * uwx, FUy = uvy * f1 + f2 * uwy, FUz = uvz * f1 + f2 * uwz;
            double FVx = uvx * f3 + f4 * vwx, FVy = uvy * f3 + f4 * vwy, FVz = uvz * f3 + f4 * vwz;
            cU = make_double4(-FUx, -FUy, -FUz, 0);
            cV = make_double4(-FVx, -FVy, -FVz, 0);
            cI = make_double4(FUx + FVx, FUy + FVy, FUz + FVz, 0);
        } else if (type == 2) {
            double rx = vx + wx, ry = vy + wy, rz = vz + wz;
            double sx, sy, sz;
            crs(ux, uy, uz, rx, ry, rz, sx, sy, sz);
            nrm3(rx, ry, rz);
            nrm3(sx, sy, sz);
            double urx, ury, urz, usx, usy, usz;
            crs(rx, ry, rz, ux, uy, uz, urx, ury, urz);
            crs(sx, sy, sz, ux, uy, uz, usx, usy, usz);
            nrm3(urx, ury, urz);
            nrm3(usx, usy, usz);
            double cosUR = ux * rx + uy * ry + uz * rz, sinUR = sqrt(1.0 - cosUR * cosUR);
            double cosVS = vx * sx + vy * sy + vz * sz, sinVS = sqrt(1.0 - cosVS * cosVS);
            double cosWS = wx * sx + wy * sy + wz * sz, sinWS = sqrt(1.0 - cosWS * cosWS);
            double t1x = vx - sx * cosVS, t1y = vy - sy * cosVS, t1z = vz - sz * cosVS;
            double t2x = wx - sx * cosWS, t2y = wy - sy * cosWS, t2z = wz - sz * cosWS;
            nrm3(t1x, t1y, t1z);
            nrm3(t2x, t2y, t2z);
            double ut1cos = ux * t1x + uy * t1y + uz * t1z, ut1sin = sqrt(1.0 - ut1cos * ut1cos);
            double ut2cos = ux * t2x + uy * t2y + uz * t2z, ut2sin = sqrt(1.0 - ut2cos * ut2cos);
            double dphiR = -(rx * t.x + ry * t.y + rz * t.z);
            double dphiS = -(sx * t.x + sy * t.y + sz * t.z);
            double f1 = dphiR / (normU * sinUR);
            double f2 = dphiS / normU;
            double f3 = dphi0 / (normV * (ut1sin + ut2sin));
            double f4 = dphi0 / (normW * (ut1sin + ut2sin));
            double FUx = urx * f1 + usx * f2, FUy = ury * f1 + usy * f2, FUz = urz * f1 + usz * f2;
            double FVx = (sx * sinVS - t1x * cosVS) * f3, FVy = (sy * sinVS - t1y * cosVS) * f3,
                   FVz = (sz * sinVS - t1z * cosVS) * f3;
            double FWx = (sx * sinWS - t2x * cosWS) * f4, FWy = (sy * sinWS - t2y * cosWS) * f4,
                   FWz = (sz * sinWS - t2z * cosWS) * f4;
            cU = make_double4(-FUx, -FUy, -FUz, 0);
            cV = make_double4(-FVx, -FVy, -FVz, 0);
            cW = make_double4(-FWx, -FWy, -FWz, 0);
            cI = make_double4(FUx + FVx + FWx, FUy + FVy + FWy, FUz + FVz + FWz, 0);
        } else if (type == 3) {
            double du[3], dv[3], dw[3];
            const double UW[3] = {uwx, uwy, uwz}, UV[3] = {uvx, uvy, uvz}, VW[3] = {vwx, vwy, vwz};
            for (int k = 0; k < 3; k++) {
                du[k] = (UW[k] * dphi2 / (normU * sinUW) + UV[k] * dphi1 / (normU * sinUV) -
                         UW[k] * dphi0 / (normU * sinUW) - UV[k] * dphi0 / (normU * sinUV)) / 3.0;
                dv[k] = (VW[k] * dphi2 / (normV * sinVW) - UV[k] * dphi0 / (normV * sinUV) -
                         VW[k] * dphi1 / (normV * sinVW) + UV[k] * dphi1 / (normV * sinUV)) / 3.0;
                dw[k] = (-UW[k] * dphi0 / (normW * sinUW) - VW[k] * dphi1 / (normW * sinVW) +
                         UW[k] * dphi2 / (normW * sinUW) + VW[k] * dphi2 / (normW * sinVW)) / 3.0;
            }
            cU = make_double4(-du[0], -du[1], -du[2], 0);
            cV = make_double4(-dv[0], -dv[1], -dv[2], 0);
            if (W >= 0) cW = make_double4(-dw[0], -dw[1], -dw[2], 0);
            cI = make_double4(du[0] + dv[0] + dw[0], du[1] + dv[1] + dw[1], du[2] + dv[2] + dw[2], 0);
        } else if (type == 4) {
            double dux = uvx * dphi1 / (normU * sinUV) + uwx * dphi2 / normU;
            double duy = uvy * dphi1 / (normU * sinUV) + uwy * dphi2 / normU;
            double duz = uvz * dphi1 / (normU * sinUV) + uwz * dphi2 / normU;
            cU = make_double4(-dux, -duy, -duz, 0);
            cI = make_double4(dux, duy, duz, 0);
        }
    }
    contrib[4 * I + 0] = cI;
    contrib[4 * I + 1] = cU;
    contrib[4 * I + 2] = cV;
    contrib[4 * I + 3] = cW;
}

// Gather of the torque contributions: atom a receives the entries listed in gather[goff[a]..goff[a+1])
// (indices into contrib, host-sorted so the sum order is fixed), added to frc.
__global__ void k_gather_torque(int n, const int* __restrict__ goff, const int* __restrict__ gidx,
                                const double4* __restrict__ contrib, double* __restrict__ frc) {
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    double x = 0, y = 0, z = 0;
    for (int k = goff[a]; k < goff[a + 1]; k++) {
        double4 c = contrib[gidx[k]];
        x += c.x; y += c.y; z += c.z;
    }
    frc[3 * a] += x;
    frc[3 * a + 1] += y;
    frc[3 * a + 2] += z;
}

// dst[perm[i]] += src[i] per component: forces in the internal order are accumulated into the
// caller's array in its own order.
// With a guard, only once every system's solve has stopped (*done >= nseg): the launch can then be
// enqueued before the host has checked the solve.
__global__ void k_scatter_accumulate(int n, const int* __restrict__ perm, const double* __restrict__ src,
                                     double* __restrict__ dst, const int* __restrict__ done, int nseg) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n || (done && *done < nseg)) return;
    int o = perm[k];
    if (o < 0) return;
    dst[3 * o] += src[3 * k];
    dst[3 * o + 1] += src[3 * k + 1];
    dst[3 * o + 2] += src[3 * k + 2];
}

// ------------------------------------------------------------------ potential (next row)
// calculateElectrostaticPotential (ReferenceTholeDipoleForce.cpp:1315-1372):
// phi(g) = sum_i q_i/r - (d_i + mu_i).Delta/r^3, Delta = x_i - g (min-imaged).
template <bool PBC>
__global__ __launch_bounds__(64) void k_potential(int n, int ngrid, const double* __restrict__ grid,
                                                  const double4* __restrict__ posa, const double4* __restrict__ qd,
                                                  const double4* __restrict__ mu, BoxT64 box, double* __restrict__ out) {
    __shared__ double4 s_p[kTile];
    __shared__ double4 s_q[kTile];
    __shared__ double4 s_u[kTile];
    int lane = threadIdx.x;
    int g = blockIdx.x * kTile + lane;
    double gx = 0, gy = 0, gz = 0;
    if (g < ngrid) { gx = grid[3 * g]; gy = grid[3 * g + 1]; gz = grid[3 * g + 2]; }
    double phi = 0.0;
    int nblk = (n + kTile - 1) / kTile;
    for (int J = 0; J < nblk; J++) {
        int jb = J * kTile;
        __syncthreads();
        s_p[lane] = posa[jb + lane];
        s_q[lane] = qd[jb + lane];
        s_u[lane] = mu[jb + lane];
        __syncthreads();
        int jn = min(kTile, n - jb);
        for (int jj = 0; jj < jn; jj++) {
            double4 p = s_p[jj], q = s_q[jj], u = s_u[jj];
            double dx = p.x - gx, dy = p.y - gy, dz = p.z - gz;
            if (PBC) min_image(dx, dy, dz, box.bx, box.by, box.bz, box.ibx, box.iby, box.ibz);
            double r2 = dx * dx + dy * dy + dz * dz;
            double rr1 = 1.0 / sqrt(r2);
            double rr3 = rr1 * rr1 * rr1;
            double v = q.w * rr1;
            double sc = (q.x + u.x) * dx + (q.y + u.y) * dy + (q.z + u.z) * dz;
            phi += v - sc * rr3;
        }
    }
    if (g < ngrid) out[g] = phi;
}

// ------------------------------------------------------------------ multipole moments (next row)
// calculateTholeDipoleSystemMultipoleMoments (ReferenceTholeDipoleForce.cpp:1225-1313): raw sums
// in one 256-thread block (fixed-order reduction), unit conversion on the host.
__global__ __launch_bounds__(256) void k_moments(int n, const double4* __restrict__ posa, const double4* __restrict__ qd,
                                                 const double4* __restrict__ mu, const double* __restrict__ mass,
                                                 double* __restrict__ out /* 4 + 10 */) {
    __shared__ double s[256][10];
    __shared__ double com[4];
    int t = threadIdx.x;
    double acc[10];
    for (int k = 0; k < 4; k++) acc[k] = 0.0;
    for (int i = t; i < n; i += 256) {
        double m = mass[i];
        double4 p = posa[i];
        acc[0] += m; acc[1] += p.x * m; acc[2] += p.y * m; acc[3] += p.z * m;
    }
    for (int k = 0; k < 4; k++) s[t][k] = acc[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w)
            for (int k = 0; k < 4; k++) s[t][k] += s[t + w][k];
        __syncthreads();
    }
    if (t == 0) {
        double tm = s[0][0];
        double cx = s[0][1], cy = s[0][2], cz = s[0][3];
        if (tm > 0.0) { double inv = 1.0 / tm; cx *= inv; cy *= inv; cz *= inv; }
        com[0] = cx; com[1] = cy; com[2] = cz; com[3] = tm;
    }
    __syncthreads();
    for (int k = 0; k < 10; k++) acc[k] = 0.0;
    for (int i = t; i < n; i += 256) {
        double4 p = posa[i], q = qd[i], u = mu[i];
        double x = p.x - com[0], y = p.y - com[1], z = p.z - com[2];
        double c = q.w;
        double nx = q.x + u.x, ny = q.y + u.y, nz = q.z + u.z;
        acc[0] += c;
        acc[1] += x * c + nx; acc[2] += y * c + ny; acc[3] += z * c + nz;
        acc[4] += x * x * c + 2.0 * x * nx;
        acc[5] += x * y * c + x * ny + y * nx;
        acc[6] += x * z * c + x * nz + z * nx;
        acc[7] += y * y * c + 2.0 * y * ny;
        acc[8] += y * z * c + y * nz + z * ny;
        acc[9] += z * z * c + 2.0 * z * nz;
    }
    for (int k = 0; k < 10; k++) s[t][k] = acc[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w)
            for (int k = 0; k < 10; k++) s[t][k] += s[t + w][k];
        __syncthreads();
    }
    if (t == 0)
        for (int k = 0; k < 10; k++) out[k] = s[0][k];
}

}  // namespace

// ====================================================================== launch wrappers
namespace tdp {

static BoxT64 mkbox(const double* b3) {
    BoxT64 b;
    b.bx = b3[0]; b.by = b3[1]; b.bz = b3[2];
    b.ibx = b3[0] > 0 ? 1.0 / b3[0] : 0.0;
    b.iby = b3[1] > 0 ? 1.0 / b3[1] : 0.0;
    b.ibz = b3[2] > 0 ? 1.0 / b3[2] : 0.0;
    return b;
}

void launch_load_positions(hipStream_t s, int npad, const double* pos3, const int* perm, const double* thole,
                           double4* posa) {
    hipLaunchKernelGGL(k_load_positions, dim3((npad + 255) / 256), dim3(256), 0, s, npad, pos3, perm, thole, posa);
}

void launch_tile_pairs(hipStream_t s, const PairLaunch& L, TilePair* tp) {
    if (L.nwork <= 0) return;
    BoxT64 b = mkbox(L.box);
    const dim3 grid((L.nwork + 63) / 64);
    if (L.pbc)
        hipLaunchKernelGGL(k_tile_pairs<true>, grid, dim3(64), 0, s, L.tm, L.work, L.nwork, L.tb_c, L.tb_h, L.tb_o, b, tp);
    else
        hipLaunchKernelGGL(k_tile_pairs<false>, grid, dim3(64), 0, s, L.tm, L.work, L.nwork, L.tb_c, L.tb_h, L.tb_o, b,
                           tp);
}

void launch_tile_bounds(hipStream_t s, const int* tcnt, int nblk, const double4* posa, const double* box3, bool pbc,
                        float4* tb_c, float4* tb_h, double4* tb_o, float4* posl, double4* posl64) {
    hipLaunchKernelGGL(k_tile_bounds, dim3(nblk), dim3(kTile), 0, s, tcnt, posa, box3[0], box3[1], box3[2],
                       pbc ? 1 : 0, tb_c, tb_h, tb_o, posl, posl64);
}

void launch_gather_vec3(hipStream_t s, int npad, const int* perm, const double* src3, double4* dst4, float4* dst32) {
    hipLaunchKernelGGL(k_gather_vec3, dim3((npad + 255) / 256), dim3(256), 0, s, npad, perm, src3, dst4, dst32);
}

void launch_scatter_vec3(hipStream_t s, int npad, const int* perm, const double4* src4, double* dst3) {
    hipLaunchKernelGGL(k_scatter_vec3, dim3((npad + 255) / 256), dim3(256), 0, s, npad, perm, src4, dst3);
}

void launch_rotate(hipStream_t s, int n, const double4* posa, const double* dmol, const double* charge,
                   const int4* axis, double4* qd) {
    hipLaunchKernelGGL(k_rotate_dipoles, dim3((n + 255) / 256), dim3(256), 0, s, n, posa, dmol, charge, axis, qd);
}

void launch_col_reduce(hipStream_t s, const PairLaunch& L, const double4* Pi, const double* Pj,
                       const double* scalars, double4* out4, double* out3, const float* Pjf, const PcgQ& pq) {
    if (Pjf)
        hipLaunchKernelGGL(k_col_reduce<true>, dim3(L.nblk), dim3(kRedWaves * kTile), 0, s, L.tm, L.row_start, Pi, Pj,
                           Pjf, L.slab, scalars, out4, out3, pq);
    else
        hipLaunchKernelGGL(k_col_reduce<false>, dim3(L.nblk), dim3(kRedWaves * kTile), 0, s, L.tm, L.row_start, Pi, Pj,
                           Pjf, L.slab, scalars, out4, out3, pq);
}

void launch_sym_matvec64(hipStream_t s, const PairLaunch& L, const double4* posa, const double4* mu, const double* inva,
                         const double* scalars, double4* Pi, double* Pj) {
    BoxT64 b = mkbox(L.box);
    if (L.pbc)
        hipLaunchKernelGGL(k_sym_matvec<true>, dim3(L.nwork), dim3(kTile), 0, s, L.tm, L.nblk, L.work, posa, mu, inva,
                           L.exc_start, L.exc_J, L.exc_mask, L.tb_c, L.tb_h, L.tp, L.posl64, b, scalars, Pi, Pj,
                           L.slab);
    else
        hipLaunchKernelGGL(k_sym_matvec<false>, dim3(L.nwork), dim3(kTile), 0, s, L.tm, L.nblk, L.work, posa, mu, inva,
                           L.exc_start, L.exc_J, L.exc_mask, L.tb_c, L.tb_h, L.tp, L.posl64, b, scalars, Pi, Pj,
                           L.slab);
}

void launch_sym_matvec_mixed(hipStream_t s, const PairLaunch& L, const double4* posa, const double4* mu,
                             const float4* mu32, const double* inva, const double* scalars, float* Pjf) {
    BoxT64 b = mkbox(L.box);
    if (L.pbc)
        hipLaunchKernelGGL(k_sym_matvec_mixed<true>, dim3(L.nwork), dim3(kTile), 0, s, L.tm, L.nblk, L.work, posa, mu,
                           mu32, inva, L.exc_start, L.exc_J, L.exc_mask, L.tb_c, L.tb_h, L.tp, L.tb_o, L.posl, b, scalars,
                           Pjf, L.slab);
    else
        hipLaunchKernelGGL(k_sym_matvec_mixed<false>, dim3(L.nwork), dim3(kTile), 0, s, L.tm, L.nblk, L.work, posa, mu,
                           mu32, inva, L.exc_start, L.exc_J, L.exc_mask, L.tb_c, L.tb_h, L.tp, L.tb_o, L.posl, b, scalars,
                           Pjf, L.slab);
}

void launch_sym_field(hipStream_t s, const PairLaunch& L, const double4* posa, const double4* qd, const double* inva,
                      double4* Pi, double* Pj, double4* out) {
    BoxT64 b = mkbox(L.box);
#define TDP_SFIELD(P, Q)                                                                                          \
    hipLaunchKernelGGL((k_sym_field<P, Q>), dim3(L.nwork), dim3(kTile), 0, s, L.tm, L.nblk, L.work, posa, qd, inva, \
                       L.exc_start, L.exc_J, L.exc_mask, L.tb_c, L.tb_h, L.tp, b, Pi, Pj, L.slab)
    if (L.pbc) {
        if (L.perm) TDP_SFIELD(true, true); else TDP_SFIELD(true, false);
    } else {
        if (L.perm) TDP_SFIELD(false, true); else TDP_SFIELD(false, false);
    }
#undef TDP_SFIELD
    launch_col_reduce(s, L, Pi, Pj, nullptr, out, nullptr);
}

void launch_sym_elec(hipStream_t s, const PairLaunch& L, bool mutual, bool mixed, const double4* posa, const double4* qd,
                     const double4* mu, const double* inva, const int* orig, double4* PiF, double* PjF, double4* PiT,
                     double* PjT, double* part_e, double* frc, double4* torque, float* slab32) {
    BoxT64 b = mkbox(L.box);
#define TDP_SELEC(P, M, X, Q)                                                                                         \
    hipLaunchKernelGGL((k_sym_elec<P, M, X, Q>), dim3(L.nwork), dim3(kTile), 0, s, L.tm, L.nblk, L.work, posa, qd, \
                       mu, inva, orig, L.exc_start, L.exc_J, L.exc_mask, L.tb_c, L.tb_h, L.tp, L.tb_o, L.posl, b, PiF, PjF,  \
                       PiT, PjT, L.slab, part_e, slab32)
#define TDP_SELEC_Q(P, M, X)              \
    if (L.perm) TDP_SELEC(P, M, X, true); \
    else TDP_SELEC(P, M, X, false)
#define TDP_SELEC_M(P, X)                                   \
    if (mutual) { TDP_SELEC_Q(P, true, X); }                \
    else { TDP_SELEC_Q(P, false, X); }
    if (mixed) {
        if (L.pbc) { TDP_SELEC_M(true, true) } else { TDP_SELEC_M(false, true) }
    } else {
        if (L.pbc) { TDP_SELEC_M(true, false) } else { TDP_SELEC_M(false, false) }
    }
#undef TDP_SELEC_M
#undef TDP_SELEC_Q
#undef TDP_SELEC
    launch_col_reduce(s, L, PiF, PjF, nullptr, nullptr, frc, mixed && !L.perm ? slab32 : nullptr);
    if (L.perm) launch_col_reduce(s, L, PiT, PjT, nullptr, torque, nullptr);
}

void launch_field_init(hipStream_t s, const SegLaunch& G, const double4* part, const double* alpha, double4* E0,
                       double4* mu, float4* mu32) {
    hipLaunchKernelGGL(k_field_init, dim3((G.npad + 255) / 256), dim3(256), 0, s, G.npad, part, alpha, E0, mu, mu32);
}

void launch_jacobi_update(hipStream_t s, const SegLaunch& G, const double4* part, const double4* E0,
                          const double* alpha, double4* mu, float4* mu32, double* scalars, double* red) {
    hipLaunchKernelGGL(k_jacobi_update, dim3(G.ntile), dim3(kTile), 0, s, G.tseg, part, E0, alpha, mu, mu32,
                       (const double*)scalars, red, G.ntile);
    hipLaunchKernelGGL(k_jacobi_finalize, dim3(G.nseg), dim3(kTile), 0, s, G.seg, (const double*)red, G.ntile, scalars,
                       G.done);
}

void launch_pcg_init(hipStream_t s, const SegLaunch& G, const double4* part, const double4* E0, const double* alpha,
                     const double4* x, double4* r, double4* p, float4* p32, double* red, double* scalars) {
    hipLaunchKernelGGL(k_pcg_init, dim3(G.ntile), dim3(kTile), 0, s, part, E0, alpha, x, r, p, p32, red, G.ntile);
    hipLaunchKernelGGL(k_pcg_init_finalize, dim3(G.nseg), dim3(kTile), 0, s, G.seg, (const double*)red, G.ntile,
                       scalars, G.done);
}

void launch_pcg_iteration_tail(hipStream_t s, const SegLaunch& G, const double* alpha, double4* x, double4* r,
                               double4* z, double4* p, float4* p32, const double4* q, double* red, double* scalars) {
    hipLaunchKernelGGL(k_pcg_update, dim3(G.ntile), dim3(kTile), 0, s, G.tseg, G.seg, alpha, (const double4*)p, q, x,
                       r, z, scalars, (const double*)(red + 2 * (size_t)G.ntile), red, G.ntile);
    hipLaunchKernelGGL(k_pcg_direction, dim3(G.ntile), dim3(kTile), 0, s, G.tseg, G.seg, (const double4*)z, p, p32,
                       scalars, (const double*)red, G.ntile, G.done);
}

void launch_extrap_step(hipStream_t s, const SegLaunch& G, const double4* part, const double4* E0,
                        const double* alpha, double4* mu, float4* mu32, double4* hist_o) {
    hipLaunchKernelGGL(k_extrap_step, dim3((G.npad + 255) / 256), dim3(256), 0, s, G.npad, part, E0, alpha, mu, mu32,
                       hist_o);
}

void launch_extrap_combine(hipStream_t s, const SegLaunch& G, int order, const double4* hist, const double* coef,
                           double4* mu, float4* mu32) {
    hipLaunchKernelGGL(k_extrap_combine, dim3((G.npad + 255) / 256), dim3(256), 0, s, G.npad, order, hist, coef, mu,
                       mu32);
}

void launch_extrap_error(hipStream_t s, const SegLaunch& G, const double4* part, const double4* E0,
                         const double* alpha, const double4* mu, double* red, double* scalars) {
    hipLaunchKernelGGL(k_extrap_error, dim3(G.ntile), dim3(kTile), 0, s, part, E0, alpha, mu, red, G.ntile);
    hipLaunchKernelGGL(k_sum_eps, dim3(G.nseg), dim3(kTile), 0, s, G.seg, (const double*)red, G.ntile, scalars);
}

void launch_reduce_energy(hipStream_t s, const SegLaunch& G, const double* part_e, double* out) {
    hipLaunchKernelGGL(k_reduce_energy, dim3(G.nseg), dim3(kEnergyThreads), 0, s, G.segw, part_e, out);
}

void launch_map_torque(hipStream_t s, int n, const double4* posa, const int4* axis, const double4* torque,
                       double4* contrib) {
    hipLaunchKernelGGL(k_map_torque, dim3((n + 255) / 256), dim3(256), 0, s, n, posa, axis, torque, contrib);
}

void launch_gather_torque(hipStream_t s, int n, const int* goff, const int* gidx, const double4* contrib, double* frc) {
    hipLaunchKernelGGL(k_gather_torque, dim3((n + 255) / 256), dim3(256), 0, s, n, goff, gidx, contrib, frc);
}

__global__ void k_copy_guarded(int n, const double* __restrict__ src, double* __restrict__ dst,
                               const int* __restrict__ done, int nseg) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n || *done < nseg) return;
    dst[k] = src[k];
}

void launch_scatter_accumulate(hipStream_t s, int n, const int* perm, const double* src, double* dst, const int* done,
                               int nseg) {
    hipLaunchKernelGGL(k_scatter_accumulate, dim3((n + 255) / 256), dim3(256), 0, s, n, perm, src, dst, done, nseg);
}

void launch_copy_guarded(hipStream_t s, int n, const double* src, double* dst, const int* done, int nseg) {
    hipLaunchKernelGGL(k_copy_guarded, dim3((n + 255) / 256), dim3(256), 0, s, n, src, dst, done, nseg);
}

void launch_potential(hipStream_t s, int n, int ngrid, const double* grid, const double4* posa, const double4* qd,
                      const double4* mu, bool pbc, const double* box3, double* out) {
    BoxT64 b = mkbox(box3);
    dim3 g((ngrid + kTile - 1) / kTile);
    if (pbc)
        hipLaunchKernelGGL(k_potential<true>, g, dim3(kTile), 0, s, n, ngrid, grid, posa, qd, mu, b, out);
    else
        hipLaunchKernelGGL(k_potential<false>, g, dim3(kTile), 0, s, n, ngrid, grid, posa, qd, mu, b, out);
}

void launch_moments(hipStream_t s, int n, const double4* posa, const double4* qd, const double4* mu,
                    const double* mass, double* out) {
    hipLaunchKernelGGL(k_moments, dim3(1), dim3(256), 0, s, n, posa, qd, mu, mass, out);
}

}  // namespace tdp
