// tdp_device.hpp -- per-pair arithmetic of the TholeDipoleForce path for gfx950.
//
// Each function restates one piece of platforms/reference/src/ReferenceTholeDipoleForce.cpp
// (cited per function) for the GPU: the same formulas and branch thresholds, written for
// 64-lane wavefronts where one lane owns one particle i and all lanes see the same j.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tdp {

constexpr int kTile = 64;  // particles per tile = lanes per wavefront

template <typename T> struct Vec4;
template <> struct Vec4<double> { using type = double4; };
template <> struct Vec4<float> { using type = float4; };

__device__ __forceinline__ double rsqrt_t(double x) { return rsqrt(x); }
__device__ __forceinline__ float rsqrt_t(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ double exp_t(double x) { return exp(x); }
__device__ __forceinline__ float exp_t(float x) { return __expf(x); }
__device__ __forceinline__ double floor_t(double x) { return floor(x); }
__device__ __forceinline__ float floor_t(float x) { return floorf(x); }

// getPeriodicDelta (ReferenceTholeDipoleForce.cpp:1414-1423): d -= L*floor(d/L + 0.5) on the box
// diagonal.  d/L is formed as d*(1/L); the two agree except when d/L + 0.5 lies within one
// rounding of an integer, i.e. a pair exactly half a box apart.
template <typename T>
__device__ __forceinline__ void min_image(T& dx, T& dy, T& dz, T bx, T by, T bz, T ibx, T iby, T ibz) {
    dx -= bx * floor_t(dx * ibx + T(0.5));
    dy -= by * floor_t(dy * iby + T(0.5));
    dz -= bz * floor_t(dz * ibz + T(0.5));
}

// getAndScaleInverseRs with two terms (ReferenceTholeDipoleForce.cpp:646-676):
// rr3 = 1/r^3, rr5 = 3/r^5, Thole-damped with damp = a_i*a_j, gamma = min(a_i,a_j),
// s = -gamma*(r/damp)^3 when damp != 0 and s > -50.  r/damp is formed as r*(1/a_i)*(1/a_j).
template <typename T>
__device__ __forceinline__ void damped_rr(T r2, T ai, T aj, T iai, T iaj, T& rr3, T& rr5) {
    T rinv = rsqrt_t(r2);
    T r = r2 * rinv;
    T rinv2 = rinv * rinv;
    rr3 = rinv * rinv2;
    rr5 = T(3) * rr3 * rinv2;
    T damp = ai * aj;
    if (damp != T(0)) {
        T g = ai < aj ? ai : aj;
        T ratio = r * (iai * iaj);
        T s = -g * (ratio * ratio * ratio);
        if (s > T(-50)) {
            T e = exp_t(s);
            rr3 *= T(1) - e;
            rr5 *= T(1) - (T(1) - s) * e;
        }
    }
}

// Inputs of one side of calculateElectrostaticPairIxn.
struct ElecSite {
    double x, y, z;      // raw position (the QI frame's axis choice compares raw y and z)
    double q;
    double dx, dy, dz;   // lab-frame permanent dipole
    double ux, uy, uz;   // induced dipole
    double a;            // Thole parameter
    double ia;           // 1/a
};

struct ElecOut {
    double e;                  // pair energy
    double fx, fy, fz;         // force on K (force on I is the negative)
    double tix, tiy, tiz;      // torque on I
    double tkx, tky, tkz;      // torque on K
};

// Inputs of one side of calculateElectrostaticPairIxn in precision T (positions enter through the
// separation and the frame's axis choice).
template <typename T>
struct ElecSiteT {
    T q;
    T dx, dy, dz;   // lab-frame permanent dipole
    T ux, uy, uz;   // induced dipole
    T a;            // Thole parameter
    T ia;           // 1/a
};

template <typename T>
struct ElecOutT {
    T e;                  // pair energy
    T fx, fy, fz;         // force on K (force on I is the negative)
    T tix, tiy, tiz;      // torque on I
    T tkx, tky, tkz;      // torque on K
};

// calculateElectrostaticPairIxn + formQIRotationMatrix (ReferenceTholeDipoleForce.cpp:129-366,
// :573-607) for the ordered pair (I,K), I < K, with _electric = _dielectric = 1, in precision T.
// yz_differ: the raw positions of I and K differ in y or z (the frame's axis choice, :584-588).
// delta = min-imaged posK - posI.  The QI frame has rows [X, Y, Z] with Z along delta; the
// first rotated component is used as the bond-axis component and forces are rotated back with
// the reference's index pattern (:150-159).  rInvVec[4] (read past a 4-element vector in the
// reference, :196 vs :264-333) is taken as 1/r^4.  Thole factors: when |a_I a_K| <= 1e-5 or
// a*u^3 >= 50 every factor is exactly 1 in FP64, so the exponential is skipped there; DAMP = false
// is the specialisation for pairs known to lie beyond that range (all factors 1).  PERM = false is
// the specialisation for systems without permanent dipoles (every term with one is zero, and so
// are the torques).
template <typename T, bool MUTUAL, bool DAMP, bool PERM = true>
__device__ __forceinline__ void elec_pair_t(const ElecSiteT<T>& I, const ElecSiteT<T>& K, T dx, T dy, T dz,
                                            bool yz_differ, T mS, T iS, ElecOutT<T>& o) {
    T r2 = dx * dx + dy * dy + dz * dz;
    T rInv = rsqrt_t(r2);
    T r = r2 * rInv;
    // formQIRotationMatrix
    T zx = dx * rInv, zy = dy * rInv, zz = dz * rInv;
    T xx = zx, xy = zy, xz = zz;
    if (yz_differ) xx += T(1); else xy += T(1);
    T dt = zx * xx + zy * xy + zz * xz;
    xx -= zx * dt; xy -= zy * dt; xz -= zz * dt;
    T xn2 = xx * xx + xy * xy + xz * xz;
    if (xn2 > T(0)) { T s = rsqrt_t(xn2); xx *= s; xy *= s; xz *= s; }
    T yx = zy * xz - zz * xy;
    T yy = zz * xx - zx * xz;
    T yz = zx * xy - zy * xx;
    // rotated components: index 0 <- X, 1 <- Y, 2 <- Z (rows of the QI matrix)
    T uI0 = xx * I.ux + xy * I.uy + xz * I.uz, uI1 = yx * I.ux + yy * I.uy + yz * I.uz,
           uI2 = zx * I.ux + zy * I.uy + zz * I.uz;
    T uJ0 = xx * K.ux + xy * K.uy + xz * K.uz, uJ1 = yx * K.ux + yy * K.uy + yz * K.uz,
           uJ2 = zx * K.ux + zy * K.uy + zz * K.uz;
    T QI0 = I.q, QJ0 = K.q;
    T QI1 = T(0), QI2 = T(0), QI3 = T(0), QJ1 = T(0), QJ2 = T(0), QJ3 = T(0);
    if (PERM) {
        QI1 = xx * I.dx + xy * I.dy + xz * I.dz;
        QI2 = yx * I.dx + yy * I.dy + yz * I.dz;
        QI3 = zx * I.dx + zy * I.dy + zz * I.dz;
        QJ1 = xx * K.dx + xy * K.dy + xz * K.dz;
        QJ2 = yx * K.dx + yy * K.dy + yz * K.dz;
        QJ3 = zx * K.dx + zy * K.dy + zz * K.dz;
    }

    T rI1 = rInv, rI2 = rI1 * rInv, rI3 = rI2 * rInv, rI4 = rI3 * rInv;

    T dmp = I.a * K.a;
    T am = I.a < K.a ? I.a : K.a;
    T expau3 = T(0.0), au3 = T(0.0);
    if (DAMP && fabs(dmp) > T(1.0e-5)) {
        T u = r * (I.ia * K.ia);
        au3 = am * u * u * u;
        if (au3 < T(50)) expau3 = exp_t(-au3);
    }
    T thole_c = T(1.0) - expau3;
    T thole_d0 = T(1.0) - expau3 * (T(1.0) + T(1.5) * au3);
    T thole_d1 = thole_c;
    T dthole_c = thole_d0;
    T dthole_d0 = T(1.0) - expau3 * (T(1.0) + au3 + T(1.5) * au3 * au3);
    T dthole_d1 = T(1.0) - expau3 * (T(1.0) + au3);

    if constexpr (!PERM) {
        // no permanent dipoles: QI1..3 = QJ1..3 = 0, so only the charge, charge-induced and
        // induced-induced terms remain and every torque is zero (the same sums with the zero terms left out)
        T ePerm = rI1 * mS;
        T dPerm = -T(0.5) * mS * rI2;
        T Vij0 = ePerm * QJ0, Vji0 = ePerm * QI0;
        T VijR0 = dPerm * QJ0, VjiR0 = dPerm * QI0;
        T eU = rI2 * iS * thole_c;
        T dU = -T(2.0) * rI3 * iS * dthole_c;
        Vij0 += -(eU * uJ0);
        VijR0 += -(dU * uJ0);
        T Vjid0 = -(eU * QI0);
        Vji0 += eU * uI0;
        VjiR0 += dU * uI0;
        T Vijd0 = eU * QJ0;
        T energy = T(0.5) * (QI0 * Vij0 + QJ0 * Vji0);
        T fIZ = QI0 * VijR0;
        T fJZ = QJ0 * VjiR0;
        T iEIX = uI2 * Vijd0;
        T iEJX = uJ2 * Vjid0;
        T iEIY = -(uI1 * Vijd0);
        T iEJY = -(uJ1 * Vjid0);
        if (MUTUAL) {
            T eC = -T(4.0) * rI3 * iS * thole_d0;
            T dC = T(6.0) * rI4 * iS * dthole_d0;
            iEIX += eC * uI2 * uJ0;
            iEJX += eC * uJ2 * uI0;
            iEIY -= eC * uI1 * uJ0;
            iEJY -= eC * uJ1 * uI0;
            fIZ += dC * uI0 * uJ0;
            fJZ += dC * uJ0 * uI0;
            eC = T(2.0) * rI3 * iS * thole_d1;
            dC = -T(3.0) * rI4 * iS * dthole_d1;
            iEIX -= eC * uI0 * uJ2;
            iEJX -= eC * uJ0 * uI2;
            iEIY += eC * uI0 * uJ1;
            iEJY += eC * uJ0 * uI1;
            fIZ += dC * (uI1 * uJ1 + uI2 * uJ2);
            fJZ += dC * (uJ1 * uI1 + uJ2 * uI2);
        }
        T qF0 = rInv * (iEIY + iEJY);
        T qF1 = -rInv * (iEIX + iEJX);
        T qF2 = -(fJZ + fIZ);
        o.e = energy;
        o.fx = yy * qF0 + zy * qF1 + xy * qF2;
        o.fy = yz * qF0 + zz * qF1 + xz * qF2;
        o.fz = yx * qF0 + zx * qF1 + xx * qF2;
        o.tix = o.tiy = o.tiz = o.tkx = o.tky = o.tkz = T(0);
    } else {
        // charge-charge (m=0)
        T ePerm = rI1 * mS;
        T dPerm = -T(0.5) * mS * rI2;
        T Vij0 = ePerm * QJ0, Vji0 = ePerm * QI0;
        T VijR0 = dPerm * QJ0, VjiR0 = dPerm * QI0;
        // C-D, C-Uind (m=0)
        ePerm = rI2 * mS;
        T eU = rI2 * iS * thole_c;
        dPerm = -rI3 * mS;
        T dU = -T(2.0) * rI3 * iS * dthole_c;
        Vij0 += -(ePerm * QJ1 + eU * uJ0);
        T Vji1 = -(ePerm * QI0);
        VijR0 += -(dPerm * QJ1 + dU * uJ0);
        T VjiR1 = -(dPerm * QI0);
        T Vjid0 = -(eU * QI0);
        // D-C, Uind-C (m=0)
        T Vij1 = ePerm * QJ0;
        Vji0 += ePerm * QI1 + eU * uI0;
        T VijR1 = dPerm * QJ0;
        VjiR0 += dPerm * QI1 + dU * uI0;
        T Vijd0 = eU * QJ0;
        // D-D, D-Uind (m=0)
        ePerm = -T(2.0) * rI3 * mS;
        eU = -T(2.0) * rI3 * iS * thole_d0;
        dPerm = T(3.0) * rI4 * mS;
        dU = T(6.0) * rI4 * iS * dthole_d0;
        Vij1 += ePerm * QJ1 + eU * uJ0;
        Vji1 += ePerm * QI1 + eU * uI0;
        VijR1 += dPerm * QJ1 + dU * uJ0;
        VjiR1 += dPerm * QI1 + dU * uI0;
        Vijd0 += eU * QJ1;
        Vjid0 += eU * QI1;
        // D-D, D-Uind (m=1)
        ePerm = rI3 * mS;
        eU = rI3 * iS * thole_d1;
        dPerm = -T(1.5) * rI4 * mS;
        dU = -T(3.0) * rI4 * iS * dthole_d1;
        T Vij2 = ePerm * QJ2 + eU * uJ1;
        T Vji2 = ePerm * QI2 + eU * uI1;
        T VijR2 = dPerm * QJ2 + dU * uJ1;
        T VjiR2 = dPerm * QI2 + dU * uI1;
        T Vijd1 = eU * QJ2;
        T Vjid1 = eU * QI2;
        T Vij3 = ePerm * QJ3 + eU * uJ2;
        T Vji3 = ePerm * QI3 + eU * uI2;
        T VijR3 = dPerm * QJ3 + dU * uJ2;
        T VjiR3 = dPerm * QI3 + dU * uI2;
        T Vijd2 = eU * QJ3;
        T Vjid2 = eU * QI3;

        T energy = T(0.5) * (QI0 * Vij0 + QJ0 * Vji0);
        energy += T(0.5) * (QI1 * Vij1 + QJ1 * Vji1);
        energy += T(0.5) * (QI2 * Vij2 + QJ2 * Vji2);
        energy += T(0.5) * (QI3 * Vij3 + QJ3 * Vji3);
        T fIZ = QI0 * VijR0 + QI1 * VijR1 + QI2 * VijR2 + QI3 * VijR3;
        T fJZ = QJ0 * VjiR0 + QJ1 * VjiR1 + QJ2 * VjiR2 + QJ3 * VjiR3;
        // qiQIX = {0, Q3, 0, -Q1}, qiQIY = {0, -Q2, Q1, 0}, qiQIZ = {0, 0, -Q3, Q2}
        T EIX = QI3 * Vij1 - QI1 * Vij3;
        T EIY = -QI2 * Vij1 + QI1 * Vij2;
        T EIZ = -QI3 * Vij2 + QI2 * Vij3;
        T EJX = QJ3 * Vji1 - QJ1 * Vji3;
        T EJY = -QJ2 * Vji1 + QJ1 * Vji2;
        T EJZ = -QJ3 * Vji2 + QJ2 * Vji3;
        T iEIX = uI2 * Vijd0 - uI0 * Vijd2;
        T iEJX = uJ2 * Vjid0 - uJ0 * Vjid2;
        T iEIY = uI0 * Vijd1 - uI1 * Vijd0;
        T iEJY = uJ0 * Vjid1 - uJ1 * Vjid0;
        if (MUTUAL) {
            T eC = -T(4.0) * rI3 * iS * thole_d0;
            T dC = T(6.0) * rI4 * iS * dthole_d0;
            iEIX += eC * uI2 * uJ0;
            iEJX += eC * uJ2 * uI0;
            iEIY -= eC * uI1 * uJ0;
            iEJY -= eC * uJ1 * uI0;
            fIZ += dC * uI0 * uJ0;
            fJZ += dC * uJ0 * uI0;
            eC = T(2.0) * rI3 * iS * thole_d1;
            dC = -T(3.0) * rI4 * iS * dthole_d1;
            iEIX -= eC * uI0 * uJ2;
            iEJX -= eC * uJ0 * uI2;
            iEIY += eC * uI0 * uJ1;
            iEJY += eC * uJ0 * uI1;
            fIZ += dC * (uI1 * uJ1 + uI2 * uJ2);
            fJZ += dC * (uJ1 * uI1 + uJ2 * uI2);
        }
        T qF0 = rInv * (EIY + EJY + iEIY + iEJY);
        T qF1 = -rInv * (EIX + EJX + iEIX + iEJX);
        T qF2 = -(fJZ + fIZ);
        // forceRotationMatrix rows: (Y_y, Z_y, X_y), (Y_z, Z_z, X_z), (Y_x, Z_x, X_x)
        o.e = energy;
        o.fx = yy * qF0 + zy * qF1 + xy * qF2;
        o.fy = yz * qF0 + zz * qF1 + xz * qF2;
        o.fz = yx * qF0 + zx * qF1 + xx * qF2;
        o.tix = -(yy * EIX + zy * EIY + xy * EIZ);
        o.tiy = -(yz * EIX + zz * EIY + xz * EIZ);
        o.tiz = -(yx * EIX + zx * EIY + xx * EIZ);
        o.tkx = -(yy * EJX + zy * EJY + xy * EJZ);
        o.tky = -(yz * EJX + zz * EJY + xz * EJZ);
        o.tkz = -(yx * EJX + zx * EJY + xx * EJZ);
    }
}


// FP64 form on full sites (the frame's axis choice compares the raw y and z).
template <bool MUTUAL, bool DAMP = true, bool PERM = true>
__device__ __forceinline__ void elec_pair(const ElecSite& I, const ElecSite& K, double dx, double dy, double dz,
                                          double mS, double iS, ElecOut& o) {
    const ElecSiteT<double> a = {I.q, I.dx, I.dy, I.dz, I.ux, I.uy, I.uz, I.a, I.ia};
    const ElecSiteT<double> b = {K.q, K.dx, K.dy, K.dz, K.ux, K.uy, K.uz, K.a, K.ia};
    ElecOutT<double> r;
    elec_pair_t<double, MUTUAL, DAMP, PERM>(a, b, dx, dy, dz, I.y != K.y || I.z != K.z, mS, iS, r);
    o.e = r.e;
    o.fx = r.fx; o.fy = r.fy; o.fz = r.fz;
    o.tix = r.tix; o.tiy = r.tiy; o.tiz = r.tiz;
    o.tkx = r.tkx; o.tky = r.tky; o.tkz = r.tkz;
}

__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

}  // namespace tdp
