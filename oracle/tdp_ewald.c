/*
 * tdp_ewald.c -- FP64 CPU oracle of the Ewald ("P-mode") electrostatics of the HIP library:
 * point charges and induced dipoles with Thole damping, summed over all periodic images by Ewald
 * splitting (real space within the cutoff; reciprocal space by smooth PME or, for pinning, by the
 * exact Ewald sum over k vectors).
 *
 * TEST INFRASTRUCTURE ONLY (see tdp_oracle.h): the checker of the HIP library's Ewald path.
 *
 * The reference has no Ewald sum (its "PME" method is the minimum image over all pairs,
 * ReferenceTholeDipoleForce.cpp:1375-1423); this restates published methods:
 *   - Ewald splitting of charge and dipole interactions with the screened radial functions
 *     B_0 = erfc(a r)/r, B_n = ((2n-1) B_{n-1} + (2a^2)^n/(a sqrt(pi)) e^{-a^2 r^2}) / r^2
 *     (Smith 1998, "Point multipoles in the Ewald summation"; Sagui, Pedersen & Darden 2004);
 *   - smooth PME with cardinal B-splines of order p and the Euler exponential-spline moduli
 *     (Essmann et al. 1995, J. Chem. Phys. 103, 8577), dipoles spread as derivatives of the
 *     B-splines (Sagui, Pedersen & Darden 2004);
 *   - self terms -a/sqrt(pi) sum q^2 (energy) and +4a^3/(3 sqrt(pi)) mu_i (field), and the
 *     neutralising-background energy -pi Q^2 / (2 V a^2) for a net charge Q;
 *   - Thole damping of the reference (getAndScaleInverseRs, ReferenceTholeDipoleForce.cpp:646-676):
 *     s = -min(a_i,a_j) (r/(a_i a_j))^3, lambda3 = 1 - e^s, lambda5 = 1 - (1-s) e^s and, for the
 *     gradient of the dipole field, lambda7 = 1 - (1 - s + 3/5 s^2) e^s (the next member of the
 *     same family: d/dr(lambda_{2n+1}/r^{2n+1}) = -(2n+1) lambda_{2n+3}/r^{2n+2}).
 *
 * Model (the library's Ewald option, documented with it):
 *   E0_i   = field at i of all charges, pair factor mScale * lambda3 (reference :678-733);
 *   (T mu)_i = field at i of all induced dipoles, pair factor iScale * (lambda3, lambda5) (:766-818),
 *            including the Ewald self-field;
 *   mu solves (alpha^-1 - T) mu = E0 (Mutual, PCG with the stopping rule RMS(alpha r) <= target) or
 *   mu = alpha E0 (Direct);
 *   energy  U = U_qq - 1/2 sum_i mu_i . E0_i, U_qq the Ewald charge-charge energy with pair factor
 *            mScale (no damping);
 *   forces  F = -dU/dr at fixed mu: charge-charge, charge-dipole and (Mutual) dipole-dipole terms,
 *            which is the exact gradient of U for both polarization types (U is stationary in mu).
 * A pair contributes in real space if its minimum-image distance is below the cutoff; a pair whose
 * scale differs from 1 also contributes the correction (scale*lambda - 1) * (bare kernel) beyond
 * the cutoff, so that the scale is applied to the whole interaction.
 */
#define _DEFAULT_SOURCE   /* M_PI */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "tdp_oracle.h"

#define TDPO_MAX_ORDER 8

/* ------------------------------------------------------------------ FFT */
/* Mixed-radix decimation-in-time DFT of length n (any n; prime factors beyond 5 by the direct
 * sum): out[k] = sum_j in[j*stride] exp(sign 2 pi i j k / n), unnormalised. */
static int smallest_factor(int n) {
    if (n % 4 == 0) return 4;
    for (int p = 2; p * p <= n; p++)
        if (n % p == 0) return p;
    return n;
}

static void dft_rec(const double complex* in, int stride, double complex* out, int n, int sign,
                    double complex* tmp) {
    if (n == 1) {
        out[0] = in[0];
        return;
    }
    const int p = smallest_factor(n), m = n / p;
    for (int r = 0; r < p; r++) dft_rec(in + (size_t)r * stride, stride * p, out + (size_t)r * m, m, sign, tmp);
    /* out[r*m + k] holds sub-transform r at frequency k; combine into tmp then copy back */
    const double w0 = sign * 2.0 * M_PI / n;
    for (int k = 0; k < m; k++) {
        for (int q = 0; q < p; q++) {
            const int f = k + q * m;
            double complex acc = 0.0;
            for (int r = 0; r < p; r++) {
                const long long e = ((long long)r * f) % n;
                acc += out[(size_t)r * m + k] * cexp(I * (w0 * (double)e));
            }
            tmp[f] = acc;
        }
    }
    memcpy(out, tmp, sizeof(double complex) * (size_t)n);
}

static void fft_line(double complex* data, int n, int stride, int sign, double complex* buf) {
    /* buf: 3n scratch */
    double complex* in = buf;
    double complex* out = buf + n;
    double complex* tmp = buf + 2 * n;
    for (int j = 0; j < n; j++) in[j] = data[(size_t)j * stride];
    dft_rec(in, 1, out, n, sign, tmp);
    for (int j = 0; j < n; j++) data[(size_t)j * stride] = out[j];
}

/* 3D unnormalised DFT of a complex grid g[(x*ny + y)*nz + z] (x slowest). */
static void fft3(double complex* g, int nx, int ny, int nz, int sign) {
    int nmax = nx > ny ? nx : ny;
    if (nz > nmax) nmax = nz;
    double complex* buf = (double complex*)malloc(sizeof(double complex) * 3 * (size_t)nmax);
    for (int x = 0; x < nx; x++)
        for (int y = 0; y < ny; y++) fft_line(g + ((size_t)x * ny + y) * nz, nz, 1, sign, buf);
    for (int x = 0; x < nx; x++)
        for (int z = 0; z < nz; z++) fft_line(g + (size_t)x * ny * nz + z, ny, nz, sign, buf);
    for (int y = 0; y < ny; y++)
        for (int z = 0; z < nz; z++) fft_line(g + (size_t)y * nz + z, nx, ny * nz, sign, buf);
    free(buf);
}

void tdpo_fft3(double* re_im, int nx, int ny, int nz, int sign) {
    fft3((double complex*)re_im, nx, ny, nz, sign >= 0 ? 1 : -1);
}

/* ------------------------------------------------------------------ B-splines */
/* th[j] = M_p(w + p-1-j), j = 0..p-1 (the weight of grid point floor(u) - (p-1) + j for the
 * fractional grid coordinate u, w = u - floor(u)), and its first and second derivatives with
 * respect to u.  M_n(x) = (x M_{n-1}(x) + (n - x) M_{n-1}(x-1)) / (n-1), M_1 = 1 on [0,1);
 * M_n' = M_{n-1}(x) - M_{n-1}(x-1). */
static void bspline_level(double w, int n, const double* prev, double* cur) {
    for (int j = 0; j < n; j++) {
        const double a = j >= 1 ? prev[j - 1] : 0.0;   /* M_{n-1}(w + n-1-j) */
        const double b = j <= n - 2 ? prev[j] : 0.0;    /* M_{n-1}(w + n-2-j) */
        cur[j] = ((w + n - 1 - j) * a + (j + 1 - w) * b) / (n - 1);
    }
}

void tdpo_bspline(double w, int p, double* th, double* d1, double* d2) {
    double lv[TDPO_MAX_ORDER + 1][TDPO_MAX_ORDER];
    lv[1][0] = 1.0;
    for (int n = 2; n <= p; n++) bspline_level(w, n, lv[n - 1], lv[n]);
    for (int j = 0; j < p; j++) th[j] = lv[p][j];
    if (d1) {
        const double* m1 = lv[p - 1];
        for (int j = 0; j < p; j++) d1[j] = (j >= 1 ? m1[j - 1] : 0.0) - (j <= p - 2 ? m1[j] : 0.0);
    }
    if (d2) {
        const double* m2 = lv[p - 2];
        for (int j = 0; j < p; j++) {
            const double a = (j >= 2 && j - 2 <= p - 3) ? m2[j - 2] : 0.0;
            const double b = (j >= 1 && j - 1 <= p - 3) ? m2[j - 1] : 0.0;
            const double c = (j <= p - 3) ? m2[j] : 0.0;
            d2[j] = a - 2.0 * b + c;
        }
    }
}

/* |b(m)|^2 of the Euler exponential spline for one axis (Essmann et al. eq. 4.4):
 * 1 / |sum_{k=0}^{p-2} M_p(k+1) e^{2 pi i m k / K}|^2; where the denominator vanishes (odd p at
 * m = K/2) the value is the mean of its neighbours, as in the original smooth PME code. */
static void bspline_moduli(int K, int p, double* mod) {
    double th[TDPO_MAX_ORDER];
    tdpo_bspline(0.0, p, th, NULL, NULL);   /* th[j] = M_p(p-1-j): M_p(k+1) = th[p-2-k] */
    for (int m = 0; m < K; m++) {
        double re = 0.0, im = 0.0;
        for (int k = 0; k <= p - 2; k++) {
            const double v = th[p - 2 - k];
            re += v * cos(2.0 * M_PI * m * k / K);
            im += v * sin(2.0 * M_PI * m * k / K);
        }
        mod[m] = re * re + im * im;
    }
    for (int m = 0; m < K; m++)
        if (mod[m] < 1e-7) mod[m] = 0.5 * (mod[(m - 1 + K) % K] + mod[(m + 1) % K]);
    for (int m = 0; m < K; m++) mod[m] = 1.0 / mod[m];
}

/* ------------------------------------------------------------------ context */
typedef struct {
    const tdpo_system* s;
    const tdpo_ewald* e;
    int n;
    double L[3], V, a, rc;
    const double* pos;
    /* scaled pairs (i < j) with their scales, CSR by i and by j */
    int nsc;
    int *sc_i, *sc_j;
    double *sc_m, *sc_is;
    int *row_off, *row_idx;   /* for particle i: scaled pair ids with i as either member */
    /* SPME */
    int K[3], p;
    double* G;   /* influence function [K0*K1*K2] */
} ewctx;

static void min_image(const ewctx* c, double* d) {
    for (int k = 0; k < 3; k++) d[k] -= c->L[k] * floor(d[k] / c->L[k] + 0.5);
}

static int ew_setup(ewctx* c, const tdpo_system* s, const tdpo_ewald* e, const double* pos) {
    memset(c, 0, sizeof *c);
    c->s = s;
    c->e = e;
    c->n = s->n;
    c->pos = pos;
    for (int k = 0; k < 3; k++) c->L[k] = s->box[k];
    if (s->method != TDPO_PME || !(c->L[0] > 0) || !(c->L[1] > 0) || !(c->L[2] > 0)) return -3;
    c->V = c->L[0] * c->L[1] * c->L[2];
    c->a = e->alpha;
    c->rc = e->cutoff;
    if (!(c->a > 0) || !(c->rc > 0)) return -3;
    for (int k = 0; k < 3; k++)
        if (c->rc > 0.5 * c->L[k]) return -3;
    if (s->polarization == TDPO_EXTRAPOLATED) return -3;
    for (int i = 0; i < 3 * s->n; i++)
        if (s->dipole[i] != 0.0) return -3;   /* charges and induced dipoles only */
    for (int i = 0; i < s->n; i++) {
        for (int t = 0; t < 2; t++)
            if (s->cov_offsets[4 * i + t + 1] - s->cov_offsets[4 * i + t] > 5) return -2;
    }
    /* scaled pairs: every (i, a) with a > i in i's 1-2 or 1-3 list whose scales differ from 1 */
    int cap = 16;
    c->sc_i = (int*)malloc(sizeof(int) * cap);
    c->sc_j = (int*)malloc(sizeof(int) * cap);
    c->sc_m = (double*)malloc(sizeof(double) * cap);
    c->sc_is = (double*)malloc(sizeof(double) * cap);
    for (int i = 0; i < s->n; i++) {
        for (int t = 0; t < 2; t++)
            for (int q = s->cov_offsets[4 * i + t]; q < s->cov_offsets[4 * i + t + 1]; q++) {
                const int j = s->cov_atoms[q];
                if (j <= i || j >= s->n) continue;
                int dup = 0;
                for (int k = 0; k < c->nsc; k++)
                    if (c->sc_i[k] == i && c->sc_j[k] == j) dup = 1;
                if (dup) continue;
                double m, is;
                tdpo_pair_scales(s, i, j, &m, &is);
                if (m == 1.0 && is == 1.0) continue;
                if (c->nsc == cap) {
                    cap *= 2;
                    c->sc_i = (int*)realloc(c->sc_i, sizeof(int) * cap);
                    c->sc_j = (int*)realloc(c->sc_j, sizeof(int) * cap);
                    c->sc_m = (double*)realloc(c->sc_m, sizeof(double) * cap);
                    c->sc_is = (double*)realloc(c->sc_is, sizeof(double) * cap);
                }
                c->sc_i[c->nsc] = i;
                c->sc_j[c->nsc] = j;
                c->sc_m[c->nsc] = m;
                c->sc_is[c->nsc] = is;
                c->nsc++;
            }
    }
    c->row_off = (int*)calloc((size_t)s->n + 1, sizeof(int));
    c->row_idx = (int*)malloc(sizeof(int) * (2 * (size_t)c->nsc + 1));
    for (int k = 0; k < c->nsc; k++) {
        c->row_off[c->sc_i[k] + 1]++;
        c->row_off[c->sc_j[k] + 1]++;
    }
    for (int i = 0; i < s->n; i++) c->row_off[i + 1] += c->row_off[i];
    int* fill = (int*)calloc((size_t)s->n, sizeof(int));
    for (int k = 0; k < c->nsc; k++) {
        c->row_idx[c->row_off[c->sc_i[k]] + fill[c->sc_i[k]]++] = k;
        c->row_idx[c->row_off[c->sc_j[k]] + fill[c->sc_j[k]]++] = k;
    }
    free(fill);
    /* reciprocal space */
    if (e->kmax[0] <= 0) {
        c->p = e->order;
        if (c->p < 3 || c->p > TDPO_MAX_ORDER) return -3;
        for (int k = 0; k < 3; k++) {
            c->K[k] = e->grid[k];
            if (c->K[k] < c->p) return -3;
        }
        double* mod[3];
        for (int k = 0; k < 3; k++) {
            mod[k] = (double*)malloc(sizeof(double) * c->K[k]);
            bspline_moduli(c->K[k], c->p, mod[k]);
        }
        const size_t ng = (size_t)c->K[0] * c->K[1] * c->K[2];
        c->G = (double*)malloc(sizeof(double) * ng);
        const double f = M_PI * M_PI / (c->a * c->a);
        for (int x = 0; x < c->K[0]; x++)
            for (int y = 0; y < c->K[1]; y++)
                for (int z = 0; z < c->K[2]; z++) {
                    const int mx = x <= c->K[0] / 2 ? x : x - c->K[0];
                    const int my = y <= c->K[1] / 2 ? y : y - c->K[1];
                    const int mz = z <= c->K[2] / 2 ? z : z - c->K[2];
                    const double kx = mx / c->L[0], ky = my / c->L[1], kz = mz / c->L[2];
                    const double m2 = kx * kx + ky * ky + kz * kz;
                    const size_t g = ((size_t)x * c->K[1] + y) * c->K[2] + z;
                    c->G[g] = (x == 0 && y == 0 && z == 0)
                                  ? 0.0
                                  : exp(-f * m2) / (M_PI * c->V * m2) * mod[0][x] * mod[1][y] * mod[2][z];
                }
        for (int k = 0; k < 3; k++) free(mod[k]);
    }
    return 0;
}

static void ew_free(ewctx* c) {
    free(c->sc_i); free(c->sc_j); free(c->sc_m); free(c->sc_is);
    free(c->row_off); free(c->row_idx); free(c->G);
}

/* scales of the pair (i, j): 1, 1 unless listed */
static void pair_scale(const ewctx* c, int i, int j, double* m, double* is, int* listed) {
    *m = 1.0;
    *is = 1.0;
    *listed = 0;
    const int lo = i < j ? i : j, hi = i < j ? j : i;
    for (int q = c->row_off[lo]; q < c->row_off[lo + 1]; q++) {
        const int k = c->row_idx[q];
        if (c->sc_i[k] == lo && c->sc_j[k] == hi) {
            *m = c->sc_m[k];
            *is = c->sc_is[k];
            *listed = 1;
            return;
        }
    }
}

/* ------------------------------------------------------------------ real space */
/* Radial coefficients of one pair at distance r: qq[0..1] (charge-charge, scale m, no damping),
 * qm[1..2] (charge-dipole, scale m, lambda3/lambda5), mm[1..3] (dipole-dipole, scale is,
 * lambda3/5/7).  Index n multiplies the n-th kernel: 1/r, 1/r^3, 3/r^5, 15/r^7. */
typedef struct { double qq[2], qm[3], mm[4]; } pair_coef;

static void real_coeffs(const ewctx* c, double r, int inside, double ai, double aj, double m, double is,
                        pair_coef* pc) {
    const double a = c->a, ri = 1.0 / r, r2i = ri * ri;
    double B[4] = {0, 0, 0, 0};
    if (inside) {
        const double ex = exp(-a * a * r * r), sp = 1.0 / sqrt(M_PI);
        B[0] = erfc(a * r) * ri;
        double fac = 2.0 * a * a;   /* (2a^2)^n / (a sqrt(pi)) * e^{-a^2 r^2} with n = 1.. */
        double pw = fac;
        for (int n = 1; n <= 3; n++) {
            B[n] = ((2 * n - 1) * B[n - 1] + pw / a * sp * ex) * r2i;
            pw *= fac;
        }
    }
    const double b[4] = {ri, ri * r2i, 3.0 * ri * r2i * r2i, 15.0 * ri * r2i * r2i * r2i};
    double l3 = 1.0, l5 = 1.0, l7 = 1.0;
    const double damp = ai * aj;
    if (damp != 0.0) {
        const double g = ai < aj ? ai : aj;
        const double u = r / damp;
        const double s = -g * u * u * u;
        if (s > -50.0) {
            const double es = exp(s);
            l3 = 1.0 - es;
            l5 = 1.0 - (1.0 - s) * es;
            l7 = 1.0 - (1.0 - s + 0.6 * s * s) * es;
        }
    }
    const double lam[4] = {1.0, l3, l5, l7};
    const int cm = inside || m != 1.0, ci = inside || is != 1.0;
    for (int n = 0; n < 2; n++) pc->qq[n] = B[n] + (cm ? (m - 1.0) * b[n] : 0.0);
    pc->qm[0] = 0.0;
    for (int n = 1; n < 3; n++) pc->qm[n] = B[n] + (cm ? (m * lam[n] - 1.0) * b[n] : 0.0);
    pc->mm[0] = 0.0;
    for (int n = 1; n < 4; n++) pc->mm[n] = B[n] + (ci ? (is * lam[n] - 1.0) * b[n] : 0.0);
}

/* What a real-space pass accumulates: fields E0 (charges), T mu (dipoles), charge-charge energy,
 * forces. */
typedef struct {
    const double* mu;   /* NULL: no dipole terms */
    double* e0;         /* [3n] field of the charges */
    double* tmu;        /* [3n] field of the dipoles */
    double* f;          /* [3n] forces */
    double* uqq;        /* charge-charge energy */
    int mutual;         /* dipole-dipole forces */
} real_out;

static void real_pair(const ewctx* c, int i, int j, const double* d, double r2, real_out* o, double* e0,
                      double* tmu, double* f, double* uqq) {
    const tdpo_system* s = c->s;
    double m, is;
    int listed;
    pair_scale(c, i, j, &m, &is, &listed);
    const int inside = r2 < c->rc * c->rc;
    if (!inside && !listed) return;
    pair_coef pc;
    real_coeffs(c, sqrt(r2), inside, s->thole[i], s->thole[j], m, is, &pc);
    const double qi = s->charge[i], qj = s->charge[j];
    if (e0)
        for (int k = 0; k < 3; k++) {
            e0[3 * i + k] += qj * pc.qm[1] * d[k];
            e0[3 * j + k] -= qi * pc.qm[1] * d[k];
        }
    const double* mi = o->mu ? &o->mu[3 * i] : NULL;
    const double* mj = o->mu ? &o->mu[3 * j] : NULL;
    if (tmu && o->mu) {
        const double mjd = mj[0] * d[0] + mj[1] * d[1] + mj[2] * d[2];
        const double mid = mi[0] * d[0] + mi[1] * d[1] + mi[2] * d[2];
        for (int k = 0; k < 3; k++) {
            tmu[3 * i + k] += pc.mm[2] * mjd * d[k] - pc.mm[1] * mj[k];
            tmu[3 * j + k] += pc.mm[2] * mid * d[k] - pc.mm[1] * mi[k];
        }
    }
    if (uqq) *uqq += qi * qj * pc.qq[0];
    if (f) {
        double g[3];   /* grad_d U, d = r_i - r_j; F_i = -g, F_j = +g */
        for (int k = 0; k < 3; k++) g[k] = -qi * qj * pc.qq[1] * d[k];
        if (o->mu) {
            double v[3];
            for (int k = 0; k < 3; k++) v[k] = qi * mj[k] - qj * mi[k];
            const double vd = v[0] * d[0] + v[1] * d[1] + v[2] * d[2];
            for (int k = 0; k < 3; k++) g[k] += pc.qm[1] * v[k] - pc.qm[2] * vd * d[k];
            if (o->mutual) {
                const double mid = mi[0] * d[0] + mi[1] * d[1] + mi[2] * d[2];
                const double mjd = mj[0] * d[0] + mj[1] * d[1] + mj[2] * d[2];
                const double mm = mi[0] * mj[0] + mi[1] * mj[1] + mi[2] * mj[2];
                for (int k = 0; k < 3; k++)
                    g[k] += -pc.mm[2] * mm * d[k] + pc.mm[3] * mid * mjd * d[k] -
                            pc.mm[2] * (mi[k] * mjd + mj[k] * mid);
            }
        }
        for (int k = 0; k < 3; k++) {
            f[3 * i + k] -= g[k];
            f[3 * j + k] += g[k];
        }
    }
}

/* Every pair within the cutoff (cell list when at least three cells fit per axis, else all pairs),
 * then the scaled pairs beyond it. */
static void real_space(const ewctx* c, real_out* o) {
    const int n = c->n;
    int nc[3], use_cells = 1;
    for (int k = 0; k < 3; k++) {
        nc[k] = (int)floor(c->L[k] / c->rc);
        if (nc[k] < 3) use_cells = 0;
    }
    int* cell = NULL;
    int* cstart = NULL;
    int* corder = NULL;
    if (use_cells) {
        const int ncell = nc[0] * nc[1] * nc[2];
        cell = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
        cstart = (int*)calloc((size_t)ncell + 1, sizeof(int));
        corder = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
        for (int i = 0; i < n; i++) {
            int ci[3];
            for (int k = 0; k < 3; k++) {
                double x = c->pos[3 * i + k];
                x -= c->L[k] * floor(x / c->L[k]);
                ci[k] = (int)(x / c->L[k] * nc[k]);
                if (ci[k] >= nc[k]) ci[k] = nc[k] - 1;
                if (ci[k] < 0) ci[k] = 0;
            }
            cell[i] = (ci[0] * nc[1] + ci[1]) * nc[2] + ci[2];
            cstart[cell[i] + 1]++;
        }
        for (int k = 0; k < ncell; k++) cstart[k + 1] += cstart[k];
        int* fill = (int*)calloc((size_t)ncell, sizeof(int));
        for (int i = 0; i < n; i++) corder[cstart[cell[i]] + fill[cell[i]]++] = i;
        free(fill);
    }
    int nt = 1;
#ifdef _OPENMP
    nt = tdpo_get_threads();
#endif
    const size_t N3 = 3 * (size_t)(n > 0 ? n : 1);
    double* acc = (double*)calloc((size_t)nt * 3 * N3, sizeof(double));
    double* eacc = (double*)calloc((size_t)nt, sizeof(double));
#pragma omp parallel num_threads(nt)
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        double* e0 = o->e0 ? acc + (size_t)t * 3 * N3 : NULL;
        double* tmu = o->tmu ? acc + (size_t)t * 3 * N3 + N3 : NULL;
        double* f = o->f ? acc + (size_t)t * 3 * N3 + 2 * N3 : NULL;
        double uqq = 0.0;
#pragma omp for schedule(dynamic, 16)
        for (int i = 0; i < n; i++) {
            if (use_cells) {
                const int cx = cell[i] / (nc[1] * nc[2]), cy = (cell[i] / nc[2]) % nc[1], cz = cell[i] % nc[2];
                for (int ox = -1; ox <= 1; ox++)
                    for (int oy = -1; oy <= 1; oy++)
                        for (int oz = -1; oz <= 1; oz++) {
                            const int nb = (((cx + ox + nc[0]) % nc[0]) * nc[1] + (cy + oy + nc[1]) % nc[1]) * nc[2] +
                                           (cz + oz + nc[2]) % nc[2];
                            for (int q = cstart[nb]; q < cstart[nb + 1]; q++) {
                                const int j = corder[q];
                                if (j <= i) continue;
                                double d[3];
                                for (int k = 0; k < 3; k++) d[k] = c->pos[3 * i + k] - c->pos[3 * j + k];
                                min_image(c, d);
                                const double r2 = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2];
                                if (r2 < c->rc * c->rc) real_pair(c, i, j, d, r2, o, e0, tmu, f, o->uqq ? &uqq : NULL);
                            }
                        }
            } else {
                for (int j = i + 1; j < n; j++) {
                    double d[3];
                    for (int k = 0; k < 3; k++) d[k] = c->pos[3 * i + k] - c->pos[3 * j + k];
                    min_image(c, d);
                    const double r2 = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2];
                    if (r2 < c->rc * c->rc) real_pair(c, i, j, d, r2, o, e0, tmu, f, o->uqq ? &uqq : NULL);
                }
            }
        }
        eacc[t] = uqq;
    }
    /* scaled pairs beyond the cutoff: their corrections only (serial) */
    double uqq = 0.0;
    double* e0 = o->e0 ? acc : NULL;
    double* tmu = o->tmu ? acc + N3 : NULL;
    double* f = o->f ? acc + 2 * N3 : NULL;
    for (int k = 0; k < c->nsc; k++) {
        const int i = c->sc_i[k], j = c->sc_j[k];
        double d[3];
        for (int q = 0; q < 3; q++) d[q] = c->pos[3 * i + q] - c->pos[3 * j + q];
        min_image(c, d);
        const double r2 = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2];
        if (!(r2 < c->rc * c->rc)) real_pair(c, i, j, d, r2, o, e0, tmu, f, o->uqq ? &uqq : NULL);
    }
    for (int t = 0; t < nt; t++) {
        const double* a = acc + (size_t)t * 3 * N3;
        for (size_t q = 0; q < 3 * (size_t)n; q++) {
            if (o->e0) o->e0[q] += a[q];
            if (o->tmu) o->tmu[q] += a[N3 + q];
            if (o->f) o->f[q] += a[2 * N3 + q];
        }
        uqq += eacc[t];
    }
    if (o->uqq) *o->uqq += uqq;
    free(acc);
    free(eacc);
    free(cell);
    free(cstart);
    free(corder);
}

/* ------------------------------------------------------------------ reciprocal space */
/* Potential, field and field gradient at every particle from the reciprocal-space sum of the
 * density of the charges q (may be NULL) and dipoles mu (may be NULL); energy = 1/2 sum rho phi.
 * phi [n], E [3n], dE [6n] (xx, yy, zz, xy, xz, yz), each may be NULL. */
static void recip_spme(const ewctx* c, const double* q, const double* mu, double* phi, double* E, double* dE,
                       double* energy) {
    const int n = c->n, p = c->p;
    const int K0 = c->K[0], K1 = c->K[1], K2 = c->K[2];
    const size_t ng = (size_t)K0 * K1 * K2;
    double complex* grid = (double complex*)calloc(ng, sizeof(double complex));
    double* Q = (double*)calloc(ng, sizeof(double));
    for (int i = 0; i < n; i++) {
        double th[3][TDPO_MAX_ORDER], d1[3][TDPO_MAX_ORDER];
        int base[3];
        for (int k = 0; k < 3; k++) {
            double fr = c->pos[3 * i + k] / c->L[k];
            fr -= floor(fr);
            double u = fr * c->K[k];
            int iu = (int)floor(u);
            double w = u - iu;
            if (iu >= c->K[k]) iu -= c->K[k];
            base[k] = iu - (p - 1);
            tdpo_bspline(w, p, th[k], d1[k], NULL);
        }
        const double qi = q ? q[i] : 0.0;
        const double mx = mu ? mu[3 * i] * c->K[0] / c->L[0] : 0.0;
        const double my = mu ? mu[3 * i + 1] * c->K[1] / c->L[1] : 0.0;
        const double mz = mu ? mu[3 * i + 2] * c->K[2] / c->L[2] : 0.0;
        for (int a = 0; a < p; a++) {
            const int gx = (base[0] + a + K0) % K0;
            for (int b = 0; b < p; b++) {
                const int gy = (base[1] + b + K1) % K1;
                for (int cc = 0; cc < p; cc++) {
                    const int gz = (base[2] + cc + K2) % K2;
                    const double v = qi * th[0][a] * th[1][b] * th[2][cc] + mx * d1[0][a] * th[1][b] * th[2][cc] +
                                     my * th[0][a] * d1[1][b] * th[2][cc] + mz * th[0][a] * th[1][b] * d1[2][cc];
                    Q[((size_t)gx * K1 + gy) * K2 + gz] += v;
                }
            }
        }
    }
    for (size_t g = 0; g < ng; g++) grid[g] = Q[g];
    fft3(grid, K0, K1, K2, -1);
    for (size_t g = 0; g < ng; g++) grid[g] *= c->G[g];
    fft3(grid, K0, K1, K2, +1);
    if (energy) {
        double en = 0.0;
        for (size_t g = 0; g < ng; g++) en += Q[g] * creal(grid[g]);
        *energy = 0.5 * en;
    }
    for (int i = 0; i < n; i++) {
        double th[3][TDPO_MAX_ORDER], d1[3][TDPO_MAX_ORDER], d2[3][TDPO_MAX_ORDER];
        int base[3];
        double sc[3];
        for (int k = 0; k < 3; k++) {
            double fr = c->pos[3 * i + k] / c->L[k];
            fr -= floor(fr);
            double u = fr * c->K[k];
            int iu = (int)floor(u);
            double w = u - iu;
            if (iu >= c->K[k]) iu -= c->K[k];
            base[k] = iu - (p - 1);
            tdpo_bspline(w, p, th[k], d1[k], d2[k]);
            sc[k] = c->K[k] / c->L[k];
        }
        double ph = 0, gx_ = 0, gy_ = 0, gz_ = 0, hxx = 0, hyy = 0, hzz = 0, hxy = 0, hxz = 0, hyz = 0;
        for (int a = 0; a < p; a++) {
            const int gx = (base[0] + a + K0) % K0;
            for (int b = 0; b < p; b++) {
                const int gy = (base[1] + b + K1) % K1;
                for (int cc = 0; cc < p; cc++) {
                    const int gz = (base[2] + cc + K2) % K2;
                    const double v = creal(grid[((size_t)gx * K1 + gy) * K2 + gz]);
                    ph += v * th[0][a] * th[1][b] * th[2][cc];
                    gx_ += v * d1[0][a] * th[1][b] * th[2][cc];
                    gy_ += v * th[0][a] * d1[1][b] * th[2][cc];
                    gz_ += v * th[0][a] * th[1][b] * d1[2][cc];
                    hxx += v * d2[0][a] * th[1][b] * th[2][cc];
                    hyy += v * th[0][a] * d2[1][b] * th[2][cc];
                    hzz += v * th[0][a] * th[1][b] * d2[2][cc];
                    hxy += v * d1[0][a] * d1[1][b] * th[2][cc];
                    hxz += v * d1[0][a] * th[1][b] * d1[2][cc];
                    hyz += v * th[0][a] * d1[1][b] * d1[2][cc];
                }
            }
        }
        if (phi) phi[i] = ph;
        if (E) {
            E[3 * i] = -gx_ * sc[0];
            E[3 * i + 1] = -gy_ * sc[1];
            E[3 * i + 2] = -gz_ * sc[2];
        }
        if (dE) {
            dE[6 * i + 0] = -hxx * sc[0] * sc[0];
            dE[6 * i + 1] = -hyy * sc[1] * sc[1];
            dE[6 * i + 2] = -hzz * sc[2] * sc[2];
            dE[6 * i + 3] = -hxy * sc[0] * sc[1];
            dE[6 * i + 4] = -hxz * sc[0] * sc[2];
            dE[6 * i + 5] = -hyz * sc[1] * sc[2];
        }
    }
    free(grid);
    free(Q);
}

/* The exact Ewald reciprocal sum over k = 2 pi (nx/Lx, ny/Ly, nz/Lz), |n_d| <= kmax[d], n != 0:
 * phi(r) = (4 pi / V) sum_k g(k) Re[S(k) e^{-i k.r}], g = e^{-k^2/4a^2}/k^2,
 * S(k) = sum_j (q_j + i k.mu_j) e^{i k.r_j}. */
static void recip_exact(const ewctx* c, const double* q, const double* mu, double* phi, double* E, double* dE,
                        double* energy) {
    const int n = c->n;
    const int kx = c->e->kmax[0], ky = c->e->kmax[1], kz = c->e->kmax[2];
    if (phi) memset(phi, 0, sizeof(double) * n);
    if (E) memset(E, 0, sizeof(double) * 3 * n);
    if (dE) memset(dE, 0, sizeof(double) * 6 * n);
    double en = 0.0;
    double complex* ph = (double complex*)malloc(sizeof(double complex) * (size_t)(n > 0 ? n : 1));
    for (int a = -kx; a <= kx; a++)
        for (int b = -ky; b <= ky; b++)
            for (int cc = -kz; cc <= kz; cc++) {
                if (a == 0 && b == 0 && cc == 0) continue;
                const double k[3] = {2 * M_PI * a / c->L[0], 2 * M_PI * b / c->L[1], 2 * M_PI * cc / c->L[2]};
                const double k2 = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
                const double g = 4.0 * M_PI / c->V * exp(-k2 / (4 * c->a * c->a)) / k2;
                double complex S = 0.0;
                for (int j = 0; j < n; j++) {
                    const double kr = k[0] * c->pos[3 * j] + k[1] * c->pos[3 * j + 1] + k[2] * c->pos[3 * j + 2];
                    ph[j] = cexp(I * kr);
                    double complex w = q ? q[j] : 0.0;
                    if (mu) w += I * (k[0] * mu[3 * j] + k[1] * mu[3 * j + 1] + k[2] * mu[3 * j + 2]);
                    S += w * ph[j];
                }
                en += 0.5 * g * creal(S * conj(S));   /* E = 1/2 sum_k g |S(k)|^2 */
                for (int i = 0; i < n; i++) {
                    const double complex t = S * conj(ph[i]);
                    if (phi) phi[i] += g * creal(t);
                    if (E)
                        for (int d = 0; d < 3; d++) E[3 * i + d] -= g * k[d] * cimag(t);
                    if (dE) {
                        const double re = creal(t);
                        dE[6 * i + 0] += g * k[0] * k[0] * re;
                        dE[6 * i + 1] += g * k[1] * k[1] * re;
                        dE[6 * i + 2] += g * k[2] * k[2] * re;
                        dE[6 * i + 3] += g * k[0] * k[1] * re;
                        dE[6 * i + 4] += g * k[0] * k[2] * re;
                        dE[6 * i + 5] += g * k[1] * k[2] * re;
                    }
                }
            }
    free(ph);
    if (energy) *energy = en;
}

static void recip(const ewctx* c, const double* q, const double* mu, double* phi, double* E, double* dE,
                  double* energy) {
    if (c->e->kmax[0] > 0) recip_exact(c, q, mu, phi, E, dE, energy);
    else recip_spme(c, q, mu, phi, E, dE, energy);
}

/* ------------------------------------------------------------------ fields and solve */
static void permanent_field(const ewctx* c, double* e0, double* phi_q, double* uqq_recip) {
    const int n = c->n;
    memset(e0, 0, sizeof(double) * 3 * n);
    real_out o = {NULL, e0, NULL, NULL, NULL, 0};
    real_space(c, &o);
    double* Er = (double*)malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
    recip(c, c->s->charge, NULL, phi_q, Er, NULL, uqq_recip);
    for (int k = 0; k < 3 * n; k++) e0[k] += Er[k];
    free(Er);
}

static void induced_field(const ewctx* c, const double* mu, double* tmu) {
    const int n = c->n;
    memset(tmu, 0, sizeof(double) * 3 * n);
    real_out o = {mu, NULL, tmu, NULL, NULL, 0};
    real_space(c, &o);
    double* Er = (double*)malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
    recip(c, NULL, mu, NULL, Er, NULL, NULL);
    const double self = 4.0 * c->a * c->a * c->a / (3.0 * sqrt(M_PI));
    for (int k = 0; k < 3 * n; k++) tmu[k] += Er[k] + self * mu[k];
    free(Er);
}

/* Jacobi-preconditioned CG on (alpha^-1 - T) mu = E0 from mu = alpha E0, stopped when
 * RMS(alpha r) <= target or after max_iter iterations (the HIP library's PCG rule). */
static void solve(const ewctx* c, const double* e0, double* mu, tdpo_stats* st) {
    const tdpo_system* s = c->s;
    const int n = c->n;
    for (int i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) mu[3 * i + k] = s->polarity[i] * e0[3 * i + k];
    if (s->polarization == TDPO_DIRECT) {
        st->iterations = 0;
        st->epsilon = 1.0e50;
        st->converged = 1;
        return;
    }
    const size_t N3 = 3 * (size_t)(n > 0 ? n : 1);
    double *r = (double*)malloc(sizeof(double) * N3), *z = (double*)malloc(sizeof(double) * N3),
           *pv = (double*)malloc(sizeof(double) * N3), *tp = (double*)malloc(sizeof(double) * N3);
    induced_field(c, mu, tp);
    double rz = 0.0, zz = 0.0;
    for (int i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) {
            const size_t q = 3 * (size_t)i + k;
            const double a = s->polarity[i];
            r[q] = e0[q] - (a > 0 ? mu[q] / a : 0.0) + tp[q];
            z[q] = a * r[q];
            pv[q] = z[q];
            rz += r[q] * z[q];
            zz += z[q] * z[q];
        }
    double eps = sqrt(zz / n);
    int it = 0;
    while (it < s->max_iter && eps > s->target_eps) {
        induced_field(c, pv, tp);
        double pq = 0.0;
        for (int i = 0; i < n; i++)
            for (int k = 0; k < 3; k++) {
                const size_t q = 3 * (size_t)i + k;
                const double a = s->polarity[i];
                tp[q] = (a > 0 ? pv[q] / a : 0.0) - tp[q];   /* A p */
                pq += pv[q] * tp[q];
            }
        const double gam = rz / pq;
        double rzn = 0.0;
        zz = 0.0;
        for (int i = 0; i < n; i++)
            for (int k = 0; k < 3; k++) {
                const size_t q = 3 * (size_t)i + k;
                mu[q] += gam * pv[q];
                r[q] -= gam * tp[q];
                z[q] = s->polarity[i] * r[q];
                rzn += r[q] * z[q];
                zz += z[q] * z[q];
            }
        const double beta = rzn / rz;
        rz = rzn;
        for (size_t q = 0; q < 3 * (size_t)n; q++) pv[q] = z[q] + beta * pv[q];
        eps = sqrt(zz / n);
        it++;
    }
    st->iterations = it;
    st->epsilon = eps;
    st->converged = eps < s->target_eps ? 1 : 0;
    free(r); free(z); free(pv); free(tp);
}

/* Energy and forces at given dipoles; e0, phi_q and the reciprocal charge energy from
 * permanent_field. */
static double energy_forces(const ewctx* c, const double* mu, const double* e0, double uqq_recip, double* forces) {
    const tdpo_system* s = c->s;
    const int n = c->n;
    double uqq = 0.0;
    const size_t N3 = 3 * (size_t)(n > 0 ? n : 1);
    double* f = (double*)calloc(N3, sizeof(double));
    real_out o = {mu, NULL, NULL, f, &uqq, s->polarization == TDPO_MUTUAL};
    real_space(c, &o);
    /* reciprocal forces: F_i = q_i E_i + (grad E)_i mu_i, E from charges and dipoles; the
     * gradient of the dipoles' own field only for Mutual */
    double* Eq = (double*)malloc(sizeof(double) * N3);
    double* Em = (double*)malloc(sizeof(double) * N3);
    double* dEq = (double*)malloc(sizeof(double) * 2 * N3);
    double* dEm = (double*)malloc(sizeof(double) * 2 * N3);
    recip(c, s->charge, NULL, NULL, Eq, dEq, NULL);
    recip(c, NULL, mu, NULL, Em, dEm, NULL);
    const int mutual = s->polarization == TDPO_MUTUAL;
    for (int i = 0; i < n; i++) {
        const double qi = s->charge[i];
        double h[6];
        for (int k = 0; k < 6; k++) h[k] = dEq[6 * i + k] + (mutual ? dEm[6 * i + k] : 0.0);
        const double* m = &mu[3 * i];
        f[3 * i + 0] += qi * (Eq[3 * i + 0] + Em[3 * i + 0]) + h[0] * m[0] + h[3] * m[1] + h[4] * m[2];
        f[3 * i + 1] += qi * (Eq[3 * i + 1] + Em[3 * i + 1]) + h[3] * m[0] + h[1] * m[1] + h[5] * m[2];
        f[3 * i + 2] += qi * (Eq[3 * i + 2] + Em[3 * i + 2]) + h[4] * m[0] + h[5] * m[1] + h[2] * m[2];
    }
    double qsum = 0.0, q2 = 0.0, upol = 0.0;
    for (int i = 0; i < n; i++) {
        qsum += s->charge[i];
        q2 += s->charge[i] * s->charge[i];
        for (int k = 0; k < 3; k++) upol += mu[3 * i + k] * e0[3 * i + k];
    }
    const double self = -c->a / sqrt(M_PI) * q2;
    const double bg = -M_PI * qsum * qsum / (2.0 * c->V * c->a * c->a);
    if (forces)
        for (size_t k = 0; k < 3 * (size_t)n; k++) forces[k] += f[k];
    free(f); free(Eq); free(Em); free(dEq); free(dEm);
    return uqq + uqq_recip + self + bg - 0.5 * upol;
}

/* ------------------------------------------------------------------ entry points */
int tdpo_ewald_fields(const tdpo_system* s, const tdpo_ewald* e, const double* pos, const double* mu, double* e0,
                      double* tmu) {
    ewctx c;
    int rc = ew_setup(&c, s, e, pos);
    if (rc) { ew_free(&c); return rc; }
    if (e0) permanent_field(&c, e0, NULL, NULL);
    if (tmu && mu) induced_field(&c, mu, tmu);
    ew_free(&c);
    return 0;
}

int tdpo_ewald_energy_at(const tdpo_system* s, const tdpo_ewald* e, const double* pos, const double* mu,
                         double* forces, double* energy, double* e0_out) {
    ewctx c;
    int rc = ew_setup(&c, s, e, pos);
    if (rc) { ew_free(&c); return rc; }
    const int n = s->n;
    double* e0 = (double*)malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
    double urec = 0.0;
    permanent_field(&c, e0, NULL, &urec);
    const double u = energy_forces(&c, mu, e0, urec, forces);
    if (energy) *energy = u;
    if (e0_out) memcpy(e0_out, e0, sizeof(double) * 3 * n);
    free(e0);
    ew_free(&c);
    return 0;
}

int tdpo_ewald_evaluate(const tdpo_system* s, const tdpo_ewald* e, const double* pos, double* forces,
                        double* energy, double* induced, double* e0_out, tdpo_stats* st) {
    ewctx c;
    int rc = ew_setup(&c, s, e, pos);
    if (rc) { ew_free(&c); return rc; }
    const int n = s->n;
    const size_t N3 = 3 * (size_t)(n > 0 ? n : 1);
    double* e0 = (double*)malloc(sizeof(double) * N3);
    double* mu = (double*)malloc(sizeof(double) * N3);
    double urec = 0.0;
    tdpo_stats local = {0};
    permanent_field(&c, e0, NULL, &urec);
    solve(&c, e0, mu, &local);
    local.energy = energy_forces(&c, mu, e0, urec, forces);
    if (energy) *energy = local.energy;
    if (induced) memcpy(induced, mu, sizeof(double) * 3 * n);
    if (e0_out) memcpy(e0_out, e0, sizeof(double) * 3 * n);
    if (st) *st = local;
    free(e0);
    free(mu);
    ew_free(&c);
    return 0;
}

/* Every pair i < j whose minimum-image distance is below the cutoff, by the criterion of the HIP
 * library's real-space passes (correctly rounded FP64 operations in this order, built without
 * contraction): per axis d = x_i - x_j, d -= L * floor(d / L + 0.5); (dx^2 + dy^2) + dz^2 < rc^2.
 * pairs[2k], pairs[2k+1] for k < min(count, cap), i ascending then j; returns the count. */
long long tdpo_cutoff_pairs(const tdpo_system* s, const double* pos, double cutoff, long long cap, int* pairs) {
    const double rc2 = cutoff * cutoff;
    long long cnt = 0;
    for (int i = 0; i < s->n; i++)
        for (int j = i + 1; j < s->n; j++) {
            double d[3];
            for (int k = 0; k < 3; k++) {
                d[k] = pos[3 * i + k] - pos[3 * j + k];
                d[k] -= s->box[k] * floor(d[k] / s->box[k] + 0.5);
            }
            if ((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2] < rc2) {
                if (cnt < cap) {
                    pairs[2 * cnt] = i;
                    pairs[2 * cnt + 1] = j;
                }
                cnt++;
            }
        }
    return cnt;
}
