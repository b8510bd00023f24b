/*
 * tdp_oracle.c -- CPU FP64 restatement of platforms/reference/src/ReferenceTholeDipoleForce.cpp.
 *
 * TEST INFRASTRUCTURE ONLY (the checker): loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py, never by the product library.
 *
 * Each function cites the reference function it restates.  The arithmetic order inside each
 * pair function follows the reference so that the serial path agrees with the reference to
 * rounding; the loops over pairs run i < j exactly as the reference does.
 */
#include "tdp_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static int g_threads = 1;

void tdpo_set_threads(int nthreads) { g_threads = nthreads < 1 ? 1 : nthreads; }
int tdpo_get_threads(void) { return g_threads; }

/* ---------------------------------------------------------------- small vector helpers */
static inline double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline void cross3(const double* a, const double* b, double* c) {
    double x = a[1] * b[2] - a[2] * b[1];
    double y = a[2] * b[0] - a[0] * b[2];
    double z = a[0] * b[1] - a[1] * b[0];
    c[0] = x; c[1] = y; c[2] = z;
}
/* normalizeVec3 (ReferenceTholeDipoleForce.cpp:469-475): leaves a zero vector unchanged. */
static inline double normalize3(double* v) {
    double norm = sqrt(dot3(v, v));
    if (norm > 0.0) {
        double s = 1.0 / norm;
        v[0] *= s; v[1] *= s; v[2] *= s;
    }
    return norm;
}

/* getPeriodicDelta (ReferenceTholeDipoleForce.h:552 no-op; PME override
 * ReferenceTholeDipoleForce.cpp:1414-1423): delta -= L * floor(delta/L + 0.5) per axis,
 * using only the box diagonal. */
static inline void periodic_delta(const tdpo_system* s, double* d) {
    if (s->method != TDPO_PME) return;
    for (int k = 0; k < 3; k++) {
        double L = s->box[k];
        if (L > 0) d[k] -= L * floor(d[k] / L + 0.5);
    }
}

/* ---------------------------------------------------------------- scale maps */
/* setupScaleMaps (ReferenceTholeDipoleForce.cpp:609-635): M values come from list 0 (1-2) with
 * _mScale = {0,0,0,0.5,1} indexed by POSITION in the list (:37-41, :618-620); I values from
 * list 1 (1-3) with _iScale = {0,1,1,1,1} (:43-47, :627-629).  A std::map insert overwrites,
 * so for a duplicated atom the last position wins.  getScaleFactor (:637-644) returns 1.0 for
 * atoms not in the map.  Positions >= 5 index past the reference's 5-element tables; the
 * oracle rejects such lists. */
static const double kMScale[5] = {0.0, 0.0, 0.0, 0.5, 1.0};
static const double kIScale[5] = {0.0, 1.0, 1.0, 1.0, 1.0};

static int check_cov_lists(const tdpo_system* s) {
    for (int i = 0; i < s->n; i++) {
        for (int t = 0; t < 2; t++) {
            int len = s->cov_offsets[4 * i + t + 1] - s->cov_offsets[4 * i + t];
            if (len > 5) return -2;
        }
    }
    return 0;
}

static double lookup_scale(const tdpo_system* s, int owner, int atom, int type) {
    int b = s->cov_offsets[4 * owner + type];
    int e = s->cov_offsets[4 * owner + type + 1];
    for (int p = e - 1; p >= b; p--) {
        if (s->cov_atoms[p] == atom) return type == 0 ? kMScale[p - b] : kIScale[p - b];
    }
    return 1.0;
}

/* The reference queries the maps only as map[i][j] with i < j
 * (ReferenceTholeDipoleForce.cpp:382-385, :725-728, :805-808). */
int tdpo_pair_scales(const tdpo_system* s, int i, int j, double* m_scale, double* i_scale) {
    if (check_cov_lists(s)) return -2;
    int lo = i < j ? i : j, hi = i < j ? j : i;
    *m_scale = lookup_scale(s, lo, hi, 0);
    *i_scale = lookup_scale(s, lo, hi, 1);
    return 0;
}

/* ---------------------------------------------------------------- Thole damping */
/* getAndScaleInverseRs (ReferenceTholeDipoleForce.cpp:646-676) with rrI of size 2:
 * rr3 = 1/r^3, rr5 = 3/r^5; damp = a_i*a_j, gamma = min(a_i,a_j), s = -gamma*(r/damp)^3;
 * if damp != 0 and s > -50: rr3 *= 1-e^s, rr5 *= 1-(1-s)e^s. */
void tdpo_damped_rr(double thole_i, double thole_j, double r, double* rr3, double* rr5) {
    double rI = 1.0 / r;
    double r2I = rI * rI;
    double a3 = rI * r2I;
    double a5 = 3.0 * a3 * r2I;
    double damp = thole_i * thole_j;
    if (damp != 0.0) {
        double pgamma = thole_i < thole_j ? thole_i : thole_j;
        double ratio = r / damp;
        ratio = ratio * ratio * ratio;
        damp = -pgamma * ratio;
        if (damp > -50.0) {
            double dampExp = exp(damp);
            a3 *= 1.0 - dampExp;
            a5 *= 1.0 - (1.0 - damp) * dampExp;
        }
    }
    *rr3 = a3;
    *rr5 = a5;
}

/* ---------------------------------------------------------------- particle setup */
typedef struct {
    double pos[3];
    double q;
    double d[3]; /* lab-frame permanent dipole after setup */
    double alpha;
    double thole;
} pdata;

static int validate_axes(const tdpo_system* s) {
    int n = s->n;
    for (int i = 0; i < n; i++) {
        int at = s->axis_type[i], z = s->atom_z[i], x = s->atom_x[i], y = s->atom_y[i];
        /* The rotation dereferences Z, and X unless ZOnly, whenever Z>=0 && Z!=i
         * (ReferenceTholeDipoleForce.cpp:556-568, :499); ZBisect/ThreeFold also Y (:509, :516).
         * Torque mapping dereferences Z and X for every axis type but NoAxisType (:1104-1108). */
        if (z >= n || x >= n || y >= n) return -3;
        if (z >= 0 && z != i) {
            if (at != TDPO_ZONLY && x < 0) return -3;
            if ((at == TDPO_ZBISECT || at == TDPO_THREEFOLD) && y < 0) return -3;
        }
        if (at != TDPO_NOAXIS && (z < 0 || x < 0)) return -3;
        if (y >= 0 && (z < 0 || x < 0) && at == TDPO_ZTHENX) return -3;
    }
    return 0;
}

/* checkChiralCenterAtParticle (ReferenceTholeDipoleForce.cpp:432-452): for ZThenX with a
 * Y atom, flip the molecular-frame y component if the signed volume is negative. */
static void check_chiral(const tdpo_system* s, pdata* p) {
    for (int i = 0; i < s->n; i++) {
        if (s->atom_y[i] > -1) { /* :460 */
            if (s->axis_type[i] != TDPO_ZTHENX) continue;
            const double* pz = p[s->atom_z[i]].pos;
            const double* px = p[s->atom_x[i]].pos;
            const double* py = p[s->atom_y[i]].pos;
            double ad[3], bd[3], cd[3], c[3];
            for (int k = 0; k < 3; k++) {
                ad[k] = p[i].pos[k] - py[k];
                bd[k] = pz[k] - py[k];
                cd[k] = px[k] - py[k];
            }
            cross3(bd, cd, c);
            double volume = dot3(c, ad);
            if (volume < 0.0) p[i].d[1] *= -1.0;
        }
    }
}

/* applyRotationMatrixToParticle (ReferenceTholeDipoleForce.cpp:477-546): rows (X,Y,Z) of the
 * local frame; lab = sum_j d_j * R[j][i].  Applied whenever Z>=0 && Z!=i, regardless of the
 * axis type (:556). */
static void rotate_one(const tdpo_system* s, pdata* p, int i) {
    int at = s->axis_type[i];
    const double* pi = p[i].pos;
    double vz[3], vx[3], vy[3];
    const double* pz = p[s->atom_z[i]].pos;
    for (int k = 0; k < 3; k++) vz[k] = pz[k] - pi[k];
    normalize3(vz);
    if (at == TDPO_ZONLY) {
        if (fabs(vz[0]) < 0.866) { vx[0] = 1.0; vx[1] = 0.0; vx[2] = 0.0; }
        else { vx[0] = 0.0; vx[1] = 1.0; vx[2] = 0.0; }
    } else {
        const double* px = p[s->atom_x[i]].pos;
        for (int k = 0; k < 3; k++) vx[k] = px[k] - pi[k];
        if (at == TDPO_BISECTOR) {
            normalize3(vx);
            for (int k = 0; k < 3; k++) vz[k] += vx[k];
            normalize3(vz);
        } else if (at == TDPO_ZBISECT) {
            normalize3(vx);
            const double* py = p[s->atom_y[i]].pos;
            for (int k = 0; k < 3; k++) vy[k] = py[k] - pi[k];
            normalize3(vy);
            for (int k = 0; k < 3; k++) vx[k] += vy[k];
            normalize3(vx);
        } else if (at == TDPO_THREEFOLD) {
            normalize3(vx);
            const double* py = p[s->atom_y[i]].pos;
            for (int k = 0; k < 3; k++) vy[k] = py[k] - pi[k];
            normalize3(vy);
            for (int k = 0; k < 3; k++) vz[k] += vx[k] + vy[k];
            normalize3(vz);
        }
    }
    double dt = dot3(vz, vx);
    for (int k = 0; k < 3; k++) vx[k] -= vz[k] * dt;
    normalize3(vx);
    cross3(vz, vx, vy);
    const double* R[3] = {vx, vy, vz};
    double lab[3];
    for (int a = 0; a < 3; a++) {
        lab[a] = 0.0;
        for (int b = 0; b < 3; b++) lab[a] += p[i].d[b] * R[b][a];
    }
    p[i].d[0] = lab[0]; p[i].d[1] = lab[1]; p[i].d[2] = lab[2];
}

/* loadParticleData + checkChiral + applyRotationMatrix (ReferenceTholeDipoleForce.cpp:395-416,
 * :454-467, :548-571, called from setup :1130-1148). */
static int load_and_rotate(const tdpo_system* s, const double* pos, pdata* p) {
    int r = validate_axes(s);
    if (r) return r;
    for (int i = 0; i < s->n; i++) {
        for (int k = 0; k < 3; k++) {
            p[i].pos[k] = pos[3 * i + k];
            p[i].d[k] = s->dipole[3 * i + k];
        }
        p[i].q = s->charge[i];
        p[i].alpha = s->polarity[i];
        p[i].thole = s->thole[i];
    }
    check_chiral(s, p);
    for (int i = 0; i < s->n; i++) {
        if (s->atom_z[i] >= 0 && s->atom_z[i] != i) rotate_one(s, p, i);
    }
    return 0;
}

/* ---------------------------------------------------------------- permanent field */
/* calculateFixedDipoleFieldPairIxn (ReferenceTholeDipoleForce.cpp:678-713). */
static inline void fixed_field_pair(const tdpo_system* s, const pdata* pi, const pdata* pj, double m,
                                    double* Ei, double* Ej) {
    double d[3] = {pj->pos[0] - pi->pos[0], pj->pos[1] - pi->pos[1], pj->pos[2] - pi->pos[2]};
    periodic_delta(s, d);
    double r = sqrt(dot3(d, d));
    double rr3, rr5;
    tdpo_damped_rr(pi->thole, pj->thole, r, &rr3, &rr5);
    double dDj = dot3(pj->d, d);
    double fj = rr3 * pj->q - rr5 * dDj;
    double dDi = dot3(pi->d, d);
    double fi = rr3 * pi->q + rr5 * dDi;
    for (int k = 0; k < 3; k++) {
        Ei[k] -= (d[k] * fj + pj->d[k] * rr3) * m;
        Ej[k] += (d[k] * fi - pi->d[k] * rr3) * m;
    }
}

/* calculateFixedDipoleField (ReferenceTholeDipoleForce.cpp:715-733). */
static void fixed_field(const tdpo_system* s, const pdata* p, double* E0) {
    int n = s->n;
    memset(E0, 0, sizeof(double) * 3 * n);
    if (g_threads <= 1) {
        for (int i = 0; i < n; i++)
            for (int j = i + 1; j < n; j++) {
                double m = lookup_scale(s, i, j, 0);
                fixed_field_pair(s, &p[i], &p[j], m, &E0[3 * i], &E0[3 * j]);
            }
        return;
    }
#ifdef _OPENMP
    int nt = g_threads;
    double* priv = (double*)calloc((size_t)nt * 3 * n, sizeof(double));
#pragma omp parallel num_threads(nt)
    {
        double* acc = priv + (size_t)omp_get_thread_num() * 3 * n;
#pragma omp for schedule(dynamic, 16)
        for (int i = 0; i < n; i++)
            for (int j = i + 1; j < n; j++) {
                double m = lookup_scale(s, i, j, 0);
                fixed_field_pair(s, &p[i], &p[j], m, &acc[3 * i], &acc[3 * j]);
            }
    }
    for (int t = 0; t < nt; t++)
        for (int k = 0; k < 3 * n; k++) E0[k] += priv[(size_t)t * 3 * n + k];
    free(priv);
#endif
}

/* ---------------------------------------------------------------- T * mu */
/* calculateInducedDipoleFields + calculateInducedDipolePairIxn
 * (ReferenceTholeDipoleForce.cpp:782-818, :766-780): rr3' = -rr3*iScale, rr5' = rr5*iScale;
 * E_i += rr3' mu_j + rr5'(mu_j.d) d and symmetrically; pairs with iScale == 0 are skipped. */
static inline void induced_pair(const tdpo_system* s, const pdata* p, int i, int j, const double* mu,
                                double* Ei, double* Ej) {
    double d[3] = {p[j].pos[0] - p[i].pos[0], p[j].pos[1] - p[i].pos[1], p[j].pos[2] - p[i].pos[2]};
    periodic_delta(s, d);
    double r = sqrt(dot3(d, d));
    double rr3, rr5;
    tdpo_damped_rr(p[i].thole, p[j].thole, r, &rr3, &rr5);
    rr3 = -rr3;
    double iscale = lookup_scale(s, i, j, 1);
    if (iscale != 0.0) {
        double a3 = rr3 * iscale, a5 = rr5 * iscale;
        const double* uj = &mu[3 * j];
        const double* ui = &mu[3 * i];
        double dd = a5 * dot3(uj, d);
        for (int k = 0; k < 3; k++) Ei[k] += uj[k] * a3 + d[k] * dd;
        dd = a5 * dot3(ui, d);
        for (int k = 0; k < 3; k++) Ej[k] += ui[k] * a3 + d[k] * dd;
    }
}

static void induced_field(const tdpo_system* s, const pdata* p, const double* mu, double* E) {
    int n = s->n;
    memset(E, 0, sizeof(double) * 3 * n);
    if (g_threads <= 1) {
        for (int i = 0; i < n; i++)
            for (int j = i + 1; j < n; j++) induced_pair(s, p, i, j, mu, &E[3 * i], &E[3 * j]);
        return;
    }
#ifdef _OPENMP
    int nt = g_threads;
    double* priv = (double*)calloc((size_t)nt * 3 * n, sizeof(double));
#pragma omp parallel num_threads(nt)
    {
        double* acc = priv + (size_t)omp_get_thread_num() * 3 * n;
#pragma omp for schedule(dynamic, 16)
        for (int i = 0; i < n; i++)
            for (int j = i + 1; j < n; j++) induced_pair(s, p, i, j, mu, &acc[3 * i], &acc[3 * j]);
    }
    for (int t = 0; t < nt; t++)
        for (int k = 0; k < 3 * n; k++) E[k] += priv[(size_t)t * 3 * n + k];
    free(priv);
#endif
}

/* ---------------------------------------------------------------- induced dipoles */
/* calculateInducedDipoles (ReferenceTholeDipoleForce.cpp:735-764), with
 * convergeInducedDipolesByDIIS (:820-851, a plain Jacobi iteration) and
 * convergeInducedDipolesByExtrapolation (:853-909). */
static void solve_induced(const tdpo_system* s, const pdata* p, double* E0, double* mu, tdpo_stats* st) {
    int n = s->n;
    fixed_field(s, p, E0);
    for (int i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) mu[3 * i + k] = E0[3 * i + k] * p[i].alpha;
    double* Eind = (double*)malloc(sizeof(double) * 3 * (size_t)n);
    if (s->polarization == TDPO_DIRECT) {
        st->converged = 1;
        st->iterations = 0;
        st->epsilon = 1.0e+50; /* _mutualInducedDipoleEpsilon keeps its initial value (:24) */
    } else if (s->polarization == TDPO_MUTUAL) {
        double eps = 1.0;
        int it;
        for (it = 0; it < s->max_iter && eps > s->target_eps; it++) {
            induced_field(s, p, mu, Eind);
            eps = 0.0;
            for (int i = 0; i < n; i++) {
                double nd[3], dl[3];
                for (int k = 0; k < 3; k++) {
                    nd[k] = (E0[3 * i + k] + Eind[3 * i + k]) * p[i].alpha;
                    dl[k] = nd[k] - mu[3 * i + k];
                }
                eps += dot3(dl, dl);
                for (int k = 0; k < 3; k++) mu[3 * i + k] = nd[k];
            }
            eps = sqrt(eps / n);
        }
        st->iterations = it;
        st->epsilon = eps;
        st->converged = eps < s->target_eps ? 1 : 0;
    } else {
        static const double opt4[4] = {-0.154, 0.017, 0.658, 0.474}; /* :859-863 */
        const double* c = s->extrap;
        int order = s->n_extrap;
        if (c == NULL || order == 0) { c = opt4; order = 4; }
        double* hist = (double*)malloc(sizeof(double) * 3 * (size_t)n * order);
        memcpy(hist, mu, sizeof(double) * 3 * n);
        for (int o = 1; o < order; o++) {
            induced_field(s, p, mu, Eind);
            for (int i = 0; i < n; i++)
                for (int k = 0; k < 3; k++) {
                    mu[3 * i + k] = (E0[3 * i + k] + Eind[3 * i + k]) * p[i].alpha;
                    hist[(size_t)o * 3 * n + 3 * i + k] = mu[3 * i + k];
                }
        }
        for (int i = 0; i < n; i++)
            for (int k = 0; k < 3; k++) {
                double v = 0.0;
                for (int o = 0; o < order; o++) v += hist[(size_t)o * 3 * n + 3 * i + k] * c[o];
                mu[3 * i + k] = v;
            }
        induced_field(s, p, mu, Eind);
        double eps = 0.0;
        for (int i = 0; i < n; i++) {
            double dl[3];
            for (int k = 0; k < 3; k++) dl[k] = (E0[3 * i + k] + Eind[3 * i + k]) * p[i].alpha - mu[3 * i + k];
            eps += dot3(dl, dl);
        }
        st->epsilon = sqrt(eps / n);
        st->converged = 1;
        st->iterations = order;
        free(hist);
    }
    free(Eind);
}

/* ---------------------------------------------------------------- QI frame */
/* formQIRotationMatrix (ReferenceTholeDipoleForce.cpp:573-607): Z along delta; X from Z + e_x,
 * or Z + e_y when the RAW positions share y and z; rows [X, Y, Z]. */
void tdpo_qi_frame(const double* pi, const double* pj, const double* delta, double r, double* rot9) {
    double vz[3] = {delta[0] / r, delta[1] / r, delta[2] / r};
    double vx[3] = {vz[0], vz[1], vz[2]};
    if (pi[1] != pj[1] || pi[2] != pj[2]) vx[0] += 1.0;
    else vx[1] += 1.0;
    double dt = dot3(vz, vx);
    for (int k = 0; k < 3; k++) vx[k] -= vz[k] * dt;
    normalize3(vx);
    double vy[3];
    cross3(vz, vx, vy);
    for (int k = 0; k < 3; k++) {
        rot9[0 + k] = vx[k];
        rot9[3 + k] = vy[k];
        rot9[6 + k] = vz[k];
    }
}

/* calculateElectrostaticPairIxn (ReferenceTholeDipoleForce.cpp:129-366), with _electric =
 * _dielectric = 1 (:20-21).  The force rotation uses the reference's index pattern (:150-159)
 * and the first rotated component is treated as the bond-axis (m=0) component (:161-185). */
static double pair_elec(const double* posI, const double* posK, double qI, double qK, const double* dI,
                        const double* dK, const double* uI, const double* uK, double aI, double aK,
                        double mScale, double iScale, int polarization, const double* box_or_null,
                        double* fK_out, double* tI_out, double* tK_out) {
    double deltaR[3] = {posK[0] - posI[0], posK[1] - posI[1], posK[2] - posI[2]};
    if (box_or_null) {
        for (int k = 0; k < 3; k++) {
            double L = box_or_null[k];
            if (L > 0) deltaR[k] -= L * floor(deltaR[k] / L + 0.5);
        }
    }
    double r2 = dot3(deltaR, deltaR);
    double r = sqrt(r2);
    double R[9];
    tdpo_qi_frame(posI, posK, deltaR, r, R);
#define QR(a, b) R[3 * (a) + (b)]
    double F[3][3];
    F[0][0] = QR(1, 1); F[0][1] = QR(2, 1); F[0][2] = QR(0, 1);
    F[1][0] = QR(1, 2); F[1][1] = QR(2, 2); F[1][2] = QR(0, 2);
    F[2][0] = QR(1, 0); F[2][1] = QR(2, 0); F[2][2] = QR(0, 0);
    double qiUindI[3], qiUindJ[3];
    for (int i = 0; i < 3; i++) {
        qiUindI[i] = 0.0;
        qiUindJ[i] = 0.0;
        for (int j = 0; j < 3; j++) {
            qiUindI[i] += QR(i, j) * uI[j];
            qiUindJ[i] += QR(i, j) * uK[j];
        }
    }
    double qiQI[4], qiQJ[4];
    qiQI[0] = qI;
    qiQJ[0] = qK;
    for (int i = 0; i < 3; i++) {
        qiQI[i + 1] = 0.0;
        qiQJ[i + 1] = 0.0;
        for (int j = 0; j < 3; j++) {
            qiQI[i + 1] += QR(i, j) * dI[j];
            qiQJ[i + 1] += QR(i, j) * dK[j];
        }
    }
#undef QR
    double qiQIX[4] = {0.0, qiQI[3], 0.0, -qiQI[1]};
    double qiQIY[4] = {0.0, -qiQI[2], qiQI[1], 0.0};
    double qiQIZ[4] = {0.0, 0.0, -qiQI[3], qiQI[2]};
    double qiQJX[4] = {0.0, qiQJ[3], 0.0, -qiQJ[1]};
    double qiQJY[4] = {0.0, -qiQJ[2], qiQJ[1], 0.0};
    double qiQJZ[4] = {0.0, 0.0, -qiQJ[3], qiQJ[2]};

    double rInvVec[5];
    double rInv = 1.0 / r;
    rInvVec[1] = 1.0 * rInv;
    for (int i = 2; i < 5; i++) rInvVec[i] = rInvVec[i - 1] * rInv;
    /* NOTE: the reference sizes rInvVec as 4 (:196) but reads rInvVec[4] (:264-278), one past
     * the end; the value there is whatever the allocator left.  The oracle continues the
     * recurrence (rInvVec[4] = 1/r^4), which is what every use site's dimensions require. */

    double dmp = aI * aK;
    double a = aI < aK ? aI : aK;
    double u = r / dmp;
    double au3 = fabs(dmp) > 1.0e-5 ? a * u * u * u : 0.0;
    double expau3 = fabs(dmp) > 1.0e-5 ? exp(-au3) : 0.0;

    double thole_c = 1.0 - expau3;
    double thole_d0 = 1.0 - expau3 * (1.0 + 1.5 * au3);
    double thole_d1 = 1.0 - expau3;
    double dthole_c = 1.0 - expau3 * (1.0 + 1.5 * au3);
    double dthole_d0 = 1.0 - expau3 * (1.0 + au3 + 1.5 * au3 * au3);
    double dthole_d1 = 1.0 - expau3 * (1.0 + au3);

    double Vij[4] = {0, 0, 0, 0}, Vji[4] = {0, 0, 0, 0}, VijR[4] = {0, 0, 0, 0}, VjiR[4] = {0, 0, 0, 0};
    double Vijd[3] = {0, 0, 0}, Vjid[3] = {0, 0, 0};

    double ePermCoef = rInvVec[1] * mScale;
    double dPermCoef = -0.5 * mScale * rInvVec[2];
    Vij[0] = ePermCoef * qiQJ[0];
    Vji[0] = ePermCoef * qiQI[0];
    VijR[0] = dPermCoef * qiQJ[0];
    VjiR[0] = dPermCoef * qiQI[0];

    ePermCoef = rInvVec[2] * mScale;
    double eUIndCoef = rInvVec[2] * iScale * thole_c;
    dPermCoef = -rInvVec[3] * mScale;
    double dUIndCoef = -2.0 * rInvVec[3] * iScale * dthole_c;
    Vij[0] += -(ePermCoef * qiQJ[1] + eUIndCoef * qiUindJ[0]);
    Vji[1] = -(ePermCoef * qiQI[0]);
    VijR[0] += -(dPermCoef * qiQJ[1] + dUIndCoef * qiUindJ[0]);
    VjiR[1] = -(dPermCoef * qiQI[0]);
    Vjid[0] = -(eUIndCoef * qiQI[0]);

    Vij[1] = ePermCoef * qiQJ[0];
    Vji[0] += ePermCoef * qiQI[1] + eUIndCoef * qiUindI[0];
    VijR[1] = dPermCoef * qiQJ[0];
    VjiR[0] += dPermCoef * qiQI[1] + dUIndCoef * qiUindI[0];
    Vijd[0] = eUIndCoef * qiQJ[0];

    ePermCoef = -2.0 * rInvVec[3] * mScale;
    eUIndCoef = -2.0 * rInvVec[3] * iScale * thole_d0;
    dPermCoef = 3.0 * rInvVec[4] * mScale;
    dUIndCoef = 6.0 * rInvVec[4] * iScale * dthole_d0;
    Vij[1] += ePermCoef * qiQJ[1] + eUIndCoef * qiUindJ[0];
    Vji[1] += ePermCoef * qiQI[1] + eUIndCoef * qiUindI[0];
    VijR[1] += dPermCoef * qiQJ[1] + dUIndCoef * qiUindJ[0];
    VjiR[1] += dPermCoef * qiQI[1] + dUIndCoef * qiUindI[0];
    Vijd[0] += eUIndCoef * qiQJ[1];
    Vjid[0] += eUIndCoef * qiQI[1];

    ePermCoef = rInvVec[3] * mScale;
    eUIndCoef = rInvVec[3] * iScale * thole_d1;
    dPermCoef = -1.5 * rInvVec[4] * mScale;
    dUIndCoef = -3.0 * rInvVec[4] * iScale * dthole_d1;
    Vij[2] = ePermCoef * qiQJ[2] + eUIndCoef * qiUindJ[1];
    Vji[2] = ePermCoef * qiQI[2] + eUIndCoef * qiUindI[1];
    VijR[2] = dPermCoef * qiQJ[2] + dUIndCoef * qiUindJ[1];
    VjiR[2] = dPermCoef * qiQI[2] + dUIndCoef * qiUindI[1];
    Vijd[1] = eUIndCoef * qiQJ[2];
    Vjid[1] = eUIndCoef * qiQI[2];
    Vij[3] = ePermCoef * qiQJ[3] + eUIndCoef * qiUindJ[2];
    Vji[3] = ePermCoef * qiQI[3] + eUIndCoef * qiUindI[2];
    VijR[3] = dPermCoef * qiQJ[3] + dUIndCoef * qiUindJ[2];
    VjiR[3] = dPermCoef * qiQI[3] + dUIndCoef * qiUindI[2];
    Vijd[2] = eUIndCoef * qiQJ[3];
    Vjid[2] = eUIndCoef * qiQI[3];

    double energy = 0.5 * (qiQI[0] * Vij[0] + qiQJ[0] * Vji[0]);
    double fIZ = qiQI[0] * VijR[0];
    double fJZ = qiQJ[0] * VjiR[0];
    double EIX = 0.0, EIY = 0.0, EIZ = 0.0, EJX = 0.0, EJY = 0.0, EJZ = 0.0;
    for (int i = 1; i < 4; i++) {
        energy += 0.5 * (qiQI[i] * Vij[i] + qiQJ[i] * Vji[i]);
        fIZ += qiQI[i] * VijR[i];
        fJZ += qiQJ[i] * VjiR[i];
        EIX += qiQIX[i] * Vij[i];
        EIY += qiQIY[i] * Vij[i];
        EIZ += qiQIZ[i] * Vij[i];
        EJX += qiQJX[i] * Vji[i];
        EJY += qiQJY[i] * Vji[i];
        EJZ += qiQJZ[i] * Vji[i];
    }
    double iEIX = qiUindI[2] * Vijd[0] - qiUindI[0] * Vijd[2];
    double iEJX = qiUindJ[2] * Vjid[0] - qiUindJ[0] * Vjid[2];
    double iEIY = qiUindI[0] * Vijd[1] - qiUindI[1] * Vijd[0];
    double iEJY = qiUindJ[0] * Vjid[1] - qiUindJ[1] * Vjid[0];
    if (polarization == TDPO_MUTUAL) {
        double eCoef = -4.0 * rInvVec[3] * iScale * thole_d0;
        double dCoef = 6.0 * rInvVec[4] * iScale * dthole_d0;
        iEIX += eCoef * qiUindI[2] * qiUindJ[0];
        iEJX += eCoef * qiUindJ[2] * qiUindI[0];
        iEIY -= eCoef * qiUindI[1] * qiUindJ[0];
        iEJY -= eCoef * qiUindJ[1] * qiUindI[0];
        fIZ += dCoef * qiUindI[0] * qiUindJ[0];
        fJZ += dCoef * qiUindJ[0] * qiUindI[0];
        eCoef = 2.0 * rInvVec[3] * iScale * thole_d1;
        dCoef = -3.0 * rInvVec[4] * iScale * dthole_d1;
        iEIX -= eCoef * qiUindI[0] * qiUindJ[2];
        iEJX -= eCoef * qiUindJ[0] * qiUindI[2];
        iEIY += eCoef * qiUindI[0] * qiUindJ[1];
        iEJY += eCoef * qiUindJ[0] * qiUindI[1];
        fIZ += dCoef * (qiUindI[1] * qiUindJ[1] + qiUindI[2] * qiUindJ[2]);
        fJZ += dCoef * (qiUindJ[1] * qiUindI[1] + qiUindJ[2] * qiUindI[2]);
    }
    double qiForce[3] = {rInv * (EIY + EJY + iEIY + iEJY), -rInv * (EIX + EJX + iEIX + iEJX), -(fJZ + fIZ)};
    double qiTorqueI[3] = {-EIX, -EIY, -EIZ};
    double qiTorqueJ[3] = {-EJX, -EJY, -EJZ};
    for (int i = 0; i < 3; i++) {
        double fv = 0.0, ti = 0.0, tj = 0.0;
        for (int j = 0; j < 3; j++) {
            fv += F[i][j] * qiForce[j];
            ti += F[i][j] * qiTorqueI[j];
            tj += F[i][j] * qiTorqueJ[j];
        }
        fK_out[i] = fv;
        tI_out[i] = ti;
        tK_out[i] = tj;
    }
    return energy;
}

double tdpo_pair_electrostatic(const double* pos_i, const double* pos_k, double q_i, double q_k,
                               const double* d_i, const double* d_k, const double* u_i,
                               const double* u_k, double a_i, double a_k, double m_scale,
                               double i_scale, int polarization, const double* box_or_null,
                               double* force_k, double* torque_i, double* torque_k) {
    return pair_elec(pos_i, pos_k, q_i, q_k, d_i, d_k, u_i, u_k, a_i, a_k, m_scale, i_scale, polarization,
                     box_or_null, force_k, torque_i, torque_k);
}

/* calculateElectrostatic (ReferenceTholeDipoleForce.cpp:368-393): forces -= f on I, += f on K. */
static double electrostatic(const tdpo_system* s, const pdata* p, const double* mu, double* torques,
                            double* forces) {
    int n = s->n;
    const double* box = s->method == TDPO_PME ? s->box : NULL;
    double energy = 0.0;
    if (g_threads <= 1) {
        for (int i = 0; i < n; i++)
            for (int j = i + 1; j < n; j++) {
                double m = lookup_scale(s, i, j, 0), is = lookup_scale(s, i, j, 1);
                double f[3], ti[3], tj[3];
                energy += pair_elec(p[i].pos, p[j].pos, p[i].q, p[j].q, p[i].d, p[j].d, &mu[3 * i], &mu[3 * j],
                                    p[i].thole, p[j].thole, m, is, s->polarization, box, f, ti, tj);
                for (int k = 0; k < 3; k++) {
                    torques[3 * i + k] += ti[k];
                    torques[3 * j + k] += tj[k];
                    forces[3 * i + k] -= f[k];
                    forces[3 * j + k] += f[k];
                }
            }
        return energy;
    }
#ifdef _OPENMP
    int nt = g_threads;
    double* priv = (double*)calloc((size_t)nt * 6 * n, sizeof(double));
#pragma omp parallel num_threads(nt) reduction(+ : energy)
    {
        double* accf = priv + (size_t)omp_get_thread_num() * 6 * n;
        double* acct = accf + 3 * n;
#pragma omp for schedule(dynamic, 16)
        for (int i = 0; i < n; i++)
            for (int j = i + 1; j < n; j++) {
                double m = lookup_scale(s, i, j, 0), is = lookup_scale(s, i, j, 1);
                double f[3], ti[3], tj[3];
                energy += pair_elec(p[i].pos, p[j].pos, p[i].q, p[j].q, p[i].d, p[j].d, &mu[3 * i], &mu[3 * j],
                                    p[i].thole, p[j].thole, m, is, s->polarization, box, f, ti, tj);
                for (int k = 0; k < 3; k++) {
                    acct[3 * i + k] += ti[k];
                    acct[3 * j + k] += tj[k];
                    accf[3 * i + k] -= f[k];
                    accf[3 * j + k] += f[k];
                }
            }
    }
    for (int t = 0; t < nt; t++)
        for (int k = 0; k < 3 * n; k++) {
            forces[k] += priv[(size_t)t * 6 * n + k];
            torques[k] += priv[(size_t)t * 6 * n + 3 * n + k];
        }
    free(priv);
#endif
    return energy;
}

/* ---------------------------------------------------------------- torque -> force */
/* mapTorqueToForceForParticle (ReferenceTholeDipoleForce.cpp:915-1090). */
static void torque_one(const pdata* p, int I, int U, int V, int W /* -1 = none */, int axisType,
                       const double* torque, double* forces) {
    if (axisType == TDPO_NOAXIS) return;
    double vU[3], vV[3], vW[3];
    for (int k = 0; k < 3; k++) {
        vU[k] = p[U].pos[k] - p[I].pos[k];
        vV[k] = p[V].pos[k] - p[I].pos[k];
    }
    double normU = normalize3(vU);
    double normV = normalize3(vV);
    if (W >= 0 && (axisType == TDPO_ZBISECT || axisType == TDPO_THREEFOLD)) {
        for (int k = 0; k < 3; k++) vW[k] = p[W].pos[k] - p[I].pos[k];
    } else {
        cross3(vU, vV, vW);
    }
    double normW = normalize3(vW);
    double vUV[3], vUW[3], vVW[3];
    cross3(vV, vU, vUV);
    cross3(vW, vU, vUW);
    cross3(vW, vV, vVW);
    normalize3(vUV);
    normalize3(vUW);
    normalize3(vVW);
    double cosUV = dot3(vU, vV), sinUV = sqrt(1.0 - cosUV * cosUV);
    double cosUW = dot3(vU, vW), sinUW = sqrt(1.0 - cosUW * cosUW);
    double cosVW = dot3(vV, vW), sinVW = sqrt(1.0 - cosVW * cosVW);
    double dphi[3] = {dot3(vU, torque), dot3(vV, torque), dot3(vW, torque)};
    dphi[0] *= -1.0; dphi[1] *= -1.0; dphi[2] *= -1.0;

    if (axisType == TDPO_ZTHENX || axisType == TDPO_BISECTOR) {
        double factor1 = dphi[1] / (normU * sinUV);
        double factor2 = dphi[2] / normU;
        double factor3 = -dphi[0] / (normV * sinUV);
        double factor4;
        if (axisType == TDPO_BISECTOR) {
            factor2 *= 0.5;
            factor4 = 0.5 * dphi[2] / normV;
        } else {
            factor4 = 0.0;
        }
        for (int i = 0; i < 3; i++) {
            double forceU = vUV[i] * factor1 + factor2 * vUW[i];
            forces[3 * U + i] -= forceU;
            double forceV = vUV[i] * factor3 + factor4 * vVW[i];
            forces[3 * V + i] -= forceV;
            forces[3 * I + i] += (forceU + forceV);
        }
    } else if (axisType == TDPO_ZBISECT) {
        double vR[3] = {vV[0] + vW[0], vV[1] + vW[1], vV[2] + vW[2]};
        double vS[3];
        cross3(vU, vR, vS);
        normalize3(vR);
        normalize3(vS);
        double vUR[3], vUS[3], vVS[3], vWS[3];
        cross3(vR, vU, vUR);
        cross3(vS, vU, vUS);
        cross3(vS, vV, vVS);
        cross3(vS, vW, vWS);
        normalize3(vUR);
        normalize3(vUS);
        normalize3(vVS);
        normalize3(vWS);
        double cosUR = dot3(vU, vR), sinUR = sqrt(1.0 - cosUR * cosUR);
        double cosUS = dot3(vU, vS), sinUS = sqrt(1.0 - cosUS * cosUS);
        double cosVS = dot3(vV, vS), sinVS = sqrt(1.0 - cosVS * cosVS);
        double cosWS = dot3(vW, vS), sinWS = sqrt(1.0 - cosWS * cosWS);
        (void)sinUS;
        double t1[3], t2[3];
        for (int k = 0; k < 3; k++) {
            t1[k] = vV[k] - vS[k] * cosVS;
            t2[k] = vW[k] - vS[k] * cosWS;
        }
        normalize3(t1);
        normalize3(t2);
        double ut1cos = dot3(vU, t1), ut1sin = sqrt(1.0 - ut1cos * ut1cos);
        double ut2cos = dot3(vU, t2), ut2sin = sqrt(1.0 - ut2cos * ut2cos);
        double dphiR = dot3(vR, torque) * (-1.0);
        double dphiS = dot3(vS, torque) * (-1.0);
        double factor1 = dphiR / (normU * sinUR);
        double factor2 = dphiS / normU;
        double factor3 = dphi[0] / (normV * (ut1sin + ut2sin));
        double factor4 = dphi[0] / (normW * (ut1sin + ut2sin));
        double fU[3], fV[3], fW[3];
        for (int k = 0; k < 3; k++) {
            fU[k] = vUR[k] * factor1 + vUS[k] * factor2;
            fV[k] = (vS[k] * sinVS - t1[k] * cosVS) * factor3;
            fW[k] = (vS[k] * sinWS - t2[k] * cosWS) * factor4;
        }
        for (int k = 0; k < 3; k++) {
            forces[3 * U + k] -= fU[k];
            forces[3 * V + k] -= fV[k];
            forces[3 * W + k] -= fW[k];
            forces[3 * I + k] += (fU[k] + fV[k] + fW[k]);
        }
    } else if (axisType == TDPO_THREEFOLD) {
        for (int i = 0; i < 3; i++) {
            double du = vUW[i] * dphi[2] / (normU * sinUW) + vUV[i] * dphi[1] / (normU * sinUV) -
                        vUW[i] * dphi[0] / (normU * sinUW) - vUV[i] * dphi[0] / (normU * sinUV);
            double dv = vVW[i] * dphi[2] / (normV * sinVW) - vUV[i] * dphi[0] / (normV * sinUV) -
                        vVW[i] * dphi[1] / (normV * sinVW) + vUV[i] * dphi[1] / (normV * sinUV);
            double dw = -vUW[i] * dphi[0] / (normW * sinUW) - vVW[i] * dphi[1] / (normW * sinVW) +
                        vUW[i] * dphi[2] / (normW * sinUW) + vVW[i] * dphi[2] / (normW * sinVW);
            du /= 3.0;
            dv /= 3.0;
            dw /= 3.0;
            forces[3 * U + i] -= du;
            forces[3 * V + i] -= dv;
            if (W >= 0) forces[3 * W + i] -= dw;
            forces[3 * I + i] += (du + dv + dw);
        }
    } else if (axisType == TDPO_ZONLY) {
        for (int i = 0; i < 3; i++) {
            double du = vUV[i] * dphi[1] / (normU * sinUV) + vUW[i] * dphi[2] / normU;
            forces[3 * U + i] -= du;
            forces[3 * I + i] += du;
        }
    }
}

/* mapTorqueToForce (ReferenceTholeDipoleForce.cpp:1092-1115). */
static void map_torques(const tdpo_system* s, const pdata* p, const double* torques, double* forces) {
    for (int ii = 0; ii < s->n; ii++) {
        if (s->axis_type[ii] != TDPO_NOAXIS)
            torque_one(p, ii, s->atom_z[ii], s->atom_x[ii], s->atom_y[ii] > -1 ? s->atom_y[ii] : -1,
                       s->axis_type[ii], &torques[3 * ii], forces);
    }
}

/* ---------------------------------------------------------------- public entry points */
int tdpo_force_energy(const tdpo_system* s, const double* pos, double* forces, tdpo_stats* st,
                      double* induced, double* labdip, double* torques_out, double* fixed_field_out) {
    int n = s->n;
    int rc = check_cov_lists(s);
    if (rc) return rc;
    pdata* p = (pdata*)malloc(sizeof(pdata) * (size_t)(n > 0 ? n : 1));
    rc = load_and_rotate(s, pos, p);
    if (rc) { free(p); return rc; }
    double* E0 = (double*)malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
    double* mu = (double*)malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
    double* tq = (double*)calloc(3 * (size_t)(n > 0 ? n : 1), sizeof(double));
    tdpo_stats local;
    solve_induced(s, p, E0, mu, &local);
    local.energy = electrostatic(s, p, mu, tq, forces);
    map_torques(s, p, tq, forces);
    if (st) *st = local;
    if (induced) memcpy(induced, mu, sizeof(double) * 3 * n);
    if (labdip)
        for (int i = 0; i < n; i++)
            for (int k = 0; k < 3; k++) labdip[3 * i + k] = p[i].d[k];
    if (torques_out) memcpy(torques_out, tq, sizeof(double) * 3 * n);
    if (fixed_field_out) memcpy(fixed_field_out, E0, sizeof(double) * 3 * n);
    free(p); free(E0); free(mu); free(tq);
    return 0;
}

/* The three pair passes of calculateForceAndEnergy (ReferenceTholeDipoleForce.cpp:100-127) at
 * GIVEN induced dipoles mu (3n) instead of solved ones: fixed field (:715-733), T*mu (:782-818),
 * calculateElectrostatic (:368-393) and mapTorqueToForce (:1092-1115).  Outputs (may be NULL):
 * e0 and eind (3n, overwritten), forces (3n, accumulated), energy. */
int tdpo_evaluate_at(const tdpo_system* s, const double* pos, const double* mu, double* e0, double* eind,
                     double* forces, double* energy) {
    int n = s->n;
    int rc = check_cov_lists(s);
    if (rc) return rc;
    pdata* p = (pdata*)malloc(sizeof(pdata) * (size_t)(n > 0 ? n : 1));
    rc = load_and_rotate(s, pos, p);
    if (rc) { free(p); return rc; }
    if (e0) fixed_field(s, p, e0);
    if (eind) induced_field(s, p, mu, eind);
    if (forces || energy) {
        double* tq = (double*)calloc(3 * (size_t)(n > 0 ? n : 1), sizeof(double));
        double* f = (double*)calloc(3 * (size_t)(n > 0 ? n : 1), sizeof(double));
        double e = electrostatic(s, p, mu, tq, f);
        map_torques(s, p, tq, f);
        if (forces)
            for (int k = 0; k < 3 * n; k++) forces[k] += f[k];
        if (energy) *energy = e;
        free(tq);
        free(f);
    }
    free(p);
    return 0;
}

int tdpo_induced_dipoles(const tdpo_system* s, const double* pos, double* induced, double* labdip,
                         double* fixed_field_out, tdpo_stats* st) {
    int n = s->n;
    int rc = check_cov_lists(s);
    if (rc) return rc;
    pdata* p = (pdata*)malloc(sizeof(pdata) * (size_t)(n > 0 ? n : 1));
    rc = load_and_rotate(s, pos, p);
    if (rc) { free(p); return rc; }
    double* E0 = (double*)malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
    double* mu = (double*)malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
    tdpo_stats local;
    solve_induced(s, p, E0, mu, &local);
    local.energy = 0.0;
    if (st) *st = local;
    if (induced) memcpy(induced, mu, sizeof(double) * 3 * n);
    if (labdip)
        for (int i = 0; i < n; i++)
            for (int k = 0; k < 3; k++) labdip[3 * i + k] = p[i].d[k];
    if (fixed_field_out) memcpy(fixed_field_out, E0, sizeof(double) * 3 * n);
    free(p); free(E0); free(mu);
    return 0;
}

int tdpo_lab_dipoles(const tdpo_system* s, const double* pos, double* labdip) {
    int n = s->n;
    pdata* p = (pdata*)malloc(sizeof(pdata) * (size_t)(n > 0 ? n : 1));
    int rc = load_and_rotate(s, pos, p);
    if (rc == 0)
        for (int i = 0; i < n; i++)
            for (int k = 0; k < 3; k++) labdip[3 * i + k] = p[i].d[k];
    free(p);
    return rc;
}

int tdpo_map_torques(const tdpo_system* s, const double* pos, const double* torques, double* forces) {
    int n = s->n;
    pdata* p = (pdata*)malloc(sizeof(pdata) * (size_t)(n > 0 ? n : 1));
    int rc = validate_axes(s);
    if (rc == 0) {
        for (int i = 0; i < n; i++)
            for (int k = 0; k < 3; k++) p[i].pos[k] = pos[3 * i + k];
        map_torques(s, p, torques, forces);
    }
    free(p);
    return rc;
}

/* calculateElectrostaticPotential + ...ForParticleGridPoint (ReferenceTholeDipoleForce.cpp:1315-1372):
 * phi(g) = sum_i q_i/r - (d_i + mu_i).(x_i - g)/r^3, with the minimum image in PME mode. */
int tdpo_potential(const tdpo_system* s, const double* pos, const double* grid, int ngrid, double* out) {
    int n = s->n;
    double* mu = (double*)malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
    double* lab = (double*)malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
    int rc = tdpo_induced_dipoles(s, pos, mu, lab, NULL, NULL);
    if (rc == 0) {
        for (int g = 0; g < ngrid; g++) {
            double phi = 0.0;
            for (int i = 0; i < n; i++) {
                double d[3] = {pos[3 * i] - grid[3 * g], pos[3 * i + 1] - grid[3 * g + 1], pos[3 * i + 2] - grid[3 * g + 2]};
                periodic_delta(s, d);
                double r2 = dot3(d, d);
                double r = sqrt(r2);
                double rr1 = 1.0 / r, rr2 = rr1 * rr1, rr3 = rr1 * rr2;
                double v = s->charge[i] * rr1;
                double scd = dot3(&lab[3 * i], d);
                double scu = dot3(&mu[3 * i], d);
                v -= (scd + scu) * rr3;
                phi += v;
            }
            out[g] = phi;
        }
    }
    free(mu); free(lab);
    return rc;
}

/* calculateTholeDipoleSystemMultipoleMoments (ReferenceTholeDipoleForce.cpp:1225-1313). */
int tdpo_multipole_moments(const tdpo_system* s, const double* pos, const double* masses, double* out) {
    int n = s->n;
    double* mu = (double*)malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
    double* lab = (double*)malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
    int rc = tdpo_induced_dipoles(s, pos, mu, lab, NULL, NULL);
    if (rc == 0) {
        double totalMass = 0.0, com[3] = {0, 0, 0};
        for (int i = 0; i < n; i++) {
            totalMass += masses[i];
            for (int k = 0; k < 3; k++) com[k] += pos[3 * i + k] * masses[i];
        }
        if (totalMass > 0.0)
            for (int k = 0; k < 3; k++) com[k] *= 1.0 / totalMass;
        double netchg = 0.0, dpl[3] = {0, 0, 0};
        double xx = 0, xy = 0, xz = 0, yy = 0, yz = 0, zz = 0;
        for (int i = 0; i < n; i++) {
            double q = s->charge[i];
            double x[3] = {pos[3 * i] - com[0], pos[3 * i + 1] - com[1], pos[3 * i + 2] - com[2]};
            double nd[3] = {lab[3 * i] + mu[3 * i], lab[3 * i + 1] + mu[3 * i + 1], lab[3 * i + 2] + mu[3 * i + 2]};
            netchg += q;
            for (int k = 0; k < 3; k++) dpl[k] += x[k] * q + nd[k];
            xx += x[0] * x[0] * q + 2.0 * x[0] * nd[0];
            xy += x[0] * x[1] * q + x[0] * nd[1] + x[1] * nd[0];
            xz += x[0] * x[2] * q + x[0] * nd[2] + x[2] * nd[0];
            yy += x[1] * x[1] * q + 2.0 * x[1] * nd[1];
            yz += x[1] * x[2] * q + x[1] * nd[2] + x[2] * nd[1];
            zz += x[2] * x[2] * q + 2.0 * x[2] * nd[2];
        }
        double qave = (xx + yy + zz) / 3.0;
        out[4] = 0.5 * (xx - qave);
        out[5] = 0.5 * xy;
        out[6] = 0.5 * xz;
        out[8] = 0.5 * (yy - qave);
        out[9] = 0.5 * yz;
        out[12] = 0.5 * (zz - qave);
        out[7] = out[5];
        out[10] = out[6];
        out[11] = out[9];
        double debye = 4.80321;
        out[0] = netchg;
        for (int k = 0; k < 3; k++) dpl[k] *= 10.0 * debye;
        out[1] = dpl[0];
        out[2] = dpl[1];
        out[3] = dpl[2];
        debye *= 3.0;
        for (int k = 4; k < 13; k++) out[k] *= 100.0 * debye;
    }
    free(mu); free(lab);
    return rc;
}

/* Rows of the all-pairs sums for selected particles, with the induced dipoles given:
 * fixed field (calculateFixedDipoleField), T*mu (calculateInducedDipoleFields), and the
 * electrostatic force, torque and pair-energy share of each row (calculateElectrostatic;
 * the pair energy is attributed to the lower index, as the reference loop visits i < j).
 * Each row sums over all j in index order with the pair evaluated in the reference's (i<j)
 * orientation, so a row equals what the reference accumulates for that particle. */
int tdpo_rows(const tdpo_system* s, const double* pos, const double* mu, const int* rows, int nrows,
              double* e0_out, double* eind_out, double* force_out, double* torque_out, double* energy_out) {
    int n = s->n;
    int rc = check_cov_lists(s);
    if (rc) return rc;
    pdata* p = (pdata*)malloc(sizeof(pdata) * (size_t)(n > 0 ? n : 1));
    rc = load_and_rotate(s, pos, p);
    if (rc) { free(p); return rc; }
    const double* box = s->method == TDPO_PME ? s->box : NULL;
#pragma omp parallel for num_threads(g_threads) schedule(dynamic, 1)
    for (int r = 0; r < nrows; r++) {
        int i = rows[r];
        double e0[3] = {0, 0, 0}, ei[3] = {0, 0, 0}, f[3] = {0, 0, 0}, t[3] = {0, 0, 0}, en = 0.0;
        double scratch[3] = {0, 0, 0};
        for (int j = 0; j < n; j++) {
            if (j == i) continue;
            int lo = i < j ? i : j, hi = i < j ? j : i;
            double m = lookup_scale(s, lo, hi, 0);
            if (i < j) fixed_field_pair(s, &p[i], &p[j], m, e0, scratch);
            else fixed_field_pair(s, &p[j], &p[i], m, scratch, e0);
            if (i < j) induced_pair(s, p, i, j, mu, ei, scratch);
            else induced_pair(s, p, j, i, mu, scratch, ei);
            double is = lookup_scale(s, lo, hi, 1);
            double fk[3], tlo[3], thi[3];
            double e = pair_elec(p[lo].pos, p[hi].pos, p[lo].q, p[hi].q, p[lo].d, p[hi].d, &mu[3 * lo], &mu[3 * hi],
                                 p[lo].thole, p[hi].thole, m, is, s->polarization, box, fk, tlo, thi);
            if (i == lo) {
                for (int k = 0; k < 3; k++) { f[k] -= fk[k]; t[k] += tlo[k]; }
                en += e;
            } else {
                for (int k = 0; k < 3; k++) { f[k] += fk[k]; t[k] += thi[k]; }
            }
        }
        for (int k = 0; k < 3; k++) {
            if (e0_out) e0_out[3 * r + k] = e0[k];
            if (eind_out) eind_out[3 * r + k] = ei[k];
            if (force_out) force_out[3 * r + k] = f[k];
            if (torque_out) torque_out[3 * r + k] = t[k];
        }
        if (energy_out) energy_out[r] = en;
    }
    free(p);
    return 0;
}

/* One slice of a reference pair pass: the outer-loop rows i in [i0, i1) with partners
 * j in (i, n), exactly the loop bodies of calculateFixedDipoleField (which == 0),
 * calculateInducedDipoleFields (which == 1) or calculateElectrostatic (which == 2)
 * (ReferenceTholeDipoleForce.cpp:715-733, :782-818, :368-393).  Used to time the CPU baseline
 * on a bounded sample of a large system; returns the number of pairs visited.  out accumulates
 * the field or force (3n); mu is the induced-dipole input (3n). */
double tdpo_pass_slice(const tdpo_system* s, const double* pos, const double* mu, int which, int i0, int i1,
                       double* out) {
    int n = s->n;
    if (check_cov_lists(s)) return -1.0;
    pdata* p = (pdata*)malloc(sizeof(pdata) * (size_t)(n > 0 ? n : 1));
    if (load_and_rotate(s, pos, p)) { free(p); return -1.0; }
    const double* box = s->method == TDPO_PME ? s->box : NULL;
    double* tq = (double*)calloc(3 * (size_t)(n > 0 ? n : 1), sizeof(double));
    double pairs = 0.0, energy = 0.0;
    int nt = g_threads;
    double* priv = (double*)calloc((size_t)nt * 3 * n, sizeof(double));
    double* privt = (double*)calloc((size_t)nt * 3 * n, sizeof(double));
#pragma omp parallel num_threads(nt) reduction(+ : pairs, energy)
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        double* acc = priv + (size_t)t * 3 * n;
        double* acct = privt + (size_t)t * 3 * n;
#pragma omp for schedule(dynamic, 4)
        for (int i = i0; i < i1; i++) {
            for (int j = i + 1; j < n; j++) {
                if (which == 0) {
                    double m = lookup_scale(s, i, j, 0);
                    fixed_field_pair(s, &p[i], &p[j], m, &acc[3 * i], &acc[3 * j]);
                } else if (which == 1) {
                    induced_pair(s, p, i, j, mu, &acc[3 * i], &acc[3 * j]);
                } else {
                    double m = lookup_scale(s, i, j, 0), is = lookup_scale(s, i, j, 1);
                    double f[3], ti[3], tj[3];
                    energy += pair_elec(p[i].pos, p[j].pos, p[i].q, p[j].q, p[i].d, p[j].d, &mu[3 * i], &mu[3 * j],
                                        p[i].thole, p[j].thole, m, is, s->polarization, box, f, ti, tj);
                    for (int k = 0; k < 3; k++) {
                        acct[3 * i + k] += ti[k];
                        acct[3 * j + k] += tj[k];
                        acc[3 * i + k] -= f[k];
                        acc[3 * j + k] += f[k];
                    }
                }
                pairs += 1.0;
            }
        }
    }
    for (int t = 0; t < nt; t++)
        for (int k = 0; k < 3 * n; k++) out[k] += priv[(size_t)t * 3 * n + k];
    free(priv); free(privt); free(tq); free(p);
    (void)energy;
    return pairs;
}
