/*
 * tdp_oracle.h -- CPU FP64 restatement of the reference platform's TholeDipoleForce
 * arithmetic (platforms/reference/src/ReferenceTholeDipoleForce.cpp).
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker for the HIP library: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The product
 * (libtholedipole_hip.so) never links, loads or calls anything in this directory.
 *
 * Every function follows the reference function named in its comment, including the
 * reference's quirks (one-sided position-indexed scale maps, the QI-frame index pattern of
 * calculateElectrostaticPairIxn, min-image on the box diagonal only).  Where the reference
 * has undefined behaviour (covalent lists longer than 5 entries, missing axis atoms) the
 * oracle returns an error code instead.
 */
#ifndef TDP_ORACLE_H
#define TDP_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Enum values follow openmmapi/include/TholeDipoleForce.h:58-109. */
enum { TDPO_NOCUTOFF = 0, TDPO_PME = 1 };
enum { TDPO_DIRECT = 0, TDPO_MUTUAL = 1, TDPO_EXTRAPOLATED = 2 };
enum { TDPO_ZTHENX = 0, TDPO_BISECTOR = 1, TDPO_ZBISECT = 2, TDPO_THREEFOLD = 3,
       TDPO_ZONLY = 4, TDPO_NOAXIS = 5 };

typedef struct {
    int n;                       /* number of particles */
    const double* charge;        /* [n] */
    const double* dipole;        /* [3n] molecular-frame dipoles, interleaved xyz */
    const double* polarity;      /* [n] isotropic polarizability, nm^3 */
    const double* thole;         /* [n] Thole damping parameter */
    const int* axis_type;        /* [n] */
    const int* atom_z;           /* [n] */
    const int* atom_x;           /* [n] */
    const int* atom_y;           /* [n] */
    const int* cov_offsets;      /* [4n+1]: list (i,t) is cov_atoms[cov_offsets[4i+t] .. cov_offsets[4i+t+1]) */
    const int* cov_atoms;
    int method;                  /* TDPO_NOCUTOFF or TDPO_PME */
    double box[3];               /* box diagonal, used for the minimum image when method == TDPO_PME */
    int polarization;            /* TDPO_DIRECT / TDPO_MUTUAL / TDPO_EXTRAPOLATED */
    int max_iter;                /* mutual: maximum Jacobi iterations */
    double target_eps;           /* mutual: RMS dipole-update target, e*nm */
    const double* extrap;        /* extrapolated: coefficients (NULL/0 -> OPT4 default) */
    int n_extrap;
} tdpo_system;

typedef struct {
    double energy;
    int iterations;
    double epsilon;
    int converged;
} tdpo_stats;

/* Worker threads for the pair loops (1 = the reference's serial loop order).  With more
 * than one thread each thread accumulates into private arrays that are summed afterwards. */
void tdpo_set_threads(int nthreads);
int tdpo_get_threads(void);

/* Full force/energy evaluation (ReferenceTholeDipoleForce::calculateForceAndEnergy,
 * ReferenceTholeDipoleForce.cpp:100-127).  forces (3n) are ACCUMULATED.  Optional outputs
 * (may be NULL): induced dipoles (3n), lab-frame permanent dipoles (3n), torques (3n),
 * fixed field (3n).  Returns 0 on success, a negative code on invalid input. */
int tdpo_force_energy(const tdpo_system* s, const double* pos, double* forces, tdpo_stats* st,
                      double* induced, double* labdip, double* torques, double* fixed_field);

/* setup() only (ReferenceTholeDipoleForce.cpp:1117-1155): chirality, rotation and the
 * induced-dipole solve; outputs induced and lab-frame dipoles. */
int tdpo_induced_dipoles(const tdpo_system* s, const double* pos, double* induced, double* labdip,
                         double* fixed_field, tdpo_stats* st);

/* The reference's three pair passes at GIVEN induced dipoles mu (3n): fixed field, T*mu, and the
 * electrostatic energy/forces with torques mapped (ReferenceTholeDipoleForce.cpp:100-127 without the
 * solve).  e0, eind (3n) overwritten; forces (3n) accumulated; energy; any may be NULL. */
int tdpo_evaluate_at(const tdpo_system* s, const double* pos, const double* mu, double* e0, double* eind,
                     double* forces, double* energy);

/* calculateElectrostaticPotential (ReferenceTholeDipoleForce.cpp:1315-1372). */
int tdpo_potential(const tdpo_system* s, const double* pos, const double* grid, int ngrid, double* out);

/* calculateTholeDipoleSystemMultipoleMoments (ReferenceTholeDipoleForce.cpp:1225-1313), 13 values. */
int tdpo_multipole_moments(const tdpo_system* s, const double* pos, const double* masses, double* out13);

/* Effective (mScale, iScale) of pair (i,j) from setupScaleMaps/getScaleFactor
 * (ReferenceTholeDipoleForce.cpp:609-644); queried as map[min][max]. */
int tdpo_pair_scales(const tdpo_system* s, int i, int j, double* m_scale, double* i_scale);

/* Selected rows of the all-pairs sums at full size, for given induced dipoles mu (3n):
 * fixed field, T*mu, electrostatic force and torque on each row particle, and the pair energy
 * summed over partners j > i.  Outputs are [3*nrows] (energy_out [nrows]); any may be NULL. */
int tdpo_rows(const tdpo_system* s, const double* pos, const double* mu, const int* rows, int nrows,
              double* e0_out, double* eind_out, double* force_out, double* torque_out, double* energy_out);

/* Rows [i0, i1) of one reference pair pass (which: 0 fixed field, 1 T*mu, 2 electrostatics),
 * partners j > i, for timing a bounded sample; returns the number of pairs visited. */
double tdpo_pass_slice(const tdpo_system* s, const double* pos, const double* mu, int which, int i0, int i1,
                       double* out);

/* Single-pair building blocks, exported for the pinning tests. */
void tdpo_damped_rr(double thole_i, double thole_j, double r, double* rr3, double* rr5);
void tdpo_qi_frame(const double* pi, const double* pj, const double* delta, double r, double* rot9);
/* One calculateElectrostaticPairIxn call for the ordered pair (I,K): positions, charges,
 * lab-frame permanent dipoles, induced dipoles, Thole parameters, scales, polarization type.
 * Outputs: returned energy; force on K (force on I is its negative), torques on I and K. */
double tdpo_pair_electrostatic(const double* pos_i, const double* pos_k, double q_i, double q_k,
                               const double* d_i, const double* d_k, const double* u_i,
                               const double* u_k, double a_i, double a_k, double m_scale,
                               double i_scale, int polarization, const double* box_or_null,
                               double* force_k, double* torque_i, double* torque_k);
/* applyRotationMatrixToParticle + checkChiralCenterAtParticle for particle i. */
int tdpo_lab_dipoles(const tdpo_system* s, const double* pos, double* labdip);
/* mapTorqueToForceForParticle for every particle (ReferenceTholeDipoleForce.cpp:1092-1115). */
int tdpo_map_torques(const tdpo_system* s, const double* pos, const double* torques, double* forces);

/* ---- Ewald sums (tdp_ewald.c): charges and induced dipoles with Thole damping summed over all
 * periodic images; the model is described at the top of tdp_ewald.c.  Requires method PME, an
 * orthorhombic box (s->box), zero molecular dipoles, Direct or Mutual polarization. */
typedef struct {
    double cutoff;   /* real-space cutoff, nm (at most half the box) */
    double alpha;    /* Ewald splitting parameter, 1/nm */
    int grid[3];     /* smooth PME grid */
    int order;       /* B-spline order (5) */
    int kmax[3];     /* kmax[0] > 0: exact Ewald sum over |n_d| <= kmax[d] instead of smooth PME */
} tdpo_ewald;

/* E0 (3n) and T mu (3n, including the Ewald self-field) at given dipoles; either may be NULL. */
int tdpo_ewald_fields(const tdpo_system* s, const tdpo_ewald* e, const double* pos, const double* mu, double* e0,
                      double* tmu);
/* Energy and forces (accumulated, 3n) at GIVEN induced dipoles mu; e0_out (3n) may be NULL. */
int tdpo_ewald_energy_at(const tdpo_system* s, const tdpo_ewald* e, const double* pos, const double* mu,
                         double* forces, double* energy, double* e0_out);
/* Full evaluation: E0, the induced-dipole solve (PCG, stop at RMS(alpha r) <= target_eps), energy and
 * forces (accumulated).  Any output may be NULL. */
int tdpo_ewald_evaluate(const tdpo_system* s, const tdpo_ewald* e, const double* pos, double* forces,
                        double* energy, double* induced, double* e0_out, tdpo_stats* st);
/* Building blocks exported for the pinning tests: 3D unnormalised complex DFT in place
 * (re_im: interleaved complex, x slowest; sign -1 forward, +1 inverse) and the order-p B-spline
 * weights th[j] = M_p(w + p-1-j) with first and second derivatives (d1, d2 may be NULL). */
void tdpo_fft3(double* re_im, int nx, int ny, int nz, int sign);
/* Every pair i < j within the cutoff by the HIP library's FP64 neighbour criterion (see tdp_ewald.c);
 * returns the count, writes the first cap pairs. */
long long tdpo_cutoff_pairs(const tdpo_system* s, const double* pos, double cutoff, long long cap, int* pairs);
void tdpo_bspline(double w, int p, double* th, double* d1, double* d2);

#ifdef __cplusplus
}
#endif
#endif
