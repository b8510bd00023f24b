/*
 * tholedipole_hip.h -- C ABI of libtholedipole_hip.so, the MI355X (gfx950) implementation of
 * the TholeDipoleForce force + induction path.
 *
 * The ABI replaces, below openmmapi/include/TholeDipoleKernels.h:46-122
 * (class CalcTholeDipoleForceKernel), the reference platform's kernel
 * (platforms/reference/src/ReferenceTholeDipoleKernels.{h,cpp}).  A host-side kernel class
 * (C++, Python, ...) implements the ten CalcTholeDipoleForceKernel virtuals by calling these
 * functions; each entry point below names the virtual it serves.
 *
 * Conventions
 *   - plain C types only; no HIP, torch or OpenMM types cross the boundary;
 *   - every function returns an int status (TDP_OK == 0) and never throws; the message of the
 *     last failure is available from tdp_last_error() (or tdp_last_create_error() when
 *     tdp_create itself failed);
 *   - positions, forces and dipoles are arrays of 3*n doubles (x,y,z per particle, the memory
 *     layout of std::vector<OpenMM::Vec3>); units nm, e, e*nm, energies e^2/nm (the reference
 *     uses _electric = 1, ReferenceTholeDipoleForce.cpp:20);
 *   - forces are ACCUMULATED into the caller's array (+=), as the reference does
 *     (ReferenceTholeDipoleForce.cpp:359-362);
 *   - a context is not re-entrant: one thread at a time, as an OpenMM kernel.
 */
#ifndef THOLEDIPOLE_HIP_H
#define THOLEDIPOLE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TDP_OK 0
#define TDP_ERR_INVALID_ARGUMENT 1   /* bad parameters (axis atoms, covalent lists, sizes) */
#define TDP_ERR_HIP 2                /* a HIP runtime call failed */
#define TDP_ERR_NOT_PME 3            /* getPMEParameters on a NoCutoff context */
#define TDP_ERR_BOX_TOO_SMALL 4      /* box < 2*cutoff (ReferenceTholeDipoleKernels.cpp:170-173) */
#define TDP_ERR_PARTICLE_COUNT 5     /* copyParameters with a different particle count */
#define TDP_ERR_NO_DEVICE 6          /* no usable gfx950 device */
#define TDP_ERR_NOT_CONVERGED 7      /* a solve ended above its tolerance (responses; Mutual solves when
                                        TDP_FLAG_REQUIRE_CONVERGENCE is set); results are still written */

/* nonbonded_method (TholeDipoleForce.h:58-70) */
#define TDP_NO_CUTOFF 0
#define TDP_PME 1
/* polarization_type (TholeDipoleForce.h:72-92) */
#define TDP_DIRECT 0
#define TDP_MUTUAL 1
#define TDP_EXTRAPOLATED 2
/* precision */
#define TDP_PRECISION_DOUBLE 0       /* every pair term in FP64 */
#define TDP_PRECISION_MIXED 1        /* far tile pairs of T*mu and energy/force in FP32, FP64 sums; FP64 elsewhere */
/* solver for Mutual polarization */
#define TDP_SOLVER_JACOBI 0          /* the reference's fixed-point iteration and stopping rule */
#define TDP_SOLVER_PCG 1             /* Jacobi-preconditioned conjugate gradients, same stopping rule */
/* dipole kinds for tdp_get_dipoles */
#define TDP_DIPOLES_INDUCED 0
#define TDP_DIPOLES_LAB_PERMANENT 1
#define TDP_DIPOLES_TOTAL 2

typedef struct tdp_context tdp_context;

/* Everything CalcTholeDipoleForceKernel::initialize reads from TholeDipoleForce and System
 * (ReferenceTholeDipoleKernels.cpp:83-158).  Arrays are borrowed for the duration of the call
 * and deep-copied. */
typedef struct tdp_force_desc {
    int32_t num_particles;
    const double* charges;            /* [n] */
    const double* dipoles;            /* [3n] molecular-frame dipoles */
    const double* polarizabilities;   /* [n] nm^3 */
    const double* thole;              /* [n] */
    const int32_t* axis_types;        /* [n] MultipoleAxisTypes */
    const int32_t* atom_z;            /* [n] */
    const int32_t* atom_x;            /* [n] */
    const int32_t* atom_y;            /* [n] */
    const int32_t* covalent_offsets;  /* [4n+1]: list t of particle i is covalent_atoms[off[4i+t] .. off[4i+t+1]) */
    const int32_t* covalent_atoms;
    int32_t nonbonded_method;
    int32_t polarization_type;
    double cutoff;
    double ewald_error_tolerance;
    double pme_alpha;                 /* setPMEParameters; 0 selects from the tolerance */
    int32_t pme_grid[3];
    int32_t mutual_max_iterations;
    double mutual_target_epsilon;     /* RMS dipole update, e*nm */
    const double* extrapolation_coefficients;
    int32_t num_extrapolation_coefficients;
    double default_box[9];            /* System default periodic box vectors (row-major a,b,c) */
} tdp_force_desc;

/* option flags (tdp_options.flags) */
#define TDP_FLAG_REQUIRE_CONVERGENCE 1   /* tdp_execute, tdp_execute_device and tdp_get_dipoles return
                                            TDP_ERR_NOT_CONVERGED when a Mutual solve stops at its iteration
                                            limit above its target; the reference records this silently
                                            (ReferenceTholeDipoleForce.cpp:849-850) */
#define TDP_FLAG_EWALD 2                 /* PME method: Ewald-summed electrostatics (no counterpart in the
                                            reference, whose PME method is the minimum image over all pairs,
                                            ReferenceTholeDipoleForce.cpp:1375-1423): pairs within the cutoff
                                            in real space plus smooth PME of order 5 on the grid and Ewald
                                            parameter of tdp_get_pme_parameters, for charges and induced
                                            dipoles (molecular dipoles zero), Direct or Mutual polarization,
                                            rectangular boxes.  Energy: U_qq - 1/2 sum mu.E0; forces: -grad
                                            of it at fixed mu (the exact gradient: U is stationary in mu) */

typedef struct tdp_options {
    int32_t precision;                /* TDP_PRECISION_* */
    int32_t solver;                   /* TDP_SOLVER_* */
    int32_t device;                   /* HIP device ordinal */
    int32_t flags;                    /* TDP_FLAG_* bits, 0 for the reference's behaviour */
} tdp_options;

typedef struct tdp_solver_stats {
    int32_t iterations;               /* matvec iterations of the last solve */
    double epsilon;                   /* final RMS dipole update, e*nm */
    int32_t converged;
    double matvec_ms;                 /* GPU time of the T*mu kernels since timing was enabled (HIP events) */
    int32_t matvec_launches;          /* T*mu launches timed since timing was enabled */
    double matvec_ms_full;            /* the same over the full passes: launches taking at least half the
                                         median (launches after a solve stopped return at once) */
    int32_t matvec_full_launches;
} tdp_solver_stats;

/* Replaces CalcTholeDipoleForceKernel::initialize (TholeDipoleKernels.h:59,
 * ReferenceTholeDipoleKernels.cpp:83-158): validates, uploads parameters and builds the
 * exclusion tiles.  Errors as TholeDipoleForceImpl::initialize (TholeDipoleForceImpl.cpp:52-108). */
int tdp_create(const tdp_force_desc* desc, const tdp_options* opts, tdp_context** out);
void tdp_destroy(tdp_context* ctx);
const char* tdp_last_error(const tdp_context* ctx);
const char* tdp_last_create_error(void);

/* Replaces copyParametersToContext (TholeDipoleKernels.h:75, ReferenceTholeDipoleKernels.cpp:317-339):
 * per-particle parameters only; covalent maps and solver settings are kept. */
int tdp_copy_parameters(tdp_context* ctx, const tdp_force_desc* desc);

/* Replaces execute (TholeDipoleKernels.h:68, ReferenceTholeDipoleKernels.cpp:201-215).
 * positions: host [3n]; box: host [9] (ignored for NoCutoff); forces: host [3n], accumulated
 * (may be NULL when include_forces == 0); energy: host scalar (may be NULL).  With include_forces
 * == 0 (or forces NULL) the torque mapping is skipped and forces are left untouched; with
 * include_energy == 0 (or energy NULL) the energy reduction is skipped; with neither, only the
 * induced dipoles are solved. */
int tdp_execute(tdp_context* ctx, const double* positions, const double* box, int32_t include_forces,
                int32_t include_energy, double* forces, double* energy);

/* Same evaluation with device-resident buffers on the caller's HIP stream (a hipStream_t
 * passed as void*, NULL = the null stream): d_positions [3n] doubles, d_forces [3n] doubles
 * accumulated on the device (may be NULL: no forces), d_energy one double written on the device
 * (may be NULL: no energy; a batch context writes one energy per system). */
int tdp_execute_device(tdp_context* ctx, const double* d_positions, const double* box, double* d_forces,
                       double* d_energy, void* stream);

/* No counterpart in the reference: tdp_execute_device for callers that cannot allocate device or
 * pinned memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * allocation).  Same arguments and results as tdp_execute_device; the evaluation's partial sums
 * live in the library's static device scratch of TDP_NOALLOC_SCRATCH_BYTES (one copy per GPU, held
 * until the process ends) instead of the context's buffers, and the call makes no hipMalloc,
 * hipHostMalloc, hipFree, event creation or graph instantiation.  Limits and differences:
 *   - contexts whose partial sums exceed the scratch are rejected with TDP_ERR_INVALID_ARGUMENT
 *     (about 7,000 particles in one system at the default schedule);
 *   - the context's internal spatial order is kept as it is (the caller's order until a
 *     tdp_execute* call has computed one), so results match tdp_execute_device to rounding;
 *   - solver iterations are enqueued one launch at a time (no captured graphs), still with one host
 *     read of the stop count per chunk of iterations for a Mutual solve;
 *   - timing must be off (TDP_ERR_INVALID_ARGUMENT otherwise).
 * Every call on one GPU uses that GPU's one scratch: calls that overlap on one GPU (several
 * streams or threads) race; serialise them. */
#define TDP_NOALLOC_SCRATCH_BYTES (64u << 20)
int tdp_execute_device_noalloc(tdp_context* ctx, const double* d_positions, const double* box, double* d_forces,
                               double* d_energy, void* stream);

/* No counterpart in the reference (its PME method has no cutoff): every pair i < j (caller indices)
 * whose minimum-image distance the Ewald option's real-space passes find below the cutoff, for the
 * given positions: pairs[2k], pairs[2k+1] for k < min(*count, capacity), in no particular order;
 * *count receives the number of pairs.  The criterion is evaluated in FP64 with correctly rounded
 * operations: per axis d = x_i - x_j, d -= L * floor(d / L + 0.5); r2 = (dx^2 + dy^2) + dz^2 < cutoff^2. */
int tdp_get_neighbor_pairs(tdp_context* ctx, const double* positions, const double* box, int64_t capacity,
                           int32_t* pairs, int64_t* count);

/* No counterpart in the reference: the Ewald option's layout after the last evaluation, for
 * diagnostics and the memory-traffic figure of the stored T v pass.  entries: real-space blocks of
 * T held per direction (i, j) and (j, i), FP32 and FP64; slots: the same lists with their padding
 * to the width of a wave (what one T v pass streams); stored 0 when the lists did not fit in
 * memory and the T v passes run matrix-free. */
typedef struct tdp_ewald_info {
    double alpha;                     /* Ewald parameter, nm^-1 */
    int32_t grid[3];                  /* PME grid */
    double cutoff;                    /* real-space cutoff, nm */
    double r_fp64;                    /* mixed precision: pairs closer than this stay in FP64 */
    int32_t stored;
    int64_t entries32, entries64;
    int64_t slots32, slots64;
    int32_t bytes_per_entry32, bytes_per_entry64;
    int32_t tiles;                    /* tiles of 64 slots */
    int32_t max_neighbor_tiles;       /* largest neighbour-tile count of a tile */
    double mean_neighbor_tiles;
} tdp_ewald_info;
int tdp_get_ewald_info(const tdp_context* ctx, tdp_ewald_info* info);

/* Replaces getInducedDipoles / getLabFramePermanentDipoles / getTotalDipoles
 * (TholeDipoleKernels.h:91-105, ReferenceTholeDipoleKernels.cpp:217-282): recomputes from the
 * given positions, like the reference.  out: host [3n]. */
int tdp_get_dipoles(tdp_context* ctx, const double* positions, const double* box, int32_t kind, double* out);

/* Replaces getElectrostaticPotential (TholeDipoleKernels.h:113-114, ReferenceTholeDipoleKernels.cpp:284-307). */
int tdp_get_electrostatic_potential(tdp_context* ctx, const double* positions, const double* box,
                                    const double* grid, int32_t num_grid, double* out);

/* Replaces getSystemMultipoleMoments (TholeDipoleKernels.h:121, ReferenceTholeDipoleKernels.cpp:309-315):
 * masses [n]; out [13] (charge, dipole xyz in Debye, traceless quadrupole 3x3 in Debye*Angstrom). */
int tdp_get_system_multipole_moments(tdp_context* ctx, const double* positions, const double* box,
                                     const double* masses, double* out);

/* Replaces getPMEParameters (TholeDipoleKernels.h:84, ReferenceTholeDipoleKernels.cpp:341-348): the
 * values given to setPMEParameters, or when alpha or the grid is 0 the closed form of OpenMM's
 * calcPMEParameters (alpha = sqrt(-ln(2 tol))/cutoff, grid sizes rounded up to prime factors 2, 3, 5
 * and 7); the Ewald option computes with these. */
int tdp_get_pme_parameters(const tdp_context* ctx, double* alpha, int32_t* nx, int32_t* ny, int32_t* nz);

/* Solver diagnostics of the last evaluation (the reference records these silently,
 * ReferenceTholeDipoleForce.cpp:849-850). */
int tdp_get_solver_stats(const tdp_context* ctx, tdp_solver_stats* stats);

/* Enables (and resets) GPU timing of every T*mu launch with HIP events recorded on the
 * context's stream around the launch; the totals are reported by tdp_get_solver_stats. */
int tdp_set_timing(tdp_context* ctx, int32_t enable);

/* Warm start (no counterpart in the reference, which starts every Mutual solve from alpha E0,
 * ReferenceTholeDipoleForce.cpp:735-764): with enable != 0 each later Mutual solve of the context
 * starts from the induced dipoles of its previous solve, as an MD loop would use it; the stopping
 * rule is unchanged, so the result differs from a cold start by less than the solver tolerance.
 * Enabling or disabling forgets the stored dipoles. */
int tdp_set_warm_start(tdp_context* ctx, int32_t enable);

/* Schedule autotuner (no counterpart in the reference): times the T*mu pass on the given positions
 * (host [3N], caller's order; box host [9] or NULL as for tdp_execute) for each candidate length of
 * the pair kernels' work items (1, 2, 3, 4, 6, 8, 12 or 16 tiles of one tile row) and keeps the
 * fastest for the context.  The systems of a context fall into size buckets by their count of 64-site
 * tiles (1-2, 3-4, 5-8, 9-16, 17 and more), each with its own length: the tuner first sweeps one
 * length for all buckets, then, when the systems of a batch occupy more than one bucket, sweeps each
 * occupied bucket in turn (largest systems first) with the others held.  *tiles_per_item (may be NULL)
 * receives the length of the bucket of the largest system.  Without it a context uses
 * min(3, max(1, tile pairs / 16,384)) tiles in every bucket.  The choice changes only the order in
 * which partial sums are added, i.e. results at the rounding level; after it, evaluations of the
 * context are deterministic again.  Contexts with the Ewald option have no work items: no effect. */
int tdp_autotune(tdp_context* ctx, const double* positions, const double* box, int32_t* tiles_per_item);

/* The work-item schedule of a context (no counterpart in the reference): tiles_per_item[5] receives the
 * length of each size bucket; tiles_per_system_bound[5] (may be NULL) the largest tile count of each
 * bucket (2, 4, 8, 16, INT32_MAX). */
int tdp_get_work_schedule(const tdp_context* ctx, int32_t* tiles_per_item, int32_t* tiles_per_system_bound);

/* Sets the work-item schedule (no counterpart in the reference): tiles_per_item[5], one length in
 * [1, 4096] per size bucket, as tdp_autotune would choose it; the same effect on results. */
int tdp_set_work_schedule(tdp_context* ctx, const int32_t* tiles_per_item);

/* ---- Batches of independent systems.  No counterpart in the reference, whose kernel holds one
 * system: these entry points evaluate many NoCutoff systems (clusters) at once, each with the
 * same semantics as its own tdp_create context (same pair terms, scale maps, solver and stopping
 * rule per system), in one sequence of launches over all of them.
 * descs[count]: one description per system, all NoCutoff and of one polarization type; each
 * system keeps its own solver limits.  Particles are numbered consecutively across the systems
 * (system g's particle i is particle p0(g) + i, p0 = running sum of num_particles). */
int tdp_batch_create(const tdp_force_desc* descs, int32_t count, const tdp_options* opts, tdp_context** out);
int32_t tdp_batch_num_systems(const tdp_context* ctx);
/* positions: host [3N], N = total particles; forces: host [3N], accumulated (may be NULL);
 * energies: host [count] (may be NULL). */
int tdp_batch_execute(tdp_context* ctx, const double* positions, double* forces, double* energies);
/* device-resident form: d_positions [3N], d_forces [3N] accumulated, d_energies [count]. */
int tdp_batch_execute_device(tdp_context* ctx, const double* d_positions, double* d_forces, double* d_energies,
                             void* stream);
/* solver diagnostics of one system of the last evaluation */
int tdp_batch_get_solver_stats(const tdp_context* ctx, int32_t system, tdp_solver_stats* stats);

/* Linear responses of every system of a batch (no counterpart in the reference; it is the dense
 * direct solve of small systems): the induced dipoles for four fields, the permanent field E0
 * of calculateFixedDipoleField and unit uniform fields along x, y and z, i.e. the solutions of
 * (alpha^-1 - T) mu = b with the Mutual operator of calculateInducedDipolePairIxn.
 * method 0: FP16 Cholesky of D^1/2 (alpha^-1 - T) D^1/2 on MFMA (FP32 accumulation) refined in FP64
 *           until |D^1/2 (b - A mu)| <= tolerance |D^1/2 b| for every system and field;
 * method 1: the same with a BF16 factor;
 * method 2: Jacobi-preconditioned FP64 CG per field, stopped at RMS(alpha r) <= tolerance RMS(alpha b).
 * Systems of up to 1,024 particles (methods 0 and 1).  max_steps bounds refinement steps or CG
 * iterations.  out: host [4][3N] (field, particle, component), caller's order, or NULL for the
 * statistics only. */
typedef struct {
    int32_t steps;                  /* refinement steps (0, 1) or the largest CG iteration count (2) */
    double max_relative_residual;   /* largest final relative residual over systems and fields */
    int32_t factor_failed;          /* nonzero: a pivot of the 16-bit factorization was not positive */
    double assemble_ms;             /* GPU time of the 16-bit assembly (methods 0, 1; HIP events) */
    double factor_ms;               /* GPU time of the factorization (methods 0, 1; HIP events) */
} tdp_dense_stats;
int tdp_batch_solve_responses(tdp_context* ctx, const double* positions, int32_t method, double tolerance,
                              int32_t max_steps, double* out, tdp_dense_stats* stats);

#ifdef __cplusplus
}
#endif
#endif
