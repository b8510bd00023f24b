// tdp_context.hip -- the C ABI of libtholedipole_hip.so (include/tholedipole_hip.h).
//
// A tdp_context is the device-side state of one CalcTholeDipoleForceKernel
// (platforms/reference/src/ReferenceTholeDipoleKernels.{h,cpp}): parameters uploaded once at
// create time, the exclusion tiles derived from the covalent maps, and the work buffers of one
// evaluation.  Every evaluation restarts the induced dipoles from alpha*E0, as the reference
// does (a new ReferenceTholeDipoleForce per call, ReferenceTholeDipoleKernels.cpp:201-215).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <functional>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/tholedipole_hip.h"
#include "tdp_kernels.h"
#include "tdp_tables.h"

using namespace tdp;

namespace {

thread_local std::string g_create_error;

}  // namespace

// Scratch of tdp_execute_device_noalloc: the partial sums of one evaluation (pair kernels' work-item
// partials and tile slabs, torque and energy partials) for callers that cannot allocate.  One copy
// per GPU, held until the process ends; calls that overlap on one GPU share it and race.
constexpr size_t kNoallocScratchBytes = (size_t)TDP_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};   // device address of scratch, per device ordinal

// Device address of this device's scratch (the device must be current); called at context
// creation, once per device.
hipError_t scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr[device]) return hipSuccess;
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = ptr;
    return hipSuccess;
}

void* scratch_ptr(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    return g_scratch_ptr[device];
}

// Zero fill of n 16-byte words (the scratch is cleared by a kernel, not a memset, so that a caller
// may capture the call).
__global__ void k_zero16(uint4* __restrict__ p, size_t n) {
    const uint4 z = make_uint4(0u, 0u, 0u, 0u);
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = z;
}

}  // namespace

// Evaluations between recomputations of the internal spatial order.
constexpr int kReorderInterval = 100;

// Size buckets of the systems of a context by tile count (1-2, 3-4, 5-8, 9-16, 17 and more): each
// bucket has its own work-item length, which the autotuner may choose separately.
constexpr int kSizeBuckets = 5;
inline int size_bucket(int tiles) { return tiles <= 2 ? 0 : tiles <= 4 ? 1 : tiles <= 8 ? 2 : tiles <= 16 ? 3 : 4; }

struct tdp_context {
    // ---- parameters (host copies)
    // n particles in the caller's order; nblk tiles of 64 slots (npad = 64 nblk) in the internal
    // order.  A context holds nseg independent systems: system s owns the caller's particles
    // [seg_p0[s], seg_p0[s] + seg_n[s]) and the tiles [seg_t0[s], seg_t1[s]); its last tile is
    // padded.  A context made by tdp_create holds one system.
    int n = 0, npad = 0, nblk = 0;
    int nseg = 1;
    std::vector<int> seg_n, seg_p0, seg_t0, seg_t1, seg_w0, seg_w1, seg_maxit;
    std::vector<double> seg_target;
    std::vector<int> tcnt, tfirst, tseg;
    std::vector<size_t> tcol;
    int method = TDP_NO_CUTOFF, polarization = TDP_MUTUAL, precision = TDP_PRECISION_DOUBLE;
    int solver = TDP_SOLVER_JACOBI, device = 0;
    int flags = 0;   // TDP_FLAG_* of the options
    double cutoff = 1.0, ewald_tol = 1e-4, pme_alpha = 0.0;
    int grid[3] = {0, 0, 0};
    std::vector<double> extrap;
    double default_box[9] = {0};
    std::vector<double> charge, dmol, polarity, thole;
    std::vector<int> axis_type, atom_z, atom_x, atom_y;
    // ---- derived host tables
    std::vector<ExcPair> exc_pairs;
    std::vector<int> perm, inv;  // internal slot -> caller index, caller index -> internal slot
    bool ordered = false;
    int evals_since_order = 0;
    size_t cap_exc_J = 0, cap_exc_mask = 0, cap_gidx = 0;
    size_t cap_work = 0, cap_pbuf = 0, cap_pbuf2 = 0, cap_part_e = 0, cap_slab32 = 0;
    int work_C = 0;   // tiles per upper-triangle work item (of the largest system's size bucket)
    int work_Cb[kSizeBuckets] = {};   // tiles per work item, per size bucket of the systems (size_bucket)
    std::vector<int> exc_start, exc_J;
    std::vector<unsigned long long> exc_mask;
    std::vector<int> goff, gidx;
    // ---- device state
    hipStream_t stream = nullptr;
    // Ewald option: the reciprocal-space chain of each T v runs on stream2 beside the real-space pass
    // (fork and join by events, also inside captured graphs); both streams at the default priority
    hipStream_t stream2 = nullptr;
    float4 *d_mu32 = nullptr, *d_slot32 = nullptr;   // the force pass's FP32 slot records (mixed)
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // the solver's host check waits on this event (recorded after its copy of the solver scalars), so
    // that work enqueued behind the copy (an evaluation's forces, ahead of the check) does not hold
    // the check up
    hipEvent_t ev_stage = nullptr;
    // Captured chunks of solver iterations (hipGraph), replayed while the launch parameters hold:
    // same chunk length, solver and precision (the pair kernel and the vectors it reads), same box,
    // same device tables (graph_gen counts table uploads).
    struct IterGraph {
        int iters;
        bool jacobi;
        int precision;
        unsigned gen;
        double box[3];
        hipGraphExec_t exec;
    };
    std::vector<IterGraph> graphs;
    // Longer captured sequences (the field and the start of a solve, or the forces and energy),
    // replayed while the tables (graph_gen), box, precision and key are unchanged.
    struct SeqGraph {
        int tag;
        long long key;
        const void* p0;
        const void* p1;
        unsigned gen;
        int precision;
        double box[3];
        hipGraphExec_t exec;
    };
    std::vector<SeqGraph> seq_graphs;
    unsigned graph_gen = 0;
    bool use_graphs = true;
    std::vector<void*> allocations;
    double4 *posa = nullptr, *qd = nullptr, *mu = nullptr, *E0 = nullptr, *r = nullptr, *z = nullptr, *p = nullptr,
            *q = nullptr, *torque = nullptr,
            *contrib = nullptr, *hist = nullptr;
    bool has_perm = true;   // some particle has a nonzero molecular dipole
    float* d_slab32 = nullptr;   // mixed T*mu: FP32 tile partial slab (3 * ntri * 64), then item partials
    float4 *vec32 = nullptr, *tb_c = nullptr, *tb_h = nullptr, *posl = nullptr;
    double4* posl64 = nullptr;   // FP64 offsets from the tile origins (the FP64 far path's frames)
    TilePair* d_tp = nullptr;   // [ntri] tile pair classification, recomputed every evaluation
    double4* tb_o = nullptr;
    // upper-triangle sweep of the pair kernels: work items, their row starts, partial sums
    int nwork = 0;
    size_t ntri = 0;
    int4* d_work = nullptr;
    int* d_row_start = nullptr;
    void* d_pbuf = nullptr;   // item partials (nwork * 64 double4), then the tile partial slab (3 * ntri * 64 doubles)
    double4* d_pbuf2 = nullptr;  // second buffer of the same shape (torque partials)
    double4* Emv = nullptr;   // T*mu of the last matvec, reduced
    double *d_pos3 = nullptr, *d_frc = nullptr, *d_alpha = nullptr, *d_inva = nullptr, *d_thole = nullptr,
           *d_charge = nullptr, *d_dmol = nullptr, *d_red = nullptr, *d_scalars = nullptr, *d_part_e = nullptr,
           *d_coef = nullptr, *d_energy = nullptr, *d_mass = nullptr, *d_out = nullptr;
    int4* d_axis = nullptr;
    int *d_exc_start = nullptr, *d_exc_J = nullptr, *d_goff = nullptr, *d_gidx = nullptr, *d_perm = nullptr;
    unsigned long long* d_exc_mask = nullptr;
    int hist_order = 0, extrap_order = 0;
    size_t out_bytes = 0;
    int *d_tcnt = nullptr, *d_tfirst = nullptr, *d_tseg = nullptr, *d_done = nullptr;
    // ---- dense path (allocated on first use)
    bool dense_ready = false;
    int dense_max_nt = 0;
    DenseSys* d_dsys = nullptr;
    unsigned short* d_dmat = nullptr;
    float* d_dinv = nullptr;
    float* d_dacc = nullptr;        // FP32 running sums of the diagonal blocks during the factorization
    size_t dense_dwords = 0;
    double *d_db = nullptr, *d_dx = nullptr, *d_dres = nullptr, *d_dy = nullptr, *d_drel = nullptr;
    double4* d_dmu = nullptr;
    double* d_rout = nullptr;   // responses in the caller's order, [4][n][3]
    int *d_dactive = nullptr, *d_dstatus = nullptr;
    size_t* d_tcol = nullptr;
    int4* d_seg = nullptr;
    int2* d_segw = nullptr;
    // ---- results of the last solve (per system)
    std::vector<int> iterations, converged;
    // pinned host copy of the stop count and the solver scalars, read in one round trip
    void* h_stage = nullptr;
    // set for the duration of a tdp_execute_device_noalloc call: no reorder, no graph capture, no timing
    bool noalloc = false;
    // ---- Ewald option (TDP_FLAG_EWALD): parameters, FFT plans, grids, neighbour tiles, scaled pairs
    bool ewald = false;
    double ew_alpha = 0.0, ew_box[3] = {0.0, 0.0, 0.0};
    double ew_r64 = 0.25;   // mixed precision: real-space pairs closer than this stay in FP64 (nm)
    int ew_K[3] = {0, 0, 0};
    int ew_radix[3][kEwMaxStages] = {};
    int ew_nst[3] = {0, 0, 0};
    double ew_qsum = 0.0;
    double2* d_tw64[3] = {nullptr, nullptr, nullptr};
    float2* d_tw32[3] = {nullptr, nullptr, nullptr};
    double* d_G64 = nullptr;
    float* d_G32 = nullptr;
    unsigned long long* d_qfix = nullptr;
    double2* d_cgrid = nullptr;
    double *d_phiq = nullptr, *d_phiv = nullptr;
    int *d_nbr = nullptr, *d_nbr_cnt = nullptr, *d_nbr_exc = nullptr;
    int *d_sc_off = nullptr, *d_sc_j = nullptr;
    double2* d_sc_s = nullptr;
    size_t cap_sc = 0;
    double *d_epart = nullptr, *d_eout = nullptr;
    double4* d_rs = nullptr;
    double4* d_rsw = nullptr;
    // stored real-space blocks of T (sliced ELL lists per tile row and wave, built by the E0 pass)
    bool ew_stored = false;
    int ew_cap32 = 0, ew_cap64 = 0;
    unsigned short* d_idx16 = nullptr;
    int *d_idx64 = nullptr, *d_w32 = nullptr, *d_w64 = nullptr, *d_need = nullptr;
    int ew_lds_tiles = 0;   // neighbour tiles the stored T v pass stages in LDS (0: gathers from HBM)
    int ew_lds_slots = 0;   // mixed: slot-table entries the T v pass stages in LDS
    int* d_slot_tab = nullptr;   // mixed: per FP32 list, the slots its entries name
    int* d_slot_cnt = nullptr;
    int ew_tcap = 0;
    float2* d_t32 = nullptr;
    double4* d_t64 = nullptr;
    unsigned *d_cand32 = nullptr, *d_cand64 = nullptr;   // the E0 pass's pair candidates
    int2* d_ccnt = nullptr;
    int64_t ew_entries[2] = {0, 0}, ew_slots[2] = {0, 0};   // stored entries and padded slots of the last build
    void* h_need = nullptr;   // pinned copy of d_need
    // the list build's counts are examined at the solver's first host check instead of right after
    // the build (ew_defer, set by the caller for one evaluation); ew_need_pending: h_need holds the
    // counts of a build not yet examined
    bool ew_defer = false, ew_need_pending = false;
    int last_max_iter = 0;   // largest iteration count of the previous Mutual solve
    // warm start of the Mutual solve from the previous evaluation's induced dipoles (caller's order)
    bool warm_start = false, have_prev = false;
    double* d_mu_prev = nullptr;
    std::vector<double> epsilon;
    // ---- timing
    bool timing = false;
    double matvec_ms = 0.0;
    int matvec_launches = 0;
    std::vector<float> matvec_each;   // every timed launch, ms
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
    std::string err;
};

namespace {

int fail(tdp_context* c, int code, const std::string& msg) {
    if (c) c->err = msg; else g_create_error = msg;
    return code;
}

#define TDP_HIP(ctx, call)                                                                              \
    do {                                                                                                \
        hipError_t e_ = (call);                                                                         \
        if (e_ != hipSuccess)                                                                           \
            return fail(ctx, TDP_ERR_HIP, std::string("HIP error in ") + #call + ": " + hipGetErrorString(e_)); \
    } while (0)

template <typename T>
int dalloc(tdp_context* c, T** ptr, size_t count) {
    void* p = nullptr;
    size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) return fail(c, TDP_ERR_HIP, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    // zero on the context's stream, so that later work on that stream is ordered after it
    e = hipMemsetAsync(p, 0, bytes, c->stream);
    if (e != hipSuccess) return fail(c, TDP_ERR_HIP, std::string("hipMemset failed: ") + hipGetErrorString(e));
    c->allocations.push_back(p);
    *ptr = reinterpret_cast<T*>(p);
    return TDP_OK;
}

// max that keeps a NaN (std::max drops a NaN second argument)
double nan_max(double a, double b) { return (std::isnan(a) || std::isnan(b)) ? NAN : std::max(a, b); }

// Validation that TholeDipoleForceImpl::initialize performs (TholeDipoleForceImpl.cpp:52-108),
// plus the inputs the reference math would dereference out of bounds.
int validate(tdp_context* c, const tdp_force_desc* d) {
    int n = d->num_particles;
    if (n < 0) return fail(c, TDP_ERR_INVALID_ARGUMENT, "TholeDipoleForce: negative particle count");
    if (d->nonbonded_method != TDP_NO_CUTOFF && d->nonbonded_method != TDP_PME)
        return fail(c, TDP_ERR_INVALID_ARGUMENT, "TholeDipoleForce: Illegal value for nonbonded method");
    if (d->polarization_type < 0 || d->polarization_type > 2)
        return fail(c, TDP_ERR_INVALID_ARGUMENT, "Polarization type not recognized.");
    if (d->nonbonded_method == TDP_PME) {
        const double* b = d->default_box;
        if (d->cutoff > 0.5 * b[0] || d->cutoff > 0.5 * b[4] || d->cutoff > 0.5 * b[8])
            return fail(c, TDP_ERR_INVALID_ARGUMENT,
                        "TholeDipoleForce: The cutoff distance cannot be greater than half the periodic box size.");
    }
    for (int i = 0; i < n; i++) {
        int at = d->axis_types[i], z = d->atom_z[i], x = d->atom_x[i], y = d->atom_y[i];
        if (at < 0 || at > 5) {
            char buf[320];
            snprintf(buf, sizeof buf,
                     "TholeDipoleForce: axis type=%d not currently handled - only axisTypes[ 0, 1, 2, 3, 4, 5] "
                     "(ZThenX, Bisector, Z-Bisect, ThreeFold, ZOnly, NoAxisType) currently handled.", at);
            return fail(c, TDP_ERR_INVALID_ARGUMENT, buf);
        }
        if (at != 5 && (z < 0 || z >= n))
            return fail(c, TDP_ERR_INVALID_ARGUMENT, "TholeDipoleForce: invalid z axis particle: " + std::to_string(z));
        if (at != 5 && at != 4 && (x < 0 || x >= n))
            return fail(c, TDP_ERR_INVALID_ARGUMENT, "TholeDipoleForce: invalid x axis particle: " + std::to_string(x));
        if ((at == 2 || at == 3) && (y < 0 || y >= n))
            return fail(c, TDP_ERR_INVALID_ARGUMENT, "TholeDipoleForce: invalid y axis particle: " + std::to_string(y));
        // The reference also dereferences these (applyRotationMatrix :556-568 for any axis type
        // with Z>=0, mapTorqueToForce :1104-1108 reads X even for ZOnly); out-of-range values
        // would read outside its particle array, so they are rejected here.
        if (z >= n || x >= n || y >= n)
            return fail(c, TDP_ERR_INVALID_ARGUMENT, "TholeDipoleForce: axis particle index out of range for particle " +
                                                         std::to_string(i));
        if (z >= 0 && z != i && at != 4 && x < 0)
            return fail(c, TDP_ERR_INVALID_ARGUMENT,
                        "TholeDipoleForce: particle " + std::to_string(i) + " has a z axis particle but no x axis particle");
        if (at == 4 && x < 0)
            return fail(c, TDP_ERR_INVALID_ARGUMENT,
                        "TholeDipoleForce: ZOnly particle " + std::to_string(i) + " needs a valid x axis particle index");
        if (y > -1 && at == 0 && (z < 0 || x < 0))
            return fail(c, TDP_ERR_INVALID_ARGUMENT, "TholeDipoleForce: invalid axis particles for particle " + std::to_string(i));
    }
    for (int i = 0; i < n; i++)
        for (int t = 0; t < 2; t++) {
            int len = d->covalent_offsets[4 * i + t + 1] - d->covalent_offsets[4 * i + t];
            if (len > 5)
                return fail(c, TDP_ERR_INVALID_ARGUMENT,
                            "TholeDipoleForce: covalent list " + std::to_string(t) + " of particle " + std::to_string(i) +
                                " has " + std::to_string(len) + " entries; the scale tables hold 5");
        }
    return TDP_OK;
}

void copy_particles(tdp_context* c, const tdp_force_desc* d) {
    int n = d->num_particles;
    c->charge.assign(d->charges, d->charges + n);
    c->dmol.assign(d->dipoles, d->dipoles + 3 * n);
    c->polarity.assign(d->polarizabilities, d->polarizabilities + n);
    c->thole.assign(d->thole, d->thole + n);
    c->axis_type.assign(d->axis_types, d->axis_types + n);
    c->atom_z.assign(d->atom_z, d->atom_z + n);
    c->atom_x.assign(d->atom_x, d->atom_x + n);
    c->atom_y.assign(d->atom_y, d->atom_y + n);
}

// setupScaleMaps/getScaleFactor (ReferenceTholeDipoleForce.cpp:609-644) -> the scaled pairs.
void build_exclusion_pairs(tdp_context* c, const tdp_force_desc* d) {
    c->exc_pairs = scaled_pairs(c->n, d->covalent_offsets, d->covalent_atoms);
}

// Exclusion tiles of the current internal order (tdp_tables.cpp).
void build_tiles(tdp_context* c) {
    ExclusionTiles t;
    exclusion_tiles(c->n, c->nblk, c->inv.data(), c->exc_pairs, t);
    c->exc_start = std::move(t.start);
    c->exc_J = std::move(t.J);
    c->exc_mask = std::move(t.mask);
}

// Gather lists for mapTorqueToForce (ReferenceTholeDipoleForce.cpp:1092-1115): each particle's
// contributions (to itself and its axis atoms) are summed in the reference's order: particle
// index ascending, and within one particle U, V, W, then I.  contrib slots and receiving atoms
// are internal indices.
void build_torque_gather(tdp_context* c) {
    const int n = c->n;
    struct Entry { int atom, ii, role, slot; };
    std::vector<Entry> es;
    for (int ii = 0; ii < n; ii++) {
        int at = c->axis_type[ii];
        if (at == 5) continue;
        int k = c->inv[ii];
        es.push_back({c->inv[c->atom_z[ii]], ii, 0, 4 * k + 1});
        if (at != 4) es.push_back({c->inv[c->atom_x[ii]], ii, 1, 4 * k + 2});
        if ((at == 2 || at == 3) && c->atom_y[ii] > -1) es.push_back({c->inv[c->atom_y[ii]], ii, 2, 4 * k + 3});
        es.push_back({k, ii, 3, 4 * k + 0});
    }
    std::stable_sort(es.begin(), es.end(), [](const Entry& a, const Entry& b) {
        if (a.atom != b.atom) return a.atom < b.atom;
        if (a.ii != b.ii) return a.ii < b.ii;
        return a.role < b.role;
    });
    const int npad = c->npad;
    c->goff.assign(npad + 1, 0);
    c->gidx.resize(es.size());
    for (size_t k = 0; k < es.size(); k++) {
        c->goff[es[k].atom + 1]++;
        c->gidx[k] = es[k].slot;
    }
    for (int a = 0; a < npad; a++) c->goff[a + 1] += c->goff[a];
}

// Replaces a device buffer with one of at least count elements.
template <typename T>
int dresize(tdp_context* c, T** ptr, size_t& cap, size_t count) {
    if (*ptr && count <= cap) return TDP_OK;
    if (*ptr) {
        auto it = std::find(c->allocations.begin(), c->allocations.end(), (void*)*ptr);
        if (it != c->allocations.end()) c->allocations.erase(it);
        TDP_HIP(c, hipFree(*ptr));
        *ptr = nullptr;
    }
    int rc = dalloc(c, ptr, std::max<size_t>(count, 1));
    if (rc) return rc;
    cap = std::max<size_t>(count, 1);
    return TDP_OK;
}

// Per-particle parameters in the internal order, with axis atoms renumbered.
int upload_particles(tdp_context* c) {
    c->graph_gen++;
    const int n = c->n, npad = c->npad;
    std::vector<double> alpha(npad, 0.0), inva(npad, 0.0), thole(npad, 0.0), charge(npad, 0.0), dmol(3 * npad, 0.0);
    std::vector<int4> axis(npad, make_int4(5, -1, -1, -1));
    std::vector<int> perm(npad, -1);
    auto ren = [c](int a) { return a >= 0 ? c->inv[a] : a; };
    bool has_perm = false;
    (void)n;
    for (int k = 0; k < npad; k++) {
        int i = c->perm[k];
        if (i < 0) continue;   // padding slot: no charge, dipole or polarizability
        perm[k] = i;
        alpha[k] = c->polarity[i];
        thole[k] = c->thole[i];
        charge[k] = c->charge[i];
        inva[k] = c->thole[i] != 0.0 ? 1.0 / c->thole[i] : 0.0;
        for (int d = 0; d < 3; d++) {
            dmol[3 * k + d] = c->dmol[3 * i + d];
            if (dmol[3 * k + d] != 0.0) has_perm = true;
        }
        axis[k] = make_int4(c->axis_type[i], ren(c->atom_z[i]), ren(c->atom_x[i]), ren(c->atom_y[i]));
    }
    TDP_HIP(c, hipMemcpyAsync(c->d_alpha, alpha.data(), npad * sizeof(double), hipMemcpyHostToDevice, c->stream));
    TDP_HIP(c, hipMemcpyAsync(c->d_inva, inva.data(), npad * sizeof(double), hipMemcpyHostToDevice, c->stream));
    TDP_HIP(c, hipMemcpyAsync(c->d_thole, thole.data(), npad * sizeof(double), hipMemcpyHostToDevice, c->stream));
    TDP_HIP(c, hipMemcpyAsync(c->d_charge, charge.data(), npad * sizeof(double), hipMemcpyHostToDevice, c->stream));
    TDP_HIP(c, hipMemcpyAsync(c->d_dmol, dmol.data(), 3 * npad * sizeof(double), hipMemcpyHostToDevice, c->stream));
    TDP_HIP(c, hipMemcpyAsync(c->d_axis, axis.data(), npad * sizeof(int4), hipMemcpyHostToDevice, c->stream));
    TDP_HIP(c, hipMemcpyAsync(c->d_perm, perm.data(), npad * sizeof(int), hipMemcpyHostToDevice, c->stream));
    TDP_HIP(c, hipStreamSynchronize(c->stream));  // the host arrays go out of scope
    c->has_perm = has_perm;
    return TDP_OK;
}

int ew_upload_scaled(tdp_context* c);

// Exclusion tiles and torque-gather lists of the current order.
int upload_tables(tdp_context* c) {
    c->graph_gen++;   // device tables may move: captured iteration graphs are stale
    build_tiles(c);
    build_torque_gather(c);
    const size_t ntiles = c->exc_J.size();
    int rc;
    if ((rc = dresize(c, &c->d_exc_J, c->cap_exc_J, ntiles)) ||
        (rc = dresize(c, &c->d_exc_mask, c->cap_exc_mask, ntiles * 3 * 64)) ||
        (rc = dresize(c, &c->d_gidx, c->cap_gidx, c->gidx.size())))
        return rc;
    TDP_HIP(c, hipMemcpyAsync(c->d_exc_start, c->exc_start.data(), c->exc_start.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
    if (ntiles) {
        TDP_HIP(c, hipMemcpyAsync(c->d_exc_J, c->exc_J.data(), ntiles * sizeof(int), hipMemcpyHostToDevice, c->stream));
        TDP_HIP(c, hipMemcpyAsync(c->d_exc_mask, c->exc_mask.data(), c->exc_mask.size() * sizeof(unsigned long long),
                                  hipMemcpyHostToDevice, c->stream));
    }
    TDP_HIP(c, hipMemcpyAsync(c->d_goff, c->goff.data(), c->goff.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
    if (!c->gidx.empty())
        TDP_HIP(c, hipMemcpyAsync(c->d_gidx, c->gidx.data(), c->gidx.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
    TDP_HIP(c, hipStreamSynchronize(c->stream));
    if (c->ewald) return ew_upload_scaled(c);
    return TDP_OK;
}

// Internal spatial order: recursive bisection.  The particles (PME: wrapped into the box) are
// split along the longest extent of their bounding box into a part holding floor(t/2) whole tiles
// of 64 and the rest, t = ceil(count/64), until each part is one tile; a tile is therefore a compact,
// nearly cubic group, and a j tile is either near or far from a whole i tile, so the Thole
// exponential and the masked exclusion path run for few tiles.  Splits order particles by
// (coordinate, index), so the order is deterministic; inside a tile particles are in index order.
// The order only changes which thread evaluates which pair; results are returned in the caller's order.
void compute_order(tdp_context* c, const double* pos3, const double* box) {
    const int n = c->n;
    const bool pbc = c->method == TDP_PME && box;
    std::vector<double> w(3 * (size_t)n);
    for (int i = 0; i < n; i++)
        for (int d = 0; d < 3; d++) {
            double x = pos3[3 * i + d];
            if (pbc && box[4 * d] > 0.0) {
                double L = box[4 * d];
                x -= L * std::floor(x / L);
            }
            w[3 * (size_t)i + d] = std::isfinite(x) ? x : 0.0;
        }
    std::fill(c->perm.begin(), c->perm.end(), -1);
    struct Range { int b, e; };
    for (int sg = 0; sg < c->nseg; sg++) {
        const int p0 = c->seg_p0[sg], ns = c->seg_n[sg];
        std::vector<int> ids(ns);
        for (int k = 0; k < ns; k++) ids[k] = p0 + k;
        std::vector<Range> stack{{0, ns}};
        while (!stack.empty()) {
            Range r = stack.back();
            stack.pop_back();
            const int m = r.e - r.b;
            if (m <= 64) {
                std::sort(ids.begin() + r.b, ids.begin() + r.e);
                if (c->ewald && m == 64) {
                    // full tiles of the Ewald option in transposed order: the real-space passes give
                    // wave w the slots 8w..8w+7, which then hold the sorted positions w, w+8, ...,
                    // w+56 (an even sample of the tile, read as one 128-byte line per wave)
                    int t[64];
                    for (int k = 0; k < 64; k++) t[8 * (k & 7) + (k >> 3)] = ids[r.b + k];
                    std::copy(t, t + 64, ids.begin() + r.b);
                }
                continue;
            }
            double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
            for (int k = r.b; k < r.e; k++)
                for (int d = 0; d < 3; d++) {
                    double x = w[3 * (size_t)ids[k] + d];
                    lo[d] = std::min(lo[d], x);
                    hi[d] = std::max(hi[d], x);
                }
            int ax = 0;
            for (int d = 1; d < 3; d++)
                if (hi[d] - lo[d] > hi[ax] - lo[ax]) ax = d;
            const int tiles = (m + 63) / 64;
            const int left = (tiles / 2) * 64;
            std::nth_element(ids.begin() + r.b, ids.begin() + r.b + left, ids.begin() + r.e, [&](int a, int b2) {
                double xa = w[3 * (size_t)a + ax], xb = w[3 * (size_t)b2 + ax];
                return xa < xb || (xa == xb && a < b2);
            });
            stack.push_back({r.b + left, r.e});
            stack.push_back({r.b, r.b + left});
        }
        const int slot0 = 64 * c->seg_t0[sg];
        for (int k = 0; k < ns; k++) {
            c->perm[slot0 + k] = ids[k];
            c->inv[ids[k]] = slot0 + k;
        }
    }
}

int reorder(tdp_context* c, const double* host_pos3, const double* box) {
    compute_order(c, host_pos3, box);
    int rc = upload_particles(c);
    if (rc) return rc;
    rc = upload_tables(c);
    if (rc) return rc;
    c->ordered = true;
    c->evals_since_order = 0;
    return TDP_OK;
}

// ---------------------------------------------------------------- Ewald option helpers
// Smallest n >= n0 whose prime factors are 2, 3, 5 and 7 (the FFT sizes OpenMM accepts).
int legal_fft_size(int n0) {
    for (int n = std::max(n0, 1);; n++) {
        int m = n;
        for (int p : {2, 3, 5, 7})
            while (m % p == 0) m /= p;
        if (m == 1) return n;
    }
}

// alpha and grid of the PME method: setPMEParameters' values, or when alpha or the grid is 0, the
// closed form of OpenMM's NonbondedForceImpl::calcPMEParameters (the reference asks it for them,
// ReferenceTholeDipoleKernels.cpp:143-152): alpha = sqrt(-ln(2 tol))/cutoff,
// n_d = legal size >= max(6, ceil(2 alpha L_d / (3 tol^(1/5)))).
void pme_parameters(const tdp_context* c, double& a, int g[3]) {
    a = c->pme_alpha;
    for (int k = 0; k < 3; k++) g[k] = c->grid[k];
    if (g[0] == 0 || a == 0.0) {
        const double tol = c->ewald_tol;
        a = std::sqrt(-std::log(2.0 * tol)) / c->cutoff;
        const double L[3] = {c->default_box[0], c->default_box[4], c->default_box[8]};
        for (int k = 0; k < 3; k++)
            g[k] = legal_fft_size(std::max(6, (int)std::ceil(2 * a * L[k] / (3 * std::pow(tol, 0.2)))));
    }
}

// Radix stages of an FFT of length n (4s first, then 2, 3, 5, 7); false if another factor remains.
bool fft_plan(int n, int* r, int& nst) {
    nst = 0;
    int m = n;
    while (m % 4 == 0 && nst < kEwMaxStages) { r[nst++] = 4; m /= 4; }
    for (int p : {2, 3, 5, 7})
        while (m % p == 0 && nst < kEwMaxStages) { r[nst++] = p; m /= p; }
    return m == 1 && n > 1;
}

// |b(m)|^2 of the order-5 Euler exponential spline of one axis (Essmann et al. 1995, eq. 4.4):
// 1 / |sum_{k=0}^{3} M_5(k+1) e^{2 pi i m k / K}|^2, the vanishing denominator at m = K/2 replaced by
// the mean of its neighbours.
std::vector<double> bspline_moduli5(int K) {
    const double M5[4] = {1.0 / 24.0, 11.0 / 24.0, 11.0 / 24.0, 1.0 / 24.0};   // M_5(1..4)
    std::vector<double> mod(K);
    for (int m = 0; m < K; m++) {
        double re = 0.0, im = 0.0;
        for (int k = 0; k < 4; k++) {
            re += M5[k] * std::cos(2.0 * M_PI * m * k / K);
            im += M5[k] * std::sin(2.0 * M_PI * m * k / K);
        }
        mod[m] = re * re + im * im;
    }
    std::vector<double> fixed = mod;
    for (int m = 0; m < K; m++)
        if (mod[m] < 1e-7) fixed[m] = 0.5 * (mod[(m - 1 + K) % K] + mod[(m + 1) % K]);
    for (int m = 0; m < K; m++) fixed[m] = 1.0 / fixed[m];
    return fixed;
}

// Influence function of the box (when it changed): G(m) = exp(-pi^2 |m/L|^2 / a^2) / (pi V |m/L|^2)
// times the three B-spline moduli, G(0) = 0, uploaded in FP64 and FP32.
int ew_update_box(tdp_context* c, const double* box9) {
    const double L[3] = {box9[0], box9[4], box9[8]};
    if (L[0] == c->ew_box[0] && L[1] == c->ew_box[1] && L[2] == c->ew_box[2]) return TDP_OK;
    const int K0 = c->ew_K[0], K1 = c->ew_K[1], K2 = c->ew_K[2];
    const std::vector<double> b0 = bspline_moduli5(K0), b1 = bspline_moduli5(K1), b2 = bspline_moduli5(K2);
    const size_t ng = (size_t)K0 * K1 * K2;
    std::vector<double> G(ng);
    std::vector<float> G32(ng);
    const double V = L[0] * L[1] * L[2], f = M_PI * M_PI / (c->ew_alpha * c->ew_alpha);
    for (int x = 0; x < K0; x++)
        for (int y = 0; y < K1; y++)
            for (int z = 0; z < K2; z++) {
                const int mx = x <= K0 / 2 ? x : x - K0, my = y <= K1 / 2 ? y : y - K1, mz = z <= K2 / 2 ? z : z - K2;
                const double kx = mx / L[0], ky = my / L[1], kz = mz / L[2];
                const double m2 = kx * kx + ky * ky + kz * kz;
                const size_t g = ((size_t)x * K1 + y) * K2 + z;
                G[g] = (x == 0 && y == 0 && z == 0) ? 0.0 : std::exp(-f * m2) / (M_PI * V * m2) * b0[x] * b1[y] * b2[z];
                G32[g] = (float)G[g];
            }
    TDP_HIP(c, hipStreamSynchronize(c->stream));   // no kernel may still read the old table
    TDP_HIP(c, hipMemcpyAsync(c->d_G64, G.data(), ng * sizeof(double), hipMemcpyHostToDevice, c->stream));
    TDP_HIP(c, hipMemcpyAsync(c->d_G32, G32.data(), ng * sizeof(float), hipMemcpyHostToDevice, c->stream));
    TDP_HIP(c, hipStreamSynchronize(c->stream));
    for (int k = 0; k < 3; k++) c->ew_box[k] = L[k];
    c->graph_gen++;
    return TDP_OK;
}

// Per-particle terms of the Ewald option: the net charge (background energy) and the FP64 radius of
// the mixed real-space passes, where a u^3 reaches 4 for the strongest damping (u = r / (a_i a_j)),
// at least 0.25 nm.  The radius splits the stored T lists: the next build uses it.
void ew_particle_terms(tdp_context* c) {
    c->ew_qsum = 0.0;
    for (double q : c->charge) c->ew_qsum += q;
    double amax = 0.0, amin = 1e300;
    for (double a : c->thole) {
        amax = std::max(amax, a);
        if (a > 0.0) amin = std::min(amin, a);
    }
    c->ew_r64 = 0.25;
    if (amax > 0.0) c->ew_r64 = std::max(0.25, std::cbrt(4.0 / amin) * amax * amax);
}

// (Re)allocates the stored T lists for the given per-wave capacities.  Returns TDP_OK with
// ew_stored false when the device cannot hold them (the T v passes then stay matrix-free).
int ew_alloc_lists(tdp_context* c, int cap32, int cap64, int tcap) {
    auto release = [c](void* p) {
        if (!p) return;
        auto it = std::find(c->allocations.begin(), c->allocations.end(), p);
        if (it != c->allocations.end()) c->allocations.erase(it);
        hipFree(p);
    };
    release(c->d_idx16); release(c->d_t32);
    release(c->d_idx64); release(c->d_t64);
    release(c->d_cand32); release(c->d_cand64); release(c->d_ccnt); release(c->d_slot_tab);
    c->d_slot_tab = nullptr;
    c->d_idx16 = nullptr;
    c->d_idx64 = nullptr;
    c->d_t32 = nullptr;
    c->d_t64 = nullptr;
    c->d_cand32 = c->d_cand64 = nullptr;
    c->d_ccnt = nullptr;
    c->ew_stored = false;
    // capacities in whole candidate blocks of 8 (FP32 entries are stored in pairs)
    cap32 = (cap32 + 7) / 8 * 8;
    cap64 = (cap64 + 7) / 8 * 8;
    const bool mixed = c->precision == TDP_PRECISION_MIXED;
    const size_t rows = (size_t)c->nblk * kEwWaves * 64;   // 64 slots per tile
    const size_t n32 = mixed ? rows * cap32 : 1, n64 = rows * cap64;
    tcap = mixed ? (tcap + 63) / 64 * 64 : 0;
    const size_t ntab = mixed ? (size_t)c->nblk * kEwWaves * tcap : 1;
    void* p[8] = {};
    const size_t bytes[8] = {n32 * sizeof(unsigned short), n32 * sizeof(float2), n64 * sizeof(int),
                             n64 * sizeof(double4), n32 * sizeof(unsigned), n64 * sizeof(unsigned), rows * sizeof(int2),
                             ntab * sizeof(int)};
    for (int k = 0; k < 8; k++)
        if (hipMalloc(&p[k], bytes[k]) != hipSuccess) {
            for (int q = 0; q < k; q++) hipFree(p[q]);
            (void)hipGetLastError();
            return TDP_OK;   // matrix-free fallback
        }
    for (int k = 0; k < 8; k++) c->allocations.push_back(p[k]);
    c->d_slot_tab = (int*)p[7];
    c->ew_tcap = tcap;
    c->d_idx16 = (unsigned short*)p[0];
    c->d_t32 = (float2*)p[1];
    c->d_idx64 = (int*)p[2];
    c->d_t64 = (double4*)p[3];
    c->d_cand32 = (unsigned*)p[4];
    c->d_cand64 = (unsigned*)p[5];
    c->d_ccnt = (int2*)p[6];
    c->ew_cap32 = mixed ? cap32 : 0;
    c->ew_cap64 = cap64;
    c->ew_stored = true;
    c->graph_gen++;
    return TDP_OK;
}

EwLaunch ew_launch(const tdp_context* c, const double* box9) {
    EwLaunch E;
    E.npad = c->npad;
    E.nblk = c->nblk;
    E.tcnt = c->d_tcnt;
    for (int k = 0; k < 3; k++) {
        E.L[k] = box9[4 * k];
        E.iL[k] = 1.0 / E.L[k];
    }
    E.rc2 = c->cutoff * c->cutoff;
    const float cm = (float)(c->cutoff * 1.0001 + 1e-4);
    E.cut2_tiles = cm * cm;
    const double a = c->ew_alpha, sp = 1.0 / std::sqrt(M_PI);
    E.alpha = a;
    E.c1 = 2.0 * a * sp;
    E.c2 = 4.0 * a * a * a * sp;
    E.c3 = 8.0 * a * a * a * a * a * sp;
    E.self_field = 4.0 * a * a * a * sp / 3.0;
    E.self_energy = -a * sp;
    E.tb_c = c->tb_c;
    E.tb_h = c->tb_h;
    E.nbr = c->d_nbr;
    E.nbr_cnt = c->d_nbr_cnt;
    E.nbr_exc = c->d_nbr_exc;
    E.exc_start = c->d_exc_start;
    E.exc_J = c->d_exc_J;
    E.exc_mask = c->d_exc_mask;
    E.sc_off = c->d_sc_off;
    E.sc_j = c->d_sc_j;
    E.sc_s = c->d_sc_s;
    for (int k = 0; k < 3; k++) {
        E.K[k] = c->ew_K[k];
        E.ax[k].n = c->ew_K[k];
        E.ax[k].nst = c->ew_nst[k];
        for (int q = 0; q < kEwMaxStages; q++) E.ax[k].r[q] = c->ew_radix[k][q];
        E.ax[k].tw64 = c->d_tw64[k];
        E.ax[k].tw32 = c->d_tw32[k];
    }
    E.G64 = c->d_G64;
    E.G32 = c->d_G32;
    E.qfix = c->d_qfix;
    E.cgrid = c->d_cgrid;
    E.phi_q = c->d_phiq;
    E.phi_v = c->d_phiv;
    E.rs = c->d_rs;
    E.rsw = c->d_rsw;
    E.cap32 = c->ew_cap32;
    E.cap64 = c->ew_cap64;
    E.idx16 = c->d_idx16;
    E.t32 = c->d_t32;
    E.idx64 = c->ew_stored ? c->d_idx64 : nullptr;
    E.t64 = c->d_t64;
    E.cand32 = c->d_cand32;
    E.cand64 = c->d_cand64;
    E.ccnt = c->d_ccnt;
    E.w32 = c->d_w32;
    E.w64 = c->d_w64;
    E.need = c->d_need;
    E.slot_tab = c->d_slot_tab;
    E.slot_cnt = c->d_slot_cnt;
    E.tcap = c->ew_tcap;
    E.lds_slots = c->ew_lds_slots;
    E.lds_tiles = c->ew_lds_tiles;
    E.tb_o = c->tb_o;
    E.posl = c->posl;
    E.mu32 = c->d_mu32;
    E.slot32 = c->d_slot32;
    // reciprocal space in FP64 in both precisions: FP32 grids and FFTs put 2.4e-8 e*nm into the
    // fixed point of the 36,864-site box, above the 1e-6 D (2.08e-8 e*nm) bar
    E.mixed = false;
    E.pairs32 = c->precision == TDP_PRECISION_MIXED;
    E.r64sq = c->ew_r64 * c->ew_r64;
    return E;
}

// Scaled partners of every slot (both members of each scaled pair), sorted by slot and partner.
int ew_upload_scaled(tdp_context* c) {
    struct E3 { int slot, partner; double m, is; };
    std::vector<E3> es;
    es.reserve(2 * c->exc_pairs.size());
    for (const ExcPair& e : c->exc_pairs) {
        const int a = c->inv[e.i], b = c->inv[e.j];
        es.push_back({a, b, e.m, e.is});
        es.push_back({b, a, e.m, e.is});
    }
    std::sort(es.begin(), es.end(), [](const E3& x, const E3& y) { return x.slot != y.slot ? x.slot < y.slot : x.partner < y.partner; });
    std::vector<int> off(c->npad + 1, 0), js(es.size());
    std::vector<double2> ss(es.size());
    for (size_t k = 0; k < es.size(); k++) {
        off[es[k].slot + 1]++;
        js[k] = es[k].partner;
        ss[k] = make_double2(es[k].m, es[k].is);
    }
    for (int k = 0; k < c->npad; k++) off[k + 1] += off[k];
    int rc;
    size_t cap_j = c->cap_sc, cap_s = c->cap_sc;
    if ((rc = dresize(c, &c->d_sc_j, cap_j, es.size())) || (rc = dresize(c, &c->d_sc_s, cap_s, es.size()))) return rc;
    c->cap_sc = std::max(cap_j, (size_t)1);
    TDP_HIP(c, hipMemcpyAsync(c->d_sc_off, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
    if (!es.empty()) {
        TDP_HIP(c, hipMemcpyAsync(c->d_sc_j, js.data(), js.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
        TDP_HIP(c, hipMemcpyAsync(c->d_sc_s, ss.data(), ss.size() * sizeof(double2), hipMemcpyHostToDevice, c->stream));
    }
    TDP_HIP(c, hipStreamSynchronize(c->stream));
    return TDP_OK;
}

PairLaunch pair_launch(const tdp_context* c, const double* box9) {
    PairLaunch L;
    L.npad = c->npad;
    L.tm.cnt = c->d_tcnt;
    L.tm.first = c->d_tfirst;
    L.tm.seg = c->d_tseg;
    L.tm.col = c->d_tcol;
    L.nblk = c->nblk;
    L.pbc = c->method == TDP_PME;
    L.perm = c->has_perm;
    L.box[0] = box9 ? box9[0] : 0.0;
    L.box[1] = box9 ? box9[4] : 0.0;
    L.box[2] = box9 ? box9[8] : 0.0;
    L.exc_start = c->d_exc_start;
    L.exc_J = c->d_exc_J;
    L.exc_mask = c->d_exc_mask;
    L.tb_c = c->tb_c;
    L.tb_h = c->tb_h;
    L.tb_o = c->tb_o;
    L.posl = c->posl;
    L.posl64 = c->posl64;
    L.tp = c->d_tp;
    L.nwork = c->nwork;
    L.work = c->d_work;
    L.row_start = c->d_row_start;
    L.slab = c->ntri * 64;
    return L;
}

SegLaunch seg_launch(const tdp_context* c) {
    SegLaunch G;
    G.nseg = c->nseg;
    G.ntile = c->nblk;
    G.npad = c->npad;
    G.tseg = c->d_tseg;
    G.seg = c->d_seg;
    G.segw = c->d_segw;
    G.done = c->d_done;
    return G;
}

hipEvent_t take_event(tdp_context* c) {
    if (!c->ev_pool.empty()) {
        hipEvent_t e = c->ev_pool.back();
        c->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    hipEventCreate(&e);
    return e;
}

void harvest_timing(tdp_context* c) {
    for (auto& pr : c->pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, pr.first, pr.second) == hipSuccess) {
            c->matvec_ms += ms;
            c->matvec_launches++;
            c->matvec_each.push_back(ms);
        }
        c->ev_pool.push_back(pr.first);
        c->ev_pool.push_back(pr.second);
    }
    c->pending.clear();
}

// T * v for the solver: v64 is mu or p; in mixed precision vec32 holds its FP32 copy.  With
// timing on, events bracket the pair kernel (not the reduction that follows it).
void matvec(tdp_context* c, const PairLaunch& L, const double4* v64, const double* scalars,
            const PcgQ& pq = PcgQ()) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const bool timed = c->timing && !c->noalloc;
    if (timed) {
        e0 = take_event(c);
        e1 = take_event(c);
        if (!c->ewald) hipEventRecord(e0, c->stream);
    }
    if (c->ewald) {   // reciprocal potential of v, then the real-space pass with the gather in its epilogue
        const double box9[9] = {L.box[0], 0, 0, 0, L.box[1], 0, 0, 0, L.box[2]};
        const EwLaunch E = ew_launch(c, box9);
        if (c->ew_stored) {
            // the reciprocal chain (spreading, FFTs) beside the stored-block T v pass, on two streams;
            // the epilogue, which gathers from its potential grid, waits for both, on the main
            // stream.  In mixed precision the chain, whose passes run in the pass's leftover
            // resources, takes the main stream and the pass the other (4.00 vs 4.05 ms per
            // evaluation at 36,864 sites); in FP64 mode the pass (2.7x longer) ends last by far and
            // keeps the main stream.
            const bool mixed = c->precision == TDP_PRECISION_MIXED;
            hipStream_t ss = mixed ? c->stream2 : c->stream;   // the T v pass
            hipStream_t sr = mixed ? c->stream : c->stream2;   // the chain
            hipEventRecord(c->ev_fork, c->stream);
            hipStreamWaitEvent(c->stream2, c->ev_fork, 0);
            launch_ew_recip(sr, E, c->posa, nullptr, v64, E.phi_v, scalars, true);
            if (timed) hipEventRecord(e0, ss);
            launch_ew_spmv(ss, E, v64, mixed ? c->vec32 : nullptr, scalars);
            if (timed) hipEventRecord(e1, ss);
            hipEventRecord(c->ev_join, c->stream2);
            hipStreamWaitEvent(c->stream, c->ev_join, 0);
            launch_ew_tmu_epilogue(c->stream, E, c->posa, c->d_inva, c->d_charge, v64, scalars, c->Emv, pq);
        } else {
            launch_ew_recip(c->stream, E, c->posa, nullptr, v64, E.phi_v, scalars);
            if (timed) hipEventRecord(e0, c->stream);
            launch_ew_pairs(c->stream, E, kEwTmu, false, c->posa, c->d_inva, c->d_charge, v64, nullptr, scalars, c->Emv,
                            pq, nullptr, nullptr, nullptr);
            if (timed) hipEventRecord(e1, c->stream);
        }
        if (timed) c->pending.push_back({e0, e1});
        return;
    }
    double4* pi = reinterpret_cast<double4*>(c->d_pbuf);
    double* pj = reinterpret_cast<double*>(pi + (size_t)c->nwork * 64);
    const bool mixed = c->precision == TDP_PRECISION_MIXED;
    if (mixed)
        launch_sym_matvec_mixed(c->stream, L, c->posa, v64, c->vec32, c->d_inva, scalars, c->d_slab32);
    else
        launch_sym_matvec64(c->stream, L, c->posa, v64, c->d_inva, scalars, pi, pj);
    if (timed) {
        hipEventRecord(e1, c->stream);
        c->pending.push_back({e0, e1});
    }
    launch_col_reduce(c->stream, L, pi, pj, scalars, c->Emv, nullptr, mixed ? c->d_slab32 : nullptr, pq);
}

int check_box(tdp_context* c, const double* box) {
    if (c->method != TDP_PME) return TDP_OK;
    if (!box) return fail(c, TDP_ERR_INVALID_ARGUMENT, "a periodic box is required for PME");
    double minAllowed = 1.999999 * c->cutoff;  // ReferenceTholeDipoleKernels.cpp:170-173
    if (box[0] < minAllowed || box[4] < minAllowed || box[8] < minAllowed)
        return fail(c, TDP_ERR_BOX_TOO_SMALL, "The periodic box size has decreased to less than twice the nonbonded cutoff.");
    if (c->ewald && (box[1] != 0.0 || box[2] != 0.0 || box[3] != 0.0 || box[5] != 0.0 || box[6] != 0.0 || box[7] != 0.0))
        return fail(c, TDP_ERR_INVALID_ARGUMENT, "the Ewald option needs a rectangular box");
    if (c->ewald && (2.0 * c->cutoff > box[0] || 2.0 * c->cutoff > box[4] || 2.0 * c->cutoff > box[8]))
        return fail(c, TDP_ERR_BOX_TOO_SMALL, "the Ewald option needs a cutoff of at most half the box");
    return TDP_OK;
}

// Pinned staging buffer of a context, allocated once at create: the download area (the stop count,
// then the solver scalars of every system) and the upload area (the solver scalars, then the stop
// count and a zero).  Every copy through it is followed by a stream synchronisation before the host
// touches it again.
size_t stage_download_doubles(const tdp_context* c) { return (size_t)SC_COUNT * c->nseg + 1; }
size_t stage_upload_doubles(const tdp_context* c) { return (size_t)SC_COUNT * c->nseg + 2; }
double* stage_upload(tdp_context* c) { return static_cast<double*>(c->h_stage) + stage_download_doubles(c); }

// Copies the solver scalars to the pinned staging buffer (one copy, one host round trip); done is
// the number of systems whose stop flag is set.  ahead (may be null) enqueues work behind the copy
// before the host waits, for the copy alone (an event).
int stage_solver_state(tdp_context* c, int& done, const std::function<int()>* ahead = nullptr) {
    double* h = static_cast<double*>(c->h_stage);
    TDP_HIP(c, hipMemcpyAsync(h + 1, c->d_scalars, sizeof(double) * SC_COUNT * c->nseg, hipMemcpyDeviceToHost,
                              c->stream));
    if (ahead && c->ev_stage) {
        TDP_HIP(c, hipEventRecord(c->ev_stage, c->stream));
        int rc = (*ahead)();
        if (rc) return rc;
        TDP_HIP(c, hipEventSynchronize(c->ev_stage));
    } else {
        TDP_HIP(c, hipStreamSynchronize(c->stream));
    }
    done = 0;
    for (int g = 0; g < c->nseg; g++) done += h[1 + (size_t)SC_COUNT * g + SC_DONE] != 0.0 ? 1 : 0;
    return TDP_OK;
}

// Per-system solver results from the device scalars; staged: stage_solver_state has just copied
// them (no launch since).
int read_solver_results(tdp_context* c, bool jacobi, bool staged = false) {
    if (!staged) {
        int unused;
        int rc = stage_solver_state(c, unused);
        if (rc) return rc;
    }
    const double* sc = static_cast<const double*>(c->h_stage) + 1;
    for (int g = 0; g < c->nseg; g++) {
        const double* x = &sc[(size_t)SC_COUNT * g];
        c->iterations[g] = (int)x[SC_ITER];
        c->epsilon[g] = x[SC_EPS];
        c->converged[g] = c->epsilon[g] < c->seg_target[g] ? 1 : 0;
        if (jacobi && c->iterations[g] == 0) c->epsilon[g] = 1.0;
    }
    return TDP_OK;
}

// Starts a Mutual solve: per-system iteration limit and target, epsilon 1, and for the Jacobi
// iteration the reference's loop condition (:829-832) as the initial stop flag.
int start_solve_host(tdp_context* c, bool jacobi) {
    double* sc = stage_upload(c);
    const size_t nsc = (size_t)SC_COUNT * c->nseg;
    std::fill(sc, sc + nsc, 0.0);
    int done = 0;
    for (int g = 0; g < c->nseg; g++) {
        double* x = &sc[(size_t)SC_COUNT * g];
        x[SC_MAXIT] = c->seg_maxit[g];
        x[SC_TARGET] = c->seg_target[g];
        x[SC_EPS] = 1.0;
        if (jacobi && (c->seg_maxit[g] <= 0 || !(1.0 > c->seg_target[g]))) {
            x[SC_DONE] = 1.0;
            done++;
        }
    }
    if (!jacobi) done = 0;   // the CG start kernel sets every system's flag afresh
    int* counts = reinterpret_cast<int*>(sc + nsc);
    counts[0] = done;
    counts[1] = 0;
    return TDP_OK;
}

// The copies of the staged start values to the device (launches only: the values are read from
// the pinned staging buffer when the copies run).
int start_solve_copies(tdp_context* c) {
    double* sc = stage_upload(c);
    const size_t nsc = (size_t)SC_COUNT * c->nseg;
    TDP_HIP(c, hipMemcpyAsync(c->d_scalars, sc, nsc * sizeof(double), hipMemcpyHostToDevice, c->stream));
    TDP_HIP(c, hipMemcpyAsync(c->d_done, sc + nsc, 2 * sizeof(int), hipMemcpyHostToDevice, c->stream));
    return TDP_OK;
}

int start_solve(tdp_context* c, bool jacobi) {
    int rc = start_solve_host(c, jacobi);
    return rc ? rc : start_solve_copies(c);
}


// setup() and the permanent field (ReferenceTholeDipoleForce.cpp:1117-1155, :735-746) from
// Internal status of a solve whose list build was examined late and missed pairs (never returned
// to a caller: the evaluation runs again).
constexpr int kEwRedo = -1000;

// The counts of the last list build (h_need, copied after it): stored entries and slots, the
// neighbour tiles the T v pass stages in LDS, and whether the lists held every pair.  Returns true
// when they did; else either the lists' indices cannot address the rows' neighbour tiles
// (ew_stored is now false: matrix-free T v passes from here on) or c32, c64 are the capacities a
// new build needs.
bool ew_take_need(tdp_context* c, int& c32, int& c64, int& tcap) {
    const int* hn = static_cast<const int*>(c->h_need);
    c->ew_need_pending = false;
    for (int k = 0; k < 2; k++) {
        c->ew_entries[k] = hn[2 + k];
        c->ew_slots[k] = hn[4 + k];
    }
    const bool mixed_lists = c->precision == TDP_PRECISION_MIXED;
    if (hn[6] > kEwMaxLocalTiles || (mixed_lists && hn[7] > kEwMaxSlots)) {
        // the list indices (FP32 lists, and the FP64 lists of FP64 mode) cannot address this many
        // neighbour tiles or slots
        c->ew_stored = false;
        c->graph_gen++;
        return false;
    }
    {   // LDS copies of the T v pass (24 B per slot, 64 KB at most): the FP32 lists' slot tables
        // (mixed), the wave's slots of up to lds_tiles tiles (FP64 mode; and the force pass's tile
        // shifts).  A pass launched with fewer stages a longer row from HBM instead.
        int lt = (hn[6] + 7) / 8 * 8;
        if ((size_t)lt * 8 * kEwSlotLds > 65536) lt = 0;
        int ls = mixed_lists ? (hn[7] + 7) / 8 * 8 : 0;
        if ((size_t)ls * kEwSlotLds > 65536) ls = 0;
        if (lt != c->ew_lds_tiles || ls != c->ew_lds_slots) {
            c->ew_lds_tiles = lt;
            c->ew_lds_slots = ls;
            c->graph_gen++;
        }
    }
    if ((!mixed_lists || (hn[0] <= c->ew_cap32 && hn[7] <= c->ew_tcap)) && hn[1] <= c->ew_cap64) return true;
    c32 = mixed_lists ? std::max(c->ew_cap32, hn[0] + hn[0] / 4 + 8) : 0;
    c64 = std::max(c->ew_cap64, hn[1] + hn[1] / 4 + 8);
    tcap = mixed_lists ? std::max(c->ew_tcap, hn[7] + hn[7] / 4 + 64) : 0;
    return false;
}

// positions already in d_pos3 (caller's order): internal order, lab-frame dipoles, E0 and
// mu = alpha E0.  host_pos3 (may be NULL) are the same positions on the host, used when the
// internal order is recomputed.
// The host part of prepare_field: the internal order, recomputed every kReorderInterval evaluations.
int prepare_order(tdp_context* c, const double* box, const double* host_pos3) {
    hipStream_t s = c->stream;
    const int n = c->n;
    // (an allocation-free call keeps the current order: recomputing it needs host work and may grow
    // the exclusion tables)
    if (!c->noalloc && (!c->ordered || c->evals_since_order >= kReorderInterval)) {
        std::vector<double> tmp;
        const double* hp = host_pos3;
        if (!hp) {
            tmp.resize(3 * (size_t)n);
            TDP_HIP(c, hipMemcpyAsync(tmp.data(), c->d_pos3, 3 * n * sizeof(double), hipMemcpyDeviceToHost, s));
            hp = tmp.data();
        }
        TDP_HIP(c, hipStreamSynchronize(s));  // no kernel may still read the tables being replaced
        int rc = reorder(c, hp, box);
        if (rc) return rc;
    }
    c->evals_since_order++;
    return TDP_OK;
}

// The launches of prepare_field without the Ewald option: positions in the internal order, tile
// boxes and pair classes, lab-frame dipoles, E0 and mu = alpha E0.
void enqueue_field(tdp_context* c, const PairLaunch& L, const SegLaunch& G) {
    hipStream_t s = c->stream;
    const int npad = c->npad;
    const bool mixed = c->precision == TDP_PRECISION_MIXED;
    launch_load_positions(s, npad, c->d_pos3, c->d_perm, c->d_thole, c->posa);
    launch_tile_bounds(s, c->d_tcnt, c->nblk, c->posa, L.box, L.pbc, c->tb_c, c->tb_h, c->tb_o, c->posl, c->posl64);
    launch_tile_pairs(s, L, c->d_tp);
    launch_rotate(s, npad, c->posa, c->d_dmol, c->d_charge, c->d_axis, c->qd);
    double4* pi = reinterpret_cast<double4*>(c->d_pbuf);
    launch_sym_field(s, L, c->posa, c->qd, c->d_inva, pi, reinterpret_cast<double*>(pi + (size_t)c->nwork * 64), c->Emv);
    launch_field_init(s, G, c->Emv, c->d_alpha, c->E0, c->mu, mixed ? c->vec32 : nullptr);
}

int prepare_field(tdp_context* c, const double* box, const double* host_pos3) {
    hipStream_t s = c->stream;
    const int npad = c->npad;
    int rc0 = prepare_order(c, box, host_pos3);
    if (rc0) return rc0;
    PairLaunch L = pair_launch(c, box);
    SegLaunch G = seg_launch(c);
    const bool mixed = c->precision == TDP_PRECISION_MIXED;
    if (!c->ewald) {
        enqueue_field(c, L, G);
        return TDP_OK;
    }
    launch_load_positions(s, npad, c->d_pos3, c->d_perm, c->d_thole, c->posa);
    launch_tile_bounds(s, c->d_tcnt, c->nblk, c->posa, L.box, L.pbc, c->tb_c, c->tb_h, c->tb_o, c->posl, c->posl64);
    {
        // neighbour tiles, the charges' potential grid, E0 = real space + reciprocal + scaled pairs
        int rc = ew_update_box(c, box);
        if (rc) return rc;
        const EwLaunch E = ew_launch(c, box);
        launch_rotate(s, npad, c->posa, c->d_dmol, c->d_charge, c->d_axis, c->qd);
        launch_ew_neighbors(s, E);
        // the charges' potential grid on stream2 beside the E0 pass, joined before its epilogue
        TDP_HIP(c, hipEventRecord(c->ev_fork, s));
        TDP_HIP(c, hipStreamWaitEvent(c->stream2, c->ev_fork, 0));
        launch_ew_recip(c->stream2, E, c->posa, c->d_charge, nullptr, E.phi_q, nullptr);
        TDP_HIP(c, hipEventRecord(c->ev_join, c->stream2));
        for (int attempt = 0;; attempt++) {
            const EwLaunch Eb = ew_launch(c, box);
            if (c->ew_stored) TDP_HIP(c, hipMemsetAsync(c->d_need, 0, 8 * sizeof(int), s));
            if (c->ew_stored) {
                launch_ew_e0_stored(s, Eb, c->posa, c->d_inva, c->d_charge, c->Emv, c->d_epart, c->ev_join);
            } else {
                TDP_HIP(c, hipStreamWaitEvent(s, c->ev_join, 0));
                launch_ew_pairs(s, Eb, kEwE0, false, c->posa, c->d_inva, c->d_charge, nullptr, nullptr, nullptr, c->Emv,
                                PcgQ(), c->d_epart, nullptr, nullptr);
            }
            if (!c->ew_stored) break;
            // the lists must have held every pair: else grow them and build again
            TDP_HIP(c, hipMemcpyAsync(c->h_need, c->d_need, 8 * sizeof(int), hipMemcpyDeviceToHost, s));
            if (c->ew_defer) {   // examined at the solver's first host check (ew_take_need)
                c->ew_need_pending = true;
                break;
            }
            TDP_HIP(c, hipStreamSynchronize(s));
            int c32 = 0, c64 = 0, tc = 0;
            if (ew_take_need(c, c32, c64, tc)) break;
            if (!c->ew_stored) break;   // matrix-free T v passes from now on (this pass's E0 is complete)
            if (attempt > 0) return fail(c, TDP_ERR_HIP, "Ewald option: the stored T lists could not be sized");
            if ((rc = ew_alloc_lists(c, c32, c64, tc))) return rc;
        }
        launch_field_init(s, G, c->Emv, c->d_alpha, c->E0, c->mu, mixed ? c->vec32 : nullptr);
        return TDP_OK;
    }
    return TDP_OK;
}

// Mutual solve of every system from the current E0 and mu (ReferenceTholeDipoleForce.cpp:820-851).
// Enqueues n solver iterations, as a replay of a captured graph of them when possible (the
// iteration kernels then go to the device as one submission).  Events for per-launch timing are
// not captured: with timing on, the launches are enqueued one by one.
// Replays the captured launches of `enqueue` (which returns a status) for (tag, key, p0, p1) under
// the current tables, box and precision, capturing them on first use.  Callers use it only where
// every launch's arguments are fixed by those: no host decision inside depends on device results.
template <class F>
int replay_or_capture(tdp_context* c, int tag, long long key, const void* p0, const void* p1, const double* box3,
                      F&& enqueue) {
    hipStream_t s = c->stream;
    hipGraphExec_t ex = nullptr;
    for (size_t k = 0; k < c->seq_graphs.size();) {
        auto& g = c->seq_graphs[k];
        if (g.gen != c->graph_gen || g.box[0] != box3[0] || g.box[1] != box3[1] || g.box[2] != box3[2]) {
            hipGraphExecDestroy(g.exec);
            c->seq_graphs.erase(c->seq_graphs.begin() + k);
            continue;
        }
        if (g.tag == tag && g.key == key && g.p0 == p0 && g.p1 == p1 && g.precision == c->precision) ex = g.exec;
        k++;
    }
    if (!ex) {
        hipGraph_t graph = nullptr;
        TDP_HIP(c, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        const int rc = enqueue();
        const hipError_t ee = hipStreamEndCapture(s, &graph);
        if (rc) {
            if (graph) hipGraphDestroy(graph);
            return rc;
        }
        if (ee != hipSuccess) return fail(c, TDP_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ee));
        hipError_t e = hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0);
        hipGraphDestroy(graph);
        if (e != hipSuccess) return fail(c, TDP_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        c->seq_graphs.push_back({tag, key, p0, p1, c->graph_gen, c->precision, {box3[0], box3[1], box3[2]}, ex});
    }
    TDP_HIP(c, hipGraphLaunch(ex, s));
    return TDP_OK;
}

// Enqueues n solver iterations (launches only).
void enqueue_iterations(tdp_context* c, const PairLaunch& L, const SegLaunch& G, bool jacobi, int n) {
    hipStream_t s = c->stream;
    const bool mixed = c->precision == TDP_PRECISION_MIXED;
    for (int k = 0; k < n; k++) {
        if (jacobi) {
            matvec(c, L, c->mu, c->d_scalars);
            launch_jacobi_update(s, G, c->Emv, c->E0, c->d_alpha, c->mu, mixed ? c->vec32 : nullptr, c->d_scalars,
                                 c->d_red);
        } else {
            PcgQ pq;
            pq.p = c->p;
            pq.alpha = c->d_alpha;
            pq.q = c->q;
            pq.red = c->d_red + 2 * (size_t)c->nblk;
            matvec(c, L, c->p, c->d_scalars, pq);
            launch_pcg_iteration_tail(s, G, c->d_alpha, c->mu, c->r, c->z, c->p, mixed ? c->vec32 : nullptr, c->q,
                                      c->d_red, c->d_scalars);
        }
    }
}

int run_iterations(tdp_context* c, const PairLaunch& L, const SegLaunch& G, bool jacobi, int n) {
    hipStream_t s = c->stream;
    auto enqueue = [&]() { enqueue_iterations(c, L, G, jacobi, n); };
    if (!c->use_graphs || c->timing || c->noalloc) {
        enqueue();
        return TDP_OK;
    }
    hipGraphExec_t ex = nullptr;
    for (size_t k = 0; k < c->graphs.size();) {
        auto& g = c->graphs[k];
        if (g.gen != c->graph_gen || g.box[0] != L.box[0] || g.box[1] != L.box[1] || g.box[2] != L.box[2]) {
            hipGraphExecDestroy(g.exec);
            c->graphs.erase(c->graphs.begin() + k);
            continue;
        }
        if (g.iters == n && g.jacobi == jacobi && g.precision == c->precision) ex = g.exec;
        k++;
    }
    if (!ex) {
        hipGraph_t graph = nullptr;
        TDP_HIP(c, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        enqueue();
        TDP_HIP(c, hipStreamEndCapture(s, &graph));
        hipError_t e = hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0);
        hipGraphDestroy(graph);
        if (e != hipSuccess) return fail(c, TDP_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        c->graphs.push_back({n, jacobi, c->precision, c->graph_gen, {L.box[0], L.box[1], L.box[2]}, ex});
    }
    TDP_HIP(c, hipGraphLaunch(ex, s));
    return TDP_OK;
}

// ahead (may be null): the rest of the evaluation, enqueued behind the first chunk of iterations
// before the host check; *ahead_stands tells whether that chunk finished the solve (else the caller
// enqueues the rest again).  With the list build's counts pending (ew_defer), the check examines
// them too and returns kEwRedo when the lists missed pairs: the evaluation must run again.
int solve_mutual(tdp_context* c, const double* box, const std::function<int()>* ahead = nullptr,
                 bool* ahead_stands = nullptr) {
    hipStream_t s = c->stream;
    PairLaunch L = pair_launch(c, box);
    SegLaunch G = seg_launch(c);
    const bool mixed = c->precision == TDP_PRECISION_MIXED;
    // every system iterates until its own stop flag is set; a launch skips the tiles and work
    // items of stopped systems.  The host checks the count of stopped systems every chunk.
    const bool jacobi = c->solver == TDP_SOLVER_JACOBI;
    // Iterations enqueued between host checks of the stop count.  The first chunk is the largest
    // iteration count of the previous solve of this context (repeated evaluations of similar
    // configurations stop after about as many), later chunks are short; launches enqueued after
    // every system has stopped return at once.
    const int chunk = c->n > 20000 ? 2 : 4;
    const int first_chunk = c->last_max_iter > 0 ? c->last_max_iter : (c->n > 20000 ? 4 : 16);
    int rc = start_solve(c, jacobi);
    if (rc) return rc;
    int done = 0;
    if (!jacobi) {
        // x = mu; r = E0 - x/alpha + T x; z = alpha r; p = z
        matvec(c, L, c->mu, nullptr);
        launch_pcg_init(s, G, c->Emv, c->E0, c->d_alpha, c->mu, c->r, c->p, mixed ? c->vec32 : nullptr, c->d_red,
                        c->d_scalars);
    }
    // no host check before the first chunk: systems already stopped make its launches return at once
    for (int pass = 0;; pass++) {
        if ((rc = run_iterations(c, L, G, jacobi, pass == 0 ? first_chunk : chunk))) return rc;
        if ((rc = stage_solver_state(c, done, pass == 0 ? ahead : nullptr))) return rc;
        if (pass == 0 && ahead_stands) *ahead_stands = ahead && done >= c->nseg;
        if (pass == 0 && c->ew_need_pending) {
            int c32 = 0, c64 = 0, tc = 0;
            if (!ew_take_need(c, c32, c64, tc)) {
                if (c->ew_stored && (rc = ew_alloc_lists(c, c32, c64, tc))) return rc;
                if (ahead_stands) *ahead_stands = false;
                return kEwRedo;
            }
        }
        if (done >= c->nseg) break;
    }
    if ((rc = read_solver_results(c, jacobi, true))) return rc;
    c->last_max_iter = *std::max_element(c->iterations.begin(), c->iterations.end());
    if (c->timing && !c->noalloc) harvest_timing(c);
    return TDP_OK;
}

constexpr int kSeqFront = 1, kSeqBack = 2;

// Whether an evaluation runs as captured sequences: the field, the start of the Mutual solve and its
// first chunk of iterations as one replay (no host step between them), then the host check of the
// stop count, then the forces and energy as one replay.  Not with the Ewald option (its list build
// checks capacities on the host), event timing or the allocation-free entry point.
bool sequence_graphs(const tdp_context* c) {
    return c->use_graphs && !c->timing && !c->noalloc && !c->ewald && c->polarization == TDP_MUTUAL && c->n > 0;
}

// setup() and the Mutual solve as sequence_graphs describes.  back (may be empty) enqueues the rest
// of the evaluation with its results guarded by the device stop count: it is launched right after
// the first chunk, before the host check, and *back_ran tells whether that launch stands (every
// system had stopped by then; else the caller enqueues the rest again after the solve).
int solve_captured(tdp_context* c, const double* box, const double* host_pos3,
                   const std::function<int()>& back = nullptr, bool* back_ran = nullptr) {
    int rc = prepare_order(c, box, host_pos3);
    if (rc) return rc;
    hipStream_t s = c->stream;
    const int npad = c->npad;
    PairLaunch L = pair_launch(c, box);
    SegLaunch G = seg_launch(c);
    const bool mixed = c->precision == TDP_PRECISION_MIXED;
    const bool jacobi = c->solver == TDP_SOLVER_JACOBI;
    const int chunk = c->n > 20000 ? 2 : 4;
    const int first_chunk = c->last_max_iter > 0 ? c->last_max_iter : (c->n > 20000 ? 4 : 16);
    const bool warm = c->warm_start && c->have_prev;
    if ((rc = start_solve_host(c, jacobi))) return rc;
    const long long key = (long long)first_chunk | ((long long)jacobi << 32) | ((long long)warm << 33);
    rc = replay_or_capture(c, kSeqFront, key, nullptr, nullptr, L.box, [&]() -> int {
        enqueue_field(c, L, G);
        if (warm)   // start from the previous solution instead of alpha E0
            launch_gather_vec3(s, npad, c->d_perm, c->d_mu_prev, c->mu, mixed ? c->vec32 : nullptr);
        int r = start_solve_copies(c);
        if (r) return r;
        if (!jacobi) {
            // x = mu; r = E0 - x/alpha + T x; z = alpha r; p = z
            matvec(c, L, c->mu, nullptr);
            launch_pcg_init(s, G, c->Emv, c->E0, c->d_alpha, c->mu, c->r, c->p, mixed ? c->vec32 : nullptr, c->d_red,
                            c->d_scalars);
        }
        enqueue_iterations(c, L, G, jacobi, first_chunk);
        return TDP_OK;
    });
    if (rc) return rc;
    if (back && (rc = back())) return rc;
    int done = 0;
    if ((rc = stage_solver_state(c, done))) return rc;
    if (back_ran) *back_ran = back && done >= c->nseg;
    while (done < c->nseg) {
        if ((rc = run_iterations(c, L, G, jacobi, chunk))) return rc;
        if ((rc = stage_solver_state(c, done))) return rc;
    }
    if ((rc = read_solver_results(c, jacobi, true))) return rc;
    c->last_max_iter = *std::max_element(c->iterations.begin(), c->iterations.end());
    if (c->warm_start) {
        launch_scatter_vec3(s, npad, c->d_perm, c->mu, c->d_mu_prev);
        c->have_prev = true;
    }
    return TDP_OK;
}

// setup() + calculateInducedDipoles (ReferenceTholeDipoleForce.cpp:1117-1155, :735-764).
int setup_and_solve(tdp_context* c, const double* box, const double* host_pos3,
                    const std::function<int()>* ahead = nullptr, bool* ahead_stands = nullptr) {
    if (ahead_stands) *ahead_stands = false;
    if (c->n == 0) {
        std::fill(c->iterations.begin(), c->iterations.end(), 0);
        std::fill(c->converged.begin(), c->converged.end(), 1);
        std::fill(c->epsilon.begin(), c->epsilon.end(), c->polarization == TDP_DIRECT ? 1.0e50 : 0.0);
        return TDP_OK;
    }
    if (sequence_graphs(c)) return solve_captured(c, box, host_pos3);
    int rc = prepare_field(c, box, host_pos3);
    if (rc) return rc;
    hipStream_t s = c->stream;
    const int npad = c->npad;
    PairLaunch L = pair_launch(c, box);
    SegLaunch G = seg_launch(c);
    const bool mixed = c->precision == TDP_PRECISION_MIXED;
    if (c->polarization == TDP_DIRECT) {
        std::fill(c->iterations.begin(), c->iterations.end(), 0);
        std::fill(c->converged.begin(), c->converged.end(), 1);
        std::fill(c->epsilon.begin(), c->epsilon.end(), 1.0e50);  // the reference leaves its initial value (:24, :749-753)
        return TDP_OK;
    }
    if (c->polarization == TDP_EXTRAPOLATED) {
        // convergeInducedDipolesByExtrapolation (:853-909); the coefficients are on the device since create
        const int order = c->extrap_order;
        TDP_HIP(c, hipMemcpyAsync(c->hist, c->mu, npad * sizeof(double4), hipMemcpyDeviceToDevice, s));
        for (int o = 1; o < order; o++) {
            matvec(c, L, c->mu, nullptr);
            launch_extrap_step(s, G, c->Emv, c->E0, c->d_alpha, c->mu, mixed ? c->vec32 : nullptr,
                               c->hist + (size_t)o * npad);
        }
        launch_extrap_combine(s, G, order, c->hist, c->d_coef, c->mu, mixed ? c->vec32 : nullptr);
        matvec(c, L, c->mu, nullptr);
        launch_extrap_error(s, G, c->Emv, c->E0, c->d_alpha, c->mu, c->d_red, c->d_scalars);
        int rc = read_solver_results(c, false);
        if (rc) return rc;
        std::fill(c->converged.begin(), c->converged.end(), 1);
        std::fill(c->iterations.begin(), c->iterations.end(), order);
        if (c->timing && !c->noalloc) harvest_timing(c);
        return TDP_OK;
    }
    if (c->warm_start && c->have_prev)   // start from the previous solution instead of alpha E0
        launch_gather_vec3(s, npad, c->d_perm, c->d_mu_prev, c->mu, mixed ? c->vec32 : nullptr);
    rc = solve_mutual(c, box, ahead, ahead_stands);
    if (rc) return rc;
    if (c->warm_start) {
        launch_scatter_vec3(s, npad, c->d_perm, c->mu, c->d_mu_prev);
        c->have_prev = true;
    }
    return TDP_OK;
}

// calculateElectrostatic + mapTorqueToForce (ReferenceTholeDipoleForce.cpp:118-124) into d_frc
// and d_energy.  The pair pass forms energies and forces together; a caller that wants no forces
// skips the torque mapping, one that wants no energy skips the energy reduction (includeForces,
// includeEnergy of TholeDipoleKernels.h:64-68).
void forces_and_energy(tdp_context* c, const double* box, bool want_forces = true, bool want_energy = true) {
    hipStream_t s = c->stream;
    if (c->n == 0) {
        hipMemsetAsync(c->d_energy, 0, sizeof(double), s);
        return;
    }
    if (!want_forces && !want_energy) return;
    const int npad = c->npad;
    if (c->ewald) {   // the dipoles' potential grid, then forces and energy per slot
        const EwLaunch E = ew_launch(c, box);
        if (c->ew_stored) {   // the E0 pass of this evaluation left the pair candidates
            // the dipoles' potential grid on stream2 beside the force pass, joined before its epilogue
            hipEventRecord(c->ev_fork, s);
            hipStreamWaitEvent(c->stream2, c->ev_fork, 0);
            launch_ew_recip(c->stream2, E, c->posa, nullptr, c->mu, E.phi_v, nullptr);
            hipEventRecord(c->ev_join, c->stream2);
            launch_ew_force_stored(s, E, c->polarization == TDP_MUTUAL, c->posa, c->d_inva, c->d_charge, c->mu, c->E0,
                                   c->d_epart, c->d_frc, c->d_eout, c->ev_join);
        } else {
            launch_ew_recip(s, E, c->posa, nullptr, c->mu, E.phi_v, nullptr);
            launch_ew_pairs(s, E, kEwForce, c->polarization == TDP_MUTUAL, c->posa, c->d_inva, c->d_charge, c->mu,
                            c->E0, nullptr, nullptr, PcgQ(), c->d_epart, c->d_frc, c->d_eout);
        }
        if (want_energy) {
            const double V = box[0] * box[4] * box[8];
            const double bg = -M_PI * c->ew_qsum * c->ew_qsum / (2.0 * V * c->ew_alpha * c->ew_alpha);
            launch_ew_energy(s, npad, c->d_eout, bg, c->d_energy);
        }
        return;
    }
    PairLaunch L = pair_launch(c, box);
    SegLaunch G = seg_launch(c);
    double4* pf = reinterpret_cast<double4*>(c->d_pbuf);
    double4* pt = c->d_pbuf2;
    const size_t off = (size_t)c->nwork * 64;
    launch_sym_elec(s, L, c->polarization == TDP_MUTUAL, c->precision == TDP_PRECISION_MIXED, c->posa, c->qd, c->mu,
                    c->d_inva, c->d_perm, pf,
                    reinterpret_cast<double*>(pf + off), pt, reinterpret_cast<double*>(pt + off), c->d_part_e, c->d_frc,
                    c->torque, c->d_slab32);
    if (want_energy) launch_reduce_energy(s, G, c->d_part_e, c->d_energy);
    if (want_forces && !c->gidx.empty() && c->has_perm) {   // without permanent dipoles every torque is zero
        launch_map_torque(s, npad, c->posa, c->d_axis, c->torque, c->contrib);
        launch_gather_torque(s, npad, c->d_goff, c->d_gidx, c->contrib, c->d_frc);
    }
}

// With TDP_FLAG_REQUIRE_CONVERGENCE, a Mutual solve of any system that stopped at its iteration
// limit above its target is an error (the reference records non-convergence silently,
// ReferenceTholeDipoleForce.cpp:849-850); the results of the evaluation are still written.
int convergence_status(tdp_context* c) {
    if (!(c->flags & TDP_FLAG_REQUIRE_CONVERGENCE) || c->polarization != TDP_MUTUAL) return TDP_OK;
    for (int g = 0; g < c->nseg; g++)
        if (!c->converged[g]) {
            char buf[200];
            snprintf(buf, sizeof buf, "Mutual induced dipoles did not converge: system %d stopped after %d iterations "
                     "at epsilon %.3e (target %.3e)", g, c->iterations[g], c->epsilon[g], c->seg_target[g]);
            return fail(c, TDP_ERR_NOT_CONVERGED, buf);
        }
    return TDP_OK;
}

int ensure_out(tdp_context* c, size_t bytes) {
    if (bytes <= c->out_bytes) return TDP_OK;
    if (c->d_out) {
        hipFree(c->d_out);
        auto it = std::find(c->allocations.begin(), c->allocations.end(), (void*)c->d_out);
        if (it != c->allocations.end()) c->allocations.erase(it);
        c->d_out = nullptr;
    }
    int rc = dalloc(c, &c->d_out, bytes / sizeof(double) + 1);
    if (rc) return rc;
    c->out_bytes = bytes;
    return TDP_OK;
}

// Forces (accumulated into the caller's order) and the energy of every system.
int download_results(tdp_context* c, double* forces, double* energies) {
    const int npad = c->npad;
    std::vector<double> f;
    std::vector<double> e(c->nseg, 0.0);
    if (forces && c->n > 0) {
        f.resize(3 * (size_t)npad);
        TDP_HIP(c, hipMemcpyAsync(f.data(), c->d_frc, 3 * (size_t)npad * sizeof(double), hipMemcpyDeviceToHost,
                                  c->stream));
    }
    TDP_HIP(c, hipMemcpyAsync(e.data(), c->d_energy, c->nseg * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    TDP_HIP(c, hipStreamSynchronize(c->stream));
    if (!f.empty())
        for (int k = 0; k < npad; k++) {
            const int o = c->perm[k];
            if (o < 0) continue;
            for (int d = 0; d < 3; d++) forces[3 * o + d] += f[3 * k + d];
        }
    if (energies)
        for (int g = 0; g < c->nseg; g++) energies[g] = e[g];
    return TDP_OK;
}

int upload_positions(tdp_context* c, const double* positions) {
    if (c->n > 0)
        TDP_HIP(c, hipMemcpyAsync(c->d_pos3, positions, 3 * c->n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    return TDP_OK;
}

// Upper-triangle work items: tile row I of a system split into chunks of C tiles (J >= I, same
// system), with the item partial buffers sized for them.  Short items keep the many one-wave
// workgroups evenly spread over the CUs to the end of a launch; long items need fewer item
// partials in the reduction.
// Work items of the pair kernels with Cb[b] tiles per item for the systems of size bucket b.
int set_work_items_buckets(tdp_context* c, const int* Cb) {
    const int nb = c->nblk, ns = c->nseg;
    std::vector<int4> work;
    std::vector<int> row_start(nb + 1, 0);
    int largest = 0;
    for (int g = 0; g < ns; g++) {
        c->seg_w0[g] = (int)work.size();
        const int t1 = c->seg_t1[g];
        const int C = std::max(1, Cb[size_bucket(t1 - c->seg_t0[g])]);
        largest = std::max(largest, t1 - c->seg_t0[g]);
        for (int I = c->seg_t0[g]; I < t1; I++) {
            row_start[I] = (int)work.size();
            for (int J0 = I; J0 < t1; J0 += C) work.push_back(make_int4(I, J0, std::min(t1, J0 + C), 0));
        }
        c->seg_w1[g] = (int)work.size();
    }
    row_start[nb] = (int)work.size();
    c->nwork = (int)work.size();
    for (int b = 0; b < kSizeBuckets; b++) c->work_Cb[b] = std::max(1, Cb[b]);
    c->work_C = c->work_Cb[size_bucket(std::max(largest, 1))];
    c->graph_gen++;
    std::vector<int2> segw(ns);
    for (int g = 0; g < ns; g++) segw[g] = make_int2(c->seg_w0[g], c->seg_w1[g]);
    const size_t pb = ((size_t)c->nwork + c->ntri) * 64;
    int rc;
    // (the Ewald option's passes keep no pair partials: its contexts get token buffers)
    const size_t pbn = c->ewald ? 1 : pb;
    if ((rc = dresize(c, &c->d_work, c->cap_work, work.size())) ||
        (rc = dresize(c, (double4**)&c->d_pbuf, c->cap_pbuf, pbn)) || (rc = dresize(c, &c->d_pbuf2, c->cap_pbuf2, pbn)) ||
        (rc = dresize(c, &c->d_part_e, c->cap_part_e, c->ewald ? 1 : (size_t)c->nwork + 1)) ||
        (c->precision == TDP_PRECISION_MIXED &&
         (rc = dresize(c, &c->d_slab32, c->cap_slab32, c->ewald ? 1 : 3 * ((size_t)c->ntri + c->nwork) * 64))))
        return rc;
    if (!c->d_row_start && (rc = dalloc(c, &c->d_row_start, nb + 1))) return rc;
    if (!c->d_segw && (rc = dalloc(c, &c->d_segw, ns))) return rc;
    TDP_HIP(c, hipMemcpyAsync(c->d_work, work.data(), work.size() * sizeof(int4), hipMemcpyHostToDevice, c->stream));
    TDP_HIP(c, hipMemcpyAsync(c->d_row_start, row_start.data(), row_start.size() * sizeof(int), hipMemcpyHostToDevice,
                              c->stream));
    TDP_HIP(c, hipMemcpyAsync(c->d_segw, segw.data(), ns * sizeof(int2), hipMemcpyHostToDevice, c->stream));
    TDP_HIP(c, hipStreamSynchronize(c->stream));
    return TDP_OK;
}

int set_work_items(tdp_context* c, int C) {
    int Cb[kSizeBuckets];
    for (int b = 0; b < kSizeBuckets; b++) Cb[b] = C;
    return set_work_items_buckets(c, Cb);
}

// Tiles, work items and partial-slab layout of the systems (sizes in the caller's order).
void build_layout(tdp_context* c, const std::vector<int>& sizes) {
    c->nseg = (int)sizes.size();
    c->seg_n = sizes;
    c->seg_p0.assign(c->nseg, 0);
    c->seg_t0.assign(c->nseg, 0);
    c->seg_t1.assign(c->nseg, 0);
    c->seg_w0.assign(c->nseg, 0);
    c->seg_w1.assign(c->nseg, 0);
    int p0 = 0, t0 = 0;
    for (int g = 0; g < c->nseg; g++) {
        c->seg_p0[g] = p0;
        c->seg_t0[g] = t0;
        t0 += (sizes[g] + 63) / 64;
        c->seg_t1[g] = t0;
        p0 += sizes[g];
    }
    c->n = p0;
    c->nblk = t0;
    c->npad = std::max(64, 64 * c->nblk);
    c->tcnt.assign(c->nblk, 0);
    c->tfirst.assign(c->nblk, 0);
    c->tseg.assign(c->nblk, 0);
    c->tcol.assign(c->nblk, 0);
    size_t base = 0;
    for (int g = 0; g < c->nseg; g++) {
        const int a = c->seg_t0[g], b = c->seg_t1[g];
        for (int T = a; T < b; T++) {
            c->tcnt[T] = std::min(64, sizes[g] - 64 * (T - a));
            c->tfirst[T] = a;
            c->tseg[T] = g;
            c->tcol[T] = base + (size_t)(T - a) * (T - a + 1) / 2;
        }
        base += (size_t)(b - a) * (b - a + 1) / 2;
    }
    c->ntri = base;
    // identity order within each system until the first evaluation computes the spatial one
    c->perm.assign(c->npad, -1);
    c->inv.assign(c->n, 0);
    for (int g = 0; g < c->nseg; g++)
        for (int k = 0; k < sizes[g]; k++) {
            c->perm[64 * c->seg_t0[g] + k] = c->seg_p0[g] + k;
            c->inv[c->seg_p0[g] + k] = 64 * c->seg_t0[g] + k;
        }
}

int create_context(const tdp_force_desc* d, const std::vector<int>& sizes, const std::vector<int>& maxit,
                   const std::vector<double>& target, const tdp_options* opts, tdp_context** out) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return fail(nullptr, TDP_ERR_NO_DEVICE, "no HIP device is available");
    tdp_context* c = new tdp_context();
    c->device = opts ? opts->device : 0;
    c->precision = opts ? opts->precision : TDP_PRECISION_DOUBLE;
    c->solver = opts ? opts->solver : TDP_SOLVER_JACOBI;
    c->flags = opts ? opts->flags : 0;
    if (c->device < 0 || c->device >= ndev) {
        delete c;
        return fail(nullptr, TDP_ERR_NO_DEVICE, "invalid device ordinal");
    }
    int rc = 0;
    auto bail = [&](int code) {
        g_create_error = c->err;
        tdp_destroy(c);
        return code;
    };
    if (hipSetDevice(c->device) != hipSuccess) {
        c->err = "hipSetDevice failed";
        return bail(TDP_ERR_HIP);
    }
    build_layout(c, sizes);
    c->method = d->nonbonded_method;
    c->polarization = d->polarization_type;
    c->cutoff = d->cutoff;
    c->ewald_tol = d->ewald_error_tolerance;
    c->pme_alpha = d->pme_alpha;
    for (int k = 0; k < 3; k++) c->grid[k] = d->pme_grid[k];
    std::memcpy(c->default_box, d->default_box, sizeof c->default_box);
    // Solver settings are read only for the active polarization type (ReferenceTholeDipoleKernels.cpp:127-134).
    c->seg_maxit.assign(c->nseg, 60);
    c->seg_target.assign(c->nseg, 1e-5);
    if (c->polarization == TDP_MUTUAL) {
        c->seg_maxit = maxit;
        c->seg_target = target;
    } else if (c->polarization == TDP_EXTRAPOLATED) {
        if (d->num_extrapolation_coefficients > 0)
            c->extrap.assign(d->extrapolation_coefficients,
                             d->extrapolation_coefficients + d->num_extrapolation_coefficients);
    }
    c->iterations.assign(c->nseg, 0);
    c->converged.assign(c->nseg, 0);
    c->epsilon.assign(c->nseg, 1.0e50);
    copy_particles(c, d);
    build_exclusion_pairs(c, d);
    if (c->flags & TDP_FLAG_EWALD) {
        c->ewald = true;
        const char* why = nullptr;
        const double* b = c->default_box;
        if (c->method != TDP_PME) why = "the Ewald option needs the PME method";
        else if (c->nseg != 1) why = "the Ewald option is not available for batches";
        else if (c->polarization == TDP_EXTRAPOLATED) why = "the Ewald option supports Direct and Mutual polarization";
        else if (b[1] != 0.0 || b[2] != 0.0 || b[3] != 0.0 || b[5] != 0.0 || b[6] != 0.0 || b[7] != 0.0)
            why = "the Ewald option needs a rectangular box";
        for (double v : c->dmol)
            if (v != 0.0) why = "the Ewald option supports charges and induced dipoles only (molecular dipoles must be zero)";
        if (!why) {
            double a;
            int g[3];
            pme_parameters(c, a, g);
            c->ew_alpha = a;
            for (int k = 0; k < 3; k++) {
                c->ew_K[k] = g[k];
                if (g[k] < kEwOrder || !fft_plan(g[k], c->ew_radix[k], c->ew_nst[k]))
                    why = "the Ewald option needs PME grid sizes of at least 5 with prime factors 2, 3, 5 and 7 only";
            }
            if (!(a > 0.0)) why = "the Ewald option needs a positive Ewald parameter";
        }
        if (why) {
            c->err = std::string("TholeDipoleForce: ") + why;
            return bail(TDP_ERR_INVALID_ARGUMENT);
        }
        ew_particle_terms(c);
    }
    if (hipStreamCreate(&c->stream) != hipSuccess) {
        c->err = "hipStreamCreate failed";
        return bail(TDP_ERR_HIP);
    }
    if (c->ewald && (hipStreamCreate(&c->stream2) != hipSuccess ||
                     hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
                     hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess ||
                     hipEventCreateWithFlags(&c->ev_stage, hipEventDisableTiming) != hipSuccess)) {
        c->err = "hipStreamCreate / hipEventCreate failed";
        return bail(TDP_ERR_HIP);
    }
    {
        const char* g = std::getenv("TDP_ITERATION_GRAPHS");   // "0": enqueue iterations one launch at a time
        c->use_graphs = !(g && g[0] == '0');
    }
    const int npad = c->npad, nb = c->nblk, ns = c->nseg;
    // Default work items: C = 3 tiles on large systems (measured at 36,864 sites, C = 3 beats C = 20
    // by 25% on the T*mu pass, and C = 1 loses more in the reduction than it gains); tdp_autotune
    // times the candidates on the caller's configuration instead.
    if ((rc = set_work_items(c, (int)std::max<size_t>(1, std::min<size_t>(3, c->ntri / 16384))))) return bail(rc);
    {
        std::vector<int4> seg(ns);
        for (int g = 0; g < ns; g++) seg[g] = make_int4(c->seg_t0[g], c->seg_t1[g], c->seg_n[g], 0);
        if ((rc = dalloc(c, &c->Emv, npad)) || (rc = dalloc(c, &c->d_tcnt, nb)) || (rc = dalloc(c, &c->d_tfirst, nb)) ||
            (rc = dalloc(c, &c->d_tseg, nb)) || (rc = dalloc(c, &c->d_tcol, nb)) || (rc = dalloc(c, &c->d_seg, ns)) ||
            (rc = dalloc(c, &c->d_done, 2)))
            return bail(rc);
        bool ok = true;
        auto up = [&](void* dst, const void* src, size_t bytes) {
            if (bytes && hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream) != hipSuccess) ok = false;
        };
        up(c->d_tcnt, c->tcnt.data(), nb * sizeof(int));
        up(c->d_tfirst, c->tfirst.data(), nb * sizeof(int));
        up(c->d_tseg, c->tseg.data(), nb * sizeof(int));
        up(c->d_tcol, c->tcol.data(), nb * sizeof(size_t));
        up(c->d_seg, seg.data(), ns * sizeof(int4));
        if (!ok || hipStreamSynchronize(c->stream) != hipSuccess) {
            c->err = "hipMemcpy of the tile tables failed";
            return bail(TDP_ERR_HIP);
        }
    }
    if ((rc = dalloc(c, &c->posa, npad)) || (rc = dalloc(c, &c->qd, npad)) || (rc = dalloc(c, &c->mu, npad)) ||
        (rc = dalloc(c, &c->E0, npad)) || (rc = dalloc(c, &c->r, npad)) || (rc = dalloc(c, &c->z, npad)) ||
        (rc = dalloc(c, &c->p, npad)) || (rc = dalloc(c, &c->q, npad)) || (rc = dalloc(c, &c->torque, npad)) ||
        (rc = dalloc(c, &c->contrib, 4 * (size_t)npad)) || (rc = dalloc(c, &c->vec32, npad)) ||
        (rc = dalloc(c, &c->d_pos3, 3 * (size_t)std::max(c->n, 1))) || (rc = dalloc(c, &c->d_frc, 3 * (size_t)npad)) ||
        (rc = dalloc(c, &c->d_alpha, npad)) || (rc = dalloc(c, &c->d_inva, npad)) ||
        (rc = dalloc(c, &c->d_thole, npad)) || (rc = dalloc(c, &c->d_charge, npad)) ||
        (rc = dalloc(c, &c->d_dmol, 3 * (size_t)npad)) || (rc = dalloc(c, &c->d_red, 3 * (size_t)nb + 2)) ||
        (rc = dalloc(c, &c->d_scalars, (size_t)SC_COUNT * ns)) || (rc = dalloc(c, &c->d_coef, 64)) ||
        (rc = dalloc(c, &c->d_energy, ns)) || (rc = dalloc(c, &c->d_axis, npad)) ||
        (rc = dalloc(c, &c->d_exc_start, nb + 1)) || (rc = dalloc(c, &c->d_goff, npad + 1)) ||
        (rc = dalloc(c, &c->d_perm, npad)) || (rc = dalloc(c, &c->d_mass, npad)) ||
        (rc = dalloc(c, &c->tb_c, nb + 1)) || (rc = dalloc(c, &c->tb_h, nb + 1)) ||
        (rc = dalloc(c, &c->tb_o, nb + 1)) || (rc = dalloc(c, &c->posl, npad)) || (rc = dalloc(c, &c->posl64, npad)) ||
        (rc = dalloc(c, &c->d_tp, c->ewald ? 1 : c->ntri)))
        return bail(rc);
    if (c->ewald) {
        const size_t ng = (size_t)c->ew_K[0] * c->ew_K[1] * c->ew_K[2];
        if ((rc = dalloc(c, &c->d_nbr, (size_t)nb * nb)) || (rc = dalloc(c, &c->d_nbr_cnt, nb)) ||
            (rc = dalloc(c, &c->d_nbr_exc, (size_t)nb * nb)) || (rc = dalloc(c, &c->d_mu32, npad)) ||
            (rc = dalloc(c, &c->d_slot32, npad)) ||
            (rc = dalloc(c, &c->d_qfix, ng)) || (rc = dalloc(c, &c->d_cgrid, ng)) || (rc = dalloc(c, &c->d_phiq, ng)) ||
            (rc = dalloc(c, &c->d_phiv, ng)) || (rc = dalloc(c, &c->d_G64, ng)) || (rc = dalloc(c, &c->d_G32, ng)) ||
            (rc = dalloc(c, &c->d_sc_off, (size_t)npad + 1)) || (rc = dalloc(c, &c->d_epart, npad)) ||
            (rc = dalloc(c, &c->d_eout, npad)) || (rc = dalloc(c, &c->d_rs, (size_t)npad * std::max(kEwParts, kEwE0Parts))) ||
            (rc = dalloc(c, &c->d_rsw, (size_t)npad * kEwWaves)) ||
            (rc = dalloc(c, &c->d_w32, (size_t)nb * kEwWaves)) || (rc = dalloc(c, &c->d_w64, (size_t)nb * kEwWaves)) ||
            (rc = dalloc(c, &c->d_need, 8)) || (rc = dalloc(c, &c->d_slot_cnt, (size_t)nb * kEwWaves)))
            return bail(rc);
        if (hipHostMalloc(&c->h_need, 8 * sizeof(int), hipHostMallocDefault) != hipSuccess) {
            c->err = "hipHostMalloc failed";
            return bail(TDP_ERR_HIP);
        }
        {   // list capacities per wave from the mean density: 1.5x the expected pairs within the cutoff
            const double V = c->default_box[0] * c->default_box[4] * c->default_box[8];
            const double rho = V > 0.0 ? c->n / V : 0.0;
            const double ball = 4.0 / 3.0 * M_PI;
            const double est = rho * ball * c->cutoff * c->cutoff * c->cutoff;
            const double est64 = rho * ball * c->ew_r64 * c->ew_r64 * c->ew_r64;
            const bool mixed_lists = c->precision == TDP_PRECISION_MIXED;
            const int c32 = mixed_lists ? (int)std::ceil(1.5 * est / kEwWaves) + 32 : 0;
            const int c64 = mixed_lists ? (int)std::ceil(1.5 * est64 / kEwWaves) + 16
                                        : (int)std::ceil(1.5 * est / kEwWaves) + 32;
            // slot tables: the wave's share of the slots within the cutoff of a tile (a ball of radius
            // r_c plus twice the radius of a tile's volume), 1.5x
            const double rt = rho > 0.0 ? std::cbrt(3.0 * 64 / (4.0 * M_PI * rho)) : 0.0;
            const double rr = c->cutoff + 2.0 * rt;
            const int tc = mixed_lists ? (int)std::min(8.0 * c->nblk, std::ceil(1.5 * rho * ball * rr * rr * rr / kEwWaves) + 64) : 0;
            if ((rc = ew_alloc_lists(c, c32, c64, tc))) return bail(rc);
        }
        for (int k = 0; k < 3; k++) {
            const int K = c->ew_K[k];
            std::vector<double2> t64(K);
            std::vector<float2> t32(K);
            for (int q = 0; q < K; q++) {
                const double ang = -2.0 * M_PI * q / K;
                t64[q] = make_double2(std::cos(ang), std::sin(ang));
                t32[q] = make_float2((float)t64[q].x, (float)t64[q].y);
            }
            if ((rc = dalloc(c, &c->d_tw64[k], K)) || (rc = dalloc(c, &c->d_tw32[k], K))) return bail(rc);
            if (hipMemcpy(c->d_tw64[k], t64.data(), K * sizeof(double2), hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(c->d_tw32[k], t32.data(), K * sizeof(float2), hipMemcpyHostToDevice) != hipSuccess) {
                c->err = "hipMemcpy of the FFT twiddles failed";
                return bail(TDP_ERR_HIP);
            }
        }
    }
    if (c->polarization == TDP_EXTRAPOLATED) {
        static const double opt4[4] = {-0.154, 0.017, 0.658, 0.474};   // TholeDipoleForce.cpp default
        std::vector<double> coef = c->extrap;
        if (coef.empty()) coef.assign(opt4, opt4 + 4);
        if (coef.size() > 64) {
            c->err = "TholeDipoleForce: more than 64 extrapolation coefficients";
            return bail(TDP_ERR_INVALID_ARGUMENT);
        }
        c->extrap_order = (int)coef.size();
        c->hist_order = c->extrap_order;
        if ((rc = dalloc(c, &c->hist, (size_t)c->hist_order * npad))) return bail(rc);
        if (hipMemcpyAsync(c->d_coef, coef.data(), coef.size() * sizeof(double), hipMemcpyHostToDevice, c->stream) !=
                hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess) {
            c->err = "hipMemcpy of the extrapolation coefficients failed";
            return bail(TDP_ERR_HIP);
        }
    }
    // pinned staging of the solver state and the events of one evaluation, so that an evaluation
    // allocates nothing
    if (hipHostMalloc(&c->h_stage, sizeof(double) * (stage_download_doubles(c) + stage_upload_doubles(c)),
                      hipHostMallocDefault) != hipSuccess) {
        c->err = "hipHostMalloc of the staging buffer failed";
        return bail(TDP_ERR_HIP);
    }
    for (int k = 0; k < 4; k++) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) {
            c->err = "hipEventCreate failed";
            return bail(TDP_ERR_HIP);
        }
        c->ev_pool.push_back(e);
    }
    if (scratch_init(c->device) != hipSuccess) {
        c->err = "hipGetSymbolAddress of the allocation-free scratch failed";
        return bail(TDP_ERR_HIP);
    }
    if ((rc = upload_tables(c)) || (rc = upload_particles(c))) return bail(rc);
    *out = c;
    return TDP_OK;
}

// Dense-path buffers of a batch context: the 16-bit matrices, diagonal-block inverses and the
// refinement vectors of every system.
int dense_alloc(tdp_context* c) {
    if (c->dense_ready) return TDP_OK;
    const int ns = c->nseg;
    std::vector<DenseSys> sys(ns);
    size_t moff = 0, voff = 0, doff = 0;
    int maxnt = 0;
    for (int g = 0; g < ns; g++) {
        const int nt = c->seg_t1[g] - c->seg_t0[g];
        if (nt > dense_max_tiles())
            return fail(c, TDP_ERR_INVALID_ARGUMENT, "dense solve: system " + std::to_string(g) + " has more than " +
                                                         std::to_string(64 * dense_max_tiles()) + " particles");
        const size_t m = (size_t)192 * nt;
        sys[g].nt = nt;
        sys[g].t0 = c->seg_t0[g];
        sys[g].moff = moff;
        sys[g].voff = voff;
        sys[g].doff = doff;
        moff += m * m;
        voff += 4 * m;
        doff += (size_t)3 * nt * 64 * 64;
        maxnt = std::max(maxnt, nt);
    }
    int rc;
    if ((rc = dalloc(c, &c->d_dsys, ns)) || (rc = dalloc(c, &c->d_dmat, moff)) || (rc = dalloc(c, &c->d_dinv, doff)) || (rc = dalloc(c, &c->d_dacc, doff)) ||
        (rc = dalloc(c, &c->d_db, voff)) || (rc = dalloc(c, &c->d_dx, voff)) || (rc = dalloc(c, &c->d_dres, voff)) ||
        (rc = dalloc(c, &c->d_dy, voff)) || (rc = dalloc(c, &c->d_drel, 4 * (size_t)ns)) ||
        (rc = dalloc(c, &c->d_dmu, 4 * (size_t)c->npad)) || (rc = dalloc(c, &c->d_dactive, ns)) ||
        (rc = dalloc(c, &c->d_dstatus, 1)))
        return rc;
    TDP_HIP(c, hipMemcpyAsync(c->d_dsys, sys.data(), ns * sizeof(DenseSys), hipMemcpyHostToDevice, c->stream));
    TDP_HIP(c, hipStreamSynchronize(c->stream));
    c->dense_max_nt = maxnt;
    c->dense_dwords = doff;
    c->dense_ready = true;
    return TDP_OK;
}

// Fields q = 0..nfield-1 (nfield <= 4) of double4 vectors in the internal order (src + q npad) to the
// caller's order out[q][n][3]: one gather per field on the device, then one download.
int download_fields(tdp_context* c, const double4* src, int nfield, double* out) {
    hipStream_t s = c->stream;
    const size_t n3 = (size_t)3 * c->n;
    int rc;
    if (!c->d_rout && (rc = dalloc(c, &c->d_rout, 4 * n3))) return rc;
    for (int q = 0; q < nfield; q++)
        launch_scatter_vec3(s, c->npad, c->d_perm, src + (size_t)q * c->npad, c->d_rout + q * n3);
    TDP_HIP(c, hipMemcpyAsync(out, c->d_rout, nfield * n3 * sizeof(double), hipMemcpyDeviceToHost, s));
    TDP_HIP(c, hipStreamSynchronize(s));
    return TDP_OK;
}

int cg_responses(tdp_context* c, double tol, int max_iter, double* out, tdp_dense_stats* st);

// Induced dipoles of every system for four fields: E0 (the permanent field) and unit uniform
// fields along x, y and z.  Low-precision Cholesky of A^ = I - D^1/2 T D^1/2 on MFMA, then FP64
// iterative refinement with matrix-free FP64 residuals until every system's relative residual
// |D^1/2 (b - A x)| / |D^1/2 b| is at most tol.  out: [4][3n] in the caller's order.
int dense_responses(tdp_context* c, bool bf16, double tol, int max_steps, double* out, tdp_dense_stats* st) {
    hipStream_t s = c->stream;
    int rc = dense_alloc(c);
    if (rc) return rc;
    PairLaunch L = pair_launch(c, nullptr);
    SegLaunch G = seg_launch(c);
    const int ns = c->nseg, npad = c->npad;
    int zero = 0;
    TDP_HIP(c, hipMemcpyAsync(c->d_dstatus, &zero, sizeof zero, hipMemcpyHostToDevice, s));
    hipEvent_t ev[3];
    for (auto& e : ev) e = take_event(c);
    TDP_HIP(c, hipEventRecord(ev[0], s));
    launch_dense_assemble(s, bf16, ns, c->dense_max_nt, c->d_dsys, c->d_tcnt, c->posa, c->d_alpha, c->d_inva,
                          c->d_exc_start, c->d_exc_J, c->d_exc_mask, c->d_dmat);
    TDP_HIP(c, hipEventRecord(ev[1], s));
    TDP_HIP(c, hipMemsetAsync(c->d_dacc, 0, c->dense_dwords * sizeof(float), s));
    launch_dense_factor(s, bf16, ns, c->dense_max_nt, c->d_dsys, c->d_dmat, c->d_dinv, c->d_dacc, c->d_dstatus);
    TDP_HIP(c, hipEventRecord(ev[2], s));
    launch_dense_rhs(s, G, c->d_dsys, c->d_tcnt, c->E0, c->d_alpha, c->d_db, c->d_dx, c->d_dres);
    std::vector<int> active(ns, 1);
    std::vector<double> rel(4 * (size_t)ns, 0.0);
    TDP_HIP(c, hipMemcpyAsync(c->d_dactive, active.data(), ns * sizeof(int), hipMemcpyHostToDevice, s));
    double4* pi = reinterpret_cast<double4*>(c->d_pbuf);
    double* pj = reinterpret_cast<double*>(pi + (size_t)c->nwork * 64);
    int steps = 0;
    double worst = 0.0;
    for (; steps < max_steps;) {
        launch_dense_solve(s, bf16, ns, c->d_dsys, c->d_dmat, c->d_dinv, c->d_dres, c->d_dy, c->d_dactive);
        launch_dense_update(s, G, c->d_dsys, c->d_alpha, c->d_dy, c->d_dx, c->d_dmu, (size_t)npad, c->d_dactive);
        steps++;
        for (int q = 0; q < 4; q++) {
            launch_sym_matvec64(s, L, c->posa, c->d_dmu + (size_t)q * npad, c->d_inva, nullptr, pi, pj);
            launch_col_reduce(s, L, pi, pj, nullptr, c->Emv, nullptr);
            launch_dense_residual(s, G, q, c->d_dsys, c->d_alpha, c->Emv, c->d_db, c->d_dx, c->d_dres, c->d_red,
                                  c->d_drel);
        }
        TDP_HIP(c, hipMemcpyAsync(rel.data(), c->d_drel, rel.size() * sizeof(double), hipMemcpyDeviceToHost, s));
        TDP_HIP(c, hipStreamSynchronize(s));
        worst = 0.0;
        int left = 0;
        for (int g = 0; g < ns; g++) {
            double w = 0.0;
            for (int q = 0; q < 4; q++) w = nan_max(w, rel[4 * g + q]);
            worst = nan_max(worst, w);
            active[g] = w <= tol ? 0 : 1;   // a NaN residual keeps the system active
            left += active[g];
        }
        if (!left) break;
        TDP_HIP(c, hipMemcpyAsync(c->d_dactive, active.data(), ns * sizeof(int), hipMemcpyHostToDevice, s));
    }
    int status = 0;
    TDP_HIP(c, hipMemcpyAsync(&status, c->d_dstatus, sizeof status, hipMemcpyDeviceToHost, s));
    if (out) {
        if ((rc = download_fields(c, c->d_dmu, 4, out))) return rc;
    } else {
        TDP_HIP(c, hipStreamSynchronize(s));
    }
    float t_asm = 0.f, t_fac = 0.f;
    hipEventElapsedTime(&t_asm, ev[0], ev[1]);
    hipEventElapsedTime(&t_fac, ev[1], ev[2]);
    for (auto& e : ev) c->ev_pool.push_back(e);
    if (st) {
        st->steps = steps;
        st->max_relative_residual = worst;
        st->factor_failed = status;
        st->assemble_ms = t_asm;
        st->factor_ms = t_fac;
    }
    if (worst <= tol) return TDP_OK;   // accepted on the FP64 residual, whatever the factor's pivots were
    // The refinement did not reach the tolerance (an indefinite or badly rounded factor, or too few
    // steps): the whole batch is solved again by FP64 CG, and fails if that does not converge either.
    tdp_dense_stats cg{};
    const int rc_cg = cg_responses(c, tol, std::max(max_steps, 200), out, &cg);
    if (st) {
        st->steps = steps + cg.steps;
        st->max_relative_residual = cg.max_relative_residual;
    }
    if (rc_cg == TDP_ERR_NOT_CONVERGED)
        return fail(c, TDP_ERR_NOT_CONVERGED,
                    std::string("responses: refinement of the ") + (bf16 ? "BF16" : "FP16") +
                        " factor and the FP64 CG fallback both ended above the tolerance" +
                        (status ? " (a pivot of the factorization was not positive)" : ""));
    return rc_cg;
}

// The same four responses by Jacobi-preconditioned FP64 CG of every system (the comparator of the
// dense path): the Mutual solver run once per field, stopped when RMS(alpha r) <= tol RMS(alpha b)
// for the system.
int cg_responses(tdp_context* c, double tol, int max_iter, double* out, tdp_dense_stats* st) {
    hipStream_t s = c->stream;
    const int ns = c->nseg, npad = c->npad;
    SegLaunch G = seg_launch(c);
    std::vector<double4> e0(npad);
    TDP_HIP(c, hipMemcpyAsync(e0.data(), c->E0, npad * sizeof(double4), hipMemcpyDeviceToHost, s));
    TDP_HIP(c, hipStreamSynchronize(s));
    std::vector<double4> e0save = e0;
    const std::vector<int> maxit0 = c->seg_maxit;
    const std::vector<double> target0 = c->seg_target;
    const int solver0 = c->solver, prec0 = c->precision;
    c->solver = TDP_SOLVER_PCG;
    c->precision = TDP_PRECISION_DOUBLE;
    int total_iter = 0;
    double worst = 0.0;
    int rc = TDP_OK;
    if (out && !c->d_rout && (rc = dalloc(c, &c->d_rout, (size_t)12 * c->n))) return rc;
    for (int q = 0; q < 4 && rc == TDP_OK; q++) {
        if (q > 0) launch_uniform_field(s, npad, c->d_perm, q - 1, c->E0);
        else TDP_HIP(c, hipMemcpyAsync(c->E0, e0save.data(), npad * sizeof(double4), hipMemcpyHostToDevice, s));
        // per system: target = tol * RMS(alpha b)
        for (int g = 0; g < ns; g++) {
            double sum = 0.0;
            for (int k = 64 * c->seg_t0[g]; k < 64 * c->seg_t1[g]; k++) {
                const int o = c->perm[k];
                if (o < 0) continue;
                const double a = c->polarity[o];
                const double4 b = e0save[k];
                const double bx = q == 0 ? b.x : (q == 1), by = q == 0 ? b.y : (q == 2), bz = q == 0 ? b.z : (q == 3);
                sum += a * a * (bx * bx + by * by + bz * bz);
            }
            c->seg_target[g] = tol * std::sqrt(sum / c->seg_n[g]);
            c->seg_maxit[g] = max_iter;
        }
        launch_field_init(s, G, c->E0, c->d_alpha, c->E0, c->mu, nullptr);
        rc = solve_mutual(c, nullptr);
        if (rc) break;
        for (int g = 0; g < ns; g++) {
            total_iter = std::max(total_iter, c->iterations[g]);
            worst = nan_max(worst, c->epsilon[g] / std::max(c->seg_target[g] / tol, 1e-300));
        }
        // field q in the caller's order; downloaded with the others after the last solve
        if (out) launch_scatter_vec3(s, npad, c->d_perm, c->mu, c->d_rout + (size_t)q * 3 * c->n);
    }
    c->seg_maxit = maxit0;
    c->seg_target = target0;
    c->solver = solver0;
    c->precision = prec0;
    if (rc) return rc;
    TDP_HIP(c, hipMemcpyAsync(c->E0, e0save.data(), npad * sizeof(double4), hipMemcpyHostToDevice, s));
    if (out) TDP_HIP(c, hipMemcpyAsync(out, c->d_rout, (size_t)12 * c->n * sizeof(double), hipMemcpyDeviceToHost, s));
    TDP_HIP(c, hipStreamSynchronize(s));
    if (st) {
        st->steps = total_iter;
        st->max_relative_residual = worst;
        st->factor_failed = 0;
        st->assemble_ms = 0.0;
        st->factor_ms = 0.0;
    }
    if (!(worst <= tol))
        return fail(c, TDP_ERR_NOT_CONVERGED, "responses: FP64 CG ended above the tolerance after " +
                                                  std::to_string(total_iter) + " iterations");
    return TDP_OK;
}

}  // namespace

// ====================================================================== C ABI
extern "C" {

const char* tdp_last_create_error(void) { return g_create_error.c_str(); }

const char* tdp_last_error(const tdp_context* ctx) { return ctx ? ctx->err.c_str() : g_create_error.c_str(); }

int tdp_create(const tdp_force_desc* d, const tdp_options* opts, tdp_context** out) {
    if (!d || !out) return fail(nullptr, TDP_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    int rc = validate(nullptr, d);
    if (rc) return rc;
    std::vector<int> sizes{d->num_particles};
    std::vector<int> maxit{d->mutual_max_iterations};
    std::vector<double> target{d->mutual_target_epsilon};
    return create_context(d, sizes, maxit, target, opts, out);
}

int tdp_batch_create(const tdp_force_desc* descs, int32_t count, const tdp_options* opts, tdp_context** out) {
    if (!descs || !out || count <= 0) return fail(nullptr, TDP_ERR_INVALID_ARGUMENT, "null argument or empty batch");
    *out = nullptr;
    for (int g = 0; g < count; g++) {
        const tdp_force_desc& d = descs[g];
        if (d.nonbonded_method != TDP_NO_CUTOFF)
            return fail(nullptr, TDP_ERR_INVALID_ARGUMENT, "batch: every system must use the NoCutoff method");
        if (d.polarization_type != descs[0].polarization_type)
            return fail(nullptr, TDP_ERR_INVALID_ARGUMENT, "batch: every system must use the same polarization type");
        if (d.num_particles <= 0)
            return fail(nullptr, TDP_ERR_INVALID_ARGUMENT, "batch: system " + std::to_string(g) + " has no particles");
        int rc = validate(nullptr, &d);
        if (rc) return rc;
    }
    // one force description with the systems' particles concatenated and their indices offset
    std::vector<int> sizes(count), maxit(count);
    std::vector<double> target(count);
    size_t ntot = 0, ncov = 0;
    for (int g = 0; g < count; g++) {
        sizes[g] = descs[g].num_particles;
        maxit[g] = descs[g].mutual_max_iterations;
        target[g] = descs[g].mutual_target_epsilon;
        ntot += sizes[g];
        ncov += descs[g].covalent_offsets[4 * sizes[g]];
    }
    if (ntot > (size_t)INT32_MAX / 4)
        return fail(nullptr, TDP_ERR_INVALID_ARGUMENT, "batch: too many particles");
    std::vector<double> q, dip, pol, th;
    std::vector<int32_t> at, az, ax, ay, off(1, 0), atoms;
    q.reserve(ntot); dip.reserve(3 * ntot); pol.reserve(ntot); th.reserve(ntot);
    at.reserve(ntot); az.reserve(ntot); ax.reserve(ntot); ay.reserve(ntot); off.reserve(4 * ntot + 1);
    atoms.reserve(ncov);
    int p0 = 0;
    for (int g = 0; g < count; g++) {
        const tdp_force_desc& d = descs[g];
        const int n = d.num_particles;
        auto shift = [p0](int32_t a) { return a >= 0 ? a + p0 : a; };
        for (int i = 0; i < n; i++) {
            q.push_back(d.charges[i]);
            for (int k = 0; k < 3; k++) dip.push_back(d.dipoles[3 * i + k]);
            pol.push_back(d.polarizabilities[i]);
            th.push_back(d.thole[i]);
            at.push_back(d.axis_types[i]);
            az.push_back(shift(d.atom_z[i]));
            ax.push_back(shift(d.atom_x[i]));
            ay.push_back(shift(d.atom_y[i]));
            for (int t = 0; t < 4; t++) {
                for (int k = d.covalent_offsets[4 * i + t]; k < d.covalent_offsets[4 * i + t + 1]; k++)
                    atoms.push_back(shift(d.covalent_atoms[k]));
                off.push_back((int32_t)atoms.size());
            }
        }
        p0 += n;
    }
    tdp_force_desc all = descs[0];
    all.num_particles = (int32_t)ntot;
    all.charges = q.data();
    all.dipoles = dip.data();
    all.polarizabilities = pol.data();
    all.thole = th.data();
    all.axis_types = at.data();
    all.atom_z = az.data();
    all.atom_x = ax.data();
    all.atom_y = ay.data();
    all.covalent_offsets = off.data();
    all.covalent_atoms = atoms.data();
    return create_context(&all, sizes, maxit, target, opts, out);
}

void tdp_destroy(tdp_context* c) {
    if (!c) return;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    for (auto& pr : c->pending) {
        hipEventDestroy(pr.first);
        hipEventDestroy(pr.second);
    }
    for (auto e : c->ev_pool) hipEventDestroy(e);
    for (auto& g : c->graphs) hipGraphExecDestroy(g.exec);
    for (auto& g : c->seq_graphs) hipGraphExecDestroy(g.exec);
    for (void* p : c->allocations) hipFree(p);
    if (c->h_stage) hipHostFree(c->h_stage);
    if (c->h_need) hipHostFree(c->h_need);
    if (c->stream2) hipStreamSynchronize(c->stream2);
    if (c->ev_fork) hipEventDestroy(c->ev_fork);
    if (c->ev_join) hipEventDestroy(c->ev_join);
    if (c->ev_stage) hipEventDestroy(c->ev_stage);
    if (c->stream2) hipStreamDestroy(c->stream2);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

int tdp_copy_parameters(tdp_context* c, const tdp_force_desc* d) {
    if (!c || !d) return fail(c, TDP_ERR_INVALID_ARGUMENT, "null argument");
    if (c->nseg != 1) return fail(c, TDP_ERR_INVALID_ARGUMENT, "updateParametersInContext: not available for a batch");
    if (d->num_particles != c->n)
        return fail(c, TDP_ERR_PARTICLE_COUNT, "updateParametersInContext: The number of particles has changed");
    // Only the per-particle parameters are copied (ReferenceTholeDipoleKernels.cpp:317-339).
    int rc = validate(c, d);
    if (rc) return rc;
    TDP_HIP(c, hipSetDevice(c->device));
    TDP_HIP(c, hipStreamSynchronize(c->stream));
    if (c->ewald)
        for (int k = 0; k < 3 * d->num_particles; k++)
            if (d->dipoles[k] != 0.0)
                return fail(c, TDP_ERR_INVALID_ARGUMENT,
                            "TholeDipoleForce: the Ewald option supports charges and induced dipoles only (molecular "
                            "dipoles must be zero)");
    copy_particles(c, d);
    if (c->ewald) ew_particle_terms(c);
    if ((rc = upload_tables(c))) return rc;  // axis atoms may have changed: torque gather lists
    return upload_particles(c);
}

int tdp_execute(tdp_context* c, const double* positions, const double* box, int32_t include_forces,
                int32_t include_energy, double* forces, double* energy) {
    if (!c || (!positions && c->n > 0)) return fail(c, TDP_ERR_INVALID_ARGUMENT, "null argument");
    int rc = check_box(c, box);
    if (rc) return rc;
    hipSetDevice(c->device);
    if ((rc = upload_positions(c, positions))) return rc;
    if ((rc = setup_and_solve(c, box, positions))) return rc;
    const bool want_f = include_forces && forces, want_e = include_energy && energy;
    forces_and_energy(c, box, want_f, want_e);
    if ((rc = download_results(c, want_f ? forces : nullptr, want_e ? energy : nullptr))) return rc;
    return convergence_status(c);
}

int tdp_execute_device(tdp_context* c, const double* d_positions, const double* box, double* d_forces,
                       double* d_energy, void* stream) {
    if (!c) return fail(c, TDP_ERR_INVALID_ARGUMENT, "null context");
    int rc = check_box(c, box);
    if (rc) return rc;
    hipSetDevice(c->device);
    hipStream_t user = reinterpret_cast<hipStream_t>(stream);
    // order the context's stream after the caller's work, run, then order the caller's stream after ours
    hipEvent_t ev = take_event(c);
    TDP_HIP(c, hipEventRecord(ev, user));
    TDP_HIP(c, hipStreamWaitEvent(c->stream, ev, 0));
    if (c->n > 0)
        TDP_HIP(c, hipMemcpyAsync(c->d_pos3, d_positions, 3 * c->n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    // forces and energy into the caller's buffers, guarded by the stop count on the captured path
    const bool seq = sequence_graphs(c);
    auto compute = [&]() -> int {
        forces_and_energy(c, box, d_forces != nullptr, d_energy != nullptr);
        return TDP_OK;
    };
    auto back = [&]() -> int {
        forces_and_energy(c, box, d_forces != nullptr, d_energy != nullptr);
        if (d_forces && c->n > 0)
            launch_scatter_accumulate(c->stream, c->npad, c->d_perm, c->d_frc, d_forces, seq ? c->d_done : nullptr,
                                      c->nseg);
        if (d_energy) {
            if (seq) launch_copy_guarded(c->stream, c->nseg, c->d_energy, d_energy, c->d_done, c->nseg);
            else
                TDP_HIP(c, hipMemcpyAsync(d_energy, c->d_energy, c->nseg * sizeof(double), hipMemcpyDeviceToDevice,
                                          c->stream));
        }
        return TDP_OK;
    };
    bool back_ran = false;
    c->ew_need_pending = false;
    if (seq) {
        const PairLaunch L = pair_launch(c, box);
        auto back_graph = [&]() -> int { return replay_or_capture(c, kSeqBack, 0, d_forces, d_energy, L.box, back); };
        if ((rc = solve_captured(c, box, nullptr, back_graph, &back_ran)) == TDP_OK && !back_ran) rc = back_graph();
    } else if (c->ewald && c->ew_stored && c->polarization == TDP_MUTUAL && !c->timing) {
        // Ewald option with stored lists: the list build's counts are examined at the solver's first
        // host check, and the forces and energy (into the context's buffers) are enqueued behind the
        // first chunk of iterations before that check; they stand when that chunk finished the solve
        // (else they run again).  The copies to the caller follow the check.
        const std::function<int()> ahead = compute;
        c->ew_defer = true;
        rc = setup_and_solve(c, box, nullptr, &ahead, &back_ran);
        c->ew_defer = false;
        if (rc == kEwRedo) rc = setup_and_solve(c, box, nullptr);   // lists regrown: build again, synchronously
        if (rc == TDP_OK && !back_ran) rc = compute();
        if (rc == TDP_OK && d_forces && c->n > 0)
            launch_scatter_accumulate(c->stream, c->npad, c->d_perm, c->d_frc, d_forces, nullptr, c->nseg);
        if (rc == TDP_OK && d_energy)
            TDP_HIP(c, hipMemcpyAsync(d_energy, c->d_energy, c->nseg * sizeof(double), hipMemcpyDeviceToDevice,
                                      c->stream));
    } else {
        rc = setup_and_solve(c, box, nullptr);
        if (rc == TDP_OK) rc = back();
    }
    if (rc) {
        c->ev_pool.push_back(ev);
        return rc;
    }
    TDP_HIP(c, hipEventRecord(ev, c->stream));
    TDP_HIP(c, hipStreamWaitEvent(user, ev, 0));
    c->ev_pool.push_back(ev);
    return convergence_status(c);
}

int tdp_execute_device_noalloc(tdp_context* c, const double* d_positions, const double* box, double* d_forces,
                               double* d_energy, void* stream) {
    if (!c) return fail(c, TDP_ERR_INVALID_ARGUMENT, "null context");
    if (c->n > 0 && !d_positions) return fail(c, TDP_ERR_INVALID_ARGUMENT, "null positions");
    if (c->timing) return fail(c, TDP_ERR_INVALID_ARGUMENT, "allocation-free evaluation: disable timing first");
    if (c->ewald) return fail(c, TDP_ERR_INVALID_ARGUMENT, "allocation-free evaluation: not available with the Ewald option");
    int rc = check_box(c, box);
    if (rc) return rc;
    hipSetDevice(c->device);
    unsigned char* base = static_cast<unsigned char*>(scratch_ptr(c->device));
    if (!base) return fail(c, TDP_ERR_HIP, "allocation-free evaluation: no scratch address for this device");
    // the context's evaluation partials, carved from the scratch in 256-byte aligned pieces
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t pb = ((size_t)c->nwork + c->ntri) * 64 * sizeof(double4);
    const size_t pe = ((size_t)c->nwork + 1) * sizeof(double);
    const size_t p32 = c->precision == TDP_PRECISION_MIXED ? 3 * ((size_t)c->ntri + c->nwork) * 64 * sizeof(float) : 0;
    const size_t need = up(pb) + up(pb) + up(pe) + up(p32);
    if (need > kNoallocScratchBytes) {
        char buf[240];
        snprintf(buf, sizeof buf, "allocation-free evaluation: the partial sums of this context need %zu bytes, more "
                 "than the %zu bytes of scratch (about 7,000 particles in one system)", need, kNoallocScratchBytes);
        return fail(c, TDP_ERR_INVALID_ARGUMENT, buf);
    }
    hipStream_t user = reinterpret_cast<hipStream_t>(stream);
    hipEvent_t ev = take_event(c);   // from the pool filled at create
    TDP_HIP(c, hipEventRecord(ev, user));
    TDP_HIP(c, hipStreamWaitEvent(c->stream, ev, 0));
    void* saved_pbuf = c->d_pbuf;
    double4* saved_pbuf2 = c->d_pbuf2;
    double* saved_part_e = c->d_part_e;
    float* saved_slab32 = c->d_slab32;
    c->d_pbuf = base;
    c->d_pbuf2 = reinterpret_cast<double4*>(base + up(pb));
    c->d_part_e = reinterpret_cast<double*>(base + 2 * up(pb));
    c->d_slab32 = p32 ? reinterpret_cast<float*>(base + 2 * up(pb) + up(pe)) : nullptr;
    c->noalloc = true;
    // the previous call's contents: start from zero, as a fresh context's buffers do
    hipLaunchKernelGGL(k_zero16, dim3(1024), dim3(256), 0, c->stream, reinterpret_cast<uint4*>(base), need / 16);
    rc = TDP_OK;
    if (c->n > 0) {
        hipError_t e = hipMemcpyAsync(c->d_pos3, d_positions, 3 * c->n * sizeof(double), hipMemcpyDeviceToDevice,
                                      c->stream);
        if (e != hipSuccess) rc = fail(c, TDP_ERR_HIP, std::string("hipMemcpyAsync: ") + hipGetErrorString(e));
    }
    if (!rc) rc = setup_and_solve(c, box, nullptr);
    if (!rc) {
        forces_and_energy(c, box, d_forces != nullptr, d_energy != nullptr);
        if (d_forces && c->n > 0) launch_scatter_accumulate(c->stream, c->npad, c->d_perm, c->d_frc, d_forces);
        if (d_energy) {
            hipError_t e = hipMemcpyAsync(d_energy, c->d_energy, c->nseg * sizeof(double), hipMemcpyDeviceToDevice,
                                          c->stream);
            if (e != hipSuccess) rc = fail(c, TDP_ERR_HIP, std::string("hipMemcpyAsync: ") + hipGetErrorString(e));
        }
    }
    // the caller's stream waits for this call's work, which reads the scratch until it ends
    hipEventRecord(ev, c->stream);
    hipStreamWaitEvent(user, ev, 0);
    c->ev_pool.push_back(ev);
    c->noalloc = false;
    c->d_pbuf = saved_pbuf;
    c->d_pbuf2 = saved_pbuf2;
    c->d_part_e = saved_part_e;
    c->d_slab32 = saved_slab32;
    if (rc) return rc;
    return convergence_status(c);
}

int tdp_get_neighbor_pairs(tdp_context* c, const double* positions, const double* box, int64_t capacity,
                           int32_t* pairs, int64_t* count) {
    if (!c || !count || (!positions && c->n > 0) || capacity < 0 || (capacity > 0 && !pairs))
        return fail(c, TDP_ERR_INVALID_ARGUMENT, "null argument");
    if (!c->ewald) return fail(c, TDP_ERR_INVALID_ARGUMENT, "neighbour pairs: only the Ewald option has a cutoff");
    int rc = check_box(c, box);
    if (rc) return rc;
    hipSetDevice(c->device);
    *count = 0;
    if (c->n == 0) return TDP_OK;
    if ((rc = upload_positions(c, positions))) return rc;
    hipStream_t s = c->stream;
    if (!c->ordered || c->evals_since_order >= kReorderInterval) {
        TDP_HIP(c, hipStreamSynchronize(s));
        if ((rc = reorder(c, positions, box))) return rc;
    }
    PairLaunch L = pair_launch(c, box);
    launch_load_positions(s, c->npad, c->d_pos3, c->d_perm, c->d_thole, c->posa);
    launch_tile_bounds(s, c->d_tcnt, c->nblk, c->posa, L.box, L.pbc, c->tb_c, c->tb_h, c->tb_o, c->posl, c->posl64);
    const EwLaunch E = ew_launch(c, box);
    launch_ew_neighbors(s, E);
    int* d_pairs = nullptr;
    unsigned long long* d_count = nullptr;
    if ((rc = dalloc(c, &d_pairs, 2 * (size_t)std::max<int64_t>(capacity, 1))) || (rc = dalloc(c, &d_count, 1))) return rc;
    launch_ew_dump_pairs(s, E, c->posa, c->d_perm, d_pairs, capacity, d_count);
    unsigned long long n = 0;
    TDP_HIP(c, hipMemcpyAsync(&n, d_count, sizeof n, hipMemcpyDeviceToHost, s));
    TDP_HIP(c, hipStreamSynchronize(s));
    if (capacity > 0)
        TDP_HIP(c, hipMemcpy(pairs, d_pairs, 2 * sizeof(int32_t) * (size_t)std::min<int64_t>((int64_t)n, capacity),
                             hipMemcpyDeviceToHost));
    for (void* p : {(void*)d_pairs, (void*)d_count}) {
        auto it = std::find(c->allocations.begin(), c->allocations.end(), p);
        if (it != c->allocations.end()) c->allocations.erase(it);
        hipFree(p);
    }
    *count = (int64_t)n;
    return TDP_OK;
}

int tdp_get_dipoles(tdp_context* c, const double* positions, const double* box, int32_t kind, double* out) {
    if (!c || !out || (!positions && c->n > 0)) return fail(c, TDP_ERR_INVALID_ARGUMENT, "null argument");
    if (kind < 0 || kind > 2) return fail(c, TDP_ERR_INVALID_ARGUMENT, "unknown dipole kind");
    int rc = check_box(c, box);
    if (rc) return rc;
    hipSetDevice(c->device);
    if ((rc = upload_positions(c, positions))) return rc;
    if ((rc = setup_and_solve(c, box, positions))) return rc;
    const int npad = c->npad;
    std::vector<double4> mu(npad), qd(npad);
    if (c->n > 0) {
        TDP_HIP(c, hipMemcpyAsync(mu.data(), c->mu, npad * sizeof(double4), hipMemcpyDeviceToHost, c->stream));
        TDP_HIP(c, hipMemcpyAsync(qd.data(), c->qd, npad * sizeof(double4), hipMemcpyDeviceToHost, c->stream));
    }
    TDP_HIP(c, hipStreamSynchronize(c->stream));
    for (int i = 0; i < (c->n > 0 ? npad : 0); i++) {
        double a[3] = {mu[i].x, mu[i].y, mu[i].z}, b[3] = {qd[i].x, qd[i].y, qd[i].z};
        int o = c->perm[i];
        if (o < 0) continue;
        for (int k = 0; k < 3; k++)
            out[3 * o + k] = kind == TDP_DIPOLES_INDUCED ? a[k] : (kind == TDP_DIPOLES_LAB_PERMANENT ? b[k] : b[k] + a[k]);
    }
    return convergence_status(c);
}

int tdp_get_electrostatic_potential(tdp_context* c, const double* positions, const double* box, const double* grid,
                                    int32_t num_grid, double* out) {
    if (!c || (!positions && c->n > 0) || (num_grid > 0 && (!grid || !out)) || num_grid < 0)
        return fail(c, TDP_ERR_INVALID_ARGUMENT, "null argument");
    if (c->nseg != 1) return fail(c, TDP_ERR_INVALID_ARGUMENT, "getElectrostaticPotential: not available for a batch");
    if (c->ewald)
        return fail(c, TDP_ERR_INVALID_ARGUMENT, "getElectrostaticPotential: not available with the Ewald option");
    int rc = check_box(c, box);
    if (rc) return rc;
    hipSetDevice(c->device);
    if ((rc = upload_positions(c, positions))) return rc;
    if ((rc = setup_and_solve(c, box, positions))) return rc;
    if (num_grid == 0) return TDP_OK;
    size_t gb = (size_t)num_grid * 3 * sizeof(double);
    if ((rc = ensure_out(c, gb + num_grid * sizeof(double)))) return rc;
    double* d_grid = c->d_out;
    double* d_pot = c->d_out + 3 * (size_t)num_grid;
    TDP_HIP(c, hipMemcpyAsync(d_grid, grid, gb, hipMemcpyHostToDevice, c->stream));
    double b3[3] = {box ? box[0] : 0.0, box ? box[4] : 0.0, box ? box[8] : 0.0};
    launch_potential(c->stream, c->n, num_grid, d_grid, c->posa, c->qd, c->mu, c->method == TDP_PME, b3, d_pot);
    TDP_HIP(c, hipMemcpyAsync(out, d_pot, num_grid * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    TDP_HIP(c, hipStreamSynchronize(c->stream));
    return TDP_OK;
}

int tdp_get_system_multipole_moments(tdp_context* c, const double* positions, const double* box, const double* masses,
                                     double* out) {
    if (!c || !out || !masses || (!positions && c->n > 0)) return fail(c, TDP_ERR_INVALID_ARGUMENT, "null argument");
    if (c->nseg != 1) return fail(c, TDP_ERR_INVALID_ARGUMENT, "getSystemMultipoleMoments: not available for a batch");
    int rc = check_box(c, box);
    if (rc) return rc;
    hipSetDevice(c->device);
    if ((rc = upload_positions(c, positions))) return rc;
    if ((rc = setup_and_solve(c, box, positions))) return rc;
    std::vector<double> mperm(c->n);
    for (int k = 0; k < c->n; k++) mperm[k] = masses[c->perm[k]];
    if (c->n > 0)
        TDP_HIP(c, hipMemcpyAsync(c->d_mass, mperm.data(), c->n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if ((rc = ensure_out(c, 16 * sizeof(double)))) return rc;
    launch_moments(c->stream, c->n, c->posa, c->qd, c->mu, c->d_mass, c->d_out);
    double raw[10];
    TDP_HIP(c, hipMemcpyAsync(raw, c->d_out, sizeof raw, hipMemcpyDeviceToHost, c->stream));
    TDP_HIP(c, hipStreamSynchronize(c->stream));
    // unit conversion of calculateTholeDipoleSystemMultipoleMoments (:1285-1312)
    double xx = raw[4], xy = raw[5], xz = raw[6], yy = raw[7], yz = raw[8], zz = raw[9];
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
    out[0] = raw[0];
    out[1] = raw[1] * 10.0 * debye;
    out[2] = raw[2] * 10.0 * debye;
    out[3] = raw[3] * 10.0 * debye;
    debye *= 3.0;
    for (int k = 4; k < 13; k++) out[k] *= 100.0 * debye;
    return TDP_OK;
}

int tdp_get_pme_parameters(const tdp_context* cc, double* alpha, int32_t* nx, int32_t* ny, int32_t* nz) {
    tdp_context* c = const_cast<tdp_context*>(cc);
    if (!c || !alpha || !nx || !ny || !nz) return fail(c, TDP_ERR_INVALID_ARGUMENT, "null argument");
    if (c->method != TDP_PME)
        return fail(c, TDP_ERR_NOT_PME, "getPMEParametersInContext: This Context is not using PME");
    double a;
    int g[3];
    pme_parameters(c, a, g);
    *alpha = a;
    *nx = g[0];
    *ny = g[1];
    *nz = g[2];
    return TDP_OK;
}

int tdp_get_ewald_info(const tdp_context* cc, tdp_ewald_info* info) {
    tdp_context* c = const_cast<tdp_context*>(cc);
    if (!c || !info) return fail(c, TDP_ERR_INVALID_ARGUMENT, "null argument");
    if (!c->ewald) return fail(c, TDP_ERR_INVALID_ARGUMENT, "the context does not use the Ewald option");
    *info = tdp_ewald_info{};
    info->alpha = c->ew_alpha;
    for (int k = 0; k < 3; k++) info->grid[k] = c->ew_K[k];
    info->cutoff = c->cutoff;
    info->r_fp64 = c->precision == TDP_PRECISION_MIXED ? c->ew_r64 : c->cutoff;
    info->stored = c->ew_stored ? 1 : 0;
    info->entries32 = c->ew_entries[0];
    info->entries64 = c->ew_entries[1];
    info->slots32 = c->ew_slots[0];
    info->slots64 = c->ew_slots[1];
    info->bytes_per_entry32 = (int32_t)(sizeof(unsigned short) + sizeof(float2));
    info->bytes_per_entry64 = (int32_t)(sizeof(int) + sizeof(double4));
    info->tiles = c->nblk;
    TDP_HIP(c, hipSetDevice(c->device));
    std::vector<int> cnt(c->nblk);
    TDP_HIP(c, hipMemcpyAsync(cnt.data(), c->d_nbr_cnt, c->nblk * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    TDP_HIP(c, hipStreamSynchronize(c->stream));
    double sum = 0.0;
    for (int v : cnt) {
        info->max_neighbor_tiles = std::max(info->max_neighbor_tiles, v);
        sum += v;
    }
    info->mean_neighbor_tiles = c->nblk > 0 ? sum / c->nblk : 0.0;
    return TDP_OK;
}

int tdp_get_solver_stats(const tdp_context* cc, tdp_solver_stats* st) {
    tdp_context* c = const_cast<tdp_context*>(cc);
    if (!c || !st) return fail(c, TDP_ERR_INVALID_ARGUMENT, "null argument");
    if (c->timing) {
        hipStreamSynchronize(c->stream);
        harvest_timing(c);
    }
    st->iterations = c->iterations[0];
    st->epsilon = c->epsilon[0];
    st->converged = c->converged[0];
    st->matvec_ms = c->matvec_ms;
    st->matvec_launches = c->matvec_launches;
    // launches enqueued after a solve stopped return at once: the full passes are those taking at
    // least half the median launch
    st->matvec_ms_full = 0.0;
    st->matvec_full_launches = 0;
    if (!c->matvec_each.empty()) {
        std::vector<float> t = c->matvec_each;
        std::nth_element(t.begin(), t.begin() + t.size() / 2, t.end());
        const float half = 0.5f * t[t.size() / 2];
        for (float ms : c->matvec_each)
            if (ms >= half) {
                st->matvec_ms_full += ms;
                st->matvec_full_launches++;
            }
    }
    return TDP_OK;
}

int tdp_batch_get_solver_stats(const tdp_context* cc, int32_t system, tdp_solver_stats* st) {
    tdp_context* c = const_cast<tdp_context*>(cc);
    if (!c || !st) return fail(c, TDP_ERR_INVALID_ARGUMENT, "null argument");
    if (system < 0 || system >= c->nseg) return fail(c, TDP_ERR_INVALID_ARGUMENT, "system index out of range");
    int rc = tdp_get_solver_stats(cc, st);
    if (rc) return rc;
    st->iterations = c->iterations[system];
    st->epsilon = c->epsilon[system];
    st->converged = c->converged[system];
    return TDP_OK;
}

int32_t tdp_batch_num_systems(const tdp_context* c) { return c ? c->nseg : 0; }

int tdp_batch_solve_responses(tdp_context* c, const double* positions, int32_t method, double tolerance,
                              int32_t max_steps, double* out, tdp_dense_stats* stats) {
    if (!c || (!positions && c->n > 0)) return fail(c, TDP_ERR_INVALID_ARGUMENT, "null argument");
    if (c->method != TDP_NO_CUTOFF) return fail(c, TDP_ERR_INVALID_ARGUMENT, "responses: NoCutoff systems only");
    if (method < 0 || method > 2) return fail(c, TDP_ERR_INVALID_ARGUMENT, "responses: unknown method");
    if (!(tolerance > 0.0) || max_steps <= 0) return fail(c, TDP_ERR_INVALID_ARGUMENT, "responses: bad tolerance");
    hipSetDevice(c->device);
    int rc;
    if ((rc = upload_positions(c, positions))) return rc;
    if (c->n == 0) return TDP_OK;
    if ((rc = prepare_field(c, nullptr, positions))) return rc;
    if (method == 2) return cg_responses(c, tolerance, max_steps, out, stats);
    return dense_responses(c, method == 1, tolerance, max_steps, out, stats);
}

int tdp_batch_execute(tdp_context* c, const double* positions, double* forces, double* energies) {
    if (!c || (!positions && c->n > 0)) return fail(c, TDP_ERR_INVALID_ARGUMENT, "null argument");
    return tdp_execute(c, positions, nullptr, 1, 1, forces, energies);
}

int tdp_batch_execute_device(tdp_context* c, const double* d_positions, double* d_forces, double* d_energies,
                             void* stream) {
    return tdp_execute_device(c, d_positions, nullptr, d_forces, d_energies, stream);
}

int tdp_autotune(tdp_context* c, const double* positions, const double* box, int32_t* tiles_per_item) {
    if (!c || (!positions && c->n > 0)) return fail(c, TDP_ERR_INVALID_ARGUMENT, "null argument");
    int rc = check_box(c, box);
    if (rc) return rc;
    hipSetDevice(c->device);
    if (c->n > 0 && !c->ewald) {   // (the Ewald option's passes have no work-item schedule)
        if ((rc = upload_positions(c, positions))) return rc;
        if ((rc = prepare_field(c, box, positions))) return rc;   // internal order, tile tables, mu = alpha E0
        int longest[kSizeBuckets] = {};   // longest system (tiles) of each bucket, 0 when the bucket is empty
        for (int g = 0; g < c->nseg; g++) {
            const int t = c->seg_t1[g] - c->seg_t0[g];
            if (t > 0) longest[size_bucket(t)] = std::max(longest[size_bucket(t)], t);
        }
        static const int cand[] = {1, 2, 3, 4, 6, 8, 12, 16};
        constexpr int kCand = sizeof(cand) / sizeof(cand[0]);
        hipEvent_t e0 = take_event(c), e1 = take_event(c);
        // one schedule's time: a warm-up pass, then three timed passes
        auto time_schedule = [&](const int* Cb, float* ms) -> int {
            int r = set_work_items_buckets(c, Cb);
            if (r) return r;
            PairLaunch L = pair_launch(c, box);
            matvec(c, L, c->mu, nullptr);
            TDP_HIP(c, hipEventRecord(e0, c->stream));
            for (int k = 0; k < 3; k++) matvec(c, L, c->mu, nullptr);
            TDP_HIP(c, hipEventRecord(e1, c->stream));
            TDP_HIP(c, hipEventSynchronize(e1));
            *ms = 0.f;
            hipEventElapsedTime(ms, e0, e1);
            return TDP_OK;
        };
        // Sweep of the candidates for the buckets in `which` (all buckets set to the same length when
        // which < 0), the others held at Cb: two rounds, a candidate's time the smaller of its two, so that
        // one disturbed measurement does not decide.  Leaves the best length in Cb.
        auto sweep = [&](int which, int* Cb) -> int {
            int lim = 0;
            for (int b = 0; b < kSizeBuckets; b++)
                if (which < 0 || which == b) lim = std::max(lim, longest[b]);
            float cand_ms[kCand];
            for (int q = 0; q < kCand; q++) cand_ms[q] = 0.f;
            int trial[kSizeBuckets];
            for (int round = 0; round < 2; round++)
                for (int q = 0; q < kCand; q++) {
                    if (cand[q] > lim && cand[q] != 1) break;
                    for (int b = 0; b < kSizeBuckets; b++) trial[b] = (which < 0 || which == b) ? cand[q] : Cb[b];
                    float ms;
                    int r = time_schedule(trial, &ms);
                    if (r) return r;
                    if (round == 0 || ms < cand_ms[q]) cand_ms[q] = ms;
                }
            int best = -1;
            for (int q = 0; q < kCand; q++)
                if (cand_ms[q] > 0.f && (best < 0 || cand_ms[q] < cand_ms[best])) best = q;
            if (best >= 0)
                for (int b = 0; b < kSizeBuckets; b++)
                    if (which < 0 || which == b) Cb[b] = cand[best];
            return TDP_OK;
        };
        int Cb[kSizeBuckets];
        for (int b = 0; b < kSizeBuckets; b++) Cb[b] = c->work_Cb[b];
        // one length for every system first, then, for a batch whose systems fall into several buckets,
        // one pass of coordinate descent over the occupied buckets from the largest systems down
        if ((rc = sweep(-1, Cb))) return rc;
        int occupied = 0;
        for (int b = 0; b < kSizeBuckets; b++) occupied += longest[b] > 0;
        if (occupied > 1)
            for (int b = kSizeBuckets - 1; b >= 0; b--)
                if (longest[b] > 0 && (rc = sweep(b, Cb))) return rc;
        c->ev_pool.push_back(e0);
        c->ev_pool.push_back(e1);
        if ((rc = set_work_items_buckets(c, Cb))) return rc;
    }
    if (tiles_per_item) *tiles_per_item = c->work_C;
    return TDP_OK;
}

int tdp_get_work_schedule(const tdp_context* c, int32_t* tiles_per_item, int32_t* tiles_per_system_bound) {
    if (!c || !tiles_per_item) return TDP_ERR_INVALID_ARGUMENT;
    for (int b = 0; b < kSizeBuckets; b++) {
        tiles_per_item[b] = c->work_Cb[b];
        if (tiles_per_system_bound) tiles_per_system_bound[b] = b == kSizeBuckets - 1 ? INT32_MAX : 2 << b;
    }
    return TDP_OK;
}

int tdp_set_work_schedule(tdp_context* c, const int32_t* tiles_per_item) {
    if (!c || !tiles_per_item) return fail(c, TDP_ERR_INVALID_ARGUMENT, "null argument");
    int Cb[kSizeBuckets];
    for (int b = 0; b < kSizeBuckets; b++) {
        if (tiles_per_item[b] < 1 || tiles_per_item[b] > 4096)
            return fail(c, TDP_ERR_INVALID_ARGUMENT, "tiles per work item must lie in [1, 4096]");
        Cb[b] = tiles_per_item[b];
    }
    if (c->ewald) return TDP_OK;   // (the Ewald option's passes have no work-item schedule)
    hipSetDevice(c->device);
    return set_work_items_buckets(c, Cb);
}

int tdp_set_warm_start(tdp_context* c, int32_t enable) {
    if (!c) return fail(c, TDP_ERR_INVALID_ARGUMENT, "null context");
    hipSetDevice(c->device);
    if (enable && !c->d_mu_prev) {
        int rc = dalloc(c, &c->d_mu_prev, 3 * (size_t)std::max(c->n, 1));
        if (rc) return rc;
    }
    c->warm_start = enable != 0;
    c->have_prev = false;
    return TDP_OK;
}

int tdp_set_timing(tdp_context* c, int32_t enable) {
    if (!c) return fail(c, TDP_ERR_INVALID_ARGUMENT, "null context");
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    harvest_timing(c);
    c->timing = enable != 0;
    c->matvec_ms = 0.0;
    c->matvec_launches = 0;
    c->matvec_each.clear();
    return TDP_OK;
}

}  // extern "C"
