// tdp_kernels.h -- launch wrappers of the gfx950 kernels (tdp_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace tdp {

// Solver scalars kept in device memory so an iteration needs no host round trip.
enum SolverScalars {
    SC_DONE = 0,    // != 0: the solve has stopped; iteration kernels return immediately
    SC_ITER = 1,    // iterations done
    SC_MAXIT = 2,   // maximum iterations
    SC_TARGET = 3,  // target RMS dipole update (e*nm)
    SC_EPS = 4,     // current RMS dipole update
    SC_RZ = 5,      // PCG r.z
    SC_RZOLD = 6,   // PCG r.z of the previous iteration (kept by the update kernel for the direction kernel)
    SC_COUNT = 7
};

// Fused PCG step in the reduction of T*p: q = p/alpha - T p (alpha = 0: q = -T p) to q and the tile
// sums of p.q to red[T]; off when q is null.
struct PcgQ {
    const double4* p = nullptr;
    const double* alpha = nullptr;
    double4* q = nullptr;
    double* red = nullptr;
};

// Tiles of a context that may hold several independent systems (segments), each in a
// contiguous range of whole 64-slot tiles; slots past a system's particles are padding.
struct TileMap {
    const int* cnt;      // [nblk] valid particles in the tile (64 except a system's last tile)
    const int* first;    // [nblk] first tile of the tile's system
    const int* seg;      // [nblk] system of the tile
    const size_t* col;   // [nblk] tile-partial slab index of tile (first[J], J)
};

// Systems for the solver kernels: tiles [seg.x, seg.y) and seg.z particles of system s, its pair
// work items [segw.x, segw.y); done[0] counts systems whose solve has stopped.
struct SegLaunch {
    int nseg, ntile, npad;
    const int* tseg;
    const int4* seg;
    const int2* segw;
    int* done;
};

// One system of the dense path: nt particle tiles from tile t0; its 16-bit matrix (192 nt)^2 at
// moff, its diagonal-block inverses (3 nt blocks of 64x64 FP32) at doff, its refinement vectors
// ((192 nt) x 4 doubles) at voff.
struct DenseSys {
    int nt, t0;
    size_t moff, voff, doff;
};

// Geometry of one all-pairs launch: nblk tiles of 64 particles.
// Classification of tile pair (I, J), I <= J, computed once per evaluation (launch_tile_pairs) and
// read by the pair kernels: the FP32 frame (shifts sh, undecidable bands b of the straddling axes),
// whether every pair lies beyond the Thole range, and the squared radii beyond which pairs carry
// no damping in FP32 (r2min) or may be evaluated in FP32 with their damping (r2d).  Indexed like
// the tile partial slab.
struct TilePair {
    double shx, shy, shz;
    float bx, by, bz;
    unsigned flags;   // bits 0-2: straddling axes; kTpGeneral: no FP32 frame; kTpFar: beyond the Thole range
    float r2min, r2d;
};
constexpr unsigned kTpStraddle = 7u, kTpGeneral = 8u, kTpFar = 16u;

struct PairLaunch {
    int npad, nblk;
    TileMap tm;
    bool pbc;
    bool perm;     // some particle has a permanent dipole (else the energy pass forms no torques)
    double box[3];
    const int* exc_start;                 // [nblk+1]
    const int* exc_J;                     // [ntiles]
    const unsigned long long* exc_mask;   // [ntiles][3][64]
    const float4* tb_c;                   // [nblk] tile bounding-box centre, largest Thole parameter
    const float4* tb_h;                   // [nblk] tile bounding-box half extent, smallest Thole parameter
    const double4* tb_o;                  // [nblk] tile origin (FP64) of the FP32 far-field paths
    const float4* posl;                   // [npad] particle offset from its tile origin (FP32)
    const double4* posl64;                // [npad] the same offset in FP64 (the FP64 far path's frames)
    const TilePair* tp;                   // [ntri] tile pair classification (launch_tile_pairs)
    // upper-triangle sweep: work item w = tiles (I, J0..J1) as int4 (I, J0, J1, 0); items of tile row I
    // are row_start[I] .. row_start[I+1]
    int nwork;
    const int4* work;
    const int* row_start;
    size_t slab;   // 64 * number of tile pairs: stride between components of the tile partials
};

// Pairwise passes over the upper-triangle work items.  Pi: nwork*64 item partials; Pj: the
// column-major tile partial slab, 3 * L.slab doubles.  The T*mu passes leave their partials for
// launch_col_reduce, which writes T*mu to out4 (double4 per particle) or out3 (3 doubles).  The mixed
// pass keeps its partials in FP32 in Pjf: the tile slab (3 * L.slab floats), then the item partials
// (3 * nwork * 64 floats, SoA); the reduction after it reads them (Pjf non-null) in place of Pj, Pi.
void launch_sym_matvec64(hipStream_t s, const PairLaunch& L, const double4* posa, const double4* mu, const double* inva,
                         const double* scalars, double4* Pi, double* Pj);
void launch_col_reduce(hipStream_t s, const PairLaunch& L, const double4* Pi, const double* Pj, const double* scalars,
                       double4* out4, double* out3, const float* Pjf = nullptr, const PcgQ& pq = PcgQ());
void launch_sym_matvec_mixed(hipStream_t s, const PairLaunch& L, const double4* posa, const double4* mu,
                             const float4* mu32, const double* inva, const double* scalars, float* Pjf);

// Classification of every tile pair of the work items into L.tp (after launch_tile_bounds).
void launch_tile_pairs(hipStream_t s, const PairLaunch& L, TilePair* tp);
// tb_o: tile origins (FP64), posl: FP32 offsets of the particles from their tile origin, posl64 (may be
// null) the same offsets in FP64
void launch_tile_bounds(hipStream_t s, const int* tcnt, int nblk, const double4* posa, const double* box3, bool pbc,
                        float4* tb_c, float4* tb_h, double4* tb_o, float4* posl, double4* posl64);

// posa: FP64 positions (internal order) and Thole parameter; slots with perm < 0 are padding.
void launch_load_positions(hipStream_t s, int npad, const double* pos3, const int* perm, const double* thole,
                           double4* posa);
// caller's order (3 doubles per particle) <-> internal order (double4 per slot; dst32 may be null)
void launch_gather_vec3(hipStream_t s, int npad, const int* perm, const double* src3, double4* dst4, float4* dst32);
void launch_scatter_vec3(hipStream_t s, int npad, const int* perm, const double4* src4, double* dst3);
void launch_rotate(hipStream_t s, int n, const double4* posa, const double* dmol, const double* charge,
                   const int4* axis, double4* qd);
void launch_sym_field(hipStream_t s, const PairLaunch& L, const double4* posa, const double4* qd, const double* inva,
                      double4* Pi, double* Pj, double4* out);
// Energy, forces (accumulated into frc, 3 doubles per slot) and, when L.perm, torques (torque).
// mixed without permanent dipoles: the force partials go to slab32 (FP32, the layout of
// launch_sym_matvec_mixed's Pjf) instead of PiF/PjF.
void launch_sym_elec(hipStream_t s, const PairLaunch& L, bool mutual, bool mixed, const double4* posa, const double4* qd,
                     const double4* mu, const double* inva, const int* orig, double4* PiF, double* PjF, double4* PiT,
                     double* PjT, double* part_e, double* frc, double4* torque, float* slab32);
void launch_field_init(hipStream_t s, const SegLaunch& G, const double4* part, const double* alpha, double4* E0,
                       double4* mu, float4* mu32);
// red: 3 * ntile doubles; scalars: SC_COUNT per system
void launch_jacobi_update(hipStream_t s, const SegLaunch& G, const double4* part, const double4* E0,
                          const double* alpha, double4* mu, float4* mu32, double* scalars, double* red);
void launch_pcg_init(hipStream_t s, const SegLaunch& G, const double4* part, const double4* E0, const double* alpha,
                     const double4* x, double4* r, double4* p, float4* p32, double* red, double* scalars);
// The rest of a PCG iteration after the reduction with PcgQ (q and its p.q tile sums in
// red + 2 ntile): update of x, r, z and the direction p, two launches.
void launch_pcg_iteration_tail(hipStream_t s, const SegLaunch& G, const double* alpha, double4* x, double4* r,
                               double4* z, double4* p, float4* p32, const double4* q, double* red, double* scalars);
void launch_extrap_step(hipStream_t s, const SegLaunch& G, const double4* part, const double4* E0,
                        const double* alpha, double4* mu, float4* mu32, double4* hist_o);
void launch_extrap_combine(hipStream_t s, const SegLaunch& G, int order, const double4* hist, const double* coef,
                           double4* mu, float4* mu32);
void launch_extrap_error(hipStream_t s, const SegLaunch& G, const double4* part, const double4* E0,
                         const double* alpha, const double4* mu, double* red, double* scalars);
void launch_reduce_energy(hipStream_t s, const SegLaunch& G, const double* part_e, double* out);
void launch_map_torque(hipStream_t s, int n, const double4* posa, const int4* axis, const double4* torque,
                       double4* contrib);
void launch_gather_torque(hipStream_t s, int n, const int* goff, const int* gidx, const double4* contrib, double* frc);
// dst[perm[k]] += src[k]; with done (device stop count) only once *done >= nseg.
void launch_scatter_accumulate(hipStream_t s, int n, const int* perm, const double* src, double* dst,
                               const int* done = nullptr, int nseg = 0);
// dst[k] = src[k] (k < n) once *done >= nseg.
void launch_copy_guarded(hipStream_t s, int n, const double* src, double* dst, const int* done, int nseg);
void launch_potential(hipStream_t s, int n, int ngrid, const double* grid, const double4* posa, const double4* qd,
                      const double4* mu, bool pbc, const double* box3, double* out);
void launch_moments(hipStream_t s, int n, const double4* posa, const double4* qd, const double4* mu,
                    const double* mass, double* out);

// ---- Ewald option (tdp_ewald.hip): real space within the cutoff over neighbour tiles, smooth PME
// (order 5) in reciprocal space.
constexpr int kEwOrder = 5;
constexpr int kEwMaxStages = 12;
constexpr int kEwWaves = 8;   // waves per workgroup of the real-space passes
constexpr int kEwParts = 2;   // workgroups per tile row of the T v and force passes (part h takes the
                              // neighbour tiles h, h + kEwParts, ...); the E0 pass has one
constexpr int ew_parts(int mode) { return mode == 0 ? 1 : kEwParts; }   // mode: EwMode (kEwE0 is 0)
constexpr int kEwE0Waves = 4;   // waves per workgroup of the list-building E0 pass
constexpr int kEwE0Parts = kEwWaves / kEwE0Waves;   // its workgroups per tile row
constexpr int kEwSpmvU = 8;   // entry pairs per step of the stored T v pass (FP32 lists)
constexpr int kEwMaxLocalTiles = 4096;   // neighbour tiles an FP32 list index can address
constexpr int kEwMaxSlots = 16383;        // slots an FP32 list's table can hold (14-bit candidate field)
constexpr int kEwSlotLds = 24;   // LDS bytes per staged slot of the FP32 T v pass: (v, y_x), (y_y, y_z)
// FFT plan of one axis: n = prod r[0..nst), twiddles tw[q] = exp(-2 pi i q / n) (FP64 and FP32).
struct EwAxis {
    int n, nst;
    int r[kEwMaxStages];
    const double2* tw64;
    const float2* tw32;
};

// Real-space pass modes: the charges' field (E0), the dipoles' field (T v), or energy and forces.
enum EwMode { kEwE0 = 0, kEwTmu = 1, kEwForce = 2 };

struct EwLaunch {
    int npad, nblk;
    const int* tcnt;
    double L[3];            // box diagonal
    double iL[3];           // 1 / L
    double rc2;             // cutoff squared (FP64 neighbour criterion)
    double r64sq;           // mixed precision: pairs closer than this (squared) stay in FP64
    bool pairs32;           // mixed precision: real-space pairs beyond r64 in FP32 (T v and forces)
    float cut2_tiles;       // tile culling radius squared (cutoff plus margin), FP32
    double alpha, c1, c2, c3, self_field, self_energy;   // Ewald constants
    const float4* tb_c;
    const float4* tb_h;
    int* nbr;               // [nblk][nblk] neighbour tiles of each tile row (J ascending, J == I included)
    int* nbr_cnt;           // [nblk]
    int* nbr_exc;           // [nblk][nblk] beside nbr: the tile pair's exclusion-tile entry, or -1
    const int* exc_start;
    const int* exc_J;
    const unsigned long long* exc_mask;
    const int* sc_off;      // [npad+1] scaled partners of each slot (internal order)
    const int* sc_j;
    const double2* sc_s;    // (mScale, iScale)
    int K[3];
    EwAxis ax[3];
    const double* G64;      // influence function [K0][K1][K2]
    const float* G32;
    unsigned long long* qfix;   // fixed-point spreading grid (int64, scale 2^40), zero between uses
    void* cgrid;            // complex work grid (double2 or float2)
    void* phi_q;            // potential grid of the charges (kept from the E0 pass)
    void* phi_v;            // potential grid of the last dipole pass
    double4* rs;            // [kEwParts npad] real-space sums of the last pass per row part (field or force,
                            // energy in w)
    double4* rsw;           // [nblk kEwWaves 64] partial T v sums of the stored lists
    // Stored real-space blocks of T (built by the E0 pass, applied by every T v pass): per tile row I
    // and wave w a list of (j, T_ij) for each of the row's 64 slots, sliced ELL: entry n of slot
    // lane at ((I kEwWaves + w) cap + n) 64 + lane, T_ij = C2 d d^T - C1 1.
    // FP64 lists (every pair in FP64 mode, the pairs closer than r64 in mixed precision, and in
    // mixed precision the pairs of tile pairs whose image is not that of their tiles' centres) hold
    // (-C1, u) with u = sqrt(|C2|) d and an index with the sign of C2 in bit 31, so that
    // T_ij v = -C1 v + sign(C2) u (u . v): 4 + 32 bytes per entry.  The index is j itself in mixed
    // precision and, in FP64 mode, j's place among the wave's slots (as for the FP32 lists; the T v
    // pass stages those v_j in LDS).
    // FP32 lists (the other pairs in mixed precision) hold (C1, C2) only and index j by its place
    // in the list's slot table (slot_tab: the slots the list's entries name, in ascending order of
    // 8 t + m for slot 8w + m of tile nbr[I][t]; uint16): 2 + 8 bytes per entry.  The T v
    // pass forms d = u_i - y_j in FP32 from the slots' offsets: u_i from tile I's origin (posl),
    // y_j = u_j + D_IJ in FP64 rounded once, D_IJ the minimum image of the tile origins, which every
    // pair of the tile pair within the cutoff takes (ew_tile_image_ok).  FP32 entries are stored in
    // pairs per lane and their list widths are even.  The lists are jagged (row n holds only the
    // lanes whose list reaches it, packed; the per-lane lengths are ccnt): the T v pass streams no
    // padding.
    int cap32, cap64;
    unsigned short* idx16;
    float2* t32;
    int* idx64;
    double4* t64;
    int* w32;               // [nblk kEwWaves] list widths
    int* w64;
    int* need;              // [8] largest width (FP32, FP64), entries (FP32, FP64), stored slots
                            // (FP32, FP64), largest neighbour-tile count, most slots an FP32 list names
    int lds_tiles;          // neighbour tiles the FP64-mode T v pass stages in LDS, and the force pass's
                            // tile shifts (0: from HBM)
    // Mixed precision: the slots an FP32 list's entries name (of the wave's slots of the row's
    // neighbour tiles), in ascending order, as global slot indices; the entries index this table,
    // which the T v pass copies into LDS (about half the wave's slots of the neighbour tiles)
    int* slot_tab;          // [nblk kEwWaves tcap]
    int* slot_cnt;          // [nblk kEwWaves] table length
    int tcap;
    int lds_slots;          // table entries the T v pass stages in LDS (0: FP32 lists gather from HBM)
    const double4* tb_o;    // [nblk] tile origins (FP64)
    const float4* posl;     // [npad] each slot's FP32 offset from its tile's origin
    float4* mu32;           // [npad] mixed: the force pass's FP32 records, dipole
    float4* slot32;         // [npad] and (charge, Thole a, 1/a)
    int e0_parts = 1;       // row parts whose sums the E0 epilogue adds (kEwE0Parts after the list build)
    unsigned* cand32;       // [nblk kEwWaves 64 cap32] the E0 pass's pair candidates with FP32 blocks,
    unsigned* cand64;       // [nblk kEwWaves 64 cap64] and with FP64 blocks, per lane in blocks of 8
    int2* ccnt;             // [nblk kEwWaves 64] the lanes' candidate counts (FP32, FP64 stream)
    bool mixed;             // FP32 grids and FFTs (FP64 spreading and sums)
};

// Neighbour tiles of every tile row (after launch_tile_bounds).
void launch_ew_neighbors(hipStream_t s, const EwLaunch& E);
// Reciprocal-space potential grid of the charges (vec == null) or of the dipoles v (charge == null)
// into phi (E.phi_q or E.phi_v): spreading, forward FFT, influence function, inverse FFT.
// Skipped when scalars says the solve has stopped.  small_fft: the FFT passes as one-wave
// workgroups of two lines (4 KB of LDS), which start beside a running real-space pass; else
// workgroups of eight lines, which wait for its workgroups to retire.
void launch_ew_recip(hipStream_t s, const EwLaunch& E, const double4* posa, const double* charge,
                     const double4* vec, void* phi, const double* scalars, bool small_fft = false);
// Real-space pass + reciprocal gather + self and scaled-pair terms for the 64 slots of every tile.
// kEwE0: out4 = E0, epart = 1/2 q phi_q - a/sqrt(pi) q^2.  kEwTmu: out4 = T v, or with pq the PCG
// q and p.q tile sums.  kEwForce (v = mu): frc (3 doubles per slot, overwritten) and eout (energy
// per slot, with epart and -1/2 mu.E0).
void launch_ew_pairs(hipStream_t s, const EwLaunch& E, int mode, bool mutual, const double4* posa,
                     const double* inva, const double* charge, const double4* vec, const double4* E0,
                     const double* scalars, double4* out4, const PcgQ& pq, double* epart, double* frc,
                     double* eout);
// The E0 pass that also builds the stored blocks (when the lists are stored): the real-space field
// and the lists, then the epilogue of kEwE0 (reciprocal gather, scaled pairs) into out4 and epart.
void launch_ew_e0_stored(hipStream_t s, const EwLaunch& E, const double4* posa, const double* inva,
                         const double* charge, double4* out4, double* epart, hipEvent_t join = nullptr);
// The force pass from the E0 pass's candidates (lists stored), then the epilogue of kEwForce (which
// reads the E0 epilogue's per-slot energy terms from epart).
void launch_ew_force_stored(hipStream_t s, const EwLaunch& E, bool mutual, const double4* posa, const double* inva,
                            const double* charge, const double4* vec, const double4* E0, double* epart, double* frc,
                            double* eout, hipEvent_t join = nullptr);
// T v of the solver from the stored blocks (v in FP64, and its FP32 copy v32 for the FP32 lists),
// into the list partials E.rsw ...
void launch_ew_spmv(hipStream_t s, const EwLaunch& E, const double4* vec, const float4* v32, const double* scalars);
// ... then the epilogue of kEwTmu over them (reciprocal gather, scaled pairs, self field, PCG q).
void launch_ew_tmu_epilogue(hipStream_t s, const EwLaunch& E, const double4* posa, const double* inva,
                            const double* charge, const double4* vec, const double* scalars, double4* out4,
                            const PcgQ& pq);
// Energy of the system: sum of eout over slots in a fixed order, plus the background term.
void launch_ew_energy(hipStream_t s, int npad, const double* eout, double background, double* energy);
// Every pair (i < j, caller indices) within the cutoff, as the real-space passes decide it:
// pairs[2k], pairs[2k+1]; count[0] (may exceed cap: pairs past cap are not written).
void launch_ew_dump_pairs(hipStream_t s, const EwLaunch& E, const double4* posa, const int* perm, int* pairs,
                          long long cap, unsigned long long* count);

// ---- dense path (tdp_dense.hip)
int dense_max_tiles();
void launch_dense_assemble(hipStream_t s, bool bf16, int nsys, int max_nt, const DenseSys* sys, const int* tcnt,
                           const double4* posa, const double* alpha, const double* inva, const int* exc_start,
                           const int* exc_J, const unsigned long long* exc_mask, unsigned short* mat);
// dacc: FP32 running sums of the diagonal blocks (same layout as dinv), zero on entry
void launch_dense_factor(hipStream_t s, bool bf16, int nsys, int max_nt, const DenseSys* sys, unsigned short* mat,
                         float* dinv, float* dacc, int* status);
void launch_dense_solve(hipStream_t s, bool bf16, int nsys, const DenseSys* sys, const unsigned short* mat,
                        const float* dinv, const double* rhs, double* y, const int* active);
void launch_dense_rhs(hipStream_t s, const SegLaunch& G, const DenseSys* sys, const int* tcnt, const double4* E0,
                      const double* alpha, double* b, double* x, double* res);
void launch_dense_update(hipStream_t s, const SegLaunch& G, const DenseSys* sys, const double* alpha, const double* y,
                         double* x, double4* mu4, size_t mu_stride, const int* active);
void launch_dense_residual(hipStream_t s, const SegLaunch& G, int q, const DenseSys* sys, const double* alpha,
                           const double4* Tmu, const double* b, const double* x, double* res, double* red,
                           double* rel);
void launch_uniform_field(hipStream_t s, int npad, const int* perm, int axis, double4* E0);

}  // namespace tdp
