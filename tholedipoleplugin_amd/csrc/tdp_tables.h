// tdp_tables.h -- host-side tables of the pair kernels: scaled pairs from the covalent maps and
// their 64x64 tile bit masks.  Plain C++ (no HIP): compiled into libtholedipole_hip.so, and on its
// own into libtdp_tables.so so that the masks can be checked pair by pair on a machine without GPU.
#pragma once
#include <stdint.h>

#ifdef __cplusplus
#include <vector>

namespace tdp {

// A pair whose scales differ from 1, in the caller's indices (i < j).
struct ExcPair {
    int i, j;
    double m, is;
};

// setupScaleMaps/getScaleFactor (ReferenceTholeDipoleForce.cpp:609-644): pairs with a scale below 1,
// sorted by (i, j).  cov_off [4n+1], cov_atoms: the covalent lists of tdp_force_desc.
std::vector<ExcPair> scaled_pairs(int n, const int32_t* cov_off, const int32_t* cov_atoms);

// Exclusion tiles in the internal order (inv: caller index -> slot, nblk tiles of 64 slots): tiles
// (I, J) holding a scaled pair, both orientations, sorted by (I, J); start [nblk+1] indexes J;
// per tile three 64-bit masks per row: mScale == 0, mScale == 0.5, iScale == 0.  Self pairs set the
// mScale == 0 and iScale == 0 masks.
struct ExclusionTiles {
    std::vector<int> start, J;
    std::vector<unsigned long long> mask;   // [tile][3][64]
};
void exclusion_tiles(int n, int nblk, const int* inv, const std::vector<ExcPair>& pairs, ExclusionTiles& out);

}  // namespace tdp

extern "C" {
#endif

// C entry point of libtdp_tables.so (not part of the product ABI): builds the tiles of n particles
// placed at slots inv[0..n) of nblk tiles and returns them through the caller's buffers.  Returns the
// number of tiles, or -(needed) when max_tiles is too small; start [nblk+1], J [max_tiles],
// mask [max_tiles * 192].
int tdp_tables_exclusion_tiles(int n, const int32_t* cov_off, const int32_t* cov_atoms, int nblk, const int32_t* inv,
                               int max_tiles, int32_t* start, int32_t* J, unsigned long long* mask);

#ifdef __cplusplus
}
#endif
