// tdp_tables.cpp -- scaled pairs and exclusion tiles (see tdp_tables.h).
#include "tdp_tables.h"

#include <algorithm>
#include <cstring>
#include <unordered_map>
#include <utility>

namespace tdp {

namespace {
const double kMScale[5] = {0.0, 0.0, 0.0, 0.5, 1.0};  // ReferenceTholeDipoleForce.cpp:37-41
const double kIScale[5] = {0.0, 1.0, 1.0, 1.0, 1.0};  // ReferenceTholeDipoleForce.cpp:43-47
}  // namespace

// Pair (i<j) takes its scales from particle i's own lists only: mScale from list 0 at the position
// of j, iScale from list 1, the last occurrence of a duplicated atom winning.  Pairs whose scales
// are both 1 are plain pairs; the rest are returned (in the caller's indices).  Lists are at most 5
// long (checked by the caller's validation).
std::vector<ExcPair> scaled_pairs(int n, const int32_t* cov_off, const int32_t* cov_atoms) {
    std::unordered_map<long long, std::pair<double, double>> pairs;
    auto key = [n](int i, int j) { return (long long)i * n + j; };
    for (int i = 0; i < n; i++) {
        for (int t = 0; t < 2; t++) {
            const int b = cov_off[4 * i + t], e = cov_off[4 * i + t + 1];
            for (int p = b; p < e; p++) {
                const int a = cov_atoms[p];
                if (a <= i || a >= n) continue;  // never looked up by the reference
                auto it = pairs.find(key(i, a));
                if (it == pairs.end()) it = pairs.emplace(key(i, a), std::make_pair(1.0, 1.0)).first;
                if (t == 0) it->second.first = kMScale[p - b];
                else it->second.second = kIScale[p - b];
            }
        }
    }
    std::vector<ExcPair> out;
    for (auto& kv : pairs) {
        if (kv.second.first == 1.0 && kv.second.second == 1.0) continue;
        out.push_back({(int)(kv.first / n), (int)(kv.first % n), kv.second.first, kv.second.second});
    }
    std::sort(out.begin(), out.end(),
              [](const ExcPair& a, const ExcPair& b) { return a.i != b.i ? a.i < b.i : a.j < b.j; });
    return out;
}

void exclusion_tiles(int n, int nblk, const int* inv, const std::vector<ExcPair>& pairs, ExclusionTiles& out) {
    std::vector<std::unordered_map<int, int>> tile_of(nblk);
    std::vector<unsigned long long> masks;  // [tile][3][64], in order of first use
    std::vector<std::pair<int, int>> tile_ij;
    auto tile = [&](int I, int J) {
        auto it = tile_of[I].find(J);
        if (it != tile_of[I].end()) return it->second;
        const int id = (int)tile_ij.size();
        tile_of[I][J] = id;
        tile_ij.push_back({I, J});
        masks.resize(masks.size() + 3 * 64, 0ull);
        return id;
    };
    auto setbits = [&](int i, int j, double m, double is) {
        const int t = tile(i / 64, j / 64);
        const int row = i % 64, col = j % 64;
        const unsigned long long bit = 1ull << col;
        if (m == 0.0) masks[((size_t)t * 3 + 0) * 64 + row] |= bit;
        if (m == 0.5) masks[((size_t)t * 3 + 1) * 64 + row] |= bit;
        if (is == 0.0) masks[((size_t)t * 3 + 2) * 64 + row] |= bit;
    };
    for (int i = 0; i < n; i++) setbits(inv[i], inv[i], 0.0, 0.0);
    for (const ExcPair& e : pairs) {
        const int a = inv[e.i], b = inv[e.j];
        setbits(a, b, e.m, e.is);
        setbits(b, a, e.m, e.is);
    }
    std::vector<int> order(tile_ij.size());
    for (size_t k = 0; k < order.size(); k++) order[k] = (int)k;
    std::sort(order.begin(), order.end(), [&](int a, int b) { return tile_ij[a] < tile_ij[b]; });
    out.start.assign(nblk + 1, 0);
    out.J.resize(order.size());
    out.mask.resize(masks.size());
    for (size_t k = 0; k < order.size(); k++) {
        const int t = order[k];
        out.start[tile_ij[t].first + 1]++;
        out.J[k] = tile_ij[t].second;
        std::memcpy(&out.mask[k * 3 * 64], &masks[(size_t)t * 3 * 64], 3 * 64 * sizeof(unsigned long long));
    }
    for (int I = 0; I < nblk; I++) out.start[I + 1] += out.start[I];
}

}  // namespace tdp

extern "C" int tdp_tables_exclusion_tiles(int n, const int32_t* cov_off, const int32_t* cov_atoms, int nblk,
                                          const int32_t* inv, int max_tiles, int32_t* start, int32_t* J,
                                          unsigned long long* mask) {
    if (n < 0 || nblk < 0 || (n > 0 && (!cov_off || !cov_atoms || !inv))) return 0;
    for (int i = 0; i < n; i++)
        if (inv[i] < 0 || inv[i] >= 64 * nblk) return 0;
    const std::vector<tdp::ExcPair> pairs = tdp::scaled_pairs(n, cov_off, cov_atoms);
    tdp::ExclusionTiles t;
    tdp::exclusion_tiles(n, nblk, inv, pairs, t);
    const int nt = (int)t.J.size();
    if (nt > max_tiles) return -nt;
    std::copy(t.start.begin(), t.start.end(), start);
    std::copy(t.J.begin(), t.J.end(), J);
    std::copy(t.mask.begin(), t.mask.end(), mask);
    return nt;
}
