// Test systems shared by the C++ drivers of tests/cpp: the reference's ammonia pair, a small
// periodic water box, and the packing of a TholeDipoleForce into the oracle's tdpo_system.
#ifndef TDP_TEST_CASES_HPP
#define TDP_TEST_CASES_HPP
#include <cmath>
#include <vector>

#include "../../oracle/tdp_oracle.h"
#include "../../tholedipoleplugin_amd/host/TholeDipoleForce.h"

namespace tdp_test {
using TholeDipolePlugin::TholeDipoleForce;
using OpenMM::Vec3;

// Two ammonia molecules (TestReferenceTholeDipoleForce.cpp:62-199), NoCutoff.
inline void ammonia(TholeDipoleForce& f, std::vector<Vec3>& pos) {
    f.setNonbondedMethod(TholeDipoleForce::NoCutoff);
    f.setPolarizationType(TholeDipoleForce::Mutual);
    f.setMutualInducedTargetEpsilon(1e-12);
    f.setMutualInducedMaxIterations(1000);
    std::vector<double> dN = {8.3832254e-3, 0.0, 3.4232474e-3}, dH = {-1.7388763e-3, 0.0, -4.6837475e-3};
    f.addParticle(-5.796e-1, dN, 1.073e-3, 3.9e-1, 2, 1, 2, 3);
    f.addParticle(1.932e-1, dH, 4.96e-4, 3.9e-1, 2, 0, 2, 3);
    f.addParticle(1.932e-1, dH, 4.96e-4, 3.9e-1, 2, 0, 1, 3);
    f.addParticle(1.932e-1, dH, 4.96e-4, 3.9e-1, 2, 0, 1, 2);
    f.addParticle(-5.796e-1, dN, 1.073e-3, 3.9e-1, 2, 5, 6, 7);
    f.addParticle(1.932e-1, dH, 4.96e-4, 3.9e-1, 2, 4, 6, 7);
    f.addParticle(1.932e-1, dH, 4.96e-4, 3.9e-1, 2, 4, 5, 7);
    f.addParticle(1.932e-1, dH, 4.96e-4, 3.9e-1, 2, 4, 5, 6);
    for (int m = 0; m < 2; m++) {
        int n0 = 4 * m;
        f.setCovalentMap(n0, TholeDipoleForce::Covalent12, {n0 + 1, n0 + 2, n0 + 3});
        for (int h = 1; h <= 3; h++) {
            f.setCovalentMap(n0 + h, TholeDipoleForce::Covalent12, {n0});
            std::vector<int> others;
            for (int k = 1; k <= 3; k++)
                if (k != h) others.push_back(n0 + k);
            f.setCovalentMap(n0 + h, TholeDipoleForce::Covalent13, others);
        }
    }
    pos = {Vec3(1.5927280e-01, 1.7000000e-06, 1.6491000e-03), Vec3(2.0805540e-01, -8.1258800e-02, 3.7282500e-02),
           Vec3(2.0843610e-01, 8.0953200e-02, 3.7462200e-02), Vec3(1.7280780e-01, 2.0730000e-04, -9.8741700e-02),
           Vec3(-1.6743680e-01, 1.5900000e-05, -6.6149000e-03), Vec3(-2.0428260e-01, 8.1071500e-02, 4.1343900e-02),
           Vec3(-6.7308300e-02, 1.2800000e-05, 1.0623300e-02), Vec3(-2.0426290e-01, -8.1231400e-02, 4.1033500e-02)};
}

struct OracleInput {
    std::vector<double> q, d, a, t;
    std::vector<int> ax, z, x, y, off, atoms;
    tdpo_system s;
};

inline void pack(const TholeDipoleForce& f, OracleInput& o) {
    int n = f.getNumParticles();
    o.off.assign(4 * n + 1, 0);
    for (int i = 0; i < n; i++) {
        double q, a, t;
        int at, z, x, y;
        std::vector<double> d;
        f.getParticleParameters(i, q, d, a, t, at, z, x, y);
        o.q.push_back(q); o.a.push_back(a); o.t.push_back(t);
        o.d.insert(o.d.end(), d.begin(), d.end());
        o.ax.push_back(at); o.z.push_back(z); o.x.push_back(x); o.y.push_back(y);
        std::vector<std::vector<int> > lists;
        f.getCovalentMaps(i, lists);
        for (int k = 0; k < 4; k++) {
            o.off[4 * i + k] = (int) o.atoms.size();
            o.atoms.insert(o.atoms.end(), lists[k].begin(), lists[k].end());
        }
    }
    o.off[4 * n] = (int) o.atoms.size();
    o.atoms.push_back(0);
    o.s = tdpo_system{n, o.q.data(), o.d.data(), o.a.data(), o.t.data(), o.ax.data(), o.z.data(), o.x.data(),
                      o.y.data(), o.off.data(), o.atoms.data(), f.getNonbondedMethod(), {0, 0, 0},
                      f.getPolarizationType(), f.getMutualInducedMaxIterations(), f.getMutualInducedTargetEpsilon(),
                      nullptr, 0};
}


// n_mol rigid 3-site waters (O-H 0.09572 nm, H-O-H 104.52 degrees, charges -0.834/+0.417,
// polarizabilities 1.44e-3/0.41e-3 nm^3, Thole 0.39) on a jittered cubic lattice in a box of side L,
// with the 1-2 and 1-3 maps of each molecule.
inline void water_box(int side, double L, TholeDipoleForce& f, std::vector<Vec3>& pos, unsigned seed = 7) {
    f.setNonbondedMethod(TholeDipoleForce::PME);
    f.setPolarizationType(TholeDipoleForce::Mutual);
    f.setMutualInducedTargetEpsilon(1e-10);
    f.setMutualInducedMaxIterations(500);
    unsigned s = seed;
    auto uni = [&s]() {   // linear congruential, [0, 1)
        s = 1664525u * s + 1013904223u;
        return (s >> 8) * (1.0 / 16777216.0);
    };
    const double h = L / side, b = 0.09572, th = 104.52 * M_PI / 180.0;
    std::vector<double> zero = {0.0, 0.0, 0.0};
    for (int a = 0; a < side; a++)
        for (int c = 0; c < side; c++)
            for (int d = 0; d < side; d++) {
                const int o = f.getNumParticles();
                Vec3 O((a + 0.5) * h + 0.02 * (uni() - 0.5), (c + 0.5) * h + 0.02 * (uni() - 0.5),
                       (d + 0.5) * h + 0.02 * (uni() - 0.5));
                const double phi = 2.0 * M_PI * uni();
                Vec3 e1(std::cos(phi), std::sin(phi), 0.0), e2(-std::sin(phi) * 0.6, std::cos(phi) * 0.6, 0.8);
                Vec3 H1(O[0] + b * e1[0], O[1] + b * e1[1], O[2] + b * e1[2]);
                Vec3 H2(O[0] + b * (std::cos(th) * e1[0] + std::sin(th) * e2[0]),
                        O[1] + b * (std::cos(th) * e1[1] + std::sin(th) * e2[1]),
                        O[2] + b * (std::cos(th) * e1[2] + std::sin(th) * e2[2]));
                f.addParticle(-0.834, zero, 1.44e-3, 0.39, TholeDipoleForce::NoAxisType, -1, -1, -1);
                f.addParticle(0.417, zero, 0.41e-3, 0.39, TholeDipoleForce::NoAxisType, -1, -1, -1);
                f.addParticle(0.417, zero, 0.41e-3, 0.39, TholeDipoleForce::NoAxisType, -1, -1, -1);
                f.setCovalentMap(o, TholeDipoleForce::Covalent12, {o + 1, o + 2});
                f.setCovalentMap(o + 1, TholeDipoleForce::Covalent12, {o});
                f.setCovalentMap(o + 1, TholeDipoleForce::Covalent13, {o + 2});
                f.setCovalentMap(o + 2, TholeDipoleForce::Covalent12, {o});
                f.setCovalentMap(o + 2, TholeDipoleForce::Covalent13, {o + 1});
                pos.push_back(O);
                pos.push_back(H1);
                pos.push_back(H2);
            }
}

}  // namespace tdp_test
#endif
