// C++ test of the host-side kernel mirror over the C ABI, in the style of the reference's
// platforms/reference/tests/TestReferenceTholeDipoleForce.cpp (standalone main, returns 0/1).
// The expected values come from the CPU oracle (oracle/tdp_oracle.c), linked as the checker.
#include <cmath>
#include <cstdio>
#include <iostream>
#include <vector>

#include "../../oracle/tdp_oracle.h"
#include "../../tholedipoleplugin_amd/host/TholeDipoleForce.h"

using namespace TholeDipolePlugin;
using OpenMM::Vec3;

#define ASSERT(cond)                                                                       \
    do {                                                                                   \
        if (!(cond)) {                                                                     \
            std::cout << "FAIL " << __FILE__ << ":" << __LINE__ << " " #cond << std::endl; \
            return 1;                                                                      \
        }                                                                                  \
    } while (0)

// Two ammonia molecules (TestReferenceTholeDipoleForce.cpp:62-199), NoCutoff.
static void ammonia(TholeDipoleForce& f, std::vector<Vec3>& pos) {
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

static void pack(const TholeDipoleForce& f, OracleInput& o) {
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

int main() {
    try {
        TholeDipoleForce force;
        std::vector<Vec3> pos;
        ammonia(force, pos);
        Vec3 box[3] = {Vec3(0.6, 0, 0), Vec3(0, 0.6, 0), Vec3(0, 0, 0.6)};
        HipCalcTholeDipoleForceKernel kernel;
        kernel.initialize(8, box, force);
        HipContextState ctx;
        ctx.positions = pos;
        for (int k = 0; k < 3; k++) ctx.box[k] = box[k];
        ctx.masses = {14.007, 1.008, 1.008, 1.008, 14.007, 1.008, 1.008, 1.008};
        std::vector<Vec3> forces(8);
        double energy = kernel.execute(ctx, forces, true, true);

        OracleInput o;
        pack(force, o);
        std::vector<double> rf(24, 0.0), ind(24), lab(24), tq(24), e0(24);
        tdpo_stats st;
        ASSERT(tdpo_force_energy(&o.s, &pos[0][0], rf.data(), &st, ind.data(), lab.data(), tq.data(), e0.data()) == 0);
        ASSERT(std::fabs(energy - st.energy) <= 1e-6 * std::fabs(st.energy));
        double num = 0, den = 0;
        for (int i = 0; i < 8; i++)
            for (int k = 0; k < 3; k++) {
                num += (forces[i][k] - rf[3 * i + k]) * (forces[i][k] - rf[3 * i + k]);
                den += rf[3 * i + k] * rf[3 * i + k];
            }
        ASSERT(std::sqrt(num / den) <= 1e-5);
        std::vector<Vec3> mu;
        kernel.getInducedDipoles(ctx, mu);
        for (int i = 0; i < 8; i++)
            for (int k = 0; k < 3; k++) ASSERT(std::fabs(mu[i][k] - ind[3 * i + k]) <= 2.08e-8);

        // getPMEParameters on a NoCutoff kernel throws (ReferenceTholeDipoleKernels.cpp:341-343)
        bool thrown = false;
        try {
            double a;
            int nx, ny, nz;
            kernel.getPMEParameters(a, nx, ny, nz);
        } catch (const OpenMM::OpenMMException& e) {
            thrown = std::string(e.what()).find("not using PME") != std::string::npos;
        }
        ASSERT(thrown);

        // updateParametersInContext changes the result; a fresh kernel with the new parameters agrees
        for (int i = 0; i < 8; i++) {
            double q, a, t;
            int at, z, x, y;
            std::vector<double> d;
            force.getParticleParameters(i, q, d, a, t, at, z, x, y);
            d[0] *= 0.7;
            force.setParticleParameters(i, 1.1 * q, d, 1.5 * a, 1.3 * t, at, z, x, y);
        }
        HipCalcTholeDipoleForceKernel kernel2;
        kernel2.initialize(8, box, force);
        std::vector<Vec3> f2(8), f1(8);
        double e2 = kernel2.execute(ctx, f2, true, true);
        double e1 = kernel.execute(ctx, f1, true, true);
        ASSERT(std::fabs(e1 - e2) > 1e-4 * std::fabs(e2));
        kernel.copyParametersToContext(ctx, force);
        std::fill(f1.begin(), f1.end(), Vec3());
        e1 = kernel.execute(ctx, f1, true, true);
        ASSERT(std::fabs(e1 - e2) <= 1e-10 * std::fabs(e2));

        // invalid axis particle (TholeDipoleForceImpl.cpp:92-95)
        TholeDipoleForce bad;
        bad.addParticle(0.0, {0, 0, 0}, 0.001, 0.39, TholeDipoleForce::ZThenX, 1, -1, -1);
        bad.addParticle(0.0, {0, 0, 0}, 0.001, 0.39);
        HipCalcTholeDipoleForceKernel kernel3;
        thrown = false;
        try {
            kernel3.initialize(2, box, bad);
        } catch (const OpenMM::OpenMMException& e) {
            thrown = std::string(e.what()).find("invalid x axis particle: -1") != std::string::npos;
        }
        ASSERT(thrown);
    } catch (const std::exception& e) {
        std::cout << "exception: " << e.what() << std::endl;
        std::cout << "FAIL - ERROR.  Test failed." << std::endl;
        return 1;
    }
    std::cout << "Done" << std::endl;
    return 0;
}
