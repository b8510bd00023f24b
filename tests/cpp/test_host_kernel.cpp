// C++ test of the host-side kernel mirror over the C ABI, in the style of the reference's
// platforms/reference/tests/TestReferenceTholeDipoleForce.cpp (standalone main, returns 0/1).
// The expected values come from the CPU oracle (oracle/tdp_oracle.c), linked as the checker.
#include <cmath>
#include <cstdio>
#include <iostream>
#include <vector>

#include "cases.hpp"

using namespace TholeDipolePlugin;
using namespace tdp_test;
using OpenMM::Vec3;

#define ASSERT(cond)                                                                       \
    do {                                                                                   \
        if (!(cond)) {                                                                     \
            std::cout << "FAIL " << __FILE__ << ":" << __LINE__ << " " #cond << std::endl; \
            return 1;                                                                      \
        }                                                                                  \
    } while (0)

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
