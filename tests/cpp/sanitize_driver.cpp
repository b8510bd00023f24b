// CPU driver for the address and undefined-behaviour sanitizers: runs the oracle (tdp_oracle.c,
// tdp_ewald.c), the host tables (tdp_tables.cpp) and the C++ host mirror (TholeDipoleForce.cpp) on
// small systems, all compiled with -fsanitize=address,undefined into this executable.  The kernel
// mirror's initialize is called as well; without a GPU it stops at the library's device check.
// Prints "sanitize driver: all passed" and returns 0; any sanitizer report aborts the process.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "../../tholedipoleplugin_amd/csrc/tdp_tables.h"
#include "cases.hpp"

using namespace TholeDipolePlugin;
using namespace tdp_test;
using OpenMM::Vec3;

#define CHECK(cond)                                                                        \
    do {                                                                                   \
        if (!(cond)) {                                                                     \
            std::cout << "FAIL " << __FILE__ << ":" << __LINE__ << " " #cond << std::endl; \
            return 1;                                                                      \
        }                                                                                  \
    } while (0)

static bool throws_with(const std::string& what, void (*fn)(const TholeDipoleForce&), const TholeDipoleForce& f) {
    try {
        fn(f);
    } catch (const OpenMM::OpenMMException& e) {
        return std::string(e.what()).find(what) != std::string::npos;
    }
    return false;
}

static void validate_8(const TholeDipoleForce& f) {
    Vec3 box[3] = {Vec3(2, 0, 0), Vec3(0, 2, 0), Vec3(0, 0, 2)};
    validateTholeDipoleForce(f, f.getNumParticles(), box);
}

static int oracle_ammonia() {
    TholeDipoleForce f;
    std::vector<Vec3> pos;
    ammonia(f, pos);
    OracleInput o;
    pack(f, o);
    const int n = 8;
    std::vector<double> frc(3 * n, 0.0), ind(3 * n), lab(3 * n), tq(3 * n), e0(3 * n), mu(3 * n);
    tdpo_stats st;
    tdpo_set_threads(1);
    CHECK(tdpo_force_energy(&o.s, &pos[0][0], frc.data(), &st, ind.data(), lab.data(), tq.data(), e0.data()) == 0);
    CHECK(st.converged == 1 && std::isfinite(st.energy));
    const double energy = st.energy;
    CHECK(tdpo_induced_dipoles(&o.s, &pos[0][0], mu.data(), lab.data(), e0.data(), &st) == 0);
    for (int k = 0; k < 3 * n; k++) CHECK(std::fabs(mu[k] - ind[k]) < 1e-12);
    std::vector<double> ee0(3 * n), eind(3 * n), f2(3 * n, 0.0);
    double en = 0.0;
    CHECK(tdpo_evaluate_at(&o.s, &pos[0][0], ind.data(), ee0.data(), eind.data(), f2.data(), &en) == 0);
    CHECK(std::fabs(en - energy) <= 1e-9 * std::fabs(energy));
    const double grid[6] = {0.5, 0.1, 0.0, -0.3, 0.2, 0.4};
    double pot[2];
    CHECK(tdpo_potential(&o.s, &pos[0][0], grid, 2, pot) == 0);
    const double masses[8] = {14.007, 1.008, 1.008, 1.008, 14.007, 1.008, 1.008, 1.008};
    double mom[13];
    CHECK(tdpo_multipole_moments(&o.s, &pos[0][0], masses, mom) == 0);
    std::vector<double> mapped(3 * n, 0.0);
    CHECK(tdpo_map_torques(&o.s, &pos[0][0], tq.data(), mapped.data()) == 0);
    const int rows[3] = {0, 3, 7};
    std::vector<double> re0(9), rind(9), rf(9), rt(9), ren(3);
    CHECK(tdpo_rows(&o.s, &pos[0][0], ind.data(), rows, 3, re0.data(), rind.data(), rf.data(), rt.data(), ren.data()) ==
          0);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            double ms, is;
            CHECK(tdpo_pair_scales(&o.s, i, j, &ms, &is) == 0);
        }
    return 0;
}

static int oracle_water_box() {
    TholeDipoleForce f;
    std::vector<Vec3> pos;
    water_box(3, 0.93, f, pos);
    f.setCutoffDistance(0.45);
    OracleInput o;
    pack(f, o);
    o.s.method = TDPO_PME;
    for (int k = 0; k < 3; k++) o.s.box[k] = 0.93;
    const int n = f.getNumParticles();
    std::vector<double> frc(3 * n, 0.0), ind(3 * n), lab(3 * n), tq(3 * n), e0(3 * n);
    tdpo_stats st;
    tdpo_set_threads(2);
    CHECK(tdpo_force_energy(&o.s, &pos[0][0], frc.data(), &st, ind.data(), lab.data(), tq.data(), e0.data()) == 0);
    CHECK(std::isfinite(st.energy));
    std::vector<double> out(3 * n, 0.0);
    for (int which = 0; which < 3; which++) (void)tdpo_pass_slice(&o.s, &pos[0][0], ind.data(), which, 0, n / 2, out.data());
    // Ewald: smooth PME and the exact reciprocal sum
    for (int exact = 0; exact < 2; exact++) {
        tdpo_ewald e;
        std::memset(&e, 0, sizeof(e));
        e.cutoff = 0.45;
        e.alpha = 7.0;
        e.order = 5;
        if (exact) {
            e.kmax[0] = e.kmax[1] = e.kmax[2] = 6;
        } else {
            e.grid[0] = 12;
            e.grid[1] = 14;
            e.grid[2] = 15;
        }
        std::vector<double> ef(3 * n, 0.0), emu(3 * n), ee0(3 * n), tmu(3 * n);
        double en = 0.0;
        CHECK(tdpo_ewald_evaluate(&o.s, &e, &pos[0][0], ef.data(), &en, emu.data(), ee0.data(), &st) == 0);
        CHECK(st.converged == 1 && std::isfinite(en));
        CHECK(tdpo_ewald_fields(&o.s, &e, &pos[0][0], emu.data(), ee0.data(), tmu.data()) == 0);
        std::vector<double> f2(3 * n, 0.0);
        double en2 = 0.0;
        CHECK(tdpo_ewald_energy_at(&o.s, &e, &pos[0][0], emu.data(), f2.data(), &en2, nullptr) == 0);
        CHECK(std::fabs(en2 - en) <= 1e-9 * std::fabs(en));
    }
    const long long np = tdpo_cutoff_pairs(&o.s, &pos[0][0], 0.45, 0, nullptr);
    std::vector<int> pairs(2 * (size_t)np + 2);
    CHECK(tdpo_cutoff_pairs(&o.s, &pos[0][0], 0.45, np, pairs.data()) == np);
    std::vector<double> a(2 * 6 * 7 * 10);
    for (size_t k = 0; k < a.size(); k++) a[k] = std::sin(0.37 * k);
    std::vector<double> b = a;
    tdpo_fft3(b.data(), 6, 7, 10, -1);
    tdpo_fft3(b.data(), 6, 7, 10, +1);
    for (size_t k = 0; k < a.size(); k++) CHECK(std::fabs(b[k] / 420.0 - a[k]) < 1e-12);
    double th[6], d1[6], d2[6];
    tdpo_bspline(0.3, 5, th, d1, d2);
    tdpo_bspline(0.0, 6, th, d1, d2);
    tdpo_set_threads(1);
    return 0;
}

static int tables(const TholeDipoleForce& f) {
    OracleInput o;
    pack(f, o);
    const int n = f.getNumParticles();
    const int nblk = (n + 63) / 64 + 1;
    std::vector<int32_t> inv(n);
    for (int i = 0; i < n; i++) inv[i] = (i * 37) % (64 * nblk);   // a scattered placement into the slots
    std::vector<int32_t> start(nblk + 1), J;
    std::vector<unsigned long long> mask;
    int need = tdp_tables_exclusion_tiles(n, o.off.data(), o.atoms.data(), nblk, inv.data(), 0, start.data(), nullptr,
                                          nullptr);
    CHECK(need <= 0);
    J.resize(-need + 1);
    mask.resize(((size_t)-need + 1) * 192);
    const int got = tdp_tables_exclusion_tiles(n, o.off.data(), o.atoms.data(), nblk, inv.data(), -need, start.data(),
                                               J.data(), mask.data());
    CHECK(got == -need);
    return 0;
}

int main() {
    try {
        // host mirror: the API, validation messages and the packing of initialize
        TholeDipoleForce f;
        std::vector<Vec3> pos;
        ammonia(f, pos);
        CHECK(f.getNumParticles() == 8);
        validate_8(f);
        TholeDipoleForce bad = f;
        bad.setParticleParameters(0, -0.5796, {8.38e-3, 0.0, 3.42e-3}, 1.073e-3, 0.39, TholeDipoleForce::ZThenX, 1, 9,
                                  -1);
        CHECK(throws_with("invalid x axis particle", validate_8, bad));
        TholeDipoleForce big;
        big.addParticle(0.0, {0.0, 0.0, 0.0}, 1e-3, 0.39, TholeDipoleForce::NoAxisType, -1, -1, -1);
        big.setCovalentMap(0, TholeDipoleForce::Covalent12, {0, 0, 0, 0, 0, 0});
        (void)throws_with("", validate_8, big);
        Vec3 box[3] = {Vec3(0.6, 0, 0), Vec3(0, 0.6, 0), Vec3(0, 0, 0.6)};
        try {
            HipCalcTholeDipoleForceKernel kernel;
            kernel.initialize(8, box, f);
            HipContextState ctx;
            ctx.positions = pos;
            for (int k = 0; k < 3; k++) ctx.box[k] = box[k];
            std::vector<Vec3> forces(8);
            (void)kernel.execute(ctx, forces, true, true);
        } catch (const OpenMM::OpenMMException& e) {
            std::cout << "kernel mirror stopped at: " << e.what() << std::endl;   // no device on this host
        }
        if (oracle_ammonia()) return 1;
        if (oracle_water_box()) return 1;
        if (tables(f)) return 1;
        TholeDipoleForce w;
        std::vector<Vec3> wpos;
        water_box(4, 1.24, w, wpos);
        if (tables(w)) return 1;
    } catch (const std::exception& e) {
        std::cout << "FAIL exception: " << e.what() << std::endl;
        return 1;
    }
    std::cout << "sanitize driver: all passed" << std::endl;
    return 0;
}
