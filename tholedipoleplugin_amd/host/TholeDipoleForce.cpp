// C++ host-side TholeDipoleForce and HipCalcTholeDipoleForceKernel over libtholedipole_hip.so.
// Parameter defaults and messages follow openmmapi/src/TholeDipoleForce.cpp:43-235,
// openmmapi/src/TholeDipoleForceImpl.cpp:52-108 and platforms/reference/src/ReferenceTholeDipoleKernels.cpp.
#include "TholeDipoleForce.h"

#include <sstream>

#include "../../include/tholedipole_hip.h"

using namespace TholeDipolePlugin;
using OpenMM::OpenMMException;
using OpenMM::Vec3;

TholeDipoleForce::TholeDipoleForce()
    : nonbondedMethod(NoCutoff), polarizationType(Mutual), cutoffDistance(1.0), alpha(0.0), pmeBSplineOrder(5), nx(0),
      ny(0), nz(0), mutualInducedMaxIterations(60), extrapolationCoefficients{-0.154, 0.017, 0.658, 0.474},
      mutualInducedTargetEpsilon(1e-5), ewaldErrorTol(1e-4), forceGroup(0) {}

void TholeDipoleForce::setNonbondedMethod(NonbondedMethod method) {
    if (method < 0 || method > 1) throw OpenMMException("TholeDipoleForce: Illegal value for nonbonded method");
    nonbondedMethod = method;
}

void TholeDipoleForce::setForceGroup(int group) {
    if (group < 0 || group > 31) throw OpenMMException("Force group must be between 0 and 31");
    forceGroup = group;
}

void TholeDipoleForce::getPMEParameters(double& a, int& x, int& y, int& z) const {
    a = alpha;
    x = nx;
    y = ny;
    z = nz;
}

void TholeDipoleForce::setPMEParameters(double a, int x, int y, int z) {
    alpha = a;
    nx = x;
    ny = y;
    nz = z;
}

int TholeDipoleForce::addParticle(double charge, const std::vector<double>& d, double polarizability, double thole,
                                  int axisType, int z, int x, int y) {
    ParticleInfo p;
    p.charge = charge;
    p.molecularDipole = {d[0], d[1], d[2]};
    p.polarizability = polarizability;
    p.tholeDamping = thole;
    p.axisType = axisType;
    p.multipoleAtomZ = z;
    p.multipoleAtomX = x;
    p.multipoleAtomY = y;
    p.covalentInfo.resize(CovalentEnd);
    particles.push_back(p);
    return (int) particles.size() - 1;
}

void TholeDipoleForce::getParticleParameters(int index, double& charge, std::vector<double>& d, double& polarizability,
                                             double& thole, int& axisType, int& z, int& x, int& y) const {
    const ParticleInfo& p = particles[index];
    charge = p.charge;
    polarizability = p.polarizability;
    thole = p.tholeDamping;
    d = p.molecularDipole;
    axisType = p.axisType;
    z = p.multipoleAtomZ;
    x = p.multipoleAtomX;
    y = p.multipoleAtomY;
}

void TholeDipoleForce::setParticleParameters(int index, double charge, const std::vector<double>& d,
                                             double polarizability, double thole, int axisType, int z, int x, int y) {
    ParticleInfo& p = particles[index];
    p.charge = charge;
    p.polarizability = polarizability;
    p.tholeDamping = thole;
    p.molecularDipole = {d[0], d[1], d[2]};
    p.axisType = axisType;
    p.multipoleAtomZ = z;
    p.multipoleAtomX = x;
    p.multipoleAtomY = y;
}

void TholeDipoleForce::setCovalentMap(int index, CovalentType typeId, const std::vector<int>& atoms) {
    particles[index].covalentInfo[typeId] = atoms;
}

void TholeDipoleForce::getCovalentMap(int index, CovalentType typeId, std::vector<int>& atoms) const {
    atoms = particles[index].covalentInfo[typeId];
}

void TholeDipoleForce::getCovalentMaps(int index, std::vector<std::vector<int> >& lists) const {
    lists = particles[index].covalentInfo;
}

void TholeDipolePlugin::validateTholeDipoleForce(const TholeDipoleForce& force, int numParticles, const Vec3 box[3]) {
    if (force.getNumParticles() != numParticles)
        throw OpenMMException("TholeDipoleForce must have exactly as many particles as the System it belongs to.");
    if (force.getNonbondedMethod() == TholeDipoleForce::PME) {
        double cutoff = force.getCutoffDistance();
        if (cutoff > 0.5 * box[0][0] || cutoff > 0.5 * box[1][1] || cutoff > 0.5 * box[2][2])
            throw OpenMMException(
                "TholeDipoleForce: The cutoff distance cannot be greater than half the periodic box size.");
    }
    for (int ii = 0; ii < numParticles; ii++) {
        int axisType, z, x, y;
        double charge, polarizability, thole;
        std::vector<double> d;
        force.getParticleParameters(ii, charge, d, polarizability, thole, axisType, z, x, y);
        if (axisType < TholeDipoleForce::ZThenX || axisType > TholeDipoleForce::NoAxisType) {
            std::stringstream buffer;
            buffer << "TholeDipoleForce: axis type=" << axisType << " not currently handled - only axisTypes[ 0, 1, 2, 3, 4, 5] "
                   << "(ZThenX, Bisector, Z-Bisect, ThreeFold, ZOnly, NoAxisType) currently handled.";
            throw OpenMMException(buffer.str());
        }
        if (axisType != TholeDipoleForce::NoAxisType && (z < 0 || z >= numParticles))
            throw OpenMMException("TholeDipoleForce: invalid z axis particle: " + std::to_string(z));
        if (axisType != TholeDipoleForce::NoAxisType && axisType != TholeDipoleForce::ZOnly && (x < 0 || x >= numParticles))
            throw OpenMMException("TholeDipoleForce: invalid x axis particle: " + std::to_string(x));
        if ((axisType == TholeDipoleForce::ZBisect || axisType == TholeDipoleForce::ThreeFold) &&
            (y < 0 || y >= numParticles))
            throw OpenMMException("TholeDipoleForce: invalid y axis particle: " + std::to_string(y));
    }
}

namespace {

// tdp_force_desc over arrays owned by this object (the fields CalcTholeDipoleForceKernel::initialize
// reads, ReferenceTholeDipoleKernels.cpp:83-158).
struct Packed {
    std::vector<double> q, d, alpha, thole, extrap;
    std::vector<int32_t> axis, az, ax, ay, off, atoms;
    tdp_force_desc desc;

    Packed(const TholeDipoleForce& force, const Vec3 box[3]) {
        int n = force.getNumParticles();
        q.resize(n); d.resize(3 * n); alpha.resize(n); thole.resize(n);
        axis.resize(n); az.resize(n); ax.resize(n); ay.resize(n); off.assign(4 * n + 1, 0);
        for (int i = 0; i < n; i++) {
            std::vector<double> dip;
            int at, z, x, y;
            force.getParticleParameters(i, q[i], dip, alpha[i], thole[i], at, z, x, y);
            for (int k = 0; k < 3; k++) d[3 * i + k] = dip[k];
            axis[i] = at; az[i] = z; ax[i] = x; ay[i] = y;
            std::vector<std::vector<int> > lists;
            force.getCovalentMaps(i, lists);
            for (int t = 0; t < 4; t++) {
                off[4 * i + t] = (int32_t) atoms.size();
                atoms.insert(atoms.end(), lists[t].begin(), lists[t].end());
            }
        }
        off[4 * n] = (int32_t) atoms.size();
        if (atoms.empty()) atoms.push_back(0);
        extrap = force.getExtrapolationCoefficients();
        desc.num_particles = n;
        desc.charges = q.data(); desc.dipoles = d.data(); desc.polarizabilities = alpha.data(); desc.thole = thole.data();
        desc.axis_types = axis.data(); desc.atom_z = az.data(); desc.atom_x = ax.data(); desc.atom_y = ay.data();
        desc.covalent_offsets = off.data(); desc.covalent_atoms = atoms.data();
        desc.nonbonded_method = force.getNonbondedMethod();
        desc.polarization_type = force.getPolarizationType();
        desc.cutoff = force.getCutoffDistance();
        desc.ewald_error_tolerance = force.getEwaldErrorTolerance();
        int gx, gy, gz;
        force.getPMEParameters(desc.pme_alpha, gx, gy, gz);
        desc.pme_grid[0] = gx; desc.pme_grid[1] = gy; desc.pme_grid[2] = gz;
        desc.mutual_max_iterations = force.getMutualInducedMaxIterations();
        desc.mutual_target_epsilon = force.getMutualInducedTargetEpsilon();
        desc.extrapolation_coefficients = extrap.empty() ? nullptr : extrap.data();
        desc.num_extrapolation_coefficients = (int32_t) extrap.size();
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) desc.default_box[3 * r + c] = box[r][c];
    }
};

void boxArray(const HipContextState& c, double out[9]) {
    for (int r = 0; r < 3; r++)
        for (int k = 0; k < 3; k++) out[3 * r + k] = c.box[r][k];
}

}  // namespace

HipCalcTholeDipoleForceKernel::HipCalcTholeDipoleForceKernel(int precision, int solver, int device)
    : ctx(nullptr), precision(precision), solver(solver), device(device) {}

HipCalcTholeDipoleForceKernel::~HipCalcTholeDipoleForceKernel() { tdp_destroy(static_cast<tdp_context*>(ctx)); }

void HipCalcTholeDipoleForceKernel::check(int rc) const {
    if (rc != TDP_OK) throw OpenMMException(tdp_last_error(static_cast<tdp_context*>(ctx)));
}

void HipCalcTholeDipoleForceKernel::initialize(int numSystemParticles, const Vec3 box[3], const TholeDipoleForce& force) {
    validateTholeDipoleForce(force, numSystemParticles, box);
    for (int k = 0; k < 3; k++) defaultBox[k] = box[k];
    Packed pk(force, box);
    tdp_options opts = {precision, solver, device};
    tdp_context* c = nullptr;
    if (tdp_create(&pk.desc, &opts, &c) != TDP_OK) throw OpenMMException(tdp_last_create_error());
    ctx = c;
}

double HipCalcTholeDipoleForceKernel::execute(HipContextState& context, std::vector<Vec3>& forces, bool includeForces,
                                              bool includeEnergy) {
    double box[9], energy = 0.0;
    boxArray(context, box);
    forces.resize(context.positions.size());
    check(tdp_execute(static_cast<tdp_context*>(ctx), &context.positions[0][0], box, includeForces, includeEnergy,
                      &forces[0][0], &energy));
    return energy;
}

void HipCalcTholeDipoleForceKernel::copyParametersToContext(HipContextState&, const TholeDipoleForce& force) {
    Packed pk(force, defaultBox);
    check(tdp_copy_parameters(static_cast<tdp_context*>(ctx), &pk.desc));
}

void HipCalcTholeDipoleForceKernel::getPMEParameters(double& alpha, int& nx, int& ny, int& nz) const {
    int32_t x, y, z;
    check(tdp_get_pme_parameters(static_cast<tdp_context*>(ctx), &alpha, &x, &y, &z));
    nx = x;
    ny = y;
    nz = z;
}

void HipCalcTholeDipoleForceKernel::dipoles(HipContextState& context, std::vector<Vec3>& out, int kind) {
    double box[9];
    boxArray(context, box);
    out.resize(context.positions.size());
    check(tdp_get_dipoles(static_cast<tdp_context*>(ctx), &context.positions[0][0], box, kind, &out[0][0]));
}

void HipCalcTholeDipoleForceKernel::getInducedDipoles(HipContextState& c, std::vector<Vec3>& d) {
    dipoles(c, d, TDP_DIPOLES_INDUCED);
}

void HipCalcTholeDipoleForceKernel::getLabFramePermanentDipoles(HipContextState& c, std::vector<Vec3>& d) {
    dipoles(c, d, TDP_DIPOLES_LAB_PERMANENT);
}

void HipCalcTholeDipoleForceKernel::getTotalDipoles(HipContextState& c, std::vector<Vec3>& d) {
    dipoles(c, d, TDP_DIPOLES_TOTAL);
}

void HipCalcTholeDipoleForceKernel::getElectrostaticPotential(HipContextState& context, const std::vector<Vec3>& grid,
                                                              std::vector<double>& out) {
    double box[9];
    boxArray(context, box);
    out.resize(grid.size());
    if (grid.empty()) return;
    check(tdp_get_electrostatic_potential(static_cast<tdp_context*>(ctx), &context.positions[0][0], box, &grid[0][0],
                                          (int32_t) grid.size(), out.data()));
}

void HipCalcTholeDipoleForceKernel::getSystemMultipoleMoments(HipContextState& context, std::vector<double>& out) {
    double box[9];
    boxArray(context, box);
    out.resize(13);
    check(tdp_get_system_multipole_moments(static_cast<tdp_context*>(ctx), &context.positions[0][0], box,
                                           context.masses.data(), out.data()));
}
