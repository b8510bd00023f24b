#ifndef THOLEDIPOLE_AMD_THOLEDIPOLEFORCE_H_
#define THOLEDIPOLE_AMD_THOLEDIPOLEFORCE_H_

// C++ host-side mirror of the reference's public force API (openmmapi/include/TholeDipoleForce.h:54-428)
// and of its kernel interface (openmmapi/include/TholeDipoleKernels.h:46-122), implemented over the
// C ABI of libtholedipole_hip.so (include/tholedipole_hip.h).
//
// OpenMM itself (Force, Context, Platform) is not part of this package: the few OpenMM value
// types the interface needs (Vec3, OpenMMException) are declared below with OpenMM's public
// signatures, and the role of ContextImpl (positions, box, masses) is played by HipContextState.

#include <cmath>
#include <exception>
#include <string>
#include <vector>

namespace OpenMM {

class Vec3 {
public:
    Vec3() : data{0.0, 0.0, 0.0} {}
    Vec3(double x, double y, double z) : data{x, y, z} {}
    const double& operator[](int i) const { return data[i]; }
    double& operator[](int i) { return data[i]; }
    Vec3 operator+(const Vec3& o) const { return Vec3(data[0] + o[0], data[1] + o[1], data[2] + o[2]); }
    Vec3 operator-(const Vec3& o) const { return Vec3(data[0] - o[0], data[1] - o[1], data[2] - o[2]); }
    Vec3 operator*(double s) const { return Vec3(data[0] * s, data[1] * s, data[2] * s); }
    double dot(const Vec3& o) const { return data[0] * o[0] + data[1] * o[1] + data[2] * o[2]; }

private:
    double data[3];
};

class OpenMMException : public std::exception {
public:
    explicit OpenMMException(const std::string& message) : message(message) {}
    const char* what() const noexcept override { return message.c_str(); }

private:
    std::string message;
};

}  // namespace OpenMM

namespace TholeDipolePlugin {

class TholeDipoleForce {
public:
    enum NonbondedMethod { NoCutoff = 0, PME = 1 };
    enum PolarizationType { Direct = 0, Mutual = 1, Extrapolated = 2 };
    enum MultipoleAxisTypes { ZThenX = 0, Bisector = 1, ZBisect = 2, ThreeFold = 3, ZOnly = 4, NoAxisType = 5,
                              LastAxisTypeIndex = 6 };
    enum CovalentType { Covalent12 = 0, Covalent13 = 1, Covalent14 = 2, Covalent15 = 3, CovalentEnd = 4 };

    TholeDipoleForce();

    int getNumParticles() const { return (int) particles.size(); }
    NonbondedMethod getNonbondedMethod() const { return nonbondedMethod; }
    void setNonbondedMethod(NonbondedMethod method);
    PolarizationType getPolarizationType() const { return polarizationType; }
    void setPolarizationType(PolarizationType type) { polarizationType = type; }
    double getCutoffDistance() const { return cutoffDistance; }
    void setCutoffDistance(double distance) { cutoffDistance = distance; }
    void getPMEParameters(double& alpha, int& nx, int& ny, int& nz) const;
    void setPMEParameters(double alpha, int nx, int ny, int nz);
    int getPmeBSplineOrder() const { return pmeBSplineOrder; }
    int addParticle(double charge, const std::vector<double>& molecularDipole, double polarizability, double tholeDamping,
                    int axisType = NoAxisType, int multipoleAtomZ = -1, int multipoleAtomX = -1, int multipoleAtomY = -1);
    void getParticleParameters(int index, double& charge, std::vector<double>& molecularDipole, double& polarizability,
                               double& tholeDamping, int& axisType, int& multipoleAtomZ, int& multipoleAtomX,
                               int& multipoleAtomY) const;
    void setParticleParameters(int index, double charge, const std::vector<double>& molecularDipole,
                               double polarizability, double tholeDamping, int axisType, int multipoleAtomZ,
                               int multipoleAtomX, int multipoleAtomY);
    void setCovalentMap(int index, CovalentType typeId, const std::vector<int>& covalentAtoms);
    void getCovalentMap(int index, CovalentType typeId, std::vector<int>& covalentAtoms) const;
    void getCovalentMaps(int index, std::vector<std::vector<int> >& covalentLists) const;
    int getMutualInducedMaxIterations() const { return mutualInducedMaxIterations; }
    void setMutualInducedMaxIterations(int n) { mutualInducedMaxIterations = n; }
    double getMutualInducedTargetEpsilon() const { return mutualInducedTargetEpsilon; }
    void setMutualInducedTargetEpsilon(double eps) { mutualInducedTargetEpsilon = eps; }
    void setExtrapolationCoefficients(const std::vector<double>& coefficients) { extrapolationCoefficients = coefficients; }
    const std::vector<double>& getExtrapolationCoefficients() const { return extrapolationCoefficients; }
    double getEwaldErrorTolerance() const { return ewaldErrorTol; }
    void setEwaldErrorTolerance(double tol) { ewaldErrorTol = tol; }
    int getForceGroup() const { return forceGroup; }
    void setForceGroup(int group);
    bool usesPeriodicBoundaryConditions() const { return nonbondedMethod == PME; }

private:
    struct ParticleInfo {
        int axisType, multipoleAtomZ, multipoleAtomX, multipoleAtomY;
        double charge, polarizability, tholeDamping;
        std::vector<double> molecularDipole;
        std::vector<std::vector<int> > covalentInfo;
    };
    NonbondedMethod nonbondedMethod;
    PolarizationType polarizationType;
    double cutoffDistance;
    double alpha;
    int pmeBSplineOrder, nx, ny, nz;
    int mutualInducedMaxIterations;
    std::vector<double> extrapolationCoefficients;
    double mutualInducedTargetEpsilon;
    double ewaldErrorTol;
    int forceGroup;
    std::vector<ParticleInfo> particles;
};

// What the kernel reads from the OpenMM ContextImpl: positions, box vectors and masses.
struct HipContextState {
    std::vector<OpenMM::Vec3> positions;
    OpenMM::Vec3 box[3];
    std::vector<double> masses;
};

// CalcTholeDipoleForceKernel (TholeDipoleKernels.h:46-122) over the C ABI.  Nonzero ABI status
// codes are raised as OpenMMException with the library's message.
class HipCalcTholeDipoleForceKernel {
public:
    static std::string Name() { return "CalcTholeDipoleForce"; }
    explicit HipCalcTholeDipoleForceKernel(int precision = 0, int solver = 0, int device = 0);
    ~HipCalcTholeDipoleForceKernel();
    HipCalcTholeDipoleForceKernel(const HipCalcTholeDipoleForceKernel&) = delete;
    HipCalcTholeDipoleForceKernel& operator=(const HipCalcTholeDipoleForceKernel&) = delete;

    // numParticles and the default box stand in for the System argument of the reference.
    void initialize(int numSystemParticles, const OpenMM::Vec3 defaultBox[3], const TholeDipoleForce& force);
    double execute(HipContextState& context, std::vector<OpenMM::Vec3>& forces, bool includeForces, bool includeEnergy);
    void copyParametersToContext(HipContextState& context, const TholeDipoleForce& force);
    void getPMEParameters(double& alpha, int& nx, int& ny, int& nz) const;
    void getInducedDipoles(HipContextState& context, std::vector<OpenMM::Vec3>& dipoles);
    void getLabFramePermanentDipoles(HipContextState& context, std::vector<OpenMM::Vec3>& dipoles);
    void getTotalDipoles(HipContextState& context, std::vector<OpenMM::Vec3>& dipoles);
    void getElectrostaticPotential(HipContextState& context, const std::vector<OpenMM::Vec3>& inputGrid,
                                   std::vector<double>& outputElectrostaticPotential);
    void getSystemMultipoleMoments(HipContextState& context, std::vector<double>& outputMultipoleMoments);

private:
    void dipoles(HipContextState& context, std::vector<OpenMM::Vec3>& out, int kind);
    void check(int rc) const;
    void* ctx;  // tdp_context*
    int precision, solver, device;
    OpenMM::Vec3 defaultBox[3];
};

// The checks of TholeDipoleForceImpl::initialize (openmmapi/src/TholeDipoleForceImpl.cpp:52-108).
void validateTholeDipoleForce(const TholeDipoleForce& force, int numSystemParticles, const OpenMM::Vec3 defaultBox[3]);

}  // namespace TholeDipolePlugin

#endif
