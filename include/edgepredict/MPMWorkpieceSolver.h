#pragma once
// MI355X MPM workpiece solver behind the EdgePredict solver interfaces.
//
// Same public surface as MPMWorkpieceSolver in include/MPMWorkpieceSolver.cuh
// (:92-179) of Up25-pathan/edegpredict-engine-rotary, so it drops in where
// that class is constructed, initialised and stepped (main.cpp:307,362-363;
// the engine's sub-cycling call SimulationEngine.cpp:517-523).  The work is
// done by libmpmw_hip.so through its C ABI (mpmw.h); no HIP type appears here.
//
// Error behaviour follows the reference: initialize() reports failure with
// its return value, step() returns silently when there is nothing to step,
// and a device (HIP) failure prints a message and ends the process with
// EXIT_FAILURE, as CUDA_CHECK does (CudaUtils.cuh:35-59).

#include <string>
#include <vector>

#include "IPhysicsSolver.h"
#include "Types.h"

struct mpmw_solver;

namespace edgepredict {

class Config;

class MPMWorkpieceSolver : public IPhysicsSolver, public IMetricsProvider, public IThermalCoupling {
public:
    MPMWorkpieceSolver() = default;
    ~MPMWorkpieceSolver() override;

    MPMWorkpieceSolver(const MPMWorkpieceSolver&) = delete;
    MPMWorkpieceSolver& operator=(const MPMWorkpieceSolver&) = delete;

    // IPhysicsSolver
    std::string getName() const override { return "MPMWorkpieceSolver"; }
    bool initialize(const Config& config) override;
    void step(double dt) override;
    double getStableTimeStep() const override;
    double getCurrentTime() const override;
    bool isInitialized() const override { return m_solver != nullptr; }
    void reset() override;
    void getBounds(double& minX, double& minY, double& minZ, double& maxX, double& maxY,
                   double& maxZ) const override;
    int getParticleCount() const override;

    // IMetricsProvider
    double getMaxStress() const override;
    double getMaxTemperature() const override;
    double getTotalKineticEnergy() const override;
    void syncMetrics() override;

    // IThermalCoupling
    void applyHeatFlux(double x, double y, double z, double heatFlux) override;
    double getTemperatureAt(double x, double y, double z) const override;
    void setParticleTemperatures(const double* temperatures, int count) override;

    // MPMWorkpieceSolver
    void initializeParticleBox(const Vec3& minBounds, const Vec3& maxBounds, double spacing);
    void initializeCylindricalWorkpiece(const Vec3& center, double radius, double length, double spacing,
                                        int axis = 2);
    void setToolMesh(const Mesh& mesh, double contactRadius);
    void applyExternalForce(const Vec3& center, double radius, const Vec3& force);
    std::vector<MPMParticle> getParticlesForExport();
    double getSmoothingRadius() const { return m_cellSize; }
    double getTotalContactForce() const;
    void compactParticles();
    void setToolTipPosition(double x, double y, double z);

    // MI355X extras: arithmetic precision and the deterministic mode must be
    // chosen before initialize(); the device ordinal too.
    void setPrecisionFP64(bool fp64) { m_fp64 = fp64; }
    void setDeterministic(bool on);
    void setDevice(int device) { m_device = device; }
    mpmw_solver* handle() const { return m_solver; }

private:
    void check(int rc, const char* what) const;

    mpmw_solver* m_solver = nullptr;
    double m_cellSize = 0.0001;
    bool m_fp64 = false;
    bool m_deterministic = false;
    int m_device = 0;
};

}  // namespace edgepredict
