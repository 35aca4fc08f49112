#pragma once
// Solver plug-in interfaces of the EdgePredict engine, with the signatures of
// include/IPhysicsSolver.h:24-149 of Up25-pathan/edegpredict-engine-rotary
// (IPhysicsSolver :24-82, ISolverPtr :87, IMetricsProvider :92-120,
// IThermalCoupling :125-149), so an engine written against them can hold the
// MI355X solver through the same base classes.

#include <memory>
#include <string>

namespace edgepredict {

class Config;

class IPhysicsSolver {
public:
    virtual ~IPhysicsSolver() = default;
    virtual std::string getName() const = 0;
    virtual bool initialize(const Config& config) = 0;
    virtual void step(double dt) = 0;
    virtual double getStableTimeStep() const = 0;
    virtual double getCurrentTime() const = 0;
    virtual bool isInitialized() const = 0;
    virtual void getBounds(double& minX, double& minY, double& minZ,
                           double& maxX, double& maxY, double& maxZ) const = 0;
    virtual int getNodeCount() const { return 0; }
    virtual int getParticleCount() const { return 0; }
    virtual void reset() = 0;
};

using ISolverPtr = std::unique_ptr<IPhysicsSolver>;

class IMetricsProvider {
public:
    virtual ~IMetricsProvider() = default;
    virtual double getMaxStress() const = 0;
    virtual double getMaxTemperature() const = 0;
    virtual double getTotalKineticEnergy() const = 0;
    virtual double getTotalInternalEnergy() const { return 0.0; }
    virtual void syncMetrics() = 0;
};

class IThermalCoupling {
public:
    virtual ~IThermalCoupling() = default;
    virtual void applyHeatFlux(double x, double y, double z, double heatFlux) = 0;
    virtual double getTemperatureAt(double x, double y, double z) const = 0;
    virtual void setParticleTemperatures(const double* temperatures, int count) { (void)temperatures; (void)count; }
};

}  // namespace edgepredict
