#pragma once
// The configuration values the MPM workpiece step reads, grouped and named as
// in include/Config.h of Up25-pathan/edegpredict-engine-rotary
// (MaterialProperties :23-45, SimulationParams :74-80, MachiningParams
// :101-111, MPMParams :128-157, PhysicsSafety :343-346, MachineSetupParams
// :366-391) with the same defaults.  Callers fill the structs directly;
// parsing a JSON file into them is left to the caller.

#include <string>

namespace edgepredict {

struct MaterialProperties {
    std::string name = "Ti-6Al-4V";
    double density = 4430.0;
    double youngsModulus = 113.8e9;
    double poissonsRatio = 0.34;
    double yieldStrength = 880e6;
    double specificHeat = 526.3;
    double thermalConductivity = 6.7;
    double meltingPoint = 1660.0;
    double jc_A = 880e6;
    double jc_B = 290e6;
    double jc_n = 0.47;
    double jc_C = 0.015;
    double jc_m = 1.0;
    double failureStrain = 0.3;
};

struct SimulationParams {
    double minTimeStep = 1e-12;
    double maxTimeStep = 1e-4;
};

struct MachiningParams {
    double ambientTemperature = 25.0;
};

struct MPMParams {
    double smoothingRadius = 0.0001;
    int maxParticles = 100000;
    double particleSpacingFactor = 0.8;
    double criticalPlasticStrain = 0.15;
    double thermalSofteningChipRatio = 0.25;
    double damageThreshold = 1.0;
    double referenceStrainRate = 1.0;
};

struct PhysicsSafety {
    double maxContactTempRisePerStep = 50.0;
    bool thermalClampEnabled = true;
};

struct MachineSetupParams {
    double fixtureLayerThickness = 0.002;
};

class Config {
public:
    const MaterialProperties& getMaterial() const { return m_material; }
    const SimulationParams& getSimulation() const { return m_simulation; }
    const MachiningParams& getMachining() const { return m_machining; }
    const MPMParams& getSPH() const { return m_sph; }
    const PhysicsSafety& getPhysicsSafety() const { return m_safety; }
    const MachineSetupParams& getMachineSetup() const { return m_machineSetup; }

    MaterialProperties& material() { return m_material; }
    SimulationParams& simulation() { return m_simulation; }
    MachiningParams& machining() { return m_machining; }
    MPMParams& sph() { return m_sph; }
    PhysicsSafety& physicsSafety() { return m_safety; }
    MachineSetupParams& machineSetup() { return m_machineSetup; }

private:
    MaterialProperties m_material;
    SimulationParams m_simulation;
    MachiningParams m_machining;
    MPMParams m_sph;
    PhysicsSafety m_safety;
    MachineSetupParams m_machineSetup;
};

}  // namespace edgepredict
