// Drives edgepredict::MPMWorkpieceSolver through the reference's
// IPhysicsSolver / IMetricsProvider / IThermalCoupling interfaces the way the
// engine does (main.cpp:307,362-363; SimulationEngine.cpp:517-523) and checks
// invariants of the result.  Exit code 0 on success.
#include <cmath>
#include <cstdio>
#include <memory>

#include "edgepredict/Config.h"
#include "edgepredict/MPMWorkpieceSolver.h"

using namespace edgepredict;

static int failures = 0;
#define EXPECT(cond)                                                             \
    do {                                                                         \
        if (!(cond)) {                                                           \
            std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #cond);           \
            ++failures;                                                          \
        }                                                                        \
    } while (0)

int main() {
    Config cfg;
    cfg.sph().smoothingRadius = 2.5e-4;
    cfg.sph().maxParticles = 1000000;
    cfg.machining().ambientTemperature = 293.0;
    cfg.material().meltingPoint = 1878.0;
    cfg.machineSetup().fixtureLayerThickness = 5e-4;

    ISolverPtr base = std::make_unique<MPMWorkpieceSolver>();
    auto* mpm = dynamic_cast<MPMWorkpieceSolver*>(base.get());
    EXPECT(base->getName() == "MPMWorkpieceSolver");
    EXPECT(!base->isInitialized());
    base->step(1e-8);                                   // uninitialised: silent return
    EXPECT(base->initialize(cfg));
    EXPECT(base->isInitialized());
    mpm->initializeParticleBox(Vec3(0, 0, 0), Vec3(4e-3, 4e-3, 2e-3), 2e-4);
    const int n = base->getParticleCount();
    EXPECT(n == 21 * 21 * 11);

    Mesh tool;                                          // a flat patch above the box
    for (int i = 0; i <= 40; ++i)
        for (int j = 0; j <= 40; ++j) tool.nodes.push_back({Vec3(i * 1e-4, j * 1e-4, 2.05e-3)});
    mpm->setToolMesh(tool, 3.75e-4);

    const double dt = base->getStableTimeStep();
    EXPECT(std::fabs(dt - 0.35 * 2.5e-4 / std::sqrt(113.8e9 / 4430.0)) < 1e-20);
    auto parts = mpm->getParticlesForExport();
    double m0 = 0;
    for (auto& p : parts) m0 += p.mass;
    // push the top layer into the tool
    mpm->applyExternalForce(Vec3(2e-3, 2e-3, 2e-3), 3e-4, Vec3(0, 0, 1e-3));
    for (int s = 0; s < 50; ++s) base->step(dt);
    EXPECT(std::fabs(base->getCurrentTime() - 50 * dt) < 1e-18);

    auto* metrics = dynamic_cast<IMetricsProvider*>(mpm);
    metrics->syncMetrics();
    EXPECT(metrics->getMaxTemperature() >= 293.0);
    EXPECT(std::isfinite(metrics->getMaxStress()));
    EXPECT(metrics->getTotalKineticEnergy() >= 0.0);

    parts = mpm->getParticlesForExport();
    EXPECT((int)parts.size() == n);
    double m1 = 0;
    int fixed = 0;
    for (auto& p : parts) {
        m1 += p.mass;
        EXPECT(std::isfinite(p.x) && std::isfinite(p.vx) && std::isfinite(p.stress_xx));
        if (p.status == ParticleStatus::FIXED_BOUNDARY) {
            ++fixed;
            EXPECT(p.vx == 0 && p.vy == 0 && p.vz == 0);
        }
        EXPECT(std::fabs(p.pressure + (p.stress_xx + p.stress_yy + p.stress_zz) / 3.0) <= 1e-6 * (1 + std::fabs(p.pressure)));
    }
    EXPECT(fixed > 0);
    EXPECT(std::fabs(m1 - m0) <= 1e-12 * m0);

    double b[6];
    base->getBounds(b[0], b[1], b[2], b[3], b[4], b[5]);
    EXPECT(b[0] < b[3] && b[1] < b[4] && b[2] < b[5]);

    auto* thermal = dynamic_cast<IThermalCoupling*>(mpm);
    const double T0 = thermal->getTemperatureAt(2e-3, 2e-3, 1e-3);
    thermal->applyHeatFlux(2e-3, 2e-3, 1e-3, 1e-3);
    std::vector<double> temps(10, 500.0);
    thermal->setParticleTemperatures(temps.data(), (int)temps.size());
    EXPECT(T0 >= 293.0);
    mpm->compactParticles();
    EXPECT(base->getParticleCount() == n);

    base->reset();
    EXPECT(base->getParticleCount() == 0);
    std::printf("solver_api_test: %d particles, %d failures\n", n, failures);
    return failures == 0 ? 0 : 1;
}
