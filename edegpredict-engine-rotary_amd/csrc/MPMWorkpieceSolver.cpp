// MPMWorkpieceSolver.cpp -- the reference's MPMWorkpieceSolver class
// (src/MPMWorkpieceSolver.cu:418-946) over the MI355X C ABI.
#include "edgepredict/MPMWorkpieceSolver.h"

#include <cstdio>
#include <cstdlib>
#include <vector>

#include "edgepredict/Config.h"
#include "mpmw.h"

namespace edgepredict {

MPMWorkpieceSolver::~MPMWorkpieceSolver() {
    if (m_solver) mpmw_destroy(m_solver);
    m_solver = nullptr;
}

// HIP failures end the process as CUDA_CHECK does; argument and state errors
// are reported and the call returns.
void MPMWorkpieceSolver::check(int rc, const char* what) const {
    if (rc == MPMW_OK) return;
    std::fprintf(stderr, "[MPM] %s failed: %s\n", what, mpmw_last_error());
    if (rc == MPMW_E_HIP || rc == MPMW_E_NOMEM) std::exit(EXIT_FAILURE);
}

// initialize, MPMWorkpieceSolver.cu:422-458 (values are taken from the
// Config when called).
bool MPMWorkpieceSolver::initialize(const Config& config) {
    const auto& material = config.getMaterial();
    const auto& sph = config.getSPH();
    const auto& safety = config.getPhysicsSafety();
    mpmw_params p;
    mpmw_params_default(&p);
    p.density = material.density;
    p.young = material.youngsModulus;
    p.poisson = material.poissonsRatio;
    p.specific_heat = material.specificHeat;
    p.melt_temp = material.meltingPoint;
    p.ref_temp = config.getMachining().ambientTemperature;
    p.jc_A = material.jc_A;
    p.jc_B = material.jc_B;
    p.jc_n = material.jc_n;
    p.jc_C = material.jc_C;
    p.jc_m = material.jc_m;
    p.ref_strain_rate = sph.referenceStrainRate;
    p.failure_strain = material.failureStrain;
    p.damage_threshold = sph.damageThreshold;
    p.critical_plastic_strain = sph.criticalPlasticStrain;
    p.thermal_chip_ratio = sph.thermalSofteningChipRatio;
    p.max_contact_temp_rise = safety.maxContactTempRisePerStep;
    p.thermal_clamp_enabled = safety.thermalClampEnabled ? 1 : 0;
    p.cell_size = sph.smoothingRadius;
    p.particle_spacing_factor = sph.particleSpacingFactor;
    p.tool_contact_radius = sph.smoothingRadius * 1.5;
    p.has_config = 1;
    p.max_particles = sph.maxParticles;
    p.fixture_thickness = config.getMachineSetup().fixtureLayerThickness;
    p.min_time_step = config.getSimulation().minTimeStep;
    p.max_time_step = config.getSimulation().maxTimeStep;
    p.precision = m_fp64 ? MPMW_FP64 : MPMW_FP32;
    p.deterministic = m_deterministic ? 1 : 0;
    p.device = m_device;
    if (m_solver) mpmw_destroy(m_solver);
    m_solver = nullptr;
    const int rc = mpmw_create(&p, &m_solver);
    if (rc != MPMW_OK) {
        std::fprintf(stderr, "[MPM] initialize failed: %s\n", mpmw_last_error());
        m_solver = nullptr;
        return false;
    }
    m_cellSize = p.cell_size;
    std::printf("[MPM] MI355X solver initialized, dx=%g mm\n", m_cellSize * 1000.0);
    return true;
}

void MPMWorkpieceSolver::step(double dt) {
    if (!m_solver) return;  // :728
    check(mpmw_step(m_solver, dt), "step");
}

double MPMWorkpieceSolver::getStableTimeStep() const {
    double dt = 0.0;
    if (m_solver) check(mpmw_stable_dt(m_solver, &dt), "getStableTimeStep");
    return dt;
}

double MPMWorkpieceSolver::getCurrentTime() const {
    if (!m_solver) return 0.0;
    mpmw_metrics m;
    check(mpmw_get_metrics(m_solver, &m), "getCurrentTime");
    return m.time;
}

void MPMWorkpieceSolver::reset() {
    if (m_solver) check(mpmw_reset(m_solver), "reset");
}

void MPMWorkpieceSolver::getBounds(double& minX, double& minY, double& minZ, double& maxX, double& maxY,
                                   double& maxZ) const {
    double b[6] = {0, 0, 0, 0, 0, 0};
    if (m_solver) check(mpmw_get_bounds(m_solver, b), "getBounds");
    minX = b[0]; minY = b[1]; minZ = b[2];
    maxX = b[3]; maxY = b[4]; maxZ = b[5];
}

int MPMWorkpieceSolver::getParticleCount() const {
    int32_t n = 0;
    if (m_solver) check(mpmw_particle_count(m_solver, &n), "getParticleCount");
    return n;
}

double MPMWorkpieceSolver::getMaxStress() const {
    mpmw_metrics m{};
    if (m_solver) check(mpmw_get_metrics(m_solver, &m), "getMaxStress");
    return m.max_stress;
}

double MPMWorkpieceSolver::getMaxTemperature() const {
    mpmw_metrics m{};
    if (m_solver) check(mpmw_get_metrics(m_solver, &m), "getMaxTemperature");
    return m.max_temperature;
}

double MPMWorkpieceSolver::getTotalKineticEnergy() const {
    mpmw_metrics m{};
    if (m_solver) check(mpmw_get_metrics(m_solver, &m), "getTotalKineticEnergy");
    return m.kinetic_energy;
}

double MPMWorkpieceSolver::getTotalContactForce() const {
    mpmw_metrics m{};
    if (m_solver) check(mpmw_get_metrics(m_solver, &m), "getTotalContactForce");
    return m.contact_force;
}

void MPMWorkpieceSolver::syncMetrics() {
    if (m_solver) check(mpmw_sync_metrics(m_solver), "syncMetrics");
}

void MPMWorkpieceSolver::applyHeatFlux(double x, double y, double z, double heatFlux) {
    if (m_solver) check(mpmw_apply_heat_flux(m_solver, x, y, z, heatFlux), "applyHeatFlux");
}

double MPMWorkpieceSolver::getTemperatureAt(double x, double y, double z) const {
    double t = 0.0;
    if (m_solver) check(mpmw_temperature_at(m_solver, x, y, z, &t), "getTemperatureAt");
    return t;
}

void MPMWorkpieceSolver::setParticleTemperatures(const double* temperatures, int count) {
    if (m_solver) check(mpmw_set_particle_temperatures(m_solver, temperatures, count), "setParticleTemperatures");
}

void MPMWorkpieceSolver::initializeParticleBox(const Vec3& minBounds, const Vec3& maxBounds, double spacing) {
    if (!m_solver) return;
    const double mn[3] = {minBounds.x, minBounds.y, minBounds.z};
    const double mx[3] = {maxBounds.x, maxBounds.y, maxBounds.z};
    check(mpmw_init_particle_box(m_solver, mn, mx, spacing), "initializeParticleBox");
    std::printf("[MPM] particles initialized in box: %d\n", getParticleCount());
}

void MPMWorkpieceSolver::initializeCylindricalWorkpiece(const Vec3& center, double radius, double length,
                                                        double spacing, int axis) {
    if (!m_solver) return;
    const double c[3] = {center.x, center.y, center.z};
    check(mpmw_init_cylinder(m_solver, c, radius, length, spacing, axis), "initializeCylindricalWorkpiece");
    std::printf("[MPM] particles initialized in cylinder: %d\n", getParticleCount());
}

void MPMWorkpieceSolver::setToolMesh(const Mesh& mesh, double contactRadius) {
    if (!m_solver) return;
    std::vector<double> xyz;
    xyz.reserve(mesh.nodes.size() * 3);
    for (const auto& n : mesh.nodes) {
        xyz.push_back(n.position.x);
        xyz.push_back(n.position.y);
        xyz.push_back(n.position.z);
    }
    check(mpmw_set_tool_samples(m_solver, xyz.data(), (int32_t)mesh.nodes.size(), contactRadius), "setToolMesh");
}

void MPMWorkpieceSolver::applyExternalForce(const Vec3& center, double radius, const Vec3& force) {
    if (!m_solver) return;
    const double c[3] = {center.x, center.y, center.z};
    const double f[3] = {force.x, force.y, force.z};
    check(mpmw_apply_external_force(m_solver, c, radius, f), "applyExternalForce");
}

// getParticlesForExport, MPMWorkpieceSolver.cu:870-894
std::vector<MPMParticle> MPMWorkpieceSolver::getParticlesForExport() {
    std::vector<MPMParticle> out;
    if (!m_solver) return out;
    const int n = getParticleCount();
    std::vector<double> pos(3 * n), vel(3 * n), fext(3 * n), mass(n), vol(n), dens(n), T(n), eps(n), rate(n),
        dmg(n), sres(n), stress(6 * n);
    std::vector<int32_t> id(n), status(n);
    mpmw_particle_arrays a{};
    a.n = n;
    a.pos = pos.data(); a.vel = vel.data(); a.fext = fext.data();
    a.mass = mass.data(); a.volume = vol.data(); a.density = dens.data(); a.temperature = T.data();
    a.plastic_strain = eps.data(); a.strain_rate = rate.data(); a.damage = dmg.data();
    a.residual_stress = sres.data(); a.stress = stress.data(); a.backstress = nullptr;
    a.id = id.data(); a.status = status.data();
    check(mpmw_get_particles(m_solver, &a), "getParticlesForExport");
    out.resize(a.n);
    for (int i = 0; i < a.n; ++i) {
        MPMParticle& sp = out[i];
        sp.x = pos[3 * i]; sp.y = pos[3 * i + 1]; sp.z = pos[3 * i + 2];
        sp.vx = vel[3 * i]; sp.vy = vel[3 * i + 1]; sp.vz = vel[3 * i + 2];
        sp.fx = fext[3 * i]; sp.fy = fext[3 * i + 1]; sp.fz = fext[3 * i + 2];
        sp.mass = mass[i];  // volume is not exported (left at its default), as in the reference
        sp.density = dens[i];
        sp.temperature = T[i];
        sp.plasticStrain = eps[i];
        sp.strainRate = rate[i];
        sp.damage = dmg[i];
        sp.residualStress = sres[i];
        sp.stress_xx = stress[6 * i]; sp.stress_yy = stress[6 * i + 1]; sp.stress_zz = stress[6 * i + 2];
        sp.stress_xy = stress[6 * i + 3]; sp.stress_xz = stress[6 * i + 4]; sp.stress_yz = stress[6 * i + 5];
        sp.pressure = -(sp.stress_xx + sp.stress_yy + sp.stress_zz) / 3.0;
        sp.id = id[i];
        sp.status = static_cast<ParticleStatus>(status[i]);
    }
    return out;
}

void MPMWorkpieceSolver::compactParticles() {
    if (m_solver) check(mpmw_compact(m_solver), "compactParticles");
}

void MPMWorkpieceSolver::setToolTipPosition(double x, double y, double z) {
    if (m_solver) check(mpmw_set_tool_tip(m_solver, x, y, z), "setToolTipPosition");
}

void MPMWorkpieceSolver::setDeterministic(bool on) {
    m_deterministic = on;
    if (m_solver) check(mpmw_set_deterministic(m_solver, on ? 1 : 0), "setDeterministic");
}

}  // namespace edgepredict
