#pragma once
// The core types the MPM workpiece step exchanges with its caller, with the
// names, fields and values of include/Types.h of
// Up25-pathan/edegpredict-engine-rotary: ParticleStatus (:37-43), Vec3
// (:163-194), the MPMParticle export record (:199-266, the fields
// getParticlesForExport fills) and the Mesh node list setToolMesh samples
// (:463-488).

#include <cmath>
#include <cstdint>
#include <vector>

namespace edgepredict {

enum class ParticleStatus : int32_t {
    ACTIVE = 0,
    INACTIVE = 1,
    CHIP = 2,
    BOUNDARY = 3,
    FIXED_BOUNDARY = 4
};

struct Vec3 {
    double x, y, z;
    Vec3() : x(0), y(0), z(0) {}
    Vec3(double x_, double y_, double z_) : x(x_), y(y_), z(z_) {}
    Vec3 operator+(const Vec3& o) const { return {x + o.x, y + o.y, z + o.z}; }
    Vec3 operator-(const Vec3& o) const { return {x - o.x, y - o.y, z - o.z}; }
    Vec3 operator*(double s) const { return {x * s, y * s, z * s}; }
    Vec3 operator/(double s) const { return {x / s, y / s, z / s}; }
    double dot(const Vec3& o) const { return x * o.x + y * o.y + z * o.z; }
    double lengthSq() const { return x * x + y * y + z * z; }
    double length() const { return std::sqrt(lengthSq()); }
    Vec3 normalized() const {
        const double len = length();
        return len > 1e-12 ? (*this / len) : Vec3(0, 0, 0);
    }
    static Vec3 zero() { return {0, 0, 0}; }
};

// Export record of one particle (the MPMParticle fields getParticlesForExport
// writes, MPMWorkpieceSolver.cu:870-894).
struct MPMParticle {
    double x = 0, y = 0, z = 0;
    double vx = 0, vy = 0, vz = 0;
    double fx = 0, fy = 0, fz = 0;
    double density = 0, pressure = 0, volume = 0, mass = 0, temperature = 25.0;
    double stress_xx = 0, stress_yy = 0, stress_zz = 0;
    double stress_xy = 0, stress_xz = 0, stress_yz = 0;
    double plasticStrain = 0, strainRate = 0, damage = 0, residualStress = 0;
    int32_t id = -1;
    ParticleStatus status = ParticleStatus::ACTIVE;
};

struct FEMNode {
    Vec3 position;
};

// Surface mesh: setToolMesh samples `nodes` (every stride-th node).
struct Mesh {
    std::vector<FEMNode> nodes;
    bool empty() const { return nodes.empty(); }
    size_t nodeCount() const { return nodes.size(); }
};

}  // namespace edgepredict
