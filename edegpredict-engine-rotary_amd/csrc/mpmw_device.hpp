// mpmw_device.hpp -- device-side data views and per-particle / per-node math
// of the MPM workpiece step, templated on the arithmetic type R (float for the
// FP32 bulk path, double for the full-FP64 path).
//
// Semantics follow src/MPMWorkpieceSolver.cu of the reference (cited per
// function); layout is MI355X-first: structure-of-arrays particle state,
// structure-of-arrays grid accumulators, a float4/double4 node record for the
// gather, and all reductions done as fixed-order block partials.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mpmw {

enum : int32_t { kActive = 0, kInactive = 1, kChip = 2, kBoundary = 3, kFixed = 4 };
enum : int32_t { kToolNone = 0, kToolPoints = 1, kToolWedge = 2, kToolDrill = 3, kToolVoxel = 4 };

// A pointer into device memory whose element accesses are global-memory
// instructions.  A plain pointer read from a work item in memory (a batch's
// per-solver records) is generic to the compiler, which then emits flat
// loads and stores: they count in both the vector-memory and the LDS /
// scalar wait counters, so every wait for an LDS result or a scalar load
// also waits for the outstanding particle loads.  Indexing through gp casts
// to the global address space, which gives global instructions with a
// scalar base and a 32-bit vector offset.
template <typename T> struct gp {
    T *p = nullptr;
    gp() = default;
    __host__ __device__ gp(T *q) : p(q) {}
    template <typename U> __host__ __device__ gp(const gp<U> &o) : p(o.p) {}
    __host__ __device__ operator T *() const { return p; }
#if defined(__HIP_DEVICE_COMPILE__)
    __device__ __forceinline__ __attribute__((address_space(1))) T &operator[](int64_t i) const {
        return ((__attribute__((address_space(1))) T *)p)[i];
    }
#endif
};

template <typename R> struct Vec4;
template <> struct Vec4<float> { using type = float4; };
template <> struct Vec4<double> { using type = double4; };
template <typename R> using vec4_t = typename Vec4<R>::type;

// Material constants after MPMWorkpieceSolver::initialize
// (MPMWorkpieceSolver.cu:428-451), in kernel precision.
// The divisors of the update that are material constants are passed with
// their reciprocals (FP32 multiplies by the reciprocal, FP64 divides, see
// divc), and flow_floor is a lower bound of johnsonCookStress over all
// arguments (-inf when none is known), below which a trial stress is elastic
// without evaluating the flow stress.
template <typename R> struct Material {
    R shear, bulk, density, cp, melt, ref_temp;
    R jcA, jcB, jcN, jcC, jcM, ref_rate;
    R failure_strain, dth, eps_crit, chip_ratio;
    R rc, c_damp, dT_max, H;
    R d_ref_rate, inv_ref_rate;   // max(ref_rate, 1e-9)
    R d_t_span, inv_t_span;       // max(melt - ref_temp, 1)
    R d_3g, inv_3g;               // max(3 shear, 1)
    R d_rho_cp, inv_rho_cp;       // max(density cp, 1)
    R d_fail, inv_fail;           // max(failure_strain, 1e-3)
    R d_deg, inv_deg;             // dth 0.3
    R flow_floor;
    int32_t clamp;
};

// a / d for a constant divisor d with reciprocal inv: a multiply in FP32 (one
// more rounding, far below the FP32 tolerance), the division in FP64 (the
// full-precision mode keeps the reference's operations).
__device__ __forceinline__ float divc(float a, float d, float inv) { (void)d; return a * inv; }
__device__ __forceinline__ double divc(double a, double d, double inv) { (void)inv; return a / d; }

// MPMGridParams (MPMWorkpieceSolver.cuh:59-63) plus the integer fixture
// threshold: nodes with k <= kfix satisfy the reference's FP64 test
// pos.z <= minZ + fixture + h (MPMWorkpieceSolver.cu:153), evaluated on the host.
// Origin and h are kept in FP64 as well: particle positions are FP64 in every
// mode, so the cell index and the weight fractions are formed in FP64.
template <typename R> struct GridView {
    double ox, oy, oz, hd;
    double inv_hd;    // 1 / h, correctly rounded
    double guard;     // |r - round(r)| below which (x - o) * inv_hd may floor differently from (x - o) / h
    R inv_h;
    int32_t nx, ny, nz, kfix;
    int64_t nn;
    // APIC: C = B * apic_k / h^2, then C *= apic_decay (the reference's 3 and
    // 0.99; the textbook quadratic B-spline's 4 and 1)
    double apic_k, apic_decay;
};

// Particle state, structure of arrays.  a0 == nullptr means all backstress is
// zero; fx == nullptr means no external force.  Positions are FP64 in every
// mode: a step moves a particle by ~1e-8 m, a few FP32 ulps of a cm-scale
// coordinate, and the reference's near-empty free-surface nodes turn weight
// errors of that size into large velocity errors.
template <typename R> struct Particles {
    gp<double> x, y, z;
    gp<R> vx, vy, vz;
    gp<R> s0, s1, s2, s3, s4, s5;
    gp<R> a0, a1, a2, a3, a4, a5;
    gp<R> fx, fy, fz;
    gp<R> eps, rate, dmg, sres, T, m, V;
    gp<R> aff[9];              // APIC affine matrix, row-major (transfer variant only)
    R m0, V0;                  // mass and volume of every particle when m == nullptr
    gp<int32_t> status, id;
    int32_t n;
};

// Grid arrays.  acc: 8 SoA channels (mass, p.xyz, f.xyz, T*m) that P2G
// accumulates into.  node: (v.xyz, T) for the gather, with T = NaN marking a
// node of mass <= 1e-18 (skipped by G2P, MPMWorkpieceSolver.cu:286).
// gm/gtm: post-update mass and T*m read by the diffusion stencil.
template <typename R> struct Grid {
    gp<R> acc;
    gp<vec4_t<R>> node;
    gp<R> gm, gtm;
    gp<R> snap;  // optional copy of acc (8 channels) for parity inspection
};

// Tool state for the grid update.  Tool geometry (node positions, sample
// distances, SDF and its gradient) is evaluated in FP64 in every mode.
struct Tool {
    int32_t kind, nsamples;
    gp<const double> sx, sy, sz;      // point samples (SoA)
    double lo[3], hi[3];              // sample AABB grown by r_c
    double rc2;
    // analytic
    double cx, cy, cz, cos_a, sin_a;  // pose
    double vx, vy, vz, omega;         // tool velocity
    double band, mu, eps;
    double sin_rake, cos_rake, sin_clear, cos_clear;
    double radius, tan_helix, sin_half_point, cos_half_point, flute_offset, flute_radius, flute_length;
    // voxel SDF (rest frame): value of voxel (i, j, k) at origin + (i, j, k
    // + 0.5) * voxel, index i + nx * (j + ny * k)
    gp<const float> sdf;
    int32_t snx, sny, snz, pad_;
    double sox, soy, soz, svox;
};

// --------------------------------------------------------------- helpers --

template <typename R> __device__ __forceinline__ R rmax(R a, R b) { return fmax(a, b); }
template <typename R> __device__ __forceinline__ R rmin(R a, R b) { return fmin(a, b); }

// Square root and quotient of the particle update.  FP32: the hardware
// square root and reciprocal (within 1 ulp each; the IEEE-rounded forms cost
// a correction sequence of ~15 instructions apiece); FP64: IEEE.
template <typename R> __device__ __forceinline__ R vsqrt(R x) { return sqrt(x); }
template <> __device__ __forceinline__ float vsqrt<float>(float x) { return __builtin_amdgcn_sqrtf(x); }
template <typename R> __device__ __forceinline__ R vdiv(R a, R b) { return a / b; }
template <> __device__ __forceinline__ float vdiv<float>(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

// equivalentStress, MPMWorkpieceSolver.cu:18-24
template <typename R>
__device__ __forceinline__ R von_mises(R xx, R yy, R zz, R xy, R xz, R yz) {
    R s1 = xx - yy, s2 = yy - zz, s3 = zz - xx;
    return vsqrt(R(0.5) * (s1 * s1 + s2 * s2 + s3 * s3 + R(6) * (xy * xy + xz * xz + yz * yz)));
}

// johnsonCookStress, MPMWorkpieceSolver.cu:43-50
// pow(x, y) for x >= 0, with the common x == 0, y > 0 case (no plastic
// strain yet, room temperature) answered without the library call.
template <typename R> __device__ __forceinline__ R pow_nonneg(R x, R y) {
    return (x == R(0) && y > R(0)) ? R(0) : pow(x, y);
}

template <typename R>
__device__ __forceinline__ R jc_flow(const Material<R> &M, R eps, R rate, R T) {
    R strainTerm = M.jcA + M.jcB * pow_nonneg(rmax(eps, R(0)), M.jcN);
    R rateRatio = rmax(divc(rate, M.d_ref_rate, M.inv_ref_rate), R(1));
    R rateTerm = R(1) + M.jcC * log(rateRatio);
    R tStar = rmax(R(0), rmin(R(1), divc(T - M.ref_temp, M.d_t_span, M.inv_t_span)));
    R thermalTerm = rmax(R(0.01), R(1) - pow_nonneg(tStar, M.jcM));
    return strainTerm * rateTerm * thermalTerm;
}

// Base cell and the 1-D weights of computeWeights, MPMWorkpieceSolver.cu:54-62
// and :77-79: w1 = r - floor(r), w0 = 1 - w1, both formed in FP64 and then
// rounded once to R.
template <typename R> struct Frac {
    int b[3];
    R w0[3], w1[3];
};

//
// FP32 mode: r = (x - o) * inv_hd is within two ulps of the correctly rounded
// quotient (x - o) / h, so its floor can differ from the quotient's only when
// r lies within g.guard of an integer; there the quotient is formed by
// division.  The base cell is therefore always floor((x - o) / h) in IEEE
// FP64, the reference's (and the binning's) decision; the fraction differs
// from the divided one by at most two ulps of r before its rounding to FP32.
// FP64 mode divides.
template <typename R>
__device__ __forceinline__ Frac<R> base_and_frac(const GridView<R> &g, double x, double y, double z) {
    Frac<R> f;
    const double d[3] = {x - g.ox, y - g.oy, z - g.oz};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        double r = sizeof(R) == 4 ? d[a] * g.inv_hd : d[a] / g.hd;
        double fl = floor(r);
        double t = r - fl;
        if (sizeof(R) == 4 && !(t > g.guard && t < 1.0 - g.guard)) {
            r = d[a] / g.hd;
            fl = floor(r);
            t = r - fl;
        }
        f.b[a] = (int)fl;
        f.w1[a] = (R)t;
        f.w0[a] = (R)(1.0 - t);
    }
    return f;
}

// One of the 8 trilinear nodes (dz outer, dy, dx inner), :64-89.
template <typename R>
__device__ __forceinline__ bool corner(const GridView<R> &g, const Frac<R> &f, int n, int64_t &idx, R &w, R &gx,
                                       R &gy, R &gz) {
    int dx = n & 1, dy = (n >> 1) & 1, dz = (n >> 2) & 1;
    int i = f.b[0] + dx, j = f.b[1] + dy, k = f.b[2] + dz;
    if (i < 0 || j < 0 || k < 0 || i >= g.nx || j >= g.ny || k >= g.nz) return false;
    R wx = dx ? f.w1[0] : f.w0[0];
    R wy = dy ? f.w1[1] : f.w0[1];
    R wz = dz ? f.w1[2] : f.w0[2];
    R dwx = dx ? g.inv_h : -g.inv_h;
    R dwy = dy ? g.inv_h : -g.inv_h;
    R dwz = dz ? g.inv_h : -g.inv_h;
    idx = ((int64_t)k * g.ny + j) * g.nx + i;
    w = wx * wy * wz;
    gx = dwx * wy * wz;
    gy = wx * dwy * wz;
    gz = wx * wy * dwz;
    return w > R(0);
}

// --------------------------------------------------------------- tools ----

__device__ __forceinline__ void tool_frame(const Tool &t, double x, double y, double z, double q[3]) {
    double dx = x - t.cx, dy = y - t.cy;
    q[0] = dx * t.cos_a + dy * t.sin_a;
    q[1] = -dx * t.sin_a + dy * t.cos_a;
    q[2] = z - t.cz;
}

__device__ __forceinline__ double phi_local(const Tool &t, const double q[3]) {
    if (t.kind == kToolWedge) {
        double pr = t.cos_rake * q[0] + t.sin_rake * q[2];
        double pf = -t.sin_clear * q[0] - t.cos_clear * q[2];
        return fmax(pr, pf);
    }
    // two-flute helical drill
    double r = sqrt(q[0] * q[0] + q[1] * q[1]);
    double tt = r * t.sin_half_point + q[2] * t.cos_half_point;
    double cone = (tt < 0.0) ? sqrt(r * r + q[2] * q[2]) : (r * t.cos_half_point - q[2] * t.sin_half_point);
    double body = fmax(fmax(cone, r - t.radius), q[2] - t.flute_length);
    double psi = q[2] * t.tan_helix / t.radius;
    double gmin = 1e300;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        double ang = psi + k * 3.14159265358979323846;
        double fcx = t.flute_offset * cos(ang), fcy = t.flute_offset * sin(ang);
        double ex = q[0] - fcx, ey = q[1] - fcy;
        gmin = fmin(gmin, sqrt(ex * ex + ey * ey) - t.flute_radius);
    }
    return fmax(body, -gmin);
}

// Voxel SDF sampler (the ToolSDF texture, ToolSDF.cu:534-567): trilinear over
// voxel centres with the edge voxels repeated outside (clamp addressing),
// exact weights in FP64 over the FP32 voxel values.
__device__ __forceinline__ double sdf_sample(const Tool &t, const double q[3]) {
    const double org[3] = {t.sox, t.soy, t.soz};
    const int dims[3] = {t.snx, t.sny, t.snz};
    int i0[3], i1[3];
    double f[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double u = (q[k] - org[k]) / t.svox - 0.5;
        const double fl = floor(u);
        f[k] = u - fl;
        const int a = fl < -1.0 ? -1 : (fl > (double)dims[k] ? dims[k] : (int)fl);
        i0[k] = min(max(a, 0), dims[k] - 1);
        i1[k] = min(max(a + 1, 0), dims[k] - 1);
    }
    double v = 0.0;
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        const int dx = n & 1, dy = (n >> 1) & 1, dz = (n >> 2) & 1;
        const double w = (dx ? f[0] : 1.0 - f[0]) * (dy ? f[1] : 1.0 - f[1]) * (dz ? f[2] : 1.0 - f[2]);
        const int64_t idx = (dx ? i1[0] : i0[0]) +
                            (int64_t)t.snx * ((dy ? i1[1] : i0[1]) + (int64_t)t.sny * (dz ? i1[2] : i0[2]));
        v += w * (double)t.sdf[idx];
    }
    return v;
}

// The sampler's range test (ContactSolver.cu:837-840): texel coordinate
// (q - origin) / voxel in [0, n - 1] on every axis.
__device__ __forceinline__ bool sdf_in_range(const Tool &t, const double q[3]) {
    const double org[3] = {t.sox, t.soy, t.soz};
    const int dims[3] = {t.snx, t.sny, t.snz};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double u = (q[k] - org[k]) / t.svox;
        if (!(u >= 0.0) || u > (double)(dims[k] - 1)) return false;
    }
    return true;
}

// phi at world point (x, y, z); the voxel tool has no contact outside its
// texture range.
__device__ __forceinline__ double tool_phi(const Tool &t, double x, double y, double z) {
    double q[3];
    tool_frame(t, x, y, z, q);
    if (t.kind == kToolVoxel) return sdf_in_range(t, q) ? sdf_sample(t, q) : 1e300;
    return phi_local(t, q);
}

// phi for the normal's central differences: the voxel sampler clamps
// instead of the range test, as the reference's gradient samples do
// (ContactSolver.cu:848-859).
__device__ __forceinline__ double tool_phi_grad(const Tool &t, double x, double y, double z) {
    double q[3];
    tool_frame(t, x, y, z, q);
    if (t.kind == kToolVoxel) return sdf_sample(t, q);
    return phi_local(t, q);
}

// True when the drill's phi at (x, y, z) is certainly >= band (no contact),
// from lower bounds of phi_local: phi >= body >= r - radius, >= q_z -
// flute_length, and below the tip plane (q_z < 0) phi >= cone >= -q_z
// sin(point/2) (the cone is sqrt(r^2 + q_z^2) >= |q_z| or r cos - q_z sin
// with r cos >= 0).  A 1e-9 m margin keeps the decision clear of rounding.
__device__ __forceinline__ bool drill_far(const Tool &t, double x, double y, double z) {
    double q[3];
    tool_frame(t, x, y, z, q);
    const double m = t.band + 1e-9;
    const double rr = t.radius + m;
    if (rr <= 0.0 || q[0] * q[0] + q[1] * q[1] >= rr * rr) return true;
    if (q[2] - t.flute_length >= m) return true;
    return -q[2] * t.sin_half_point >= m;
}

// ---------------------------------------------------------- constitutive --

// Register copy of one particle's mutable state.  `sres` holds the largest
// candidate value for the residual stress met during the step (-inf if
// none); the stored value becomes max(stored, sres), which is the
// reference's chain of fmax(residualStress, x) updates.
template <typename R> struct PState {
    double x, y, z;
    R vx, vy, vz;
    R s[6], a[6];
    R eps, rate, dmg, sres, T;
    int32_t status;
};

// ACTIVE-particle block of mpmG2PKernel, MPMWorkpieceSolver.cu:308-382.
template <typename R, bool HAS_BACK>
__device__ __forceinline__ void constitutive_active(const Material<R> &M, PState<R> &p, const R L[9], R dt) {
    R Dxx = L[0], Dyy = L[4], Dzz = L[8];
    R Dxy = R(0.5) * (L[1] + L[3]);
    R Dxz = R(0.5) * (L[2] + L[6]);
    R Dyz = R(0.5) * (L[5] + L[7]);
    R tr = Dxx + Dyy + Dzz;
    const R tr3 = divc(tr, R(3), R(1) / R(3));
    R devXX = Dxx - tr3, devYY = Dyy - tr3, devZZ = Dzz - tr3;
    R eqRate = vsqrt(R(2) / R(3) *
                    (devXX * devXX + devYY * devYY + devZZ * devZZ + R(2) * (Dxy * Dxy + Dxz * Dxz + Dyz * Dyz)));
    p.rate = eqRate;
    R pressureInc = M.bulk * tr * dt;
    p.s[0] += (R(2) * M.shear * devXX) * dt + pressureInc;
    p.s[1] += (R(2) * M.shear * devYY) * dt + pressureInc;
    p.s[2] += (R(2) * M.shear * devZZ) * dt + pressureInc;
    p.s[3] += R(2) * M.shear * Dxy * dt;
    p.s[4] += R(2) * M.shear * Dxz * dt;
    p.s[5] += R(2) * M.shear * Dyz * dt;
    R sh[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) sh[c] = HAS_BACK ? p.s[c] - p.a[c] : p.s[c];
    R seq = von_mises(sh[0], sh[1], sh[2], sh[3], sh[4], sh[5]);
    // seq <= flow_floor <= flow: elastic, the flow stress is not needed
    const R flow = (seq > M.flow_floor && seq > R(1e-9)) ? jc_flow(M, p.eps, p.rate, p.T) : R(0);
    if (seq > flow && seq > R(1e-9) && seq > M.flow_floor) {
        R scale = flow / seq;
#pragma unroll
        for (int c = 0; c < 6; ++c) p.s[c] = HAS_BACK ? p.a[c] + scale * sh[c] : scale * sh[c];
        R dp = divc(seq - flow, M.d_3g, M.inv_3g);
        if (HAS_BACK && M.H > R(0) && dp > R(0)) {
            R ns = M.H * dp / seq;
#pragma unroll
            for (int c = 0; c < 6; ++c) p.a[c] += ns * sh[c];
        }
        p.eps += dp;
        R dT = divc(R(0.9) * flow * dp, M.d_rho_cp, M.inv_rho_cp);
        if (M.clamp) dT = rmin(dT, M.dT_max);
        p.T = rmin(p.T + dT, M.melt);
    }
    if (eqRate > R(0)) p.dmg += divc(eqRate * dt, M.d_fail, M.inv_fail);
    if (p.dmg > R(0.7) * M.dth && p.dmg < M.dth) {
        R degrade = divc(M.dth - p.dmg, M.d_deg, M.inv_deg);
        degrade = rmax(R(0), degrade);
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            p.s[c] *= degrade;
            if (HAS_BACK) p.a[c] *= degrade;
        }
        p.sres = rmax(p.sres, von_mises(p.s[0], p.s[1], p.s[2], p.s[3], p.s[4], p.s[5]));
    }
}

// Chip transition, MPMWorkpieceSolver.cu:384-397.
template <typename R> __device__ __forceinline__ void chip_transition(const Material<R> &M, PState<R> &p) {
    if (p.status == kActive && (p.dmg >= M.dth || p.eps >= M.eps_crit || p.T >= M.melt * M.chip_ratio)) {
        p.status = kChip;
        p.sres = rmax(p.sres, von_mises(p.s[0], p.s[1], p.s[2], p.s[3], p.s[4], p.s[5]));
#pragma unroll
        for (int c = 0; c < 6; ++c) { p.s[c] = R(0); p.a[c] = R(0); }
        p.dmg = rmax(p.dmg, M.dth);
    }
}

// chipFractureAndSeparationKernel, MPMWorkpieceSolver.cu:667-725.  The
// particle id (seed of the fracture kick) is read only when a kick happens.
template <typename R>
__device__ __forceinline__ void chip_pass(const Material<R> &M, PState<R> &p, gp<const int32_t> ids, int i,
                                          double zcut) {
    if (p.status == kActive) {
        if (p.z < zcut && p.eps > R(1e-6)) {
            p.status = kChip;
            p.sres = rmax(p.sres, sqrt(p.s[0] * p.s[0] + p.s[1] * p.s[1] + p.s[2] * p.s[2]));
#pragma unroll
            for (int c = 0; c < 6; ++c) { p.s[c] = R(0); p.a[c] = R(0); }
            p.dmg = rmax(p.dmg, M.dth);
        }
        return;
    }
    if (p.status == kChip) {
        if (p.dmg < M.dth * R(1.5)) return;
        uint32_t seed = (uint32_t)ids[i] * 2654435761u;
        R rx = (R)((seed >> 16) & 0xFFu) / R(255) - R(0.5);
        R ry = (R)((seed >> 8) & 0xFFu) / R(255) - R(0.5);
        R rz = (R)(seed & 0xFFu) / R(255) - R(0.5);
        R speed = sqrt(p.vx * p.vx + p.vy * p.vy + p.vz * p.vz);
        R mag = speed * R(0.3) + R(0.5);
        p.vx += rx * mag;
        p.vy += ry * mag;
        p.vz += rz * mag;
        p.dmg = M.dth * R(0.5);
    }
}

}  // namespace mpmw
