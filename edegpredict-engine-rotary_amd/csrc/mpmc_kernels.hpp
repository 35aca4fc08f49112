// mpmc_kernels.hpp -- device side of the tool FEM step (the reference's
// FEMSolver, src/FEMSolver.cu, dynamic mode) and of the spatial-hash XPBD
// contact between F-tracking particles and the tool's surface nodes
// (src/ContactSolver.cu, src/SharedSpatialHash.cu).
//
// Layout: tool nodes in structure-of-arrays, FP64 (the tool is small: a
// 50 k-tet tool is ~10 k nodes); tets as 96-byte records (4 node indices,
// volume, inverse reference shape) so a thread reads its element in one
// contiguous run.  The hash keeps, besides the reference's cell start / end
// table, a sorted copy of each surface node's contact record (position,
// temperature, velocity, status and original index: 64 bytes, one half
// cache line) so a particle's probe of a cell reads consecutive records.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mpms_kernels.hpp"

namespace mpmc {

constexpr int kBlock = 256;
constexpr double kPi = 3.14159265358979323846;
enum : int32_t { kNodeOk = 0, kNodeWorn = 1, kNodeFailed = 2 };
enum : int32_t { kRegionShank = 8 };

struct Nodes {
    double *x, *y, *z, *ox, *oy, *oz, *vx, *vy, *vz, *fx, *fy, *fz;
    double *mass, *T, *vm, *wear, *heat, *cond, *fn, *pen;
    double *s[6], *sw, *loff[3];
    int32_t *status, *fixed, *in_contact, *driven, *region;
    int32_t n;
};

struct Elem {
    int32_t n[4];
    double volume;
    double inv_dm[9];
};
static_assert(sizeof(Elem) == 96, "element record");

struct Embed {
    int32_t surf, tet[4], pad_;
    double w[4];
};

inline int blocks(int64_t n) { return (int)((n + kBlock - 1) / kBlock > 0 ? (n + kBlock - 1) / kBlock : 1); }

// ------------------------------------------------------------- tool FEM ---

// applyRotationKernel (FEMSolver.cu:685-725): Rodrigues rotation of the
// current and the original position about the axis through c.
__global__ void __launch_bounds__(kBlock) k_rotate(Nodes N, double cx, double cy, double cz, double ax, double ay,
                                                   double az, double angle) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= N.n) return;
    const double co = cos(angle), si = sin(angle), omc = 1.0 - co;
    double px = N.x[i] - cx, py = N.y[i] - cy, pz = N.z[i] - cz;
    double dot = ax * px + ay * py + az * pz;
    double qx = ay * pz - az * py, qy = az * px - ax * pz, qz = ax * py - ay * px;
    N.x[i] = px * co + qx * si + ax * dot * omc + cx;
    N.y[i] = py * co + qy * si + ay * dot * omc + cy;
    N.z[i] = pz * co + qz * si + az * dot * omc + cz;
    px = N.ox[i] - cx; py = N.oy[i] - cy; pz = N.oz[i] - cz;
    dot = ax * px + ay * py + az * pz;
    qx = ay * pz - az * py; qy = az * px - ax * pz; qz = ax * py - ay * px;
    N.ox[i] = px * co + qx * si + ax * dot * omc + cx;
    N.oy[i] = py * co + qy * si + ay * dot * omc + cy;
    N.oz[i] = pz * co + qz * si + az * dot * omc + cz;
}

// setRigidBodyVelocitiesKernel (:787-798)
__global__ void __launch_bounds__(kBlock) k_rigid_velocity(Nodes N, double vx, double vy, double vz, double wz,
                                                           double cx, double cy) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= N.n) return;
    const double rx = N.x[i] - cx, ry = N.y[i] - cy;
    N.vx[i] = vx - wz * ry;
    N.vy[i] = vy + wz * rx;
    N.vz[i] = vz;
}

// transferForcesToTetNodesKernel (:423-450)
__global__ void __launch_bounds__(kBlock) k_transfer_forces(Nodes S, Nodes Tt, const Embed *__restrict__ emb, int ne) {
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= ne) return;
    const Embed E = emb[e];
    if (E.surf < 0) return;
    const int s = E.surf;
    const double fx = S.fx[s], fy = S.fy[s], fz = S.fz[s];
    if (fabs(fx) > 1e-12 || fabs(fy) > 1e-12 || fabs(fz) > 1e-12) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int t = E.tet[k];
            if (t >= 0) {
                unsafeAtomicAdd(Tt.fx + t, fx * E.w[k]);
                unsafeAtomicAdd(Tt.fy + t, fy * E.w[k]);
                unsafeAtomicAdd(Tt.fz + t, fz * E.w[k]);
            }
        }
        S.fx[s] = 0.0; S.fy[s] = 0.0; S.fz[s] = 0.0;
    }
}

// computeSpindleForcesKernel (:385-420)
__global__ void __launch_bounds__(kBlock) k_spindle(Nodes T, double px, double py, double pz, double vx, double vy,
                                                    double vz, double k, double c) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= T.n) return;
    if (!T.driven[i] || T.fixed[i]) return;
    const double fx = (px + T.loff[0][i] - T.x[i]) * k + (vx - T.vx[i]) * c;
    const double fy = (py + T.loff[1][i] - T.y[i]) * k + (vy - T.vy[i]) * c;
    const double fz = (pz + T.loff[2][i] - T.z[i]) * k + (vz - T.vz[i]) * c;
    // no other writer of these nodes' forces runs concurrently in this launch
    T.fx[i] += fx; T.fy[i] += fy; T.fz[i] += fz;
}

// resetNodeStressKernel (:372-380)
__global__ void __launch_bounds__(kBlock) k_reset_stress(Nodes T) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= T.n) return;
#pragma unroll
    for (int c = 0; c < 6; ++c) T.s[c][i] = 0.0;
    T.sw[i] = 0.0;
}

// The element polar iteration (:217-278): R <- (g R + g^-1 Y) / 2 with
// Y = R^-1 (the reference) or R^-T (textbook), at most 20 steps, stop when
// no entry moves by 1e-6 or more, or when |det R| < 1e-12.
__device__ __forceinline__ void elem_polar(const double F[9], bool textbook, double R[9]) {
#pragma unroll
    for (int i = 0; i < 9; ++i) R[i] = F[i];
    for (int it = 0; it < 20; ++it) {
        double cof[9];
        cof[0] = R[4] * R[8] - R[5] * R[7];
        cof[1] = -(R[3] * R[8] - R[5] * R[6]);
        cof[2] = R[3] * R[7] - R[4] * R[6];
        cof[3] = -(R[1] * R[8] - R[2] * R[7]);
        cof[4] = R[0] * R[8] - R[2] * R[6];
        cof[5] = -(R[0] * R[7] - R[1] * R[6]);
        cof[6] = R[1] * R[5] - R[2] * R[4];
        cof[7] = -(R[0] * R[5] - R[2] * R[3]);
        cof[8] = R[0] * R[4] - R[1] * R[3];
        const double det = R[0] * cof[0] + R[1] * cof[1] + R[2] * cof[2];
        if (fabs(det) < 1e-12) break;
        const double idet = 1.0 / det;
        double nR1 = 0, nRinf = 0, nC1 = 0, nCinf = 0;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const double sr = fabs(R[3 * i]) + fabs(R[3 * i + 1]) + fabs(R[3 * i + 2]);
            const double sc = (fabs(cof[i]) + fabs(cof[3 + i]) + fabs(cof[6 + i])) * idet;
            nRinf = sr > nRinf ? sr : nRinf;
            nC1 = sc > nC1 ? sc : nC1;
        }
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const double sr = fabs(R[j]) + fabs(R[3 + j]) + fabs(R[6 + j]);
            const double sc = (fabs(cof[3 * j]) + fabs(cof[3 * j + 1]) + fabs(cof[3 * j + 2])) * idet;
            nR1 = sr > nR1 ? sr : nR1;
            nCinf = sc > nCinf ? sc : nCinf;
        }
        const double g = pow((nC1 * nCinf) / fmax(1e-12, nR1 * nRinf), 0.25), ig = 1.0 / g;
        double md = 0;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                // g^-1 * cofactor * (1 / det), left to right as the reference
                const double y = ig * (textbook ? cof[3 * i + j] : cof[3 * j + i]) * idet;
                const double nv = 0.5 * (g * R[3 * i + j] + y);
                const double d = fabs(R[3 * i + j] - nv);
                md = d > md ? d : md;
                R[3 * i + j] = nv;
            }
        if (md < 1e-6) break;
    }
}

// computeElementForcesKernel (:167-342): corotational linear tet, erosion
// of inverted elements, nodal forces -V0 P Dm^-T, volume-weighted Cauchy
// stress to the nodes.
__global__ void __launch_bounds__(kBlock) k_elem_forces(Nodes T, const Elem *__restrict__ el, int ne, double young,
                                                        double nu, int textbook) {
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= ne) return;
    const Elem E = el[e];
    const int N = T.n;
    const int n0 = E.n[0], n1 = E.n[1], n2 = E.n[2], n3 = E.n[3];
    if (n0 < 0 || n0 >= N || n1 >= N || n2 >= N || n3 >= N) return;
    const int nd[4] = {n0, n1, n2, n3};
    const double x0 = T.x[n0], y0 = T.y[n0], z0 = T.z[n0];
    double Ds[9];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        Ds[c] = T.x[nd[c + 1]] - x0;
        Ds[3 + c] = T.y[nd[c + 1]] - y0;
        Ds[6 + c] = T.z[nd[c + 1]] - z0;
    }
    const double *M = E.inv_dm;
    double F[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) F[3 * i + j] = Ds[3 * i] * M[j] + Ds[3 * i + 1] * M[3 + j] + Ds[3 * i + 2] * M[6 + j];
    const double detF = F[0] * (F[4] * F[8] - F[5] * F[7]) - F[1] * (F[3] * F[8] - F[5] * F[6]) +
                        F[2] * (F[3] * F[7] - F[4] * F[6]);
    double erosion = 1.0;
    if (detF < 0.1) {
        erosion = detF / 0.1;
        erosion = erosion * erosion;
        const double sc = cbrt(0.1 / fmax(detF, 1e-9));
#pragma unroll
        for (int i = 0; i < 9; ++i) F[i] *= sc;
    }
    if (erosion < 1e-6) return;
    double R[9];
    elem_polar(F, textbook != 0, R);
    double eps[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            double s = 0;
#pragma unroll
            for (int k = 0; k < 3; ++k) s += R[3 * k + i] * F[3 * k + j];
            eps[3 * i + j] = s - (i == j ? 1.0 : 0.0);
        }
    const double Ee = young * erosion;
    const double lambda = (Ee * nu) / ((1.0 + nu) * (1.0 - 2.0 * nu));
    const double mu = Ee / (2.0 * (1.0 + nu));
    const double tr = eps[0] + eps[4] + eps[8];
    double sig[9], P[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) sig[3 * i + j] = 2.0 * mu * eps[3 * i + j] + (i == j ? lambda * tr : 0.0);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) P[3 * i + j] = R[3 * i] * sig[j] + R[3 * i + 1] * sig[3 + j] + R[3 * i + 2] * sig[6 + j];
    const double V0 = E.volume;
    double f[4][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            f[c + 1][a] = -V0 * (P[3 * a] * M[3 * c] + P[3 * a + 1] * M[3 * c + 1] + P[3 * a + 2] * M[3 * c + 2]);
#pragma unroll
    for (int a = 0; a < 3; ++a) f[0][a] = -(f[1][a] + f[2][a] + f[3][a]);
    const double w = V0 * 0.25;
    const double sv[6] = {P[0] * w, P[4] * w, P[8] * w, P[1] * w, P[2] * w, P[5] * w};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int q = nd[k];
        unsafeAtomicAdd(T.fx + q, f[k][0]);
        unsafeAtomicAdd(T.fy + q, f[k][1]);
        unsafeAtomicAdd(T.fz + q, f[k][2]);
#pragma unroll
        for (int c = 0; c < 6; ++c) unsafeAtomicAdd(T.s[c] + q, sv[c]);
        unsafeAtomicAdd(T.sw + q, w);
    }
}

// normalizeNodeStressKernel (:346-369)
__global__ void __launch_bounds__(kBlock) k_norm_stress(Nodes T) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= T.n) return;
    const double sw = T.sw[i];
    if (!(sw > 1e-30)) return;
    const double iw = 1.0 / sw;
    double s[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) s[c] = T.s[c][i] * iw, T.s[c][i] = s[c];
    const double a = s[0] - s[1], b = s[1] - s[2], c3 = s[2] - s[0];
    const double seq = sqrt(0.5 * (a * a + b * b + c3 * c3 + 6.0 * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5])));
    T.vm[i] = 0.7 * seq + 0.3 * T.vm[i];
}

// integrateNodesKernel (:525-571): force cap at 1e8 m/s^2, symplectic
// Euler, exp(-damping dt) velocity damping
__global__ void __launch_bounds__(kBlock) k_integrate(Nodes T, double dt, double damping) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= T.n) return;
    if (T.fixed[i] || T.status[i] == kNodeFailed) {
        T.vx[i] = T.vy[i] = T.vz[i] = 0.0;
        T.fx[i] = T.fy[i] = T.fz[i] = 0.0;
        return;
    }
    double fx = T.fx[i], fy = T.fy[i], fz = T.fz[i];
    const double m = T.mass[i];
    const double fm = sqrt(fx * fx + fy * fy + fz * fz), maxf = m * 1e8;
    if (fm > maxf && fm > 0) {
        const double sc = maxf / fm;
        fx *= sc; fy *= sc; fz *= sc;
    }
    const double im = 1.0 / fmax(m, 1e-12);
    double vx = T.vx[i] + fx * im * dt, vy = T.vy[i] + fy * im * dt, vz = T.vz[i] + fz * im * dt;
    T.x[i] += vx * dt; T.y[i] += vy * dt; T.z[i] += vz * dt;
    const double df = exp(-damping * dt);
    T.vx[i] = vx * df; T.vy[i] = vy * df; T.vz[i] = vz * df;
    T.fx[i] = T.fy[i] = T.fz[i] = 0.0;
}

// updateSurfaceNodesKernel (:452-485)
__global__ void __launch_bounds__(kBlock) k_update_surface(Nodes S, Nodes T, const Embed *__restrict__ emb, int ne) {
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= ne) return;
    const Embed E = emb[e];
    if (E.surf < 0) return;
    double p[3] = {0, 0, 0}, v[3] = {0, 0, 0};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int t = E.tet[k];
        if (t < 0) continue;
        const double w = E.w[k];
        p[0] += T.x[t] * w; p[1] += T.y[t] * w; p[2] += T.z[t] * w;
        v[0] += T.vx[t] * w; v[1] += T.vy[t] * w; v[2] += T.vz[t] * w;
    }
    const int s = E.surf;
    S.x[s] = p[0]; S.y[s] = p[1]; S.z[s] = p[2];
    S.vx[s] = v[0]; S.vy[s] = v[1]; S.vz[s] = v[2];
}

// transferSurfaceHeatToTetNodesKernel (:487-523)
__global__ void __launch_bounds__(kBlock) k_surface_heat(Nodes S, Nodes T, const Embed *__restrict__ emb, int ne,
                                                         double cp, double fraction) {
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= ne) return;
    const Embed E = emb[e];
    if (E.surf < 0) return;
    const int s = E.surf;
    const double Ts = S.T[s];
    if (Ts <= 25.01) return;
    double tt = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (E.tet[k] >= 0) tt += T.T[E.tet[k]] * E.w[k];
    const double dT = Ts - tt;
    if (dT <= 0.0) return;
    const double m = fmax(S.mass[s], 1e-12);
    const double q = dT * m * cp * fraction;
    S.T[s] = Ts - q / (m * cp);
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (E.tet[k] >= 0) unsafeAtomicAdd(T.heat + E.tet[k], q * E.w[k]);
}

// applyTetHeatKernel (:87-97)
__global__ void __launch_bounds__(kBlock) k_apply_tet_heat(Nodes T, double cp) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= T.n) return;
    const double m = T.mass[i];
    if (m > 1e-15 && cp > 0) {
        T.T[i] += (T.heat[i] + T.cond[i]) / (m * cp);
        T.heat[i] = 0.0;
        T.cond[i] = 0.0;
    }
}

// thermalConductionTetKernel (:38-84)
__global__ void __launch_bounds__(kBlock) k_conduction(Nodes T, const Elem *__restrict__ el, int ne, double k,
                                                       double cp, double dt) {
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= ne) return;
    const Elem E = el[e];
    const int N = T.n;
    if (E.n[0] < 0 || E.n[0] >= N || E.n[3] >= N) return;
    const double *M = E.inv_dm;
    double g[3][4];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        g[i][1] = M[i]; g[i][2] = M[3 + i]; g[i][3] = M[6 + i];
        g[i][0] = -(g[i][1] + g[i][2] + g[i][3]);
    }
    double Tn[4], Q[4] = {0, 0, 0, 0};
#pragma unroll
    for (int q = 0; q < 4; ++q) Tn[q] = T.T[E.n[q]];
    const double kVdt = k * E.volume * dt;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) Q[i] -= kVdt * (g[0][i] * g[0][j] + g[1][i] * g[1][j] + g[2][i] * g[2][j]) * Tn[j];
#pragma unroll
    for (int q = 0; q < 4; ++q)
        if (T.mass[E.n[q]] > 1e-15 && cp > 0) unsafeAtomicAdd(T.cond + E.n[q], Q[q]);
}

// applyThermalBCKernel (:112-165) with ambient 25, Stefan-Boltzmann
// 5.670374419e-8 and emissivity 0.7 (:1276-1288)
__global__ void __launch_bounds__(kBlock) k_thermal_bc(Nodes T, double htc, double density, double cp, double dt) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= T.n) return;
    const double amb = 25.0;
    if (T.region[i] == kRegionShank) {
        T.T[i] = amb;
        return;
    }
    const double m = T.mass[i];
    const double area = pow(fmax(m / density, 1e-15), 2.0 / 3.0);
    double t = T.T[i];
    if (htc > 0.0 && t > amb + 0.1) {
        const double dT = t - amb;
        const double q = htc * area * dT * dt;
        const double mc = fmax(m * cp, 1e-15);
        t -= fmin(q / mc, dT * 0.5);
    }
    if (t > amb + 10.0) {
        const double Tn = t + 273.15, Ta = amb + 273.15;
        const double q = 0.7 * 5.670374419e-8 * area * (Tn * Tn * Tn * Tn - Ta * Ta * Ta * Ta) * dt;
        const double mc = fmax(m * cp, 1e-15);
        t -= fmin(q / mc, (t - amb) * 0.3);
    }
    T.T[i] = t;
}

// updateWearKernel (:573-588) with A 1e-9, B 1000 (:1291-1292)
__global__ void __launch_bounds__(kBlock) k_wear(Nodes S, double dt) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= S.n) return;
    if (!S.in_contact[i]) return;
    const double v = sqrt(S.vx[i] * S.vx[i] + S.vy[i] * S.vy[i] + S.vz[i] * S.vz[i]);
    double TK = S.T[i] + 273.15;
    if (TK < 300) TK = 300;
    const double w = S.wear[i] + 1e-9 * S.vm[i] * v * exp(-1000.0 / TK) * dt;
    S.wear[i] = w;
    if (w > 0.3e-3) S.status[i] = kNodeWorn;
}

// ----------------------------------------------------------- the hash ---

struct ContactParams {
    double r, k, zeta, friction, hp, shear_yield, max_p;
    double A, B, n, C, m, rate_ref, T_ref, T_melt;
    double tool_cp, work_cp, work_rho, tool_melt, work_melt, htc, emissivity, sb, ambient, tool_rho;
    double latent, solidus, liquidus;
    int32_t jc_enabled, pad_;
};

// the cell mix of computeNodeHashKernel and hashToolCell
__device__ __forceinline__ uint64_t cell_mix(int cx, int cy, int cz) {
    return ((uint64_t)cx * 73856093ULL) ^ ((uint64_t)cy * 19349663ULL) ^ ((uint64_t)cz * 83492791ULL);
}

// A double's bits mapped to an unsigned key that orders like the double
// (for min / max by integer atomics).
__device__ __forceinline__ unsigned long long ordered_bits(double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    return (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
}
__device__ __forceinline__ double from_ordered_bits(unsigned long long k) {
    const unsigned long long b = (k & 0x8000000000000000ull) ? (k & 0x7fffffffffffffffull) : ~k;
    return __longlong_as_double((long long)b);
}

// resetContactAccumulatorsKernel (ContactSolver.cu:636-648), and the tool
// box's min / max keys set to their identities (box[0..2] min, [3..5] max)
__global__ void __launch_bounds__(kBlock) k_reset_contact(Nodes S, unsigned long long *box) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < 6 && blockIdx.x == 0) box[i] = i < 3 ? ~0ull : 0ull;
    if (i >= S.n) return;
    S.fx[i] = 0.0; S.fy[i] = 0.0; S.fz[i] = 0.0;
    S.heat[i] = 0.0; S.fn[i] = 0.0; S.pen[i] = 0.0;
    S.in_contact[i] = 0;
}

// computeNodeHashKernel (SharedSpatialHash.cu:9-31): the slot of the node's
// cell, with the node index as the sort's payload; and the tool box
__global__ void __launch_bounds__(kBlock) k_node_hash(Nodes S, double ox, double oy, double oz, double cell,
                                                      uint32_t table, uint32_t *keys, int32_t *idx,
                                                      unsigned long long *box) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    if (i < S.n) {
        const double x = S.x[i], y = S.y[i], z = S.z[i];
        int cx = (int)floor((x - ox) / cell), cy = (int)floor((y - oy) / cell), cz = (int)floor((z - oz) / cell);
        cx = cx > 0 ? cx : 0;
        cy = cy > 0 ? cy : 0;
        cz = cz > 0 ? cz : 0;
        const uint64_t h = cell_mix(cx, cy, cz), T = table;
        keys[i] = (uint32_t)((h % T + T) % T);
        idx[i] = i;
        lo[0] = hi[0] = x; lo[1] = hi[1] = y; lo[2] = hi[2] = z;
    }
    // wave min / max, then one atomic per wave and bound
#pragma unroll
    for (int a = 0; a < 3; ++a) {
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) {
            lo[a] = fmin(lo[a], __shfl_xor(lo[a], d, 64));
            hi[a] = fmax(hi[a], __shfl_xor(hi[a], d, 64));
        }
    }
    if ((threadIdx.x & 63) == 0 && lo[0] <= hi[0]) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            atomicMin(box + a, ordered_bits(lo[a]));
            atomicMax(box + 3 + a, ordered_bits(hi[a]));
        }
    }
}

__global__ void __launch_bounds__(kBlock) k_fill_i32(int32_t *p, int32_t v, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) p[i] = v;
}

// findToolCellBoundsKernel (ContactSolver.cu:123-137) over the sorted keys
__global__ void __launch_bounds__(kBlock) k_cell_bounds(const uint32_t *__restrict__ keys, int n, int32_t *start,
                                                        int32_t *end) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t h = keys[i];
    if (i == 0 || keys[i - 1] != h) start[h] = i;
    if (i == n - 1 || keys[i + 1] != h) end[h] = i + 1;
}

// The sorted contact records: (x, y, z, T) and (vx, vy, vz, status | index)
struct NodeRec {
    double4 a, b;
};

__global__ void __launch_bounds__(kBlock) k_pack_records(Nodes S, const int32_t *__restrict__ order, int n,
                                                         NodeRec *rec) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const int j = order[i];
    NodeRec r;
    r.a = make_double4(S.x[j], S.y[j], S.z[j], S.T[j]);
    const long long tag = ((long long)S.status[j] << 32) | (long long)(uint32_t)j;
    r.b = make_double4(S.vx[j], S.vy[j], S.vz[j], __longlong_as_double(tag));
    rec[i] = r;
}

// deviceJCFlowStress (ContactSolver.cu:143-157)
__device__ __forceinline__ double jc_flow(const ContactParams &P, double eps, double rate, double T) {
    const double st = P.A + P.B * pow(fmax(eps, 0.0), P.n);
    const double rr = fmax(rate / fmax(P.rate_ref, 1e-9), 1.0);
    const double rt = 1.0 + P.C * log(rr);
    const double ts = T > P.T_ref ? fmin(1.0, (T - P.T_ref) / fmax(P.T_melt - P.T_ref, 1.0)) : 0.0;
    const double th = fmax(0.01, 1.0 - pow(ts, P.m));
    return fmax(st * rt * th, 1.0);
}

struct ContactArgs {
    double ox, oy, oz, dt, max_pen, scale;
    uint32_t table;
    int32_t nn;
    const int32_t *start, *end;
    const NodeRec *rec;
    const unsigned long long *box;
    unsigned long long *counts;   // [0] contacts, [1] particles that probed the hash (inside the box)
    double *totals;               // [0] heat, [1] force
};

// The particles that can touch a tool node: not INACTIVE or
// FIXED_BOUNDARY, and inside the tool's box grown by the contact radius
// (a particle outside it can touch no node).  Their indices are appended to
// cand (*ncand of them, one atomic per wave that has any, in no fixed
// order); each thread tests kCullPer particles, their loads issued first,
// so the pass over the whole workpiece streams its 28 bytes per particle.
constexpr int kCullPer = 4;
template <typename R>
__global__ void __launch_bounds__(kBlock) k_contact_cull(mpms::Particles<R> P, double r,
                                                         const unsigned long long *__restrict__ box,
                                                         int32_t *__restrict__ cand, int32_t *__restrict__ ncand) {
    const int lane = threadIdx.x & 63;
    const int base = blockIdx.x * (kBlock * kCullPer) + threadIdx.x;
    int32_t st[kCullPer];
    double x[kCullPer], y[kCullPer], z[kCullPer];
#pragma unroll
    for (int u = 0; u < kCullPer; ++u) {
        const int i = base + u * kBlock;
        st[u] = mpms::kInactive;
        x[u] = y[u] = z[u] = 0.0;
        if (i < P.n) { st[u] = P.status[i]; x[u] = P.x[i]; y[u] = P.y[i]; z[u] = P.z[i]; }
    }
    const double g = r * (1.0 + 1e-9) + 1e-12;
    const double bx0 = from_ordered_bits(box[0]), by0 = from_ordered_bits(box[1]), bz0 = from_ordered_bits(box[2]),
                 bx1 = from_ordered_bits(box[3]), by1 = from_ordered_bits(box[4]), bz1 = from_ordered_bits(box[5]);
#pragma unroll
    for (int u = 0; u < kCullPer; ++u) {
        const bool c = st[u] != mpms::kInactive && st[u] != mpms::kFixed &&
                       !(x[u] < bx0 - g || x[u] > bx1 + g || y[u] < by0 - g || y[u] > by1 + g || z[u] < bz0 - g ||
                         z[u] > bz1 + g);
        const unsigned long long m = __ballot(c);
        if (m) {   // uniform over the wave
            const int leader = __ffsll((long long)m) - 1;
            int b = 0;
            if (lane == leader) b = atomicAdd(ncand, __popcll(m));
            b = __shfl(b, leader, 64);
            if (c) cand[b + __popcll(m & ((1ull << lane) - 1ull))] = base + u * kBlock;
        }
    }
}

// xpbdContactSpatialHashKernel (ContactSolver.cu:159-408) on the particles
// k_contact_cull kept (cand[0, *ncand)); the others can touch no node and
// skip the 27-cell probe (same results).  One candidate per 32 lanes: lane
// c < 27 probes cell c of the particle's 3 x 3 x 3 neighbourhood (c = 9 (dz
// + 1) + 3 (dy + 1) + dx + 1), so the cells' dependent loads (bucket, its
// range, its records) run side by side instead of one after another; the
// lanes' position corrections, contact counts and heat for the particle are
// summed over the 32 lanes and the group's first lane updates the
// particle.  Node accumulators go to the original node index (the
// reference adds into its sorted copy, then scatters with atomics: the same
// terms).
template <typename R>
__global__ void __launch_bounds__(kBlock) k_contact(mpms::Particles<R> P, Nodes S, ContactParams C, ContactArgs A,
                                                    const int32_t *__restrict__ cand,
                                                    const int32_t *__restrict__ ncand) {
    const int nc = *ncand;
    const int sub = threadIdx.x & 31;
    const int groups = gridDim.x * (kBlock / 32);
    int counted = 0, probed = 0;
    double heat_gen = 0.0, force_sum = 0.0;
    for (int k = (blockIdx.x * kBlock + (int)threadIdx.x) >> 5; k < nc; k += groups) {   // uniform per 32 lanes
        const int i = cand[k];
        const double px = P.x[i], py = P.y[i], pz = P.z[i];
        const double vx = (double)P.v[0][i], vy = (double)P.v[1][i], vz = (double)P.v[2][i];
        const double mass = (double)P.m[i], vol = (double)P.V[i], Tp = (double)P.T[i];
        // the particle volume (its density is not held: the reference's
        // fallback mass / max(density, 1) with density 0)
        const double pv = vol > 1e-24 ? vol : mass;
        double cx = 0, cy = 0, cz = 0, heat_to_p = 0;
        int count = 0;
        if (sub < 27) {
            const double cell = C.r, r = C.r;
            const int pcx = (int)floor((px - A.ox) / cell), pcy = (int)floor((py - A.oy) / cell),
                      pcz = (int)floor((pz - A.oz) / cell);
            const double kN = fmax(C.k, 0.0);
            const double cN = 2.0 * fmax(C.zeta, 0.0) * sqrt(kN * fmax(mass, 1e-15));
            const double prad = pow(3.0 * pv / (4.0 * kPi), 1.0 / 3.0);
            const double flow =
                C.jc_enabled ? jc_flow(C, (double)P.eps[i], (double)P.rate[i], Tp) : C.shear_yield;
            const double shear_yield = flow / 1.7320508;
            const double mina = 0.01 * pow(fmax(pv, 1e-24), 2.0 / 3.0);
            const int qx = pcx + sub % 3 - 1, qy = pcy + (sub / 3) % 3 - 1, qz = pcz + sub / 9 - 1;
            const uint32_t h = (uint32_t)(cell_mix(qx > 0 ? qx : 0, qy > 0 ? qy : 0, qz > 0 ? qz : 0) %
                                          (uint64_t)A.table);
            const int s0 = A.start[h];
            const int s1 = s0 < 0 ? s0 : A.end[h];
            for (int j = s0; j < s1; ++j) {
                const NodeRec rec = A.rec[j];
                const long long tag = __double_as_longlong(rec.b.w);
                if ((int32_t)(tag >> 32) == kNodeFailed) continue;
                const int node = (int)(uint32_t)(tag & 0xffffffff);
                const double ndx = px - rec.a.x, ndy = py - rec.a.y, ndz = pz - rec.a.z;
                const double dist = sqrt(ndx * ndx + ndy * ndy + ndz * ndz);
                if (dist >= r || dist < 1e-12 || isnan(dist)) continue;
                double pen = r - dist;
                if (pen > A.max_pen) pen = A.max_pen;
                const double id = 1.0 / dist;
                const double nx = ndx * id, ny = ndy * id, nz = ndz * id;
                cx += pen * nx; cy += pen * ny; cz += pen * nz;
                count++;
                const double rvx = vx - rec.b.x, rvy = vy - rec.b.y, rvz = vz - rec.b.z;
                const double vn = rvx * nx + rvy * ny + rvz * nz;
                const double fn = (kN * pen + cN * fmax(0.0, -vn)) * A.scale;
                unsafeAtomicAdd(S.fx + node, -fn * nx);
                unsafeAtomicAdd(S.fy + node, -fn * ny);
                unsafeAtomicAdd(S.fz + node, -fn * nz);
                unsafeAtomicAdd(S.fn + node, fn);
                // penetrations are positive: their bits order like them
                atomicMax((unsigned long long *)(S.pen + node), (unsigned long long)__double_as_longlong(pen));
                const double vtx = rvx - vn * nx, vty = rvy - vn * ny, vtz = rvz - vn * nz;
                const double vt = sqrt(vtx * vtx + vty * vty + vtz * vtz);
                const double Tavg = 0.5 * (Tp + rec.a.w);
                const double Tr = fmin(1.0, fmax(0.0, (Tavg - 200.0) / 600.0));
                const double mu_eff = C.friction * (1.0 - 0.4 * Tr);
                const double cpp = pen > 0.0 ? pen : 0.0;
                const double ea = kPi * (2.0 * prad * cpp - cpp * cpp);
                const double pa = flow > 1.0 ? fn / flow : 0.0;
                double area = ea + pa;
                if (area < mina) area = mina;
                const double sn = fn / fmax(area, 1e-24);
                const double scrit = shear_yield / fmax(mu_eff, 1e-6);
                const double stick = 0.5 * (1.0 + tanh(10.0 * (sn / fmax(scrit, 1e-12) - 1.0)));
                const double tau = (1.0 - stick) * (mu_eff * sn) + stick * shear_yield;
                double ft = tau * area;
                const double fst = vt > 1e-10 ? mass * vt / A.dt : ft;
                ft = fmin(ft, fst);
                const double htc = C.htc * (1.0 + 5.0 * (1.0 - exp(-sn / flow)));
                const double hs = ft * vt * A.dt;
                const double qc = htc * area * (Tp - rec.a.w) * A.dt;
                heat_to_p += (hs * (1.0 - C.hp) - qc) * A.scale;
                unsafeAtomicAdd(S.heat + node, (hs * C.hp + qc) * A.scale);
                if (vt > 1e-15) {
                    unsafeAtomicAdd(S.fx + node, ft * (vtx / vt));
                    unsafeAtomicAdd(S.fy + node, ft * (vty / vt));
                    unsafeAtomicAdd(S.fz + node, ft * (vtz / vt));
                }
                S.in_contact[node] = 1;
                heat_gen += hs * A.scale;
                force_sum += fn;
            }
        }
        // the particle's sums over its 32 lanes
#pragma unroll
        for (int d = 16; d >= 1; d >>= 1) {
            cx += __shfl_xor(cx, d, 32);
            cy += __shfl_xor(cy, d, 32);
            cz += __shfl_xor(cz, d, 32);
            heat_to_p += __shfl_xor(heat_to_p, d, 32);
            count += __shfl_xor(count, d, 32);
        }
        if (sub == 0) {
            ++probed;
            if (count > 0) {
                const double w = 1.0 / count;
                cx *= w; cy *= w; cz *= w;
                P.x[i] = px + cx; P.y[i] = py + cy; P.z[i] = pz + cz;
                const double cm = sqrt(cx * cx + cy * cy + cz * cz);
                if (cm > 1e-15) {
                    const double ic = 1.0 / cm;
                    const double nx = cx * ic, ny = cy * ic, nz = cz * ic;
                    const double vd = vx * nx + vy * ny + vz * nz;
                    if (vd < 0.0) {
                        P.v[0][i] = (R)(vx - vd * nx);
                        P.v[1][i] = (R)(vy - vd * ny);
                        P.v[2][i] = (R)(vz - vd * nz);
                    }
                }
                if (heat_to_p > 0.0) {
                    const double pm = fmax(C.work_rho, 1.0) * fmax(pv, 1e-24);
                    const double dT = heat_to_p / (fmax(pm, 1e-15) * C.work_cp);
                    P.T[i] = (R)fmin(Tp + dT, C.work_melt);
                }
                ++counted;
            }
        }
    }
    // per-wave totals, one atomic per wave that has any
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        heat_gen += __shfl_xor(heat_gen, d, 64);
        force_sum += __shfl_xor(force_sum, d, 64);
        counted += __shfl_xor(counted, d, 64);
        probed += __shfl_xor(probed, d, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        if (counted) {
            atomicAdd(A.counts, (unsigned long long)counted);
            unsafeAtomicAdd(A.totals, heat_gen);
            unsafeAtomicAdd(A.totals + 1, force_sum);
        }
        if (probed) atomicAdd(A.counts + 1, (unsigned long long)probed);
    }
}

// applyNodeHeatAndHertzStressKernel (ContactSolver.cu:665-791), EMA 0.9
__global__ void __launch_bounds__(kBlock) k_node_heat(Nodes S, ContactParams C, double dt) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= S.n) return;
    double heat = S.heat[i], T = S.T[i];
    const double m = S.mass[i];
    if (T > C.ambient + 1.0) {
        const double vol = fmax(m / C.tool_rho, 1e-15);
        const double area = pow(vol, 2.0 / 3.0);
        const double Tn = T + 273.15, Ta = C.ambient + 273.15;
        const double q = C.emissivity * C.sb * area * (Tn * Tn * Tn * Tn - Ta * Ta * Ta * Ta) * dt;
        heat = fmax(heat - q, 0.0);
    }
    const bool phase = C.latent > 0.0;
    const bool mushy = phase && C.solidus > 0.0 && C.liquidus > C.solidus;
    const double Ts = mushy ? C.solidus : C.tool_melt, Tl = mushy ? C.liquidus : C.tool_melt;
    if (heat > 0.0 && phase) {
        const double mm = fmax(m, 1e-12), cp = fmax(C.tool_cp, 1e-12), L = C.latent;
        double H0;
        if (T < Ts) H0 = cp * T;
        else if (T < Tl) H0 = cp * Ts + L * ((T - Ts) / (Tl - Ts));
        else H0 = cp * Ts + L + cp * (T - Tl);
        const double H1 = H0 + heat / mm, Hs = cp * Ts, Hl = cp * Ts + L;
        double Tnew;
        if (H1 <= Hs) Tnew = H1 / cp;
        else if (H1 < Hl) Tnew = Ts + ((H1 - Hs) / (Hl - Hs)) * (Tl - Ts);
        else Tnew = Tl + (H1 - Hl) / C.tool_cp;
        T = fmin(T + (Tnew - T), C.tool_melt);
        if (T >= C.tool_melt) S.status[i] = kNodeFailed;
    } else if (heat > 0.0) {
        const double dT = heat / (fmax(m, 1e-12) * fmax(C.tool_cp, 1e-12));
        T = fmin(T + dT, C.tool_melt);
        if (T >= C.tool_melt) S.status[i] = kNodeFailed;
    }
    S.T[i] = T;
    S.heat[i] = 0.0;
    double sh = 0.0;
    const double fn = S.fn[i], pen = S.pen[i];
    if (fn > 0.0 && pen > 1e-12) {
        const double a2 = C.r * pen;
        if (a2 > 1e-18) sh = fmin((3.0 * fn) / (2.0 * kPi * a2), C.max_p);
    }
    S.vm[i] = 0.9 * S.vm[i] + (1.0 - 0.9) * sh;
    S.fn[i] = 0.0;
    S.pen[i] = 0.0;
}

// MPMSolver::setToolNodes with the tool's surface nodes: interleaved xyz and
// velocity into the F-tracking solver's tool-node buffers
__global__ void __launch_bounds__(kBlock) k_link_tool(Nodes S, double *xyz, double *vel) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= S.n) return;
    xyz[3 * i] = S.x[i]; xyz[3 * i + 1] = S.y[i]; xyz[3 * i + 2] = S.z[i];
    vel[3 * i] = S.vx[i]; vel[3 * i + 1] = S.vy[i]; vel[3 * i + 2] = S.vz[i];
}

}  // namespace mpmc
