// mpms_kernels.hpp -- device side of the F-tracking MPM step (the reference's
// MPMSolver, src/MPMSolver.cu), templated on the bulk arithmetic type R
// (float: FP32 bulk with FP64 islands; double: FP64 throughout).
//
// Layout: particles in structure-of-arrays (positions FP64 in every mode),
// the deformation gradients held as their differences from the identity,
// H = F - I and Hp = Fp - I: near the identity an FP32 F keeps only ~7
// digits of 1 + strain, so the strain of a 1e-3 deformation would carry a
// 1e-4 relative error into the stress; H keeps the strain's own digits and
// every quantity the stress needs (Fe - I, J - 1, B - I) is formed from it
// without cancellation;
// sorted by their stencil's base node at every grid fit so that a wave's
// P2G atomics and G2P loads fall on nearby nodes; the grid holds only a
// sub-box of the reference's lattice, channels in one contiguous block
// (mass, momentum xyz, force xyz, then the tool claims) cleared by one fill.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mpms {

constexpr int kBlock = 256;
enum : int32_t { kActive = 0, kInactive = 1, kChip = 2, kBoundary = 3, kFixed = 4 };
enum : int32_t { kLodActive = 0, kLodNear = 1, kLodFar = 2 };

template <typename R> struct Particles {
    double *x, *y, *z;
    R *v[3], *ef[3], *m, *V, *T;
    R *F[9], *Fp[9], *C[9], *s[6], *a[6];   // F and Fp hold H = F - I and Hp = Fp - I
    R *eps, *rate, *dmg, *sres;
    int32_t *id, *status, *lod;
    int32_t n;
};

// The lattice (origin, spacing, whole-lattice dims) and the held sub-box.
template <typename R> struct Grid {
    double ox, oy, oz, dx, inv_dx;
    int32_t dims[3];
    int32_t lo[3], nb[3];
    int64_t nn;
    R *mass, *p[3], *f[3], *v[3];
    int32_t *claim;              // index + 1 of the tool node that marked the node, 0 for none
    unsigned long long *overflow;  // particle visits of nodes outside the sub-box
};

// MPMKernelConfig in kernel precision plus the islands' FP64 copies.
template <typename R> struct Mat {
    R mu, lambda, rho_phys, cp, tq, melt, ambient, A, B, n, C, m, H, D1, D2, D3, D4, D5, dth, ref_rate;
    double mu_d, lambda_d, A_d, B_d, n_d, C_d, m_d;
    int32_t corotated, newton, newton_fp64, polar_fp64;
    double island_det;
};

struct StepArgs {
    double dt, tool[3], lod_active, lod_near;
    int32_t lod_enabled, chip_step;
    double fixture, fixture_damping, gravity;
    const double *tool_vel;      // 3 per tool node
    unsigned long long *islands;   // [0] kinematics, [1] Newton, [2] P2G particles outside their tile's window
};

// The step's scalars live in device memory, written by this one-thread
// kernel before the step's kernels: the kernels that read them can then be
// replayed from a captured graph while dt, the tool position and the chip
// step change.
[[maybe_unused]] static __global__ void k_step_args(StepArgs *__restrict__ d, StepArgs v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *d = v;
}

// ------------------------------------------------------------ matrices ---

template <typename K> __device__ __forceinline__ K det3(const K A[9]) {
    return A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) +
           A[2] * (A[3] * A[7] - A[4] * A[6]);
}

template <typename K> __device__ __forceinline__ void mul3(const K A[9], const K B[9], K C[9]) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) C[i * 3 + j] = A[i * 3] * B[j] + A[i * 3 + 1] * B[3 + j] + A[i * 3 + 2] * B[6 + j];
}

// inv3x3 (MPMSolver.cu:50-69): the identity when |det| < 1e-12
template <typename K> __device__ __forceinline__ void inv3(const K A[9], K inv[9]) {
    const K d = det3(A);
    if (fabs(d) < K(1e-12)) {
#pragma unroll
        for (int i = 0; i < 9; ++i) inv[i] = (i == 0 || i == 4 || i == 8) ? K(1) : K(0);
        return;
    }
    const K id = K(1) / d;
    inv[0] = (A[4] * A[8] - A[5] * A[7]) * id;
    inv[1] = -(A[1] * A[8] - A[2] * A[7]) * id;
    inv[2] = (A[1] * A[5] - A[2] * A[4]) * id;
    inv[3] = -(A[3] * A[8] - A[5] * A[6]) * id;
    inv[4] = (A[0] * A[8] - A[2] * A[6]) * id;
    inv[5] = -(A[0] * A[5] - A[2] * A[3]) * id;
    inv[6] = (A[3] * A[7] - A[4] * A[6]) * id;
    inv[7] = -(A[0] * A[7] - A[1] * A[6]) * id;
    inv[8] = (A[0] * A[4] - A[1] * A[3]) * id;
}

// The rotation of F = R U by Higham's scaled Newton iteration
// X <- (g X + X^-T / g) / 2, g = (|X^-1|_1 |X^-1|_inf / (|X|_1 |X|_inf))^(1/4)
// (Higham 1986), until no entry moves by more than tol (30 iterations at
// most).  The reference's FEM tool adds X^-1 where this adds X^-T
// (FEMSolver.cu:225-276); they agree for symmetric X only.
template <typename K> __device__ __forceinline__ void polar(const K F[9], K R[9]) {
    const K tol = sizeof(K) == 8 ? K(1e-14) : K(2e-7);
#pragma unroll
    for (int i = 0; i < 9; ++i) R[i] = F[i];
    for (int it = 0; it < 30; ++it) {
        K Xi[9];
        if (fabs(det3(R)) < K(1e-30)) break;
        inv3(R, Xi);
        K n1 = 0, ninf = 0, m1 = 0, minf = 0;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const K r = fabs(R[i * 3]) + fabs(R[i * 3 + 1]) + fabs(R[i * 3 + 2]);
            const K c = fabs(R[i]) + fabs(R[3 + i]) + fabs(R[6 + i]);
            const K ri = fabs(Xi[i * 3]) + fabs(Xi[i * 3 + 1]) + fabs(Xi[i * 3 + 2]);
            const K ci = fabs(Xi[i]) + fabs(Xi[3 + i]) + fabs(Xi[6 + i]);
            ninf = fmax(ninf, r); n1 = fmax(n1, c); minf = fmax(minf, ri); m1 = fmax(m1, ci);
        }
        const K g = sqrt(sqrt((m1 * minf) / (n1 * ninf))), ig = K(1) / g;
        K diff = 0;
        K Rn[9];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                Rn[i * 3 + j] = K(0.5) * (g * R[i * 3 + j] + ig * Xi[j * 3 + i]);
                diff = fmax(diff, fabs(Rn[i * 3 + j] - R[i * 3 + j]));
            }
#pragma unroll
        for (int i = 0; i < 9; ++i) R[i] = Rn[i];
        if (diff <= tol) break;
    }
}

// det(I + X) - 1 = tr X + (tr^2 X - tr X^2) / 2 + det X, with no
// cancellation for a small X.
template <typename K> __device__ __forceinline__ K det1m(const K X[9]) {
    const K t = X[0] + X[4] + X[8];
    K t2 = 0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) t2 += X[i * 3 + j] * X[j * 3 + i];
    return t + K(0.5) * (t * t - t2) + det3(X);
}

// (I + X)^-1 - I = (-(1 + tr X) X + X^2 - det(X) I) / det(I + X) (the
// adjugate of I + X minus det(I + X) I, divided); the identity when
// |det(I + X)| < 1e-12, as inv3x3 (MPMSolver.cu:50-69) returns.
template <typename K> __device__ __forceinline__ void inv_i_plus_m1(const K X[9], K G[9]) {
    const K d = K(1) + det1m(X);
    if (fabs(d) < K(1e-12)) {
#pragma unroll
        for (int c = 0; c < 9; ++c) G[c] = K(0);
        return;
    }
    const K t = X[0] + X[4] + X[8], dx = det3(X), id = K(1) / d;
    K X2[9];
    mul3(X, X, X2);
#pragma unroll
    for (int c = 0; c < 9; ++c) G[c] = (-(K(1) + t) * X[c] + X2[c] - ((c % 4 == 0) ? dx : K(0))) * id;
}

// The elastic part of p2gKernel (:164-196) from H = F - I and Hp = Fp - I:
// A = Fe - I with Fe = F Fp^-1, J = det Fe clamped to [0.9, 1.1], and the
// Cauchy stress (row-major) of the compressible Neo-Hookean law
// mu/J (B - I) + lambda (J - 1) I (the reference's mu/J B + (lambda (J - 1)
// - mu/J) I with B - I = A + A^T + A A^T), or of the fixed-corotated law
// 2 mu/J (Fe - R) Fe^T + lambda (J - 1) I with R the rotation of Fe, whose
// polar decomposition runs in P (FP64 when the island asks for it).
template <typename K, typename P>
__device__ __forceinline__ void elastic_stress(const K H[9], const K Hp[9], K mu, K lambda, bool corotated, K sig[9],
                                               K &J) {
    K G[9], HG[9], A[9];
    inv_i_plus_m1(Hp, G);
    mul3(H, G, HG);
#pragma unroll
    for (int c = 0; c < 9; ++c) A[c] = H[c] + G[c] + HG[c];
    K Jm1 = det1m(A);
    if (Jm1 < K(-0.1)) Jm1 = K(-0.1);
    if (Jm1 > K(0.1)) Jm1 = K(0.1);
    J = K(1) + Jm1;
    const K Js = fmax(J, K(1e-12));
    if (corotated) {
        P Fe[9], R[9];
#pragma unroll
        for (int c = 0; c < 9; ++c) Fe[c] = (P)A[c] + ((c % 4 == 0) ? P(1) : P(0));
        polar(Fe, R);
        K D[9];   // Fe - R = A - (R - I)
#pragma unroll
        for (int c = 0; c < 9; ++c) D[c] = A[c] - (K)(R[c] - ((c % 4 == 0) ? P(1) : P(0)));
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                K v = D[i * 3 + j];   // (D (I + A)^T)_ij = D_ij + sum_k D_ik A_jk
#pragma unroll
                for (int k = 0; k < 3; ++k) v += D[i * 3 + k] * A[j * 3 + k];
                sig[i * 3 + j] = K(2) * mu / Js * v + (i == j ? lambda * Jm1 : K(0));
            }
    } else {
        const K muJ = mu / Js;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const K bm = A[i * 3 + j] + A[j * 3 + i] + A[i * 3] * A[j * 3] + A[i * 3 + 1] * A[j * 3 + 1] +
                             A[i * 3 + 2] * A[j * 3 + 2];
                sig[i * 3 + j] = muJ * bm + (i == j ? lambda * Jm1 : K(0));
            }
    }
}

// Johnson-Cook flow stress of p2gKernel (:225-232) at plastic strain e.
template <typename K> __device__ __forceinline__ K jc_flow(K A, K B, K n, K C, K m, K e, K edot, K ts) {
    return (A + B * pow(e, n)) * (K(1) + C * log(edot)) * (K(1) - pow(ts, m));
}

// The consistent return (extension): d in [0, q / 3mu] with
// q - 3 mu d - Y(e + d) = 0, Newton from the radial estimate kept in the
// residual's sign bracket (bisection when a step leaves it).
template <typename K>
__device__ __forceinline__ K newton_return(K A, K B, K n, K C, K m, K q, K mu3, K e, K edot, K ts, K y0) {
    K lo = 0, hi = q / mu3;
    K d = fmin(fmax((q - y0) / mu3, K(0)), hi);
    const K rate = (K(1) + C * log(edot)) * (K(1) - pow(ts, m));
    const K tol = sizeof(K) == 8 ? K(1e-13) : K(1e-6);
    for (int it = 0; it < 60; ++it) {
        const K g = q - mu3 * d - (A + B * pow(e + d, n)) * rate;
        if (g > 0) lo = d; else hi = d;
        if (fabs(g) <= tol * q || !(hi > lo)) break;
        const K dg = -mu3 - (B == K(0) ? K(0) : B * n * pow(e + d, n - K(1)) * rate);
        K nd = d - g / dg;
        if (!(nd > lo && nd < hi)) nd = K(0.5) * (lo + hi);
        d = nd;
    }
    return d;
}

// ---------------------------------------------------------------- P2G ---

// The quadratic B-spline of one axis (getWeightsAndGradients, :19-34): the
// fraction is formed in FP64 from the FP64 position.
template <typename R> struct Axis {
    int b;
    R w[3], dw[3];
};

template <typename R> __device__ __forceinline__ Axis<R> axis_weights(double xp, double inv_dx) {
    Axis<R> a;
    const double xf = xp * inv_dx;
    a.b = (int)floor(xf - 0.5);
    const double d0 = xf - a.b, d1 = d0 - 1.0, d2 = d0 - 2.0;
    a.w[0] = (R)(0.5 * (1.5 - d0) * (1.5 - d0));
    a.w[1] = (R)(0.75 - d1 * d1);
    a.w[2] = (R)(0.5 * (1.5 + d2) * (1.5 + d2));
    a.dw[0] = (R)(-(1.5 - d0) * inv_dx);
    a.dw[1] = (R)(-2.0 * d1 * inv_dx);
    a.dw[2] = (R)((1.5 + d2) * inv_dx);
    return a;
}

template <typename R> __device__ __forceinline__ int64_t box_index(const Grid<R> &g, int gx, int gy, int gz) {
    const int i = gx - g.lo[0], j = gy - g.lo[1], k = gz - g.lo[2];
    if ((unsigned)i >= (unsigned)g.nb[0] || (unsigned)j >= (unsigned)g.nb[1] || (unsigned)k >= (unsigned)g.nb[2])
        return -1;
    return ((int64_t)k * g.nb[1] + j) * g.nb[0] + i;
}

// The constitutive part of p2gKernel (:156-343) on one ACTIVE particle in
// kernel precision, with the FP64 islands of the FP32 path: the elastic
// kinematics when |det F| < island_det, the Newton return when newton_fp64.
template <typename R>
__device__ __forceinline__ void constitutive(const Particles<R> &P, int i, const Mat<R> &M, R sig[9], R &J,
                                             bool &skip, unsigned long long *islands) {
    R F[9], Fp[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) { F[c] = P.F[c][i]; Fp[c] = P.Fp[c][i]; }
    // F and Fp here are H = F - I and Hp = Fp - I
    bool isl = false;
    if (sizeof(R) == 4 && M.island_det > 0.0) isl = fabs(1.0 + (double)det1m(F)) < M.island_det;
    if (isl) {
        double Fd[9], Fpd[9], sd[9], Jd;
#pragma unroll
        for (int c = 0; c < 9; ++c) { Fd[c] = F[c]; Fpd[c] = Fp[c]; }
        elastic_stress<double, double>(Fd, Fpd, M.mu_d, M.lambda_d, M.corotated != 0, sd, Jd);
#pragma unroll
        for (int c = 0; c < 9; ++c) sig[c] = (R)sd[c];
        J = (R)Jd;
        atomicAdd(&islands[0], 1ull);
    } else if (M.polar_fp64) {
        elastic_stress<R, double>(F, Fp, M.mu, M.lambda, M.corotated != 0, sig, J);
    } else {
        elastic_stress<R, R>(F, Fp, M.mu, M.lambda, M.corotated != 0, sig, J);
    }
    const R tr = sig[0] + sig[4] + sig[8];
    R s[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) s[c] = sig[c];
    s[0] -= tr / R(3); s[4] -= tr / R(3); s[8] -= tr / R(3);
    R al[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) al[c] = P.a[c][i];
    const R ex = s[0] - al[0], ey = s[4] - al[1], ez = s[8] - al[2];
    const R exy = s[1] - al[3], exz = s[2] - al[4], eyz = s[5] - al[5];
    const R q = sqrt(R(1.5) * (ex * ex + ey * ey + ez * ez + R(2) * (exy * exy + exz * exz + eyz * eyz)));
    R T = P.T[i];
    R ts = (T - M.ambient) / fmax(M.melt - M.ambient, R(1));
    ts = fmin(fmax(ts, R(0)), R(1));
    const R edot = fmax(P.rate[i] / fmax(M.ref_rate, R(1e-12)), R(1e-5));
    R eps = P.eps[i];
    R Y = jc_flow<R>(M.A, M.B, M.n, M.C, M.m, eps, edot, ts);
    if (q > Y && q > R(1e-12)) {
        R de, scale;
        if (M.newton) {
            if (sizeof(R) == 4 && M.newton_fp64) {
                const double dd = newton_return<double>(M.A_d, M.B_d, M.n_d, M.C_d, M.m_d, (double)q, 3.0 * M.mu_d,
                                                        (double)eps, (double)edot, (double)ts, (double)Y);
                const double d5 = fmin(dd, 0.05);
                Y = (R)jc_flow<double>(M.A_d, M.B_d, M.n_d, M.C_d, M.m_d, (double)eps + d5, (double)edot, (double)ts);
                scale = (R)(((double)q - 3.0 * M.mu_d * d5) / (double)q);
                de = (R)d5;
                atomicAdd(&islands[1], 1ull);
            } else {
                de = newton_return<R>(M.A, M.B, M.n, M.C, M.m, q, R(3) * M.mu, eps, edot, ts, Y);
                if (de > R(0.05)) de = R(0.05);
                Y = jc_flow<R>(M.A, M.B, M.n, M.C, M.m, eps + de, edot, ts);
                scale = (q - R(3) * M.mu * de) / q;
            }
        } else {
            scale = Y / q;
            de = (q - Y) / fmax(R(3) * M.mu, R(1e-12));
            if (de > R(0.05)) de = R(0.05);
        }
        s[0] = al[0] + scale * ex; s[4] = al[1] + scale * ey; s[8] = al[2] + scale * ez;
        s[1] = al[3] + scale * exy; s[2] = al[4] + scale * exz; s[5] = al[5] + scale * eyz;
        // the lower triangle keeps the trial deviator (:246-249)
#pragma unroll
        for (int c = 0; c < 9; ++c) sig[c] = s[c] + ((c % 4 == 0) ? tr / R(3) : R(0));
        if (de > R(0)) {
            if (M.H > R(0)) {
                const R ns = M.H * de / q;
                al[0] += ns * ex; al[1] += ns * ey; al[2] += ns * ez;
                al[3] += ns * exy; al[4] += ns * exz; al[5] += ns * eyz;
#pragma unroll
                for (int c = 0; c < 6; ++c) P.a[c][i] = al[c];
            }
            eps += de;
            P.eps[i] = eps;
            {   // Fp <- (I + 1.5 d / |s| s) Fp / det^(1/3) (:277-301), in Hp form:
                // K = Hp + a s (I + Hp), Hp <- (K - (c - 1) I) / c with
                // c = det(I + K)^(1/3) and c - 1 = (c^3 - 1) / (c^2 + c + 1)
                const R sn = sqrt(s[0] * s[0] + s[4] * s[4] + s[8] * s[8] + R(2) * (s[1] * s[1] + s[2] * s[2] + s[5] * s[5]));
                const R alpha = sn > R(1e-12) ? R(1.5) * de / sn : R(0);
                R As[9], AsHp[9], Kx[9];
#pragma unroll
                for (int c = 0; c < 9; ++c) As[c] = alpha * s[c];
                mul3(As, Fp, AsHp);
#pragma unroll
                for (int c = 0; c < 9; ++c) Kx[c] = Fp[c] + As[c] + AsHp[c];
                const R dm = det1m(Kx);
                R cb, cm1;
                if (R(1) + dm >= R(1e-12)) {
                    cb = cbrt(R(1) + dm);
                    cm1 = dm / (cb * cb + cb + R(1));
                } else {
                    cb = cbrt(R(1e-12));
                    cm1 = cb - R(1);
                }
#pragma unroll
                for (int c = 0; c < 9; ++c) P.Fp[c][i] = (Kx[c] - ((c % 4 == 0) ? cm1 : R(0))) / cb;
            }
            R dT = M.tq * Y * de / fmax(M.rho_phys * M.cp, R(1e-12));
            if (dT > R(50)) dT = R(50);
            T += dT;
            if (T > M.melt) T = M.melt;
            P.T[i] = T;
            const R tri = (tr / R(3)) / fmax(q, R(1e-12));
            const R ef = (M.D1 + M.D2 * exp(M.D3 * tri)) * (R(1) + M.D4 * log(fmax(edot, R(1)))) * (R(1) + M.D5 * ts);
            if (ef > R(1e-12)) {
                R D = P.dmg[i] + de / ef;
                P.dmg[i] = D;
                if (D > R(0.7) * M.dth) {
                    R dg = (M.dth - D) / (M.dth * R(0.3));
                    dg = fmax(R(0), fmin(R(1), dg));
#pragma unroll
                    for (int c = 0; c < 9; ++c) sig[c] *= dg;
                    const R vm = sqrt(R(1.5) * (s[0] * s[0] + s[4] * s[4] + s[8] * s[8] +
                                                R(2) * (s[1] * s[1] + s[2] * s[2] + s[5] * s[5])));
                    P.sres[i] = fmax(P.sres[i], vm);
                }
                if (D >= M.dth) {
                    P.status[i] = kChip;
                    skip = true;
                }
            }
        }
    }
    P.s[0][i] = sig[0]; P.s[1][i] = sig[4]; P.s[2][i] = sig[8];
    P.s[3][i] = sig[1]; P.s[4][i] = sig[2]; P.s[5][i] = sig[5];
}

// The value of lane l - D of the same 16-lane row (DPP row_shr:D), zero for
// the row's first D lanes.
template <int D> __device__ __forceinline__ float row_shr(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x110 + D, 0xf, 0xf,
                                                                 true));
}
template <int D> __device__ __forceinline__ double row_shr(double v) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(u & 0xffffffffull), 0x110 + D, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), 0x110 + D, 0xf, 0xf, true);
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned long long)(unsigned)lo);
}

// One step of the segmented row sum over a particle's 7 channel values:
// ch[q] += mk * (ch[q] of lane l - D in the same 16-lane row), zero from
// before the row's start.  FP32: one v_fmac_f32 per channel reading its
// first operand through DPP row_shr:D (the compiler's form is a DPP move,
// the multiply-add and a wait state: three issue slots); the s_nop ahead of
// the seven covers the wait states a DPP read needs after a VALU write of
// its source and after an EXEC write, whatever the compiler placed before.
// The same fused multiply-adds as the compiler's form, so the same sums.
template <int D, typename R> __device__ __forceinline__ void seg_step(R (&ch)[7], R mk) {
#pragma unroll
    for (int q = 0; q < 7; ++q) ch[q] = fma(row_shr<D>(ch[q]), mk, ch[q]);
}
#define MPMS_SEG_STEP_F32(D)                                                                                   \
    template <> __device__ __forceinline__ void seg_step<D, float>(float (&ch)[7], float mk) {                  \
        asm("s_nop 4\n\t"                                                                                       \
            "v_fmac_f32_dpp %0, %0, %7 row_shr:" #D " row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"               \
            "v_fmac_f32_dpp %1, %1, %7 row_shr:" #D " row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"               \
            "v_fmac_f32_dpp %2, %2, %7 row_shr:" #D " row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"               \
            "v_fmac_f32_dpp %3, %3, %7 row_shr:" #D " row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"               \
            "v_fmac_f32_dpp %4, %4, %7 row_shr:" #D " row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"               \
            "v_fmac_f32_dpp %5, %5, %7 row_shr:" #D " row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"               \
            "v_fmac_f32_dpp %6, %6, %7 row_shr:" #D " row_mask:0xf bank_mask:0xf bound_ctrl:1"                     \
            : "+v"(ch[0]), "+v"(ch[1]), "+v"(ch[2]), "+v"(ch[3]), "+v"(ch[4]), "+v"(ch[5]), "+v"(ch[6])         \
            : "v"(mk));                                                                                         \
    }
MPMS_SEG_STEP_F32(1)
MPMS_SEG_STEP_F32(2)
MPMS_SEG_STEP_F32(4)
MPMS_SEG_STEP_F32(8)
#undef MPMS_SEG_STEP_F32

// Eight channel values as one record (two 16-byte or 32-byte accesses).
__device__ __forceinline__ void store_record(float *p, const float v[8]) {
    reinterpret_cast<float4 *>(p)[0] = make_float4(v[0], v[1], v[2], v[3]);
    reinterpret_cast<float4 *>(p)[1] = make_float4(v[4], v[5], v[6], v[7]);
}
__device__ __forceinline__ void store_record(double *p, const double v[8]) {
#pragma unroll
    for (int h = 0; h < 8; h += 2) reinterpret_cast<double2 *>(p)[h / 2] = make_double2(v[h], v[h + 1]);
}
__device__ __forceinline__ void load_record(const float *p, float v[8]) {
    const float4 a = reinterpret_cast<const float4 *>(p)[0], b = reinterpret_cast<const float4 *>(p)[1];
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w; v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
}
__device__ __forceinline__ void load_record(const double *p, double v[8]) {
#pragma unroll
    for (int h = 0; h < 8; h += 2) {
        const double2 a = reinterpret_cast<const double2 *>(p)[h / 2];
        v[h] = a.x; v[h + 1] = a.y;
    }
}

// Base-node tiles of the P2G: TX x TY x TZ base nodes of the grid sub-box,
// the particles of a tile contiguous after each fit's sort (key: tile, then
// the base node inside it, x fastest), their stencils inside a window of
// (TX + 2) x (TY + 2) x (TZ + 2) nodes.
constexpr int kTX = 8, kTY = 8, kTZ = 4;
constexpr int kWX = kTX + 2, kWY = kTY + 2, kWZ = kTZ + 2, kWN = kWX * kWY * kWZ;
constexpr int kRounds = 3;   // window-add rounds of a wave's heads sharing a base-node column

// p2gKernel (MPMSolver.cu:124-412) with the LOD classification of the step
// (lodClassifyKernel :655-676) folded in: one workgroup per base-node tile,
// its particles [tile_start[t], tile_start[t + 1]) one per thread in chunks
// of the workgroup.  Consecutive lanes mostly share a base node and hence
// all 27 target nodes: each of a node's 7 channel values is summed over
// each run of equal base nodes inside a 16-lane row (a segmented prefix sum
// by DPP row shifts) and the run's last lane in the row adds the sums to its
// wave's LDS window of the tile with plain read-add-writes, a column of
// three nodes at a time, heads sharing a base-node column in successive
// rounds (a particle whose base node has left the tile since the sort adds
// to the device grid directly).  The four windows'
// sum is stored whole to tpart[t] (kWN records of 8 channels, the last
// unused), which k_grid adds up: no device atomics for the window, whose
// nodes are mostly shared with neighbouring tiles (the device atomics were
// 85 % of a per-particle form's time, and a window flush by atomics still a
// third of this kernel's).
template <typename R>
__global__ void __launch_bounds__(kBlock) k_p2g(Particles<R> P, Grid<R> g, Mat<R> M, const StepArgs *__restrict__ sap,
                                                const int32_t *__restrict__ tile_start, R *__restrict__ tpart) {
    constexpr int kWaves = kBlock / 64;
    __shared__ R s_w[kWaves * 7 * kWN];
    __shared__ int32_t s_own[kWaves * kTX * kTY];   // per wave: the head owning a base-node column this round
    const StepArgs sa = *sap;
    const int t = blockIdx.x;
    const int ts = tile_start[t], te = tile_start[t + 1];
    if (ts >= te) return;   // uniform over the workgroup
    const int ntx = (g.nb[0] + kTX - 1) / kTX, nty = (g.nb[1] + kTY - 1) / kTY;
    // the tile's first base node in lattice coordinates
    const int tx0 = g.lo[0] + (t % ntx) * kTX, ty0 = g.lo[1] + ((t / ntx) % nty) * kTY,
              tz0 = g.lo[2] + (t / (ntx * nty)) * kTZ;
    for (int w = threadIdx.x; w < kWaves * 7 * kWN; w += kBlock) s_w[w] = R(0);
    __syncthreads();
    const int lane = threadIdx.x & 63;
    R *const s_wave = s_w + (threadIdx.x / 64) * (7 * kWN);
    int missed = 0;
    for (int c0 = ts; c0 < te; c0 += kBlock) {   // uniform over the workgroup
    const int i = c0 + threadIdx.x;
    bool live = i < te;
    int32_t st = live ? P.status[i] : kInactive;
    live = live && st != kInactive;
    double x = 0, y = 0, z = 0;
    if (live) {
        x = P.x[i]; y = P.y[i]; z = P.z[i];
        int32_t lod = P.lod[i];
        if (sa.lod_enabled && (st == kActive || st == kChip)) {
            const double dx = x - sa.tool[0], dy = y - sa.tool[1], dz = z - sa.tool[2];
            const double d = sqrt(dx * dx + dy * dy + dz * dz);
            lod = d <= sa.lod_active ? kLodActive : (d <= sa.lod_near ? kLodNear : kLodFar);
            P.lod[i] = lod;
        }
        const int cx = (int)floor((x - g.ox) * g.inv_dx), cy = (int)floor((y - g.oy) * g.inv_dx),
                  cz = (int)floor((z - g.oz) * g.inv_dx);
        if (cx < 1 || cx >= g.dims[0] - 2 || cy < 1 || cy >= g.dims[1] - 2 || cz < 1 || cz >= g.dims[2] - 2) {
            P.status[i] = kInactive;
            live = false;
        }
    }
    const bool fixed = live && st == kFixed;
    bool skip = true;
    R sig[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) sig[c] = R(0);
    R J = R(1);
    Axis<R> ax, ay, az;
    R m = 0, VJ = 0, vx = 0, vy = 0, vz = 0, efx = 0, efy = 0, efz = 0, C[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) C[c] = R(0);
    int key = -1 - lane;   // a run of its own unless live
    if (live) {
        skip = P.lod[i] == kLodFar || fixed || st == kChip;
        if (!skip && st == kActive) constitutive<R>(P, i, M, sig, J, skip, sa.islands);
        ax = axis_weights<R>(x - g.ox, g.inv_dx);
        ay = axis_weights<R>(y - g.oy, g.inv_dx);
        az = axis_weights<R>(z - g.oz, g.inv_dx);
        m = P.m[i];
        VJ = P.V[i] * J;
        vx = P.v[0][i]; vy = P.v[1][i]; vz = P.v[2][i];
        efx = P.ef[0][i]; efy = P.ef[1][i]; efz = P.ef[2][i];
#pragma unroll
        for (int c = 0; c < 9; ++c) C[c] = P.C[c][i];
        key = (az.b * g.dims[1] + ay.b) * g.dims[0] + ax.b;
    } else {
#pragma unroll
        for (int c = 0; c < 3; ++c) { ax.w[c] = ay.w[c] = az.w[c] = R(0); ax.dw[c] = ay.dw[c] = az.dw[c] = R(0); }
        ax.b = ay.b = az.b = 0;
    }
    // the run structure: a run starts where the base node changes and ends
    // before the next start; same[k] when lane + 2^k is in this lane's run
    // (contiguity matters: an equal key after a different one is another
    // run, with its own head)
    const int prev = __shfl_up(key, 1, 64);
    const bool starts = lane == 0 || prev != key;
    const unsigned long long starts_mask = __ballot(starts);
    const unsigned long long later = lane == 63 ? 0ull : starts_mask & (~0ull << (lane + 1));
    const int run_end = later ? (int)__builtin_ctzll(later) : 64;
    const unsigned long long upto = starts_mask & (lane == 63 ? ~0ull : ((2ull << lane) - 1ull));
    const int run_start = 63 - __builtin_clzll(upto);
    // the run's part in this lane's 16-lane row: [seg_start, lane] summed by
    // DPP row shifts of 1, 2, 4 and 8 lanes (mk[k] = 1 when lane - 2^k is in
    // it); its last lane holds the part's sum and is its head
    const int seg_start = max(run_start, lane & ~15);
    R mk[4];
    bool need[4];   // need[k]: some lane of the wave adds at distance 2^k (wave-uniform)
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const bool same = lane - (1 << k) >= seg_start;
        mk[k] = same ? R(1) : R(0);
        need[k] = __ballot(same) != 0ull;
    }
    const bool head = live && (lane + 1 == run_end || (lane & 15) == 15);
    const R pvx = fixed ? R(0) : vx, pvy = fixed ? R(0) : vy, pvz = fixed ? R(0) : vz;
    const R mm = fixed ? R(0) : m;   // FIXED_BOUNDARY particles add no momentum
    // the run's base node in the tile: its 27 nodes are in the window
    const int lbx = ax.b - tx0, lby = ay.b - ty0, lbz = az.b - tz0;
    // A head adds its 27 nodes to the window a column (x, y offsets a, b) at
    // a time: the three nodes' reads, then their writes.  Heads whose base
    // nodes share the column (x, y) (runs split by a row boundary, a particle
    // that crossed a base-node plane since the sort, runs of two z layers)
    // would then race on a node, so each head claims its base node's column
    // in the wave's owner table, one winner per column and round, and the
    // rounds add one after another (a wave's LDS operations complete in
    // order); heads of one round target distinct nodes at every offset.  A
    // head with no round left, or whose base node has left the tile, adds
    // to the device grid.
    const bool inwin = (unsigned)lbx < (unsigned)kTX && (unsigned)lby < (unsigned)kTY && (unsigned)lbz < (unsigned)kTZ;
    int myround = head && inwin ? -1 : kRounds;
    {
        volatile int32_t *own = s_own + (threadIdx.x / 64) * (kTX * kTY) + (lby * kTX + lbx);
#pragma unroll
        for (int r = 0; r < kRounds; ++r) {
            if (myround == -1) *own = lane;
            if (myround == -1 && *own == lane) myround = r;
        }
    }
    const bool lds_ok = myround >= 0 && myround < kRounds;
    const int nrounds = __ballot(myround == 2) ? 3 : (__ballot(myround == 1) ? 2 : (__ballot(myround == 0) ? 1 : 0));
    static_assert(kRounds == 3, "round count");
    // particles whose run goes to the device grid (the host re-sorts when
    // they grow)
    if (head && !lds_ok) atomicAdd(sa.islands + 2, (unsigned long long)(lane - seg_start + 1));
    if (head) {
        // the run's stencil visits outside the sub-box: 27 less the product
        // of the stencil's in-box node counts per axis
        auto inbox = [&](int b0, int ax_) {
            int k = 0;
#pragma unroll
            for (int d = 0; d < 3; ++d) k += (unsigned)(b0 + d - g.lo[ax_]) < (unsigned)g.nb[ax_] ? 1 : 0;
            return k;
        };
        missed += 27 - inbox(ax.b, 0) * inbox(ay.b, 1) * inbox(az.b, 2);
    }
    R *const wp = s_wave + (lbz * kWY + lby) * kWX + lbx;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            R ch[3][7];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const int gx = ax.b + a, gy = ay.b + b, gz = az.b + c;
                const R w = ax.w[a] * ay.w[b] * az.w[c];
                const R wgx = ax.dw[a] * ay.w[b] * az.w[c];
                const R wgy = ax.w[a] * ay.dw[b] * az.w[c];
                const R wgz = ax.w[a] * ay.w[b] * az.dw[c];
                const R dpx = (R)((gx * g.dx + g.ox) - x), dpy = (R)((gy * g.dx + g.oy) - y),
                        dpz = (R)((gz * g.dx + g.oz) - z);
                const R wm = w * mm;
                R *const h = ch[c];
                h[0] = w * m;
                h[1] = wm * (pvx + C[0] * dpx + C[1] * dpy + C[2] * dpz);
                h[2] = wm * (pvy + C[3] * dpx + C[4] * dpy + C[5] * dpz);
                h[3] = wm * (pvz + C[6] * dpx + C[7] * dpy + C[8] * dpz);
                h[4] = w * efx; h[5] = w * efy; h[6] = w * efz;
                if (!skip) {
                    h[4] += -VJ * (sig[0] * wgx + sig[1] * wgy + sig[2] * wgz);
                    h[5] += -VJ * (sig[3] * wgx + sig[4] * wgy + sig[5] * wgz);
                    h[6] += -VJ * (sig[6] * wgx + sig[7] * wgy + sig[8] * wgz);
                }
                if (need[0]) seg_step<1, R>(ch[c], mk[0]);
                if (need[1]) seg_step<2, R>(ch[c], mk[1]);
                if (need[2]) seg_step<4, R>(ch[c], mk[2]);
                if (need[3]) seg_step<8, R>(ch[c], mk[3]);
            }
            // the column's three nodes into the window (a node outside the
            // sub-box too: k_grid reads no window entry outside it)
            if (lds_ok) {
                R *const q = wp + b * kWX + a;
#pragma unroll
                for (int r = 0; r < kRounds; ++r) {
                    if (r < nrounds && myround == r) {
                        R old[3][7];
#pragma unroll
                        for (int c = 0; c < 3; ++c)
#pragma unroll
                            for (int k = 0; k < 7; ++k) old[c][k] = q[c * (kWX * kWY) + k * kWN];
#pragma unroll
                        for (int c = 0; c < 3; ++c)
#pragma unroll
                            for (int k = 0; k < 7; ++k) q[c * (kWX * kWY) + k * kWN] = old[c][k] + ch[c][k];
                    }
                }
            }
            if (head && !lds_ok) {
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const int64_t n = box_index(g, ax.b + a, ay.b + b, az.b + c);
                    if (n >= 0) {
                        const R *const h = ch[c];
                        unsafeAtomicAdd(g.mass + n, h[0]);
                        if (h[1] != R(0) || h[2] != R(0) || h[3] != R(0)) {
                            unsafeAtomicAdd(g.p[0] + n, h[1]);
                            unsafeAtomicAdd(g.p[1] + n, h[2]);
                            unsafeAtomicAdd(g.p[2] + n, h[3]);
                        }
                        unsafeAtomicAdd(g.f[0] + n, h[4]);
                        unsafeAtomicAdd(g.f[1] + n, h[5]);
                        unsafeAtomicAdd(g.f[2] + n, h[6]);
                    }
                }
            }
        }
    }
    __syncthreads();
    // the window, its waves' parts summed in wave order (entries of nodes
    // outside the sub-box are never read: their visits were counted as
    // missed)
    R *const dst = tpart + (int64_t)t * (kWN * 8);
    for (int w = threadIdx.x; w < kWN; w += kBlock) {
        R v[8];
#pragma unroll
        for (int h = 0; h < 7; ++h) {
            v[h] = s_w[h * kWN + w];
#pragma unroll
            for (int k = 1; k < kWaves; ++k) v[h] += s_w[(k * 7 + h) * kWN + w];
        }
        v[7] = R(0);
        store_record(dst + 8 * w, v);
    }
    if (missed) atomicAdd(g.overflow, (unsigned long long)missed);
}

// Adds the windows of the non-empty tiles whose window holds lattice node
// (gx, gy, gz) (box coordinates i, j, k) to c[0..7): tile (tx, ty, tz) holds
// nodes [kTX tx, kTX tx + kWX) x ... of the box, so at most two tiles per
// axis, in a fixed order.
template <typename R>
__device__ __forceinline__ void gather_windows(const Grid<R> &g, const R *__restrict__ tpart,
                                               const int32_t *__restrict__ tile_start, int i, int j, int k, R c[7]) {
    const int ntx = (g.nb[0] + kTX - 1) / kTX, nty = (g.nb[1] + kTY - 1) / kTY;
    const int x1 = i / kTX, y1 = j / kTY, z1 = k / kTZ;
    const int nx = (i % kTX < kWX - kTX && x1 > 0) ? 2 : 1, ny = (j % kTY < kWY - kTY && y1 > 0) ? 2 : 1,
              nz = (k % kTZ < kWZ - kTZ && z1 > 0) ? 2 : 1;
    for (int dz = 0; dz < nz; ++dz)
        for (int dy = 0; dy < ny; ++dy)
            for (int dx = 0; dx < nx; ++dx) {
                const int tx = x1 - dx, ty = y1 - dy, tz = z1 - dz;
                const int t = (tz * nty + ty) * ntx + tx;
                if (tile_start[t] == tile_start[t + 1]) continue;   // an empty tile stores no window
                const int w = ((k - kTZ * tz) * kWY + (j - kTY * ty)) * kWX + (i - kTX * tx);
                R v[8];
                load_record(tpart + ((int64_t)t * kWN + w) * 8, v);
#pragma unroll
                for (int h = 0; h < 7; ++h) c[h] += v[h];
            }
}

// The tile ranges of the sorted keys: tile_start[t] = first particle of
// tile t (keys of INACTIVE particles, 0xFFFFFFFF, sort after every tile);
// thread i in [0, n] writes the starts of the tiles after its predecessor's
// up to its own, so every t in [0, ntiles] is written once.
[[maybe_unused]] static __global__ void __launch_bounds__(kBlock) k_tile_starts(const uint32_t *__restrict__ keys, int n, int ntiles,
                                                        int32_t *__restrict__ tile_start) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i > n) return;
    auto tile = [&](int k) { return keys[k] == 0xFFFFFFFFu ? ntiles : (int)(keys[k] >> 8); };
    const int ti = i < n ? tile(i) : ntiles;
    const int tp = i > 0 ? tile(i - 1) : -1;
    for (int t = tp + 1; t <= ti; ++t) tile_start[t] = i;
}

// rasterizeToolKernel (:444-473): each tool node marks the 2 x 2 x 2 nodes
// from its rounded lattice position; the highest tool index wins.
template <typename R> __global__ void __launch_bounds__(kBlock) k_raster(Grid<R> g, const double *xyz, int ntool) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= ntool) return;
    const int gx = (int)round((xyz[3 * t] - g.ox) * g.inv_dx), gy = (int)round((xyz[3 * t + 1] - g.oy) * g.inv_dx),
              gz = (int)round((xyz[3 * t + 2] - g.oz) * g.inv_dx);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int x = gx + (k & 1), y = gy + ((k >> 1) & 1), z = gz + (k >> 2);
        if (x < 0 || x >= g.dims[0] || y < 0 || y >= g.dims[1] || z < 0 || z >= g.dims[2]) continue;
        const int64_t n = box_index(g, x, y, z);
        if (n >= 0) atomicMax(g.claim + n, t + 1);
    }
}

// updateGridKernel (:424-442), applyFixtureBCKernel (:81-111) and
// applyGridCollisionKernel (:475-492) on one node of the sub-box.
template <typename R>
__global__ void __launch_bounds__(kBlock) k_grid(Grid<R> g, const StepArgs *__restrict__ sap, const R *__restrict__ tpart,
                                                 const int32_t *__restrict__ tile_start) {
    const StepArgs sa = *sap;
    const int64_t n = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (n >= g.nn) return;
    // the P2G sums: the tile windows holding the node and the device
    // atomics of the particles outside their tile; the totals are kept for
    // the G2P (mass) and the grid read-back (mass, force)
    R c[7] = {g.mass[n], g.p[0][n], g.p[1][n], g.p[2][n], g.f[0][n], g.f[1][n], g.f[2][n]};
    {
        const int i = (int)(n % g.nb[0]), j = (int)((n / g.nb[0]) % g.nb[1]), k = (int)(n / ((int64_t)g.nb[0] * g.nb[1]));
        gather_windows(g, tpart, tile_start, i, j, k, c);
    }
    g.mass[n] = c[0];
    g.f[0][n] = c[4]; g.f[1][n] = c[5]; g.f[2][n] = c[6];
    const R m = c[0];
    R px = c[1], py = c[2], pz = c[3], vx = 0, vy = 0, vz = 0;
    const R dt = (R)sa.dt;
    const bool live = m > R(1e-12);
    if (live) {
        pz += m * (R)sa.gravity * dt;
        px += c[4] * dt;
        py += c[5] * dt;
        pz += c[6] * dt;
        vx = px / m; vy = py / m; vz = pz / m;
    }
    if (sa.fixture > 0.0) {
        const int gz = g.lo[2] + (int)(n / ((int64_t)g.nb[0] * g.nb[1]));
        const double zr = ((gz * g.dx + g.oz) - g.oz) / sa.fixture;
        if (zr <= 1.0) {
            if (gz == 0) {
                vx = vy = vz = R(0);
            } else if (zr > 0.0) {
                const double sg = sa.fixture_damping * (1.0 - zr) * (1.0 - zr);
                const R f = (R)(1.0 / (1.0 + sg * sa.dt));
                vx *= f; vy *= f; vz *= f;
            }
        }
    }
    const int32_t cl = g.claim[n];
    if (live && cl > 0) {
        vx = (R)sa.tool_vel[3 * (cl - 1)];
        vy = (R)sa.tool_vel[3 * (cl - 1) + 1];
        vz = (R)sa.tool_vel[3 * (cl - 1) + 2];
        px = m * vx; py = m * vy; pz = m * vz;
    }
    g.p[0][n] = px; g.p[1][n] = py; g.p[2][n] = pz;
    g.v[0][n] = vx; g.v[1][n] = vy; g.v[2][n] = vz;
}

// A grid node's velocity and mass, as the G2P reads them.
template <typename R> struct alignas(4 * sizeof(R)) NodeVM {
    R x, y, z, m;
};

// g2pKernel (:506-605), resetExternalForceKernel (:682-688) and, on even
// steps, MPMSolver's chipFractureAndSeparationKernel (:1007-1060, critical
// damage 1e-6, cut depth 1.5 dx) on particle i.  The 27 nodes come from the
// tile window `win` (tile origin t0) when the particle's base node is in
// the tile, else from the grid; the same values in the same order either
// way.
template <typename R>
__device__ __forceinline__ void g2p_particle(const Particles<R> &P, const Grid<R> &g, const StepArgs &sa, R dth, int i,
                                             const NodeVM<R> *win, int tx0, int ty0, int tz0) {
    const int32_t st = P.status[i];
    if (st != kInactive && st != kFixed) {
        const double x = P.x[i], y = P.y[i], z = P.z[i];
        const Axis<R> ax = axis_weights<R>(x - g.ox, g.inv_dx), ay = axis_weights<R>(y - g.oy, g.inv_dx),
                      az = axis_weights<R>(z - g.oz, g.inv_dx);
        const int lbx = ax.b - tx0, lby = ay.b - ty0, lbz = az.b - tz0;
        const bool inwin = win && (unsigned)lbx < (unsigned)kTX && (unsigned)lby < (unsigned)kTY &&
                           (unsigned)lbz < (unsigned)kTZ;
        const NodeVM<R> *wp = win + (inwin ? (lbz * kWY + lby) * kWX + lbx : 0);
        R v[3] = {0, 0, 0}, B[9], L[9];
#pragma unroll
        for (int c = 0; c < 9; ++c) B[c] = L[c] = R(0);
        int missed = 0;
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b)
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const int gx = ax.b + a, gy = ay.b + b, gz = az.b + c;
                    const int64_t n = box_index(g, gx, gy, gz);
                    if (n < 0) { ++missed; continue; }
                    R gv[3], gm;
                    if (inwin) {
                        const NodeVM<R> r = wp[(c * kWY + b) * kWX + a];
                        gv[0] = r.x; gv[1] = r.y; gv[2] = r.z; gm = r.m;
                    } else {
                        gv[0] = g.v[0][n]; gv[1] = g.v[1][n]; gv[2] = g.v[2][n]; gm = g.mass[n];
                    }
                    if (gm < R(1e-12)) continue;
                    const R w = ax.w[a] * ay.w[b] * az.w[c];
                    const R gw[3] = {ax.dw[a] * ay.w[b] * az.w[c], ax.w[a] * ay.dw[b] * az.w[c],
                                     ax.w[a] * ay.w[b] * az.dw[c]};
                    const R dp[3] = {(R)((gx * g.dx + g.ox) - x), (R)((gy * g.dx + g.oy) - y),
                                     (R)((gz * g.dx + g.oz) - z)};
#pragma unroll
                    for (int r = 0; r < 3; ++r) {
                        v[r] += w * gv[r];
#pragma unroll
                        for (int q = 0; q < 3; ++q) {
                            B[r * 3 + q] += w * gv[r] * dp[q];
                            L[r * 3 + q] += gv[r] * gw[q];
                        }
                    }
                }
        if (missed) atomicAdd(g.overflow, (unsigned long long)missed);
        P.v[0][i] = v[0]; P.v[1][i] = v[1]; P.v[2][i] = v[2];
        P.x[i] = x + (double)v[0] * sa.dt;
        P.y[i] = y + (double)v[1] * sa.dt;
        const double znew = z + (double)v[2] * sa.dt;
        P.z[i] = znew;
        if (st == kChip) {
            P.rate[i] = R(0);
        } else {
            const R Dinv = (R)(3.0 * g.inv_dx * g.inv_dx);
            const bool far = P.lod[i] == kLodFar;
#pragma unroll
            for (int c = 0; c < 9; ++c) P.C[c][i] = far ? R(0) : (B[c] * Dinv) * R(0.99);
            // F <- (I + L dt) F (:594-601) as H <- H + L dt + (L dt) H
            R H[9], Ld[9], LH[9];
#pragma unroll
            for (int c = 0; c < 9; ++c) {
                H[c] = P.F[c][i];
                Ld[c] = L[c] * (R)sa.dt;
            }
            mul3(Ld, H, LH);
#pragma unroll
            for (int c = 0; c < 9; ++c) P.F[c][i] = H[c] + Ld[c] + LH[c];
            P.rate[i] = sqrt(R(2) / R(3) * (L[0] * L[0] + L[4] * L[4] + L[8] * L[8] +
                                            R(0.5) * (L[1] + L[3]) * (L[1] + L[3]) + R(0.5) * (L[2] + L[6]) * (L[2] + L[6]) +
                                            R(0.5) * (L[5] + L[7]) * (L[5] + L[7])));
        }
        if (sa.chip_step) {
            if (st == kActive) {
                const R D = P.dmg[i];
                if (znew < sa.tool[2] - g.dx * 1.5 && D > R(1e-6)) {
                    P.status[i] = kChip;
                    const R s0 = P.s[0][i], s1 = P.s[1][i], s2 = P.s[2][i];
                    P.sres[i] = fmax(P.sres[i], sqrt(s0 * s0 + s1 * s1 + s2 * s2));
#pragma unroll
                    for (int c = 0; c < 6; ++c) P.s[c][i] = R(0);
                    P.dmg[i] = fmax(D, dth);
                }
            } else if (st == kChip && P.dmg[i] >= dth * R(1.5)) {
                const uint32_t seed = (uint32_t)P.id[i] * 2654435761u;
                const R rx = (R)((double)((seed >> 16) & 0xFF) / 255.0 - 0.5);
                const R ry = (R)((double)((seed >> 8) & 0xFF) / 255.0 - 0.5);
                const R rz = (R)((double)(seed & 0xFF) / 255.0 - 0.5);
                const R sp = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
                const R mag = sp * R(0.3) + R(0.5);
                P.v[0][i] = v[0] + rx * mag;
                P.v[1][i] = v[1] + ry * mag;
                P.v[2][i] = v[2] + rz * mag;
                P.dmg[i] = dth * R(0.5);
            }
        }
    } else if (st == kFixed) {
        P.v[0][i] = P.v[1][i] = P.v[2][i] = R(0);
    }
    P.ef[0][i] = P.ef[1][i] = P.ef[2][i] = R(0);
}

// One workgroup per base-node tile (the P2G's tiles): the tile's window of
// node velocities and masses is read into LDS once, then its particles one
// per thread in chunks.  Blocks from ntiles on take the particles after
// every tile (INACTIVE at the sort: their external force is reset).
template <typename R>
__global__ void __launch_bounds__(kBlock) k_g2p(Particles<R> P, Grid<R> g, const StepArgs *__restrict__ sap, R dth,
                                                const int32_t *__restrict__ tile_start, int ntiles, int tail_blocks) {
    const StepArgs sa = *sap;
    const int t = blockIdx.x;
    if (t >= ntiles) {   // uniform over the workgroup
        for (int i = tile_start[ntiles] + (t - ntiles) * kBlock + (int)threadIdx.x; i < P.n; i += tail_blocks * kBlock)
            g2p_particle<R>(P, g, sa, dth, i, nullptr, 0, 0, 0);
        return;
    }
    __shared__ NodeVM<R> s_n[kWN];
    const int ts = tile_start[t], te = tile_start[t + 1];
    if (ts >= te) return;   // uniform over the workgroup
    const int ntx = (g.nb[0] + kTX - 1) / kTX, nty = (g.nb[1] + kTY - 1) / kTY;
    const int tx0 = g.lo[0] + (t % ntx) * kTX, ty0 = g.lo[1] + ((t / ntx) % nty) * kTY,
              tz0 = g.lo[2] + (t / (ntx * nty)) * kTZ;
    for (int w = threadIdx.x; w < kWN; w += kBlock) {
        const int64_t n = box_index(g, tx0 + w % kWX, ty0 + (w / kWX) % kWY, tz0 + w / (kWX * kWY));
        NodeVM<R> r;
        r.x = r.y = r.z = r.m = R(0);
        if (n >= 0) { r.x = g.v[0][n]; r.y = g.v[1][n]; r.z = g.v[2][n]; r.m = g.mass[n]; }
        s_n[w] = r;
    }
    __syncthreads();
    for (int c0 = ts; c0 < te; c0 += kBlock) {
        const int i = c0 + (int)threadIdx.x;
        if (i < te) g2p_particle<R>(P, g, sa, dth, i, s_n, tx0, ty0, tz0);
    }
}

// applyExternalForceKernel (:607-627)
template <typename R>
__global__ void __launch_bounds__(kBlock) k_ext_force(Particles<R> P, double cx, double cy, double cz, double r,
                                                        double fx, double fy, double fz) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= P.n) return;
    const int32_t st = P.status[i];
    if (st != kActive && st != kChip) return;
    const double dx = P.x[i] - cx, dy = P.y[i] - cy, dz = P.z[i] - cz;
    const double d = sqrt(dx * dx + dy * dy + dz * dz);
    if (d < r) {
        const double w = 1.0 - d / r;
        P.ef[0][i] = (R)(fx * w); P.ef[1][i] = (R)(fy * w); P.ef[2][i] = (R)(fz * w);
    }
}

// Block partials of the AABBs (non-INACTIVE: the grid fit; ACTIVE and CHIP:
// getBounds) and of the metrics (computeParticleStatsKernel :726-785):
// 16 doubles per block, min as -max.
constexpr int kStats = 16;
template <typename R> __global__ void __launch_bounds__(kBlock) k_stats(Particles<R> P, double *part) {
    __shared__ double sh[kStats][kBlock];
    double v[kStats];
#pragma unroll
    for (int c = 0; c < 12; ++c) v[c] = -INFINITY;
    v[12] = 0.0; v[13] = 0.0; v[14] = 0.0; v[15] = 0.0;   // max vm, max T, KE, IE
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < P.n; i += gridDim.x * kBlock) {
        const int32_t st = P.status[i];
        if (st == kInactive) continue;
        const double r[3] = {P.x[i], P.y[i], P.z[i]};
#pragma unroll
        for (int a = 0; a < 3; ++a) { v[a] = fmax(v[a], -r[a]); v[3 + a] = fmax(v[3 + a], r[a]); }
        if (st != kActive && st != kChip) continue;
#pragma unroll
        for (int a = 0; a < 3; ++a) { v[6 + a] = fmax(v[6 + a], -r[a]); v[9 + a] = fmax(v[9 + a], r[a]); }
        const double s0 = P.s[0][i], s1 = P.s[1][i], s2 = P.s[2][i], s3 = P.s[3][i], s4 = P.s[4][i], s5 = P.s[5][i];
        const double tr = s0 + s1 + s2;
        const double a0 = s0 - tr / 3.0, a1 = s1 - tr / 3.0, a2 = s2 - tr / 3.0;
        const double vm = sqrt(1.5 * (a0 * a0 + a1 * a1 + a2 * a2 + 2.0 * (s3 * s3 + s4 * s4 + s5 * s5)));
        const double m = P.m[i], vx = P.v[0][i], vy = P.v[1][i], vz = P.v[2][i];
        v[12] = fmax(v[12], vm);
        v[13] = fmax(v[13], (double)P.T[i]);
        v[14] += 0.5 * m * (vx * vx + vy * vy + vz * vz);
        v[15] += (vm * vm * m) / (6.0 * 80e9) + (double)P.V[i] * (double)P.eps[i] * vm;
    }
#pragma unroll
    for (int c = 0; c < kStats; ++c) sh[c][threadIdx.x] = v[c];
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
#pragma unroll
            for (int c = 0; c < kStats; ++c)
                sh[c][threadIdx.x] = c >= 14 ? sh[c][threadIdx.x] + sh[c][threadIdx.x + o]
                                             : fmax(sh[c][threadIdx.x], sh[c][threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x < kStats) part[blockIdx.x * kStats + threadIdx.x] = sh[threadIdx.x][0];
}

// The base node of each particle's stencil as a sort key in the sub-box
// (INACTIVE particles last).
template <typename R>
__global__ void __launch_bounds__(kBlock) k_sort_keys(Particles<R> P, Grid<R> g, uint32_t *keys, int32_t *idx) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= P.n) return;
    uint32_t k = 0xFFFFFFFFu;
    if (P.status[i] != kInactive) {
        const int bx = (int)floor((P.x[i] - g.ox) * g.inv_dx - 0.5) - g.lo[0];
        const int by = (int)floor((P.y[i] - g.oy) * g.inv_dx - 0.5) - g.lo[1];
        const int bz = (int)floor((P.z[i] - g.oz) * g.inv_dx - 0.5) - g.lo[2];
        const int cx = min(max(bx, 0), g.nb[0] - 1), cy = min(max(by, 0), g.nb[1] - 1), cz = min(max(bz, 0), g.nb[2] - 1);
        const int ntx = (g.nb[0] + kTX - 1) / kTX, nty = (g.nb[1] + kTY - 1) / kTY;
        const int tile = ((cz / kTZ) * nty + cy / kTY) * ntx + cx / kTX;
        k = (uint32_t)tile * (uint32_t)(kTX * kTY * kTZ) + (uint32_t)(((cz % kTZ) * kTY + cy % kTY) * kTX + cx % kTX);
    }
    keys[i] = k;
    idx[i] = i;
}

// Every particle field permuted in one launch: field k of particle i takes
// the value of particle perm[i]; fields [0, NW) are 8 bytes wide, [NW, NF)
// 4 bytes.  A thread issues its particle's NF loads before its stores.
template <int NF> struct FieldPtrs {
    const void *src[NF];
    void *dst[NF];
};
template <int NF, int NW>
__global__ void __launch_bounds__(kBlock) k_gather_fields(FieldPtrs<NF> G, const int32_t *__restrict__ perm, int n) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const int j = perm[i];
    double w[NW];
    int32_t v[NF - NW];
#pragma unroll
    for (int k = 0; k < NW; ++k) w[k] = static_cast<const double *>(G.src[k])[j];
#pragma unroll
    for (int k = NW; k < NF; ++k) v[k - NW] = static_cast<const int32_t *>(G.src[k])[j];
#pragma unroll
    for (int k = 0; k < NW; ++k) static_cast<double *>(G.dst[k])[i] = w[k];
#pragma unroll
    for (int k = NW; k < NF; ++k) static_cast<int32_t *>(G.dst[k])[i] = v[k - NW];
}

[[maybe_unused]] static __global__ void __launch_bounds__(kBlock) k_zero_words(uint32_t *p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) p[i] = 0u;
}

}  // namespace mpms
