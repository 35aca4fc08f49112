// mpmw_tile.hpp -- the particle pass of the MPM step over cell tiles.
//
// One workgroup per tile of TX x TY x TZ cells; thread t owns cell t.  The
// tile's node window (WX x WY x WZ nodes) is staged in LDS: the gather record
// (v, T) for G2P and eight accumulator channels for P2G.  A thread keeps the
// 8 x 8 P2G corner sums of its cell's particles in registers; after the
// particle loop the sums are combined across the wave with lane shuffles into
// node values and added to the LDS window layer by layer, then the whole
// window is stored, with plain coalesced stores, into the tile's own slot of
// a per-tile window buffer; the grid update sums the windows holding each
// node.  No float atomics are used on this path (LDS float atomics serialise
// on gfx950; device float atomics are bound by the L2 atomic rate).  A
// particle that has left its cell since the last binning is scattered with
// atomics (LDS inside the window, the node accumulators outside it), or, in
// the deterministic mode, counted for re-binning.
//
// Reference kernels: mpmP2GKernel (MPMWorkpieceSolver.cu:98-134),
// mpmG2PKernel (:264-400), chipFractureAndSeparationKernel (:667-725).
#pragma once

#include "mpmw_kernels.hpp"
#include "mpmw_sort.hpp"

namespace mpmw {

struct G2PFlags {
    int chip;      // run the chip separation / fracture pass (even step)
    int stats;     // write AABB + metrics block partials
};

// ------------------------------------------------------- particle state ---

// The node sums of the G2P gather (mpmG2PKernel :276-293): velocity and
// temperature weighted by w, the weight sum and the velocity gradient.
template <typename R> struct GatherSums {
    R v0, v1, v2, temp, wsum;
    R L[9];
};

// Gather over the 8 corners from device memory; corners outside the grid,
// with w <= 0 or with an empty node (NaN record) are skipped (:280-286).
template <typename R>
__device__ __forceinline__ void gather_global(const GridView<R> &g, gp<const vec4_t<R>> node, const Frac<R> &f,
                                              GatherSums<R> &q) {
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        int64_t idx;
        R w, gx, gy, gz;
        if (!corner(g, f, n, idx, w, gx, gy, gz)) continue;
        const vec4_t<R> nd = node[idx];
        if (nd.w != nd.w) continue;
        q.v0 += nd.x * w; q.v1 += nd.y * w; q.v2 += nd.z * w;
        q.temp += nd.w * w;
        q.wsum += w;
        q.L[0] += nd.x * gx; q.L[1] += nd.x * gy; q.L[2] += nd.x * gz;
        q.L[3] += nd.y * gx; q.L[4] += nd.y * gy; q.L[5] += nd.y * gz;
        q.L[6] += nd.z * gx; q.L[7] += nd.z * gy; q.L[8] += nd.z * gz;
    }
}

__device__ __forceinline__ int wnode(int x, int y, int z) { return (z * WY + y) * WX + x; }

// The LDS window holds node records with the empty node (mass <= 1e-18,
// or outside the grid) as (0, 0, 0, -1): a live node's temperature is
// never negative, so liveness is T >= 0 and an empty node adds zeros to
// every gather sum without a select per component.
template <typename R> __device__ __forceinline__ vec4_t<R> window_record(const vec4_t<R> &nd) {
    vec4_t<R> r = nd;
    if (!(nd.w == nd.w)) { r.x = r.y = r.z = R(0); r.w = R(-1); }
    return r;
}

template <typename R> __device__ __forceinline__ vec4_t<R> empty_record() {
    vec4_t<R> r;
    r.x = r.y = r.z = R(0);
    r.w = R(-1);
    return r;
}

// Products of two axis weights: pyz[dy][dz] = wy * wz, pxz[dx][dz],
// pxy[dx][dy].  Corner (dx, dy, dz) has weight wx[dx] * pyz[dy][dz] and
// weight gradient (sx pyz, sy pxz, sz pxy) / h with s = -1 for the near
// node and +1 for the far node of each axis.
template <typename R> struct WeightPairs {
    R wx[2], yz[2][2], xz[2][2], xy[2][2];
};

template <typename R> __device__ __forceinline__ WeightPairs<R> weight_pairs(const Frac<R> &f) {
    WeightPairs<R> P;
    const R wx[2] = {f.w0[0], f.w1[0]}, wy[2] = {f.w0[1], f.w1[1]}, wz[2] = {f.w0[2], f.w1[2]};
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        P.wx[a] = wx[a];
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            P.yz[a][b] = wy[a] * wz[b];
            P.xz[a][b] = wx[a] * wz[b];
            P.xy[a][b] = wx[a] * wy[b];
        }
    }
    return P;
}

// Every corner has a positive weight (no axis fraction is zero): then no
// corner is skipped and the factored forms below hold.
template <typename R> __device__ __forceinline__ bool all_corners_live(const Frac<R> &f) {
    return f.w1[0] > R(0) && f.w1[1] > R(0) && f.w1[2] > R(0) && f.w0[0] > R(0) && f.w0[1] > R(0) &&
           f.w0[2] > R(0);
}

// The same gather from the tile's LDS window, base corner at window index
// wb: every corner is inside the window (window records as window_record
// writes them).  When every corner weight is positive the velocity
// gradient is formed along each axis from node differences,
// L[a][x] = (1/h) sum over (dy, dz) of wy wz (v_a(1, dy, dz) - v_a(0, dy, dz))
// and likewise for y and z: the same sum as sum_c v_c grad w_c, with an
// empty node's zero velocity standing for the skipped corner, but without
// the cancellation of large equal velocities (a translating body) that
// the corner-by-corner sum suffers in FP32.  Otherwise the corners with
// w <= 0 are skipped one by one (:280-286).
template <typename R>
__device__ __forceinline__ void gather_window(const vec4_t<R> *s_node, int wb, const Frac<R> &f, R inv_h,
                                              GatherSums<R> &q) {
    if (all_corners_live(f)) {
        const WeightPairs<R> P = weight_pairs(f);
        vec4_t<R> nd[8];
#pragma unroll
        for (int n = 0; n < 8; ++n) nd[n] = s_node[wb + wnode(n & 1, (n >> 1) & 1, (n >> 2) & 1)];
#pragma unroll
        for (int n = 0; n < 8; ++n) {
            const int dx = n & 1, dy = (n >> 1) & 1, dz = (n >> 2) & 1;
            const R w = P.wx[dx] * P.yz[dy][dz];
            q.v0 = fma(nd[n].x, w, q.v0); q.v1 = fma(nd[n].y, w, q.v1); q.v2 = fma(nd[n].z, w, q.v2);
            q.temp = fma(rmax(nd[n].w, R(0)), w, q.temp);
            q.wsum += nd[n].w >= R(0) ? w : R(0);
        }
        R L[9];
#pragma unroll
        for (int c = 0; c < 9; ++c) L[c] = R(0);
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                // x pairs (0, a, b) - (1, a, b); y pairs (a, 0, b) - (a, 1, b); z pairs (a, b, 0) - (a, b, 1)
                const int x0 = 2 * a + 4 * b, y0 = a + 4 * b, z0 = a + 2 * b;
                const R px = P.yz[a][b], py = P.xz[a][b], pz = P.xy[a][b];
                L[0] = fma(nd[x0 + 1].x - nd[x0].x, px, L[0]);
                L[3] = fma(nd[x0 + 1].y - nd[x0].y, px, L[3]);
                L[6] = fma(nd[x0 + 1].z - nd[x0].z, px, L[6]);
                L[1] = fma(nd[y0 + 2].x - nd[y0].x, py, L[1]);
                L[4] = fma(nd[y0 + 2].y - nd[y0].y, py, L[4]);
                L[7] = fma(nd[y0 + 2].z - nd[y0].z, py, L[7]);
                L[2] = fma(nd[z0 + 4].x - nd[z0].x, pz, L[2]);
                L[5] = fma(nd[z0 + 4].y - nd[z0].y, pz, L[5]);
                L[8] = fma(nd[z0 + 4].z - nd[z0].z, pz, L[8]);
            }
#pragma unroll
        for (int c = 0; c < 9; ++c) q.L[c] = fma(L[c], inv_h, q.L[c]);
        return;
    }
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        const int dx = n & 1, dy = (n >> 1) & 1, dz = (n >> 2) & 1;
        const R wx = dx ? f.w1[0] : f.w0[0];
        const R wy = dy ? f.w1[1] : f.w0[1];
        const R wz = dz ? f.w1[2] : f.w0[2];
        const R w = wx * wy * wz;
        const vec4_t<R> nd = s_node[wb + wnode(dx, dy, dz)];
        const bool ok = (w > R(0)) && (nd.w >= R(0));
        const R nx = ok ? nd.x : R(0), ny = ok ? nd.y : R(0), nz = ok ? nd.z : R(0), nt = ok ? nd.w : R(0);
        const R gx = (dx ? inv_h : -inv_h) * wy * wz;
        const R gy = wx * (dy ? inv_h : -inv_h) * wz;
        const R gz = wx * wy * (dz ? inv_h : -inv_h);
        q.v0 += nx * w; q.v1 += ny * w; q.v2 += nz * w;
        q.temp += nt * w;
        q.wsum += ok ? w : R(0);
        q.L[0] += nx * gx; q.L[1] += nx * gy; q.L[2] += nx * gz;
        q.L[3] += ny * gx; q.L[4] += ny * gy; q.L[5] += ny * gz;
        q.L[6] += nz * gx; q.L[7] += nz * gy; q.L[8] += nz * gz;
    }
}

// One particle through mpmG2PKernel (MPMWorkpieceSolver.cu:264-400) and, on
// chip steps, chipFractureAndSeparationKernel (:667-725).  The caller has
// loaded status and position.  Only the fields the particle's status makes
// the reference read are loaded, and only fields that changed are stored.
// gather(q) adds the node sums of the particle's corners to q.
// The per-particle inputs of the G2P pass, loaded together (a tile pass
// loads the next slot's while it computes the current one).  Stress, plastic
// strain and damage are loaded whatever the status; g2p_particle uses them
// only where the reference reads them.
template <typename R> struct G2PIn {
    double x, y, z;
    R T, eps, dmg;
    R s[6];
    int32_t status;
};

template <typename R> __device__ __forceinline__ void load_g2p_in(const Particles<R> &P, int i, G2PIn<R> &in) {
    in.status = P.status[i];
    in.x = P.x[i]; in.y = P.y[i]; in.z = P.z[i];
    in.T = P.T[i];
    in.eps = P.eps[i];
    in.dmg = P.dmg[i];
    in.s[0] = P.s0[i]; in.s[1] = P.s1[i]; in.s[2] = P.s2[i];
    in.s[3] = P.s3[i]; in.s[4] = P.s4[i]; in.s[5] = P.s5[i];
}

// Which fields a particle's G2P changed (the stores g2p_store makes).
enum : uint32_t { kStPos = 1, kStActive = 2, kStEps = 4, kStDmg = 8, kStT = 16, kStStatus = 32, kStFext = 64 };

// The G2P update of one particle without its stores (except the rare
// residual-stress update, a read-modify-write); returns the kSt* set of
// fields to store.
template <typename R, bool HAS_BACK, typename Gather>
__device__ __forceinline__ uint32_t g2p_compute(const Particles<R> &P, int i, const G2PIn<R> &in, PState<R> &p,
                                                const Material<R> &M, R dt, double dtd, bool chip_step,
                                                bool need_stress, double zcut, Gather gather) {
    const int32_t st0 = in.status;
    const bool act = st0 == kActive, fixed = st0 == kFixed;
    p.status = st0;
    p.x = in.x; p.y = in.y; p.z = in.z;
    p.T = in.T;
    p.sres = -__builtin_huge_valf();
    p.rate = R(0);
    const bool use_s = act || fixed || need_stress;
#pragma unroll
    for (int c = 0; c < 6; ++c) { p.s[c] = use_s ? in.s[c] : R(0); p.a[c] = R(0); }
    p.eps = act ? in.eps : R(0);
    if (HAS_BACK && act) {
        p.a[0] = P.a0[i]; p.a[1] = P.a1[i]; p.a[2] = P.a2[i];
        p.a[3] = P.a3[i]; p.a[4] = P.a4[i]; p.a[5] = P.a5[i];
    }
    p.dmg = (act || st0 == kChip) ? in.dmg : R(0);
    const R eps0 = p.eps, dmg0 = p.dmg, T0 = p.T;
    if (fixed) {  // :296-300
        p.vx = p.vy = p.vz = R(0);
    } else {
        GatherSums<R> q;
        q.v0 = q.v1 = q.v2 = q.temp = q.wsum = R(0);
#pragma unroll
        for (int c = 0; c < 9; ++c) q.L[c] = R(0);
        gather(q);
        p.vx = q.v0; p.vy = q.v1; p.vz = q.v2;
        p.x += (double)p.vx * dtd;
        p.y += (double)p.vy * dtd;
        p.z += (double)p.vz * dtd;
        if (q.wsum > R(1e-8)) p.T = rmin(rmax(vdiv(q.temp, q.wsum), M.ref_temp), M.melt);
        if (act) constitutive_active<R, HAS_BACK>(M, p, q.L, dt);
        chip_transition(M, p);
    }
    if (chip_step) chip_pass(M, p, P.id, i, zcut);
    uint32_t st = 0;
    st |= fixed ? 0u : kStPos;
    st |= act ? kStActive : 0u;
    st |= p.eps != eps0 ? kStEps : 0u;
    st |= p.dmg != dmg0 ? kStDmg : 0u;
    st |= p.T != T0 ? kStT : 0u;
    st |= p.status != st0 ? kStStatus : 0u;
    st |= P.fx ? kStFext : 0u;
    if (p.sres > -__builtin_huge_valf()) {
        const R old = P.sres[i];
        if (p.sres > old) P.sres[i] = p.sres;
    }
    return st;
}

// The stores of one particle's G2P: only fields that changed, velocity
// always, the external force zeroed (:399).
template <typename R, bool HAS_BACK>
__device__ __forceinline__ void g2p_store(const Particles<R> &P, int i, const PState<R> &p, uint32_t st) {
    if (st & kStPos) { P.x[i] = p.x; P.y[i] = p.y; P.z[i] = p.z; }
    P.vx[i] = p.vx; P.vy[i] = p.vy; P.vz[i] = p.vz;
    if (st & kStActive) {
        P.s0[i] = p.s[0]; P.s1[i] = p.s[1]; P.s2[i] = p.s[2];
        P.s3[i] = p.s[3]; P.s4[i] = p.s[4]; P.s5[i] = p.s[5];
        if (HAS_BACK) {
            P.a0[i] = p.a[0]; P.a1[i] = p.a[1]; P.a2[i] = p.a[2];
            P.a3[i] = p.a[3]; P.a4[i] = p.a[4]; P.a5[i] = p.a[5];
        }
        P.rate[i] = p.rate;
    }
    if (st & kStEps) P.eps[i] = p.eps;
    if (st & kStDmg) P.dmg[i] = p.dmg;
    if (st & kStT) P.T[i] = p.T;
    if (st & kStStatus) P.status[i] = p.status;
    if (st & kStFext) { P.fx[i] = R(0); P.fy[i] = R(0); P.fz[i] = R(0); }
}

// One particle through mpmG2PKernel (MPMWorkpieceSolver.cu:264-400) and, on
// chip steps, chipFractureAndSeparationKernel (:667-725), stores included.
template <typename R, bool HAS_BACK, typename Gather>
__device__ __forceinline__ void g2p_particle(const Particles<R> &P, int i, const G2PIn<R> &in, PState<R> &p,
                                             const Material<R> &M, R dt, double dtd, bool chip_step,
                                             bool need_stress, double zcut, Gather gather) {
    const uint32_t st = g2p_compute<R, HAS_BACK>(P, i, in, p, M, dt, dtd, chip_step, need_stress, zcut, gather);
    g2p_store<R, HAS_BACK>(P, i, p, st);
}

__device__ __forceinline__ void accumulate_stats(double st[kRed], double x, double y, double z, double vm, double T,
                                                 double ke) {
    st[0] = fmin(st[0], x); st[1] = fmin(st[1], y); st[2] = fmin(st[2], z);
    st[3] = fmin(st[3], -x); st[4] = fmin(st[4], -y); st[5] = fmin(st[5], -z);
    st[6] = fmax(st[6], vm);
    st[7] = fmax(st[7], T);
    st[8] += ke;
}

// The 8 channel contributions of one particle to one corner node,
// mpmP2GKernel :113-132.  Returns the number of channels written to c[]
// (5: mass, momentum, T*m; 8: plus force).
template <typename R>
__device__ __forceinline__ int p2g_corner(R mass, R vol, R vx, R vy, R vz, R T, const R s[6], R fex, R fey, R fez,
                                          bool stress_force, bool ext_only, R w, R gx, R gy, R gz, R c[8]) {
    const R m = mass * w;
    c[0] = m;
    c[1] = vx * m;
    c[2] = vy * m;
    c[3] = vz * m;
    c[7] = T * m;
    if (stress_force) {
        const R fX = -(s[0] * gx + s[3] * gy + s[4] * gz) * vol;
        const R fY = -(s[3] * gx + s[1] * gy + s[5] * gz) * vol;
        const R fZ = -(s[4] * gx + s[5] * gy + s[2] * gz) * vol;
        c[4] = fX + fex * w;
        c[5] = fY + fey * w;
        c[6] = fZ + fez * w;
        return 8;
    }
    if (ext_only) {
        c[4] = fex * w;
        c[5] = fey * w;
        c[6] = fez * w;
        return 8;
    }
    return 5;
}

// ----------------------------------------------------------- tile kernel --

// The 8 corner contributions of the particles of one cell, 8 channels each
// (mass, momentum xyz, force xyz, T*m), held in registers.  Corner c has
// offsets dx = c & 1, dy = (c >> 1) & 1, dz = (c >> 2) & 1 (the node order
// of computeWeights, :64-66).
template <typename R> struct CellAcc {
    R v[8][8];  // [channel][corner]
};

// Add one particle's P2G contributions (mpmP2GKernel :110-133) to the
// accumulators of its base cell.  Corners with w <= 0 are skipped as the
// reference skips them (:112).  With every corner weight positive the
// stress term -V sigma grad w is formed from the weight pairs with
// -V sigma / h scaled once per particle; `ext` adds the external force
// (only when the solver has one).
template <typename R>
__device__ __forceinline__ void cell_accumulate(CellAcc<R> &A, const Frac<R> &f, R inv_h, R mass, R vol, R vx, R vy,
                                                R vz, R T, const R s[6], R fex, R fey, R fez, bool stress_force,
                                                bool ext_only, bool ext) {
    const bool fe = ext && (stress_force || ext_only);
    const R ex = fe ? fex : R(0), ey = fe ? fey : R(0), ez = fe ? fez : R(0);
    if (all_corners_live(f)) {
        const WeightPairs<R> P = weight_pairs(f);
        const R sc = stress_force ? -vol * inv_h : R(0);
        const R qxx = s[0] * sc, qyy = s[1] * sc, qzz = s[2] * sc;
        const R qxy = s[3] * sc, qxz = s[4] * sc, qyz = s[5] * sc;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const int dx = c & 1, dy = (c >> 1) & 1, dz = (c >> 2) & 1;
            const R w = P.wx[dx] * P.yz[dy][dz];
            const R gx = dx ? P.yz[dy][dz] : -P.yz[dy][dz];
            const R gy = dy ? P.xz[dx][dz] : -P.xz[dx][dz];
            const R gz = dz ? P.xy[dx][dy] : -P.xy[dx][dy];
            const R m = mass * w;
            A.v[0][c] += m;
            A.v[1][c] = fma(vx, m, A.v[1][c]);
            A.v[2][c] = fma(vy, m, A.v[2][c]);
            A.v[3][c] = fma(vz, m, A.v[3][c]);
            A.v[7][c] = fma(T, m, A.v[7][c]);
            A.v[4][c] = fma(qxz, gz, fma(qxy, gy, fma(qxx, gx, A.v[4][c])));
            A.v[5][c] = fma(qyz, gz, fma(qyy, gy, fma(qxy, gx, A.v[5][c])));
            A.v[6][c] = fma(qzz, gz, fma(qyz, gy, fma(qxz, gx, A.v[6][c])));
            if (ext) {
                A.v[4][c] = fma(ex, w, A.v[4][c]);
                A.v[5][c] = fma(ey, w, A.v[5][c]);
                A.v[6][c] = fma(ez, w, A.v[6][c]);
            }
        }
        return;
    }
    // branch-free form of p2g_corner: a corner with w <= 0 adds zeros, the
    // stress term is zero unless stress_force, the external force unless
    // stress_force or ext_only
    const R volm = stress_force ? vol : R(0);
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const int dx = c & 1, dy = (c >> 1) & 1, dz = (c >> 2) & 1;
        const R wx = dx ? f.w1[0] : f.w0[0];
        const R wy = dy ? f.w1[1] : f.w0[1];
        const R wz = dz ? f.w1[2] : f.w0[2];
        const R w = wx * wy * wz;
        const bool ok = w > R(0);
        const R gx = ok ? (dx ? inv_h : -inv_h) * wy * wz : R(0);
        const R gy = ok ? wx * (dy ? inv_h : -inv_h) * wz : R(0);
        const R gz = ok ? wx * wy * (dz ? inv_h : -inv_h) : R(0);
        const R m = ok ? mass * w : R(0);
        const R we = ok ? w : R(0);
        A.v[0][c] += m;
        A.v[1][c] += vx * m;
        A.v[2][c] += vy * m;
        A.v[3][c] += vz * m;
        A.v[7][c] += T * m;
        A.v[4][c] += -(s[0] * gx + s[3] * gy + s[4] * gz) * volm + ex * we;
        A.v[5][c] += -(s[3] * gx + s[1] * gy + s[5] * gz) * volm + ey * we;
        A.v[6][c] += -(s[4] * gx + s[5] * gy + s[2] * gz) * volm + ez * we;
    }
}

// Adds the live channels of c[8] (bit ch of live) to p[ch * stride] in LDS.
// FP64 uses the LDS atomic add.  FP32 uses compare-and-swap: gfx950 runs
// ds_add_f32 one lane at a time (0.33 lane-adds per clock per CU measured,
// against 4.6 for the compare-and-swap loop and 7.8 for ds_add_f64).  Each
// pass reads the words, then swaps in old + c and repeats for the lanes
// whose word another lane changed in between; every add is one FP32 add
// onto the current sum, as the atomic add is.
__device__ __forceinline__ void lds_add8(double *p, int stride, const double c[8], unsigned live) {
#pragma unroll
    for (int ch = 0; ch < 8; ++ch)
        if (live & (1u << ch)) atomicAdd(p + ch * stride, c[ch]);
}

__device__ __forceinline__ void lds_add8(float *p, int stride, const float c[8], unsigned live) {
    unsigned *q = reinterpret_cast<unsigned *>(p);
    unsigned old[8];
#pragma unroll
    for (int ch = 0; ch < 8; ++ch)
        old[ch] = (live & (1u << ch))
                      ? __hip_atomic_load(q + ch * stride, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)
                      : 0u;
    while (live) {
        unsigned again = 0;
#pragma unroll
        for (int ch = 0; ch < 8; ++ch) {
            if (!(live & (1u << ch))) continue;
            unsigned expect = old[ch];
            const unsigned want = __float_as_uint(__uint_as_float(expect) + c[ch]);
            if (!__hip_atomic_compare_exchange_strong(q + ch * stride, &expect, want, __ATOMIC_RELAXED,
                                                      __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) {
                old[ch] = expect;
                again |= 1u << ch;
            }
        }
        live = again;
    }
}

// Scatter of one particle whose stencil is not its home cell: through LDS
// adds when the stencil lies in the tile window, else device atomics into
// the node accumulators (then returns true).
template <typename R>
__device__ __forceinline__ bool scatter_exception(const GridView<R> &g, gp<R> acc, R *s_acc, const Frac<R> &f,
                                                  int wx0, int wy0, int wz0, R mass, R vol, R vx, R vy, R vz, R T,
                                                  const R s[6], R fex, R fey, R fez, bool stress_force,
                                                  bool ext_only) {
    const int lx = f.b[0] - wx0, ly = f.b[1] - wy0, lz = f.b[2] - wz0;
    const bool in = s_acc && (unsigned)lx < (unsigned)TX && (unsigned)ly < (unsigned)TY && (unsigned)lz < (unsigned)TZ;
    const int wb = wnode(lx, ly, lz);
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        int64_t idx;
        R w, gx, gy, gz;
        if (!corner(g, f, n, idx, w, gx, gy, gz)) continue;
        R c[8];
        const int nch = p2g_corner(mass, vol, vx, vy, vz, T, s, fex, fey, fez, stress_force, ext_only, w, gx, gy,
                                   gz, c);
        if (in) {
            lds_add8(s_acc + wb + wnode(n & 1, (n >> 1) & 1, (n >> 2) & 1), WN, c, nch == 8 ? 0xffu : 0x8fu);
        } else {
            R *a = acc + idx;
            unsafeAtomicAdd(a, c[0]);
            unsafeAtomicAdd(a + g.nn, c[1]);
            unsafeAtomicAdd(a + 2 * g.nn, c[2]);
            unsafeAtomicAdd(a + 3 * g.nn, c[3]);
            unsafeAtomicAdd(a + 7 * g.nn, c[7]);
            if (nch == 8) {
                unsafeAtomicAdd(a + 4 * g.nn, c[4]);
                unsafeAtomicAdd(a + 5 * g.nn, c[5]);
                unsafeAtomicAdd(a + 6 * g.nn, c[6]);
            }
        }
    }
    return !in;
}

// The fields a P2G-only pass reads (mpmP2GKernel :103-133): velocity,
// temperature, the stress of ACTIVE / FIXED_BOUNDARY particles and the
// external force.
template <typename R>
__device__ __forceinline__ void load_p2g_state(const Particles<R> &P, int i, PState<R> &p, R &fex, R &fey, R &fez) {
    p.vx = P.vx[i]; p.vy = P.vy[i]; p.vz = P.vz[i];
    p.T = P.T[i];
    if (p.status == kActive || p.status == kFixed) {
        p.s[0] = P.s0[i]; p.s[1] = P.s1[i]; p.s[2] = P.s2[i];
        p.s[3] = P.s3[i]; p.s[4] = P.s4[i]; p.s[5] = P.s5[i];
    } else {
#pragma unroll
        for (int c = 0; c < 6; ++c) p.s[c] = R(0);
    }
    if (P.fx) { fex = P.fx[i]; fey = P.fy[i]; fez = P.fz[i]; }
}

template <typename R> __device__ __forceinline__ R shfl_from(R v, int src) { return __shfl(v, src, kWave); }

// Combine the cell accumulators of a wave (an 8 x 8 layer of cells) into
// node values of node layers z and z+1 and add them to the LDS window.
// Node (x, y) of a layer = corner (0,0) of cell (x,y) + (1,0) of (x-1,y) +
// (0,1) of (x,y-1) + (1,1) of (x-1,y-1); the far edge nodes x = 8 / y = 8
// are formed by the lanes of the last column / row.  Fixed summation order.
template <typename R>
__device__ __forceinline__ void combine_layer(const CellAcc<R> &A, R *s_acc, int hx, int hy, int hz, int dz) {
    const int lane = threadIdx.x % kWave;
#pragma unroll
    for (int ch = 0; ch < 8; ++ch) {
        const R c00 = A.v[ch][4 * dz + 0], c10 = A.v[ch][4 * dz + 1];
        const R c01 = A.v[ch][4 * dz + 2], c11 = A.v[ch][4 * dz + 3];
        const R l1 = shfl_from(c10, lane - 1);   // (x-1, y) corner (1,0)
        const R l8 = shfl_from(c01, lane - 8);   // (x, y-1) corner (0,1)
        const R l9 = shfl_from(c11, lane - 9);   // (x-1, y-1) corner (1,1)
        const R e8 = shfl_from(c11, lane - 8);   // for node (8, y): (7, y-1) corner (1,1)
        const R e1 = shfl_from(c11, lane - 1);   // for node (x, 8): (x-1, 7) corner (1,1)
        R v = c00;
        if (hx >= 1) v += l1;
        if (hy >= 1) v += l8;
        if (hx >= 1 && hy >= 1) v += l9;
        R *a = s_acc + ch * WN;
        a[wnode(hx, hy, hz + dz)] += v;
        if (hx == TX - 1) {
            R ve = c10;
            if (hy >= 1) ve += e8;
            a[wnode(TX, hy, hz + dz)] += ve;
        }
        if (hy == TY - 1) {
            R ve = c01;
            if (hx >= 1) ve += e1;
            a[wnode(hx, TY, hz + dz)] += ve;
        }
        if (hx == TX - 1 && hy == TY - 1) a[wnode(TX, TY, hz + dz)] += c11;
    }
}

// One workgroup per tile; thread t owns cell t of the tile (x fastest) and
// walks that cell's particles slot by slot (the memory order is slot-major,
// so a full tile reads contiguously).  DO_G2P: gather, constitutive update
// and chip pass.  DO_P2G: scatter into the next step's accumulators.
// The window goes to tpart[t] (WN node records of 8 channels).  counters[2] is set
// when a particle was scattered into the node accumulators.  DET
// (deterministic P2G): a particle outside its binned cell is only counted in
// counters[1]; the caller re-bins and repeats the P2G when that count is not
// zero.  Every sum then has a fixed order: per cell in slot order, across a
// wave by the fixed shuffle tree, across tiles in tile order (k_grid_update_batch).
// Waves per SIMD the tile pass is compiled for (register budget 512 / waves).
// FP32: 2 (a budget of 3 waves spills to scratch and measured 18 % slower);
// FP64 keeps the compiler's choice.
#ifndef MPMW_TILE_WAVES_F32
#define MPMW_TILE_WAVES_F32 2
#endif
template <typename R> struct TileWaves {
    static constexpr int value = sizeof(R) == 4 ? MPMW_TILE_WAVES_F32 : 1;
};

// Slots of the per-wave count table of an uneven tile (tile_block).
constexpr int kSlotTab = 32;
// Entries of a tile's queue of particles scattered after its slot loop.
constexpr int kExcQueue = 512;

// Zeros as the window of a tile without particles: the grid update adds
// every window that holds a node without testing the tile for particles.
template <typename R, int N> __device__ __forceinline__ void zero_window(R *dst, int tid) {
    vec4_t<R> *d = reinterpret_cast<vec4_t<R> *>(dst);
    vec4_t<R> z;
    z.x = z.y = z.z = z.w = R(0);
    for (int w = tid; w < 2 * N; w += kBlock) d[w] = z;
}

// Stores a tile's LDS window (N nodes, channel stride N) into its slot of
// the window buffer node-major: node w's eight channel sums are the 32-byte
// record dst[8 w .. 8 w + 7], which the grid update reads with two vector
// loads.
template <typename R, int N> __device__ __forceinline__ void store_window(R *dst, const R *s_acc, int tid) {
    vec4_t<R> *d = reinterpret_cast<vec4_t<R> *>(dst);
    for (int w = tid; w < N; w += kBlock) {
        vec4_t<R> a, b;
        a.x = s_acc[w]; a.y = s_acc[N + w]; a.z = s_acc[2 * N + w]; a.w = s_acc[3 * N + w];
        b.x = s_acc[4 * N + w]; b.y = s_acc[5 * N + w]; b.z = s_acc[6 * N + w]; b.w = s_acc[7 * N + w];
        d[2 * w] = a;
        d[2 * w + 1] = b;
    }
}

template <typename R, bool DO_G2P, bool DO_P2G, bool HAS_BACK, bool DET>
__device__ __forceinline__ void tile_block(const Particles<R> &P, const GridView<R> &g, const Grid<R> &G,
                                           const Material<R> &M, R dt, double dtd, double zcut, G2PFlags flags,
                                           double ref_temp, double *part, const int32_t *cell_start,
                                           const TileGeom &tg, int32_t *counters, R *tpart, int t) {
    static_assert(TX == 8 && TY == 8 && TZ * TX * TY == kBlock, "one 8x8 cell layer per wave");
    constexpr int kNodeBytes = DO_G2P ? WN * (int)sizeof(vec4_t<R>) : 0;
    constexpr int kAccBytes = DO_P2G ? 8 * WN * (int)sizeof(R) : 0;
    __shared__ __align__(32) unsigned char smem[kNodeBytes + kAccBytes + 32];
    __shared__ int32_t s_cnt[2][kBlock / kWave];
    __shared__ int32_t s_red[3][kBlock / kWave];
    __shared__ int32_t s_tab[kBlock / kWave][kSlotTab];   // uneven tiles: cells per wave holding more than s
    __shared__ int32_t s_q[DO_P2G && !DET ? kExcQueue : 1];  // particles whose P2G left their home cell
    __shared__ int32_t s_qn;
    vec4_t<R> *s_node = reinterpret_cast<vec4_t<R> *>(smem);
    R *s_acc = DO_P2G ? reinterpret_cast<R *>(smem + kNodeBytes) : nullptr;

    // statistics are reduced only by the G2P-only pass (the step before a
    // rebuild); the fused pass carries no statistics registers
    const bool stats = DO_G2P && !DO_P2G && flags.stats != 0;
    const int tid = threadIdx.x;
    const int lane = tid % kWave, wv = tid / kWave;
    const int hx = tid % TX, hy = (tid / TX) % TY, hz = tid / (TX * TY);
    const int tx = t % tg.ntx, ty = (t / tg.ntx) % tg.nty, tz = t / (tg.ntx * tg.nty);
    const int wx0 = tg.ox + tx * TX, wy0 = tg.oy + ty * TY, wz0 = tg.oz + tz * TZ;
    const int key = t * kTileCells + tid;
    const int c_begin = cell_start[key];
    const int cnt = cell_start[key + 1] - c_begin;
    const int b0 = cell_start[t * kTileCells];
    const int b1 = cell_start[(t + 1) * kTileCells];
    double st[kRed];
    stats_identity(st, ref_temp);

    if (b0 < b1) {  // uniform over the workgroup
        // Prologue: the window's node records and the first slot's particle
        // inputs are requested before anything waits on them (the first
        // slot is b0 + tid when the tile turns out uniform, else unused).
        static_assert(WN <= 2 * kBlock, "two window records per thread");
        vec4_t<R> wrec[2];
        if (DO_G2P) {
#pragma unroll
            for (int r = 0; r < 2; ++r) {
                const int w = tid + r * kBlock;
                const int i = wx0 + w % WX, j = wy0 + (w / WX) % WY, k = wz0 + w / (WX * WY);
                if (w < WN && i < g.nx && j < g.ny && k < g.nz) wrec[r] = G.node[((int64_t)k * g.ny + j) * g.nx + i];
                else wrec[r] = empty_record<R>();
            }
        }
        G2PIn<R> nxt;
        if (DO_G2P) load_g2p_in(P, cnt > 0 && b0 + tid < b1 ? b0 + tid : b0, nxt);
        if (DO_P2G) {
            for (int w = tid; w < 8 * WN; w += kBlock) s_acc[w] = R(0);
            if (tid == 0) s_qn = 0;  // before the count reduction's barrier
        }
        if (DO_G2P) {
#pragma unroll
            for (int r = 0; r < 2; ++r) {
                const int w = tid + r * kBlock;
                if (w < WN) s_node[w] = window_record<R>(wrec[r]);
            }
        }
        // largest and smallest particle count of the tile's cells, and the
        // smallest of its non-empty cells
        int mx = cnt, mn = cnt, mnz = cnt > 0 ? cnt : INT32_MAX;
#pragma unroll
        for (int o = kWave / 2; o > 0; o >>= 1) {
            mx = max(mx, __shfl_xor(mx, o, kWave));
            mn = min(mn, __shfl_xor(mn, o, kWave));
            mnz = min(mnz, __shfl_xor(mnz, o, kWave));
        }
        if (lane == 0) { s_red[0][wv] = mx; s_red[1][wv] = mn; s_red[2][wv] = mnz; }
        __syncthreads();
        int maxc = s_red[0][0], minc = s_red[1][0], minnz = s_red[2][0];
        for (int w = 1; w < kBlock / kWave; ++w) {
            maxc = max(maxc, s_red[0][w]);
            minc = min(minc, s_red[1][w]);
            minnz = min(minnz, s_red[2][w]);
        }
        const bool uniform = (maxc == minc);
        // every non-empty cell holds maxc particles (a workpiece's boundary
        // tiles): slot s of the cell of rank q among the non-empty cells is
        // particle b0 + s * nnz + q, found once instead of per slot
        const bool semi = !uniform && minnz == maxc;

        // home cell: its global cell index
        const int hgx = wx0 + hx, hgy = wy0 + hy, hgz = wz0 + hz;
        const bool home_ok = hgx + 1 < g.nx && hgy + 1 < g.ny && hgz + 1 < g.nz;
        CellAcc<R> A;
        if (DO_P2G) {
#pragma unroll
            for (int ch = 0; ch < 8; ++ch)
#pragma unroll
                for (int c = 0; c < 8; ++c) A.v[ch][c] = R(0);
        }
        int my_oow = 0, my_exc = 0, my_acc = 0;
        const unsigned long long lt = lane ? (~0ull >> (64 - lane)) : 0ull;
        int base = 0;
        int nnz = 0, rank_nz = 0;
        if (semi) {  // uniform over the workgroup
            const unsigned long long m = __ballot(cnt > 0);
            if (lane == 0) s_cnt[0][wv] = __popcll(m);
            __syncthreads();
            for (int w = 0; w < kBlock / kWave; ++w) {
                const int c = s_cnt[0][w];
                rank_nz += (w < wv) ? c : 0;
                nnz += c;
            }
            rank_nz += __popcll(m & lt);
        }
        // Uneven tiles of fewer than kSlotTab slots: every wave's count of
        // cells holding more than s particles, for every s, in one table
        // behind one barrier.  (A barrier per slot holds the four waves in
        // step, so no wave's memory latency overlaps another's compute: the
        // table made the 64-workpiece batched run 8 % faster.)
        const bool tab = !uniform && !semi && maxc < kSlotTab;
        if (tab) {  // uniform over the workgroup
            for (int sl = 0; sl <= maxc; ++sl) {
                const unsigned long long m = __ballot(cnt > sl);
                if (lane == 0) s_tab[wv][sl] = __popcll(m);
            }
            __syncthreads();
        }
        // Particle index of this thread's cell in slot s: uniform tiles
        // b0 + s * kBlock + tid, tiles whose non-empty cells hold equal
        // counts b0 + s * nnz + rank; otherwise the slot-major position
        // (cells with more than s particles, in key order), from a ballot
        // and a prefix over the waves (from the table, else through a
        // barrier per slot).
        auto slot_index = [&](int sl) -> int {
            if (uniform) return b0 + sl * kBlock + tid;
            if (semi) return b0 + sl * nnz + rank_nz;
            if (tab) {
                const unsigned long long mask = __ballot(cnt > sl);
                int before = 0, total = 0;
#pragma unroll
                for (int w = 0; w < kBlock / kWave; ++w) {
                    const int c = s_tab[w][sl];
                    before += (w < wv) ? c : 0;
                    total += c;
                }
                const int idx = b0 + base + before + __popcll(mask & lt);
                base += total;
                return idx;
            }
            const unsigned long long mask = __ballot(cnt > sl);
            if (lane == 0) s_cnt[sl & 1][wv] = __popcll(mask);
            __syncthreads();
            int before = 0, total = 0;
            for (int w = 0; w < kBlock / kWave; ++w) {
                const int c = s_cnt[sl & 1][w];
                before += (w < wv) ? c : 0;
                total += c;
            }
            const int idx = b0 + base + before + __popcll(mask & lt);
            base += total;
            return idx;
        };
        // A slot's G2P inputs are loaded one slot ahead (slot 0 in the
        // prologue), unconditionally (a branch joining two versions of them
        // would move the loaded registers, waiting for the loads on the
        // spot); a slot past a cell's last particle loads particle b0 and
        // does not use it.  A slot's G2P stores are made in the next slot,
        // after that slot has taken its prefetched inputs (after the loop
        // for the last): the wait for the inputs then covers loads and
        // stores issued a slot earlier, not stores just issued (the vector-
        // memory counter is in order and counts both).
        PState<R> pend;
        int pend_i = -1;
        uint32_t pend_st = 0;
        auto flush = [&]() {
            if (DO_G2P && pend_i >= 0) {
                g2p_store<R, HAS_BACK>(P, pend_i, pend, pend_st);
                pend_i = -1;
            }
        };
        auto process = [&](const G2PIn<R> &cur, int i, bool active) {
            if (!active) return;
            PState<R> p;
            if (DO_G2P) {
                p.status = cur.status;
            } else {
                p.status = P.status[i];
            }
            if (p.status == kInactive) return;
            const R mass = P.m ? P.m[i] : P.m0, vol = P.m ? P.V[i] : P.V0;
            R fex = R(0), fey = R(0), fez = R(0);
            if (DO_G2P) {
                const Frac<R> f = base_and_frac(g, cur.x, cur.y, cur.z);
                const int lx = f.b[0] - wx0, ly = f.b[1] - wy0, lz = f.b[2] - wz0;
                const bool in = (unsigned)lx < (unsigned)TX && (unsigned)ly < (unsigned)TY &&
                                (unsigned)lz < (unsigned)TZ;
                const int wb = wnode(lx, ly, lz);
                my_oow += in ? 0 : 1;
                pend_st = g2p_compute<R, HAS_BACK>(P, i, cur, p, M, dt, dtd, flags.chip != 0, stats, zcut,
                                                   [&](GatherSums<R> &q) {
                                                       if (in) gather_window(s_node, wb, f, g.inv_h, q);
                                                       else gather_global(g, G.node, f, q);
                                                   });
                pend = p;
                pend_i = i;
                if (stats) {
                    const double vx = p.vx, vy = p.vy, vz = p.vz;
                    accumulate_stats(st, p.x, p.y, p.z,
                                     (double)von_mises<R>(p.s[0], p.s[1], p.s[2], p.s[3], p.s[4], p.s[5]),
                                     (double)p.T, 0.5 * (double)mass * (vx * vx + vy * vy + vz * vz));
                }
            } else {
                p.x = P.x[i]; p.y = P.y[i]; p.z = P.z[i];
                load_p2g_state(P, i, p, fex, fey, fez);
            }
            if (DO_P2G) {
                // after G2P the external force has been consumed (:399)
                const Frac<R> f = base_and_frac(g, p.x, p.y, p.z);
                const bool stress_force = (p.status == kActive || p.status == kFixed);
                const bool ext_only = (p.status == kChip);
                if (home_ok && f.b[0] == hgx && f.b[1] == hgy && f.b[2] == hgz) {
                    cell_accumulate(A, f, g.inv_h, mass, vol, p.vx, p.vy, p.vz, p.T, p.s, fex, fey, fez,
                                    stress_force, ext_only, P.fx != nullptr);
                } else if (DET) {
                    ++my_exc;
                } else {
                    ++my_oow;
                    // queued for the scatter after the slot loop, where the
                    // queued particles take the lanes together instead of
                    // every slot of every wave holding one waiting on its
                    // compare-and-swap adds; scattered here when the queue
                    // is full
                    const unsigned long long em = __ballot(1);
                    const int leader = __ffsll((unsigned long long)em) - 1;
                    int qb = 0;
                    if (lane == leader) qb = atomicAdd(&s_qn, __popcll(em));
                    const int q = __shfl(qb, leader, kWave) + __popcll(em & lt);
                    if (q < kExcQueue) {
                        s_q[q] = i;
                    } else {
                        my_acc |= scatter_exception(g, G.acc, s_acc, f, wx0, wy0, wz0, mass, vol, p.vx, p.vy, p.vz,
                                                    p.T, p.s, fex, fey, fez, stress_force, ext_only)
                                      ? 1
                                      : 0;
                    }
                }
            }
        };
        int i_next = slot_index(0);
        if (DO_G2P && !uniform) load_g2p_in(P, cnt > 0 ? i_next : b0, nxt);
        for (int s = 0; s < maxc; ++s) {
            G2PIn<R> cur;
            if (DO_G2P) {
                cur = nxt;
                // the prefetched inputs are complete here, before this
                // slot's loads and the previous slot's stores are issued
                // (the asm uses them and keeps memory operations on its side)
                asm volatile("" ::"v"(cur.x), "v"(cur.y), "v"(cur.z), "v"(cur.T), "v"(cur.eps), "v"(cur.dmg),
                             "v"(cur.s[0]), "v"(cur.s[1]), "v"(cur.s[2]), "v"(cur.s[3]), "v"(cur.s[4]),
                             "v"(cur.s[5]), "v"(cur.status)
                             : "memory");
            }
            flush();
            const int i = i_next;
            i_next = slot_index(s + 1);
            if (DO_G2P) load_g2p_in(P, cnt > s + 1 ? i_next : b0, nxt);
            process(cur, i, cnt > s);
        }
        flush();
        if (DO_P2G) {
            __syncthreads();  // exception atomics are in
            combine_layer(A, s_acc, hx, hy, hz, 0);
            __syncthreads();  // layer z written by wave z only
            combine_layer(A, s_acc, hx, hy, hz, 1);
            __syncthreads();  // layer z+1 written by wave z only
            if (!DET) {
                // the queued particles, one per thread: their state as the
                // G2P above stored it (the barriers order those stores
                // before these loads); the external force was consumed by
                // the G2P of a fused pass
                const int nq = min(s_qn, kExcQueue);
                if (nq > 0) {  // uniform over the workgroup
                    for (int j = tid; j < nq; j += kBlock) {
                        const int i = s_q[j];
                        PState<R> p;
                        p.status = P.status[i];
                        p.x = P.x[i]; p.y = P.y[i]; p.z = P.z[i];
                        R fex = R(0), fey = R(0), fez = R(0);
                        load_p2g_state(P, i, p, fex, fey, fez);
                        if (DO_G2P) fex = fey = fez = R(0);
                        const R mass = P.m ? P.m[i] : P.m0, vol = P.m ? P.V[i] : P.V0;
                        const Frac<R> f = base_and_frac(g, p.x, p.y, p.z);
                        my_acc |= scatter_exception(g, G.acc, s_acc, f, wx0, wy0, wz0, mass, vol, p.vx, p.vy, p.vz,
                                                    p.T, p.s, fex, fey, fez, p.status == kActive || p.status == kFixed,
                                                    p.status == kChip)
                                      ? 1
                                      : 0;
                    }
                    __syncthreads();  // the queued scatters are in
                }
            }
            store_window<R, WN>(tpart + (int64_t)t * (8 * WN), s_acc, tid);
        }
        int cntv = my_oow, cnte = my_exc, cnta = my_acc;
#pragma unroll
        for (int o = kWave / 2; o > 0; o >>= 1) {
            cntv += __shfl_xor(cntv, o, kWave);
            cnte += __shfl_xor(cnte, o, kWave);
            cnta |= __shfl_xor(cnta, o, kWave);
        }
        if (lane == 0 && cntv) atomicAdd(counters, cntv);
        if (DET && lane == 0 && cnte) atomicAdd(counters + 1, cnte);
        if (!DET && lane == 0 && cnta) atomicOr(counters + 2, 1);
    } else if (DO_P2G) {
        zero_window<R, WN>(tpart + (int64_t)t * (8 * WN), tid);
    }
    if (stats) block_reduce_stats(st, part, t);
}

template <typename R, bool DO_G2P, bool DO_P2G, bool HAS_BACK, bool DET>
__global__ void __launch_bounds__(kBlock, TileWaves<R>::value)
    k_tile(Particles<R> P, GridView<R> g, Grid<R> G, Material<R> M, R dt, double dtd, double zcut, G2PFlags flags,
           double ref_temp, double *part, const int32_t *cell_start, TileGeom tg, int32_t *counters, R *tpart) {
    tile_block<R, DO_G2P, DO_P2G, HAS_BACK, DET>(P, g, G, M, dt, dtd, zcut, flags, ref_temp, part, cell_start, tg,
                                                 counters, tpart, (int)blockIdx.x);
}

// The tile pass of many solvers in one launch (see find_item): item w's
// tiles are blocks [block0, block0 + ntiles).
template <typename R> struct TileWork {
    Particles<R> P;
    GridView<R> g;
    Grid<R> G;
    Material<R> M;
    R dt;
    double dtd, zcut, ref_temp;
    G2PFlags flags;
    double *part;
    const int32_t *cell_start;
    TileGeom tg;
    int32_t *counters;
    R *tpart;
    int32_t block0, blocks;
};

template <typename R, bool DO_G2P, bool DO_P2G, bool HAS_BACK, bool DET>
__global__ void __launch_bounds__(kBlock, TileWaves<R>::value)
    k_tile_batch(const TileWork<R> *__restrict__ items, int nitems, int order) {
    // order 0 (blockIdx.x, the default): tiles own disjoint particle ranges
    // and with neighbouring tiles spread over the dies the slab's pass
    // measured 8 % faster than with the per-die ranges of order 1; the
    // autotuner measures both on the workload at hand
    const int b = launch_block(order);
    const int w = find_item(items, nitems, b);
    const TileWork<R> &W = items[w];
    tile_block<R, DO_G2P, DO_P2G, HAS_BACK, DET>(W.P, W.g, W.G, W.M, W.dt, W.dtd, W.zcut, W.flags, W.ref_temp, W.part,
                                                 W.cell_start, W.tg, W.counters, W.tpart, b - W.block0);
}

// ------------------------------------------------------ loose particles --

// The particles binned as loose (cell outside the tiles, or a stencil node
// outside the grid), one thread each, nodes read from device memory.  P2G
// here is the atomic mode's (device atomics into the node accumulators,
// setting acc_flag); the deterministic mode runs k_loose_p2g_ordered after
// this pass instead.  Statistics partials go to part[blockIdx.x].
template <typename R, bool DO_G2P, bool DO_P2G, bool HAS_BACK>
__global__ void __launch_bounds__(kBlock) k_loose(Particles<R> P, GridView<R> g, Grid<R> G, Material<R> M, R dt,
                                                  double dtd, double zcut, G2PFlags flags, double ref_temp,
                                                  double *part, int32_t begin, int32_t end, int32_t *acc_flag) {
    double st[kRed];
    stats_identity(st, ref_temp);
    const int i = begin + (int)blockIdx.x * kBlock + (int)threadIdx.x;
    if (i < end) {
        PState<R> p;
        p.status = P.status[i];
        if (p.status != kInactive) {
            const R mass = P.m ? P.m[i] : P.m0, vol = P.m ? P.V[i] : P.V0;
            R fex = R(0), fey = R(0), fez = R(0);
            p.x = P.x[i]; p.y = P.y[i]; p.z = P.z[i];
            if (DO_G2P) {
                G2PIn<R> in;
                load_g2p_in(P, i, in);
                const Frac<R> f = base_and_frac(g, p.x, p.y, p.z);
                g2p_particle<R, HAS_BACK>(P, i, in, p, M, dt, dtd, flags.chip != 0, flags.stats != 0, zcut,
                                          [&](GatherSums<R> &q) { gather_global(g, G.node, f, q); });
                if (flags.stats) {
                    const double vx = p.vx, vy = p.vy, vz = p.vz;
                    accumulate_stats(st, p.x, p.y, p.z,
                                     (double)von_mises<R>(p.s[0], p.s[1], p.s[2], p.s[3], p.s[4], p.s[5]),
                                     (double)p.T, 0.5 * (double)mass * (vx * vx + vy * vy + vz * vz));
                }
            } else {
                load_p2g_state(P, i, p, fex, fey, fez);
            }
            if (DO_P2G) {
                const Frac<R> f = base_and_frac(g, p.x, p.y, p.z);
                scatter_exception(g, G.acc, (R *)nullptr, f, 0, 0, 0, mass, vol, p.vx, p.vy, p.vz, p.T, p.s, fex,
                                  fey, fez, p.status == kActive || p.status == kFixed, p.status == kChip);
                atomicOr(acc_flag, 1);  // loose particles are few
            }
        }
    }
    if (DO_G2P && flags.stats) block_reduce_stats(st, part, (int)blockIdx.x);
}

// Deterministic P2G of the loose particles: one wave walks them in memory
// order, lane 8 * c + ch adding channel ch of corner c to the node
// accumulators.  One particle's adds go to distinct addresses; the barrier
// after each particle (workgroup-scope release + acquire) orders them before
// the next particle's, so every node sums its loose particles in index order.
// Sets acc_flag.
template <typename R>
__global__ void __launch_bounds__(kWave) k_loose_p2g_ordered(Particles<R> P, GridView<R> g, Grid<R> G,
                                                             int32_t begin, int32_t end, int32_t *acc_flag) {
    const int n = (int)threadIdx.x >> 3, ch = (int)threadIdx.x & 7;
    if (threadIdx.x == 0 && begin < end) *acc_flag = 1;
    for (int i = begin; i < end; ++i) {
        PState<R> p;
        p.status = P.status[i];
        if (p.status != kInactive) {
            const R mass = P.m ? P.m[i] : P.m0, vol = P.m ? P.V[i] : P.V0;
            R fex = R(0), fey = R(0), fez = R(0);
            load_p2g_state(P, i, p, fex, fey, fez);
            const Frac<R> f = base_and_frac(g, P.x[i], P.y[i], P.z[i]);
            int64_t idx;
            R w, gx, gy, gz;
            if (corner(g, f, n, idx, w, gx, gy, gz)) {
                R c[8];
                const int nch = p2g_corner(mass, vol, p.vx, p.vy, p.vz, p.T, p.s, fex, fey, fez,
                                           p.status == kActive || p.status == kFixed, p.status == kChip, w, gx, gy,
                                           gz, c);
                R v = c[0];
#pragma unroll
                for (int k = 1; k < 8; ++k) v = (ch == k) ? c[k] : v;
                if (nch == 8 || ch < 4 || ch == 7) {
                    R *a = G.acc + (int64_t)ch * g.nn + idx;
                    *a = *a + v;
                }
            }
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------- binning ----

// Tile-major cell key of every particle on the current grid.  A particle is
// loose (key nkeys) when its cell (the floor used by the P2G stencil) lies
// outside the tiles or has a stencil node outside the grid (trilinear: the
// cell's corners; quadratic, `quad`: one node before the cell to two after
// it); INACTIVE particles get key nkeys + 1 and sort to the end.
template <typename R>
__global__ void k_bin_keys(Particles<R> P, GridView<R> g, TileGeom tg, int quad, uint32_t *keys, int32_t *idx,
                           int32_t *counts) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    idx[i] = i;
    uint32_t key = (uint32_t)tg.nkeys + 1u;
    if (P.status[i] != kInactive) {
        const double r[3] = {(P.x[i] - g.ox) / g.hd, (P.y[i] - g.oy) / g.hd, (P.z[i] - g.oz) / g.hd};
        const int nt[3] = {tg.ntx * TX, tg.nty * TY, tg.ntz * TZ};
        const int nodes[3] = {g.nx, g.ny, g.nz};
        const int off[3] = {tg.ox, tg.oy, tg.oz};
        int c[3];
        bool tiled = true;
        const double lo = quad ? 1.0 : 0.0, hi = quad ? 2.0 : 1.0;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const double fl = floor(r[a]);
            // in a tile, and the stencil nodes inside the grid (NaN fails too)
            tiled = tiled && fl >= (double)off[a] && fl < (double)(off[a] + nt[a]) && fl - lo >= 0.0 &&
                    fl + hi < (double)nodes[a];
            c[a] = tiled ? (int)fl - off[a] : 0;
        }
        if (tiled) {
            const uint32_t tile = (uint32_t)(((c[2] / TZ) * tg.nty + c[1] / TY) * tg.ntx + c[0] / TX);
            const uint32_t local = (uint32_t)(((c[2] % TZ) * TY + c[1] % TY) * TX + c[0] % TX);
            key = tile * kTileCells + local;
        } else {
            key = (uint32_t)tg.nkeys;
        }
    }
    keys[i] = key;
    atomicAdd(&counts[key], 1);
}

// Slot-major order inside each tile: for slot s = 0, 1, ... the s-th particle
// of every cell holding more than s particles, cells in key order.
__global__ void __launch_bounds__(kBlock) k_slot_major(const int32_t *cell_start, const int32_t *idx_sorted,
                                                       int32_t *perm, int32_t *tile_start) {
    static_assert(kTileCells == kBlock, "one cell per thread");
    __shared__ int32_t s_wave[kBlock / kWave];
    __shared__ int32_t s_max[kBlock / kWave];
    const int t = blockIdx.x, c = threadIdx.x;
    const int lane = c % kWave, wid = c / kWave;
    const int key = t * kTileCells + c;
    const int cs = cell_start[key];
    const int cnt = cell_start[key + 1] - cs;
    const int tstart = cell_start[t * kTileCells];
    if (c == 0) tile_start[t] = tstart;
    int m = cnt;
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o, kWave));
    if (lane == 0) s_max[wid] = m;
    __syncthreads();
    int maxc = s_max[0];
    for (int w = 1; w < kBlock / kWave; ++w) maxc = max(maxc, s_max[w]);
    const unsigned long long lt = lane ? (~0ull >> (64 - lane)) : 0ull;
    int base = tstart;
    for (int s = 0; s < maxc; ++s) {
        const bool flag = cnt > s;
        const unsigned long long mask = __ballot(flag);
        const int pos = __popcll(mask & lt);
        __syncthreads();
        if (lane == 0) s_wave[wid] = __popcll(mask);
        __syncthreads();
        int before = 0, total = 0;
        for (int w = 0; w < kBlock / kWave; ++w) {
            if (w < wid) before += s_wave[w];
            total += s_wave[w];
        }
        if (flag) perm[base + before + pos] = idx_sorted[cs + s];
        base += total;
    }
}

// Inactive particles (key nkeys) keep their sorted order after the tiles.
__global__ void k_tail(const int32_t *cell_start, int32_t nkeys, int32_t ntiles, const int32_t *idx_sorted,
                       int32_t *perm, int32_t *tile_start, int n) {
    const int first = cell_start[nkeys];
    const int i = first + blockIdx.x * blockDim.x + threadIdx.x;
    if (blockIdx.x == 0 && threadIdx.x == 0) tile_start[ntiles] = first;
    if (i < n) perm[i] = idx_sorted[i];
}

}  // namespace mpmw
