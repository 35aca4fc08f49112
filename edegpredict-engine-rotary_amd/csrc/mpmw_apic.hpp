// mpmw_apic.hpp -- the APIC transfer variant of the particle pass.
//
// Quadratic B-spline weights over 27 nodes with base = floor(x/h - 0.5)
// (getWeightsAndGradients, MPMSolver.cu:20-34), affine momentum
// m w (v + C (x_i - x_p)) in P2G (p2gKernel :370-421) and the affine matrix
// C = 0.99 * 3/h^2 * sum w v_i (x_i - x_p)^T from G2P (g2pKernel :554-592);
// the rest of the step (forces, grid update, contact, diffusion,
// constitutive update, chip pass) is the workpiece solver's.
//
// Same tiles and binning as the trilinear pass.  A particle of cell c has
// its base at c - 1 or c, so the node window of a tile of TX x TY x TZ cells
// spans one node before the tile and two after it: QX x QY x QZ nodes.  The
// 27 x 8 contributions of a particle go into the LDS window with LDS adds
// (lds_add8; a register accumulation per cell does not fit 27 nodes); the
// window is then stored, with plain stores, into the tile's slot of the
// per-tile window buffer, and the grid update sums the windows holding each
// node in tile order.  Particles whose stencil leaves their tile's window
// go to the node accumulators with device atomics.
#pragma once

#include "mpmw_tile.hpp"

namespace mpmw {

__device__ __forceinline__ int qnode(int x, int y, int z) { return (z * QY + y) * QX + x; }

// The per-axis quadratic weights of one particle.  d0 = r - base in
// [0.5, 1.5); node offsets x_i - x_p = (i - d0) h.
template <typename R> struct QFrac {
    int b[3];
    R w[3][3], dw[3][3], d0[3];
};

// getWeightsAndGradients (MPMSolver.cu:20-34) on the three axes; r is
// formed in FP64 as the reference forms x_frac, (x - origin) * invDx.
template <typename R>
__device__ __forceinline__ QFrac<R> quad_frac(const GridView<R> &g, double x, double y, double z) {
    QFrac<R> q;
    const double d[3] = {x - g.ox, y - g.oy, z - g.oz};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double r = d[a] * g.inv_hd;
        const double fl = floor(r - 0.5);
        q.b[a] = (int)fl;
        const R d0 = (R)(r - fl), d1 = d0 - R(1), d2 = d0 - R(2);
        q.d0[a] = d0;
        q.w[a][0] = R(0.5) * (R(1.5) - d0) * (R(1.5) - d0);
        q.w[a][1] = R(0.75) - d1 * d1;
        q.w[a][2] = R(0.5) * (R(1.5) + d2) * (R(1.5) + d2);
        q.dw[a][0] = -(R(1.5) - d0) * g.inv_h;
        q.dw[a][1] = R(-2) * d1 * g.inv_h;
        q.dw[a][2] = (R(1.5) + d2) * g.inv_h;
    }
    return q;
}

// Node (i, j, k) of the stencil: weight, gradient and offset x_i - x_p.
template <typename R>
__device__ __forceinline__ void quad_node(const QFrac<R> &q, int i, int j, int k, R h, R &w, R &gx, R &gy, R &gz,
                                          R &dx, R &dy, R &dz) {
    w = q.w[0][i] * q.w[1][j] * q.w[2][k];
    gx = q.dw[0][i] * q.w[1][j] * q.w[2][k];
    gy = q.w[0][i] * q.dw[1][j] * q.w[2][k];
    gz = q.w[0][i] * q.w[1][j] * q.dw[2][k];
    dx = (R(i) - q.d0[0]) * h;
    dy = (R(j) - q.d0[1]) * h;
    dz = (R(k) - q.d0[2]) * h;
}

// One live node's share of the G2P sums and of B.
template <typename R>
__device__ __forceinline__ void quad_gather_add(const vec4_t<R> &nd, R w, R gx, R gy, R gz, R dx, R dy, R dz,
                                                GatherSums<R> &s, R B[9]) {
    s.v0 += w * nd.x; s.v1 += w * nd.y; s.v2 += w * nd.z;
    s.temp += nd.w * w;
    s.wsum += w;
    const R wx = w * nd.x, wy = w * nd.y, wz = w * nd.z;
    B[0] += wx * dx; B[1] += wx * dy; B[2] += wx * dz;
    B[3] += wy * dx; B[4] += wy * dy; B[5] += wy * dz;
    B[6] += wz * dx; B[7] += wz * dy; B[8] += wz * dz;
    s.L[0] += nd.x * gx; s.L[1] += nd.x * gy; s.L[2] += nd.x * gz;
    s.L[3] += nd.y * gx; s.L[4] += nd.y * gy; s.L[5] += nd.y * gz;
    s.L[6] += nd.z * gx; s.L[7] += nd.z * gy; s.L[8] += nd.z * gz;
}

// G2P gather over the 27 nodes from the LDS window (base node at window
// index wb; window nodes outside the grid hold the empty NaN record) or
// from device memory (nodes outside the grid skipped).  Empty nodes are
// skipped (g2pKernel :538, with the workpiece grid's emptiness test).
template <typename R>
__device__ __forceinline__ void quad_gather_window(const vec4_t<R> *s_node, int wb, const QFrac<R> &q, R h,
                                                   GatherSums<R> &s, R B[9]) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const vec4_t<R> nd = s_node[wb + qnode(i, j, k)];
                if (nd.w != nd.w) continue;
                R w, gx, gy, gz, dx, dy, dz;
                quad_node(q, i, j, k, h, w, gx, gy, gz, dx, dy, dz);
                quad_gather_add(nd, w, gx, gy, gz, dx, dy, dz, s, B);
            }
}

template <typename R>
__device__ __forceinline__ void quad_gather_global(const GridView<R> &g, gp<const vec4_t<R>> node, const QFrac<R> &q,
                                                   R h, GatherSums<R> &s, R B[9]) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const int x = q.b[0] + i, y = q.b[1] + j, z = q.b[2] + k;
                if (x < 0 || y < 0 || z < 0 || x >= g.nx || y >= g.ny || z >= g.nz) continue;
                const vec4_t<R> nd = node[((int64_t)z * g.ny + y) * g.nx + x];
                if (nd.w != nd.w) continue;
                R w, gx, gy, gz, dx, dy, dz;
                quad_node(q, i, j, k, h, w, gx, gy, gz, dx, dy, dz);
                quad_gather_add(nd, w, gx, gy, gz, dx, dy, dz, s, B);
            }
}

// What one particle adds in the APIC P2G.
template <typename R> struct QP2G {
    R mass, vol, vx, vy, vz, T;
    R s[6];
    R C[9];
    R fex, fey, fez;
    bool momentum;   // not FIXED_BOUNDARY (p2gKernel :393-398)
    bool stress;     // ACTIVE or FIXED_BOUNDARY: stress force (mpmP2GKernel :121-127)
    bool ext;        // stress or CHIP: external force (:121-132)
};

// A particle's P2G terms in the form the 27 nodes share: with node offset
// x_i - x_p = (idx - d0) h, the affine velocity v + C (x_i - x_p) is
// u + (h C) idx with u = v - (h C) d0, so a node adds (h C) idx to u with
// its small integer idx (the sum over idx x and y is shared by the nodes of
// a column); the stress force is S grad w with S = -V sigma.
template <typename R> struct QScatter {
    R m, T, u[3], hc[9], S[6], fe[3];
};

template <typename R>
__device__ __forceinline__ QScatter<R> quad_scatter_terms(const QP2G<R> &a, const QFrac<R> &q, R h) {
    QScatter<R> z;
    z.m = a.mass;
    z.T = a.T;
#pragma unroll
    for (int c = 0; c < 9; ++c) z.hc[c] = a.C[c] * h;
    const R v[3] = {a.vx, a.vy, a.vz};
#pragma unroll
    for (int r = 0; r < 3; ++r)
        z.u[r] = v[r] - (z.hc[3 * r] * q.d0[0] + z.hc[3 * r + 1] * q.d0[1] + z.hc[3 * r + 2] * q.d0[2]);
#pragma unroll
    for (int c = 0; c < 6; ++c) z.S[c] = -a.vol * a.s[c];
    z.fe[0] = a.fex; z.fe[1] = a.fey; z.fe[2] = a.fez;
    return z;
}

// The 8 channels of node (i, j, k) of the stencil: mass, momentum, force,
// T * m.
template <typename R>
__device__ __forceinline__ void quad_node_channels(const QScatter<R> &z, int i, int j, int k, R w, R gx, R gy, R gz,
                                                   R c[8]) {
    const R m = z.m * w;
    c[0] = m;
#pragma unroll
    for (int r = 0; r < 3; ++r)
        c[1 + r] = m * (z.u[r] + z.hc[3 * r] * R(i) + z.hc[3 * r + 1] * R(j) + z.hc[3 * r + 2] * R(k));
    c[4] = z.S[0] * gx + z.S[3] * gy + z.S[4] * gz + z.fe[0] * w;
    c[5] = z.S[3] * gx + z.S[1] * gy + z.S[5] * gz + z.fe[1] * w;
    c[6] = z.S[4] * gx + z.S[5] * gy + z.S[2] * gz + z.fe[2] * w;
    c[7] = z.T * m;
}

// FP32 LDS window of the APIC pass: the 8 channels of a node as four 64-bit
// words of channel pairs (0, 7) (1, 2) (3, 4) (5, 6), word q of node w at
// index q * QN + w, so that a node takes four compare-and-swaps instead of
// eight (ds_cmpst_rtn_b64 ran at 3.2 lane-ops per clock per CU against 4.6
// for the 32-bit form).  A pair whose two channels are both skipped (no
// momentum from FIXED_BOUNDARY, no force from BOUNDARY) is not touched; in
// pair (3, 4) the skipped one of the two adds +0.
__device__ __forceinline__ void lds_add_pairs(unsigned long long *p, const float c[8], bool momentum, bool ext) {
    const float v[4][2] = {{c[0], c[7]},
                           {c[1], c[2]},
                           {momentum ? c[3] : 0.0f, ext ? c[4] : 0.0f},
                           {c[5], c[6]}};
    unsigned live = 1u | (momentum ? 2u : 0u) | ((momentum || ext) ? 4u : 0u) | (ext ? 8u : 0u);
    unsigned long long old[4];
#pragma unroll
    for (int q = 0; q < 4; ++q)
        old[q] = (live & (1u << q)) ? __hip_atomic_load(p + q * QN, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)
                                    : 0ull;
    while (live) {
        unsigned again = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (!(live & (1u << q))) continue;
            unsigned long long expect = old[q];
            const float lo = __uint_as_float((unsigned)expect) + v[q][0];
            const float hi = __uint_as_float((unsigned)(expect >> 32)) + v[q][1];
            const unsigned long long want =
                (unsigned long long)__float_as_uint(lo) | ((unsigned long long)__float_as_uint(hi) << 32);
            if (!__hip_atomic_compare_exchange_strong(p + q * QN, &expect, want, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                      __HIP_MEMORY_SCOPE_WORKGROUP)) {
                old[q] = expect;
                again |= 1u << q;
            }
        }
        live = again;
    }
}

// The FP32 pair window stored node-major (the layout store_window writes).
__device__ __forceinline__ void store_window_pairs(float *dst, const float *s_acc, int tid) {
    const unsigned long long *p = reinterpret_cast<const unsigned long long *>(s_acc);
    float4 *d = reinterpret_cast<float4 *>(dst);
    for (int w = tid; w < QN; w += kBlock) {
        const unsigned long long p0 = p[w], p1 = p[QN + w], p2 = p[2 * QN + w], p3 = p[3 * QN + w];
        float4 a, b;
        a.x = __uint_as_float((unsigned)p0); b.w = __uint_as_float((unsigned)(p0 >> 32));
        a.y = __uint_as_float((unsigned)p1); a.z = __uint_as_float((unsigned)(p1 >> 32));
        a.w = __uint_as_float((unsigned)p2); b.x = __uint_as_float((unsigned)(p2 >> 32));
        b.y = __uint_as_float((unsigned)p3); b.z = __uint_as_float((unsigned)(p3 >> 32));
        d[2 * w] = a;
        d[2 * w + 1] = b;
    }
}

// Scatter into the LDS window (window index wb of the base node): FP32 as
// channel pairs (lds_add_pairs), FP64 per channel with stride QN.
template <typename R>
__device__ __forceinline__ void quad_scatter_window(R *s_acc, int wb, const QFrac<R> &q, R h, const QP2G<R> &a) {
    const unsigned live = 0x81u | (a.momentum ? 0x0eu : 0u) | (a.ext ? 0x70u : 0u);
    const QScatter<R> z = quad_scatter_terms(a, q, h);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                R w, gx, gy, gz, dx, dy, dz, c[8];
                quad_node(q, i, j, k, h, w, gx, gy, gz, dx, dy, dz);
                quad_node_channels(z, i, j, k, w, gx, gy, gz, c);
                if constexpr (sizeof(R) == 4)
                    lds_add_pairs(reinterpret_cast<unsigned long long *>(s_acc) + wb + qnode(i, j, k), c, a.momentum,
                                  a.ext);
                else
                    lds_add8(s_acc + wb + qnode(i, j, k), QN, c, live);
            }
}

// Scatter into the node accumulators with device atomics (nodes outside the
// grid skipped, as p2gKernel :371-378 skips them).
template <typename R>
__device__ __forceinline__ void quad_scatter_global(const GridView<R> &g, gp<R> acc, const QFrac<R> &q, R h,
                                                    const QP2G<R> &a) {
    const QScatter<R> zt = quad_scatter_terms(a, q, h);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const int x = q.b[0] + i, y = q.b[1] + j, z = q.b[2] + k;
                if (x < 0 || y < 0 || z < 0 || x >= g.nx || y >= g.ny || z >= g.nz) continue;
                R w, gx, gy, gz, dx, dy, dz, c[8];
                quad_node(q, i, j, k, h, w, gx, gy, gz, dx, dy, dz);
                quad_node_channels(zt, i, j, k, w, gx, gy, gz, c);
                R *p = acc + ((int64_t)z * g.ny + y) * g.nx + x;
                unsafeAtomicAdd(p, c[0]);
                if (a.momentum) {
                    unsafeAtomicAdd(p + g.nn, c[1]);
                    unsafeAtomicAdd(p + 2 * g.nn, c[2]);
                    unsafeAtomicAdd(p + 3 * g.nn, c[3]);
                }
                if (a.ext) {
                    unsafeAtomicAdd(p + 4 * g.nn, c[4]);
                    unsafeAtomicAdd(p + 5 * g.nn, c[5]);
                    unsafeAtomicAdd(p + 6 * g.nn, c[6]);
                }
                unsafeAtomicAdd(p + 7 * g.nn, c[7]);
            }
}

template <typename R> __device__ __forceinline__ void load_affine(const Particles<R> &P, int i, R C[9]) {
#pragma unroll
    for (int c = 0; c < 9; ++c) C[c] = P.aff[c][i];
}

// The APIC P2G inputs of a particle from its state p (after its G2P, or as
// loaded for a P2G-only pass) and its affine matrix.
template <typename R>
__device__ __forceinline__ QP2G<R> quad_inputs(const PState<R> &p, const R C[9], R mass, R vol, R fex, R fey,
                                               R fez) {
    QP2G<R> a;
    a.mass = mass; a.vol = vol;
    a.vx = p.vx; a.vy = p.vy; a.vz = p.vz;
    a.T = p.T;
    a.stress = p.status == kActive || p.status == kFixed;
    a.ext = a.stress || p.status == kChip;
    a.momentum = p.status != kFixed;
#pragma unroll
    for (int c = 0; c < 6; ++c) a.s[c] = a.stress ? p.s[c] : R(0);
#pragma unroll
    for (int c = 0; c < 9; ++c) a.C[c] = C[c];
    a.fex = a.ext ? fex : R(0);
    a.fey = a.ext ? fey : R(0);
    a.fez = a.ext ? fez : R(0);
    return a;
}

// G2P of one particle with the quadratic gather, then its affine matrix:
// C = B * 3/h^2, C *= 0.99 (g2pKernel :577-591; cfac and decay, which the
// textbook form sets to 4/h^2 and 1) unless the particle was CHIP
// (:570-573) or FIXED_BOUNDARY (:516-519) at entry.  C is left in C[] for a
// following P2G.
template <typename R, bool HAS_BACK, typename GatherFn>
__device__ __forceinline__ void apic_g2p(const Particles<R> &P, int i, const G2PIn<R> &in, PState<R> &p,
                                         const Material<R> &M, R dt, double dtd, bool chip_step, bool need_stress,
                                         double zcut, R cfac, R decay, R C[9], GatherFn gather_fn) {
    R B[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) B[c] = R(0);
    g2p_particle<R, HAS_BACK>(P, i, in, p, M, dt, dtd, chip_step, need_stress, zcut,
                              [&](GatherSums<R> &s) { gather_fn(s, B); });
    if (in.status != kChip && in.status != kFixed) {
#pragma unroll
        for (int c = 0; c < 9; ++c) {
            C[c] = B[c] * cfac;
            C[c] = C[c] * decay;
            P.aff[c][i] = C[c];
        }
    } else {
        load_affine(P, i, C);
    }
}

// One workgroup per tile: the APIC G2P and / or P2G of the tile's particles,
// thread tid taking particles b0 + tid, b0 + tid + kBlock, ... (slot-major
// order: consecutive particles sit in different cells).  The window goes to
// tpart[t] (QN node records of 8 channels).  counters[0] counts particles whose
// stencil left the window, counters[2] is set when one was scattered into
// the node accumulators.
template <typename R, bool DO_G2P, bool DO_P2G, bool HAS_BACK>
__device__ __forceinline__ void apic_tile_block(const Particles<R> &P, const GridView<R> &g, const Grid<R> &G,
                                                const Material<R> &M, R dt, double dtd, double zcut,
                                                G2PFlags flags, double ref_temp, double *part,
                                                const int32_t *cell_start, const TileGeom &tg, int32_t *counters,
                                                R *tpart, int t) {
    constexpr int kNodeBytes = DO_G2P ? QN * (int)sizeof(vec4_t<R>) : 0;
    constexpr int kAccBytes = DO_P2G ? 8 * QN * (int)sizeof(R) : 0;
    __shared__ __align__(32) unsigned char smem[kNodeBytes + kAccBytes + 32];
    vec4_t<R> *s_node = reinterpret_cast<vec4_t<R> *>(smem);
    R *s_acc = DO_P2G ? reinterpret_cast<R *>(smem + kNodeBytes) : nullptr;
    const bool stats = DO_G2P && !DO_P2G && flags.stats != 0;
    const int tid = threadIdx.x, lane = tid % kWave;
    const int tx = t % tg.ntx, ty = (t / tg.ntx) % tg.nty, tz = t / (tg.ntx * tg.nty);
    // window node 0 is the node before the tile's first cell
    const int qx0 = tg.ox + tx * TX - 1, qy0 = tg.oy + ty * TY - 1, qz0 = tg.oz + tz * TZ - 1;
    const int b0 = cell_start[t * kTileCells];
    const int b1 = cell_start[(t + 1) * kTileCells];
    const R h = (R)g.hd;
    const R cfac = (R)(g.apic_k * g.inv_hd * g.inv_hd), decay = (R)g.apic_decay;
    double st[kRed];
    stats_identity(st, ref_temp);
    if (b0 < b1) {  // uniform over the workgroup
        if (DO_G2P) {
            for (int w = tid; w < QN; w += kBlock) {
                const int i = qx0 + w % QX, j = qy0 + (w / QX) % QY, k = qz0 + w / (QX * QY);
                vec4_t<R> nd;
                if (i >= 0 && j >= 0 && k >= 0 && i < g.nx && j < g.ny && k < g.nz) {
                    nd = G.node[((int64_t)k * g.ny + j) * g.nx + i];
                } else {
                    nd.x = nd.y = nd.z = R(0);
                    nd.w = R(__builtin_nan(""));
                }
                s_node[w] = nd;
            }
        }
        if (DO_P2G)
            for (int w = tid; w < 8 * QN; w += kBlock) s_acc[w] = R(0);
        __syncthreads();
        int my_oow = 0, my_acc = 0;
        for (int i = b0 + tid; i < b1; i += kBlock) {
            PState<R> p;
            p.status = P.status[i];
            if (p.status == kInactive) continue;
            const R mass = P.m ? P.m[i] : P.m0, vol = P.m ? P.V[i] : P.V0;
            R fex = R(0), fey = R(0), fez = R(0);
            R C[9];
            if (DO_G2P) {
                G2PIn<R> in;
                load_g2p_in(P, i, in);
                const QFrac<R> q = quad_frac(g, in.x, in.y, in.z);
                const int lx = q.b[0] - qx0, ly = q.b[1] - qy0, lz = q.b[2] - qz0;
                const bool inw = (unsigned)lx <= (unsigned)TX && (unsigned)ly <= (unsigned)TY &&
                                 (unsigned)lz <= (unsigned)TZ;
                const int wb = qnode(lx, ly, lz);
                my_oow += inw ? 0 : 1;
                apic_g2p<R, HAS_BACK>(P, i, in, p, M, dt, dtd, flags.chip != 0, stats, zcut, cfac, decay, C,
                                      [&](GatherSums<R> &s, R B[9]) {
                                          if (inw) quad_gather_window(s_node, wb, q, h, s, B);
                                          else quad_gather_global(g, G.node, q, h, s, B);
                                      });
                if (stats) {
                    const double vx = p.vx, vy = p.vy, vz = p.vz;
                    accumulate_stats(st, p.x, p.y, p.z,
                                     (double)von_mises<R>(p.s[0], p.s[1], p.s[2], p.s[3], p.s[4], p.s[5]),
                                     (double)p.T, 0.5 * (double)mass * (vx * vx + vy * vy + vz * vz));
                }
            } else {
                p.x = P.x[i]; p.y = P.y[i]; p.z = P.z[i];
                load_p2g_state(P, i, p, fex, fey, fez);
                load_affine(P, i, C);
            }
            if (DO_P2G) {
                // after G2P the external force has been consumed (:399)
                const QFrac<R> q = quad_frac(g, p.x, p.y, p.z);
                const QP2G<R> a = quad_inputs(p, C, mass, vol, fex, fey, fez);
                const int lx = q.b[0] - qx0, ly = q.b[1] - qy0, lz = q.b[2] - qz0;
                if ((unsigned)lx <= (unsigned)TX && (unsigned)ly <= (unsigned)TY && (unsigned)lz <= (unsigned)TZ) {
                    quad_scatter_window(s_acc, qnode(lx, ly, lz), q, h, a);
                } else {
                    ++my_oow;
                    my_acc = 1;
                    quad_scatter_global(g, G.acc, q, h, a);
                }
            }
        }
        if (DO_P2G) {
            __syncthreads();  // every LDS atomic is in
            if constexpr (sizeof(R) == 4) store_window_pairs(tpart + (int64_t)t * (8 * QN), s_acc, tid);
            else store_window<R, QN>(tpart + (int64_t)t * (8 * QN), s_acc, tid);
        }
        int cntv = my_oow, cnta = my_acc;
#pragma unroll
        for (int o = kWave / 2; o > 0; o >>= 1) {
            cntv += __shfl_xor(cntv, o, kWave);
            cnta |= __shfl_xor(cnta, o, kWave);
        }
        if (lane == 0 && cntv) atomicAdd(counters, cntv);
        if (lane == 0 && cnta) atomicOr(counters + 2, 1);
    } else if (DO_P2G) {
        zero_window<R, QN>(tpart + (int64_t)t * (8 * QN), tid);
    }
    if (stats) block_reduce_stats(st, part, t);
}

template <typename R, bool DO_G2P, bool DO_P2G, bool HAS_BACK>
__global__ void __launch_bounds__(kBlock) k_apic_tile_batch(const TileWork<R> *__restrict__ items, int nitems,
                                                            int order) {
    const int b = launch_block(order);
    const int w = find_item(items, nitems, b);
    const TileWork<R> &W = items[w];
    apic_tile_block<R, DO_G2P, DO_P2G, HAS_BACK>(W.P, W.g, W.G, W.M, W.dt, W.dtd, W.zcut, W.flags, W.ref_temp,
                                                 W.part, W.cell_start, W.tg, W.counters, W.tpart, b - W.block0);
}

// The loose particles of the APIC pass, one thread each, nodes read from
// and accumulated into device memory (sets acc_flag).  Statistics partials
// go to part[blockIdx.x].
template <typename R, bool DO_G2P, bool DO_P2G, bool HAS_BACK>
__global__ void __launch_bounds__(kBlock)
    k_apic_loose(Particles<R> P, GridView<R> g, Grid<R> G, Material<R> M, R dt, double dtd, double zcut,
                 G2PFlags flags, double ref_temp, double *part, int32_t begin, int32_t end, int32_t *acc_flag) {
    double st[kRed];
    stats_identity(st, ref_temp);
    const int i = begin + (int)blockIdx.x * kBlock + (int)threadIdx.x;
    const R h = (R)g.hd;
    const R cfac = (R)(g.apic_k * g.inv_hd * g.inv_hd), decay = (R)g.apic_decay;
    if (i < end) {
        PState<R> p;
        p.status = P.status[i];
        if (p.status != kInactive) {
            const R mass = P.m ? P.m[i] : P.m0, vol = P.m ? P.V[i] : P.V0;
            R fex = R(0), fey = R(0), fez = R(0);
            R C[9];
            if (DO_G2P) {
                G2PIn<R> in;
                load_g2p_in(P, i, in);
                const QFrac<R> q = quad_frac(g, in.x, in.y, in.z);
                apic_g2p<R, HAS_BACK>(P, i, in, p, M, dt, dtd, flags.chip != 0, flags.stats != 0, zcut, cfac, decay, C,
                                      [&](GatherSums<R> &s, R B[9]) { quad_gather_global(g, G.node, q, h, s, B); });
                if (flags.stats) {
                    const double vx = p.vx, vy = p.vy, vz = p.vz;
                    accumulate_stats(st, p.x, p.y, p.z,
                                     (double)von_mises<R>(p.s[0], p.s[1], p.s[2], p.s[3], p.s[4], p.s[5]),
                                     (double)p.T, 0.5 * (double)mass * (vx * vx + vy * vy + vz * vz));
                }
            } else {
                p.x = P.x[i]; p.y = P.y[i]; p.z = P.z[i];
                load_p2g_state(P, i, p, fex, fey, fez);
                load_affine(P, i, C);
            }
            if (DO_P2G) {
                const QFrac<R> q = quad_frac(g, p.x, p.y, p.z);
                quad_scatter_global(g, G.acc, q, h, quad_inputs(p, C, mass, vol, fex, fey, fez));
                atomicOr(acc_flag, 1);
            }
        }
    }
    if (DO_G2P && flags.stats) block_reduce_stats(st, part, (int)blockIdx.x);
}

}  // namespace mpmw
