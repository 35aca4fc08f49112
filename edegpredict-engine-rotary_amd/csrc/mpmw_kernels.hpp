// mpmw_kernels.hpp -- HIP kernels of the MPM workpiece step for gfx950.
//
// Kernel map (reference kernel -> here):
//   clearMPMGridKernel          -> the grid update zeroes the accumulators it has read
//   mpmP2GKernel                -> k_tile<.., P2G> (mpmw_tile.hpp)
//   mpmGridUpdateKernel         -> k_grid_update_batch (grid_update_block)
//   mpmDiffuseTemperatureKernel -> k_post_batch (diffuse_block, race-free Jacobi form)
//   mpmG2PKernel + chipFractureAndSeparationKernel -> k_tile<.., G2P, ..>
//   updateMetrics / rebuild AABB host loops -> block partials in k_tile / k_reduce
#pragma once

#include "mpmw_device.hpp"
#include "mpmw_sort.hpp"

namespace mpmw {

constexpr int kBlock = 256;
constexpr int kWave = 64;
constexpr int kRed = 9;  // AABB min xyz, -max xyz, max von Mises, max T, kinetic energy
constexpr int kXcds = 8;  // accelerator dies of the MI355X, each with its own L2
// Nodes per thread of the diffusion pass (4: 28 -> 21 us at 2.65 M nodes;
// the grid update measured fastest at one node per thread).
constexpr int kDiffuseNodes = 4;
constexpr int kDiffuseBlock = kBlock * kDiffuseNodes;

// The block's index in a launch read as kXcds contiguous ranges, one per
// XCD.  Workgroups go to the XCDs round-robin (workgroup b to XCD b % 8), so
// with blockIdx.x neighbouring blocks run on different dies and every cache
// line two of them share is fetched into two L2s; with this index each die
// works through one contiguous range of the launch and neighbours share its
// L2.  A bijection of [0, gridDim.x): what a block computes is unchanged.
__device__ __forceinline__ int xcd_block() {
    const int nb = (int)gridDim.x, b = (int)blockIdx.x;
    const int q = nb / kXcds, r = nb % kXcds;
    const int x = b % kXcds, pos = b / kXcds;
    return x < r ? x * (q + 1) + pos : r * (q + 1) + (x - r) * q + pos;
}

// ------------------------------------------------------------ reductions --

__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, kWave));
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, kWave));
    return v;
}
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

// Fixed-order block reduction of the kRed particle statistics into
// part[blk * kRed + c].  Slots 0-5 and 6-7 reduce with min/max, slot 8
// with a sum whose order depends only on thread and block indices.
__device__ __forceinline__ void block_reduce_stats(double v[kRed], double *part, int blk) {
    __shared__ double sh[kBlock / kWave][kRed];
    const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
#pragma unroll
    for (int c = 0; c < kRed; ++c) {
        double r = (c < 6) ? wave_min(v[c]) : (c < 8 ? wave_max(v[c]) : wave_sum(v[c]));
        if (lane == 0) sh[wid][c] = r;
    }
    __syncthreads();
    if (threadIdx.x < kRed) {
        const int c = threadIdx.x;
        double r = sh[0][c];
        for (int w = 1; w < kBlock / kWave; ++w)
            r = (c < 6) ? fmin(r, sh[w][c]) : (c < 8 ? fmax(r, sh[w][c]) : r + sh[w][c]);
        part[(int64_t)blk * kRed + c] = r;
    }
}

__device__ __forceinline__ void stats_identity(double v[kRed], double ref_temp) {
#pragma unroll
    for (int c = 0; c < 6; ++c) v[c] = 1e30;
    v[6] = 0.0;
    v[7] = ref_temp;
    v[8] = 0.0;
}

// Single-block fixed-order reduction of nparts partial records of width w.
// mode per column: 0 min, 1 max, 2 sum.
__device__ __forceinline__ void finalize_block(const double *part, int nparts, int w, const int32_t *modes,
                                               double *out) {
    __shared__ double sh[kBlock / kWave];
    for (int c = 0; c < w; ++c) {
        const int mode = modes[c];
        double r = mode == 0 ? 1e300 : (mode == 1 ? -1e300 : 0.0);
        // each thread combines parts tid, tid + kBlock, ... in that order;
        // sixteen loads are issued before the first of them is combined (the
        // partials come from every die, a memory round trip each: 41 k
        // contact-force partials of a 2.65 M-node grid take ten rounds)
        int i = threadIdx.x;
        for (; i + 15 * kBlock < nparts; i += 16 * kBlock) {
            double x[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) x[u] = part[(int64_t)(i + u * kBlock) * w + c];
#pragma unroll
            for (int u = 0; u < 16; ++u) r = mode == 0 ? fmin(r, x[u]) : (mode == 1 ? fmax(r, x[u]) : r + x[u]);
        }
        for (; i < nparts; i += kBlock) {
            double x = part[(int64_t)i * w + c];
            r = mode == 0 ? fmin(r, x) : (mode == 1 ? fmax(r, x) : r + x);
        }
        r = mode == 0 ? wave_min(r) : (mode == 1 ? wave_max(r) : wave_sum(r));
        if (threadIdx.x % kWave == 0) sh[threadIdx.x / kWave] = r;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = sh[0];
            for (int k = 1; k < kBlock / kWave; ++k)
                t = mode == 0 ? fmin(t, sh[k]) : (mode == 1 ? fmax(t, sh[k]) : t + sh[k]);
            out[c] = t;
        }
        __syncthreads();
    }
}

__global__ void __launch_bounds__(kBlock) k_finalize(const double *part, int nparts, int w,
                                                     const int32_t *modes, double *out) {
    finalize_block(part, nparts, w, modes, out);
}

template <typename R> __global__ void k_fill(R *p, int64_t n, R v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

// AABB + metrics over non-INACTIVE particles (rebuildGridFromHostBounds
// :572-577 and updateMetrics :855-868).
template <typename R>
__global__ void __launch_bounds__(kBlock) k_reduce(Particles<R> P, double ref_temp, double *part) {
    double v[kRed];
    stats_identity(v, ref_temp);
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < P.n; i += gridDim.x * kBlock) {
        if (P.status[i] == kInactive) continue;
        double x = P.x[i], y = P.y[i], z = P.z[i];
        v[0] = fmin(v[0], x); v[1] = fmin(v[1], y); v[2] = fmin(v[2], z);
        v[3] = fmin(v[3], -x); v[4] = fmin(v[4], -y); v[5] = fmin(v[5], -z);
        double vm = (double)von_mises<R>(P.s0[i], P.s1[i], P.s2[i], P.s3[i], P.s4[i], P.s5[i]);
        v[6] = fmax(v[6], vm);
        v[7] = fmax(v[7], (double)P.T[i]);
        double vx = P.vx[i], vy = P.vy[i], vz = P.vz[i];
        v[8] += 0.5 * (double)(P.m ? P.m[i] : P.m0) * (vx * vx + vy * vy + vz * vz);
    }
    block_reduce_stats(v, part, (int)blockIdx.x);
}

// ---------------------------------------------------------- grid update ---

// Where the grid update finds this step's P2G sums: in the tile windows,
// part[(tile * WN + window node) * 8 + channel], written whole by every tile
// (zeros when it has no particles), plus, when *acc_flag is set, the node
// accumulators (loose particles, and in the atomic mode particles scattered
// outside their tile's window).
template <typename R> struct TileParts {
    gp<const R> part;
    gp<const int32_t> acc_flag;
    TileGeom tg;
    int32_t quad;                // windows of the quadratic stencil (QX x QY x QZ)
};

// The tiles whose window holds node coordinate i along one axis (tiles start
// at cell `off`), lower tile first, with the window-local coordinate: node i
// is local T of tile q - 1 and local 0 of tile q when i - off = q * T.
__device__ __forceinline__ int window_tiles(int i, int off, int nt, int T, int tile[2], int local[2]) {
    const int r = i - off;
    if (r < 0) return 0;
    const int q = r / T, l = r - q * T;
    int k = 0;
    if (l == 0 && q >= 1 && q - 1 < nt) { tile[k] = q - 1; local[k] = T; ++k; }
    if (q < nt) { tile[k] = q; local[k] = l; ++k; }
    return k;
}

// The same for the quadratic windows, which start one node before the tile
// and end two after it: node i is local l = i - off + 1 - q * T of tile q
// for l in [0, T + 2], so of at most two tiles when T >= 3.
__device__ __forceinline__ int window_tiles_q(int i, int off, int nt, int T, int tile[2], int local[2]) {
    const int r1 = i - off + 1;
    if (r1 < 0) return 0;
    const int q = r1 / T, l = r1 - q * T;
    int k = 0;
    if (l <= 2 && q >= 1 && q - 1 < nt) { tile[k] = q - 1; local[k] = l + T; ++k; }
    if (q < nt) { tile[k] = q; local[k] = l; ++k; }
    return k;
}

// This step's 8 P2G channel sums of node (i, j, k), in a fixed order:
// tiles by increasing z, then y, then x tile index, then the accumulators.
template <typename R>
__device__ __forceinline__ void gather_tile_sums(const TileParts<R> &tp, int i, int j, int k, R c[8]) {
    int tx[2], lx[2], ty[2], ly[2], tz[2], lz[2];
    const bool q = tp.quad != 0;
    const int nxs = q ? window_tiles_q(i, tp.tg.ox, tp.tg.ntx, TX, tx, lx) : window_tiles(i, tp.tg.ox, tp.tg.ntx, TX, tx, lx);
    const int nys = q ? window_tiles_q(j, tp.tg.oy, tp.tg.nty, TY, ty, ly) : window_tiles(j, tp.tg.oy, tp.tg.nty, TY, ty, ly);
    const int nzs = q ? window_tiles_q(k, tp.tg.oz, tp.tg.ntz, TZ, tz, lz) : window_tiles(k, tp.tg.oz, tp.tg.ntz, TZ, tz, lz);
    const int ex = q ? QX : WX, ey = q ? QY : WY, en = q ? QN : WN;
#pragma unroll
    for (int ch = 0; ch < 8; ++ch) c[ch] = R(0);
    // at most two tiles per axis; unrolled so that the loads of every
    // window are in flight together
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int d = 0; d < 2; ++d) {
                if (a >= nzs || b >= nys || d >= nxs) continue;
                // an empty tile's window holds zeros (the particle pass
                // stores them), so no emptiness test waits on a load
                const int t = (tz[a] * tp.tg.nty + ty[b]) * tp.tg.ntx + tx[d];
                // node-major window records: 8 channels, 32 bytes per node
                const gp<const vec4_t<R>> src =
                    reinterpret_cast<const vec4_t<R> *>((const R *)tp.part + (int64_t)t * (8 * en)) +
                    2 * ((lz[a] * ey + ly[b]) * ex + lx[d]);
                const vec4_t<R> u = src[0], v = src[1];
                c[0] += u.x; c[1] += u.y; c[2] += u.z; c[3] += u.w;
                c[4] += v.x; c[5] += v.y; c[6] += v.z; c[7] += v.w;
            }
}

// mpmGridUpdateKernel, MPMWorkpieceSolver.cu:136-193 (point samples) and its
// analytic-SDF form with tool velocity and Coulomb friction, for node
// (i, j, k) = idx.  Takes the P2G sums c[] of the tile windows
// (gather_tile_sums) and adds the accumulators when `tp` flags them,
// zeroing what it has read; writes node velocity, mass and T*m.  When
// `direct_T` is set (no diffusion pass follows) it also writes the gather
// temperature.  Returns the node's contact-force contribution.
template <typename R>
__device__ __forceinline__ double grid_node(const GridView<R> &g, const Grid<R> &G, const TileParts<R> &tp,
                                            const Material<R> &M, const Tool &tool, R dt, int direct_T, int64_t idx,
                                            int i, int j, int k, R c[8]) {
    double cf = 0.0;
    {
        if (*tp.acc_flag) {
#pragma unroll
            for (int ch = 0; ch < 8; ++ch) {
                c[ch] += G.acc[ch * g.nn + idx];
                G.acc[ch * g.nn + idx] = R(0);
            }
        }
        if (G.snap) {
#pragma unroll
            for (int ch = 0; ch < 8; ++ch) G.snap[ch * g.nn + idx] = c[ch];
        }
        const R mass = c[0];
        R vx = c[1], vy = c[2], vz = c[3];
        const R fxn = c[4], fyn = c[5], fzn = c[6];
        R tm = c[7];
        bool live = mass > R(1e-18);
        if (live) {
            vx = vdiv(vx, mass); vy = vdiv(vy, mass); vz = vdiv(vz, mass);
            if (k <= g.kfix) {
                vx = vy = vz = R(0);
            } else {
                vx += vdiv(fxn, mass) * dt;
                vy += vdiv(fyn, mass) * dt;
                vz += vdiv(fzn, mass) * dt;
                // node position nodePosition (:37-41), in FP64
                const double px = g.ox + i * g.hd, py = g.oy + j * g.hd, pz = g.oz + k * g.hd;
                if (tool.kind == kToolPoints) {
                    if (px >= tool.lo[0] && px <= tool.hi[0] && py >= tool.lo[1] && py <= tool.hi[1] &&
                        pz >= tool.lo[2] && pz <= tool.hi[2]) {
                        double r2 = tool.rc2;
                        double bdx = 0, bdy = 0, bdz = 0;
                        bool hit = false;
                        for (int s = 0; s < tool.nsamples; ++s) {
                            double dx = px - tool.sx[s], dy = py - tool.sy[s], dz = pz - tool.sz[s];
                            double d2 = dx * dx + dy * dy + dz * dz;
                            if (d2 < r2 && d2 > 1e-20) {
                                r2 = d2;
                                bdx = dx; bdy = dy; bdz = dz;
                                hit = true;
                            }
                        }
                        if (hit) {
                            double len = sqrt(bdx * bdx + bdy * bdy + bdz * bdz);
                            R nx = R(0), ny = R(0), nz = R(0);
                            if (len > 1e-12) { nx = (R)(bdx / len); ny = (R)(bdy / len); nz = (R)(bdz / len); }
                            R vn = vx * nx + vy * ny + vz * nz;
                            if (vn < R(0)) {
                                cf += (double)fabs(mass * vn / rmax(dt, R(1e-15)));
                                vx -= vn * nx; vy -= vn * ny; vz -= vn * nz;
                                R heat = R(0.5) * mass * vn * vn * M.c_damp;
                                R dT = heat / rmax(mass * M.cp, R(1e-12));
                                if (M.clamp) dT = rmin(dT, M.dT_max);
                                tm += dT * mass;
                            }
                        }
                    }
                } else if (tool.kind == kToolWedge || tool.kind == kToolVoxel ||
                           (tool.kind == kToolDrill && !drill_far(tool, px, py, pz))) {
                    const double phi = tool_phi(tool, px, py, pz);
                    if (phi < tool.band) {
                        const double e = tool.eps;
                        double gx = (tool_phi_grad(tool, px + e, py, pz) - tool_phi_grad(tool, px - e, py, pz)) /
                                    (2.0 * e);
                        double gy = (tool_phi_grad(tool, px, py + e, pz) - tool_phi_grad(tool, px, py - e, pz)) /
                                    (2.0 * e);
                        double gz = (tool_phi_grad(tool, px, py, pz + e) - tool_phi_grad(tool, px, py, pz - e)) /
                                    (2.0 * e);
                        double len = sqrt(gx * gx + gy * gy + gz * gz);
                        if (len > 1e-12) {
                            R nx = (R)(gx / len), ny = (R)(gy / len), nz = (R)(gz / len);
                            R tvx = (R)(tool.vx - tool.omega * (py - tool.cy));
                            R tvy = (R)(tool.vy + tool.omega * (px - tool.cx));
                            R tvz = (R)tool.vz;
                            R rx = vx - tvx, ry = vy - tvy, rz = vz - tvz;
                            R vn = rx * nx + ry * ny + rz * nz;
                            if (vn < R(0)) {
                                cf += (double)fabs(mass * vn / rmax(dt, R(1e-15)));
                                R ux = rx - vn * nx, uy = ry - vn * ny, uz = rz - vn * nz;
                                R ut2 = ux * ux + uy * uy + uz * uz;
                                R ut = sqrt(ut2);
                                R sc = R(1);
                                const R mu = (R)tool.mu;
                                if (mu > R(0) && ut > R(0)) sc = rmax(R(0), R(1) + mu * vn / ut);
                                vx = tvx + sc * ux;
                                vy = tvy + sc * uy;
                                vz = tvz + sc * uz;
                                R removed = vn * vn + (R(1) - sc * sc) * ut2;
                                R heat = R(0.5) * mass * removed * M.c_damp;
                                R dT = heat / rmax(mass * M.cp, R(1e-12));
                                if (M.clamp) dT = rmin(dT, M.dT_max);
                                tm += dT * mass;
                            }
                        }
                    }
                }
            }
        }
        vec4_t<R> nd;
        nd.x = vx; nd.y = vy; nd.z = vz;
        nd.w = live ? (direct_T ? vdiv(tm, mass) : R(0)) : R(__builtin_nan(""));
        G.node[idx] = nd;
        G.gm[idx] = mass;
        G.gtm[idx] = tm;
    }
    return cf;
}

// The grid update, one thread per node in node order (x fastest), block
// `blk` of the grid.  Contact-force partial per wave of nodes, summed in a
// fixed order.
template <typename R>
__device__ __forceinline__ void grid_update_block(const GridView<R> &g, const Grid<R> &G, const TileParts<R> &tp,
                                                  const Material<R> &M, const Tool &tool, R dt, int direct_T,
                                                  double *force_part, int nwaves, int blk) {
    double cf = 0.0;
    // node counts are below 2^31 (the grid's index type is int32 per axis
    // and the binning keys are 32-bit); the block size is any multiple of
    // the wave (the autotuner picks it)
    const int idx = blk * (int)blockDim.x + (int)threadIdx.x;
    if (idx < g.nn) {
        const int row = idx / g.nx;
        const int i = idx - row * g.nx;
        const int k = row / g.ny, j = row - k * g.ny;
        R c[8];
        gather_tile_sums(tp, i, j, k, c);
        cf = grid_node(g, G, tp, M, tool, dt, direct_T, idx, i, j, k, c);
    }
    // the contact force: one fixed-order partial per wave of 64 nodes, the
    // same partials for every block size and block order
    cf = wave_sum(cf);
    const int wv = idx / kWave;
    if (threadIdx.x % kWave == 0 && wv < nwaves) force_part[wv] = cf;
}

// mpmDiffuseTemperatureKernel, MPMWorkpieceSolver.cu:208-262, Jacobi form:
// reads the post-update field (gm, gtm), writes the gather temperature.
// kDiffuseNodes nodes per thread (node blk * kDiffuseBlock + r * kBlock +
// tid): every node's 14 loads are issued before any node's update.
template <typename R>
__device__ __forceinline__ void diffuse_block(const GridView<R> &g, const Grid<R> &G, R alpha, R dt, int blk) {
    const int first = blk * kDiffuseBlock + (int)threadIdx.x;
    const int nxy = g.nx * g.ny;
    R mass[kDiffuseNodes], tm[kDiffuseNodes], ntm[kDiffuseNodes][6], nm[kDiffuseNodes][6];
    unsigned has[kDiffuseNodes];
#pragma unroll
    for (int r = 0; r < kDiffuseNodes; ++r) {
        const int idx = first + r * kBlock;
        has[r] = 0;
        if (idx >= g.nn) continue;
        const int t = idx / g.nx;
        const int i = idx - t * g.nx;
        const int k = t / g.ny, j = t - k * g.ny;
        // the six neighbours' (T m, m), read together with the node's own (a
        // missing neighbour reads the node itself and is left out)
        const int nb[6] = {i > 0 ? idx - 1 : idx, i < g.nx - 1 ? idx + 1 : idx,
                           j > 0 ? idx - g.nx : idx, j < g.ny - 1 ? idx + g.nx : idx,
                           k > 0 ? idx - nxy : idx, k < g.nz - 1 ? idx + nxy : idx};
        has[r] = (i > 0 ? 1u : 0u) | (i < g.nx - 1 ? 2u : 0u) | (j > 0 ? 4u : 0u) | (j < g.ny - 1 ? 8u : 0u) |
                 (k > 0 ? 16u : 0u) | (k < g.nz - 1 ? 32u : 0u) | 64u;
        mass[r] = G.gm[idx];
        tm[r] = G.gtm[idx];
#pragma unroll
        for (int q = 0; q < 6; ++q) { ntm[r][q] = G.gtm[nb[q]]; nm[r][q] = G.gm[nb[q]]; }
    }
    const R inv_h2 = (R)(1.0 / (g.hd * g.hd));
#pragma unroll
    for (int r = 0; r < kDiffuseNodes; ++r) {
        const int idx = first + r * kBlock;
        if (!(has[r] & 64u)) continue;
        if (!(mass[r] > R(1e-18))) continue;  // empty node keeps its NaN marker
        R t = tm[r];
        if (mass[r] >= R(1e-15)) {
            const R T = vdiv(t, mass[r]);
            R lap = R(0);
#pragma unroll
            for (int q = 0; q < 6; ++q)
                if (has[r] & (1u << q)) lap += (vdiv(ntm[r][q], rmax(nm[r][q], R(1e-30))) - T);
            const R T_new = T + alpha * dt * lap * inv_h2;
            t = rmax(T_new, R(0)) * mass[r];
        }
        G.node[idx].w = vdiv(t, mass[r]);
    }
}

// ------------------------------------------------------- batched launches --
//
// One launch serves many solvers: work item w owns blocks [block0, block0 +
// blocks) of the launch and carries the arguments the single-solver kernel
// would get.  A block finds its item by binary search over block0 (uniform
// per block) and runs exactly the single-solver body on its local block
// index, so a solver's results do not depend on what it is batched with.

// Work item of block `b`: the last w with items[w].block0 <= b.
template <typename Item> __device__ __forceinline__ int find_item(const Item *items, int nitems, int b) {
    int lo = 0, hi = nitems - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (items[mid].block0 <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// The same over the diffusion blocks of grid work items (dblock0).
template <typename Item> __device__ __forceinline__ int find_item_d(const Item *items, int nitems, int b) {
    int lo = 0, hi = nitems - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (items[mid].dblock0 <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <typename R> struct GridWork {
    GridView<R> g;
    Grid<R> G;
    TileParts<R> tp;
    Material<R> M;
    Tool tool;
    R dt, alpha;
    int32_t direct_T, diffuse;
    double *force_part;
    int32_t block0, blocks;    // grid-update blocks of the item
    int32_t dblock0, dblocks;  // diffusion blocks of the item (post pass)
    int32_t fwaves, pad_;      // contact-force partials (one per wave of nodes)
};

// Block order of a launch: 0 = blockIdx.x (workgroups round-robin over the
// dies), 1 = xcd_block() (one contiguous range per die).  A bijection either
// way: the autotuner's choice changes no result.
__device__ __forceinline__ int launch_block(int order) { return order ? xcd_block() : (int)blockIdx.x; }

template <typename R>
__global__ void __launch_bounds__(kBlock) k_grid_update_batch(const GridWork<R> *__restrict__ items, int nitems,
                                                              int order) {
    const int b = launch_block(order);
    const int w = find_item(items, nitems, b);
    const GridWork<R> &W = items[w];
    grid_update_block(W.g, W.G, W.tp, W.M, W.tool, W.dt, W.direct_T, W.force_part, W.fwaves, b - W.block0);
}

// The contact-force partials of an item are summed in chunks of kSumChunk,
// one block per chunk, each in a fixed order; the chunk sums are added, in
// order, only when the force is read (k_finalize), which keeps the step free
// of a dependent second pass.  Chunk 0's block also resets the item's
// P2G-pass flags when `counters` is set (counters[1] and [2], read by the
// grid update, written by the next particle pass).
constexpr int kSumChunk = 16 * kBlock;

struct SumWork {
    const double *part;
    int32_t nparts, chunks;    // partials; ceil(nparts / kSumChunk)
    int32_t block0, pad_;      // first sum block of the item
    const int32_t *modes;
    double *chunk;             // chunks chunk sums
    int32_t *counters;
};

__device__ __forceinline__ void sum_block(const SumWork &W, int k) {
    const int first = k * kSumChunk;
    finalize_block(W.part + first, min(kSumChunk, W.nparts - first), 1, W.modes, W.chunk + k);
    if (k == 0 && W.counters && threadIdx.x == 0) { W.counters[1] = 0; W.counters[2] = 0; }
}

// The pass after the grid update, one launch: blocks [0, sblocks) sum the
// members' contact-force chunks, blocks sblocks + d diffuse the members'
// grids (items with diffusion).
template <typename R>
__global__ void __launch_bounds__(kBlock) k_post_batch(const GridWork<R> *__restrict__ items, int nitems,
                                                       const SumWork *__restrict__ sums, int sblocks, int order) {
    const int b = launch_block(order);
    if (b < sblocks) {
        const int r = find_item(sums, nitems, b);
        sum_block(sums[r], b - sums[r].block0);
        return;
    }
    const int d = b - sblocks;
    const int w = find_item_d(items, nitems, d);
    const GridWork<R> &W = items[w];
    if (W.diffuse) diffuse_block(W.g, W.G, W.alpha, W.dt, d - W.dblock0);
}

// ------------------------------------------------------------ small ops ---

// SplitMix64 (Steele, Lea and Flood 2014): the counter-based generator of
// the jittered slab.
__host__ __device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Jittered slab on the device: particle pi = cell * 8 + k sits in sub-cell k
// of its cell at origin + (cell + sub / 2 + u / 2) * h (u from SplitMix64 on
// counter pi * 3 + axis), rounded to float; the same values as the host
// generator (the product and sum are rounded separately, no FMA).  Every
// other field gets its initial value.  m and V are not written (uniform).
template <typename R>
__global__ void k_init_slab(Particles<R> P, R *dens, uint64_t seed, int32_t cx, int32_t cy, double ox, double oy,
                            double oz, double h, double z_fix, R vx, R vy, R vz, R T0, R rho, int32_t n) {
    const int32_t pi = blockIdx.x * blockDim.x + threadIdx.x;
    if (pi >= n) return;
    const int32_t c = pi >> 3, k = pi & 7;
    const int32_t cell[3] = {c % cx, (c / cx) % cy, c / (cx * cy)};
    const double org[3] = {ox, oy, oz};
    double x[3];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        const uint64_t ctr = (uint64_t)((int64_t)pi * 3 + ax);
        const double u = (double)(splitmix64(seed ^ (ctr * 0xD1B54A32D192ED03ull)) >> 11) * 0x1.0p-53;
        const double t = ((double)cell[ax] + 0.5 * ((k >> ax) & 1)) + 0.5 * u;
        x[ax] = (double)(float)__dadd_rn(org[ax], __dmul_rn(t, h));
    }
    const bool fixed = x[2] <= z_fix;
    P.x[pi] = x[0]; P.y[pi] = x[1]; P.z[pi] = x[2];
    P.vx[pi] = fixed ? R(0) : vx; P.vy[pi] = fixed ? R(0) : vy; P.vz[pi] = fixed ? R(0) : vz;
    P.s0[pi] = R(0); P.s1[pi] = R(0); P.s2[pi] = R(0); P.s3[pi] = R(0); P.s4[pi] = R(0); P.s5[pi] = R(0);
    if (P.a0) { P.a0[pi] = R(0); P.a1[pi] = R(0); P.a2[pi] = R(0); P.a3[pi] = R(0); P.a4[pi] = R(0); P.a5[pi] = R(0); }
    if (P.fx) { P.fx[pi] = R(0); P.fy[pi] = R(0); P.fz[pi] = R(0); }
    P.eps[pi] = R(0); P.rate[pi] = R(0); P.dmg[pi] = R(0); P.sres[pi] = R(0);
    P.T[pi] = T0;
    dens[pi] = rho;
    P.status[pi] = fixed ? kFixed : kActive;
    P.id[pi] = pi;
}

// applyMPMExternalForceKernel, MPMWorkpieceSolver.cu:402-414
template <typename R>
__global__ void k_external_force(Particles<R> P, double cx, double cy, double cz, double r2, R fx, R fy, R fz) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n || P.status[i] != kActive) return;
    double dx = P.x[i] - cx, dy = P.y[i] - cy, dz = P.z[i] - cz;
    if (dx * dx + dy * dy + dz * dz <= r2) {
        P.fx[i] += fx;
        P.fy[i] += fy;
        P.fz[i] += fz;
    }
}

// setParticleTempKernel, MPMWorkpieceSolver.cu:899-904
template <typename R> __global__ void k_set_temperature(R *T, const double *src, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) T[i] = (R)src[i];
}

// Nearest particle to a point: per-block (d2, index) with the lowest index
// winning ties, as the host loops of getTemperatureAt/applyHeatFlux do.
template <typename R>
__global__ void __launch_bounds__(kBlock) k_nearest(Particles<R> P, double qx, double qy, double qz,
                                                    double *best_d2, int32_t *best_i) {
    __shared__ double sd[kBlock];
    __shared__ int32_t si[kBlock];
    double bd = 1e30;
    int32_t bi = -1;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < P.n; i += gridDim.x * kBlock) {
        double dx = (double)P.x[i] - qx, dy = (double)P.y[i] - qy, dz = (double)P.z[i] - qz;
        double d2 = dx * dx + dy * dy + dz * dz;
        if (d2 < bd) { bd = d2; bi = i; }
    }
    sd[threadIdx.x] = bd;
    si[threadIdx.x] = bi;
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            double d = sd[threadIdx.x + o];
            int32_t k = si[threadIdx.x + o];
            if (k >= 0 && (d < sd[threadIdx.x] || (d == sd[threadIdx.x] && (si[threadIdx.x] < 0 || k < si[threadIdx.x])))) {
                sd[threadIdx.x] = d;
                si[threadIdx.x] = k;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        best_d2[blockIdx.x] = sd[0];
        best_i[blockIdx.x] = si[0];
    }
}

// Combine the per-block nearest candidates (single block) and, for
// applyHeatFlux (:928-932), heat the winner when within one cell.
template <typename R>
__global__ void k_nearest_finalize(Particles<R> P, const double *best_d2, const int32_t *best_i, int nparts,
                                   double *out_d2, int32_t *out_i, int apply, double q, double cp, double melt,
                                   double h2) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double bd = 1e30;
    int32_t bi = -1;
    for (int b = 0; b < nparts; ++b) {
        double d = best_d2[b];
        int32_t k = best_i[b];
        if (k >= 0 && (d < bd || (d == bd && (bi < 0 || k < bi)))) { bd = d; bi = k; }
    }
    *out_d2 = bd;
    *out_i = bi;
    if (apply && bi >= 0 && bd <= h2) {
        double T = (double)P.T[bi];
        double m = (double)(P.m ? P.m[bi] : P.m0);
        P.T[bi] = (R)fmin(T + q / fmax(m * cp, 1e-12), melt);
    }
}

// The particle arrays of one permutation, dst[i] = src[perm[i]], gathered
// by one launch: array blockIdx.y of the table, 4- or 8-byte elements.
constexpr int kMaxGatherFields = 48;
struct GatherTable {
    void *dst[kMaxGatherFields];
    const void *src[kMaxGatherFields];
    int32_t wide[kMaxGatherFields];  // 1: 8-byte elements
    int32_t nf;
};

__global__ void __launch_bounds__(kBlock) k_gather_fields(GatherTable t, const int32_t *perm, int n) {
    const int f = blockIdx.y;
    const int i = blockIdx.x * kBlock + (int)threadIdx.x;
    if (f >= t.nf || i >= n) return;
    const int j = perm[i];
    if (t.wide[f]) static_cast<uint64_t *>(t.dst[f])[i] = static_cast<const uint64_t *>(t.src[f])[j];
    else static_cast<uint32_t *>(t.dst[f])[i] = static_cast<const uint32_t *>(t.src[f])[j];
}

}  // namespace mpmw
