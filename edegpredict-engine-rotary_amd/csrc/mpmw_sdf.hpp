// mpmw_sdf.hpp -- voxel signed distance field of a tool mesh, built on the
// device.
//
// The reference builds it on the host (ToolSDF::build / computeSDFCPU,
// ToolSDF.cu:212-532): squared distance of every voxel centre to the
// triangles whose box (grown by -1 / +2 voxels) holds it, the square root,
// a sign within two voxels of the surface from the nearest triangle's
// normal, three flood-fill passes for the rest.  Here the first two steps
// run one workgroup per triangle (atomicMin on the FP32 bit patterns, which
// order like the values for non-negative floats) and one thread per voxel;
// the flood fill, whose in-place passes depend on the visiting order, stays
// on the host.  The arithmetic is the reference's single precision, with no
// FMA contraction, so the grid equals the host build bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mpmw {

struct SdfTri {
    float a[3], b[3], c[3], n[3], lo[3], hi[3];
    int32_t i0[3], i1[3];  // voxel index box of the scatter
};

// pointTriDistanceSq (ToolSDF.cu:14-89): the closest-point regions of
// Ericson, Real-Time Collision Detection 5.1.5, with the reference's
// interior case a + u (b - a) + v (c - a), u = 1 - v - w.
__device__ __forceinline__ float sdf_point_tri_d2(const float p[3], const float a[3], const float b[3],
                                                  const float c[3], float q[3]) {
#pragma clang fp contract(off)
    float ab[3], ac[3], ap[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        ab[k] = b[k] - a[k];
        ac[k] = c[k] - a[k];
        ap[k] = p[k] - a[k];
    }
    const float d1 = ab[0] * ap[0] + ab[1] * ap[1] + ab[2] * ap[2];
    const float d2 = ac[0] * ap[0] + ac[1] * ap[1] + ac[2] * ap[2];
    if (d1 <= 0.0f && d2 <= 0.0f) {
#pragma unroll
        for (int k = 0; k < 3; ++k) q[k] = a[k];
    } else {
        float bp[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) bp[k] = p[k] - b[k];
        const float d3 = ab[0] * bp[0] + ab[1] * bp[1] + ab[2] * bp[2];
        const float d4 = ac[0] * bp[0] + ac[1] * bp[1] + ac[2] * bp[2];
        const float vc = d1 * d4 - d3 * d2;
        float cp[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) cp[k] = p[k] - c[k];
        const float d5 = ab[0] * cp[0] + ab[1] * cp[1] + ab[2] * cp[2];
        const float d6 = ac[0] * cp[0] + ac[1] * cp[1] + ac[2] * cp[2];
        const float vb = d5 * d2 - d1 * d6;
        const float va = d3 * d6 - d5 * d4;
        if (d3 >= 0.0f && d4 <= d3) {
#pragma unroll
            for (int k = 0; k < 3; ++k) q[k] = b[k];
        } else if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
            const float v = d1 / (d1 - d3);
#pragma unroll
            for (int k = 0; k < 3; ++k) q[k] = a[k] + v * ab[k];
        } else if (d6 >= 0.0f && d5 <= d6) {
#pragma unroll
            for (int k = 0; k < 3; ++k) q[k] = c[k];
        } else if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
            const float w = d2 / (d2 - d6);
#pragma unroll
            for (int k = 0; k < 3; ++k) q[k] = a[k] + w * ac[k];
        } else if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
            const float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
#pragma unroll
            for (int k = 0; k < 3; ++k) q[k] = b[k] + w * (c[k] - b[k]);
        } else {
            const float den = 1.0f / (va + vb + vc);
            const float v = vb * den, w = vc * den, u = 1.0f - v - w;
#pragma unroll
            for (int k = 0; k < 3; ++k) q[k] = a[k] + u * (b[k] - a[k]) + v * (c[k] - a[k]);
        }
    }
    const float dx = p[0] - q[0], dy = p[1] - q[1], dz = p[2] - q[2];
    return dx * dx + dy * dy + dz * dz;
}

__device__ __forceinline__ void sdf_center(float ox, float oy, float oz, float vox, int ix, int iy, int iz,
                                           float c[3]) {
#pragma clang fp contract(off)
    c[0] = ox + ((float)ix + 0.5f) * vox;
    c[1] = oy + ((float)iy + 0.5f) * vox;
    c[2] = oz + ((float)iz + 0.5f) * vox;
}

__global__ void k_sdf_fill(uint32_t *p, int64_t n, uint32_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// One workgroup per triangle: the squared distance of every voxel of its
// index box, kept as the smallest per voxel.
__global__ void __launch_bounds__(256) k_sdf_scatter(const SdfTri *tris, float ox, float oy, float oz, float vox,
                                                     int32_t nx, int32_t ny, uint32_t *d2bits) {
    const SdfTri &T = tris[blockIdx.x];
    const int bx = T.i1[0] - T.i0[0] + 1, by = T.i1[1] - T.i0[1] + 1, bz = T.i1[2] - T.i0[2] + 1;
    if (bx <= 0 || by <= 0 || bz <= 0) return;
    const int nbox = bx * by * bz;
    for (int w = threadIdx.x; w < nbox; w += blockDim.x) {
        const int ix = T.i0[0] + w % bx, iy = T.i0[1] + (w / bx) % by, iz = T.i0[2] + w / (bx * by);
        float c[3], q[3];
        sdf_center(ox, oy, oz, vox, ix, iy, iz, c);
        const float d2 = sdf_point_tri_d2(c, T.a, T.b, T.c, q);
        atomicMin(d2bits + (ix + (int64_t)nx * (iy + (int64_t)ny * iz)), __float_as_uint(d2));
    }
}

// One thread per voxel: the distance, negative within `band` of the
// surface when the first strictly nearest triangle (among those whose box
// grown by a voxel holds the centre) has its normal facing away.  Voxels no
// triangle reached keep FLT_MAX for the flood fill.
__global__ void k_sdf_sign(const SdfTri *tris, int32_t ntris, float ox, float oy, float oz, float vox,
                           int32_t nx, int32_t ny, int64_t total, float band, const uint32_t *d2bits, float *grid) {
#pragma clang fp contract(off)
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const float big = 3.402823466e+38f;
    const uint32_t u = d2bits[idx];
    if (u == __float_as_uint(big)) {
        grid[idx] = big;
        return;
    }
    const float dist = sqrtf(__uint_as_float(u));
    if (dist > band) {
        grid[idx] = dist;
        return;
    }
    const int ix = (int)(idx % nx), iy = (int)((idx / nx) % ny), iz = (int)(idx / ((int64_t)nx * ny));
    float c[3];
    sdf_center(ox, oy, oz, vox, ix, iy, iz, c);
    float best = big, bn[3] = {0, 0, 0}, bq[3] = {0, 0, 0};
    for (int t = 0; t < ntris; ++t) {
        const SdfTri &T = tris[t];
        if (c[0] < T.lo[0] - vox || c[0] > T.hi[0] + vox || c[1] < T.lo[1] - vox || c[1] > T.hi[1] + vox ||
            c[2] < T.lo[2] - vox || c[2] > T.hi[2] + vox)
            continue;
        float q[3];
        const float d2 = sdf_point_tri_d2(c, T.a, T.b, T.c, q);
        if (d2 < best) {
            best = d2;
            bn[0] = T.n[0]; bn[1] = T.n[1]; bn[2] = T.n[2];
            bq[0] = q[0]; bq[1] = q[1]; bq[2] = q[2];
        }
    }
    float out = dist;
    if (best < big) {
        const float dot = (c[0] - bq[0]) * bn[0] + (c[1] - bq[1]) * bn[1] + (c[2] - bq[2]) * bn[2];
        if (dot < 0.0f) out = -dist;
    }
    grid[idx] = out;
}

}  // namespace mpmw
