/*
 * mpm_oracle.c -- CPU FP64 restatement of MPMWorkpieceSolver::step.
 *
 * TEST INFRASTRUCTURE ONLY (the checker for the HIP solver; never shipped,
 * never called by the product library).
 *
 * Reference: Up25-pathan/edegpredict-engine-rotary, src/MPMWorkpieceSolver.cu
 * and include/MPMWorkpieceSolver.cuh.  Every function cites the lines it
 * restates.  Deliberate, documented differences from the reference binary:
 *   - the grid heat diffusion is Jacobi (reads the pre-diffusion field, writes
 *     a new one).  The reference kernel (MPMWorkpieceSolver.cu:208-262) updates
 *     temperatureMass in place while neighbouring threads read it, a data race
 *     whose outcome depends on scheduling; its own comment (:195-207) states the
 *     Jacobi update T_new = T + alpha*dt*lap(T).
 *   - compaction keeps the surviving particles in their previous order; the
 *     reference's atomic-counter compaction (:464-473) leaves them in an
 *     unspecified order.
 *   - the analytic tool modes (wedge, drill) and tool velocity/friction are the
 *     grid-node form of ContactSolver.cu:797-923; in point-sample mode the grid
 *     update is exactly :136-193.
 *
 * Parity pins for each function are listed in tests/test_oracle.py.
 */
#include "mpm_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

struct orc_solver {
    orc_params prm;
    orc_particle *p;
    int n, cap;
    orc_grid g;
    orc_node *grid;
    orc_node *snap;       /* post-P2G copy of the grid */
    int nnodes;
    double *tm_new;       /* Jacobi buffer for the diffusion */
    double *tool_xyz;     /* point samples */
    int ntool;
    orc_tool_desc tool;
    orc_tool_pose pose;
    orc_sdf sdf;          /* voxel tool (kind ORC_TOOL_VOXEL) */
    double tip[3];
    double time;
    int64_t step;
    double max_stress, max_temp, ke, contact_force;
    int compactions;
    int transfer;         /* ORC_TRANSFER_PIC / ORC_TRANSFER_APIC */
    int apic_textbook;    /* APIC with 4/h^2 and no decay */
    int fp32_emulation;   /* sensitivity probe: FP32 P2G sums and stored state */
};

static int g_threads = 1;
void orc_set_threads(int n) { g_threads = n < 1 ? 1 : n; }

static double clampd(double v, double lo, double hi) { return fmax(lo, fmin(hi, v)); }  /* :14-16 */

/* equivalentStress, MPMWorkpieceSolver.cu:18-24 */
double orc_equivalent_stress(double xx, double yy, double zz, double xy, double xz, double yz) {
    double s1 = xx - yy, s2 = yy - zz, s3 = zz - xx;
    return sqrt(0.5 * (s1 * s1 + s2 * s2 + s3 * s3 + 6.0 * (xy * xy + xz * xz + yz * yz)));
}

/* johnsonCookStress, MPMWorkpieceSolver.cu:43-50 */
double orc_jc_flow(const orc_params *m, double eps_p, double eps_rate, double temp) {
    double strainTerm = m->jc_A + m->jc_B * pow(fmax(eps_p, 0.0), m->jc_n);
    double rateRatio = fmax(eps_rate / fmax(m->ref_strain_rate, 1e-9), 1.0);
    double rateTerm = 1.0 + m->jc_C * log(rateRatio);
    double tStar = clampd((temp - m->ref_temp) / fmax(m->melt_temp - m->ref_temp, 1.0), 0.0, 1.0);
    double thermalTerm = fmax(0.01, 1.0 - pow(tStar, m->jc_m));
    return strainTerm * rateTerm * thermalTerm;
}

/* computeWeights, MPMWorkpieceSolver.cu:52-90 (gridIndex :26-28) */
void orc_weights(const orc_grid *g, double x, double y, double z,
                 int32_t idx[8], double w[8], double grad[24]) {
    double rx = (x - g->minx) / g->cell;
    double ry = (y - g->miny) / g->cell;
    double rz = (z - g->minz) / g->cell;
    int bx = (int)floor(rx), by = (int)floor(ry), bz = (int)floor(rz);
    double fx = rx - bx, fy = ry - by, fz = rz - bz;
    int n = 0;
    for (int dz = 0; dz <= 1; ++dz)
        for (int dy = 0; dy <= 1; ++dy)
            for (int dx = 0; dx <= 1; ++dx, ++n) {
                int i = bx + dx, j = by + dy, k = bz + dz;
                if (i < 0 || j < 0 || k < 0 || i >= g->nx || j >= g->ny || k >= g->nz) {
                    idx[n] = -1;
                    w[n] = 0.0;
                    grad[3 * n] = grad[3 * n + 1] = grad[3 * n + 2] = 0.0;
                    continue;
                }
                double wx = dx ? fx : 1.0 - fx;
                double wy = dy ? fy : 1.0 - fy;
                double wz = dz ? fz : 1.0 - fz;
                double dwx = dx ? 1.0 / g->cell : -1.0 / g->cell;
                double dwy = dy ? 1.0 / g->cell : -1.0 / g->cell;
                double dwz = dz ? 1.0 / g->cell : -1.0 / g->cell;
                idx[n] = (k * g->ny + j) * g->nx + i;
                w[n] = wx * wy * wz;
                grad[3 * n] = dwx * wy * wz;
                grad[3 * n + 1] = wx * dwy * wz;
                grad[3 * n + 2] = wx * wy * dwz;
            }
}

/* getWeightsAndGradients, MPMSolver.cu:20-34, along one axis: r is the
 * position in cells, base = floor(r - 0.5), d0 = r - base in [0.5, 1.5). */
static void quad_axis(double r, double inv, int *base, double w[3], double dw[3]) {
    int b = (int)floor(r - 0.5);
    double d0 = r - b, d1 = d0 - 1.0, d2 = d0 - 2.0;
    w[0] = 0.5 * (1.5 - d0) * (1.5 - d0);
    w[1] = 0.75 - d1 * d1;
    w[2] = 0.5 * (1.5 + d2) * (1.5 + d2);
    dw[0] = -(1.5 - d0) * inv;
    dw[1] = -2.0 * d1 * inv;
    dw[2] = (1.5 + d2) * inv;
    *base = b;
}

/* The 27-node stencil (p2gKernel :370-400 / g2pKernel :526-556 loop order:
 * x outer, z inner), positions relative to the grid origin scaled by
 * invDx = 1 / h as the reference forms x_frac (:21). */
void orc_weights_quadratic(const orc_grid *g, double x, double y, double z, int32_t idx[27], double w[27],
                           double grad[81], double dpos[81]) {
    const double inv = 1.0 / g->cell;
    int b[3];
    double wa[3][3], dwa[3][3];
    quad_axis((x - g->minx) * inv, inv, &b[0], wa[0], dwa[0]);
    quad_axis((y - g->miny) * inv, inv, &b[1], wa[1], dwa[1]);
    quad_axis((z - g->minz) * inv, inv, &b[2], wa[2], dwa[2]);
    int n = 0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k, ++n) {
                const int gx = b[0] + i, gy = b[1] + j, gz = b[2] + k;
                double *gr = &grad[3 * n], *dp = &dpos[3 * n];
                if (gx < 0 || gy < 0 || gz < 0 || gx >= g->nx || gy >= g->ny || gz >= g->nz) {
                    idx[n] = -1;
                    w[n] = 0.0;
                    gr[0] = gr[1] = gr[2] = dp[0] = dp[1] = dp[2] = 0.0;
                    continue;
                }
                idx[n] = (gz * g->ny + gy) * g->nx + gx;
                w[n] = wa[0][i] * wa[1][j] * wa[2][k];
                gr[0] = dwa[0][i] * wa[1][j] * wa[2][k];
                gr[1] = wa[0][i] * dwa[1][j] * wa[2][k];
                gr[2] = wa[0][i] * wa[1][j] * dwa[2][k];
                /* node position minus particle position (:385-387) */
                dp[0] = (gx * g->cell + g->minx) - x;
                dp[1] = (gy * g->cell + g->miny) - y;
                dp[2] = (gz * g->cell + g->minz) - z;
            }
}

/* rebuildGridFromHostBounds, MPMWorkpieceSolver.cu:578-584 (given the AABB) */
void orc_rebuild_params(double cell, const double mn[3], const double mx[3], orc_grid *g) {
    double pad = cell * 4.0;
    g->cell = cell;
    g->minx = mn[0] - pad;
    g->miny = mn[1] - pad;
    g->minz = mn[2] - pad;
    int nx = (int)ceil((mx[0] - mn[0] + 2 * pad) / cell) + 1;
    int ny = (int)ceil((mx[1] - mn[1] + 2 * pad) / cell) + 1;
    int nz = (int)ceil((mx[2] - mn[2] + 2 * pad) / cell) + 1;
    g->nx = nx > 4 ? nx : 4;
    g->ny = ny > 4 ? ny : 4;
    g->nz = nz > 4 ? nz : 4;
}

/* chip-fracture pseudo-random direction, MPMWorkpieceSolver.cu:709-712 */
void orc_chip_kick(int32_t id, double r[3]) {
    uint32_t seed = (uint32_t)((uint32_t)id * 2654435761U);
    r[0] = (double)((seed >> 16) & 0xFF) / 255.0 - 0.5;
    r[1] = (double)((seed >> 8) & 0xFF) / 255.0 - 0.5;
    r[2] = (double)(seed & 0xFF) / 255.0 - 0.5;
}

/* ---------------------------------------------------------------- tools -- */

/* World -> tool frame: translate by -center, rotate by -angle about z
 * (ContactSolver.cu:825-829 uses the same Z-rotation convention). */
static void to_tool_frame(const orc_tool_pose *p, double x, double y, double z, double q[3]) {
    double c = cos(p->angle), s = sin(p->angle);
    double dx = x - p->center[0], dy = y - p->center[1];
    q[0] = dx * c + dy * s;
    q[1] = -dx * s + dy * c;
    q[2] = z - p->center[2];
}

/* Wedge: rake plane through the edge (normal (cos g, 0, sin g)) intersected
 * with the flank plane (normal (-sin a, 0, -cos a)); cutting direction +x,
 * edge along y.  Interior where both plane distances are negative. */
static double phi_wedge(const orc_tool_desc *d, const double q[3]) {
    double g = d->rake_angle, a = d->clearance_angle;
    double pr = cos(g) * q[0] + sin(g) * q[2];
    double pf = -sin(a) * q[0] - cos(a) * q[2];
    return fmax(pr, pf);
}

/* Two-flute helical drill, tip at the origin, body along +z.
 * body = cone(half-angle = point/2) & cylinder(r <= R) & (z <= L);
 * two helical flute grooves of radius rf at distance df from the axis,
 * phase k*pi + z*tan(helix)/R, are subtracted. */
static double phi_drill(const orc_tool_desc *d, const double q[3]) {
    double r = sqrt(q[0] * q[0] + q[1] * q[1]);
    double beta = 0.5 * d->point_angle;
    double sb = sin(beta), cb = cos(beta);
    double t = r * sb + q[2] * cb;
    double cone = (t < 0.0) ? sqrt(r * r + q[2] * q[2]) : (r * cb - q[2] * sb);
    double body = fmax(fmax(cone, r - d->radius), q[2] - d->flute_length);
    double psi = q[2] * tan(d->helix_angle) / d->radius;
    double gmin = 1e300;
    for (int k = 0; k < 2; ++k) {
        double ang = psi + k * 3.14159265358979323846;
        double cx = d->flute_offset * cos(ang), cy = d->flute_offset * sin(ang);
        double ex = q[0] - cx, ey = q[1] - cy;
        double gk = sqrt(ex * ex + ey * ey) - d->flute_radius;
        gmin = fmin(gmin, gk);
    }
    return fmax(body, -gmin);
}

/* ---------------------------------------------------------- voxel SDF --- */

/* pointTriDistanceSqHost, ToolSDF.cu:92-162: the closest-point-on-triangle
 * regions of Ericson, Real-Time Collision Detection 5.1.5, in the
 * reference's float arithmetic, including its interior case, which forms the
 * point as a + u (b - a) + v (c - a) with u = 1 - v - w (:148-155). */
float orc_point_tri_d2(const float p[3], const float a[3], const float b[3], const float c[3], float q[3]) {
    float ab[3], ac[3], ap[3], bp[3], cp[3];
    for (int k = 0; k < 3; ++k) {
        ab[k] = b[k] - a[k];
        ac[k] = c[k] - a[k];
        ap[k] = p[k] - a[k];
    }
    float d1 = ab[0] * ap[0] + ab[1] * ap[1] + ab[2] * ap[2];
    float d2 = ac[0] * ap[0] + ac[1] * ap[1] + ac[2] * ap[2];
    if (d1 <= 0.0f && d2 <= 0.0f) {
        for (int k = 0; k < 3; ++k) q[k] = a[k];
        goto out;
    }
    for (int k = 0; k < 3; ++k) bp[k] = p[k] - b[k];
    float d3 = ab[0] * bp[0] + ab[1] * bp[1] + ab[2] * bp[2];
    float d4 = ac[0] * bp[0] + ac[1] * bp[1] + ac[2] * bp[2];
    if (d3 >= 0.0f && d4 <= d3) {
        for (int k = 0; k < 3; ++k) q[k] = b[k];
        goto out;
    }
    float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
        float v = d1 / (d1 - d3);
        for (int k = 0; k < 3; ++k) q[k] = a[k] + v * ab[k];
        goto out;
    }
    for (int k = 0; k < 3; ++k) cp[k] = p[k] - c[k];
    float d5 = ab[0] * cp[0] + ab[1] * cp[1] + ab[2] * cp[2];
    float d6 = ac[0] * cp[0] + ac[1] * cp[1] + ac[2] * cp[2];
    if (d6 >= 0.0f && d5 <= d6) {
        for (int k = 0; k < 3; ++k) q[k] = c[k];
        goto out;
    }
    float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
        float w = d2 / (d2 - d6);
        for (int k = 0; k < 3; ++k) q[k] = a[k] + w * ac[k];
        goto out;
    }
    float va = d3 * d6 - d5 * d4;
    if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
        float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        for (int k = 0; k < 3; ++k) q[k] = b[k] + w * (c[k] - b[k]);
        goto out;
    }
    {
        float den = 1.0f / (va + vb + vc);
        float v = vb * den, w = vc * den, u = 1.0f - v - w;
        for (int k = 0; k < 3; ++k) q[k] = a[k] + u * (b[k] - a[k]) + v * (c[k] - a[k]);
    }
out:;
    float dx = p[0] - q[0], dy = p[1] - q[1], dz = p[2] - q[2];
    return dx * dx + dy * dy + dz * dz;
}

typedef struct {
    float a[3], b[3], c[3], n[3], lo[3], hi[3];
} sdf_tri;

int orc_sdf_build(const double *verts, int nverts, const int32_t *tris, const double *normals, int ntris,
                  double padding, int voxels_per_axis, orc_sdf *out) {
    memset(out, 0, sizeof(*out));
    if (nverts <= 0 || ntris <= 0 || voxels_per_axis <= 0) return -1;
    /* build, :233-266 */
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (int i = 0; i < nverts; ++i)
        for (int k = 0; k < 3; ++k) {
            lo[k] = fminf(lo[k], (float)verts[3 * i + k]);
            hi[k] = fmaxf(hi[k], (float)verts[3 * i + k]);
        }
    const float pad = (float)padding;
    for (int k = 0; k < 3; ++k) { lo[k] -= pad; hi[k] += pad; }
    const float ext[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
    const float vox = fmaxf(ext[0], fmaxf(ext[1], ext[2])) / (float)voxels_per_axis;
    int dims[3];
    for (int k = 0; k < 3; ++k) {
        dims[k] = (int)(ext[k] / vox) + 2;
        if (dims[k] > 256) dims[k] = 256;
    }
    const int nx = dims[0], ny = dims[1], nz = dims[2];
    const long total = (long)nx * ny * nz;
    if (nx <= 0 || ny <= 0 || nz <= 0 || total <= 0) return -1;
    out->nx = nx; out->ny = ny; out->nz = nz;
    out->ox = lo[0]; out->oy = lo[1]; out->oz = lo[2];
    out->voxel = vox;
    /* buildFromVerts, :348-378 */
    sdf_tri *T = (sdf_tri *)malloc(sizeof(sdf_tri) * (size_t)ntris);
    for (int t = 0; t < ntris; ++t) {
        for (int k = 0; k < 3; ++k) {
            T[t].a[k] = (float)verts[3 * tris[3 * t + 0] + k];
            T[t].b[k] = (float)verts[3 * tris[3 * t + 1] + k];
            T[t].c[k] = (float)verts[3 * tris[3 * t + 2] + k];
            T[t].n[k] = (float)normals[3 * t + k];
        }
        for (int k = 0; k < 3; ++k) {
            T[t].lo[k] = fminf(T[t].a[k], fminf(T[t].b[k], T[t].c[k]));
            T[t].hi[k] = fmaxf(T[t].a[k], fmaxf(T[t].b[k], T[t].c[k]));
        }
    }
    float *G = (float *)malloc(sizeof(float) * (size_t)total);
    for (long i = 0; i < total; ++i) G[i] = FLT_MAX;
    const float org[3] = {out->ox, out->oy, out->oz};
    const float inv = 1.0f / vox;
#define SDF_CENTER(ix, iy, iz, c) do { (c)[0] = org[0] + ((float)(ix) + 0.5f) * vox; \
                                       (c)[1] = org[1] + ((float)(iy) + 0.5f) * vox; \
                                       (c)[2] = org[2] + ((float)(iz) + 0.5f) * vox; } while (0)
    /* squared distance to each triangle over its voxel box grown by one, :394-420 */
    for (int t = 0; t < ntris; ++t) {
        int i0[3], i1[3];
        for (int k = 0; k < 3; ++k) {
            i0[k] = (int)((T[t].lo[k] - org[k]) * inv) - 1;
            if (i0[k] < 0) i0[k] = 0;
            i1[k] = (int)((T[t].hi[k] - org[k]) * inv) + 2;
            if (i1[k] > dims[k] - 1) i1[k] = dims[k] - 1;
        }
        for (int iz = i0[2]; iz <= i1[2]; ++iz)
            for (int iy = i0[1]; iy <= i1[1]; ++iy)
                for (int ix = i0[0]; ix <= i1[0]; ++ix) {
                    float c[3], q[3];
                    SDF_CENTER(ix, iy, iz, c);
                    const float d2 = orc_point_tri_d2(c, T[t].a, T[t].b, T[t].c, q);
                    const long idx = ix + (long)nx * (iy + (long)ny * iz);
                    if (d2 < G[idx]) G[idx] = d2;
                }
    }
    /* distances, :424-428 */
    for (long i = 0; i < total; ++i)
        if (G[i] < FLT_MAX) G[i] = sqrtf(G[i]);
    /* sign within two voxels of the surface from the nearest triangle's
     * normal (first strictly nearest among triangles whose box grown by one
     * voxel holds the centre), :442-483 */
    const float band = vox * 2.0f;
    for (int iz = 0; iz < nz; ++iz)
        for (int iy = 0; iy < ny; ++iy)
            for (int ix = 0; ix < nx; ++ix) {
                const long idx = ix + (long)nx * (iy + (long)ny * iz);
                const float dist = G[idx];
                if (dist > band) continue;
                float c[3];
                SDF_CENTER(ix, iy, iz, c);
                float best = FLT_MAX, bn[3] = {0, 0, 0}, bq[3] = {0, 0, 0};
                for (int t = 0; t < ntris; ++t) {
                    if (c[0] < T[t].lo[0] - vox || c[0] > T[t].hi[0] + vox || c[1] < T[t].lo[1] - vox ||
                        c[1] > T[t].hi[1] + vox || c[2] < T[t].lo[2] - vox || c[2] > T[t].hi[2] + vox)
                        continue;
                    float q[3];
                    const float d2 = orc_point_tri_d2(c, T[t].a, T[t].b, T[t].c, q);
                    if (d2 < best) {
                        best = d2;
                        for (int k = 0; k < 3; ++k) { bn[k] = T[t].n[k]; bq[k] = q[k]; }
                    }
                }
                if (best < FLT_MAX) {
                    const float dot = (c[0] - bq[0]) * bn[0] + (c[1] - bq[1]) * bn[1] + (c[2] - bq[2]) * bn[2];
                    if (dot < 0.0f) G[idx] = -dist;
                }
            }
    /* unknown voxels take the smallest |value| of their 26 neighbours plus
     * half a voxel, three in-place passes over the interior, :489-520 */
    for (int pass = 0; pass < 3; ++pass)
        for (int iz = 1; iz < nz - 1; ++iz)
            for (int iy = 1; iy < ny - 1; ++iy)
                for (int ix = 1; ix < nx - 1; ++ix) {
                    const long idx = ix + (long)nx * (iy + (long)ny * iz);
                    if (G[idx] < FLT_MAX) continue;
                    float near = FLT_MAX;
                    for (int dz = -1; dz <= 1; ++dz)
                        for (int dy = -1; dy <= 1; ++dy)
                            for (int dx = -1; dx <= 1; ++dx) {
                                if (dx == 0 && dy == 0 && dz == 0) continue;
                                const long ni = (ix + dx) + (long)nx * ((iy + dy) + (long)ny * (iz + dz));
                                const float v = fabsf(G[ni]);
                                if (v < near) near = v;
                            }
                    if (near < FLT_MAX) G[idx] = near + vox * 0.5f;
                }
    /* the rest, :523-527 */
    const int mx = nx > ny ? (nx > nz ? nx : nz) : (ny > nz ? ny : nz);
    for (long i = 0; i < total; ++i)
        if (G[i] >= FLT_MAX) G[i] = vox * (float)mx;
#undef SDF_CENTER
    free(T);
    out->grid = G;
    return 0;
}

void orc_sdf_free(orc_sdf *sdf) {
    if (!sdf) return;
    free(sdf->grid);
    memset(sdf, 0, sizeof(*sdf));
}

int orc_sdf_in_range(const orc_sdf *sdf, const double q[3]) {
    const double org[3] = {sdf->ox, sdf->oy, sdf->oz};
    const int dims[3] = {sdf->nx, sdf->ny, sdf->nz};
    for (int k = 0; k < 3; ++k) {
        const double t = (q[k] - org[k]) / (double)sdf->voxel;
        if (!(t >= 0.0) || t > (double)(dims[k] - 1)) return 0;
    }
    return 1;
}

double orc_sdf_sample(const orc_sdf *sdf, const double q[3]) {
    const double org[3] = {sdf->ox, sdf->oy, sdf->oz};
    const int dims[3] = {sdf->nx, sdf->ny, sdf->nz};
    int i0[3], i1[3];
    double f[3];
    for (int k = 0; k < 3; ++k) {
        const double u = (q[k] - org[k]) / (double)sdf->voxel - 0.5;
        const double fl = floor(u);
        f[k] = u - fl;
        int a = fl < -1.0 ? -1 : (fl > (double)dims[k] ? dims[k] : (int)fl);
        i0[k] = a < 0 ? 0 : (a > dims[k] - 1 ? dims[k] - 1 : a);
        i1[k] = a + 1 < 0 ? 0 : (a + 1 > dims[k] - 1 ? dims[k] - 1 : a + 1);
    }
    double v = 0.0;
    for (int n = 0; n < 8; ++n) {
        const int dx = n & 1, dy = (n >> 1) & 1, dz = (n >> 2) & 1;
        const double w = (dx ? f[0] : 1.0 - f[0]) * (dy ? f[1] : 1.0 - f[1]) * (dz ? f[2] : 1.0 - f[2]);
        const long idx = (dx ? i1[0] : i0[0]) + (long)sdf->nx * ((dy ? i1[1] : i0[1]) + (long)sdf->ny * (dz ? i1[2] : i0[2]));
        v += w * (double)sdf->grid[idx];
    }
    return v;
}

void orc_set_tool_sdf(orc_solver *s, const orc_sdf *sdf, const orc_tool_desc *d) {
    orc_sdf_free(&s->sdf);
    s->sdf = *sdf;
    const size_t n = (size_t)sdf->nx * sdf->ny * sdf->nz;
    s->sdf.grid = (float *)malloc(sizeof(float) * n);
    memcpy(s->sdf.grid, sdf->grid, sizeof(float) * n);
    s->tool = *d;
    s->tool.kind = ORC_TOOL_VOXEL;
}

/* phi of the voxel tool at world point (x, y, z): the sampled value inside
 * the texture range, no contact (a large value) outside it. */
static double phi_voxel_world(const orc_solver *s, double x, double y, double z) {
    double q[3];
    to_tool_frame(&s->pose, x, y, z, q);
    if (!orc_sdf_in_range(&s->sdf, q)) return 1e300;
    return orc_sdf_sample(&s->sdf, q);
}

/* Normal of the voxel tool: central differences of the clamped sampler
 * along the world axes (the reference differences the texture along its
 * rest-frame axes, ContactSolver.cu:848-859; along the world axes the
 * gradient is the same vector rotated into the world frame). */
static void normal_voxel_world(const orc_solver *s, double x, double y, double z, double n[3]) {
    const double e = s->tool.grad_eps;
    double qp[3], qm[3], g[3];
    const double dd[3][3] = {{e, 0, 0}, {0, e, 0}, {0, 0, e}};
    for (int k = 0; k < 3; ++k) {
        to_tool_frame(&s->pose, x + dd[k][0], y + dd[k][1], z + dd[k][2], qp);
        to_tool_frame(&s->pose, x - dd[k][0], y - dd[k][1], z - dd[k][2], qm);
        g[k] = (orc_sdf_sample(&s->sdf, qp) - orc_sdf_sample(&s->sdf, qm)) / (2.0 * e);
    }
    const double len = sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
    if (len > 1e-12) {
        n[0] = g[0] / len; n[1] = g[1] / len; n[2] = g[2] / len;
    } else {
        n[0] = n[1] = n[2] = 0.0;
    }
}

static double phi_local(const orc_tool_desc *d, const double q[3]) {
    if (d->kind == ORC_TOOL_WEDGE) return phi_wedge(d, q);
    if (d->kind == ORC_TOOL_DRILL) return phi_drill(d, q);
    return 1e300;
}

double orc_tool_phi(const orc_tool_desc *d, const orc_tool_pose *p, double x, double y, double z) {
    double q[3];
    to_tool_frame(p, x, y, z, q);
    return phi_local(d, q);
}

/* Normal = normalised central-difference gradient of phi in world space. */
void orc_tool_normal(const orc_tool_desc *d, const orc_tool_pose *p, double x, double y, double z,
                     double n[3]) {
    double e = d->grad_eps;
    double gx = (orc_tool_phi(d, p, x + e, y, z) - orc_tool_phi(d, p, x - e, y, z)) / (2.0 * e);
    double gy = (orc_tool_phi(d, p, x, y + e, z) - orc_tool_phi(d, p, x, y - e, z)) / (2.0 * e);
    double gz = (orc_tool_phi(d, p, x, y, z + e) - orc_tool_phi(d, p, x, y, z - e)) / (2.0 * e);
    double len = sqrt(gx * gx + gy * gy + gz * gz);
    if (len > 1e-12) {
        n[0] = gx / len; n[1] = gy / len; n[2] = gz / len;
    } else {
        n[0] = n[1] = n[2] = 0.0;
    }
}

/* ------------------------------------------------------------ lifecycle -- */

/* Config defaults the hot path reads (Config.h:23-45,74-80,128-157,343-346,366-391)
 * mapped as MPMWorkpieceSolver::initialize does (MPMWorkpieceSolver.cu:428-451). */
void orc_params_default(orc_params *p) {
    memset(p, 0, sizeof(*p));
    p->density = 4430.0;
    p->young = 113.8e9;
    p->poisson = 0.34;
    p->specific_heat = 526.3;
    p->melt_temp = 1660.0;
    p->ref_temp = 25.0;
    p->jc_A = 880e6; p->jc_B = 290e6; p->jc_n = 0.47; p->jc_C = 0.015; p->jc_m = 1.0;
    p->ref_strain_rate = 1.0;
    p->failure_strain = 0.3;
    p->damage_threshold = 1.0;
    p->critical_plastic_strain = 0.15;
    p->thermal_chip_ratio = 0.25;
    p->cell_size = 0.0001;
    p->tool_contact_radius = 1.5 * p->cell_size;
    p->contact_damping = 0.15;              /* MPMWorkpieceSolver.cuh:85 */
    p->max_contact_temp_rise = 50.0;
    p->thermal_conductivity = 42.6;         /* MPMWorkpieceSolver.cuh:87 */
    p->kinematic_hardening = 0.0;           /* MPMWorkpieceSolver.cuh:88 */
    p->thermal_clamp_enabled = 1;
    p->has_config = 1;
    p->particle_spacing_factor = 0.8;
    p->fixture_thickness = 0.002;
    p->min_time_step = 1e-12;
    p->max_time_step = 1e-4;
    p->max_particles = 100000;
}

orc_solver *orc_create(const orc_params *p) {
    orc_solver *s = (orc_solver *)calloc(1, sizeof(orc_solver));
    s->prm = *p;
    s->g.cell = p->cell_size;
    s->g.nx = s->g.ny = s->g.nz = 1;
    s->max_temp = p->ref_temp;    /* reset() leaves m_maxTemperature = refTemp */
    s->tool.kind = ORC_TOOL_NONE;
    return s;
}

void orc_destroy(orc_solver *s) {
    if (!s) return;
    free(s->p); free(s->grid); free(s->snap); free(s->tm_new); free(s->tool_xyz);
    orc_sdf_free(&s->sdf);
    free(s);
}

static void ensure_particles(orc_solver *s, int n) {
    if (n > s->cap) {
        s->p = (orc_particle *)realloc(s->p, (size_t)n * sizeof(orc_particle));
        s->cap = n;
    }
}

static orc_particle *push_particle(orc_solver *s) {
    if (s->n + 1 > s->cap) ensure_particles(s, 2 * s->cap + 1024);
    return &s->p[s->n++];
}

/* allocateGrid, MPMWorkpieceSolver.cu:562-568 (reallocate on size change) */
static void alloc_grid(orc_solver *s) {
    int n = s->g.nx * s->g.ny * s->g.nz;
    if (n == s->nnodes && s->grid) return;
    free(s->grid); free(s->snap); free(s->tm_new);
    s->nnodes = n;
    s->grid = (orc_node *)calloc((size_t)n, sizeof(orc_node));
    s->snap = (orc_node *)calloc((size_t)n, sizeof(orc_node));
    s->tm_new = (double *)calloc((size_t)n, sizeof(double));
}

/* rebuildGridFromHostBounds, MPMWorkpieceSolver.cu:570-586 */
static void rebuild_grid(orc_solver *s) {
    if (s->n <= 0) return;
    double mn[3] = {1e30, 1e30, 1e30}, mx[3] = {-1e30, -1e30, -1e30};
    for (int i = 0; i < s->n; ++i) {
        const orc_particle *q = &s->p[i];
        if (q->status == ORC_INACTIVE) continue;
        mn[0] = fmin(mn[0], q->x); mn[1] = fmin(mn[1], q->y); mn[2] = fmin(mn[2], q->z);
        mx[0] = fmax(mx[0], q->x); mx[1] = fmax(mx[1], q->y); mx[2] = fmax(mx[2], q->z);
    }
    orc_rebuild_params(s->g.cell, mn, mx, &s->g);
    alloc_grid(s);
}

static void load_particles(orc_solver *s, const orc_particle_arrays *a);

int orc_set_particles(orc_solver *s, const orc_particle_arrays *a) {
    load_particles(s, a);
    rebuild_grid(s);   /* as initializeParticleBox does before uploadParticles */
    return 0;
}

/* The particle state replaced, the grid, step count and time kept: the
 * shadow oracle of a parity test continues a trajectory from another
 * solver's state.  No counterpart in the reference. */
int orc_replace_particles(orc_solver *s, const orc_particle_arrays *a) {
    load_particles(s, a);
    return 0;
}

static void load_particles(orc_solver *s, const orc_particle_arrays *a) {
    ensure_particles(s, a->n);
    s->n = a->n;
    for (int i = 0; i < a->n; ++i) {
        orc_particle *q = &s->p[i];
        memset(q, 0, sizeof(*q));
        q->x = a->pos[3 * i]; q->y = a->pos[3 * i + 1]; q->z = a->pos[3 * i + 2];
        q->vx = a->vel[3 * i]; q->vy = a->vel[3 * i + 1]; q->vz = a->vel[3 * i + 2];
        if (a->fext) { q->fx = a->fext[3 * i]; q->fy = a->fext[3 * i + 1]; q->fz = a->fext[3 * i + 2]; }
        q->mass = a->mass[i]; q->volume = a->volume[i];
        q->density = a->density ? a->density[i] : s->prm.density;
        q->temperature = a->temperature[i];
        q->plasticStrain = a->plastic_strain[i]; q->strainRate = a->strain_rate[i];
        q->damage = a->damage[i]; q->residualStress = a->residual_stress[i];
        for (int c = 0; c < 6; ++c) {
            q->s[c] = a->stress[6 * i + c];
            q->a[c] = a->backstress ? a->backstress[6 * i + c] : 0.0;
        }
        q->id = a->id[i];
        q->status = a->status[i];
    }
}

int orc_get_particles(orc_solver *s, orc_particle_arrays *a) {
    int n = a->n < s->n ? a->n : s->n;
    for (int i = 0; i < n; ++i) {
        const orc_particle *q = &s->p[i];
        a->pos[3 * i] = q->x; a->pos[3 * i + 1] = q->y; a->pos[3 * i + 2] = q->z;
        a->vel[3 * i] = q->vx; a->vel[3 * i + 1] = q->vy; a->vel[3 * i + 2] = q->vz;
        if (a->fext) { a->fext[3 * i] = q->fx; a->fext[3 * i + 1] = q->fy; a->fext[3 * i + 2] = q->fz; }
        a->mass[i] = q->mass; a->volume[i] = q->volume;
        if (a->density) a->density[i] = q->density;
        a->temperature[i] = q->temperature;
        a->plastic_strain[i] = q->plasticStrain; a->strain_rate[i] = q->strainRate;
        a->damage[i] = q->damage; a->residual_stress[i] = q->residualStress;
        for (int c = 0; c < 6; ++c) {
            a->stress[6 * i + c] = q->s[c];
            if (a->backstress) a->backstress[6 * i + c] = q->a[c];
        }
        a->id[i] = q->id;
        a->status[i] = q->status;
    }
    a->n = n;
    return n;
}

int orc_particle_count(const orc_solver *s) { return s->n; }

/* initializeParticleBox, MPMWorkpieceSolver.cu:588-614 */
void orc_init_particle_box(orc_solver *s, const double mn[3], const double mx[3], double spacing) {
    int maxp = s->prm.has_config ? s->prm.max_particles : 100000;
    double vol = spacing * spacing * spacing;
    int id = 0;
    s->n = 0;
    for (double z = mx[2]; z >= mn[2] - 1e-12 && s->n < maxp; z -= spacing)
        for (double y = mn[1]; y <= mx[1] + 1e-12 && s->n < maxp; y += spacing)
            for (double x = mn[0]; x <= mx[0] + 1e-12 && s->n < maxp; x += spacing) {
                orc_particle *q = push_particle(s);
                memset(q, 0, sizeof(*q));
                q->x = x; q->y = y; q->z = z;
                q->mass = s->prm.density * vol;
                q->volume = vol;
                q->density = s->prm.density;
                q->temperature = s->prm.ref_temp;
                q->id = id++;
                q->status = ORC_ACTIVE;
                if (s->prm.has_config && z <= mn[2] + s->prm.fixture_thickness)
                    q->status = ORC_FIXED_BOUNDARY;
            }
    rebuild_grid(s);
}

/* initializeCylindricalWorkpiece, MPMWorkpieceSolver.cu:616-648 (axis ignored as there) */
void orc_init_cylinder(orc_solver *s, const double c[3], double radius, double length,
                       double spacing, int axis) {
    (void)axis;
    int maxp = s->prm.has_config ? s->prm.max_particles : 100000;
    double vol = spacing * spacing * spacing;
    int id = 0;
    double minZ = c[2] - length * 0.5, maxZ = c[2] + length * 0.5;
    s->n = 0;
    for (double z = maxZ; z >= minZ - 1e-12 && s->n < maxp; z -= spacing)
        for (double y = c[1] - radius; y <= c[1] + radius + 1e-12 && s->n < maxp; y += spacing)
            for (double x = c[0] - radius; x <= c[0] + radius + 1e-12 && s->n < maxp; x += spacing) {
                double dx = x - c[0], dy = y - c[1];
                if (dx * dx + dy * dy > radius * radius) continue;
                orc_particle *q = push_particle(s);
                memset(q, 0, sizeof(*q));
                q->x = x; q->y = y; q->z = z;
                q->mass = s->prm.density * vol;
                q->volume = vol;
                q->density = s->prm.density;
                q->temperature = s->prm.ref_temp;
                q->id = id++;
                q->status = ORC_ACTIVE;
                if (s->prm.has_config && z <= minZ + s->prm.fixture_thickness)
                    q->status = ORC_FIXED_BOUNDARY;
            }
    rebuild_grid(s);
}

/* setToolMesh, MPMWorkpieceSolver.cu:793-810: every stride-th node, stride = max(1, N/4000) */
void orc_set_tool_samples(orc_solver *s, const double *xyz, int n, double contact_radius) {
    s->prm.tool_contact_radius = fmax(contact_radius, s->g.cell);
    int stride = n / 4000;
    if (stride < 1) stride = 1;
    free(s->tool_xyz);
    s->tool_xyz = NULL;
    s->ntool = 0;
    if (n > 0) {
        s->tool_xyz = (double *)malloc(sizeof(double) * 3 * (size_t)((n + stride - 1) / stride));
        for (int i = 0; i < n; i += stride) {
            s->tool_xyz[3 * s->ntool] = xyz[3 * i];
            s->tool_xyz[3 * s->ntool + 1] = xyz[3 * i + 1];
            s->tool_xyz[3 * s->ntool + 2] = xyz[3 * i + 2];
            s->ntool++;
        }
    }
    s->tool.kind = s->ntool > 0 ? ORC_TOOL_POINTS : ORC_TOOL_NONE;
}

void orc_set_tool_analytic(orc_solver *s, const orc_tool_desc *d) { s->tool = *d; }
void orc_set_tool_pose(orc_solver *s, const orc_tool_pose *p) { s->pose = *p; }

/* setToolTipPosition, MPMWorkpieceSolver.cuh:133-135 */
void orc_set_tool_tip(orc_solver *s, double x, double y, double z) {
    s->tip[0] = x; s->tip[1] = y; s->tip[2] = z;
}

/* applyExternalForce + applyMPMExternalForceKernel, MPMWorkpieceSolver.cu:812-818, :402-414 */
void orc_apply_external_force(orc_solver *s, const double c[3], double radius, const double f[3]) {
    for (int i = 0; i < s->n; ++i) {
        orc_particle *q = &s->p[i];
        if (q->status != ORC_ACTIVE) continue;
        double dx = q->x - c[0], dy = q->y - c[1], dz = q->z - c[2];
        if (dx * dx + dy * dy + dz * dz <= radius * radius) {
            q->fx += f[0]; q->fy += f[1]; q->fz += f[2];
        }
    }
}

/* ---------------------------------------------------------------- step --- */

/* One particle of mpmP2GKernel, MPMWorkpieceSolver.cu:103-133. */
/* x += v, rounded to FP32 after the add when f32 is set (the sensitivity
 * probe's FP32 accumulators; off for every parity use of the oracle). */
static void acc(double *x, double v, int f32) {
    *x += v;
    if (f32) *x = (double)(float)*x;
}

static void p2g_particle_pic(const orc_grid *g, const orc_particle *p, orc_node *G, int f32) {
    int32_t ids[8];
    double ws[8], gr[24];
    orc_weights(g, p->x, p->y, p->z, ids, ws, gr);
    for (int n = 0; n < 8; ++n) {
        int gid = ids[n];
        if (gid < 0 || ws[n] <= 0.0) continue;
        orc_node *nd = &G[gid];
        double m = p->mass * ws[n];
        acc(&nd->mass, m, f32);
        acc(&nd->vx, p->vx * m, f32);
        acc(&nd->vy, p->vy * m, f32);
        acc(&nd->vz, p->vz * m, f32);
        acc(&nd->tm, p->temperature * m, f32);
        if (p->status == ORC_ACTIVE || p->status == ORC_FIXED_BOUNDARY) {
            double gx = gr[3 * n], gy = gr[3 * n + 1], gz = gr[3 * n + 2];
            double fX = -(p->s[0] * gx + p->s[3] * gy + p->s[4] * gz) * p->volume;
            double fY = -(p->s[3] * gx + p->s[1] * gy + p->s[5] * gz) * p->volume;
            double fZ = -(p->s[4] * gx + p->s[5] * gy + p->s[2] * gz) * p->volume;
            acc(&nd->fx, fX + p->fx * ws[n], f32);
            acc(&nd->fy, fY + p->fy * ws[n], f32);
            acc(&nd->fz, fZ + p->fz * ws[n], f32);
        } else if (p->status == ORC_CHIP) {
            acc(&nd->fx, p->fx * ws[n], f32);
            acc(&nd->fy, p->fy * ws[n], f32);
            acc(&nd->fz, p->fz * ws[n], f32);
        }
    }
}

/* One particle of the APIC transfer (p2gKernel, MPMSolver.cu:370-421) with
 * the workpiece solver's channels and force rules: mass m w; momentum
 * m w (v + C (x_i - x_p)), none from FIXED_BOUNDARY particles (:393-398);
 * T m w; the stress force of ACTIVE and FIXED_BOUNDARY particles and the
 * external force as mpmP2GKernel forms them (MPMWorkpieceSolver.cu:121-132),
 * with the quadratic weight gradients. */
static void p2g_particle_apic(const orc_grid *g, const orc_particle *p, orc_node *G) {
    int32_t ids[27];
    double ws[27], gr[81], dp[81];
    orc_weights_quadratic(g, p->x, p->y, p->z, ids, ws, gr, dp);
    const int fixed = p->status == ORC_FIXED_BOUNDARY;
    const int stress = p->status == ORC_ACTIVE || fixed;
    for (int n = 0; n < 27; ++n) {
        if (ids[n] < 0) continue;
        orc_node *nd = &G[ids[n]];
        const double m = p->mass * ws[n];
        nd->mass += m;
        if (!fixed) {
            const double *d = &dp[3 * n];
            const double ax = p->C[0] * d[0] + p->C[1] * d[1] + p->C[2] * d[2];
            const double ay = p->C[3] * d[0] + p->C[4] * d[1] + p->C[5] * d[2];
            const double az = p->C[6] * d[0] + p->C[7] * d[1] + p->C[8] * d[2];
            nd->vx += m * (p->vx + ax);
            nd->vy += m * (p->vy + ay);
            nd->vz += m * (p->vz + az);
        }
        nd->tm += p->temperature * m;
        if (stress) {
            const double gx = gr[3 * n], gy = gr[3 * n + 1], gz = gr[3 * n + 2];
            const double fX = -(p->s[0] * gx + p->s[3] * gy + p->s[4] * gz) * p->volume;
            const double fY = -(p->s[3] * gx + p->s[1] * gy + p->s[5] * gz) * p->volume;
            const double fZ = -(p->s[4] * gx + p->s[5] * gy + p->s[2] * gz) * p->volume;
            nd->fx += fX + p->fx * ws[n];
            nd->fy += fY + p->fy * ws[n];
            nd->fz += fZ + p->fz * ws[n];
        } else if (p->status == ORC_CHIP) {
            nd->fx += p->fx * ws[n];
            nd->fy += p->fy * ws[n];
            nd->fz += p->fz * ws[n];
        }
    }
}

/* mpmP2GKernel, MPMWorkpieceSolver.cu:98-134 (serial: atomics become plain adds;
 * with several threads each thread fills a private grid that is summed after). */
static void p2g(orc_solver *s) {
    const orc_grid *g = &s->g;
    int nt = g_threads;
    orc_node *priv = NULL;
    if (nt > 1) priv = (orc_node *)calloc((size_t)(nt - 1) * s->nnodes, sizeof(orc_node));
#ifdef _OPENMP
#pragma omp parallel num_threads(nt) if (nt > 1)
#endif
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        orc_node *G = (t == 0) ? s->grid : priv + (size_t)(t - 1) * s->nnodes;
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (int pi = 0; pi < s->n; ++pi) {
            const orc_particle *p = &s->p[pi];
            if (p->status == ORC_INACTIVE) continue;
            if (s->transfer == ORC_TRANSFER_APIC) p2g_particle_apic(g, p, G);
            else p2g_particle_pic(g, p, G, s->fp32_emulation);
        }
    }
    if (priv) {
        for (int t = 1; t < nt; ++t) {
            const orc_node *P = priv + (size_t)(t - 1) * s->nnodes;
#ifdef _OPENMP
#pragma omp parallel for num_threads(nt) schedule(static)
#endif
            for (int i = 0; i < s->nnodes; ++i) {
                orc_node *d = &s->grid[i];
                d->mass += P[i].mass; d->vx += P[i].vx; d->vy += P[i].vy; d->vz += P[i].vz;
                d->fx += P[i].fx; d->fy += P[i].fy; d->fz += P[i].fz; d->tm += P[i].tm;
            }
        }
        free(priv);
    }
}

/* mpmGridUpdateKernel, MPMWorkpieceSolver.cu:136-193; analytic tools follow the
 * grid-node form of ContactSolver.cu:797-923 (normal from the SDF gradient,
 * relative velocity against the moving tool, Coulomb friction). */
static void grid_update(orc_solver *s, double fixture, double dt) {
    const orc_grid *g = &s->g;
    const orc_params *mat = &s->prm;
    double force = 0.0;
    const int nt = g_threads;
#ifdef _OPENMP
#pragma omp parallel for num_threads(nt) schedule(static) reduction(+ : force) if (nt > 1)
#endif
    for (int idx = 0; idx < s->nnodes; ++idx) {
        orc_node *node = &s->grid[idx];
        if (node->mass <= 1e-18) continue;
        node->vx /= node->mass;
        node->vy /= node->mass;
        node->vz /= node->mass;
        int i = idx % g->nx, t = idx / g->nx, j = t % g->ny, k = t / g->ny;
        double px = g->minx + i * g->cell, py = g->miny + j * g->cell, pz = g->minz + k * g->cell;
        int fixed = pz <= g->minz + fixture + g->cell;
        if (fixed) {
            node->fixed = 1;
            node->vx = node->vy = node->vz = 0.0;
            continue;
        }
        node->vx += (node->fx / node->mass) * dt;
        node->vy += (node->fy / node->mass) * dt;
        node->vz += (node->fz / node->mass) * dt;

        if (s->tool.kind == ORC_TOOL_POINTS && s->ntool > 0) {
            double r2 = mat->tool_contact_radius * mat->tool_contact_radius;
            double bn[3] = {0, 0, 0};
            int hit = 0;
            for (int q = 0; q < s->ntool; ++q) {
                double dx = px - s->tool_xyz[3 * q], dy = py - s->tool_xyz[3 * q + 1],
                       dz = pz - s->tool_xyz[3 * q + 2];
                double d2 = dx * dx + dy * dy + dz * dz;
                if (d2 < r2 && d2 > 1e-20) {
                    r2 = d2;
                    double len = sqrt(d2);   /* Vec3::normalized, Types.h:185-188 */
                    if (len > 1e-12) { bn[0] = dx / len; bn[1] = dy / len; bn[2] = dz / len; }
                    else { bn[0] = bn[1] = bn[2] = 0.0; }
                    hit = 1;
                }
            }
            if (hit) {
                double vn = node->vx * bn[0] + node->vy * bn[1] + node->vz * bn[2];
                if (vn < 0.0) {
                    force += fabs(node->mass * vn / fmax(dt, 1e-15));
                    node->vx -= vn * bn[0];
                    node->vy -= vn * bn[1];
                    node->vz -= vn * bn[2];
                    double heat = 0.5 * node->mass * vn * vn * mat->contact_damping;
                    double dT = heat / fmax(node->mass * mat->specific_heat, 1e-12);
                    if (mat->thermal_clamp_enabled) dT = fmin(dT, mat->max_contact_temp_rise);
                    node->tm += dT * node->mass;
                }
            }
        } else if (s->tool.kind == ORC_TOOL_WEDGE || s->tool.kind == ORC_TOOL_DRILL ||
                   s->tool.kind == ORC_TOOL_VOXEL) {
            const int vox = s->tool.kind == ORC_TOOL_VOXEL;
            double phi = vox ? phi_voxel_world(s, px, py, pz) : orc_tool_phi(&s->tool, &s->pose, px, py, pz);
            if (phi < s->tool.band) {
                double n[3];
                if (vox) normal_voxel_world(s, px, py, pz, n);
                else orc_tool_normal(&s->tool, &s->pose, px, py, pz, n);
                if (n[0] != 0.0 || n[1] != 0.0 || n[2] != 0.0) {
                    const orc_tool_pose *P = &s->pose;
                    double tvx = P->vel[0] - P->omega * (py - P->center[1]);
                    double tvy = P->vel[1] + P->omega * (px - P->center[0]);
                    double tvz = P->vel[2];
                    double rx = node->vx - tvx, ry = node->vy - tvy, rz = node->vz - tvz;
                    double vn = rx * n[0] + ry * n[1] + rz * n[2];
                    if (vn < 0.0) {
                        force += fabs(node->mass * vn / fmax(dt, 1e-15));
                        double ux = rx - vn * n[0], uy = ry - vn * n[1], uz = rz - vn * n[2];
                        double ut2 = ux * ux + uy * uy + uz * uz;
                        double ut = sqrt(ut2);
                        double sc = 1.0;
                        if (s->tool.friction > 0.0 && ut > 0.0)
                            sc = fmax(0.0, 1.0 + s->tool.friction * vn / ut);
                        node->vx = tvx + sc * ux;
                        node->vy = tvy + sc * uy;
                        node->vz = tvz + sc * uz;
                        double removed = vn * vn + (1.0 - sc * sc) * ut2;
                        double heat = 0.5 * node->mass * removed * mat->contact_damping;
                        double dT = heat / fmax(node->mass * mat->specific_heat, 1e-12);
                        if (mat->thermal_clamp_enabled) dT = fmin(dT, mat->max_contact_temp_rise);
                        node->tm += dT * node->mass;
                    }
                }
            }
        }
    }
    s->contact_force = force;
}

/* mpmDiffuseTemperatureKernel, MPMWorkpieceSolver.cu:208-262, in its race-free
 * Jacobi form (see the file header). */
static void diffuse(orc_solver *s, double alpha, double dt) {
    const orc_grid *g = &s->g;
    const int nt = g_threads;
    const orc_node *G = s->grid;
    double h = g->cell, inv_h2 = 1.0 / (h * h);
    int nxy = g->nx * g->ny;
#ifdef _OPENMP
#pragma omp parallel for num_threads(nt) schedule(static) if (nt > 1)
#endif
    for (int idx = 0; idx < s->nnodes; ++idx) {
        const orc_node *node = &G[idx];
        s->tm_new[idx] = node->tm;
        if (node->mass < 1e-15) continue;
        double T = node->tm / node->mass;
        int i = idx % g->nx, t = idx / g->nx, j = t % g->ny, k = t / g->ny;
        double lap = 0.0;
        if (i > 0) lap += (G[idx - 1].tm / fmax(G[idx - 1].mass, 1e-30) - T);
        if (i < g->nx - 1) lap += (G[idx + 1].tm / fmax(G[idx + 1].mass, 1e-30) - T);
        if (j > 0) lap += (G[idx - g->nx].tm / fmax(G[idx - g->nx].mass, 1e-30) - T);
        if (j < g->ny - 1) lap += (G[idx + g->nx].tm / fmax(G[idx + g->nx].mass, 1e-30) - T);
        if (k > 0) lap += (G[idx - nxy].tm / fmax(G[idx - nxy].mass, 1e-30) - T);
        if (k < g->nz - 1) lap += (G[idx + nxy].tm / fmax(G[idx + nxy].mass, 1e-30) - T);
        double T_new = T + alpha * dt * lap * inv_h2;
        s->tm_new[idx] = fmax(T_new, 0.0) * node->mass;
    }
    for (int idx = 0; idx < s->nnodes; ++idx) s->grid[idx].tm = s->tm_new[idx];
}

/* The ACTIVE-particle constitutive block of mpmG2PKernel,
 * MPMWorkpieceSolver.cu:308-397 (with f zeroing at :399 left to the caller). */
static void constitutive_active(const orc_params *mat, orc_particle *p, const double L[9], double dt) {
    double shear = mat->young / (2.0 * (1.0 + mat->poisson));
    double bulk = mat->young / (3.0 * (1.0 - 2.0 * mat->poisson));
    double Dxx = L[0], Dyy = L[4], Dzz = L[8];
    double Dxy = 0.5 * (L[1] + L[3]);
    double Dxz = 0.5 * (L[2] + L[6]);
    double Dyz = 0.5 * (L[5] + L[7]);
    double tr = Dxx + Dyy + Dzz;
    double devXX = Dxx - tr / 3.0, devYY = Dyy - tr / 3.0, devZZ = Dzz - tr / 3.0;
    double eqRate = sqrt(2.0 / 3.0 *
        (devXX * devXX + devYY * devYY + devZZ * devZZ + 2.0 * (Dxy * Dxy + Dxz * Dxz + Dyz * Dyz)));
    p->strainRate = eqRate;
    double pressureInc = bulk * tr * dt;
    p->s[0] += (2.0 * shear * devXX) * dt + pressureInc;
    p->s[1] += (2.0 * shear * devYY) * dt + pressureInc;
    p->s[2] += (2.0 * shear * devZZ) * dt + pressureInc;
    p->s[3] += 2.0 * shear * Dxy * dt;
    p->s[4] += 2.0 * shear * Dxz * dt;
    p->s[5] += 2.0 * shear * Dyz * dt;
    double sx = p->s[0] - p->a[0], sy = p->s[1] - p->a[1], sz = p->s[2] - p->a[2];
    double sxy = p->s[3] - p->a[3], sxz = p->s[4] - p->a[4], syz = p->s[5] - p->a[5];
    double seq = orc_equivalent_stress(sx, sy, sz, sxy, sxz, syz);
    double flow = orc_jc_flow(mat, p->plasticStrain, p->strainRate, p->temperature);
    if (seq > flow && seq > 1e-9) {
        double scale = flow / seq;
        p->s[0] = p->a[0] + scale * sx;
        p->s[1] = p->a[1] + scale * sy;
        p->s[2] = p->a[2] + scale * sz;
        p->s[3] = p->a[3] + scale * sxy;
        p->s[4] = p->a[4] + scale * sxz;
        p->s[5] = p->a[5] + scale * syz;
        double dp = (seq - flow) / fmax(3.0 * shear, 1.0);
        if (mat->kinematic_hardening > 0.0 && dp > 0.0) {
            double ns = mat->kinematic_hardening * dp / seq;
            p->a[0] += ns * sx; p->a[1] += ns * sy; p->a[2] += ns * sz;
            p->a[3] += ns * sxy; p->a[4] += ns * sxz; p->a[5] += ns * syz;
        }
        p->plasticStrain += dp;
        double dT = 0.9 * flow * dp / fmax(mat->density * mat->specific_heat, 1.0);
        if (mat->thermal_clamp_enabled) dT = fmin(dT, mat->max_contact_temp_rise);
        p->temperature = fmin(p->temperature + dT, mat->melt_temp);
    }
    if (eqRate > 0.0) p->damage += (eqRate * dt) / fmax(mat->failure_strain, 1e-3);
    if (p->damage > 0.7 * mat->damage_threshold && p->damage < mat->damage_threshold) {
        double degrade = (mat->damage_threshold - p->damage) / (mat->damage_threshold * 0.3);
        degrade = fmax(0.0, degrade);
        for (int c = 0; c < 6; ++c) { p->s[c] *= degrade; p->a[c] *= degrade; }
        p->residualStress = fmax(p->residualStress,
            orc_equivalent_stress(p->s[0], p->s[1], p->s[2], p->s[3], p->s[4], p->s[5]));
    }
}

/* Chip transition at the end of mpmG2PKernel, MPMWorkpieceSolver.cu:384-397 */
static void chip_transition(const orc_params *mat, orc_particle *p) {
    if (p->status == ORC_ACTIVE &&
        (p->damage >= mat->damage_threshold || p->plasticStrain >= mat->critical_plastic_strain ||
         p->temperature >= mat->melt_temp * mat->thermal_chip_ratio)) {
        p->status = ORC_CHIP;
        p->residualStress = fmax(p->residualStress,
            orc_equivalent_stress(p->s[0], p->s[1], p->s[2], p->s[3], p->s[4], p->s[5]));
        for (int c = 0; c < 6; ++c) { p->s[c] = 0.0; p->a[c] = 0.0; }
        p->damage = fmax(p->damage, mat->damage_threshold);
    }
}

void orc_constitutive(const orc_params *p, orc_particle *pt, const double L[9], double dt) {
    if (pt->status == ORC_ACTIVE) constitutive_active(p, pt, L, dt);
    chip_transition(p, pt);
}

/* One particle of mpmG2PKernel, MPMWorkpieceSolver.cu:264-400. */
static void g2p_particle_pic(const orc_solver *s, orc_particle *p, double dt) {
    const orc_grid *g = &s->g;
    const orc_params *mat = &s->prm;
    int32_t ids[8];
    double ws[8], gr[24];
    orc_weights(g, p->x, p->y, p->z, ids, ws, gr);
    double v[3] = {0, 0, 0}, temp = 0.0, wsum = 0.0;
    double L[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int n = 0; n < 8; ++n) {
        int gid = ids[n];
        if (gid < 0 || ws[n] <= 0.0) continue;
        const orc_node *nd = &s->grid[gid];
        if (nd->mass <= 1e-18) continue;
        double nv[3] = {nd->vx, nd->vy, nd->vz};
        v[0] += nv[0] * ws[n]; v[1] += nv[1] * ws[n]; v[2] += nv[2] * ws[n];
        temp += (nd->tm / nd->mass) * ws[n];
        wsum += ws[n];
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) L[3 * a + b] += nv[a] * gr[3 * n + b];
    }
    if (p->status == ORC_FIXED_BOUNDARY) {
        p->vx = p->vy = p->vz = 0.0;
        p->fx = p->fy = p->fz = 0.0;
        return;
    }
    p->vx = v[0]; p->vy = v[1]; p->vz = v[2];
    p->x += p->vx * dt;
    p->y += p->vy * dt;
    p->z += p->vz * dt;
    if (wsum > 1e-8) p->temperature = fmin(fmax(temp / wsum, mat->ref_temp), mat->melt_temp);
    if (p->status == ORC_ACTIVE) constitutive_active(mat, p, L, dt);
    chip_transition(mat, p);
    p->fx = p->fy = p->fz = 0.0;
}

/* One particle of the APIC transfer's G2P (g2pKernel, MPMSolver.cu:506-592)
 * followed by the workpiece solver's update (mpmG2PKernel,
 * MPMWorkpieceSolver.cu:296-399): the same gather over the 27 quadratic
 * nodes (nodes of mass <= 1e-18 skipped, as the workpiece grid marks them
 * empty), plus B = sum w v_i (x_i - x_p)^T; then C = B * 3 / h^2 (:577-578,
 * the reference's factor; the quadratic B-spline's D^-1 is 4 / h^2) and
 * C *= 0.99 (:591), except for particles that were CHIP at entry, whose C
 * the reference leaves unchanged (:570-573). */
static void g2p_particle_apic(const orc_solver *s, orc_particle *p, double dt) {
    const orc_grid *g = &s->g;
    const orc_params *mat = &s->prm;
    int32_t ids[27];
    double ws[27], gr[81], dp[81];
    orc_weights_quadratic(g, p->x, p->y, p->z, ids, ws, gr, dp);
    double v[3] = {0, 0, 0}, temp = 0.0, wsum = 0.0;
    double L[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, B[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int n = 0; n < 27; ++n) {
        if (ids[n] < 0) continue;
        const orc_node *nd = &s->grid[ids[n]];
        if (nd->mass <= 1e-18) continue;
        const double nv[3] = {nd->vx, nd->vy, nd->vz};
        const double w = ws[n];
        v[0] += w * nv[0]; v[1] += w * nv[1]; v[2] += w * nv[2];
        temp += (nd->tm / nd->mass) * w;
        wsum += w;
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                B[3 * a + b] += w * nv[a] * dp[3 * n + b];
                L[3 * a + b] += nv[a] * gr[3 * n + b];
            }
    }
    if (p->status == ORC_FIXED_BOUNDARY) {
        p->vx = p->vy = p->vz = 0.0;
        p->fx = p->fy = p->fz = 0.0;
        return;
    }
    const int chip0 = p->status == ORC_CHIP;
    p->vx = v[0]; p->vy = v[1]; p->vz = v[2];
    p->x += p->vx * dt;
    p->y += p->vy * dt;
    p->z += p->vz * dt;
    if (wsum > 1e-8) p->temperature = fmin(fmax(temp / wsum, mat->ref_temp), mat->melt_temp);
    if (p->status == ORC_ACTIVE) constitutive_active(mat, p, L, dt);
    chip_transition(mat, p);
    if (!chip0) {
        const double inv = 1.0 / g->cell;
        const double dinv = (s->apic_textbook ? 4.0 : 3.0) * inv * inv;
        for (int c = 0; c < 9; ++c) {
            p->C[c] = B[c] * dinv;
            if (!s->apic_textbook) p->C[c] *= 0.99;
        }
    }
    p->fx = p->fy = p->fz = 0.0;
}

/* mpmG2PKernel, MPMWorkpieceSolver.cu:264-400 */
static void g2p(orc_solver *s, double dt) {
    const int nt = g_threads;
#ifdef _OPENMP
#pragma omp parallel for num_threads(nt) schedule(static) if (nt > 1)
#endif
    for (int pi = 0; pi < s->n; ++pi) {
        orc_particle *p = &s->p[pi];
        if (p->status == ORC_INACTIVE) continue;
        if (s->transfer == ORC_TRANSFER_APIC) g2p_particle_apic(s, p, dt);
        else g2p_particle_pic(s, p, dt);
    }
}

/* chipFractureAndSeparationKernel, MPMWorkpieceSolver.cu:667-725 */
static void chip_fracture(orc_solver *s) {
    const orc_params *mat = &s->prm;
    double cell = s->g.cell, tipz = s->tip[2];
    double dth = mat->damage_threshold;
    for (int pi = 0; pi < s->n; ++pi) {
        orc_particle *p = &s->p[pi];
        if (p->status == ORC_INACTIVE) continue;
        if (p->status == ORC_ACTIVE) {
            double cutDepth = cell * 1.5;
            if (p->z < tipz - cutDepth && p->plasticStrain > 1e-6) {
                p->status = ORC_CHIP;
                p->residualStress = fmax(p->residualStress,
                    sqrt(p->s[0] * p->s[0] + p->s[1] * p->s[1] + p->s[2] * p->s[2]));
                for (int c = 0; c < 6; ++c) { p->s[c] = 0.0; p->a[c] = 0.0; }
                p->damage = fmax(p->damage, dth);
            }
            continue;
        }
        if (p->status == ORC_CHIP) {
            if (p->damage < dth * 1.5) continue;
            double r[3];
            orc_chip_kick(p->id, r);
            double speed = sqrt(p->vx * p->vx + p->vy * p->vy + p->vz * p->vz);
            double mag = speed * 0.3 + 0.5;
            p->vx += r[0] * mag;
            p->vy += r[1] * mag;
            p->vz += r[2] * mag;
            p->damage = dth * 0.5;
        }
    }
}

/* updateMetrics, MPMWorkpieceSolver.cu:855-868 */
static void update_metrics(orc_solver *s) {
    s->max_stress = 0.0;
    s->max_temp = s->prm.ref_temp;
    s->ke = 0.0;
    for (int i = 0; i < s->n; ++i) {
        const orc_particle *p = &s->p[i];
        if (p->status == ORC_INACTIVE) continue;
        double st = orc_equivalent_stress(p->s[0], p->s[1], p->s[2], p->s[3], p->s[4], p->s[5]);
        s->max_stress = fmax(s->max_stress, st);
        s->max_temp = fmax(s->max_temp, p->temperature);
        s->ke += 0.5 * p->mass * (p->vx * p->vx + p->vy * p->vy + p->vz * p->vz);
    }
}

void orc_sync_metrics(orc_solver *s) { update_metrics(s); }

/* compactParticles, MPMWorkpieceSolver.cu:475-527 (order-preserving) */
void orc_compact(orc_solver *s) {
    if (s->n <= 0) return;
    int k = 0;
    for (int i = 0; i < s->n; ++i)
        if (s->p[i].status != ORC_INACTIVE) s->p[k++] = s->p[i];
    if (k <= 0) return;
    s->n = k;
    s->compactions++;
}

void orc_set_transfer(orc_solver *s, int mode) {
    s->transfer = (mode == ORC_TRANSFER_APIC || mode == ORC_TRANSFER_APIC_TEXTBOOK) ? ORC_TRANSFER_APIC
                                                                                  : ORC_TRANSFER_PIC;
    s->apic_textbook = mode == ORC_TRANSFER_APIC_TEXTBOOK;
    for (int i = 0; i < s->n; ++i) memset(s->p[i].C, 0, sizeof(s->p[i].C));
}

int orc_get_affine(const orc_solver *s, double *C) {
    for (int i = 0; i < s->n; ++i) memcpy(C + 9 * (size_t)i, s->p[i].C, sizeof(s->p[i].C));
    return s->n;
}

void orc_set_affine(orc_solver *s, const double *C) {
    for (int i = 0; i < s->n; ++i) memcpy(s->p[i].C, C + 9 * (size_t)i, sizeof(s->p[i].C));
}

/* MPMWorkpieceSolver::step, MPMWorkpieceSolver.cu:727-791 */
static double to_f32(double v) { return (double)(float)v; }

void orc_set_fp32_emulation(orc_solver *s, int on) { s->fp32_emulation = on ? 1 : 0; }

void orc_round_state_fp32(orc_solver *s) {
    for (int i = 0; i < s->n; ++i) {
        orc_particle *q = &s->p[i];
        double *f[] = {&q->vx, &q->vy, &q->vz, &q->fx, &q->fy, &q->fz, &q->mass, &q->volume, &q->density,
                       &q->temperature, &q->plasticStrain, &q->strainRate, &q->damage, &q->residualStress};
        for (size_t k = 0; k < sizeof(f) / sizeof(f[0]); ++k) *f[k] = to_f32(*f[k]);
        for (int c = 0; c < 6; ++c) {
            q->s[c] = to_f32(q->s[c]);
            q->a[c] = to_f32(q->a[c]);
        }
        for (int c = 0; c < 9; ++c) q->C[c] = to_f32(q->C[c]);
    }
}

void orc_step(orc_solver *s, double dt) {
    if (s->n <= 0 || !s->grid) return;
    if (s->step % 25 == 0) rebuild_grid(s);
    memset(s->grid, 0, sizeof(orc_node) * (size_t)s->nnodes);   /* clearMPMGridKernel :92-96 */
    p2g(s);
    memcpy(s->snap, s->grid, sizeof(orc_node) * (size_t)s->nnodes);
    double fixture = s->prm.has_config ? s->prm.fixture_thickness : 0.002;
    grid_update(s, fixture, dt);
    if (s->prm.specific_heat > 0.0) {
        double alpha = s->prm.thermal_conductivity / (s->prm.density * s->prm.specific_heat);
        if (alpha > 0.0) diffuse(s, alpha, dt);
    }
    g2p(s, dt);
    if (s->step % 2 == 0) chip_fracture(s);
    /* kinematic tool: the pose advances with its own velocity */
    s->pose.center[0] += s->pose.vel[0] * dt;
    s->pose.center[1] += s->pose.vel[1] * dt;
    s->pose.center[2] += s->pose.vel[2] * dt;
    s->pose.angle += s->pose.omega * dt;
    s->time += dt;
    ++s->step;
    if (s->step % 25 == 0) update_metrics(s);
    if (s->step % 1000 == 0 && s->step > 0) orc_compact(s);
    if (s->fp32_emulation) orc_round_state_fp32(s);
}

void orc_run(orc_solver *s, int nsteps, double dt) {
    for (int i = 0; i < nsteps; ++i) orc_step(s, dt);
}

/* getStableTimeStep, MPMWorkpieceSolver.cu:820-828 */
double orc_stable_dt(const orc_solver *s) {
    double wave = sqrt(fmax(s->prm.young / fmax(s->prm.density, 1.0), 1.0));
    double dt = 0.35 * s->g.cell / wave;
    if (s->prm.has_config) dt = fmax(s->prm.min_time_step, fmin(s->prm.max_time_step, dt));
    return dt;
}

void orc_get_metrics(const orc_solver *s, orc_metrics *m) {
    m->max_stress = s->max_stress;
    m->max_temperature = s->max_temp;
    m->kinetic_energy = s->ke;
    m->contact_force = s->contact_force;
    m->time = s->time;
    m->step = s->step;
}

/* reset, MPMWorkpieceSolver.cu:830-841 */
void orc_reset(orc_solver *s) {
    s->n = 0;
    s->ntool = 0;
    s->tool.kind = ORC_TOOL_NONE;
    s->time = 0.0;
    s->step = 0;
    s->max_stress = 0.0;
    s->max_temp = s->prm.ref_temp;
    s->ke = 0.0;
    s->compactions = 0;
    s->tip[0] = s->tip[1] = s->tip[2] = 0.0;
}

/* getBounds, MPMWorkpieceSolver.cu:843-853 (over the current particles) */
void orc_get_bounds(const orc_solver *s, double b[6]) {
    b[0] = b[1] = b[2] = 1e30;
    b[3] = b[4] = b[5] = -1e30;
    for (int i = 0; i < s->n; ++i) {
        const orc_particle *p = &s->p[i];
        if (p->status == ORC_INACTIVE) continue;
        b[0] = fmin(b[0], p->x); b[1] = fmin(b[1], p->y); b[2] = fmin(b[2], p->z);
        b[3] = fmax(b[3], p->x); b[4] = fmax(b[4], p->y); b[5] = fmax(b[5], p->z);
    }
    if (s->n == 0) b[0] = b[1] = b[2] = b[3] = b[4] = b[5] = 0.0;
}

/* getTemperatureAt, MPMWorkpieceSolver.cu:936-946 */
double orc_temperature_at(const orc_solver *s, double x, double y, double z) {
    double best = 1e30, temp = s->prm.ref_temp;
    for (int i = 0; i < s->n; ++i) {
        const orc_particle *p = &s->p[i];
        double dx = p->x - x, dy = p->y - y, dz = p->z - z;
        double d2 = dx * dx + dy * dy + dz * dz;
        if (d2 < best) { best = d2; temp = p->temperature; }
    }
    return temp;
}

/* applyHeatFlux, MPMWorkpieceSolver.cu:918-934 */
void orc_apply_heat_flux(orc_solver *s, double x, double y, double z, double q) {
    double best = 1e30;
    int nearest = -1;
    for (int i = 0; i < s->n; ++i) {
        const orc_particle *p = &s->p[i];
        double dx = p->x - x, dy = p->y - y, dz = p->z - z;
        double d2 = dx * dx + dy * dy + dz * dz;
        if (d2 < best) { best = d2; nearest = i; }
    }
    if (nearest >= 0 && best <= s->g.cell * s->g.cell) {
        orc_particle *p = &s->p[nearest];
        p->temperature = fmin(p->temperature + q / fmax(p->mass * s->prm.specific_heat, 1e-12),
                              s->prm.melt_temp);
    }
}

/* setParticleTemperatures, MPMWorkpieceSolver.cu:906-916 */
void orc_set_particle_temperatures(orc_solver *s, const double *t, int count) {
    if (count <= 0 || count > s->n) return;
    for (int i = 0; i < count; ++i) s->p[i].temperature = t[i];
}

void orc_get_grid(const orc_solver *s, orc_grid *g) { *g = s->g; }
int orc_grid_node_count(const orc_solver *s) { return s->nnodes; }

/* Grid after the last step: post-update velocity and post-diffusion T*m. */
int orc_get_grid_state(const orc_solver *s, double *mass, double *vel, double *tm, int32_t *fixed) {
    for (int i = 0; i < s->nnodes; ++i) {
        const orc_node *n = &s->grid[i];
        if (mass) mass[i] = n->mass;
        if (vel) { vel[3 * i] = n->vx; vel[3 * i + 1] = n->vy; vel[3 * i + 2] = n->vz; }
        if (tm) tm[i] = n->tm;
        if (fixed) fixed[i] = n->fixed;
    }
    return s->nnodes;
}

int orc_get_grid_snapshot(const orc_solver *s, double *mass, double *mom, double *force, double *tm) {
    for (int i = 0; i < s->nnodes; ++i) {
        const orc_node *n = &s->snap[i];
        if (mass) mass[i] = n->mass;
        if (mom) { mom[3 * i] = n->vx; mom[3 * i + 1] = n->vy; mom[3 * i + 2] = n->vz; }
        if (force) { force[3 * i] = n->fx; force[3 * i + 1] = n->fy; force[3 * i + 2] = n->fz; }
        if (tm) tm[i] = n->tm;
    }
    return s->nnodes;
}

/* ------------------------------------------------------- input generator -- */

static uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static double uniform01(uint64_t seed, uint64_t counter) {
    return (double)(splitmix64(seed ^ (counter * 0xD1B54A32D192ED03ull)) >> 11) * 0x1.0p-53;
}

/* Slab of cx*cy*cz cells, 8 particles per cell (one per 2x2x2 sub-cell) at
 * sub-cell origin + (h/2)*U[0,1)^3, positions rounded to the nearest float. */
void orc_gen_jittered_slab(uint64_t seed, int cx, int cy, int cz, double h, const double origin[3],
                           double density, double fixture, double *pos, double *mass,
                           double *volume, int32_t *id, int32_t *status) {
    double vol = h * h * h / 8.0;
    for (int64_t c = 0; c < (int64_t)cx * cy * cz; ++c) {
        int ci = (int)(c % cx), cj = (int)((c / cx) % cy), ck = (int)(c / ((int64_t)cx * cy));
        for (int k = 0; k < 8; ++k) {
            int64_t pi = c * 8 + k;
            int sub[3] = {k & 1, (k >> 1) & 1, (k >> 2) & 1};
            int cell[3] = {ci, cj, ck};
            for (int a = 0; a < 3; ++a) {
                double u = uniform01(seed, (uint64_t)(pi * 3 + a));
                double x = origin[a] + ((double)cell[a] + 0.5 * sub[a] + 0.5 * u) * h;
                pos[3 * pi + a] = (double)(float)x;
            }
            mass[pi] = density * vol;
            volume[pi] = vol;
            id[pi] = (int32_t)pi;
            status[pi] = (pos[3 * pi + 2] <= origin[2] + fixture) ? ORC_FIXED_BOUNDARY : ORC_ACTIVE;
        }
    }
}
