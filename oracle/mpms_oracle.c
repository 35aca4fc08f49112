/*
 * mpms_oracle.c -- CPU FP64 restatement of MPMSolver::step
 * (src/MPMSolver.cu of Up25-pathan/edegpredict-engine-rotary), AoS, serial.
 * TEST INFRASTRUCTURE ONLY (the checker of the parity tests; see
 * mpms_oracle.h).  Each function cites the reference lines it restates.
 */
#include "mpms_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    double mass, px, py, pz, fx, fy, fz, vx, vy, vz, tvx, tvy, tvz;
    int32_t tool;     /* index + 1 of the tool node that marked it, 0 for none */
} oms_node;

struct oms_solver {
    oms_params prm;
    oms_particle *p;
    int n, cap;
    int32_t dims[3];          /* the whole lattice (MPMSolver.cu:869-871) */
    int32_t lo[3], nb[3];     /* the node sub-box of this step */
    oms_node *grid;
    size_t grid_cap;
    double *tool_xyz, *tool_vel;
    int ntool;
    double tool_pos[3];
    double time;
    int64_t step;
};

static const double kGravityZ = -9.81;   /* GRAVITY_Z, MPMSolver.cu:15 */

void oms_params_default(oms_params *p) {
    memset(p, 0, sizeof(*p));
    /* Config MaterialProperties (Config.h:23-45), Types.h:520-523 */
    p->young = 113.8e9;
    p->poisson = 0.34;
    p->density = 4430.0;
    p->mass_scaling = 1.0;
    p->specific_heat = 526.3;
    p->taylor_quinney = 0.9;          /* MPMSolver.cu:889 */
    p->melt_temp = 1660.0;
    p->ambient_temp = 25.0;
    p->jc_A = 880e6; p->jc_B = 290e6; p->jc_n = 0.47; p->jc_C = 0.015; p->jc_m = 1.0;
    p->kinematic_hardening = 0.0;     /* MPMKernelConfig default, not read by initialize */
    /* MPMParams (Config.h:128-157) */
    p->jc_D1 = 0.06; p->jc_D2 = 3.31; p->jc_D3 = -1.96; p->jc_D4 = 0.018; p->jc_D5 = 0.58;
    p->damage_threshold = 1.0;
    p->ref_strain_rate = 1.0;
    p->lod_active_radius = 0.002;
    p->lod_near_radius = 0.01;
    p->lod_enabled = 1;
    p->fixture_thickness = 0.002;     /* Config.h:383 */
    p->fixture_damping = 1e7;         /* MPMSolver.cu:915 */
    p->dx = fmax(0.0001 * 2.0, 0.0005);   /* MPMSolver.cu:867 */
    for (int a = 0; a < 3; ++a) { p->domain_min[a] = -0.1; p->domain_max[a] = 0.1; }
    p->cfl = 0.2;                     /* m_adaptiveCFL */
    p->max_particles = 5000000;       /* MPMSolver.cuh:199 */
    p->elastic_model = OMS_ELASTIC_NEOHOOKEAN;
    p->return_map = OMS_RETURN_RADIAL;
}

oms_solver *oms_create(const oms_params *p) {
    oms_solver *s = (oms_solver *)calloc(1, sizeof(oms_solver));
    s->prm = *p;
    const double inv = 1.0 / p->dx;
    for (int a = 0; a < 3; ++a) s->dims[a] = (int32_t)((p->domain_max[a] - p->domain_min[a]) * inv) + 1;
    return s;
}

void oms_destroy(oms_solver *s) {
    if (!s) return;
    free(s->p);
    free(s->grid);
    free(s->tool_xyz);
    free(s->tool_vel);
    free(s);
}

static void reserve(oms_solver *s, int n) {
    if (n <= s->cap) return;
    s->p = (oms_particle *)realloc(s->p, sizeof(oms_particle) * (size_t)n);
    s->cap = n;
}

void oms_set_particles(oms_solver *s, const oms_particle *p, int n) {
    reserve(s, n);
    memcpy(s->p, p, sizeof(oms_particle) * (size_t)n);
    s->n = n;
}

int oms_particle_count(const oms_solver *s) { return s->n; }

void oms_get_particles(const oms_solver *s, oms_particle *p) { memcpy(p, s->p, sizeof(oms_particle) * (size_t)s->n); }

void oms_set_tool_nodes(oms_solver *s, const double *xyz, const double *vel, int n) {
    free(s->tool_xyz);
    free(s->tool_vel);
    s->tool_xyz = s->tool_vel = NULL;
    s->ntool = n > 0 ? n : 0;
    if (!s->ntool) return;
    s->tool_xyz = (double *)malloc(sizeof(double) * 3 * (size_t)n);
    s->tool_vel = (double *)malloc(sizeof(double) * 3 * (size_t)n);
    memcpy(s->tool_xyz, xyz, sizeof(double) * 3 * (size_t)n);
    memcpy(s->tool_vel, vel, sizeof(double) * 3 * (size_t)n);
}

void oms_set_tool_position(oms_solver *s, double x, double y, double z) {
    s->tool_pos[0] = x; s->tool_pos[1] = y; s->tool_pos[2] = z;
}

/* MPMParticle default member values (Types.h:246-265) */
static void particle_defaults(oms_particle *q) {
    memset(q, 0, sizeof(*q));
    q->temperature = 25.0;
    q->id = -1;
    for (int i = 0; i < 9; ++i) q->F[i] = q->Fp[i] = (i == 0 || i == 4 || i == 8) ? 1.0 : 0.0;
}

/* MPMSolver::initializeParticleBox, MPMSolver.cu:1290-1325 (z, y, x
 * ascending; bottom 10 % FIXED_BOUNDARY; mass = rho * scaling * s^3) */
void oms_init_box(oms_solver *s, const double mn[3], const double mx[3], double spacing) {
    s->n = 0;
    const double density = s->prm.density * s->prm.mass_scaling;
    const double mass = density * spacing * spacing * spacing;
    const double fixed_z = mn[2] + (mx[2] - mn[2]) * 0.1;
    for (double z = mn[2]; z <= mx[2]; z += spacing)
        for (double y = mn[1]; y <= mx[1]; y += spacing)
            for (double x = mn[0]; x <= mx[0]; x += spacing) {
                if (s->n >= s->prm.max_particles) break;
                if (s->n + 1 > s->cap) reserve(s, 2 * s->cap + 64);
                oms_particle *q = &s->p[s->n++];
                particle_defaults(q);
                q->x = x; q->y = y; q->z = z;
                q->mass = mass;
                q->density = density;
                q->volume = spacing * spacing * spacing;
                q->temperature = s->prm.ambient_temp;
                if (z <= fixed_z) q->status = OMS_FIXED;
            }
}

/* MPMSolver::initializeCylindricalWorkpiece, MPMSolver.cu:1327-1366 (the
 * axis argument is not used there either) */
void oms_init_cylinder(oms_solver *s, const double c[3], double radius, double length, double spacing) {
    s->n = 0;
    const double density = s->prm.density * s->prm.mass_scaling;
    const double mass = density * spacing * spacing * spacing;
    const double fixed_z = c[2] - length * 0.1;
    const double r2 = radius * radius;
    for (double dz = 0; dz <= length; dz += spacing)
        for (double dy = -radius; dy <= radius; dy += spacing)
            for (double dx = -radius; dx <= radius; dx += spacing) {
                if (dx * dx + dy * dy > r2) continue;
                if (s->n >= s->prm.max_particles) break;
                if (s->n + 1 > s->cap) reserve(s, 2 * s->cap + 64);
                oms_particle *q = &s->p[s->n++];
                particle_defaults(q);
                q->x = c[0] + dx; q->y = c[1] + dy; q->z = c[2] - dz;
                q->mass = mass;
                q->density = density;
                q->volume = spacing * spacing * spacing;
                q->temperature = s->prm.ambient_temp;
                if (q->z <= fixed_z) q->status = OMS_FIXED;
            }
}

/* ---- math helpers (MPMSolver.cu:19-69) */

/* getWeightsAndGradients, :19-34 */
static void weights(double xp, double inv_dx, int *base, double w[3], double dw[3]) {
    const double xf = xp * inv_dx;
    *base = (int)floor(xf - 0.5);
    const double d0 = xf - *base, d1 = d0 - 1.0, d2 = d0 - 2.0;
    w[0] = 0.5 * (1.5 - d0) * (1.5 - d0);
    w[1] = 0.75 - d1 * d1;
    w[2] = 0.5 * (1.5 + d2) * (1.5 + d2);
    dw[0] = -(1.5 - d0) * inv_dx;
    dw[1] = -2.0 * d1 * inv_dx;
    dw[2] = (1.5 + d2) * inv_dx;
}

static double det3(const double A[9]) {   /* :36-40 */
    return A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) +
           A[2] * (A[3] * A[7] - A[4] * A[6]);
}

static void mul3(const double A[9], const double B[9], double C[9]) {   /* :42-48 */
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) C[i * 3 + j] = A[i * 3] * B[j] + A[i * 3 + 1] * B[3 + j] + A[i * 3 + 2] * B[6 + j];
}

static void inv3(const double A[9], double inv[9]) {   /* :50-69, identity when |det| < 1e-12 */
    const double d = det3(A);
    if (fabs(d) < 1e-12) {
        for (int i = 0; i < 9; ++i) inv[i] = (i == 0 || i == 4 || i == 8) ? 1.0 : 0.0;
        return;
    }
    const double id = 1.0 / d;
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

/* Rotation of the polar decomposition F = R U by Higham's scaled Newton
 * iteration X <- (g X + X^-T / g) / 2 with g = (|X^-1|_1 |X^-1|_inf /
 * (|X|_1 |X|_inf))^(1/4) (N. J. Higham, "Computing the polar
 * decomposition -- with applications", SIAM J. Sci. Stat. Comput. 7,
 * 1986), until the update changes no entry by more than 1e-14 (at most 30
 * iterations).  The reference's FEM tool iterates the same scaled form
 * (FEMSolver.cu:225-276) but adds X^-1 where Higham adds X^-T, which is the
 * same only for symmetric X; here X^-T, so a rotation gives R = F. */
void oms_polar(const double F[9], double R[9], int *iters) {
    double X[9];
    memcpy(X, F, sizeof(X));
    int it = 0;
    for (; it < 30; ++it) {
        double Xi[9];
        const double d = det3(X);
        if (fabs(d) < 1e-300) break;
        inv3(X, Xi);
        double n1 = 0, ninf = 0, m1 = 0, minf = 0;
        for (int i = 0; i < 3; ++i) {
            double r = 0, c = 0, ri = 0, ci = 0;
            for (int j = 0; j < 3; ++j) {
                r += fabs(X[i * 3 + j]); c += fabs(X[j * 3 + i]);
                ri += fabs(Xi[i * 3 + j]); ci += fabs(Xi[j * 3 + i]);
            }
            ninf = fmax(ninf, r); n1 = fmax(n1, c); minf = fmax(minf, ri); m1 = fmax(m1, ci);
        }
        const double g = pow((m1 * minf) / (n1 * ninf), 0.25), ig = 1.0 / g;
        double diff = 0;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                const double v = 0.5 * (g * X[i * 3 + j] + ig * Xi[j * 3 + i]);
                diff = fmax(diff, fabs(v - X[i * 3 + j]));
                X[i * 3 + j] = v;
            }
        if (diff <= 1e-14) { ++it; break; }
    }
    memcpy(R, X, sizeof(X));
    if (iters) *iters = it;
}

/* Johnson-Cook flow stress of p2gKernel (:225-232) at plastic strain e */
static double jc_flow(const oms_params *c, double e, double edot_star, double t_star) {
    return (c->jc_A + c->jc_B * pow(e, c->jc_n)) * (1.0 + c->jc_C * log(edot_star)) * (1.0 - pow(t_star, c->jc_m));
}

/* Its derivative in e (the Newton return). */
static double jc_flow_de(const oms_params *c, double e, double edot_star, double t_star) {
    if (c->jc_B == 0.0) return 0.0;
    return c->jc_B * c->jc_n * pow(e, c->jc_n - 1.0) * (1.0 + c->jc_C * log(edot_star)) *
           (1.0 - pow(t_star, c->jc_m));
}

/* The consistent return: d >= 0 with q - 3 mu d = Y(e + d), by Newton from
 * the radial estimate, kept inside the bracket [lo, hi] where the residual
 * changes sign (bisection when a Newton step leaves it). */
static double newton_return(const oms_params *c, double q, double mu3, double e, double edot_star, double t_star,
                            double y0) {
    double lo = 0.0, hi = q / mu3;
    double d = fmin(fmax((q - y0) / mu3, 0.0), hi);
    for (int it = 0; it < 60; ++it) {
        const double g = q - mu3 * d - jc_flow(c, e + d, edot_star, t_star);
        if (g > 0) lo = d; else hi = d;
        if (fabs(g) <= 1e-13 * q || hi - lo <= 1e-300) break;
        const double dg = -mu3 - jc_flow_de(c, e + d, edot_star, t_star);
        double nd = d - g / dg;
        if (!(nd > lo && nd < hi)) nd = 0.5 * (lo + hi);
        d = nd;
    }
    return d;
}

/* The constitutive part of p2gKernel for one particle (MPMSolver.cu:
 * 156-343): Fe = F Fp^-1, J = det Fe clamped to [0.9, 1.1], the Cauchy
 * stress mu/J B + (lambda (J - 1) - mu/J) I, the Johnson-Cook return on
 * the backstress-shifted deviator with eps_p increment capped at 0.05, the
 * backstress, the plastic deformation gradient (I + 1.5 d / |s| s) Fp /
 * det^(1/3), heating, the JC failure damage, stress degradation past 0.7 of
 * the threshold and the CHIP transition.  sigma (row-major 3x3) and J are
 * what the scatter uses; *skip is set when the particle adds no stress
 * force.  p is updated. */
void oms_p2g_stress(const oms_params *c, oms_particle *p, double sigma[9], double *Jout, int *skip) {
    int skip_stress = p->lod == OMS_LOD_FAR;
    const int fixed = p->status == OMS_FIXED, chip = p->status == OMS_CHIP;
    if (fixed || chip) skip_stress = 1;
    for (int i = 0; i < 9; ++i) sigma[i] = 0.0;
    double J = 1.0;
    if (!skip_stress && p->status == OMS_ACTIVE) {
        double iFp[9], Fe[9];
        inv3(p->Fp, iFp);
        mul3(p->F, iFp, Fe);
        J = det3(Fe);
        if (J < 0.9) J = 0.9;
        if (J > 1.1) J = 1.1;
        const double mu = c->young / (2.0 * (1.0 + c->poisson));
        const double lambda = c->young * c->poisson / ((1.0 + c->poisson) * (1.0 - 2.0 * c->poisson));
        const double Js = fmax(J, 1e-12);
        if (c->elastic_model == OMS_ELASTIC_COROTATED) {
            double R[9];
            oms_polar(Fe, R, NULL);
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) {
                    double v = 0;
                    for (int k = 0; k < 3; ++k) v += (Fe[i * 3 + k] - R[i * 3 + k]) * Fe[j * 3 + k];
                    sigma[i * 3 + j] = 2.0 * mu / Js * v + (i == j ? lambda * (Js - 1.0) : 0.0);
                }
        } else {
            double B[9];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j)
                    B[i * 3 + j] = Fe[i * 3] * Fe[j * 3] + Fe[i * 3 + 1] * Fe[j * 3 + 1] + Fe[i * 3 + 2] * Fe[j * 3 + 2];
            const double muJ = mu / Js;
            for (int i = 0; i < 9; ++i) {
                sigma[i] = muJ * B[i];
                if (i % 4 == 0) sigma[i] += lambda * (Js - 1.0) - muJ;
            }
        }
        const double tr = sigma[0] + sigma[4] + sigma[8];
        double s[9];
        for (int i = 0; i < 9; ++i) s[i] = sigma[i];
        s[0] -= tr / 3.0; s[4] -= tr / 3.0; s[8] -= tr / 3.0;
        const double ex = s[0] - p->a[0], ey = s[4] - p->a[1], ez = s[8] - p->a[2];
        const double exy = s[1] - p->a[3], exz = s[2] - p->a[4], eyz = s[5] - p->a[5];
        const double q = sqrt(1.5 * (ex * ex + ey * ey + ez * ez + 2.0 * (exy * exy + exz * exz + eyz * eyz)));
        double ts = (p->temperature - c->ambient_temp) / fmax(c->melt_temp - c->ambient_temp, 1.0);
        if (ts < 0) ts = 0;
        if (ts > 1) ts = 1;
        const double edot_star = fmax(p->strain_rate / fmax(c->ref_strain_rate, 1e-12), 1e-5);
        double Y = jc_flow(c, p->plastic_strain, edot_star, ts);
        if (q > Y && q > 1e-12) {
            double de, scale;
            if (c->return_map == OMS_RETURN_NEWTON) {
                de = newton_return(c, q, 3.0 * mu, p->plastic_strain, edot_star, ts, Y);
                if (de > 0.05) de = 0.05;
                Y = jc_flow(c, p->plastic_strain + de, edot_star, ts);
                scale = (q - 3.0 * mu * de) / q;
            } else {
                scale = Y / q;
                de = (q - Y) / fmax(3.0 * mu, 1e-12);
                if (de > 0.05) de = 0.05;
            }
            s[0] = p->a[0] + scale * ex; s[4] = p->a[1] + scale * ey; s[8] = p->a[2] + scale * ez;
            s[1] = p->a[3] + scale * exy; s[2] = p->a[4] + scale * exz; s[5] = p->a[5] + scale * eyz;
            for (int i = 0; i < 9; ++i) {   /* the lower triangle keeps the trial deviator, as :246-249 */
                sigma[i] = s[i];
                if (i % 4 == 0) sigma[i] += tr / 3.0;
            }
            if (de > 0) {
                if (c->kinematic_hardening > 0.0) {
                    const double ns = c->kinematic_hardening * de / q;
                    p->a[0] += ns * ex; p->a[1] += ns * ey; p->a[2] += ns * ez;
                    p->a[3] += ns * exy; p->a[4] += ns * exz; p->a[5] += ns * eyz;
                }
                p->plastic_strain += de;
                {   /* :277-301 */
                    const double sn = sqrt(s[0] * s[0] + s[4] * s[4] + s[8] * s[8] +
                                           2.0 * (s[1] * s[1] + s[2] * s[2] + s[5] * s[5]));
                    const double al = sn > 1e-12 ? 1.5 * de / sn : 0.0;
                    double M[9], nF[9];
                    for (int i = 0; i < 9; ++i) M[i] = al * s[i];
                    M[0] += 1.0; M[4] += 1.0; M[8] += 1.0;
                    mul3(M, p->Fp, nF);
                    const double cb = cbrt(fmax(det3(nF), 1e-12));
                    for (int i = 0; i < 9; ++i) p->Fp[i] = nF[i] / cb;
                }
                double dT = c->taylor_quinney * Y * de / fmax(c->density * c->specific_heat, 1e-12);
                if (dT > 50.0) dT = 50.0;
                p->temperature += dT;
                if (p->temperature > c->melt_temp) p->temperature = c->melt_temp;
                const double tri = (tr / 3.0) / fmax(q, 1e-12);
                const double ef = (c->jc_D1 + c->jc_D2 * exp(c->jc_D3 * tri)) *
                                  (1.0 + c->jc_D4 * log(fmax(edot_star, 1.0))) * (1.0 + c->jc_D5 * ts);
                if (ef > 1e-12) {
                    p->damage += de / ef;
                    if (p->damage > 0.7 * c->damage_threshold) {
                        double dg = (c->damage_threshold - p->damage) / (c->damage_threshold * 0.3);
                        dg = fmax(0.0, fmin(1.0, dg));
                        for (int i = 0; i < 9; ++i) sigma[i] *= dg;
                        p->residual_stress = fmax(p->residual_stress,
                                                  sqrt(1.5 * (s[0] * s[0] + s[4] * s[4] + s[8] * s[8] +
                                                              2.0 * (s[1] * s[1] + s[2] * s[2] + s[5] * s[5]))));
                    }
                    if (p->damage >= c->damage_threshold) {
                        p->status = OMS_CHIP;
                        skip_stress = 1;
                    }
                }
            }
        }
        p->s[0] = sigma[0]; p->s[1] = sigma[4]; p->s[2] = sigma[8];
        p->s[3] = sigma[1]; p->s[4] = sigma[2]; p->s[5] = sigma[5];
    }
    *Jout = J;
    *skip = skip_stress;
}

/* The node sub-box: every node of a non-INACTIVE particle's stencil. */
static void fit_box(oms_solver *s) {
    const double inv = 1.0 / s->prm.dx;
    int lo[3] = {s->dims[0], s->dims[1], s->dims[2]}, hi[3] = {0, 0, 0};
    int any = 0;
    for (int i = 0; i < s->n; ++i) {
        const oms_particle *p = &s->p[i];
        if (p->status == OMS_INACTIVE) continue;
        const double r[3] = {p->x, p->y, p->z};
        for (int a = 0; a < 3; ++a) {
            const int b = (int)floor((r[a] - s->prm.domain_min[a]) * inv - 0.5);
            if (b < lo[a]) lo[a] = b;
            if (b + 3 > hi[a]) hi[a] = b + 3;
        }
        any = 1;
    }
    for (int a = 0; a < 3; ++a) {
        if (!any) { lo[a] = 0; hi[a] = 1; }
        if (lo[a] < 0) lo[a] = 0;
        if (hi[a] > s->dims[a]) hi[a] = s->dims[a];
        if (hi[a] <= lo[a]) hi[a] = lo[a] + 1;
        s->lo[a] = lo[a];
        s->nb[a] = hi[a] - lo[a];
    }
    const size_t nn = (size_t)s->nb[0] * s->nb[1] * s->nb[2];
    if (nn > s->grid_cap) {
        free(s->grid);
        s->grid = (oms_node *)malloc(sizeof(oms_node) * nn);
        s->grid_cap = nn;
    }
    memset(s->grid, 0, sizeof(oms_node) * nn);   /* resetGridKernel, :113-122 */
}

static oms_node *node_at(oms_solver *s, int gx, int gy, int gz) {
    const int i = gx - s->lo[0], j = gy - s->lo[1], k = gz - s->lo[2];
    if (i < 0 || j < 0 || k < 0 || i >= s->nb[0] || j >= s->nb[1] || k >= s->nb[2]) return NULL;
    return &s->grid[((size_t)k * s->nb[1] + j) * s->nb[0] + i];
}

/* lodClassifyKernel, :655-676 */
static void lod_classify(oms_solver *s) {
    for (int i = 0; i < s->n; ++i) {
        oms_particle *p = &s->p[i];
        if (p->status != OMS_ACTIVE && p->status != OMS_CHIP) continue;
        const double dx = p->x - s->tool_pos[0], dy = p->y - s->tool_pos[1], dz = p->z - s->tool_pos[2];
        const double d = sqrt(dx * dx + dy * dy + dz * dz);
        p->lod = d <= s->prm.lod_active_radius ? OMS_LOD_ACTIVE
                 : (d <= s->prm.lod_near_radius ? OMS_LOD_NEAR : OMS_LOD_FAR);
    }
}

/* p2gKernel, :124-412: kill zone, constitutive update, APIC scatter */
static void p2g(oms_solver *s) {
    const oms_params *c = &s->prm;
    const double inv = 1.0 / c->dx;
    for (int pi = 0; pi < s->n; ++pi) {
        oms_particle *p = &s->p[pi];
        if (p->status == OMS_INACTIVE) continue;
        const int cx = (int)floor((p->x - c->domain_min[0]) * inv);
        const int cy = (int)floor((p->y - c->domain_min[1]) * inv);
        const int cz = (int)floor((p->z - c->domain_min[2]) * inv);
        if (cx < 1 || cx >= s->dims[0] - 2 || cy < 1 || cy >= s->dims[1] - 2 || cz < 1 || cz >= s->dims[2] - 2) {
            p->status = OMS_INACTIVE;
            continue;
        }
        int b[3];
        double w[3][3], dw[3][3];
        weights(p->x - c->domain_min[0], inv, &b[0], w[0], dw[0]);
        weights(p->y - c->domain_min[1], inv, &b[1], w[1], dw[1]);
        weights(p->z - c->domain_min[2], inv, &b[2], w[2], dw[2]);
        const int fixed = p->status == OMS_FIXED;
        double sigma[9], J;
        int skip;
        oms_p2g_stress(c, p, sigma, &J, &skip);
        const double V0 = p->volume;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (int k = 0; k < 3; ++k) {
                    const int gx = b[0] + i, gy = b[1] + j, gz = b[2] + k;
                    if (gx < 0 || gx >= s->dims[0] || gy < 0 || gy >= s->dims[1] || gz < 0 || gz >= s->dims[2])
                        continue;
                    oms_node *nd = node_at(s, gx, gy, gz);
                    const double wt = w[0][i] * w[1][j] * w[2][k];
                    const double gxw = dw[0][i] * w[1][j] * w[2][k];
                    const double gyw = w[0][i] * dw[1][j] * w[2][k];
                    const double gzw = w[0][i] * w[1][j] * dw[2][k];
                    const double dpx = (gx * c->dx + c->domain_min[0]) - p->x;
                    const double dpy = (gy * c->dx + c->domain_min[1]) - p->y;
                    const double dpz = (gz * c->dx + c->domain_min[2]) - p->z;
                    const double ax = p->C[0] * dpx + p->C[1] * dpy + p->C[2] * dpz;
                    const double ay = p->C[3] * dpx + p->C[4] * dpy + p->C[5] * dpz;
                    const double az = p->C[6] * dpx + p->C[7] * dpy + p->C[8] * dpz;
                    nd->mass += wt * p->mass;
                    if (!fixed) {
                        nd->px += wt * p->mass * (p->vx + ax);
                        nd->py += wt * p->mass * (p->vy + ay);
                        nd->pz += wt * p->mass * (p->vz + az);
                    }
                    double fx = 0, fy = 0, fz = 0;
                    if (!skip) {
                        fx = -V0 * J * (sigma[0] * gxw + sigma[1] * gyw + sigma[2] * gzw);
                        fy = -V0 * J * (sigma[3] * gxw + sigma[4] * gyw + sigma[5] * gzw);
                        fz = -V0 * J * (sigma[6] * gxw + sigma[7] * gyw + sigma[8] * gzw);
                    }
                    nd->fx += fx + wt * p->efx;
                    nd->fy += fy + wt * p->efy;
                    nd->fz += fz + wt * p->efz;
                }
    }
}

/* updateGridKernel (:424-442), applyFixtureBCKernel (:81-111),
 * rasterizeToolKernel (:444-473; a node marked by several tool nodes takes
 * the velocity of the last one, here the highest index) and
 * applyGridCollisionKernel (:475-492) */
static void grid_update(oms_solver *s, double dt) {
    const oms_params *c = &s->prm;
    const size_t nn = (size_t)s->nb[0] * s->nb[1] * s->nb[2];
    for (size_t i = 0; i < nn; ++i) {
        oms_node *g = &s->grid[i];
        if (g->mass > 1e-12) {
            g->pz += g->mass * kGravityZ * dt;
            g->px += g->fx * dt;
            g->py += g->fy * dt;
            g->pz += g->fz * dt;
            g->vx = g->px / g->mass;
            g->vy = g->py / g->mass;
            g->vz = g->pz / g->mass;
        }
    }
    if (c->fixture_thickness > 0.0) {
        for (int k = 0; k < s->nb[2]; ++k) {
            const int gz = s->lo[2] + k;
            const double wz = gz * c->dx + c->domain_min[2];
            const double zr = (wz - c->domain_min[2]) / c->fixture_thickness;
            if (zr > 1.0) continue;
            for (int j = 0; j < s->nb[1]; ++j)
                for (int i = 0; i < s->nb[0]; ++i) {
                    oms_node *g = &s->grid[((size_t)k * s->nb[1] + j) * s->nb[0] + i];
                    if (gz == 0) {
                        g->vx = g->vy = g->vz = 0.0;
                    } else if (zr > 0.0) {
                        const double sg = c->fixture_damping * (1.0 - zr) * (1.0 - zr);
                        const double f = 1.0 / (1.0 + sg * dt);
                        g->vx *= f; g->vy *= f; g->vz *= f;
                    }
                }
        }
    }
    if (s->ntool > 0) {
        const double inv = 1.0 / c->dx;
        for (int t = 0; t < s->ntool; ++t) {
            const int gx = (int)round((s->tool_xyz[3 * t] - c->domain_min[0]) * inv);
            const int gy = (int)round((s->tool_xyz[3 * t + 1] - c->domain_min[1]) * inv);
            const int gz = (int)round((s->tool_xyz[3 * t + 2] - c->domain_min[2]) * inv);
            for (int i = 0; i < 2; ++i)
                for (int j = 0; j < 2; ++j)
                    for (int k = 0; k < 2; ++k) {
                        const int x = gx + i, y = gy + j, z = gz + k;
                        if (x < 0 || x >= s->dims[0] || y < 0 || y >= s->dims[1] || z < 0 || z >= s->dims[2]) continue;
                        oms_node *g = node_at(s, x, y, z);
                        if (!g) continue;   /* no particle reaches it: nothing reads it */
                        g->tool = t + 1;
                        g->tvx = s->tool_vel[3 * t];
                        g->tvy = s->tool_vel[3 * t + 1];
                        g->tvz = s->tool_vel[3 * t + 2];
                    }
        }
        for (size_t i = 0; i < nn; ++i) {
            oms_node *g = &s->grid[i];
            if (g->mass > 1e-12 && g->tool) {
                g->vx = g->tvx; g->vy = g->tvy; g->vz = g->tvz;
                g->px = g->mass * g->vx; g->py = g->mass * g->vy; g->pz = g->mass * g->vz;
            }
        }
    }
}

/* g2pKernel (:506-605), then resetExternalForceKernel (:682-688) */
static void g2p(oms_solver *s, double dt) {
    const oms_params *c = &s->prm;
    const double inv = 1.0 / c->dx;
    for (int pi = 0; pi < s->n; ++pi) {
        oms_particle *p = &s->p[pi];
        if (p->status == OMS_INACTIVE) continue;
        if (p->status == OMS_FIXED) {
            p->vx = p->vy = p->vz = 0.0;
            continue;
        }
        int b[3];
        double w[3][3], dw[3][3];
        weights(p->x - c->domain_min[0], inv, &b[0], w[0], dw[0]);
        weights(p->y - c->domain_min[1], inv, &b[1], w[1], dw[1]);
        weights(p->z - c->domain_min[2], inv, &b[2], w[2], dw[2]);
        double v[3] = {0, 0, 0}, B[9] = {0}, L[9] = {0};
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (int k = 0; k < 3; ++k) {
                    const int gx = b[0] + i, gy = b[1] + j, gz = b[2] + k;
                    if (gx < 0 || gx >= s->dims[0] || gy < 0 || gy >= s->dims[1] || gz < 0 || gz >= s->dims[2])
                        continue;
                    const oms_node *g = node_at(s, gx, gy, gz);
                    if (!g || g->mass < 1e-12) continue;
                    const double wt = w[0][i] * w[1][j] * w[2][k];
                    const double gw[3] = {dw[0][i] * w[1][j] * w[2][k], w[0][i] * dw[1][j] * w[2][k],
                                          w[0][i] * w[1][j] * dw[2][k]};
                    const double gv[3] = {g->vx, g->vy, g->vz};
                    const double dp[3] = {(gx * c->dx + c->domain_min[0]) - p->x, (gy * c->dx + c->domain_min[1]) - p->y,
                                          (gz * c->dx + c->domain_min[2]) - p->z};
                    for (int r = 0; r < 3; ++r) {
                        v[r] += wt * gv[r];
                        for (int q = 0; q < 3; ++q) {
                            B[r * 3 + q] += wt * gv[r] * dp[q];
                            L[r * 3 + q] += gv[r] * gw[q];
                        }
                    }
                }
        p->vx = v[0]; p->vy = v[1]; p->vz = v[2];
        p->x += v[0] * dt; p->y += v[1] * dt; p->z += v[2] * dt;
        if (p->status == OMS_CHIP) {
            p->strain_rate = 0;
            continue;
        }
        const double Dinv = 3.0 * inv * inv;
        for (int i = 0; i < 9; ++i) p->C[i] = B[i] * Dinv;
        if (p->lod == OMS_LOD_FAR)
            for (int i = 0; i < 9; ++i) p->C[i] = 0.0;
        else
            for (int i = 0; i < 9; ++i) p->C[i] *= 0.99;
        double M[9], nF[9];
        for (int i = 0; i < 9; ++i) {
            M[i] = L[i] * dt;
            if (i % 4 == 0) M[i] += 1.0;
        }
        mul3(M, p->F, nF);
        memcpy(p->F, nF, sizeof(nF));
        p->strain_rate = sqrt(2.0 / 3.0 * (L[0] * L[0] + L[4] * L[4] + L[8] * L[8] + 0.5 * (L[1] + L[3]) * (L[1] + L[3]) +
                                           0.5 * (L[2] + L[6]) * (L[2] + L[6]) + 0.5 * (L[5] + L[7]) * (L[5] + L[7])));
    }
    for (int pi = 0; pi < s->n; ++pi) s->p[pi].efx = s->p[pi].efy = s->p[pi].efz = 0.0;
}

/* chipFractureAndSeparationKernel of MPMSolver (:1007-1060), run after the
 * step's kernels when the step index is even, with critical damage 1e-6 */
static void chip_pass(oms_solver *s) {
    const double tipz = s->tool_pos[2], cut = s->prm.dx * 1.5, th = s->prm.damage_threshold;
    for (int i = 0; i < s->n; ++i) {
        oms_particle *p = &s->p[i];
        if (p->status == OMS_INACTIVE) continue;
        if (p->status == OMS_ACTIVE) {
            if (p->z < tipz - cut && p->damage > 1e-6) {
                p->status = OMS_CHIP;
                p->residual_stress = fmax(p->residual_stress, sqrt(p->s[0] * p->s[0] + p->s[1] * p->s[1] + p->s[2] * p->s[2]));
                for (int k = 0; k < 6; ++k) p->s[k] = 0.0;
                p->damage = fmax(p->damage, th);
            }
            continue;
        }
        if (p->status == OMS_CHIP) {
            if (p->damage < th * 1.5) continue;
            const uint32_t seed = (uint32_t)p->id * 2654435761u;
            const double rx = (double)((seed >> 16) & 0xFF) / 255.0 - 0.5;
            const double ry = (double)((seed >> 8) & 0xFF) / 255.0 - 0.5;
            const double rz = (double)(seed & 0xFF) / 255.0 - 0.5;
            const double sp = sqrt(p->vx * p->vx + p->vy * p->vy + p->vz * p->vz);
            const double mag = sp * 0.3 + 0.5;
            p->vx += rx * mag; p->vy += ry * mag; p->vz += rz * mag;
            p->damage = th * 0.5;
        }
    }
}

/* MPMSolver::step, :1062-1168 */
void oms_step(oms_solver *s, double dt) {
    if (s->n == 0) return;
    if (s->prm.lod_enabled) lod_classify(s);
    fit_box(s);
    p2g(s);
    grid_update(s, dt);
    g2p(s, dt);
    if (s->step % 2 == 0) chip_pass(s);
    s->time += dt;
    s->step++;
}

/* getStableTimeStep, :1170-1178 (cfl dx / max(sqrt(E / rho_scaled), 1), at least 1e-10) */
double oms_stable_dt(const oms_solver *s) {
    const double c = sqrt(s->prm.young / (s->prm.density * s->prm.mass_scaling));
    return fmax(s->prm.cfl * s->prm.dx / fmax(c, 1.0), 1e-10);
}

/* computeParticleStatsKernel (:726-785) over ACTIVE and CHIP particles */
void oms_metrics(const oms_solver *s, double out[4]) {
    double ms = 0, mt = 0, ke = 0, ie = 0;
    for (int i = 0; i < s->n; ++i) {
        const oms_particle *p = &s->p[i];
        if (p->status != OMS_ACTIVE && p->status != OMS_CHIP) continue;
        const double tr = p->s[0] + p->s[1] + p->s[2];
        const double a = p->s[0] - tr / 3.0, b = p->s[1] - tr / 3.0, cc = p->s[2] - tr / 3.0;
        const double vm = sqrt(1.5 * (a * a + b * b + cc * cc + 2.0 * (p->s[3] * p->s[3] + p->s[4] * p->s[4] + p->s[5] * p->s[5])));
        ms = fmax(ms, vm);
        mt = fmax(mt, p->temperature);
        ke += 0.5 * p->mass * (p->vx * p->vx + p->vy * p->vy + p->vz * p->vz);
        ie += (vm * vm * p->mass) / (6.0 * 80e9) + p->volume * p->plastic_strain * vm;
    }
    out[0] = ms; out[1] = mt; out[2] = ke; out[3] = ie;
}

void oms_bounds(const oms_solver *s, double out[6]) {
    out[0] = out[1] = out[2] = INFINITY;
    out[3] = out[4] = out[5] = -INFINITY;
    for (int i = 0; i < s->n; ++i) {
        const oms_particle *p = &s->p[i];
        if (p->status != OMS_ACTIVE && p->status != OMS_CHIP) continue;
        const double r[3] = {p->x, p->y, p->z};
        for (int a = 0; a < 3; ++a) {
            out[a] = fmin(out[a], r[a]);
            out[3 + a] = fmax(out[3 + a], r[a]);
        }
    }
}

/* applyExternalForceKernel (:607-627): ACTIVE and CHIP particles within
 * the radius get f (1 - d / r), replacing what they held; the call does
 * nothing when |f| < 1e-12 (MPMSolver.cu:1444) */
void oms_apply_external_force(oms_solver *s, const double c[3], double radius, const double f[3]) {
    if (sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]) < 1e-12) return;
    for (int i = 0; i < s->n; ++i) {
        oms_particle *p = &s->p[i];
        if (p->status != OMS_ACTIVE && p->status != OMS_CHIP) continue;
        const double dx = p->x - c[0], dy = p->y - c[1], dz = p->z - c[2];
        const double d = sqrt(dx * dx + dy * dy + dz * dz);
        if (d < radius) {
            const double w = 1.0 - d / radius;
            p->efx = f[0] * w; p->efy = f[1] * w; p->efz = f[2] * w;
        }
    }
}

void oms_grid(const oms_solver *s, int32_t lo[3], int32_t n[3], double *mass, double *mom, double *force) {
    for (int a = 0; a < 3; ++a) { lo[a] = s->lo[a]; n[a] = s->nb[a]; }
    const size_t nn = (size_t)s->nb[0] * s->nb[1] * s->nb[2];
    for (size_t i = 0; mass && i < nn; ++i) mass[i] = s->grid[i].mass;
    /* momentum after the grid update and collision; force of the P2G */
    for (size_t i = 0; mom && i < nn; ++i) {
        mom[3 * i] = s->grid[i].px; mom[3 * i + 1] = s->grid[i].py; mom[3 * i + 2] = s->grid[i].pz;
    }
    for (size_t i = 0; force && i < nn; ++i) {
        force[3 * i] = s->grid[i].fx; force[3 * i + 1] = s->grid[i].fy; force[3 * i + 2] = s->grid[i].fz;
    }
}

oms_particle *oms_particle_data(oms_solver *s, int *n) {
    if (n) *n = s->n;
    return s->p;
}

void oms_domain_min(const oms_solver *s, double out[3]) {
    for (int a = 0; a < 3; ++a) out[a] = s->prm.domain_min[a];
}
