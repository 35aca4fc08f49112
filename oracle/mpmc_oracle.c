/*
 * mpmc_oracle.c -- CPU FP64 restatement of the tool FEM step and the
 * spatial-hash XPBD contact (src/FEMSolver.cu, src/ContactSolver.cu,
 * src/SharedSpatialHash.cu of Up25-pathan/edegpredict-engine-rotary), serial.
 * TEST INFRASTRUCTURE ONLY (the checker of the parity tests; see
 * mpmc_oracle.h).  Each function cites the reference lines it restates.
 */
#include "mpmc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static const double kPi = 3.14159265358979323846;   /* constants::PI, Types.h:514 */

typedef struct {
    int32_t n[4];
    double volume;
    double inv_dm[9];
} omt_elem;

struct omt_tool {
    omt_params prm;
    omt_node *surf, *tet;
    int ns, nt;
    omt_elem *el;
    int ne;
    int32_t *emb_surf, *emb_tet;
    double *emb_w;
    int nemb;
    double tool_pos[3], tool_axis[3], tool_angle, tool_omega;
    double spindle_pos[3], spindle_vel[3];
    double typical_len;
    double time;
    int64_t step;
};

void omt_params_default(omt_params *p) {
    memset(p, 0, sizeof(*p));
    p->young = 600e9;                 /* FEMMaterialProps, FEMSolver.cuh:42-47 */
    p->poisson = 0.22;
    p->density = 14500.0;
    p->thermal_conductivity = 80.0;
    p->specific_heat = 200.0;
    p->damping_ratio = 0.1;           /* m_globalDamping, FEMSolver.cuh:235 */
    p->mass_scaling = 1.0;
    p->coolant_htc = 1.0e3;           /* m_coolantHTC, :271 */
    p->spindle_stiffness = 5e7;       /* :265-266 */
    p->spindle_damping = 1e4;
    p->spindle_enabled = 1;           /* :264 */
    p->thermal_skip = 10;             /* :205 */
    p->polar_textbook = 0;
}

omt_tool *omt_create(const omt_params *p) {
    omt_tool *t = (omt_tool *)calloc(1, sizeof(omt_tool));
    t->prm = *p;
    t->tool_axis[2] = 1.0;            /* FEMSolver() ctor, :841 */
    t->typical_len = 0.001;           /* :252 */
    return t;
}

void omt_destroy(omt_tool *t) {
    if (!t) return;
    free(t->surf); free(t->tet); free(t->el);
    free(t->emb_surf); free(t->emb_tet); free(t->emb_w);
    free(t);
}

static void node_init(omt_node *n, double x, double y, double z, int id, int region) {
    memset(n, 0, sizeof(*n));
    n->x = n->ox = x; n->y = n->oy = y; n->z = n->oz = z;
    n->temperature = 25.0;
    n->id = id;
    n->status = OMT_OK;
    n->region = region;
}

/* initializeFromMesh (FEMSolver.cu:900-1010) and createElementsFromMesh
 * (:1012-1078) */
void omt_set_mesh(omt_tool *t, const double *surf_xyz, const int32_t *surf_region, int ns, const double *tet_xyz,
                  const int32_t *tet_region, int nt, const int32_t *elems, int ne, const int32_t *emb_surf,
                  const int32_t *emb_tet, const double *emb_w, int nemb) {
    free(t->surf); free(t->tet); free(t->el);
    free(t->emb_surf); free(t->emb_tet); free(t->emb_w);
    t->ns = ns; t->nt = nt; t->ne = ne; t->nemb = nemb;
    t->surf = (omt_node *)calloc(ns > 0 ? ns : 1, sizeof(omt_node));
    t->tet = (omt_node *)calloc(nt > 0 ? nt : 1, sizeof(omt_node));
    t->el = (omt_elem *)calloc(ne > 0 ? ne : 1, sizeof(omt_elem));
    t->emb_surf = (int32_t *)calloc(nemb > 0 ? nemb : 1, sizeof(int32_t));
    t->emb_tet = (int32_t *)calloc(nemb > 0 ? 4 * nemb : 4, sizeof(int32_t));
    t->emb_w = (double *)calloc(nemb > 0 ? 4 * nemb : 4, sizeof(double));
    const omt_params *P = &t->prm;
    for (int i = 0; i < ns; ++i) {   /* :907-925 */
        node_init(&t->surf[i], surf_xyz[3 * i], surf_xyz[3 * i + 1], surf_xyz[3 * i + 2], i,
                  surf_region ? surf_region[i] : 0);
        t->surf[i].mass = P->density * 1e-9 * P->mass_scaling;
    }
    double maxz = -1e10;
    for (int i = 0; i < nt; ++i) {   /* :928-945 */
        node_init(&t->tet[i], tet_xyz[3 * i], tet_xyz[3 * i + 1], tet_xyz[3 * i + 2], i, tet_region ? tet_region[i] : 0);
        t->tet[i].mass = 1e-12;
        t->tet[i].fixed = 1;
        if (t->tet[i].oz > maxz) maxz = t->tet[i].oz;
    }
    for (int i = 0; i < nt; ++i) t->tet[i].fixed = t->tet[i].oz > maxz - 0.005 ? 1 : 0;   /* :955-962 */
    if (nemb > 0) {
        memcpy(t->emb_surf, emb_surf, sizeof(int32_t) * nemb);
        memcpy(t->emb_tet, emb_tet, sizeof(int32_t) * 4 * nemb);
        memcpy(t->emb_w, emb_w, sizeof(double) * 4 * nemb);
    }
    /* createElementsFromMesh */
    for (int i = 0; i < nt; ++i) t->tet[i].mass = 0.0;
    for (int e = 0; e < ne; ++e) {
        omt_elem *E = &t->el[e];
        for (int k = 0; k < 4; ++k) E->n[k] = elems[4 * e + k];
        const omt_node *a = &t->tet[E->n[0]];
        double X1[3] = {t->tet[E->n[1]].ox, t->tet[E->n[1]].oy, t->tet[E->n[1]].oz};
        double X2[3] = {t->tet[E->n[2]].ox, t->tet[E->n[2]].oy, t->tet[E->n[2]].oz};
        const double X3[3] = {t->tet[E->n[3]].ox, t->tet[E->n[3]].oy, t->tet[E->n[3]].oz};
        const double X0[3] = {a->ox, a->oy, a->oz};
        double D[3][3];
        for (int i = 0; i < 3; ++i) { D[i][0] = X1[i] - X0[i]; D[i][1] = X2[i] - X0[i]; D[i][2] = X3[i] - X0[i]; }
        double det = D[0][0] * (D[1][1] * D[2][2] - D[1][2] * D[2][1]) - D[0][1] * (D[1][0] * D[2][2] - D[1][2] * D[2][0]) +
                     D[0][2] * (D[1][0] * D[2][1] - D[1][1] * D[2][0]);
        if (det < 0.0) {   /* :1041-1046 */
            int32_t s = E->n[1]; E->n[1] = E->n[2]; E->n[2] = s;
            for (int i = 0; i < 3; ++i) { double u = X1[i]; X1[i] = X2[i]; X2[i] = u; }
            for (int i = 0; i < 3; ++i) { D[i][0] = X1[i] - X0[i]; D[i][1] = X2[i] - X0[i]; D[i][2] = X3[i] - X0[i]; }
            det = -det;
        }
        E->volume = det / 6.0;
        if (E->volume < 1e-18) { E->volume = 1e-18; det = 6e-18; }
        const double id = 1.0 / det;   /* :1054-1065 */
        double *M = E->inv_dm;
        M[0] = (D[1][1] * D[2][2] - D[1][2] * D[2][1]) * id;
        M[1] = (D[0][2] * D[2][1] - D[0][1] * D[2][2]) * id;
        M[2] = (D[0][1] * D[1][2] - D[0][2] * D[1][1]) * id;
        M[3] = (D[1][2] * D[2][0] - D[1][0] * D[2][2]) * id;
        M[4] = (D[0][0] * D[2][2] - D[0][2] * D[2][0]) * id;
        M[5] = (D[0][2] * D[1][0] - D[0][0] * D[1][2]) * id;
        M[6] = (D[1][0] * D[2][1] - D[1][1] * D[2][0]) * id;
        M[7] = (D[0][1] * D[2][0] - D[0][0] * D[2][1]) * id;
        M[8] = (D[0][0] * D[1][1] - D[0][1] * D[1][0]) * id;
        double mpn = (P->density * E->volume * P->mass_scaling) / 4.0;   /* :1069-1075 */
        if (mpn < 1e-12) mpn = 1e-12;
        for (int k = 0; k < 4; ++k) t->tet[E->n[k]].mass += mpn;
    }
    for (int i = 0; i < nemb; ++i) {   /* :973-992 */
        const int s = t->emb_surf[i];
        if (s < 0 || s >= ns) continue;
        double total = 0.0;
        for (int k = 0; k < 4; ++k) {
            const int ti = t->emb_tet[4 * i + k];
            if (ti >= 0 && ti < nt) total += t->tet[ti].mass * t->emb_w[4 * i + k];
        }
        if (total > 1e-15) t->surf[s].mass = total;
    }
}

void omt_init_driven(omt_tool *t, double top_fraction) {   /* :1367-1408 */
    if (t->ns == 0) return;
    double mn = 1e30, mx = -1e30;
    for (int i = 0; i < t->nt; ++i) {
        if (t->tet[i].z < mn) mn = t->tet[i].z;
        if (t->tet[i].z > mx) mx = t->tet[i].z;
    }
    const double thr = mx - (mx - mn) * top_fraction;
    for (int i = 0; i < t->nt; ++i) {
        omt_node *n = &t->tet[i];
        if (n->z >= thr) {
            n->driven = 1;
            n->loff[0] = n->ox; n->loff[1] = n->oy; n->loff[2] = n->oz;
        } else {
            n->driven = 0;
        }
        if (t->prm.spindle_enabled) n->fixed = 0;
    }
}

void omt_set_spindle_state(omt_tool *t, const double pos[3], const double vel[3]) {
    for (int a = 0; a < 3; ++a) { t->spindle_pos[a] = pos[a]; t->spindle_vel[a] = vel[a]; }
}

void omt_set_transform(omt_tool *t, const double pos[3], const double axis[3], double omega) {   /* :1473-1478 */
    const double l = sqrt(axis[0] * axis[0] + axis[1] * axis[1] + axis[2] * axis[2]);
    for (int a = 0; a < 3; ++a) {
        t->tool_pos[a] = pos[a];
        /* Vec3::normalized (Types.h:185-188): zero for a length below 1e-12 */
        t->tool_axis[a] = l > 1e-12 ? axis[a] / l : 0.0;
    }
    t->tool_omega = omega;
}

static void rigid_velocity(omt_node *n, int count, const double v[3], double wz, double cx, double cy) {
    for (int i = 0; i < count; ++i) {   /* setRigidBodyVelocitiesKernel :787-798 */
        const double rx = n[i].x - cx, ry = n[i].y - cy;
        n[i].vx = v[0] - wz * ry;
        n[i].vy = v[1] + wz * rx;
        n[i].vz = v[2];
    }
}

void omt_set_rigid_velocity(omt_tool *t, const double v[3], double omega_z, double cx, double cy) {
    rigid_velocity(t->surf, t->ns, v, omega_z, cx, cy);
    rigid_velocity(t->tet, t->nt, v, omega_z, cx, cy);
}

void omt_apply_contact_force(omt_tool *t, int i, const double f[3], double heat) {   /* :1524-1536 */
    if (i < 0 || i >= t->ns) return;
    omt_node *n = &t->surf[i];
    n->fx += f[0]; n->fy += f[1]; n->fz += f[2];
    n->in_contact = 1;
    n->temperature += heat / (n->mass * t->prm.specific_heat);
}

/* applyRotationKernel (:685-725): Rodrigues rotation of x and of the
 * original position about the axis through c */
static void rotate(omt_node *n, int count, const double c[3], const double ax[3], double angle) {
    const double co = cos(angle), si = sin(angle), omc = 1.0 - co;
    for (int i = 0; i < count; ++i) {
        omt_node *q = &n[i];
        double px = q->x - c[0], py = q->y - c[1], pz = q->z - c[2];
        double dot = ax[0] * px + ax[1] * py + ax[2] * pz;
        double cxp = ax[1] * pz - ax[2] * py, cyp = ax[2] * px - ax[0] * pz, czp = ax[0] * py - ax[1] * px;
        q->x = px * co + cxp * si + ax[0] * dot * omc + c[0];
        q->y = py * co + cyp * si + ax[1] * dot * omc + c[1];
        q->z = pz * co + czp * si + ax[2] * dot * omc + c[2];
        px = q->ox - c[0]; py = q->oy - c[1]; pz = q->oz - c[2];
        dot = ax[0] * px + ax[1] * py + ax[2] * pz;
        cxp = ax[1] * pz - ax[2] * py; cyp = ax[2] * px - ax[0] * pz; czp = ax[0] * py - ax[1] * px;
        q->ox = px * co + cxp * si + ax[0] * dot * omc + c[0];
        q->oy = py * co + cyp * si + ax[1] * dot * omc + c[1];
        q->oz = pz * co + czp * si + ax[2] * dot * omc + c[2];
    }
}

/* The polar iteration of computeElementForcesKernel (:217-278): scaled
 * Newton steps R <- (g R + g^-1 Y) / 2 with Y = R^-1 (the reference) or
 * R^-T (textbook), g from the 1- and inf-norms, at most 20 steps, stop when
 * no entry moves by 1e-6 or more. */
void omt_polar(const double F[9], int textbook, double R[9]) {
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
        for (int i = 0; i < 3; ++i) {
            const double sr = fabs(R[3 * i]) + fabs(R[3 * i + 1]) + fabs(R[3 * i + 2]);
            const double sc = (fabs(cof[i]) + fabs(cof[3 + i]) + fabs(cof[6 + i])) * idet;
            if (sr > nRinf) nRinf = sr;
            if (sc > nC1) nC1 = sc;
        }
        for (int j = 0; j < 3; ++j) {
            const double sr = fabs(R[j]) + fabs(R[3 + j]) + fabs(R[6 + j]);
            const double sc = (fabs(cof[3 * j]) + fabs(cof[3 * j + 1]) + fabs(cof[3 * j + 2])) * idet;
            if (sr > nR1) nR1 = sr;
            if (sc > nCinf) nCinf = sc;
        }
        const double g = pow((nC1 * nCinf) / fmax(1e-12, nR1 * nRinf), 0.25), ig = 1.0 / g;
        double md = 0;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                /* R^-1[i][j] = cof[j][i] / det; R^-T[i][j] = cof[i][j] / det;
                 * g^-1 * cofactor * (1 / det) left to right as the reference */
                const double y = ig * (textbook ? cof[3 * i + j] : cof[3 * j + i]) * idet;
                const double nv = 0.5 * (g * R[3 * i + j] + y);
                const double d = fabs(R[3 * i + j] - nv);
                if (d > md) md = d;
                R[3 * i + j] = nv;
            }
        if (md < 1e-6) break;
    }
}

/* computeElementForcesKernel (:167-342) */
static void element_forces(omt_tool *t, const omt_elem *E) {
    const int N = t->nt;
    const int n0 = E->n[0], n1 = E->n[1], n2 = E->n[2], n3 = E->n[3];
    if (n0 < 0 || n0 >= N || n1 >= N || n2 >= N || n3 >= N) return;
    omt_node *nd[4] = {&t->tet[n0], &t->tet[n1], &t->tet[n2], &t->tet[n3]};
    double Ds[3][3];
    for (int c = 0; c < 3; ++c) {
        Ds[0][c] = nd[c + 1]->x - nd[0]->x;
        Ds[1][c] = nd[c + 1]->y - nd[0]->y;
        Ds[2][c] = nd[c + 1]->z - nd[0]->z;
    }
    const double *M = E->inv_dm;
    double F[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) F[3 * i + j] = Ds[i][0] * M[j] + Ds[i][1] * M[3 + j] + Ds[i][2] * M[6 + j];
    const double detF = F[0] * (F[4] * F[8] - F[5] * F[7]) - F[1] * (F[3] * F[8] - F[5] * F[6]) +
                        F[2] * (F[3] * F[7] - F[4] * F[6]);
    double erosion = 1.0;
    if (detF < 0.1) {   /* :201-214 */
        erosion = detF / 0.1;
        erosion = erosion * erosion;
        const double sc = cbrt(0.1 / fmax(detF, 1e-9));
        for (int i = 0; i < 9; ++i) F[i] *= sc;
    }
    if (erosion < 1e-6) return;
    double R[9];
    omt_polar(F, t->prm.polar_textbook, R);
    double eps[9];   /* R^T F - I */
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0;
            for (int k = 0; k < 3; ++k) s += R[3 * k + i] * F[3 * k + j];
            eps[3 * i + j] = s - (i == j ? 1.0 : 0.0);
        }
    const double E_eff = t->prm.young * erosion, nu = t->prm.poisson;
    const double lambda = (E_eff * nu) / ((1.0 + nu) * (1.0 - 2.0 * nu));
    const double mu = E_eff / (2.0 * (1.0 + nu));
    const double tr = eps[0] + eps[4] + eps[8];
    double sig[9], P[9], H[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) sig[3 * i + j] = 2.0 * mu * eps[3 * i + j] + (i == j ? lambda * tr : 0.0);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            P[3 * i + j] = R[3 * i] * sig[j] + R[3 * i + 1] * sig[3 + j] + R[3 * i + 2] * sig[6 + j];
    const double V0 = E->volume;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            H[3 * i + j] = -V0 * (P[3 * i] * M[3 * j] + P[3 * i + 1] * M[3 * j + 1] + P[3 * i + 2] * M[3 * j + 2]);
    double f[4][3];
    for (int c = 0; c < 3; ++c)
        for (int a = 0; a < 3; ++a) f[c + 1][a] = H[3 * a + c];
    for (int a = 0; a < 3; ++a) f[0][a] = -(f[1][a] + f[2][a] + f[3][a]);
    const double w = V0 * 0.25;
    for (int k = 0; k < 4; ++k) {
        nd[k]->fx += f[k][0]; nd[k]->fy += f[k][1]; nd[k]->fz += f[k][2];
        nd[k]->s[0] += P[0] * w; nd[k]->s[1] += P[4] * w; nd[k]->s[2] += P[8] * w;
        nd[k]->s[3] += P[1] * w; nd[k]->s[4] += P[2] * w; nd[k]->s[5] += P[5] * w;
        nd[k]->stress_weight += w;
    }
}

/* thermalConductionTetKernel (:38-84) into conduction buffers */
static void element_conduction(omt_tool *t, const omt_elem *E, double dt) {
    const int N = t->nt;
    const int n0 = E->n[0], n3 = E->n[3];
    if (n0 < 0 || n0 >= N || n3 >= N) return;
    const double *M = E->inv_dm;
    double g[3][4];
    for (int i = 0; i < 3; ++i) {
        g[i][1] = M[i]; g[i][2] = M[3 + i]; g[i][3] = M[6 + i];
        g[i][0] = -(g[i][1] + g[i][2] + g[i][3]);
    }
    double T[4], Q[4] = {0, 0, 0, 0};
    for (int k = 0; k < 4; ++k) T[k] = t->tet[E->n[k]].temperature;
    const double kVdt = t->prm.thermal_conductivity * E->volume * dt;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            const double d = g[0][i] * g[0][j] + g[1][i] * g[1][j] + g[2][i] * g[2][j];
            Q[i] -= kVdt * d * T[j];
        }
    for (int k = 0; k < 4; ++k) {
        omt_node *n = &t->tet[E->n[k]];
        if (n->mass > 1e-15 && t->prm.specific_heat > 0) n->conduction += Q[k];
    }
}

static void apply_tet_heat(omt_tool *t) {   /* applyTetHeatKernel (:87-97) */
    const double cp = t->prm.specific_heat;
    for (int i = 0; i < t->nt; ++i) {
        omt_node *n = &t->tet[i];
        if (n->mass > 1e-15 && cp > 0) {
            n->temperature += (n->heat + n->conduction) / (n->mass * cp);
            n->heat = 0.0;
            n->conduction = 0.0;
        }
    }
}

void omt_step(omt_tool *t, double dt) {   /* FEMSolver::step (:1112-1300) */
    if (t->ns == 0) return;
    const omt_params *P = &t->prm;
    if (fabs(t->tool_omega) > 1e-12) {   /* :1119-1138 */
        const double inc = t->tool_omega * dt;
        t->tool_angle += inc;
        rotate(t->surf, t->ns, t->tool_pos, t->tool_axis, inc);
        if (t->nt > 0) rotate(t->tet, t->nt, t->tool_pos, t->tool_axis, inc);
    }
    if (P->spindle_enabled) {
        /* transferForcesToTetNodesKernel (:423-450) */
        for (int e = 0; e < t->nemb; ++e) {
            const int s = t->emb_surf[e];
            if (s < 0) continue;
            omt_node *sn = &t->surf[s];
            if (fabs(sn->fx) > 1e-12 || fabs(sn->fy) > 1e-12 || fabs(sn->fz) > 1e-12) {
                for (int k = 0; k < 4; ++k) {
                    const int ti = t->emb_tet[4 * e + k];
                    if (ti >= 0) {
                        t->tet[ti].fx += sn->fx * t->emb_w[4 * e + k];
                        t->tet[ti].fy += sn->fy * t->emb_w[4 * e + k];
                        t->tet[ti].fz += sn->fz * t->emb_w[4 * e + k];
                    }
                }
                sn->fx = sn->fy = sn->fz = 0.0;
            }
        }
        /* computeSpindleForcesKernel (:385-420) */
        for (int i = 0; i < t->nt; ++i) {
            omt_node *n = &t->tet[i];
            if (!n->driven || n->fixed) continue;
            const double k = P->spindle_stiffness, c = P->spindle_damping;
            n->fx += (t->spindle_pos[0] + n->loff[0] - n->x) * k + (t->spindle_vel[0] - n->vx) * c;
            n->fy += (t->spindle_pos[1] + n->loff[1] - n->y) * k + (t->spindle_vel[1] - n->vy) * c;
            n->fz += (t->spindle_pos[2] + n->loff[2] - n->z) * k + (t->spindle_vel[2] - n->vz) * c;
        }
        for (int i = 0; i < t->nt; ++i) {   /* resetNodeStressKernel (:372-380) */
            memset(t->tet[i].s, 0, sizeof(t->tet[i].s));
            t->tet[i].stress_weight = 0.0;
        }
        for (int e = 0; e < t->ne; ++e) element_forces(t, &t->el[e]);
        for (int i = 0; i < t->nt; ++i) {   /* normalizeNodeStressKernel (:346-369) */
            omt_node *n = &t->tet[i];
            if (n->stress_weight > 1e-30) {
                const double iw = 1.0 / n->stress_weight;
                for (int c = 0; c < 6; ++c) n->s[c] *= iw;
                const double s1 = n->s[0] - n->s[1], s2 = n->s[1] - n->s[2], s3 = n->s[2] - n->s[0];
                const double seq = sqrt(0.5 * (s1 * s1 + s2 * s2 + s3 * s3 +
                                               6.0 * (n->s[3] * n->s[3] + n->s[4] * n->s[4] + n->s[5] * n->s[5])));
                n->von_mises = 0.7 * seq + 0.3 * n->von_mises;
            }
        }
        const double damping = P->damping_ratio * 0.01;
        for (int i = 0; i < t->nt; ++i) {   /* integrateNodesKernel (:525-571) */
            omt_node *n = &t->tet[i];
            if (n->fixed || n->status == OMT_FAILED) {
                n->vx = n->vy = n->vz = 0;
                n->fx = n->fy = n->fz = 0;
                continue;
            }
            const double fm = sqrt(n->fx * n->fx + n->fy * n->fy + n->fz * n->fz);
            const double maxf = n->mass * 1e8;
            if (fm > maxf && fm > 0) {
                const double sc = maxf / fm;
                n->fx *= sc; n->fy *= sc; n->fz *= sc;
            }
            const double im = 1.0 / fmax(n->mass, 1e-12);
            n->vx += n->fx * im * dt; n->vy += n->fy * im * dt; n->vz += n->fz * im * dt;
            n->x += n->vx * dt; n->y += n->vy * dt; n->z += n->vz * dt;
            const double df = exp(-damping * dt);
            n->vx *= df; n->vy *= df; n->vz *= df;
            n->fx = n->fy = n->fz = 0;
        }
        /* updateSurfaceNodesKernel (:452-485) */
        for (int e = 0; e < t->nemb; ++e) {
            const int s = t->emb_surf[e];
            if (s < 0) continue;
            double p[3] = {0, 0, 0}, v[3] = {0, 0, 0};
            for (int k = 0; k < 4; ++k) {
                const int ti = t->emb_tet[4 * e + k];
                if (ti < 0) continue;
                const double w = t->emb_w[4 * e + k];
                p[0] += t->tet[ti].x * w; p[1] += t->tet[ti].y * w; p[2] += t->tet[ti].z * w;
                v[0] += t->tet[ti].vx * w; v[1] += t->tet[ti].vy * w; v[2] += t->tet[ti].vz * w;
            }
            omt_node *sn = &t->surf[s];
            sn->x = p[0]; sn->y = p[1]; sn->z = p[2];
            sn->vx = v[0]; sn->vy = v[1]; sn->vz = v[2];
        }
    }
    if (t->step % 100 == 0 && t->ne > 0) t->typical_len = pow(t->el[0].volume, 1.0 / 3.0);   /* :1230-1233 */
    if (t->nemb > 0 && t->nt > 0) {
        /* transferSurfaceHeatToTetNodesKernel (:487-523), fraction 0.35 */
        const double cp = P->specific_heat;
        for (int e = 0; e < t->nemb; ++e) {
            const int s = t->emb_surf[e];
            if (s < 0) continue;
            omt_node *sn = &t->surf[s];
            if (sn->temperature <= 25.01) continue;
            double tt = 0.0;
            for (int k = 0; k < 4; ++k) {
                const int ti = t->emb_tet[4 * e + k];
                if (ti >= 0) tt += t->tet[ti].temperature * t->emb_w[4 * e + k];
            }
            const double dT = sn->temperature - tt;
            if (dT <= 0.0) continue;
            const double q = dT * fmax(sn->mass, 1e-12) * cp * 0.35;
            sn->temperature -= q / (fmax(sn->mass, 1e-12) * cp);
            for (int k = 0; k < 4; ++k) {
                const int ti = t->emb_tet[4 * e + k];
                if (ti >= 0) t->tet[ti].heat += q * t->emb_w[4 * e + k];
            }
        }
        apply_tet_heat(t);
    }
    if (t->ne > 0 && t->step % P->thermal_skip == 0) {   /* :1241-1272 */
        const double alpha = (P->density > 0.0 && P->specific_heat > 0.0)
                                 ? P->thermal_conductivity / (P->density * P->specific_heat) : 0.0;
        const double h = fmax(t->typical_len, 0.0001);
        const double dtt = dt * P->thermal_skip;
        int nsub = 1;
        if (alpha > 0.0) {
            const double cfl = alpha * dtt / (h * h);
            if (cfl > 0.25) nsub = (int)ceil(cfl / 0.25);
        }
        const double dts = dtt / nsub;
        for (int sub = 0; sub < nsub; ++sub) {
            for (int e = 0; e < t->ne; ++e) element_conduction(t, &t->el[e], dts);
            apply_tet_heat(t);
        }
    }
    /* applyThermalBCKernel (:112-165) with ambient 25, the coolant HTC,
     * Stefan-Boltzmann 5.670374419e-8, emissivity 0.7 (:1276-1288) */
    for (int i = 0; i < t->nt; ++i) {
        omt_node *n = &t->tet[i];
        const double amb = 25.0;
        if (n->region == OMT_REGION_SHANK) {
            n->temperature = amb;
            continue;
        }
        const double vol = fmax(n->mass / P->density, 1e-15);
        const double area = pow(vol, 2.0 / 3.0);
        if (P->coolant_htc > 0.0 && n->temperature > amb + 0.1) {
            const double dT = n->temperature - amb;
            const double q = P->coolant_htc * area * dT * dt;
            const double mc = fmax(n->mass * P->specific_heat, 1e-15);
            n->temperature -= fmin(q / mc, dT * 0.5);
        }
        if (0.7 > 0.0 && n->temperature > amb + 10.0) {
            const double Tn = n->temperature + 273.15, Ta = amb + 273.15;
            const double q = 0.7 * 5.670374419e-8 * area * (Tn * Tn * Tn * Tn - Ta * Ta * Ta * Ta) * dt;
            const double mc = fmax(n->mass * P->specific_heat, 1e-15);
            n->temperature -= fmin(q / mc, (n->temperature - amb) * 0.3);
        }
    }
    for (int i = 0; i < t->ns; ++i) {   /* updateWearKernel (:573-588), A 1e-9, B 1000 (:1291-1292) */
        omt_node *n = &t->surf[i];
        if (!n->in_contact) continue;
        const double v = sqrt(n->vx * n->vx + n->vy * n->vy + n->vz * n->vz);
        double TK = n->temperature + 273.15;
        if (TK < 300) TK = 300;
        n->wear += 1e-9 * n->von_mises * v * exp(-1000.0 / TK) * dt;
        if (n->wear > 0.3e-3) n->status = OMT_WORN;
    }
    t->time += dt;
    t->step++;
}

int omt_counts(const omt_tool *t, int out[4]) {
    out[0] = t->ns; out[1] = t->nt; out[2] = t->ne; out[3] = t->nemb;
    return 0;
}

void omt_get_surface(const omt_tool *t, omt_node *out) { memcpy(out, t->surf, sizeof(omt_node) * t->ns); }
void omt_get_tets(const omt_tool *t, omt_node *out) { memcpy(out, t->tet, sizeof(omt_node) * t->nt); }
void omt_set_surface(omt_tool *t, const omt_node *in) { memcpy(t->surf, in, sizeof(omt_node) * t->ns); }
void omt_set_tets(omt_tool *t, const omt_node *in) { memcpy(t->tet, in, sizeof(omt_node) * t->nt); }

void omt_get_elements(const omt_tool *t, int32_t *nodes4, double *volume, double *inv_dm9) {
    for (int e = 0; e < t->ne; ++e) {
        if (nodes4) memcpy(nodes4 + 4 * e, t->el[e].n, sizeof(int32_t) * 4);
        if (volume) volume[e] = t->el[e].volume;
        if (inv_dm9) memcpy(inv_dm9 + 9 * e, t->el[e].inv_dm, sizeof(double) * 9);
    }
}

/* ------------------------------------------------------------ contact ---- */

struct omc_contact {
    omc_params prm;
    int32_t table;                 /* hash table size (0 until the first build) */
    int32_t cap;                   /* node capacity of the hash (ensureCapacity) */
    int32_t *start, *end;          /* per table slot */
    int32_t *order;                /* sorted position -> surface node index */
    int32_t *hash;                 /* per surface node */
    int nn;
    int contacts;
    double heat, force, cumulative, scale;
    int64_t calls, last_contact_call, age;
};

void omc_params_default(omc_params *p) {   /* ContactConfig, ContactSolver.cuh:29-71 */
    memset(p, 0, sizeof(*p));
    p->contact_radius = 0.0002;
    p->contact_stiffness = 1e7;
    p->contact_damping_ratio = 0.25;
    p->friction = 0.3;
    p->heat_partition = 0.5;
    p->workpiece_shear_yield = 378.2e6;
    p->max_contact_pressure = 4.0e9;
    p->jc_A = 553.0e6; p->jc_B = 600.0e6; p->jc_n = 0.234; p->jc_C = 0.013; p->jc_m = 1.0;
    p->jc_rate_ref = 1.0; p->jc_T_ref = 25.0; p->jc_T_melt = 1520.0;
    p->jc_enabled = 1;
    p->engagement_ramp_steps = 200;
    p->max_penetration_fraction = 0.5;
    p->tool_specific_heat = 200.0;
    p->work_specific_heat = 475.0;
    p->work_physical_density = 7850.0;
    p->tool_melt_temp = 2870.0;
    p->work_melt_temp = 1460.0;
    p->contact_htc = 1.0e5;
    p->emissivity = 0.7;
    p->stefan_boltzmann = 5.670374419e-8;
    p->ambient_temp = 25.0;
    p->tool_physical_density = 14500.0;
    p->latent_heat_fusion = 0.0;
    p->solidus_temp = 0.0;
    p->liquidus_temp = 0.0;
}

omc_contact *omc_create(const omc_params *p) {
    omc_contact *c = (omc_contact *)calloc(1, sizeof(omc_contact));
    c->prm = *p;
    return c;
}

void omc_destroy(omc_contact *c) {
    if (!c) return;
    free(c->start); free(c->end); free(c->order); free(c->hash);
    free(c);
}

int32_t omc_next_prime(int32_t n) {   /* SharedSpatialHash::nextPrime (:117-128) */
    for (;; ++n) {
        if (n < 2) continue;
        if (n == 2) return n;
        if (n % 2 == 0) continue;
        int prime = 1;
        for (int i = 3; i * i <= n; i += 2)
            if (n % i == 0) { prime = 0; break; }
        if (prime) return n;
    }
}

static uint64_t cell_mix(int cx, int cy, int cz) {
    return ((uint64_t)cx * 73856093ULL) ^ ((uint64_t)cy * 19349663ULL) ^ ((uint64_t)cz * 83492791ULL);
}

int32_t omc_node_hash(double x, double y, double z, const double o[3], double cell, int32_t table) {
    /* computeNodeHashKernel (SharedSpatialHash.cu:9-31) */
    int cx = (int)floor((x - o[0]) / cell), cy = (int)floor((y - o[1]) / cell), cz = (int)floor((z - o[2]) / cell);
    if (cx < 0) cx = 0;
    if (cy < 0) cy = 0;
    if (cz < 0) cz = 0;
    const uint64_t h = cell_mix(cx, cy, cz), T = (uint64_t)table;
    return (int32_t)((h % T + T) % T);
}

/* hashToolCell (ContactSolver.cu:106-117) */
static int32_t probe_hash(int cx, int cy, int cz, int32_t table) {
    return (int32_t)(cell_mix(cx > 0 ? cx : 0, cy > 0 ? cy : 0, cz > 0 ? cz : 0) % (uint64_t)table);
}

/* deviceJCFlowStress (ContactSolver.cu:143-157) */
static double jc_flow(const omc_params *P, double eps, double rate, double T) {
    const double st = P->jc_A + P->jc_B * pow(fmax(eps, 0.0), P->jc_n);
    const double rr = fmax(rate / fmax(P->jc_rate_ref, 1e-9), 1.0);
    const double rt = 1.0 + P->jc_C * log(rr);
    const double ts = T > P->jc_T_ref ? fmin(1.0, (T - P->jc_T_ref) / fmax(P->jc_T_melt - P->jc_T_ref, 1.0)) : 0.0;
    const double th = fmax(0.01, 1.0 - pow(ts, P->jc_m));
    return fmax(st * rt * th, 1.0);
}

static void reset_accumulators(omt_tool *fem) {   /* resetContactAccumulatorsKernel (:636-648) */
    for (int i = 0; i < fem->ns; ++i) {
        omt_node *n = &fem->surf[i];
        n->fx = n->fy = n->fz = 0.0;
        n->heat = 0.0;
        n->fn = 0.0;
        n->pen = 0.0;
        n->in_contact = 0;
    }
}

static int cmp_i64(const void *a, const void *b) {
    const int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

/* SharedSpatialHash::build (:130-165): hashes, a sort by hash (ties in node
 * order) and the cell bounds (findToolCellBoundsKernel, ContactSolver.cu:
 * 123-137) */
static void build_hash(omc_contact *c, const omt_tool *fem, const double origin[3], double cell) {
    const int n = fem->ns;
    if (n > c->cap) {   /* ensureCapacity (:103-115) */
        free(c->order); free(c->hash);
        c->order = (int32_t *)malloc(sizeof(int32_t) * n);
        c->hash = (int32_t *)malloc(sizeof(int32_t) * n);
        const int32_t ts = omc_next_prime(n * 2 > 100003 ? n * 2 : 100003);
        if (ts != c->table) {
            free(c->start); free(c->end);
            c->start = (int32_t *)malloc(sizeof(int32_t) * ts);
            c->end = (int32_t *)malloc(sizeof(int32_t) * ts);
            c->table = ts;
        }
        c->cap = n;
    }
    c->nn = n;
    int64_t *kv = (int64_t *)malloc(sizeof(int64_t) * (n > 0 ? n : 1));
    for (int i = 0; i < n; ++i) {
        c->hash[i] = omc_node_hash(fem->surf[i].x, fem->surf[i].y, fem->surf[i].z, origin, cell, c->table);
        kv[i] = ((int64_t)c->hash[i] << 32) | (int64_t)i;
    }
    /* (hash, index) keys are distinct: the sort orders ties by node index */
    qsort(kv, (size_t)n, sizeof(int64_t), cmp_i64);
    for (int i = 0; i < c->table; ++i) c->start[i] = c->end[i] = -1;
    for (int i = 0; i < n; ++i) {
        c->order[i] = (int32_t)(kv[i] & 0xffffffff);
        const int32_t h = (int32_t)(kv[i] >> 32);
        if (i == 0 || (int32_t)(kv[i - 1] >> 32) != h) c->start[h] = i;
        if (i == n - 1 || (int32_t)(kv[i + 1] >> 32) != h) c->end[h] = i + 1;
    }
    free(kv);
}

/* applyNodeHeatAndHertzStressKernel (:665-791) with EMA 0.9 (:1212-1227) */
static void node_heat_hertz(const omc_params *P, omt_tool *fem, double dt) {
    const double ema = 0.9;
    for (int i = 0; i < fem->ns; ++i) {
        omt_node *n = &fem->surf[i];
        if (n->temperature > P->ambient_temp + 1.0) {
            const double vol = fmax(n->mass / P->tool_physical_density, 1e-15);
            const double area = pow(vol, 2.0 / 3.0);
            const double Tn = n->temperature + 273.15, Ta = P->ambient_temp + 273.15;
            const double q = P->emissivity * P->stefan_boltzmann * area * (Tn * Tn * Tn * Tn - Ta * Ta * Ta * Ta) * dt;
            n->heat = fmax(n->heat - q, 0.0);
        }
        const int phase = P->latent_heat_fusion > 0.0;
        const int mushy = phase && P->solidus_temp > 0.0 && P->liquidus_temp > P->solidus_temp;
        const double Ts = mushy ? P->solidus_temp : P->tool_melt_temp;
        const double Tl = mushy ? P->liquidus_temp : P->tool_melt_temp;
        if (n->heat > 0.0 && phase) {
            const double m = fmax(n->mass, 1e-12), T0 = n->temperature, cp = fmax(P->tool_specific_heat, 1e-12);
            const double L = P->latent_heat_fusion;
            double H0;
            if (T0 < Ts) H0 = cp * T0;
            else if (T0 < Tl) H0 = cp * Ts + L * ((T0 - Ts) / (Tl - Ts));
            else H0 = cp * Ts + L + cp * (T0 - Tl);
            const double H1 = H0 + n->heat / m;
            const double Hs = cp * Ts, Hl = cp * Ts + L;
            double Tnew;
            if (H1 <= Hs) Tnew = H1 / cp;
            else if (H1 < Hl) Tnew = Ts + ((H1 - Hs) / (Hl - Hs)) * (Tl - Ts);
            else Tnew = Tl + (H1 - Hl) / P->tool_specific_heat;
            const double dT = Tnew - T0;
            n->temperature = fmin(T0 + dT, P->tool_melt_temp);
            if (n->temperature >= P->tool_melt_temp) n->status = OMT_FAILED;
        } else if (n->heat > 0.0) {
            const double dT = n->heat / (fmax(n->mass, 1e-12) * fmax(P->tool_specific_heat, 1e-12));
            n->temperature = fmin(n->temperature + dT, P->tool_melt_temp);
            if (n->temperature >= P->tool_melt_temp) n->status = OMT_FAILED;
        }
        n->heat = 0.0;
        double sh = 0.0;
        if (n->fn > 0.0 && n->pen > 1e-12) {
            const double a2 = P->contact_radius * n->pen;
            if (a2 > 1e-18) {
                sh = (3.0 * n->fn) / (2.0 * kPi * a2);
                sh = fmin(sh, P->max_contact_pressure);
            }
        }
        n->von_mises = ema * n->von_mises + (1.0 - ema) * sh;
        n->fn = 0.0;
        n->pen = 0.0;
    }
}

/* xpbdContactSpatialHashKernel (:159-408) for one particle; the node
 * accumulators go straight to the surface nodes (the reference's sorted copy
 * plus scatterContactResultsKernel, SharedSpatialHash.cu:33-55, add the same
 * terms) */
static void contact_particle(omc_contact *c, oms_particle *p, omt_tool *fem, const double o[3], double dt,
                             double max_pen, double scale, double *heat_gen, double *force_sum) {
    const omc_params *P = &c->prm;
    if (p->status == OMS_INACTIVE || p->status == OMS_FIXED) return;
    const double cell = P->contact_radius, r = P->contact_radius;
    const int pcx = (int)floor((p->x - o[0]) / cell), pcy = (int)floor((p->y - o[1]) / cell),
              pcz = (int)floor((p->z - o[2]) / cell);
    double corr[3] = {0, 0, 0}, heat_to_p = 0.0, fsum = 0.0, hsum = 0.0;
    int count = 0, has = 0;
    const double kN = fmax(P->contact_stiffness, 0.0);
    const double cN = 2.0 * fmax(P->contact_damping_ratio, 0.0) * sqrt(kN * fmax(p->mass, 1e-15));
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                const int32_t h = probe_hash(pcx + dx, pcy + dy, pcz + dz, c->table);
                const int s0 = c->start[h];
                if (s0 < 0) continue;
                const int s1 = c->end[h];
                for (int j = s0; j < s1; ++j) {
                    omt_node *nd = &fem->surf[c->order[j]];
                    if (nd->status == OMT_FAILED) continue;
                    const double ndx = p->x - nd->x, ndy = p->y - nd->y, ndz = p->z - nd->z;
                    const double dist = sqrt(ndx * ndx + ndy * ndy + ndz * ndz);
                    if (dist >= r || dist < 1e-12 || isnan(dist)) continue;
                    has = 1;
                    double pen = r - dist;
                    if (pen > max_pen) pen = max_pen;
                    const double id = 1.0 / dist;
                    const double nx = ndx * id, ny = ndy * id, nz = ndz * id;
                    corr[0] += pen * nx; corr[1] += pen * ny; corr[2] += pen * nz;
                    count++;
                    const double rvx = p->vx - nd->vx, rvy = p->vy - nd->vy, rvz = p->vz - nd->vz;
                    const double vn = rvx * nx + rvy * ny + rvz * nz;
                    const double appr = fmax(0.0, -vn);
                    const double fn = (kN * pen + cN * appr) * scale;
                    nd->fx += -fn * nx; nd->fy += -fn * ny; nd->fz += -fn * nz;
                    nd->fn += fn;
                    if (pen > nd->pen) nd->pen = pen;
                    const double vtx = rvx - vn * nx, vty = rvy - vn * ny, vtz = rvz - vn * nz;
                    const double vt = sqrt(vtx * vtx + vty * vty + vtz * vtz);
                    const double Tavg = 0.5 * (p->temperature + nd->temperature);
                    const double Tr = fmin(1.0, fmax(0.0, (Tavg - 200.0) / 600.0));
                    const double mu_eff = P->friction * (1.0 - 0.4 * Tr);
                    const double pv = p->volume > 1e-24 ? p->volume : p->mass / fmax(p->density, 1.0);
                    const double prad = pow(3.0 * pv / (4.0 * kPi), 1.0 / 3.0);
                    const double flow = P->jc_enabled
                                            ? jc_flow(P, p->plastic_strain, p->strain_rate, p->temperature)
                                            : P->workpiece_shear_yield;
                    const double shear_yield = flow / 1.7320508;
                    const double cpp = pen > 0.0 ? pen : 0.0;
                    const double ea = kPi * (2.0 * prad * cpp - cpp * cpp);
                    const double pa = flow > 1.0 ? fn / flow : 0.0;
                    double area = ea + pa;
                    const double mina = 0.01 * pow(fmax(pv, 1e-24), 2.0 / 3.0);
                    if (area < mina) area = mina;
                    const double sn = fn / fmax(area, 1e-24);
                    const double scrit = shear_yield / fmax(mu_eff, 1e-6);
                    const double stick = 0.5 * (1.0 + tanh(10.0 * (sn / fmax(scrit, 1e-12) - 1.0)));
                    const double tau = (1.0 - stick) * (mu_eff * sn) + stick * shear_yield;
                    double ft = tau * area;
                    const double fst = vt > 1e-10 ? p->mass * vt / dt : ft;
                    ft = fmin(ft, fst);
                    const double htc = P->contact_htc * (1.0 + 5.0 * (1.0 - exp(-sn / flow)));
                    const double hs = ft * vt * dt;
                    const double qc = htc * area * (p->temperature - nd->temperature) * dt;
                    heat_to_p += (hs * (1.0 - P->heat_partition) - qc) * scale;
                    nd->heat += (hs * P->heat_partition + qc) * scale;
                    if (vt > 1e-15) {
                        nd->fx += ft * (vtx / vt); nd->fy += ft * (vty / vt); nd->fz += ft * (vtz / vt);
                    }
                    nd->in_contact = 1;
                    hsum += hs * scale;
                    fsum += fn;
                }
            }
    if (!has) return;
    const double w = 1.0 / count;
    const double cx = corr[0] * w, cy = corr[1] * w, cz = corr[2] * w;
    p->x += cx; p->y += cy; p->z += cz;
    const double cm = sqrt(cx * cx + cy * cy + cz * cz);
    if (cm > 1e-15) {
        const double ic = 1.0 / cm;
        const double nx = cx * ic, ny = cy * ic, nz = cz * ic;
        const double vd = p->vx * nx + p->vy * ny + p->vz * nz;
        if (vd < 0.0) { p->vx -= vd * nx; p->vy -= vd * ny; p->vz -= vd * nz; }
    }
    if (heat_to_p > 0.0) {
        const double pv = p->volume > 1e-24 ? p->volume : p->mass / fmax(p->density, 1.0);
        const double pm = fmax(P->work_physical_density, 1.0) * fmax(pv, 1e-24);
        const double dT = heat_to_p / (fmax(pm, 1e-15) * P->work_specific_heat);
        p->temperature = fmin(p->temperature + dT, P->work_melt_temp);
    }
    c->contacts++;
    *heat_gen += hsum;
    *force_sum += fsum;
}

void omc_resolve(omc_contact *c, oms_solver *sph, omt_tool *fem, double dt) {   /* :1085-1300 */
    int np = 0;
    oms_particle *p = oms_particle_data(sph, &np);
    const int nn = fem->ns;
    if (dt <= 0.0 || np <= 0 || nn <= 0) return;
    const omc_params *P = &c->prm;
    double scale = 1.0;
    if (c->age < P->engagement_ramp_steps && P->engagement_ramp_steps > 0) {   /* :1104-1112 */
        const double t = (double)c->age / (double)P->engagement_ramp_steps;
        scale = 0.1 + 0.9 * (t * t * (3.0 - 2.0 * t));
    }
    const double max_pen = P->contact_radius * P->max_penetration_fraction;
    double o[3];
    oms_domain_min(sph, o);
    reset_accumulators(fem);
    build_hash(c, fem, o, P->contact_radius);
    c->contacts = 0;
    double heat = 0.0, force = 0.0;
    for (int i = 0; i < np; ++i) contact_particle(c, &p[i], fem, o, dt, max_pen, scale, &heat, &force);
    node_heat_hertz(P, fem, dt);
    c->heat = heat;
    c->force = force;
    c->cumulative += heat;   /* transferResults (:1369-1376) */
    c->scale = scale;
    c->calls++;   /* :1260-1272 */
    if (c->contacts > 0) {
        c->age++;
        c->last_contact_call = c->calls;
    } else if (c->calls - c->last_contact_call > 50) {
        c->age = 0;
    }
    if (c->contacts > nn * 50) {   /* :1276-1296 */
        reset_accumulators(fem);
        c->contacts = 0;
        c->force = 0.0;
        c->heat = 0.0;
    }
}

void omc_results(const omc_contact *c, double out[5]) {
    out[0] = c->contacts; out[1] = c->heat; out[2] = c->force; out[3] = c->cumulative; out[4] = c->scale;
}

int omc_hash_size(const omc_contact *c) { return c->table; }

void omc_hash(const omc_contact *c, int32_t *start, int32_t *end, int32_t *order, int nn) {
    if (start) memcpy(start, c->start, sizeof(int32_t) * c->table);
    if (end) memcpy(end, c->end, sizeof(int32_t) * c->table);
    if (order) memcpy(order, c->order, sizeof(int32_t) * (nn < c->nn ? nn : c->nn));
}
