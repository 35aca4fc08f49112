/*
 * mpms_oracle.h -- CPU FP64 restatement of the reference's F-tracking MPM
 * step, MPMSolver::step (src/MPMSolver.cu:1062-1168 of
 * Up25-pathan/edegpredict-engine-rotary), for parity tests only.
 *
 * TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library, as the checker.  No product path
 * links or calls it.
 *
 * Grid: the reference's node lattice, anchored at domain_min with spacing dx
 * and (int)((max - min) / dx) + 1 nodes per axis (MPMSolver.cu:869-874).
 * Only the sub-box of nodes the particles' stencils reach is held: every
 * node outside it has no mass and is read by no particle, so a step on the
 * sub-box computes what a step on the whole lattice computes.
 *
 * Two extensions with no counterpart in the reference, off by default:
 *  - elastic model COROTATED: the fixed-corotated Cauchy stress
 *    2 mu / J (Fe - R) Fe^T + lambda (J - 1) I with R the rotation of the
 *    polar decomposition of Fe (Higham's scaled Newton iteration);
 *  - return map NEWTON: the consistent Johnson-Cook return, Newton on
 *    q - 3 mu d - Y(eps_p + d) = 0 for the plastic increment d.
 */
#ifndef MPMS_ORACLE_H
#define MPMS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OMS_ACTIVE = 0, OMS_INACTIVE = 1, OMS_CHIP = 2, OMS_BOUNDARY = 3, OMS_FIXED = 4 };
enum { OMS_LOD_ACTIVE = 0, OMS_LOD_NEAR = 1, OMS_LOD_FAR = 2 };
enum { OMS_ELASTIC_NEOHOOKEAN = 0, OMS_ELASTIC_COROTATED = 1 };
enum { OMS_RETURN_RADIAL = 0, OMS_RETURN_NEWTON = 1 };

/* MPMKernelConfig (MPMSolver.cuh:27-78) as MPMSolver::initialize fills it
 * (MPMSolver.cu:858-920), plus the step's host parameters. */
typedef struct {
    double young, poisson, density, mass_scaling, specific_heat, taylor_quinney, melt_temp, ambient_temp;
    double jc_A, jc_B, jc_n, jc_C, jc_m, kinematic_hardening;
    double jc_D1, jc_D2, jc_D3, jc_D4, jc_D5, damage_threshold, ref_strain_rate;
    double lod_active_radius, lod_near_radius;
    double fixture_thickness, fixture_damping;
    double dx;                     /* grid spacing: max(2 * smoothingRadius, 0.5 mm) in the reference */
    double domain_min[3], domain_max[3];
    double cfl;
    int32_t lod_enabled;
    int32_t max_particles;
    int32_t elastic_model;         /* OMS_ELASTIC_* */
    int32_t return_map;            /* OMS_RETURN_* */
} oms_params;

/* MPMParticle (Types.h:199-266), the fields the step reads or writes. */
typedef struct {
    double x, y, z, vx, vy, vz, efx, efy, efz;
    double mass, volume, density, temperature;
    double F[9], Fp[9], C[9];
    double s[6];                   /* stress xx yy zz xy xz yz */
    double a[6];                   /* backstress xx yy zz xy xz yz */
    double plastic_strain, strain_rate, damage, residual_stress;
    int32_t id, status, lod, pad_;
} oms_particle;

typedef struct oms_solver oms_solver;

void oms_params_default(oms_params *p);
oms_solver *oms_create(const oms_params *p);
void oms_destroy(oms_solver *s);
void oms_set_particles(oms_solver *s, const oms_particle *p, int n);
int oms_particle_count(const oms_solver *s);
void oms_get_particles(const oms_solver *s, oms_particle *p);
/* FEMNodeGPU positions and velocities of the tool (MPMSolver::setToolNodes) */
void oms_set_tool_nodes(oms_solver *s, const double *xyz, const double *vel, int n);
void oms_set_tool_position(oms_solver *s, double x, double y, double z);
void oms_apply_external_force(oms_solver *s, const double c[3], double radius, const double f[3]);
void oms_step(oms_solver *s, double dt);
double oms_stable_dt(const oms_solver *s);
/* syncMetrics (MPMSolver.cu:1236-1254): max von Mises, max T, KE, internal energy */
void oms_metrics(const oms_solver *s, double out[4]);
/* the AABB of ACTIVE and CHIP particles (updateParticleBoundsKernel) */
void oms_bounds(const oms_solver *s, double out[6]);
void oms_init_box(oms_solver *s, const double mn[3], const double mx[3], double spacing);
void oms_init_cylinder(oms_solver *s, const double c[3], double radius, double length, double spacing);
/* the post-P2G grid of the last step on its node sub-box: lo[3], n[3] and,
 * when the arrays are non-NULL, mass, momentum (xyz), force (xyz) per node */
void oms_grid(const oms_solver *s, int32_t lo[3], int32_t n[3], double *mass, double *mom, double *force);

/* The particle array and the lattice origin, for the contact restatement
 * (mpmc_oracle.c), which moves particles and changes their velocity and
 * temperature in place. */
oms_particle *oms_particle_data(oms_solver *s, int *n);
void oms_domain_min(const oms_solver *s, double out[3]);

/* Constitutive pieces, exposed for closed-form pins. */
void oms_polar(const double F[9], double R[9], int *iters);
void oms_p2g_stress(const oms_params *prm, oms_particle *p, double sigma[9], double *J, int *skip);

#ifdef __cplusplus
}
#endif
#endif
