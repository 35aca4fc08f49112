/*
 * mpm_oracle.h -- CPU FP64 restatement of the reference MPM workpiece step.
 *
 * TEST INFRASTRUCTURE ONLY.  This oracle is the checker for the HIP solver:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product library never links or calls it.
 *
 * It restates src/MPMWorkpieceSolver.cu (Up25-pathan/edegpredict-engine-rotary)
 * in plain C99, array-of-structs, FP64, serial by default (OpenMP optional for
 * the CPU baseline).  Each function in mpm_oracle.c cites the reference
 * file:line it follows.
 */
#ifndef MPM_ORACLE_H
#define MPM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ParticleStatus values, include/Types.h:37-43 */
enum { ORC_ACTIVE = 0, ORC_INACTIVE = 1, ORC_CHIP = 2, ORC_BOUNDARY = 3, ORC_FIXED_BOUNDARY = 4 };

/* Tool kinds (none / reference point samples / analytic wedge / analytic drill) */
enum { ORC_TOOL_NONE = 0, ORC_TOOL_POINTS = 1, ORC_TOOL_WEDGE = 2, ORC_TOOL_DRILL = 3, ORC_TOOL_VOXEL = 4 };

/* MPMParticleGPU, include/MPMWorkpieceSolver.cuh:18-45 (field order kept) */
typedef struct {
    double x, y, z;
    double vx, vy, vz;
    double fx, fy, fz;
    double mass, volume, density, temperature;
    double plasticStrain, strainRate, damage, residualStress;
    double s[6];   /* stress xx yy zz xy xz yz */
    double a[6];   /* backstress xx yy zz xy xz yz */
    int32_t id;
    int32_t status;
    double C[9];   /* APIC affine matrix, row-major (MPMSolver.cuh MPMParticle::C) */
} orc_particle;

/* Particle-grid transfer: the workpiece solver's PIC with trilinear weights,
 * or the APIC variant with quadratic B-splines (MPMSolver.cu:20-34,370-421,
 * 554-592).  APIC_TEXTBOOK: the same APIC with the quadratic B-spline's own
 * C = 4/h^2 * B and no 0.99 decay (no counterpart in the reference). */
enum { ORC_TRANSFER_PIC = 0, ORC_TRANSFER_APIC = 1, ORC_TRANSFER_APIC_TEXTBOOK = 2 };

/* MPMGridNodeGPU, include/MPMWorkpieceSolver.cuh:47-57 */
typedef struct {
    double mass, vx, vy, vz, fx, fy, fz, tm;
    int32_t fixed;
} orc_node;

/* MPMGridParams, include/MPMWorkpieceSolver.cuh:59-63 */
typedef struct {
    double cell;
    double minx, miny, minz;
    int32_t nx, ny, nz;
} orc_grid;

/* Solver parameters: MPMMaterialParams (MPMWorkpieceSolver.cuh:65-90) after
 * MPMWorkpieceSolver::initialize (MPMWorkpieceSolver.cu:422-458), plus the
 * Config fields the step reads through its stored Config pointer. */
typedef struct {
    double density, young, poisson, specific_heat, melt_temp, ref_temp;
    double jc_A, jc_B, jc_n, jc_C, jc_m, ref_strain_rate;
    double failure_strain, damage_threshold, critical_plastic_strain, thermal_chip_ratio;
    double tool_contact_radius, contact_damping, max_contact_temp_rise;
    double thermal_conductivity, kinematic_hardening;
    int32_t thermal_clamp_enabled;
    int32_t has_config;            /* m_config != nullptr */
    double cell_size;              /* smoothingRadius */
    double particle_spacing_factor;
    double fixture_thickness;      /* MachineSetupParams::fixtureLayerThickness */
    double min_time_step, max_time_step;
    int32_t max_particles;
    int32_t pad_;
} orc_params;

/* Analytic tool description (grid-node SDF collision, extension of
 * ContactSolver.cu:797-923 semantics to grid nodes). */
typedef struct {
    int32_t kind;
    int32_t pad_;
    double band;             /* contact when phi < band */
    double friction;         /* Coulomb coefficient */
    double grad_eps;         /* central-difference step for the normal */
    /* wedge */
    double rake_angle, clearance_angle;
    /* drill */
    double radius, helix_angle, point_angle, flute_offset, flute_radius, flute_length;
} orc_tool_desc;

typedef struct {
    double center[3];
    double angle;            /* rotation about +z (rad) */
    double vel[3];           /* linear velocity of the tool frame (m/s) */
    double omega;            /* spin about the +z axis through center (rad/s) */
} orc_tool_pose;

typedef struct {
    int32_t n;
    int32_t pad_;
    double *pos, *vel, *fext;               /* 3n each; fext may be NULL */
    double *mass, *volume, *density, *temperature;
    double *plastic_strain, *strain_rate, *damage, *residual_stress;
    double *stress, *backstress;            /* 6n each; backstress may be NULL */
    int32_t *id, *status;
} orc_particle_arrays;

typedef struct {
    double max_stress, max_temperature, kinetic_energy, contact_force;
    double time;
    int64_t step;
} orc_metrics;

typedef struct orc_solver orc_solver;

void orc_params_default(orc_params *p);
orc_solver *orc_create(const orc_params *p);
void orc_destroy(orc_solver *s);
void orc_set_threads(int n);

int orc_set_particles(orc_solver *s, const orc_particle_arrays *a);
int orc_get_particles(orc_solver *s, orc_particle_arrays *a);
int orc_particle_count(const orc_solver *s);

void orc_init_particle_box(orc_solver *s, const double mn[3], const double mx[3], double spacing);
void orc_init_cylinder(orc_solver *s, const double c[3], double radius, double length,
                       double spacing, int axis);

void orc_set_tool_samples(orc_solver *s, const double *xyz, int n, double contact_radius);
void orc_set_tool_analytic(orc_solver *s, const orc_tool_desc *d);
void orc_set_tool_pose(orc_solver *s, const orc_tool_pose *p);
void orc_set_tool_tip(orc_solver *s, double x, double y, double z);
void orc_apply_external_force(orc_solver *s, const double c[3], double radius, const double f[3]);

/* Select the transfer; the affine matrices start at zero. */
void orc_set_transfer(orc_solver *s, int mode);
/* Affine matrices of all particles, 9 per particle (row-major). */
int orc_get_affine(const orc_solver *s, double *C);
void orc_set_affine(orc_solver *s, const double *C);

void orc_step(orc_solver *s, double dt);
void orc_run(orc_solver *s, int nsteps, double dt);
double orc_stable_dt(const orc_solver *s);
void orc_get_metrics(const orc_solver *s, orc_metrics *m);
void orc_sync_metrics(orc_solver *s);
void orc_compact(orc_solver *s);
/* Replace the particle state, keeping the grid, step count and time (a
 * shadow oracle that continues from another solver's state).  No
 * counterpart in the reference. */
int orc_replace_particles(orc_solver *s, const orc_particle_arrays *a);
/* Round every stored particle field except the positions to the nearest
 * FP32 value (a sensitivity probe: the state an FP32 solver keeps between
 * steps, with FP64 positions).  No counterpart in the reference. */
void orc_round_state_fp32(orc_solver *s);
/* Sensitivity probe (no counterpart in the reference; off by default): the
 * PIC P2G rounds every node sum to FP32 after each add and the step ends
 * with orc_round_state_fp32, i.e. FP64 arithmetic with the rounding of FP32
 * accumulators and FP32 stored state.  The difference it makes to a run is
 * the run's sensitivity to FP32-sized errors. */
void orc_set_fp32_emulation(orc_solver *s, int on);
void orc_reset(orc_solver *s);
void orc_get_bounds(const orc_solver *s, double b[6]);
double orc_temperature_at(const orc_solver *s, double x, double y, double z);
void orc_apply_heat_flux(orc_solver *s, double x, double y, double z, double q);
void orc_set_particle_temperatures(orc_solver *s, const double *t, int count);

void orc_get_grid(const orc_solver *s, orc_grid *g);
/* Post-P2G grid snapshot of the last step (mass, momentum xyz, force xyz, T*m). */
int orc_get_grid_snapshot(const orc_solver *s, double *mass, double *mom, double *force, double *tm);
int orc_grid_node_count(const orc_solver *s);
/* Grid after the last step (post-update velocity, post-diffusion T*m). */
int orc_get_grid_state(const orc_solver *s, double *mass, double *vel, double *tm, int32_t *fixed);

/* Single-function entry points for known-answer tests. */
void orc_weights(const orc_grid *g, double x, double y, double z,
                 int32_t idx[8], double w[8], double grad[24]);
/* getWeightsAndGradients (MPMSolver.cu:20-34) on all three axes: the 27
 * nodes of the quadratic stencil, x outermost and z innermost as the APIC
 * loops visit them (:370-378), with weight, gradient and node offset
 * x_i - x_p; nodes outside the grid get index -1. */
void orc_weights_quadratic(const orc_grid *g, double x, double y, double z, int32_t idx[27], double w[27],
                           double grad[81], double dpos[81]);
double orc_equivalent_stress(double xx, double yy, double zz, double xy, double xz, double yz);
double orc_jc_flow(const orc_params *p, double eps_p, double eps_rate, double temp);
void orc_rebuild_params(double cell, const double mn[3], const double mx[3], orc_grid *g);
void orc_chip_kick(int32_t id, double r[3]);
double orc_tool_phi(const orc_tool_desc *d, const orc_tool_pose *p, double x, double y, double z);
void orc_tool_normal(const orc_tool_desc *d, const orc_tool_pose *p, double x, double y, double z,
                     double n[3]);
/* One G2P constitutive update of a single particle from a prescribed
 * velocity gradient (row = velocity component, column = d/dx_b). */
void orc_constitutive(const orc_params *p, orc_particle *pt, const double L[9], double dt);
/* Voxel signed distance field of a tool mesh at rest pose: ToolSDF::build
 * (ToolSDF.cu:212-266, buildFromVerts :348-378, computeSDFCPU :380-532),
 * restated in single precision as the reference computes it.  Voxel
 * (i, j, k) holds the value at origin + (i + 0.5, j + 0.5, k + 0.5) * voxel,
 * index i + nx * (j + ny * k). */
typedef struct {
    int32_t nx, ny, nz, pad_;
    float ox, oy, oz, voxel;
    float *grid;
} orc_sdf;

/* pointTriDistanceSqHost, ToolSDF.cu:92-162: squared distance from p to the
 * point the reference takes as closest on triangle abc (written to closest). */
float orc_point_tri_d2(const float p[3], const float a[3], const float b[3], const float c[3], float closest[3]);
/* Build from nverts xyz vertices and ntris triangles (3 vertex indices and a
 * normal each); returns 0 on success (grid allocated), -1 on bad input. */
int orc_sdf_build(const double *verts, int nverts, const int32_t *tris, const double *normals, int ntris,
                  double padding, int voxels_per_axis, orc_sdf *out);
void orc_sdf_free(orc_sdf *sdf);
/* Trilinear value at rest-frame point q over voxel centres, edge voxels
 * repeated outside the grid (the texture's clamp addressing, :544-553). */
double orc_sdf_sample(const orc_sdf *sdf, const double q[3]);
/* The sampler's range test (ContactSolver.cu:837-840): 1 when the texel
 * coordinate (q - origin) / voxel lies in [0, n - 1] on every axis. */
int orc_sdf_in_range(const orc_sdf *sdf, const double q[3]);
/* Use the SDF (copied) as the tool's phi, with the contact law of d
 * (band, friction, grad_eps); the pose comes from orc_set_tool_pose. */
void orc_set_tool_sdf(orc_solver *s, const orc_sdf *sdf, const orc_tool_desc *d);

/* Jittered slab generator (config inputs), float-exact positions. */
void orc_gen_jittered_slab(uint64_t seed, int cx, int cy, int cz, double h, const double origin[3],
                           double density, double fixture, double *pos, double *mass,
                           double *volume, int32_t *id, int32_t *status);

#ifdef __cplusplus
}
#endif
#endif
