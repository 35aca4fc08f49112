/*
 * mpms.h -- C ABI of the MI355X F-tracking MPM solver (in libmpmw_hip.so).
 *
 * Drop-in replacement for the device side of the reference's second MPM
 * solver, MPMSolver (include/MPMSolver.cuh:80-279, src/MPMSolver.cu of
 * Up25-pathan/edegpredict-engine-rotary): quadratic B-spline APIC transfer,
 * a deformation gradient F and a plastic one Fp per particle, the
 * compressible Neo-Hookean stress of Fe = F Fp^-1 with J clamped to
 * [0.9, 1.1], the Johnson-Cook return with kinematic hardening, JC failure
 * damage, gravity, fixture damping, tool rasterisation and LOD zones.
 * Plain C: opaque handle, host pointers and sizes, int status codes
 * (MPMW_OK and the MPMW_E_* values of mpmw.h; mpmw_last_error() holds the
 * text).  Not thread-safe per handle.
 *
 * Precision: MPMS_FP32 keeps the particle and grid fields in FP32 with the
 * positions in FP64 (F and Fp held as F - I and Fp - I, so the strain keeps
 * its own digits), and FP64 islands: the elastic kinematics of a particle
 * whose |det F| is below island_det (Fe = F Fp^-1, its determinant,
 * B = Fe Fe^T or the polar rotation), the polar decomposition of the
 * corotated law, and the Newton solve of the consistent return map.
 * MPMS_FP64 keeps everything in FP64.
 *
 * Grid: the reference's lattice (nodes at domain_min + i dx, (int)((max -
 * min) / dx) + 1 per axis, MPMSolver.cu:864-874); only a sub-box around the
 * particles is held, fitted every 25 steps with a margin of 4 nodes.  A
 * particle whose stencil leaves the sub-box between fits is counted
 * (mpms_get_diagnostics) and makes the next fit fail with MPMW_E_STATE.
 */
#ifndef MPMS_H
#define MPMS_H

#include <stddef.h>
#include <stdint.h>

#include "mpmw.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { MPMS_FP32 = 0, MPMS_FP64 = 1 };
enum { MPMS_LOD_ACTIVE = 0, MPMS_LOD_NEAR = 1, MPMS_LOD_FAR = 2 };
/* Elastic law: the reference's compressible Neo-Hookean (MPMSolver.cu:
 * 171-196), or the fixed-corotated Cauchy stress 2 mu/J (Fe - R) Fe^T +
 * lambda (J - 1) I with R the rotation of Fe's polar decomposition
 * (extension; no counterpart in the reference's MPM). */
enum { MPMS_ELASTIC_NEOHOOKEAN = 0, MPMS_ELASTIC_COROTATED = 1 };
/* Return map: the reference's radial return with the flow stress at the old
 * plastic strain (:214-276), or the consistent return q - 3 mu d =
 * Y(eps_p + d) solved by Newton (extension). */
enum { MPMS_RETURN_RADIAL = 0, MPMS_RETURN_NEWTON = 1 };

/* MPMKernelConfig (MPMSolver.cuh:27-78) as MPMSolver::initialize fills it
 * (MPMSolver.cu:858-920) and the step's host parameters. */
typedef struct mpms_params {
    double young, poisson, density, mass_scaling, specific_heat, taylor_quinney, melt_temp, ambient_temp;
    double jc_A, jc_B, jc_n, jc_C, jc_m, kinematic_hardening;
    double jc_D1, jc_D2, jc_D3, jc_D4, jc_D5, damage_threshold, ref_strain_rate;
    double lod_active_radius, lod_near_radius;
    double fixture_thickness, fixture_damping;
    double dx;
    double domain_min[3], domain_max[3];
    double cfl;
    int32_t lod_enabled;
    int32_t max_particles;
    int32_t elastic_model;     /* MPMS_ELASTIC_* */
    int32_t return_map;        /* MPMS_RETURN_* */
    int32_t precision;         /* MPMS_FP32 / MPMS_FP64 */
    int32_t device;
    double island_det;         /* FP32: |det F| below this takes the FP64 kinematics (0: never) */
    int32_t newton_fp64;       /* FP32: the Newton return in FP64 (1) or FP32 (0) */
    int32_t polar_fp64;        /* FP32: the corotated law's polar decomposition in FP64 (1) or FP32 (0) */
} mpms_params;

/* Host particle arrays (caller-owned), MPMParticle (Types.h:199-266) in SoA:
 * vectors xyz interleaved, 3x3 matrices row-major, stress and backstress
 * xx yy zz xy xz yz. */
typedef struct mpms_particle_arrays {
    int32_t n, pad_;
    double *pos, *vel, *fext;
    double *mass, *volume, *temperature;
    double *F, *Fp, *C;
    double *stress, *backstress;
    double *plastic_strain, *strain_rate, *damage, *residual_stress;
    int32_t *id, *status, *lod;
} mpms_particle_arrays;

typedef struct mpms_solver mpms_solver;

typedef struct mpms_diagnostics {
    int32_t box_lo[3], box_n[3];     /* the node sub-box */
    int64_t fits, sorts;
    int64_t overflow;               /* particle stencils outside the sub-box since the last fit */
    int64_t island_kinematics;      /* FP32: particles that took the FP64 kinematics, last step */
    int64_t island_newton;          /* FP32: particles that took the FP64 Newton return, last step */
    int64_t out_of_tile;            /* particles whose P2G went to the device grid directly (their base node
                                       left its tile, or their wave's runs share a base node), last step;
                                       above 5 % of the particles two, six or twelve steps after a fit,
                                       the step fits and re-sorts */
} mpms_diagnostics;

/* MPMSolver ctor / initialize (MPMSolver.cu:851-920); defaults as Config */
void mpms_params_default(mpms_params *p);
int mpms_create(const mpms_params *p, mpms_solver **out);
int mpms_destroy(mpms_solver *s);
/* copyToDevice / getParticles (MPMSolver.cu:958-966,1368-1371) in SoA */
int mpms_set_particles(mpms_solver *s, const mpms_particle_arrays *a);
int mpms_get_particles(mpms_solver *s, mpms_particle_arrays *a);
int mpms_particle_count(mpms_solver *s, int32_t *n);
/* initializeParticleBox (:1290-1325) / initializeCylindricalWorkpiece (:1327-1366) */
int mpms_init_particle_box(mpms_solver *s, const double mn[3], const double mx[3], double spacing);
int mpms_init_cylinder(mpms_solver *s, const double c[3], double radius, double length, double spacing);
/* setToolNodes (MPMSolver.cuh:93-96): tool node positions and velocities
 * (the reference passes its FEM tool's device nodes; here host arrays,
 * copied).  A node marked by several tool nodes takes the velocity of the
 * highest-indexed one (the reference's last writer). */
int mpms_set_tool_nodes(mpms_solver *s, const double *xyz, const double *vel, int32_t n);
/* setToolPosition (MPMSolver.cuh:157): LOD centre and chip-separation tip */
int mpms_set_tool_position(mpms_solver *s, double x, double y, double z);
/* applyExternalForce (:1440-1450, kernel :607-627) */
int mpms_apply_external_force(mpms_solver *s, const double c[3], double radius, const double f[3]);
/* MPMSolver::step (:1062-1168) */
int mpms_step(mpms_solver *s, double dt);
int mpms_run(mpms_solver *s, int32_t nsteps, double dt);
int mpms_synchronize(mpms_solver *s);
/* getStableTimeStep (:1170-1178) */
int mpms_stable_dt(mpms_solver *s, double *dt);
/* syncMetrics (:1236-1254): max von Mises, max T, KE, internal energy */
int mpms_metrics(mpms_solver *s, double out[4]);
/* getBounds: the AABB of ACTIVE and CHIP particles (updateParticleBoundsKernel :690-709) */
int mpms_get_bounds(mpms_solver *s, double b[6]);
/* The last step's grid on its sub-box (lo, n; mass, momentum after the
 * update, P2G force; NULL skips): for parity. */
int mpms_get_grid(mpms_solver *s, int32_t lo[3], int32_t n[3], double *mass, double *mom, double *force);
int mpms_get_diagnostics(mpms_solver *s, mpms_diagnostics *d);
/* HIP graph of the step (MPMSolver::step records its launches as a CUDA
 * graph and replays it, MPMSolver.cu:1077-1157): with mpms_set_graph(s, 1)
 * the launches of a step after its fit are captured once per launch
 * configuration (particle count, grid sub-box, buffer addresses; the two
 * most recent kept) and replayed; dt, the tool position and the chip step
 * reach the graph through device memory, so changing them does not capture
 * again.  Results equal the eager step's bit for bit.  mpms_graph_stats:
 * captures and replays so far. */
int mpms_set_graph(mpms_solver *s, int32_t on);
int mpms_graph_stats(mpms_solver *s, int64_t out[2]);
/* The hipStream_t the solver launches on, as an opaque pointer. */
int mpms_get_stream(mpms_solver *s, void **stream);

#ifdef __cplusplus
}
#endif
#endif
