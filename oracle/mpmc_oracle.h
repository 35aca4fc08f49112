/*
 * mpmc_oracle.h -- CPU FP64 restatement of the reference's tool FEM step
 * (FEMSolver::step, src/FEMSolver.cu:1112-1300, dynamic mode) and of its
 * spatial-hash XPBD contact between the F-tracking MPM particles and the
 * tool's surface nodes (ContactSolver::resolveContacts, hash path,
 * src/ContactSolver.cu:1085-1300; SharedSpatialHash.cu:9-177) of
 * Up25-pathan/edegpredict-engine-rotary, for parity tests only.
 *
 * TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library, as the checker.  No product path
 * links or calls it.
 *
 * Built into liborc_mpms.so with mpms_oracle.c: the contact acts on the
 * particles of an oms_solver.
 *
 * One extension with no counterpart in the reference, off by default:
 * polar_textbook = 1 makes the element's polar iteration add X^-T (Higham's
 * iteration, whose fixed point is the rotation of F); the reference adds
 * X^-1 (FEMSolver.cu:268-273), which agrees with it for symmetric X only.
 */
#ifndef MPMC_ORACLE_H
#define MPMC_ORACLE_H

#include <stdint.h>

#include "mpms_oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { OMT_OK = 0, OMT_WORN = 1, OMT_FAILED = 2, OMT_FIXED = 3 };   /* NodeStatus, Types.h:58-63 */
enum { OMT_REGION_SHANK = 8 };                                     /* ToolRegion::SHANK, Types.h:117 */

/* FEMMaterialProps (FEMSolver.cuh:41-48), FEMParams (Config.h:162-184) and
 * the FEMSolver members the step reads (FEMSolver.cuh:205-271). */
typedef struct {
    double young, poisson, density, thermal_conductivity, specific_heat;
    double damping_ratio, mass_scaling;
    double coolant_htc, spindle_stiffness, spindle_damping;
    int32_t spindle_enabled;
    int32_t thermal_skip;
    int32_t polar_textbook;        /* extension, 0 = the reference's iteration */
    int32_t pad_;
} omt_params;

/* FEMNodeGPU (Types.h:316-380), the fields the step and the contact use. */
typedef struct {
    double x, y, z, ox, oy, oz, vx, vy, vz, fx, fy, fz;
    double mass, temperature, von_mises, wear;
    double s[6], stress_weight;    /* xx yy zz xy xz yz, sum of weights */
    double heat, conduction, fn, pen;
    double loff[3];
    int32_t id, status, fixed, in_contact, driven, region;
} omt_node;

typedef struct omt_tool omt_tool;

void omt_params_default(omt_params *p);
omt_tool *omt_create(const omt_params *p);
void omt_destroy(omt_tool *t);
/* initializeFromMesh + createElementsFromMesh (FEMSolver.cu:900-1078):
 * surface nodes, tet nodes (xyz interleaved, tool region per node),
 * elements (4 tet-node indices each) and embed constraints (surface node,
 * 4 tet nodes, 4 weights). */
void omt_set_mesh(omt_tool *t, const double *surf_xyz, const int32_t *surf_region, int ns, const double *tet_xyz,
                  const int32_t *tet_region, int nt, const int32_t *elems, int ne, const int32_t *emb_surf,
                  const int32_t *emb_tet, const double *emb_w, int nemb);
/* initializeDrivenNodes (:1367-1408) */
void omt_init_driven(omt_tool *t, double top_fraction);
/* setVirtualSpindleState (:1410-1413), setToolTransform (:1473-1478) */
void omt_set_spindle_state(omt_tool *t, const double pos[3], const double vel[3]);
void omt_set_transform(omt_tool *t, const double pos[3], const double axis[3], double omega);
/* setRigidBodyNodeVelocities (:1491-1510, kernel :787-798) */
void omt_set_rigid_velocity(omt_tool *t, const double v[3], double omega_z, double cx, double cy);
/* applyContactForce (:1524-1536) */
void omt_apply_contact_force(omt_tool *t, int node, const double f[3], double heat);
/* FEMSolver::step (:1112-1300) */
void omt_step(omt_tool *t, double dt);
int omt_counts(const omt_tool *t, int out[4]);   /* surface nodes, tet nodes, elements, embeds */
void omt_get_surface(const omt_tool *t, omt_node *out);
void omt_get_tets(const omt_tool *t, omt_node *out);
void omt_set_surface(omt_tool *t, const omt_node *in);
void omt_set_tets(omt_tool *t, const omt_node *in);
/* element volume and inverse reference shape matrix (row-major), for pins */
void omt_get_elements(const omt_tool *t, int32_t *nodes4, double *volume, double *inv_dm9);
/* computeElementForcesKernel's polar iteration on one F (row-major) */
void omt_polar(const double F[9], int textbook, double R[9]);

/* ContactConfig (ContactSolver.cuh:29-71) */
typedef struct {
    double contact_radius, contact_stiffness, contact_damping_ratio, friction, heat_partition;
    double workpiece_shear_yield, max_contact_pressure;
    double jc_A, jc_B, jc_n, jc_C, jc_m, jc_rate_ref, jc_T_ref, jc_T_melt;
    double max_penetration_fraction;
    double tool_specific_heat, work_specific_heat, work_physical_density, tool_melt_temp, work_melt_temp;
    double contact_htc, emissivity, stefan_boltzmann, ambient_temp, tool_physical_density;
    double latent_heat_fusion, solidus_temp, liquidus_temp;
    int32_t jc_enabled, engagement_ramp_steps;
} omc_params;

typedef struct omc_contact omc_contact;

void omc_params_default(omc_params *p);
omc_contact *omc_create(const omc_params *p);
void omc_destroy(omc_contact *c);
/* resolveContacts, the spatial-hash path (ContactSolver.cu:1085-1300) */
void omc_resolve(omc_contact *c, oms_solver *sph, omt_tool *fem, double dt);
/* contacts, heat (J), force (N) of the last call, cumulative heat, engagement scale of the last call */
void omc_results(const omc_contact *c, double out[5]);
/* the hash of the last call: table size, cell start / end per slot (table
 * entries), the node order (original indices), for pins */
int omc_hash_size(const omc_contact *c);
void omc_hash(const omc_contact *c, int32_t *start, int32_t *end, int32_t *order, int nn);
/* the cell hash of a node position (computeNodeHashKernel, SharedSpatialHash.cu:9-31) */
int32_t omc_node_hash(double x, double y, double z, const double origin[3], double cell, int32_t table);
int32_t omc_next_prime(int32_t n);

#ifdef __cplusplus
}
#endif
#endif
