/*
 * mpmc.h -- C ABI of the MI355X tool FEM step and of the spatial-hash XPBD
 * contact between the F-tracking MPM particles (mpms.h) and the tool's
 * surface nodes (in libmpmw_hip.so).
 *
 * Drop-in for the device side of the reference's FEMSolver (its dynamic
 * "anchored" mode: corotational linear tets driven through spindle springs,
 * src/FEMSolver.cu:167-342,423-588,1112-1300, include/FEMSolver.cuh) and of
 * ContactSolver's spatial-hash path (src/ContactSolver.cu:106-408,636-791,
 * 1085-1300; src/SharedSpatialHash.cu:9-177) of
 * Up25-pathan/edegpredict-engine-rotary.  Plain C: opaque handles, host
 * pointers and sizes, int status codes (MPMW_OK and the MPMW_E_* values of
 * mpmw.h; mpmw_last_error() holds the text).  Not thread-safe per handle.
 *
 * Precision: the tool is FP64 in every mode (a 50 k-tet tool is a few MB);
 * the contact computes in FP64 and reads and writes the particles in the
 * F-tracking solver's precision.
 *
 * Streams: the tool steps on its own stream; mpmc_resolve and
 * mpmc_link_tool_nodes order themselves after the tool's pending work and
 * the tool's later work after them (events), and run on the F-tracking
 * solver's stream.
 */
#ifndef MPMC_H
#define MPMC_H

#include <stddef.h>
#include <stdint.h>

#include "mpms.h"
#include "mpmw.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { MPMT_OK = 0, MPMT_WORN = 1, MPMT_FAILED = 2 };   /* NodeStatus, Types.h:58-63 */
enum { MPMT_REGION_SHANK = 8 };                         /* ToolRegion::SHANK: held at ambient */

/* FEMMaterialProps (FEMSolver.cuh:41-48), FEMParams (Config.h:162-184) and
 * the FEMSolver members the step reads (FEMSolver.cuh:205-271). */
typedef struct mpmt_params {
    double young, poisson, density, thermal_conductivity, specific_heat;
    double damping_ratio, mass_scaling;
    double coolant_htc, spindle_stiffness, spindle_damping;
    int32_t spindle_enabled;    /* dynamic mode (the step runs the mechanics) */
    int32_t thermal_skip;       /* conduction every thermal_skip steps */
    /* extension, 0 = the reference: the element polar iteration adds X^-T
     * (Higham's iteration, fixed point the rotation of F) instead of the
     * reference's X^-1 (FEMSolver.cu:268-273; equal for symmetric X) */
    int32_t polar_textbook;
    int32_t device;
} mpmt_params;

/* Tool mesh as FEMSolver::initializeFromMesh takes it (FEMSolver.cu:900-
 * 1010): surface nodes and volumetric (tet) nodes, xyz interleaved, a
 * ToolRegion per node (NULL: 0), tets as 4 tet-node indices, and embed
 * constraints tying each surface node to 4 tet nodes with weights. */
typedef struct mpmt_mesh {
    int32_t n_surface, n_tet_nodes, n_elements, n_embeds;
    const double *surface_xyz;
    const int32_t *surface_region;
    const double *tet_xyz;
    const int32_t *tet_region;
    const int32_t *elements;      /* 4 per element */
    const int32_t *embed_surface; /* 1 per embed */
    const int32_t *embed_tets;    /* 4 per embed (-1: none) */
    const double *embed_weights;  /* 4 per embed */
} mpmt_mesh;

/* FEMNodeGPU (Types.h:316-380) in SoA, host arrays (NULL skips): vectors xyz
 * interleaved, stress xx yy zz xy xz yz. */
typedef struct mpmt_node_arrays {
    double *pos, *vel, *force, *mass, *temperature, *von_mises, *wear, *stress;
    int32_t *status, *in_contact;
} mpmt_node_arrays;

typedef struct mpmt_tool mpmt_tool;

void mpmt_params_default(mpmt_params *p);
/* FEMSolver::initialize (:846-867) */
int mpmt_create(const mpmt_params *p, mpmt_tool **out);
int mpmt_destroy(mpmt_tool *t);
/* initializeFromMesh + createElementsFromMesh (:900-1078): orientation,
 * volumes, inverse reference shapes, lumped masses, surface masses from the
 * embeds, tets within 5 mm of the top fixed */
int mpmt_set_mesh(mpmt_tool *t, const mpmt_mesh *m);
/* initializeDrivenNodes (:1367-1408) */
int mpmt_init_driven_nodes(mpmt_tool *t, double top_fraction);
/* setVirtualSpindleState (:1410-1413) */
int mpmt_set_spindle_state(mpmt_tool *t, const double pos[3], const double vel[3]);
/* setToolTransform (:1473-1478): rotation of the nodes about axis through
 * pos at omega rad/s during each step */
int mpmt_set_transform(mpmt_tool *t, const double pos[3], const double axis[3], double omega);
/* setRigidBodyNodeVelocities (:1491-1510) */
int mpmt_set_rigid_velocity(mpmt_tool *t, const double v[3], double omega_z, double cx, double cy);
/* applyContactForce (:1524-1536) */
int mpmt_apply_contact_force(mpmt_tool *t, int32_t node, const double f[3], double heat);
/* FEMSolver::step (:1112-1300) */
int mpmt_step(mpmt_tool *t, double dt);
int mpmt_synchronize(mpmt_tool *t);
/* counts: surface nodes, tet nodes, elements, embeds */
int mpmt_counts(mpmt_tool *t, int32_t out[4]);
/* the surface nodes (d_nodes) or the tet nodes (d_tetNodes) */
int mpmt_get_surface(mpmt_tool *t, mpmt_node_arrays *a);
int mpmt_get_tet_nodes(mpmt_tool *t, mpmt_node_arrays *a);
/* surface temperatures and forces, for tests and couplers (NULL skips) */
int mpmt_set_surface(mpmt_tool *t, const double *temperature, const double *force);

/* ContactConfig (ContactSolver.cuh:29-71) */
typedef struct mpmc_params {
    double contact_radius, contact_stiffness, contact_damping_ratio, friction, heat_partition;
    double workpiece_shear_yield, max_contact_pressure;
    double jc_A, jc_B, jc_n, jc_C, jc_m, jc_rate_ref, jc_T_ref, jc_T_melt;
    double max_penetration_fraction;
    double tool_specific_heat, work_specific_heat, work_physical_density, tool_melt_temp, work_melt_temp;
    double contact_htc, emissivity, stefan_boltzmann, ambient_temp, tool_physical_density;
    double latent_heat_fusion, solidus_temp, liquidus_temp;
    int32_t jc_enabled, engagement_ramp_steps;
    int32_t device, pad_;
} mpmc_params;

typedef struct mpmc_contact mpmc_contact;

/* Last call's results (ContactSolver getters, ContactSolver.cuh:114-118). */
typedef struct mpmc_results {
    int64_t contacts;
    double heat, force, cumulative_heat, engagement;
    int64_t probed;         /* particles inside the tool's box grown by the radius: they probed the hash */
    int32_t table_size;
    int32_t pad_;
} mpmc_results;

void mpmc_params_default(mpmc_params *p);
/* ContactSolver::initialize (:994-1011) */
int mpmc_create(const mpmc_params *p, mpmc_contact **out);
int mpmc_destroy(mpmc_contact *c);
int mpmc_set_params(mpmc_contact *c, const mpmc_params *p);
/* resolveContacts, the spatial-hash path (:1085-1300): hash the tool's
 * surface nodes (cell = contact radius, origin = the MPM lattice's
 * domain_min), then per particle the 27 neighbouring cells: XPBD position
 * correction, inward-velocity removal, Zorev friction, contact heat, node
 * reaction forces, Hertz pressure and heat on the nodes.  Reads back the
 * totals (the engagement ramp and the contact-explosion guard need them). */
int mpmc_resolve(mpmc_contact *c, mpms_solver *s, mpmt_tool *t, double dt);
int mpmc_get_results(mpmc_contact *c, mpmc_results *r);
/* MPMSolver::setToolNodes(fem->getDeviceNodes(), ...) as the engine calls it
 * before the MPM step (SimulationEngine.cpp:513): the F-tracking solver's
 * tool nodes become the tool's surface node positions and velocities
 * (device-side copy). */
int mpmc_link_tool_nodes(mpms_solver *s, mpmt_tool *t);

#ifdef __cplusplus
}
#endif
#endif
