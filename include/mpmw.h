/*
 * mpmw.h -- C ABI of the MI355X MPM workpiece solver (libmpmw_hip.so).
 *
 * Drop-in replacement for the device side of the reference's
 * MPMWorkpieceSolver (include/MPMWorkpieceSolver.cuh:92-179,
 * src/MPMWorkpieceSolver.cu:418-946 of Up25-pathan/edegpredict-engine-rotary).
 * Plain C: opaque handle, host pointers and sizes, int status codes; no HIP
 * or torch types appear here.  The C++ class edgepredict::MPMWorkpieceSolver
 * (include/edgepredict/MPMWorkpieceSolver.h) maps these onto the reference's
 * IPhysicsSolver / IMetricsProvider / IThermalCoupling signatures.
 *
 * Every entry point returns MPMW_OK (0) or a negative error code; the text of
 * the last error on the calling thread is returned by mpmw_last_error().
 * A solver handle is not thread-safe: calls on one handle must not overlap.
 */
#ifndef MPMW_H
#define MPMW_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPMW_ABI_VERSION 1

enum mpmw_status_code {
    MPMW_OK = 0,
    MPMW_E_INVAL = -1,     /* bad argument (null pointer, size, non-finite value) */
    MPMW_E_HIP = -2,       /* HIP runtime error */
    MPMW_E_STATE = -3,     /* call not valid in the solver's current state */
    MPMW_E_NOMEM = -4,     /* device allocation failed */
    MPMW_E_RANGE = -5      /* size beyond a documented limit */
};

/* ParticleStatus, include/Types.h:37-43 (same values) */
enum mpmw_particle_status {
    MPMW_ACTIVE = 0,
    MPMW_INACTIVE = 1,
    MPMW_CHIP = 2,
    MPMW_BOUNDARY = 3,
    MPMW_FIXED_BOUNDARY = 4
};

enum mpmw_precision { MPMW_FP32 = 0, MPMW_FP64 = 1 };

/* Particle-grid transfer: the workpiece solver's PIC with trilinear weights
 * (MPMWorkpieceSolver.cu:52-90), or the APIC variant with quadratic
 * B-splines (MPMSolver.cu:20-34, 370-421, 554-592).  APIC_TEXTBOOK is an
 * extension with no counterpart in the reference: the same transfer with the
 * quadratic B-spline's own C = 4/h^2 * B and no 0.99 decay, under which an
 * affine velocity field is carried exactly. */
enum mpmw_transfer { MPMW_TRANSFER_PIC = 0, MPMW_TRANSFER_APIC = 1, MPMW_TRANSFER_APIC_TEXTBOOK = 2 };

enum mpmw_tool_kind {
    MPMW_TOOL_NONE = 0,
    MPMW_TOOL_POINTS = 1,  /* reference point samples (setToolMesh) */
    MPMW_TOOL_WEDGE = 2,   /* analytic rake/flank wedge */
    MPMW_TOOL_DRILL = 3,   /* analytic two-flute helical drill */
    MPMW_TOOL_VOXEL = 4    /* voxel SDF of a tool mesh (ToolSDF) */
};

/* Solver parameters.  Material fields are MPMMaterialParams
 * (MPMWorkpieceSolver.cuh:65-90) as MPMWorkpieceSolver::initialize fills them
 * (MPMWorkpieceSolver.cu:428-451); the setup fields are the Config values the
 * solver reads through its stored Config pointer (maxParticles :591,
 * fixtureLayerThickness :604,742, min/max time step :823-825).
 * mpmw_params_default() fills the reference Config defaults. */
typedef struct mpmw_params {
    double density, young, poisson, specific_heat, melt_temp, ref_temp;
    double jc_A, jc_B, jc_n, jc_C, jc_m, ref_strain_rate;
    double failure_strain, damage_threshold, critical_plastic_strain, thermal_chip_ratio;
    double tool_contact_radius, contact_damping, max_contact_temp_rise;
    double thermal_conductivity, kinematic_hardening;
    int32_t thermal_clamp_enabled;
    int32_t has_config;            /* 0 reproduces a solver built without a Config */
    double cell_size;              /* smoothingRadius: grid cell h */
    double particle_spacing_factor;
    double fixture_thickness;
    double min_time_step, max_time_step;
    int32_t max_particles;
    int32_t precision;             /* mpmw_precision */
    int32_t deterministic;         /* 1: sorted fixed-order reductions, bit-exact reruns */
    int32_t device;                /* HIP device ordinal */
} mpmw_params;

/* Host-side particle arrays (FP64 and int32, caller-owned).  Layout of the
 * fields follows MPMParticleGPU (MPMWorkpieceSolver.cuh:18-45); vectors are
 * interleaved xyz, tensors are xx yy zz xy xz yz.  fext, backstress and
 * density may be NULL (meaning zero / zero / material density on input). */
typedef struct mpmw_particle_arrays {
    int32_t n;
    int32_t pad_;
    double *pos, *vel, *fext;
    double *mass, *volume, *density, *temperature;
    double *plastic_strain, *strain_rate, *damage, *residual_stress;
    double *stress, *backstress;
    int32_t *id, *status;
} mpmw_particle_arrays;

/* Analytic tool geometry + contact law (grid-node SDF collision; the
 * reference's particle form is ContactSolver.cu:797-923). */
typedef struct mpmw_tool_desc {
    int32_t kind;
    int32_t pad_;
    double band;               /* contact when phi < band (m) */
    double friction;           /* Coulomb coefficient */
    double grad_eps;           /* central-difference step for the normal (m) */
    double rake_angle, clearance_angle;                  /* wedge (rad) */
    double radius, helix_angle, point_angle;             /* drill (m, rad, rad) */
    double flute_offset, flute_radius, flute_length;     /* drill (m) */
} mpmw_tool_desc;

/* Tool pose; the pose advances by vel*dt and omega*dt after every step. */
typedef struct mpmw_tool_pose {
    double center[3];
    double angle;
    double vel[3];
    double omega;
} mpmw_tool_pose;

typedef struct mpmw_metrics {
    double max_stress, max_temperature, kinetic_energy, contact_force;
    double time;
    int64_t step;
} mpmw_metrics;

typedef struct mpmw_grid_info {
    double cell, minx, miny, minz;
    int32_t nx, ny, nz, pad_;
} mpmw_grid_info;

/* Per-kernel timing collected with HIP events when enabled. */
typedef struct mpmw_kernel_times {
    double ms[8];              /* accumulated milliseconds per kernel class */
    int64_t launches[8];       /* launches per kernel class */
} mpmw_kernel_times;

enum mpmw_kernel_class {
    MPMW_K_G2P_P2G = 0,        /* fused G2P + constitutive + chip + next P2G */
    MPMW_K_P2G = 1,            /* standalone P2G */
    MPMW_K_GRID = 2,           /* grid update + tool contact */
    MPMW_K_DIFFUSE = 3,        /* grid heat diffusion */
    MPMW_K_SORT = 4,           /* binning / sort / permute */
    MPMW_K_REDUCE = 5,         /* AABB / metrics / contact-force reductions */
    MPMW_K_OTHER = 6
};

typedef struct mpmw_solver mpmw_solver;
typedef struct mpmw_batch mpmw_batch;

const char *mpmw_last_error(void);
int mpmw_abi_version(void);

/* --- lifecycle: MPMWorkpieceSolver ctor/initialize/~ (MPMWorkpieceSolver.cu:418-458) */
void mpmw_params_default(mpmw_params *p);
int mpmw_create(const mpmw_params *p, mpmw_solver **out);
int mpmw_destroy(mpmw_solver *s);
/* reset(), MPMWorkpieceSolver.cu:830-841: particles, tool, time and step
 * count cleared; the tool pose returns to the one last given to
 * mpmw_set_tool_pose (the pose advances with every step). */
int mpmw_reset(mpmw_solver *s);

/* --- particles: uploadParticles/downloadParticles (:546-560) and
 *     getParticlesForExport (:870-894) in SoA form.  get_particles writes
 *     min(a->n, count) particles in device order and sets a->n. */
int mpmw_set_particles(mpmw_solver *s, const mpmw_particle_arrays *a);
int mpmw_get_particles(mpmw_solver *s, mpmw_particle_arrays *a);
int mpmw_particle_count(mpmw_solver *s, int32_t *n);
/* initializeParticleBox (:588-614) and initializeCylindricalWorkpiece (:616-648) */
int mpmw_init_particle_box(mpmw_solver *s, const double mn[3], const double mx[3], double spacing);
int mpmw_init_cylinder(mpmw_solver *s, const double c[3], double radius, double length,
                       double spacing, int32_t axis);
/* Seeded jittered slab (8 particles per cell, float-exact positions); fills
 * caller arrays of cx*cy*cz*8 entries.  No counterpart in the reference: it
 * generates the benchmark workpieces. */
int mpmw_gen_jittered_slab(uint64_t seed, int32_t cx, int32_t cy, int32_t cz, double h,
                           const double origin[3], double density, double fixture, double *pos,
                           double *mass, double *volume, int32_t *id, int32_t *status);
/* Generate the same slab straight into the solver (no host round trip). */
int mpmw_init_jittered_slab(mpmw_solver *s, uint64_t seed, int32_t cx, int32_t cy, int32_t cz,
                            const double origin[3], const double vel[3], double temperature);

/* --- tool: setToolMesh (:793-810) takes every stride-th of n mesh nodes
 *     (stride = max(1, n/4000)) and r_c = max(contact_radius, h). */
int mpmw_set_tool_samples(mpmw_solver *s, const double *xyz, int32_t n, double contact_radius);
int mpmw_set_tool_analytic(mpmw_solver *s, const mpmw_tool_desc *d);
int mpmw_set_tool_pose(mpmw_solver *s, const mpmw_tool_pose *p);
/* ToolSDF::build (ToolSDF.cu:212-280) with computeSDFCPU (:379-532): a
 * voxel SDF of the triangle mesh (verts: nverts xyz in the tool's rest
 * frame, tris: ntris vertex-index triples, normals: ntris outward unit
 * normals), bounding box grown by `padding`, voxel = longest extent /
 * voxels_per_axis, at most 256 voxels per axis.  Built on the device (the
 * flood fill of voxels beyond the scatter reach on the host); the values
 * equal the reference's single-precision host build.  The tool becomes
 * MPMW_TOOL_VOXEL with the contact law of d (band, friction, grad_eps; its
 * kind is ignored) and is sampled through the pose (setToolTransform,
 * :569-576) with the clamp-addressed trilinear filter of the texture
 * (:534-567): no contact outside the texel range [0, n - 1]. */
int mpmw_set_tool_sdf_mesh(mpmw_solver *s, const mpmw_tool_desc *d, const double *verts, int32_t nverts,
                           const int32_t *tris, const double *normals, int32_t ntris, double padding,
                           int32_t voxels_per_axis);
/* The voxel SDF in use: dims, origin and voxel size always (when non-NULL);
 * the dims[0]*dims[1]*dims[2] values (x fastest) when grid is non-NULL.
 * MPMW_E_STATE when no SDF was built.  No counterpart in the reference
 * (its grid lives only in the texture). */
int mpmw_get_tool_sdf(mpmw_solver *s, float *grid, int32_t dims[3], double origin[3], double *voxel);
/* Replaces ContactSolver::morphToolSDF (ContactSolver.cu:1053-1083): vertex i
 * of the mesh last given to mpmw_set_tool_sdf_mesh moves inward by wear[i]
 * along its vertex normal (the normalised sum of its normalised face
 * normals, ContactSolver::initSDF :1024-1044; vertices past nwear do not
 * move) and the SDF is rebuilt from the worn vertices with the mesh's
 * triangles and face normals (ToolSDF::rebuild, ToolSDF.cu:282-346).  The
 * reference takes the wear from its FEM tool's nodes; here the caller
 * passes it.  MPMW_E_STATE when no mesh was set. */
int mpmw_morph_tool_sdf(mpmw_solver *s, const double *wear, int32_t nwear);
/* setToolTipPosition, MPMWorkpieceSolver.cuh:133-135 */
int mpmw_set_tool_tip(mpmw_solver *s, double x, double y, double z);
/* applyExternalForce, MPMWorkpieceSolver.cu:812-818 */
int mpmw_apply_external_force(mpmw_solver *s, const double c[3], double radius, const double f[3]);

/* --- the step: MPMWorkpieceSolver::step(double), MPMWorkpieceSolver.cu:727-791.
 *     Enqueues on the solver's stream; returns before the GPU finishes except
 *     on grid-rebuild steps (every 25th), which read back the particle AABB. */
int mpmw_step(mpmw_solver *s, double dt);
int mpmw_run(mpmw_solver *s, int32_t nsteps, double dt);
int mpmw_synchronize(mpmw_solver *s);
/* getStableTimeStep, :820-828 */
int mpmw_stable_dt(mpmw_solver *s, double *dt);
/* getMaxStress/getMaxTemperature/getTotalKineticEnergy (refreshed every 25
 * steps, :781), getTotalContactForce (last step), time, step count. */
int mpmw_get_metrics(mpmw_solver *s, mpmw_metrics *m);
/* syncMetrics -> updateMetrics now, :855-868 */
int mpmw_sync_metrics(mpmw_solver *s);
/* compactParticles, :475-527 (order-preserving here) */
int mpmw_compact(mpmw_solver *s);
/* getBounds, :843-853: AABB of the particles at the last metrics refresh */
int mpmw_get_bounds(mpmw_solver *s, double b[6]);
/* getTemperatureAt :936-946, applyHeatFlux :918-934, setParticleTemperatures :906-916 */
int mpmw_temperature_at(mpmw_solver *s, double x, double y, double z, double *t);
int mpmw_apply_heat_flux(mpmw_solver *s, double x, double y, double z, double q);
int mpmw_set_particle_temperatures(mpmw_solver *s, const double *t, int32_t count);

/* Deterministic mode (no counterpart in the reference, whose FP64 atomicAdd
 * P2G sums in arbitrary order, MPMWorkpieceSolver.cu:114-131): every P2G sum
 * is formed in a fixed order, without float atomics, so reruns from the same
 * input agree in every bit.  Particles are re-binned whenever one has changed
 * cell since the last binning.  Takes effect at the next step. */
int mpmw_set_deterministic(mpmw_solver *s, int32_t on);

/* HIP graph of the step (no counterpart in the reference solver; its
 * MPMSolver records its step as a CUDA graph, MPMSolver.cu:1077-1157): the
 * copy of the step's work items and the step's kernels are replayed as one
 * graph while their launch configuration stays the same (captured again
 * after a re-binning changes it).  The per-step values travel in the copied
 * work items, so results are bit for bit those of the eager launches.  Off
 * by default; kernel timing (mpmw_set_kernel_timing) steps eagerly.
 * mpmw_graph_stats: captures and replays so far. */
int mpmw_set_graph(mpmw_solver *s, int32_t on);
int mpmw_graph_stats(mpmw_solver *s, int64_t out[2]);

/* Launch-configuration autotuner (no counterpart in the reference).  The
 * step's kernels take a block order (0: blockIdx.x, workgroups round-robin
 * over the 8 XCDs; 1: one contiguous range of blocks per XCD) for the tile
 * pass, the grid update and the post pass, and a block size of 64, 128 or
 * 256 for the grid update; every choice computes the same results bit for
 * bit (the orders are bijections of the blocks; the contact-force partials
 * are one per wave of nodes whatever the block size).  Defaults: tile 0,
 * grid 1 / 256, post 1.  mpmw_set_autotune(s, trials) with trials > 0 tunes
 * on the next steps: each kernel class in turn, each candidate for `trials`
 * steps without a grid rebuild, timed with events (such steps run without
 * graph replay and wait for their kernels), the median kept, the fastest
 * adopted; 0 stops tuning and keeps the configuration in use.
 * mpmw_set_launch_config sets one; mpmw_get_tuning reports the one in use,
 * the state and the medians. */
typedef struct mpmw_tuning {
    int32_t state;          /* 0 default or set, 1 tuning, 2 tuned */
    int32_t tile_order, grid_order, grid_block, post_order;
    int32_t trials;
    float tile_ms[2];       /* medians per candidate: tile order 0, 1 */
    float grid_ms[6];       /* grid (order, block): (1,256) (1,128) (1,64) (0,256) (0,128) (0,64) */
    float post_ms[2];       /* post order 1, 0 */
} mpmw_tuning;
int mpmw_set_autotune(mpmw_solver *s, int32_t trials);
int mpmw_set_launch_config(mpmw_solver *s, const mpmw_tuning *cfg);
int mpmw_get_tuning(mpmw_solver *s, mpmw_tuning *out);

/* --- APIC transfer variant: quadratic B-spline weights over 27 nodes with
 *     base = floor(x/h - 0.5) (getWeightsAndGradients, MPMSolver.cu:20-34),
 *     P2G momentum m w (v + C (x_i - x_p)) with none from FIXED_BOUNDARY
 *     particles (p2gKernel :370-421), and from G2P
 *     C = 0.99 * 3/h^2 * sum w v_i (x_i - x_p)^T, unchanged for CHIP and
 *     FIXED_BOUNDARY particles (g2pKernel :554-592); MPMW_TRANSFER_APIC_TEXTBOOK
 *     uses C = 4/h^2 * sum w v_i (x_i - x_p)^T.  Everything else in the
 *     step is the workpiece solver's.  Selecting a transfer sets every
 *     particle's C to zero; new particles start at zero.  The deterministic
 *     mode is not available with APIC (MPMW_E_STATE either way round). */
int mpmw_set_transfer(mpmw_solver *s, int32_t mode);
/* C of the first n particles, 9 values each (row-major); APIC only. */
int mpmw_get_affine(mpmw_solver *s, double *C, int32_t n);
/* Replace C of all n particles (n = particle count); APIC only. */
int mpmw_set_affine(mpmw_solver *s, const double *C, int32_t n);

/* --- parametric tool mesh: generateParametricToolMesh (main.cpp:83-233),
 * the reference application's tool surface when no mesh file is given, for
 * mpmw_set_tool_sdf_mesh.  Drill (r 5 mm, flute 25 mm, 118 deg point cone),
 * boring bar, reamer (flat bottom, 20 mm), thread mill (r 0.8 x major / 2,
 * 15 mm, V tip of the form angle), end mill (r 5 mm, 25 mm, flat).  The
 * caller's arrays hold MPMW_TOOLMESH_NODES nodes (xyz) and
 * MPMW_TOOLMESH_TRIS triangles (3 indices, xyz normal).  Host only. */
enum { MPMW_TOOLMESH_DRILL = 0, MPMW_TOOLMESH_BORING = 1, MPMW_TOOLMESH_REAMING = 2, MPMW_TOOLMESH_THREADING = 3,
       MPMW_TOOLMESH_MILLING = 4 };
enum { MPMW_TOOLMESH_NODES = 674, MPMW_TOOLMESH_TRIS = 1344 };
typedef struct mpmw_tool_mesh_params {
    int32_t kind, pad_;
    double boring_bar_length_mm, boring_bar_diameter_mm;   /* config machining_parameters, defaults 40, 10 */
    double reamer_diameter_mm;                             /* default 10 */
    double thread_major_diameter_mm, thread_form_angle_deg; /* defaults 10, 60 */
} mpmw_tool_mesh_params;
void mpmw_tool_mesh_params_default(mpmw_tool_mesh_params *p);
int mpmw_parametric_tool_mesh(const mpmw_tool_mesh_params *p, double *verts, int32_t *tris, double *normals,
                              int32_t *nverts, int32_t *ntris);

/* --- result formats of the serving loop
 * mpmw_export_ndjson replaces StdoutIPCExporter::exportParticles
 * (StdoutIPCExporter.cpp:70-113, with getParticlesForExport
 * MPMWorkpieceSolver.cu:870-894 as its input): one NDJSON line
 *   {"count":N,"positions_b64":"..","step":S,"temps_b64":"..","type":"particles"}\n
 * over the particles that are not INACTIVE, FP32 x y z and FP32 temperature
 * as little-endian bytes in standard base64 (Base64.h:13-52), encoded on the
 * device.  *len receives the line's length; with buf NULL nothing else is
 * written, with cap < *len the call fails with MPMW_E_RANGE.  No particle
 * exported: *len = 0 (the reference writes nothing).  The line is not NUL
 * terminated.
 * mpmw_export_vtk replaces VTKExporter::exportParticles(filename, particles)
 * (VTKExporter.cpp:252-336): the legacy ASCII POLYDATA file of the particles
 * that are not INACTIVE. */
int mpmw_export_ndjson(mpmw_solver *s, int32_t step, char *buf, size_t cap, size_t *len);
int mpmw_export_vtk(mpmw_solver *s, const char *path);
/* mpmw_export_ndjson with no device or pinned-host allocation, for callers
 * that cannot allocate (a serving loop that must not call hipMalloc between
 * steps, code around HIP graph capture).  No counterpart in the reference.
 * The same line, byte for byte, as mpmw_export_ndjson; its device buffers
 * live in one static scratch array of MPMW_EXPORT_SCRATCH_BYTES per device
 * (held until the process ends), which limits it to about 2.5 million
 * particles: a larger solver gets MPMW_E_RANGE and nothing is written.  The
 * line is copied straight into buf (pageable memory is slower than the
 * staged copy of mpmw_export_ndjson).  Calls on one device share the array
 * and are serialised by a lock inside the library. */
#define MPMW_EXPORT_SCRATCH_BYTES (80u << 20)
int mpmw_export_ndjson_noalloc(mpmw_solver *s, int32_t step, char *buf, size_t cap, size_t *len);

/* --- inspection (parity / measurement) */
int mpmw_get_grid_info(mpmw_solver *s, mpmw_grid_info *g);
/* Keep a copy of each step's post-P2G grid (mass, momentum, force, T*m). */
int mpmw_set_grid_snapshot(mpmw_solver *s, int32_t on);
/* Copy the last snapshot to host arrays of grid-node length (NULL skips). */
int mpmw_get_grid_snapshot(mpmw_solver *s, double *mass, double *mom, double *force, double *tm);
/* Per-kernel timing: MPMW_TIMING_OFF, MPMW_TIMING_ALL (every kernel class;
 * each timed launch adds two event records to the stream) or
 * MPMW_TIMING_PARTICLE_PASS (the fused particle pass only, the least
 * disturbance of the step). */
enum { MPMW_TIMING_OFF = 0, MPMW_TIMING_ALL = 1, MPMW_TIMING_PARTICLE_PASS = 2 };
int mpmw_set_kernel_timing(mpmw_solver *s, int32_t mode);
int mpmw_get_kernel_times(mpmw_solver *s, mpmw_kernel_times *t);
/* The hipStream_t the solver launches on, as an opaque pointer. */
int mpmw_get_stream(mpmw_solver *s, void **stream);
/* Device bytes currently held by the solver. */
int mpmw_device_bytes(mpmw_solver *s, int64_t *bytes);

/* --- batched stepping (no counterpart in the reference, which steps one
 *     solver per call).  A batch steps many solvers of one precision on one
 *     device together: each kernel of the step is launched once for all
 *     members over a flattened block list, every member running exactly the
 *     arithmetic of its own mpmw_step, so a member's results are the same
 *     alone and in any batch (bit for bit in the deterministic mode).
 *     Members keep their own state, step count and rebuild phase; they may
 *     join and leave between steps.  While in a batch a solver launches on
 *     the batch's stream.  Destroying a member removes it; destroying the
 *     batch removes every member (none is destroyed). */
int mpmw_batch_create(int32_t device, mpmw_batch **out);
int mpmw_batch_destroy(mpmw_batch *b);
int mpmw_batch_add(mpmw_batch *b, mpmw_solver *s);
int mpmw_batch_remove(mpmw_batch *b, mpmw_solver *s);
int mpmw_batch_count(mpmw_batch *b, int32_t *n);
/* One step of every member; dt[i] is the step of the i-th member in the
 * order they were added (removals close the gaps). */
int mpmw_batch_step(mpmw_batch *b, const double *dt, int32_t n);
int mpmw_batch_synchronize(mpmw_batch *b);
/* mpmw_set_graph / mpmw_graph_stats for the batch's step (the graph holds
 * every member's launches; a member joining or leaving captures anew). */
int mpmw_batch_set_graph(mpmw_batch *b, int32_t on);
int mpmw_batch_graph_stats(mpmw_batch *b, int64_t out[2]);
/* mpmw_set_autotune / mpmw_set_launch_config / mpmw_get_tuning for the
 * batch's step (one configuration for all its members). */
int mpmw_batch_set_autotune(mpmw_batch *b, int32_t trials);
int mpmw_batch_set_launch_config(mpmw_batch *b, const mpmw_tuning *cfg);
int mpmw_batch_get_tuning(mpmw_batch *b, mpmw_tuning *out);

/* Layout diagnostics of the particle binning. */
typedef struct mpmw_diagnostics {
    int64_t sorts;             /* times the particles were re-binned into tiles */
    int64_t out_of_window;     /* particle visits that took the device-memory path since the last grid rebuild */
    int64_t steps_since_rebuild;
    int32_t tiles, tile_cells; /* tiles of the current binning, cells per tile */
    int32_t max_tile_particles;
    int32_t loose_particles;   /* particles of the current binning outside every tile */
    int64_t det_rebins;        /* re-binnings the deterministic P2G forced (a particle changed cell) */
    int32_t deterministic;     /* 1 when the deterministic P2G is on */
    int32_t pad_;
} mpmw_diagnostics;
int mpmw_get_diagnostics(mpmw_solver *s, mpmw_diagnostics *d);

#ifdef __cplusplus
}
#endif
#endif
