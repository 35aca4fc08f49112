// mpmw_solver.hip -- host orchestration of the MI355X MPM workpiece step and
// the C ABI declared in include/mpmw.h.
//
// Replaces MPMWorkpieceSolver (src/MPMWorkpieceSolver.cu:418-946 of the
// reference).  Differences in how, not what:
//   * particle state lives in HBM as structure-of-arrays in the solver's
//     precision (FP32 bulk or FP64) and never round-trips to the host during
//     stepping; the grid rebuild every 25 steps reads back only the AABB
//     (the reference downloads and re-uploads every particle, :729-733);
//   * particles are binned into cell tiles; G2P, the chip pass and the next
//     step's P2G run as one particle pass per tile (k_tile) with the node
//     window staged in LDS, so each particle is read and written once per
//     step and the grid sees one atomic flush per tile;
//   * all reductions (AABB, metrics, contact force) are fixed-order block
//     partials, so their results do not depend on scheduling.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <limits>
#include <map>
#include <mutex>
#include <tuple>
#include <unordered_map>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "mpmw.h"
#include "mpmw_sdf.hpp"
#include "mpmw_apic.hpp"

using namespace mpmw;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

}  // namespace

// The error text of mpmw_last_error, for the library's other translation
// units (the F-tracking solver).
namespace mpmw_internal {
int set_error(int code, const std::string &msg) { return fail(code, msg); }
}  // namespace mpmw_internal

namespace {

#define HIP_TRY(expr)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail(MPMW_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));      \
    } while (0)

// Particle fields.  Real-valued fields first, then the two int32 fields.
enum Field {
    F_X, F_Y, F_Z, F_VX, F_VY, F_VZ,
    F_S0, F_S1, F_S2, F_S3, F_S4, F_S5,
    F_EPS, F_RATE, F_DMG, F_SRES, F_T, F_M, F_V, F_DENS,
    F_A0, F_A1, F_A2, F_A3, F_A4, F_A5,
    F_FX, F_FY, F_FZ,
    F_C0, F_C1, F_C2, F_C3, F_C4, F_C5, F_C6, F_C7, F_C8,
    F_STATUS, F_ID,
    NF
};
constexpr int kNumReal = F_STATUS;

inline int blocks_for(int64_t n, int b = kBlock) { return (int)std::max<int64_t>(1, (n + b - 1) / b); }

}  // namespace

// Staging of the work items of one step (see step_many): two pinned host
// areas used in turn, an event guarding each area's reuse, one device area.
// The rest mesh of a voxel tool, kept for wear morphing.
struct ToolMesh {
    std::vector<double> verts, normals, vnormals;  // xyz per vertex, per triangle, per vertex
    std::vector<int32_t> tris;
    double padding = 0.0;
    int32_t voxels = 0;
};

struct StepStaging {
    struct Area {
        void *h = nullptr;
        size_t cap = 0;
        hipEvent_t copied = nullptr;
        bool pending = false;
    };
    Area area[2];
    int parity = 0;
    void *d_blob = nullptr;
    size_t d_cap = 0;
    // HIP graphs of the step's launches, one per staging area: the copy of
    // that area into the device blob, then the kernels that read the blob.
    // A graph is replayed while its launch configuration (the signature)
    // stays the same, and captured again when it changes.
    bool graph_on = false;
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    std::vector<uint64_t> sig[2];
    int64_t captures = 0, replays = 0;
    // Launch configuration of the step's kernels and its autotuner.  Every
    // choice is a bijection of the launch's blocks or a block size whose
    // partials do not depend on it, so no choice changes a result.
    struct Tuner {
        int32_t tile_order = 0, grid_order = 1, grid_block = kBlock, post_order = 1;   // in use
        int32_t trials = 0;        // steps per candidate; 0: not tuning
        int32_t state = 0;         // 0 default / set, 1 tuning, 2 tuned
        int32_t phase = 0, cand = 0, trial = 0;
        float ms[3][6] = {};       // median per candidate: tile (2), grid (6), post (2)
        float sample[16] = {};
        hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    } tune;
};

struct mpmw_solver {
    mpmw_params prm{};
    hipStream_t stream = nullptr;
    bool fp64 = false;
    size_t rsz = 4;

    // particles
    int32_t n = 0, cap = 0;
    void *f[NF] = {};
    void *alt[NF] = {};    // second buffer set for permutations
    int32_t alt_cap = 0;
    bool has_back = false, has_fext = false;
    bool apic = false;     // APIC transfer variant (affine fields F_C0..F_C8 live)
    bool apic_textbook = false;  // APIC with 4/h^2 and no decay instead of 3/h^2 and 0.99
    ToolMesh mesh;         // rest mesh of the voxel tool (wear morphing)

    // grid
    double minx = 0, miny = 0, minz = 0, h = 0;
    int32_t nx = 1, ny = 1, nz = 1, kfix = -1;
    int32_t tile_off[3] = {4, 4, 4};  // first tiled cell per axis (the AABB-min particle's cell)
    int64_t nn = 0, nn_cap = 0;
    void *acc = nullptr, *node = nullptr, *gm = nullptr, *gtm = nullptr, *snap = nullptr;
    bool snap_on = false, snap_valid = false;

    // reductions
    double *part = nullptr;        // k_tile / k_reduce partials
    int64_t part_cap = 0;
    double *fpart = nullptr;       // contact-force partials, then the sums of their chunks
    int64_t fpart_cap = 0;
    int32_t cf_first = 0, cf_chunks = 0;   // chunk sums of the last step not yet added (fpart + cf_first)
    double *d_stats = nullptr;     // kRed finalized statistics
    double *d_cf = nullptr;        // contact force of the last step
    int32_t *d_modes = nullptr;    // finalize modes for kRed columns, then one "sum"
    double *h_stats = nullptr;     // pinned copy of d_stats
    double *h_cf = nullptr;        // pinned copy of d_cf
    double *d_near = nullptr;      // nearest-particle partials
    int32_t *d_near_i = nullptr;
    double *d_near_out = nullptr;
    int32_t *d_near_out_i = nullptr;

    // tool
    int32_t tool_kind = MPMW_TOOL_NONE;
    void *samples = nullptr;       // 3 * nsamples R (SoA)
    int32_t nsamples = 0;
    double sample_lo[3] = {0, 0, 0}, sample_hi[3] = {0, 0, 0};
    mpmw_tool_desc desc{};
    mpmw_tool_pose pose{};
    mpmw_tool_pose pose0{};        // the pose as last set (restored by reset)
    void *sdf = nullptr;           // voxel SDF of the tool mesh (float, x fastest)
    int32_t sdf_dims[3] = {0, 0, 0};
    double sdf_origin[3] = {0, 0, 0}, sdf_voxel = 0.0;
    double tip[3] = {0, 0, 0};

    // stepping state
    double time = 0.0;
    int64_t step = 0;
    bool p2g_ready = false;        // accumulators hold this step's P2G
    bool stats_fresh = false;      // d_stats/h_stats hold the current AABB
    double max_stress = 0, max_temp = 0, ke = 0, contact_force = 0;
    bool metrics_pending = false, cf_pending = false;
    double bounds[6] = {0, 0, 0, 0, 0, 0};
    bool deterministic = false;
    bool uniform_mv = false;       // every particle has mass m0 and volume V0
    double m0 = 0.0, V0 = 0.0;

    // sorting
    SortState sort{};
    // deterministic mode: per-tile P2G windows (ntiles x 8 x WN of R)
    void *tpart = nullptr;
    int64_t tpart_cap = 0;         // bytes

    // timing
    bool timing = false;
    bool timing_all = false;       // every kernel class, else the fused particle pass only
    struct Ev { hipEvent_t a, b; int cls; };
    std::vector<Ev> events;
    std::vector<hipEvent_t> ev_pool;
    double t_ms[8] = {};
    int64_t t_n[8] = {};

    int64_t device_bytes = 0;

    // batch membership: while in a batch the solver launches on the batch's
    // stream; own_stream is its own, restored when it leaves
    mpmw_batch *batch = nullptr;
    hipStream_t own_stream = nullptr;
    StepStaging staging;
};

template <typename R> int permute_particles(mpmw_solver *s, const int32_t *perm, int32_t count, bool host_perm);

namespace {

// ------------------------------------------------------------ allocation --

// Device and pinned host blocks are kept for reuse after release (per
// device, by size class), and streams come from a small per-device pool:
// creating a stream or allocating and freeing memory costs milliseconds,
// which a server admitting and retiring simulations would otherwise pay per
// request.  Callers release a block only after the stream that used it has
// drained.  Size classes: 256 B, then multiples of 1/8 of the enclosing power
// of two (at most 12.5 % over the request).
namespace pool {

std::mutex mu;
std::map<std::tuple<int, int, size_t>, std::vector<void *>> free_blocks;
std::unordered_map<void *, std::tuple<int, int, size_t>> live_blocks;
std::map<int, std::vector<hipStream_t>> streams;
std::map<int, size_t> next_stream;
constexpr size_t kStreamsPerDevice = 4;

size_t size_class(size_t b) {
    if (b <= 256) return 256;
    size_t p = 1;
    while (p <= b / 2) p *= 2;
    const size_t step = p / 8;
    return (b + step - 1) / step * step;
}

hipError_t get(int dev, bool host, size_t bytes, void **out) {
    const size_t c = size_class(bytes);
    const auto key = std::make_tuple(dev, host ? 1 : 0, c);
    {
        std::lock_guard<std::mutex> g(mu);
        auto it = free_blocks.find(key);
        if (it != free_blocks.end() && !it->second.empty()) {
            *out = it->second.back();
            it->second.pop_back();
            live_blocks[*out] = key;
            return hipSuccess;
        }
    }
    hipError_t e = host ? hipHostMalloc(out, c, hipHostMallocDefault) : hipMalloc(out, c);
    if (e != hipSuccess) {
        // free the cached blocks of this device and try once more
        std::lock_guard<std::mutex> g(mu);
        for (auto &kv : free_blocks) {
            if (std::get<0>(kv.first) != dev || std::get<1>(kv.first) != (host ? 1 : 0)) continue;
            for (void *q : kv.second) (void)(host ? hipHostFree(q) : hipFree(q));
            kv.second.clear();
        }
        (void)hipGetLastError();
        e = host ? hipHostMalloc(out, c, hipHostMallocDefault) : hipMalloc(out, c);
        if (e != hipSuccess) return e;
        live_blocks[*out] = key;
        return e;
    }
    std::lock_guard<std::mutex> g(mu);
    live_blocks[*out] = key;
    return e;
}

size_t block_size(void *p) {
    std::lock_guard<std::mutex> g(mu);
    auto it = live_blocks.find(p);
    return it == live_blocks.end() ? 0 : std::get<2>(it->second);
}

void put(void *p) {
    if (!p) return;
    std::lock_guard<std::mutex> g(mu);
    auto it = live_blocks.find(p);
    if (it == live_blocks.end()) return;
    free_blocks[it->second].push_back(p);
    live_blocks.erase(it);
}

// A stream of device `dev` (the current device) from the pool, round robin.
hipError_t stream(int dev, hipStream_t *out) {
    std::lock_guard<std::mutex> g(mu);
    auto &v = streams[dev];
    if (v.size() < kStreamsPerDevice) {
        hipStream_t st;
        hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
        if (e != hipSuccess) {
            if (v.empty()) return e;
        } else {
            v.push_back(st);
            *out = st;
            return hipSuccess;
        }
    }
    *out = v[next_stream[dev]++ % v.size()];
    return hipSuccess;
}

}  // namespace pool

int dalloc(mpmw_solver *s, void **p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0) return MPMW_OK;
    hipError_t e = pool::get(s->prm.device, false, bytes, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        return fail(MPMW_E_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e));
    }
    s->device_bytes += (int64_t)pool::block_size(*p);
    return MPMW_OK;
}

// Release a device block of the solver to the pool once the solver's stream
// has drained (work in flight may still use it).
void dfree(mpmw_solver *s, void *&p, size_t bytes) {
    (void)bytes;
    if (p) {
        if (s->stream) (void)hipStreamSynchronize(s->stream);
        s->device_bytes -= (int64_t)pool::block_size(p);
        pool::put(p);
    }
    p = nullptr;
}

// Positions are FP64 in every precision mode; ids and status are int32.
inline bool is_pos(int fld) { return fld == F_X || fld == F_Y || fld == F_Z; }

size_t field_bytes(const mpmw_solver *s, int fld, int64_t count) {
    if (fld >= kNumReal) return (size_t)count * sizeof(int32_t);
    return (size_t)count * (is_pos(fld) ? sizeof(double) : s->rsz);
}

bool field_live(const mpmw_solver *s, int fld) {
    if (fld >= F_A0 && fld <= F_A5) return s->has_back;
    if (fld >= F_FX && fld <= F_FZ) return s->has_fext;
    if (fld >= F_C0 && fld <= F_C8) return s->apic;
    return true;
}

void free_particles(mpmw_solver *s) {
    for (int k = 0; k < NF; ++k) {
        dfree(s, s->f[k], field_bytes(s, k, s->cap));
        dfree(s, s->alt[k], field_bytes(s, k, s->alt_cap));
    }
    s->cap = s->alt_cap = 0;
}

// (Re)allocate the live particle fields for `count` particles; contents are
// not preserved.
int alloc_particles(mpmw_solver *s, int32_t count) {
    if (count <= s->cap) {
        for (int k = 0; k < NF; ++k)
            if (field_live(s, k) && !s->f[k]) {
                int rc = dalloc(s, &s->f[k], field_bytes(s, k, s->cap));
                if (rc) return rc;
            }
        return MPMW_OK;
    }
    free_particles(s);
    s->cap = count;
    for (int k = 0; k < NF; ++k) {
        if (!field_live(s, k)) continue;
        int rc = dalloc(s, &s->f[k], field_bytes(s, k, count));
        if (rc) return rc;
    }
    return MPMW_OK;
}

// Allocate a field if needed and set all cap entries to zero (a buffer kept
// from an earlier particle set may hold stale values).
int ensure_field(mpmw_solver *s, int fld) {
    if (s->cap == 0) return MPMW_OK;
    if (!s->f[fld]) {
        int rc = dalloc(s, &s->f[fld], field_bytes(s, fld, s->cap));
        if (rc) return rc;
    }
    const int64_t words = (int64_t)(field_bytes(s, fld, s->cap) / 4);
    hipLaunchKernelGGL(k_fill<int32_t>, dim3(std::min(blocks_for(words), 4096)), dim3(kBlock), 0, s->stream,
                       (int32_t *)s->f[fld], words, 0);
    HIP_TRY(hipGetLastError());
    return MPMW_OK;
}

int ensure_alt(mpmw_solver *s) {
    if (s->alt_cap >= s->n && s->alt_cap == s->cap) {
        for (int k = 0; k < NF; ++k)
            if (field_live(s, k) && !s->alt[k]) {
                int rc = dalloc(s, &s->alt[k], field_bytes(s, k, s->alt_cap));
                if (rc) return rc;
            }
        return MPMW_OK;
    }
    for (int k = 0; k < NF; ++k) dfree(s, s->alt[k], field_bytes(s, k, s->alt_cap));
    s->alt_cap = s->cap;
    for (int k = 0; k < NF; ++k) {
        if (!field_live(s, k)) continue;
        int rc = dalloc(s, &s->alt[k], field_bytes(s, k, s->alt_cap));
        if (rc) return rc;
    }
    return MPMW_OK;
}

template <typename T> int grow(mpmw_solver *s, T *&p, int64_t &cap, int64_t need) {
    if (need <= cap) return MPMW_OK;
    void *q = p;
    dfree(s, q, sizeof(T) * (size_t)cap);
    p = nullptr;
    int64_t c = std::max<int64_t>(need, cap * 3 / 2);
    int rc = dalloc(s, &q, sizeof(T) * (size_t)c);
    if (rc) return rc;
    p = (T *)q;
    cap = c;
    return MPMW_OK;
}

// Adds the chunk sums of the last step's contact-force partials, in order,
// into d_cf: before the force is read, or before a larger grid replaces the
// buffer that holds them.
static int add_contact_chunks(mpmw_solver *s) {
    if (s->cf_chunks == 0) return MPMW_OK;
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kBlock), 0, s->stream, (const double *)(s->fpart + s->cf_first),
                       s->cf_chunks, 1, (const int32_t *)(s->d_modes + kRed), s->d_cf);
    HIP_TRY(hipGetLastError());
    s->cf_chunks = 0;
    return MPMW_OK;
}

// ---------------------------------------------------------------- views ---

template <typename R> Particles<R> pview(mpmw_solver *s) {
    Particles<R> P;
    auto F = [&](int k) { return (R *)s->f[k]; };
    P.x = (double *)s->f[F_X]; P.y = (double *)s->f[F_Y]; P.z = (double *)s->f[F_Z];
    P.vx = F(F_VX); P.vy = F(F_VY); P.vz = F(F_VZ);
    P.s0 = F(F_S0); P.s1 = F(F_S1); P.s2 = F(F_S2); P.s3 = F(F_S3); P.s4 = F(F_S4); P.s5 = F(F_S5);
    P.a0 = s->has_back ? F(F_A0) : nullptr; P.a1 = F(F_A1); P.a2 = F(F_A2);
    P.a3 = F(F_A3); P.a4 = F(F_A4); P.a5 = F(F_A5);
    P.fx = s->has_fext ? F(F_FX) : nullptr; P.fy = F(F_FY); P.fz = F(F_FZ);
    P.eps = F(F_EPS); P.rate = F(F_RATE); P.dmg = F(F_DMG); P.sres = F(F_SRES);
    P.T = F(F_T);
    for (int c = 0; c < 9; ++c) P.aff[c] = s->apic ? F(F_C0 + c) : nullptr;
    // uniform mass and volume are passed as scalars (the arrays stay for export)
    P.m = s->uniform_mv ? nullptr : F(F_M);
    P.V = F(F_V);
    P.m0 = (R)s->m0;
    P.V0 = (R)s->V0;
    P.status = (int32_t *)s->f[F_STATUS];
    P.id = (int32_t *)s->f[F_ID];
    P.n = s->n;
    return P;
}

template <typename R> GridView<R> gview(const mpmw_solver *s) {
    GridView<R> g;
    g.ox = s->minx; g.oy = s->miny; g.oz = s->minz;
    g.hd = s->h;
    g.inv_hd = 1.0 / s->h;
    // two ulps of the largest cell coordinate, with margin
    g.guard = 8.0 * std::ldexp(1.0, -52) * (double)std::max(std::max(s->nx, s->ny), std::max(s->nz, 1));
    g.inv_h = (R)(1.0 / s->h);
    g.nx = s->nx; g.ny = s->ny; g.nz = s->nz;
    g.kfix = s->kfix;
    g.nn = s->nn;
    g.apic_k = s->apic_textbook ? 4.0 : 3.0;
    g.apic_decay = s->apic_textbook ? 1.0 : 0.99;
    return g;
}

template <typename R> Grid<R> grid_arrays(mpmw_solver *s) {
    Grid<R> G;
    G.acc = (R *)s->acc;
    G.node = (vec4_t<R> *)s->node;
    G.gm = (R *)s->gm;
    G.gtm = (R *)s->gtm;
    G.snap = s->snap_on ? (R *)s->snap : nullptr;
    return G;
}

template <typename R> Material<R> material(const mpmw_solver *s) {
    const mpmw_params &p = s->prm;
    Material<R> M;
    M.shear = (R)(p.young / (2.0 * (1.0 + p.poisson)));
    M.bulk = (R)(p.young / (3.0 * (1.0 - 2.0 * p.poisson)));
    M.density = (R)p.density; M.cp = (R)p.specific_heat;
    M.melt = (R)p.melt_temp; M.ref_temp = (R)p.ref_temp;
    M.jcA = (R)p.jc_A; M.jcB = (R)p.jc_B; M.jcN = (R)p.jc_n; M.jcC = (R)p.jc_C; M.jcM = (R)p.jc_m;
    M.ref_rate = (R)p.ref_strain_rate;
    M.failure_strain = (R)p.failure_strain; M.dth = (R)p.damage_threshold;
    M.eps_crit = (R)p.critical_plastic_strain; M.chip_ratio = (R)p.thermal_chip_ratio;
    M.rc = (R)p.tool_contact_radius; M.c_damp = (R)p.contact_damping;
    M.dT_max = (R)p.max_contact_temp_rise; M.H = (R)p.kinematic_hardening;
    M.clamp = p.thermal_clamp_enabled;
    // divisors formed as the kernels formed them (in R), and their reciprocals
    M.d_ref_rate = std::max(M.ref_rate, (R)1e-9);
    M.d_t_span = std::max(M.melt - M.ref_temp, (R)1);
    M.d_3g = std::max((R)3 * M.shear, (R)1);
    M.d_rho_cp = std::max(M.density * M.cp, (R)1);
    M.d_fail = std::max(M.failure_strain, (R)1e-3);
    M.d_deg = M.dth * (R)0.3;
    M.inv_ref_rate = (R)(1.0 / (double)M.d_ref_rate);
    M.inv_t_span = (R)(1.0 / (double)M.d_t_span);
    M.inv_3g = (R)(1.0 / (double)M.d_3g);
    M.inv_rho_cp = (R)(1.0 / (double)M.d_rho_cp);
    M.inv_fail = (R)(1.0 / (double)M.d_fail);
    M.inv_deg = (R)(1.0 / (double)M.d_deg);
    // johnsonCookStress >= A * 1 * 0.01 when A, B, C >= 0 (strain term >= A,
    // rate term >= 1 since the rate ratio is >= 1, thermal term >= 0.01); the
    // bound is rounded like the kernel's product, which rounding keeps monotone
    if (p.jc_A >= 0.0 && p.jc_B >= 0.0 && p.jc_C >= 0.0 && p.jc_n > 0.0)
        M.flow_floor = M.jcA * (R)0.01;
    else
        M.flow_floor = -std::numeric_limits<R>::infinity();
    return M;
}

Tool tool_view(const mpmw_solver *s) {
    Tool t{};
    t.kind = s->tool_kind;
    if (t.kind == MPMW_TOOL_POINTS) {
        t.nsamples = s->nsamples;
        t.sx = (const double *)s->samples;
        t.sy = t.sx + s->nsamples;
        t.sz = t.sy + s->nsamples;
        for (int a = 0; a < 3; ++a) { t.lo[a] = s->sample_lo[a]; t.hi[a] = s->sample_hi[a]; }
        t.rc2 = s->prm.tool_contact_radius * s->prm.tool_contact_radius;
    } else if (t.kind == MPMW_TOOL_WEDGE || t.kind == MPMW_TOOL_DRILL || t.kind == MPMW_TOOL_VOXEL) {
        const mpmw_tool_desc &d = s->desc;
        const mpmw_tool_pose &p = s->pose;
        t.cx = p.center[0]; t.cy = p.center[1]; t.cz = p.center[2];
        t.cos_a = std::cos(p.angle); t.sin_a = std::sin(p.angle);
        t.vx = p.vel[0]; t.vy = p.vel[1]; t.vz = p.vel[2]; t.omega = p.omega;
        t.band = d.band; t.mu = d.friction; t.eps = d.grad_eps;
        t.sin_rake = std::sin(d.rake_angle); t.cos_rake = std::cos(d.rake_angle);
        t.sin_clear = std::sin(d.clearance_angle); t.cos_clear = std::cos(d.clearance_angle);
        t.radius = d.radius; t.tan_helix = std::tan(d.helix_angle);
        t.sin_half_point = std::sin(0.5 * d.point_angle); t.cos_half_point = std::cos(0.5 * d.point_angle);
        t.flute_offset = d.flute_offset; t.flute_radius = d.flute_radius; t.flute_length = d.flute_length;
        if (t.kind == MPMW_TOOL_VOXEL) {
            t.sdf = (const float *)s->sdf;
            t.snx = s->sdf_dims[0]; t.sny = s->sdf_dims[1]; t.snz = s->sdf_dims[2];
            t.sox = s->sdf_origin[0]; t.soy = s->sdf_origin[1]; t.soz = s->sdf_origin[2];
            t.svox = s->sdf_voxel;
        }
    }
    return t;
}

// --------------------------------------------------------------- timing ---

hipEvent_t take_event(mpmw_solver *s) {
    if (!s->ev_pool.empty()) {
        hipEvent_t e = s->ev_pool.back();
        s->ev_pool.pop_back();
        return e;
    }
    // timing only: no system-scope fence (with one, every record wrote back
    // and invalidated the caches, 30-60 us per step of the drill slab)
    hipEvent_t e;
    if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return nullptr;
    return e;
}

struct Timed {
    mpmw_solver *s;
    mpmw_solver::Ev ev{};
    bool on;
    Timed(mpmw_solver *s_, int cls)
        : s(s_), on(s_ && s_->timing && (s_->timing_all || cls == MPMW_K_G2P_P2G)) {
        if (!on) return;
        ev.a = take_event(s);
        ev.b = take_event(s);
        ev.cls = cls;
        if (!ev.a || !ev.b) { on = false; return; }
        (void)hipEventRecord(ev.a, s->stream);
    }
    ~Timed() {
        if (!on) return;
        (void)hipEventRecord(ev.b, s->stream);
        s->events.push_back(ev);
    }
};

void drain_events(mpmw_solver *s) {
    if (s->events.empty()) return;
    (void)hipStreamSynchronize(s->stream);
    for (auto &e : s->events) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, e.a, e.b) == hipSuccess) {
            s->t_ms[e.cls] += ms;
            s->t_n[e.cls] += 1;
        }
        s->ev_pool.push_back(e.a);
        s->ev_pool.push_back(e.b);
    }
    s->events.clear();
}

// ----------------------------------------------------------- reductions ---

int ensure_part(mpmw_solver *s, int64_t nblocks) { return grow(s, s->part, s->part_cap, nblocks * kRed); }

// Finalize kRed statistics from `nparts` partials into d_stats and start the
// copy to the pinned host buffer.
int finalize_stats(mpmw_solver *s, int nparts) {
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kBlock), 0, s->stream, s->part, nparts, kRed, s->d_modes,
                       s->d_stats);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(s->h_stats, s->d_stats, sizeof(double) * kRed, hipMemcpyDeviceToHost, s->stream));
    return MPMW_OK;
}

template <typename R> int reduce_particles(mpmw_solver *s) {
    int nb = std::min(blocks_for(s->n), 2048);
    int rc = ensure_part(s, nb);
    if (rc) return rc;
    {
        Timed t(s, MPMW_K_REDUCE);
        hipLaunchKernelGGL(k_reduce<R>, dim3(nb), dim3(kBlock), 0, s->stream, pview<R>(s), s->prm.ref_temp,
                           s->part);
        HIP_TRY(hipGetLastError());
        rc = finalize_stats(s, nb);
    }
    return rc;
}

// Copy finalized statistics (already synchronised) into the cached metrics.
void absorb_stats(mpmw_solver *s) {
    const double *v = s->h_stats;
    s->bounds[0] = v[0]; s->bounds[1] = v[1]; s->bounds[2] = v[2];
    s->bounds[3] = -v[3]; s->bounds[4] = -v[4]; s->bounds[5] = -v[5];
    s->max_stress = v[6];
    s->max_temp = v[7];
    s->ke = v[8];
}

// ------------------------------------------------------------ the grid ----

// rebuildGridFromHostBounds, MPMWorkpieceSolver.cu:570-586, from the AABB in
// h_stats (min xyz, -max xyz).
int rebuild_grid_from_stats(mpmw_solver *s) {
    const double *v = s->h_stats;
    double mn[3] = {v[0], v[1], v[2]}, mx[3] = {-v[3], -v[4], -v[5]};
    if (!(mn[0] <= mx[0])) {  // no non-INACTIVE particle: reference leaves +-1e30 bounds
        mn[0] = mn[1] = mn[2] = 1e30;
        mx[0] = mx[1] = mx[2] = -1e30;
    }
    const double cell = s->h;
    const double pad = cell * 4.0;
    s->minx = mn[0] - pad;
    s->miny = mn[1] - pad;
    s->minz = mn[2] - pad;
    // tiles start at the AABB-min particle's cell, computed with the kernels'
    // expression floor((x - origin) / h): 4, or 3 when the division rounds down
    const double org[3] = {s->minx, s->miny, s->minz};
    for (int a = 0; a < 3; ++a) {
        const double fl = std::floor((mn[a] - org[a]) / cell);
        s->tile_off[a] = (fl >= 0.0 && fl <= 4.0) ? (int32_t)fl : 4;
    }
    // max(4, (int)ceil(extent / h) + 1), with the int conversion kept in range
    auto dim = [&](double lo, double hi, int32_t &out) {
        const double c = std::ceil((hi - lo + 2 * pad) / cell) + 1.0;
        if (!(c < 1e9)) return false;
        out = c < 4.0 ? 4 : (int32_t)c;
        return true;
    };
    if (!dim(mn[0], mx[0], s->nx) || !dim(mn[1], mx[1], s->ny) || !dim(mn[2], mx[2], s->nz))
        return fail(MPMW_E_RANGE, "particle bounds give a grid beyond 1e9 nodes per axis");
    const int64_t nn = (int64_t)s->nx * s->ny * s->nz;
    // node indices are 32-bit in the grid kernels
    if (nn >= (int64_t)1 << 31)
        return fail(MPMW_E_RANGE, "grid of " + std::to_string(nn) + " nodes is too large (limit 2^31 - 1)");
    // fixture: nodes with pos.z <= minZ + fixture + h are clamped (:153)
    const double fixture = s->prm.has_config ? s->prm.fixture_thickness : 0.002;
    s->kfix = -1;
    for (int k = 0; k < s->nz; ++k) {
        double pz = s->minz + k * cell;
        if (pz <= s->minz + fixture + cell) s->kfix = k;
        else break;
    }
    if (nn > s->nn_cap) {
        dfree(s, s->acc, 8 * s->rsz * (size_t)s->nn_cap);
        dfree(s, s->node, 4 * s->rsz * (size_t)s->nn_cap);
        dfree(s, s->gm, s->rsz * (size_t)s->nn_cap);
        dfree(s, s->gtm, s->rsz * (size_t)s->nn_cap);
        dfree(s, s->snap, 8 * s->rsz * (size_t)s->nn_cap);
        int64_t c = std::max<int64_t>(nn, s->nn_cap * 5 / 4);
        int rc;
        if ((rc = dalloc(s, &s->acc, 8 * s->rsz * (size_t)c))) return rc;
        if ((rc = dalloc(s, &s->node, 4 * s->rsz * (size_t)c))) return rc;
        if ((rc = dalloc(s, &s->gm, s->rsz * (size_t)c))) return rc;
        if ((rc = dalloc(s, &s->gtm, s->rsz * (size_t)c))) return rc;
        if (s->snap_on && (rc = dalloc(s, &s->snap, 8 * s->rsz * (size_t)c))) return rc;
        s->nn_cap = c;
        s->snap_valid = false;
    }
    s->nn = nn;
    // one contact-force partial per wave, then the sums of their chunks
    const int64_t fwaves = (nn + kWave - 1) / kWave, need = fwaves + (fwaves + kSumChunk - 1) / kSumChunk;
    int rc = need > s->fpart_cap ? add_contact_chunks(s) : MPMW_OK;
    if (rc || (rc = grow(s, s->fpart, s->fpart_cap, need))) return rc;
    s->p2g_ready = false;
    return MPMW_OK;
}

template <typename R> int zero_acc(mpmw_solver *s) {
    hipLaunchKernelGGL(k_fill<R>, dim3(std::min(blocks_for(8 * s->nn), 4096)), dim3(kBlock), 0, s->stream,
                       (R *)s->acc, 8 * s->nn, R(0));
    HIP_TRY(hipGetLastError());
    return MPMW_OK;
}

// ------------------------------------------------------------ binning -----

TileGeom tile_geom(const mpmw_solver *s) {
    auto tiles = [](int n, int off, int T) { return std::max(1, (n - 1 - off + T - 1) / T); };
    TileGeom t;
    t.ox = s->tile_off[0];
    t.oy = s->tile_off[1];
    t.oz = s->tile_off[2];
    t.ntx = tiles(s->nx, t.ox, TX);
    t.nty = tiles(s->ny, t.oy, TY);
    t.ntz = tiles(s->nz, t.oz, TZ);
    t.ntiles = t.ntx * t.nty * t.ntz;
    t.nkeys = t.ntiles * kTileCells;
    return t;
}

template <typename T> int grow32(mpmw_solver *s, T *&p, int32_t &cap, int64_t need) {
    int64_t c = cap;
    int rc = grow(s, p, c, need);
    cap = (int32_t)c;
    return rc;
}

void free_sort_state(mpmw_solver *s) {
    SortState &S = s->sort;
    void *q;
    q = S.keys; dfree(s, q, 0);
    q = S.keys_sorted; dfree(s, q, 0);
    q = S.idx; dfree(s, q, 0);
    q = S.idx_sorted; dfree(s, q, 0);
    q = S.perm; dfree(s, q, 0);
    q = S.counts; dfree(s, q, 0);
    q = S.cell_start; dfree(s, q, 0);
    q = S.tile_start; dfree(s, q, 0);
    q = S.oow; dfree(s, q, 0);
    pool::put(S.h_oow);
    free_scratch(S.temp);
    free_scratch(S.temp2);
    S = SortState();
}

// Bin the particles into cell tiles of the current grid: stable radix sort
// by tile-major cell key, slot-major order inside each tile, then one gather
// of every particle field.  Only the order changes, never a value.
// sort_launch enqueues the binning and the copy of the loose range to the
// host; after the stream has been synchronised, sort_finish takes the range
// and marks the binning valid.
template <typename R> int sort_launch(mpmw_solver *s) {
    SortState &S = s->sort;
    const TileGeom tg = tile_geom(s);
    if ((int64_t)tg.ntiles * kTileCells >= INT32_MAX) return fail(MPMW_E_RANGE, "too many tiles for the binning");
    const int n = s->n;
    int rc;
    int32_t ncap = S.n_cap, ncap2 = S.n_cap, ncap3 = S.n_cap, ncap4 = S.n_cap, ncap5 = S.n_cap;
    if ((rc = grow32(s, S.keys, ncap, n)) || (rc = grow32(s, S.keys_sorted, ncap2, n)) ||
        (rc = grow32(s, S.idx, ncap3, n)) || (rc = grow32(s, S.idx_sorted, ncap4, n)) ||
        (rc = grow32(s, S.perm, ncap5, n)))
        return rc;
    S.n_cap = std::min(std::min(ncap, ncap2), std::min(std::min(ncap3, ncap4), ncap5));
    int32_t kc = S.key_cap, kc2 = S.key_cap;
    if ((rc = grow32(s, S.counts, kc, tg.nkeys + 2)) || (rc = grow32(s, S.cell_start, kc2, tg.nkeys + 2))) return rc;
    S.key_cap = std::min(kc, kc2);
    if ((rc = grow32(s, S.tile_start, S.tile_cap, tg.ntiles + 1))) return rc;
    if (!S.oow) {
        void *q;
        if ((rc = dalloc(s, &q, 3 * sizeof(int32_t)))) return rc;
        S.oow = (int32_t *)q;
        HIP_TRY(pool::get(s->prm.device, true, 4 * sizeof(int32_t), (void **)&S.h_oow));
        S.h_oow[0] = S.h_oow[1] = S.h_oow[2] = S.h_oow[3] = 0;
        hipLaunchKernelGGL(k_fill<int32_t>, dim3(1), dim3(64), 0, s->stream, S.oow, (int64_t)3, 0);
    }
    {
        const int64_t bytes = (int64_t)tg.ntiles * 8 * (s->apic ? QN : WN) * (int64_t)s->rsz;
        if (bytes > s->tpart_cap) {
            const int64_t c = std::max<int64_t>(bytes, s->tpart_cap * 5 / 4);
            dfree(s, s->tpart, (size_t)s->tpart_cap);
            s->tpart_cap = 0;
            if ((rc = dalloc(s, &s->tpart, (size_t)c))) return rc;
            s->tpart_cap = c;
        }
    }
    int nbits = 1;
    while (((int64_t)1 << nbits) <= tg.nkeys + 1) ++nbits;
    {
        Timed t(s, MPMW_K_SORT);
        hipLaunchKernelGGL(k_fill<int32_t>, dim3(std::min(blocks_for(tg.nkeys + 2), 4096)), dim3(kBlock), 0,
                           s->stream, S.counts, (int64_t)tg.nkeys + 2, 0);
        hipLaunchKernelGGL(k_bin_keys<R>, dim3(blocks_for(n)), dim3(kBlock), 0, s->stream, pview<R>(s),
                           gview<R>(s), tg, s->apic ? 1 : 0, S.keys, S.idx, S.counts);
        HIP_TRY(hipGetLastError());
        HIP_TRY(radix_sort_pairs(S.temp, S.keys, S.keys_sorted, S.idx, S.idx_sorted, n, nbits, s->stream));
        HIP_TRY(exclusive_scan(S.temp2, S.counts, S.cell_start, tg.nkeys + 2, s->stream));
        hipLaunchKernelGGL(k_slot_major, dim3(tg.ntiles), dim3(kBlock), 0, s->stream, S.cell_start, S.idx_sorted,
                           S.perm, S.tile_start);
        hipLaunchKernelGGL(k_tail, dim3(blocks_for(n)), dim3(kBlock), 0, s->stream, S.cell_start, tg.nkeys,
                           tg.ntiles, S.idx_sorted, S.perm, S.tile_start, n);
        HIP_TRY(hipGetLastError());
    }
    if ((rc = permute_particles<R>(s, S.perm, n, false))) return rc;
    // the loose range [cell_start[nkeys], cell_start[nkeys + 1]) for the host
    HIP_TRY(hipMemcpyAsync(S.h_oow + 2, S.cell_start + tg.nkeys, 2 * sizeof(int32_t), hipMemcpyDeviceToHost,
                           s->stream));
    S.geom = tg;
    return MPMW_OK;
}

void sort_finish(mpmw_solver *s) {
    SortState &S = s->sort;
    S.loose_begin = S.h_oow[2];
    S.loose_end = S.h_oow[3];
    S.valid = true;
    S.origin[0] = s->minx; S.origin[1] = s->miny; S.origin[2] = s->minz;
    S.sorts++;
}

template <typename R> int sort_particles(mpmw_solver *s) {
    int rc = sort_launch<R>(s);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(s->stream));
    sort_finish(s);
    return MPMW_OK;
}

// Recompute the AABB on the device, read it back, rebuild the grid and
// re-bin the particles when the tiles no longer match it.  rebuild_launch
// enqueues the reduction (unless the last pass left fresh statistics) and
// the copies to the host; after the stream has been synchronised,
// rebuild_finish rebuilds the grid and, when the binning must change,
// enqueues it (sets `sorting`; sort_finish after the next synchronisation).
template <typename R> int rebuild_launch(mpmw_solver *s) {
    int rc;
    if (!s->stats_fresh) {
        if ((rc = reduce_particles<R>(s))) return rc;
        s->stats_fresh = true;
    }
    SortState &S = s->sort;
    if (S.oow) HIP_TRY(hipMemcpyAsync(S.h_oow, S.oow, sizeof(int32_t), hipMemcpyDeviceToHost, s->stream));
    return MPMW_OK;
}

template <typename R> int rebuild_finish(mpmw_solver *s, bool &sorting) {
    int rc;
    sorting = false;
    SortState &S = s->sort;
    if ((rc = rebuild_grid_from_stats(s))) return rc;
    bool need = !S.valid;
    if (S.valid) {
        const TileGeom tg = tile_geom(s);
        need = tg.ntx != S.geom.ntx || tg.nty != S.geom.nty || tg.ntz != S.geom.ntz || tg.ox != S.geom.ox ||
               tg.oy != S.geom.oy || tg.oz != S.geom.oz;
        const double o[3] = {s->minx, s->miny, s->minz};
        // the deterministic P2G needs every particle in its binned cell: any
        // origin change re-bins (else the first P2G would be repeated)
        for (int a = 0; a < 3; ++a) {
            const double d = std::fabs(o[a] - S.origin[a]);
            need = need || (s->deterministic ? d > 0.0 : d >= 0.5 * s->h);
        }
        // particles that left their tile window since the last binning
        need = need || (double)*S.h_oow > 0.01 * (double)s->n * (double)std::max<int64_t>(1, S.steps_since);
    }
    if (S.oow) hipLaunchKernelGGL(k_fill<int32_t>, dim3(1), dim3(64), 0, s->stream, S.oow, (int64_t)1, 0);
    S.steps_since = 0;
    if (need) {
        if ((rc = sort_launch<R>(s))) return rc;
        sorting = true;
    }
    return MPMW_OK;
}

template <typename R> int compact_impl(mpmw_solver *s);

// ------------------------------------------------------------- the step ---

// Blocks of the loose-particle pass (0 when there are none).
int loose_blocks(const mpmw_solver *s) {
    const int nl = s->sort.loose_end - s->sort.loose_begin;
    return nl > 0 ? blocks_for(nl) : 0;
}

// One particle pass: the tiles, then the loose particles.  The backstress
// arrays select the kernel variant; the deterministic flag the P2G form.
// Statistics partials: tiles first, then the loose blocks.
// The loose particles of one pass (after the tile pass; statistics
// partials after the tiles').
template <typename R, bool DO_G2P, bool DO_P2G, bool HAS_BACK, bool DET>
int loose_pass_t(mpmw_solver *s, double dt, double zcut, G2PFlags fl) {
    const int nbl = loose_blocks(s);
    if (nbl <= 0) return MPMW_OK;
    const Particles<R> P = pview<R>(s);
    const GridView<R> g = gview<R>(s);
    const Grid<R> G = grid_arrays<R>(s);
    const Material<R> M = material<R>(s);
    SortState &S = s->sort;
    if (s->apic) {
        hipLaunchKernelGGL((k_apic_loose<R, DO_G2P, DO_P2G, HAS_BACK>), dim3(nbl), dim3(kBlock), 0, s->stream, P, g,
                           G, M, (R)dt, dt, zcut, fl, s->prm.ref_temp, s->part + (int64_t)S.geom.ntiles * kRed,
                           S.loose_begin, S.loose_end, S.oow + 2);
        HIP_TRY(hipGetLastError());
        return MPMW_OK;
    }
    hipLaunchKernelGGL((k_loose<R, DO_G2P, DO_P2G && !DET, HAS_BACK>), dim3(nbl), dim3(kBlock), 0, s->stream, P, g,
                       G, M, (R)dt, dt, zcut, fl, s->prm.ref_temp, s->part + (int64_t)S.geom.ntiles * kRed,
                       S.loose_begin, S.loose_end, S.oow + 2);
    if (DET && DO_P2G)
        hipLaunchKernelGGL(k_loose_p2g_ordered<R>, dim3(1), dim3(kWave), 0, s->stream, P, g, G, S.loose_begin,
                           S.loose_end, S.oow + 2);
    HIP_TRY(hipGetLastError());
    return MPMW_OK;
}

// The deterministic solver's standalone P2G pass (its re-binning check
// repeats it): the tile pass, then the loose particles in a fixed order.
// Counters [1] (particles off their cell) and [2] (node accumulators
// written) start at zero.  The standalone P2G of a step's begin otherwise
// runs batched (step_begin_many).
template <typename R, bool HAS_BACK> int det_p2g_pass(mpmw_solver *s, double dt, double zcut) {
    const TileGeom tg = s->sort.geom;
    G2PFlags fl;
    fl.chip = 0;
    fl.stats = 0;
    hipLaunchKernelGGL(k_fill<int32_t>, dim3(1), dim3(64), 0, s->stream, s->sort.oow + 1, (int64_t)2, 0);
    hipLaunchKernelGGL((k_tile<R, false, true, HAS_BACK, true>), dim3(tg.ntiles), dim3(kBlock), 0, s->stream,
                       pview<R>(s), gview<R>(s), grid_arrays<R>(s), material<R>(s), (R)dt, dt, zcut, fl,
                       s->prm.ref_temp, s->part, s->sort.cell_start, tg, s->sort.oow, (R *)s->tpart);
    HIP_TRY(hipGetLastError());
    return loose_pass_t<R, false, true, HAS_BACK, true>(s, dt, zcut, fl);
}

// P2G alone into empty sums for a deterministic solver (first step, after a
// rebuild or a re-binning).  The node accumulators are cleared: a repeated
// P2G may follow a pass that wrote them without a grid update in between.
template <typename R> int p2g_alone(mpmw_solver *s, double dt, double zcut) {
    if (!s->deterministic) return fail(MPMW_E_STATE, "p2g_alone: deterministic solvers only");
    int rc;
    if ((rc = zero_acc<R>(s))) return rc;
    Timed t(s, MPMW_K_P2G);
    return s->has_back ? det_p2g_pass<R, true>(s, dt, zcut) : det_p2g_pass<R, false>(s, dt, zcut);
}

// Deterministic mode: when a particle of the last P2G pass had left its
// binned cell, re-bin the particles and repeat the P2G.  After a binning
// every particle is in its cell or loose, so the repeat finds none.
template <typename R> int det_settle(mpmw_solver *s, double dt, double zcut) {
    SortState &S = s->sort;
    HIP_TRY(hipMemcpyAsync(S.h_oow + 1, S.oow + 1, sizeof(int32_t), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (S.h_oow[1] == 0) return MPMW_OK;
    int rc;
    if ((rc = sort_particles<R>(s))) return rc;
    S.resorts++;
    if ((rc = p2g_alone<R>(s, dt, zcut))) return rc;
    HIP_TRY(hipMemcpyAsync(S.h_oow + 1, S.oow + 1, sizeof(int32_t), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (S.h_oow[1] != 0)
        return fail(MPMW_E_STATE, "deterministic P2G: " + std::to_string(S.h_oow[1]) +
                                      " particles outside their cell right after binning");
    return MPMW_OK;
}

// Host plan of one step (MPMWorkpieceSolver::step, MPMWorkpieceSolver.cu:
// 727-791), shared by the single-solver step and the batched step: what
// step_begin_many prepared and what the launches and step_end need.
struct StepPlan {
    bool run = false;         // false: nothing to step (:728 silent return)
    bool diffusion = false;
    bool next_rebuild = false;
    double alpha = 0.0, zcut = 0.0;
    G2PFlags fl{};
};

template <typename R> TileParts<R> tile_parts(const mpmw_solver *s) {
    TileParts<R> tp;
    tp.part = (const R *)s->tpart;
    tp.acc_flag = s->sort.oow + 2;
    tp.tg = s->sort.geom;
    tp.quad = s->apic ? 1 : 0;
    return tp;
}

// Everything after the particle pass: statistics, the kinematic tool, time
// and step count, the deterministic check of the next step's P2G, and the
// compaction every 1000 steps.
template <typename R> int step_end(mpmw_solver *s, double dt, const StepPlan &sp) {
    int rc;
    s->cf_pending = true;
    s->snap_valid = s->snap_on;
    if (sp.fl.stats) {
        Timed t(s, MPMW_K_REDUCE);
        if ((rc = finalize_stats(s, s->sort.geom.ntiles + loose_blocks(s)))) return rc;
        s->stats_fresh = true;
        s->metrics_pending = true;
    } else {
        s->stats_fresh = false;
    }
    s->p2g_ready = !sp.next_rebuild;
    s->sort.steps_since++;
    // kinematic tool
    for (int a = 0; a < 3; ++a) s->pose.center[a] += s->pose.vel[a] * dt;
    s->pose.angle += s->pose.omega * dt;
    s->time += dt;
    ++s->step;
    // the next step's P2G: re-bin and repeat it if a particle changed cell
    if (s->deterministic && s->p2g_ready && (rc = det_settle<R>(s, dt, sp.zcut))) return rc;
    if (s->step % 1000 == 0) return compact_impl<R>(s);  // :788-790
    return MPMW_OK;
}

template <typename R> int step_many(mpmw_solver *const *members, size_t nm, const double *dts,
                                   hipStream_t stream, StepStaging &st);
void free_staging(StepStaging &st);
void drop_graphs(StepStaging &st);
int set_autotune(StepStaging &st, int32_t trials);
int set_launch_config(StepStaging &st, const mpmw_tuning *c);
void get_tuning(const StepStaging &st, mpmw_tuning *o);

// MPMWorkpieceSolver::step, MPMWorkpieceSolver.cu:727-791: a step of a batch
// of one, so a solver runs the same kernels alone and in a batch.
template <typename R> int step_impl(mpmw_solver *s, double dt) {
    return step_many<R>(&s, 1, &dt, s->stream, s->staging);
}

// Order-preserving removal of INACTIVE particles (compactParticles :475-527).
template <typename R> int compact_impl(mpmw_solver *s) {
    if (s->n <= 0) return MPMW_OK;
    std::vector<int32_t> st(s->n);
    HIP_TRY(hipMemcpyAsync(st.data(), s->f[F_STATUS], sizeof(int32_t) * s->n, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    std::vector<int32_t> keep;
    keep.reserve(s->n);
    for (int32_t i = 0; i < s->n; ++i)
        if (st[i] != MPMW_INACTIVE) keep.push_back(i);
    if (keep.empty() || (int32_t)keep.size() == s->n) return MPMW_OK;
    return permute_particles<R>(s, keep.data(), (int32_t)keep.size(), true);
}

}  // namespace

// Permute all live particle fields: new[i] = old[perm[i]] for i < count.
// perm is a host array when host_perm, else a device array.
template <typename R> int permute_particles(mpmw_solver *s, const int32_t *perm, int32_t count, bool host_perm) {
    int rc = ensure_alt(s);
    if (rc) return rc;
    const int32_t *dperm = perm;
    void *tmp = nullptr;
    if (host_perm) {
        if ((rc = dalloc(s, &tmp, sizeof(int32_t) * (size_t)count))) return rc;
        HIP_TRY(hipMemcpyAsync(tmp, perm, sizeof(int32_t) * count, hipMemcpyHostToDevice, s->stream));
        dperm = (const int32_t *)tmp;
    }
    {
        // every live array in one launch
        static_assert(NF <= kMaxGatherFields, "gather table too small");
        GatherTable gt{};
        for (int k = 0; k < NF; ++k) {
            if (!field_live(s, k)) continue;
            gt.dst[gt.nf] = s->alt[k];
            gt.src[gt.nf] = s->f[k];
            gt.wide[gt.nf] = (k < kNumReal && (is_pos(k) || sizeof(R) == 8)) ? 1 : 0;
            ++gt.nf;
        }
        Timed t(s, MPMW_K_SORT);
        if (gt.nf > 0 && count > 0)
            hipLaunchKernelGGL(k_gather_fields, dim3(blocks_for(count), gt.nf), dim3(kBlock), 0, s->stream, gt, dperm,
                               count);
    }
    hipError_t e = hipGetLastError();
    if (host_perm) {
        (void)hipStreamSynchronize(s->stream);
        dfree(s, tmp, sizeof(int32_t) * (size_t)count);
    }
    if (e != hipSuccess) return fail(MPMW_E_HIP, std::string("permute: ") + hipGetErrorString(e));
    for (int k = 0; k < NF; ++k) std::swap(s->f[k], s->alt[k]);
    s->n = count;
    s->p2g_ready = false;
    s->sort.valid = false;   // a binning sets it again after its permutation
    return MPMW_OK;
}

// ================================================================ C ABI ====

namespace {

int check(mpmw_solver *s) {
    if (!s) return fail(MPMW_E_INVAL, "null solver handle");
    hipError_t e = hipSetDevice(s->prm.device);
    if (e != hipSuccess) return fail(MPMW_E_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
    return MPMW_OK;
}

#define CHECK_S(s)              \
    do {                        \
        int rc_ = check(s);     \
        if (rc_) return rc_;    \
    } while (0)

#define DISPATCH(s, fn, ...) ((s)->fp64 ? fn<double>(__VA_ARGS__) : fn<float>(__VA_ARGS__))

bool finite_array(const double *a, int64_t n) {
    for (int64_t i = 0; i < n; ++i)
        if (!std::isfinite(a[i])) return false;
    return true;
}

template <typename R> int upload_field(mpmw_solver *s, int fld, const double *src, int64_t n, int comp, int c,
                                       double fill) {
    std::vector<R> tmp((size_t)n);
    for (int64_t i = 0; i < n; ++i) tmp[(size_t)i] = (R)(src ? src[i * comp + c] : fill);
    HIP_TRY(hipMemcpy(s->f[fld], tmp.data(), sizeof(R) * (size_t)n, hipMemcpyHostToDevice));
    return MPMW_OK;
}

template <typename R> int download_field(mpmw_solver *s, int fld, double *dst, int64_t n, int comp, int c) {
    std::vector<R> tmp((size_t)n);
    HIP_TRY(hipMemcpy(tmp.data(), s->f[fld], sizeof(R) * (size_t)n, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < n; ++i) dst[i * comp + c] = (double)tmp[(size_t)i];
    return MPMW_OK;
}

template <typename R> int set_particles_impl(mpmw_solver *s, const mpmw_particle_arrays *a) {
    const int32_t n = a->n;
    bool back = s->prm.kinematic_hardening > 0.0;
    if (a->backstress)
        for (int64_t i = 0; i < 6 * (int64_t)n && !back; ++i) back = a->backstress[i] != 0.0;
    bool fext = false;
    if (a->fext)
        for (int64_t i = 0; i < 3 * (int64_t)n && !fext; ++i) fext = a->fext[i] != 0.0;
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (s->metrics_pending) {
        absorb_stats(s);
        s->metrics_pending = false;
    }
    s->has_back = back;
    s->has_fext = fext;
    s->n = 0;
    int rc = alloc_particles(s, n);
    if (rc) return rc;
    s->n = n;
    if (n == 0) return MPMW_OK;
    struct { int fld; const double *src; int comp, c; double fill; } map[] = {
        {F_X, a->pos, 3, 0, 0}, {F_Y, a->pos, 3, 1, 0}, {F_Z, a->pos, 3, 2, 0},
        {F_VX, a->vel, 3, 0, 0}, {F_VY, a->vel, 3, 1, 0}, {F_VZ, a->vel, 3, 2, 0},
        {F_S0, a->stress, 6, 0, 0}, {F_S1, a->stress, 6, 1, 0}, {F_S2, a->stress, 6, 2, 0},
        {F_S3, a->stress, 6, 3, 0}, {F_S4, a->stress, 6, 4, 0}, {F_S5, a->stress, 6, 5, 0},
        {F_EPS, a->plastic_strain, 1, 0, 0}, {F_RATE, a->strain_rate, 1, 0, 0},
        {F_DMG, a->damage, 1, 0, 0}, {F_SRES, a->residual_stress, 1, 0, 0},
        {F_T, a->temperature, 1, 0, 0}, {F_M, a->mass, 1, 0, 0}, {F_V, a->volume, 1, 0, 0},
        {F_DENS, a->density, 1, 0, s->prm.density},
        {F_A0, a->backstress, 6, 0, 0}, {F_A1, a->backstress, 6, 1, 0}, {F_A2, a->backstress, 6, 2, 0},
        {F_A3, a->backstress, 6, 3, 0}, {F_A4, a->backstress, 6, 4, 0}, {F_A5, a->backstress, 6, 5, 0},
        {F_FX, a->fext, 3, 0, 0}, {F_FY, a->fext, 3, 1, 0}, {F_FZ, a->fext, 3, 2, 0},
        // new particles start with a zero affine matrix
        {F_C0, nullptr, 1, 0, 0}, {F_C1, nullptr, 1, 0, 0}, {F_C2, nullptr, 1, 0, 0},
        {F_C3, nullptr, 1, 0, 0}, {F_C4, nullptr, 1, 0, 0}, {F_C5, nullptr, 1, 0, 0},
        {F_C6, nullptr, 1, 0, 0}, {F_C7, nullptr, 1, 0, 0}, {F_C8, nullptr, 1, 0, 0},
    };
    for (auto &e : map) {
        if (!field_live(s, e.fld)) continue;
        rc = is_pos(e.fld) ? upload_field<double>(s, e.fld, e.src, n, e.comp, e.c, e.fill)
                           : upload_field<R>(s, e.fld, e.src, n, e.comp, e.c, e.fill);
        if (rc) return rc;
    }
    s->uniform_mv = true;
    s->m0 = a->mass[0];
    s->V0 = a->volume[0];
    for (int32_t i = 1; i < n && s->uniform_mv; ++i)
        s->uniform_mv = a->mass[i] == s->m0 && a->volume[i] == s->V0;
    HIP_TRY(hipMemcpy(s->f[F_STATUS], a->status, sizeof(int32_t) * n, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(s->f[F_ID], a->id, sizeof(int32_t) * n, hipMemcpyHostToDevice));
    s->p2g_ready = false;
    s->stats_fresh = false;
    s->sort.valid = false;
    // as initializeParticleBox does, size the grid for the new particles now
    rc = reduce_particles<R>(s);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->stats_fresh = true;
    {
        const double *v = s->h_stats;
        s->bounds[0] = v[0]; s->bounds[1] = v[1]; s->bounds[2] = v[2];
        s->bounds[3] = -v[3]; s->bounds[4] = -v[4]; s->bounds[5] = -v[5];
    }
    return rebuild_grid_from_stats(s);
}

template <typename R> int get_particles_impl(mpmw_solver *s, mpmw_particle_arrays *a) {
    HIP_TRY(hipStreamSynchronize(s->stream));
    const int32_t n = std::min(a->n, s->n);
    int rc;
    struct { int fld; double *dst; int comp, c; } map[] = {
        {F_X, a->pos, 3, 0}, {F_Y, a->pos, 3, 1}, {F_Z, a->pos, 3, 2},
        {F_VX, a->vel, 3, 0}, {F_VY, a->vel, 3, 1}, {F_VZ, a->vel, 3, 2},
        {F_S0, a->stress, 6, 0}, {F_S1, a->stress, 6, 1}, {F_S2, a->stress, 6, 2},
        {F_S3, a->stress, 6, 3}, {F_S4, a->stress, 6, 4}, {F_S5, a->stress, 6, 5},
        {F_EPS, a->plastic_strain, 1, 0}, {F_RATE, a->strain_rate, 1, 0},
        {F_DMG, a->damage, 1, 0}, {F_SRES, a->residual_stress, 1, 0},
        {F_T, a->temperature, 1, 0}, {F_M, a->mass, 1, 0}, {F_V, a->volume, 1, 0},
        {F_DENS, a->density, 1, 0},
        {F_A0, a->backstress, 6, 0}, {F_A1, a->backstress, 6, 1}, {F_A2, a->backstress, 6, 2},
        {F_A3, a->backstress, 6, 3}, {F_A4, a->backstress, 6, 4}, {F_A5, a->backstress, 6, 5},
        {F_FX, a->fext, 3, 0}, {F_FY, a->fext, 3, 1}, {F_FZ, a->fext, 3, 2},
    };
    for (auto &e : map) {
        if (!e.dst || n == 0) continue;
        if (!field_live(s, e.fld)) {
            for (int64_t i = 0; i < n; ++i) e.dst[i * e.comp + e.c] = 0.0;
            continue;
        }
        rc = is_pos(e.fld) ? download_field<double>(s, e.fld, e.dst, n, e.comp, e.c)
                           : download_field<R>(s, e.fld, e.dst, n, e.comp, e.c);
        if (rc) return rc;
    }
    if (n > 0 && a->status) HIP_TRY(hipMemcpy(a->status, s->f[F_STATUS], sizeof(int32_t) * n, hipMemcpyDeviceToHost));
    if (n > 0 && a->id) HIP_TRY(hipMemcpy(a->id, s->f[F_ID], sizeof(int32_t) * n, hipMemcpyDeviceToHost));
    a->n = n;
    return MPMW_OK;
}

// initializeParticleBox / initializeCylindricalWorkpiece lattice, host side.
struct Lattice {
    std::vector<double> pos, mass, vol, dens, T;
    std::vector<int32_t> id, status;
    void push(const mpmw_params &p, double x, double y, double z, double v, int32_t pid, bool fixed) {
        pos.push_back(x); pos.push_back(y); pos.push_back(z);
        mass.push_back(p.density * v);
        vol.push_back(v);
        dens.push_back(p.density);
        T.push_back(p.ref_temp);
        id.push_back(pid);
        status.push_back(fixed ? MPMW_FIXED_BOUNDARY : MPMW_ACTIVE);
    }
    int upload(mpmw_solver *s) {
        const size_t n = mass.size();
        std::vector<double> zero3(3 * n, 0.0), zero1(n, 0.0), zero6(6 * n, 0.0);
        mpmw_particle_arrays a{};
        a.n = (int32_t)n;
        a.pos = pos.data(); a.vel = zero3.data(); a.fext = nullptr;
        a.mass = mass.data(); a.volume = vol.data(); a.density = dens.data(); a.temperature = T.data();
        a.plastic_strain = zero1.data(); a.strain_rate = zero1.data(); a.damage = zero1.data();
        a.residual_stress = zero1.data(); a.stress = zero6.data(); a.backstress = nullptr;
        a.id = id.data(); a.status = status.data();
        return mpmw_set_particles(s, &a);
    }
};


}  // namespace

namespace {
template <typename R> int nearest_impl(mpmw_solver *s, double x, double y, double z, int apply, double q) {
    const int nb = std::min(blocks_for(s->n), 4096);
    hipLaunchKernelGGL(k_nearest<R>, dim3(nb), dim3(kBlock), 0, s->stream, pview<R>(s), x, y, z, s->d_near,
                       s->d_near_i);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_nearest_finalize<R>, dim3(1), dim3(1), 0, s->stream, pview<R>(s), s->d_near, s->d_near_i,
                       nb, s->d_near_out, s->d_near_out_i, apply, q, s->prm.specific_heat, s->prm.melt_temp,
                       s->h * s->h);
    HIP_TRY(hipGetLastError());
    return MPMW_OK;
}
}  // namespace

namespace {

// The slab of mpmw_gen_jittered_slab generated on the device (k_init_slab),
// then what set_particles does after an upload.
template <typename R> int init_slab_impl(mpmw_solver *s, uint64_t seed, int32_t cx, int32_t cy, int32_t cz,
                                         const double origin[3], const double vel[3], double temperature) {
    const int64_t n = (int64_t)cx * cy * cz * 8;
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (s->metrics_pending) {
        absorb_stats(s);
        s->metrics_pending = false;
    }
    s->has_back = s->prm.kinematic_hardening > 0.0;
    s->has_fext = false;
    s->n = 0;
    int rc = alloc_particles(s, (int32_t)n);
    if (rc) return rc;
    s->n = (int32_t)n;
    const double h = s->prm.cell_size;
    const double vol = h * h * h / 8.0;
    s->uniform_mv = true;
    s->m0 = s->prm.density * vol;
    s->V0 = vol;
    const double fixture = s->prm.has_config ? s->prm.fixture_thickness : 0.0;
    hipLaunchKernelGGL(k_init_slab<R>, dim3(blocks_for(n)), dim3(kBlock), 0, s->stream, pview<R>(s),
                       (R *)s->f[F_DENS], seed, cx, cy, origin[0], origin[1], origin[2], h, origin[2] + fixture,
                       (R)vel[0], (R)vel[1], (R)vel[2], (R)temperature, (R)s->prm.density, (int32_t)n);
    HIP_TRY(hipGetLastError());
    // the per-particle mass and volume arrays (kept for export)
    hipLaunchKernelGGL(k_fill<R>, dim3(std::min(blocks_for(n), 4096)), dim3(kBlock), 0, s->stream, (R *)s->f[F_M],
                       n, (R)s->m0);
    hipLaunchKernelGGL(k_fill<R>, dim3(std::min(blocks_for(n), 4096)), dim3(kBlock), 0, s->stream, (R *)s->f[F_V],
                       n, (R)s->V0);
    if (s->apic)
        for (int c = F_C0; c <= F_C8; ++c)
            hipLaunchKernelGGL(k_fill<R>, dim3(std::min(blocks_for(n), 4096)), dim3(kBlock), 0, s->stream,
                               (R *)s->f[c], n, R(0));
    HIP_TRY(hipGetLastError());
    s->p2g_ready = false;
    s->stats_fresh = false;
    s->sort.valid = false;
    if ((rc = reduce_particles<R>(s))) return rc;
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->stats_fresh = true;
    {
        const double *v = s->h_stats;
        s->bounds[0] = v[0]; s->bounds[1] = v[1]; s->bounds[2] = v[2];
        s->bounds[3] = -v[3]; s->bounds[4] = -v[4]; s->bounds[5] = -v[5];
    }
    return rebuild_grid_from_stats(s);
}

}  // namespace

extern "C" {

const char *mpmw_last_error(void) { return g_err.c_str(); }
int mpmw_abi_version(void) { return MPMW_ABI_VERSION; }

void mpmw_params_default(mpmw_params *p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    // Config.h defaults (MaterialProperties :23-45, SimulationParams :74-80,
    // MPMParams :128-157, PhysicsSafety :343-346, MachineSetupParams :366-391),
    // mapped as MPMWorkpieceSolver::initialize does (:428-451); k, c_damp and H
    // keep MPMMaterialParams' own defaults (MPMWorkpieceSolver.cuh:85-88).
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
    p->contact_damping = 0.15;
    p->max_contact_temp_rise = 50.0;
    p->thermal_conductivity = 42.6;
    p->kinematic_hardening = 0.0;
    p->thermal_clamp_enabled = 1;
    p->has_config = 1;
    p->particle_spacing_factor = 0.8;
    p->fixture_thickness = 0.002;
    p->min_time_step = 1e-12;
    p->max_time_step = 1e-4;
    p->max_particles = 100000;
    p->precision = MPMW_FP32;
    p->deterministic = 0;
    p->device = 0;
}

int mpmw_create(const mpmw_params *p, mpmw_solver **out) {
    if (!p || !out) return fail(MPMW_E_INVAL, "mpmw_create: null argument");
    *out = nullptr;
    if (!(p->cell_size > 0.0) || !std::isfinite(p->cell_size))
        return fail(MPMW_E_INVAL, "cell_size must be > 0 (smoothing_radius_m)");
    if (p->precision != MPMW_FP32 && p->precision != MPMW_FP64)
        return fail(MPMW_E_INVAL, "precision must be MPMW_FP32 or MPMW_FP64");
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev <= 0)
        return fail(MPMW_E_HIP, std::string("no HIP device available: ") + hipGetErrorString(e));
    if (p->device < 0 || p->device >= ndev) return fail(MPMW_E_INVAL, "device ordinal out of range");
    HIP_TRY(hipSetDevice(p->device));
    auto *s = new mpmw_solver();
    s->prm = *p;
    s->fp64 = p->precision == MPMW_FP64;
    s->rsz = s->fp64 ? sizeof(double) : sizeof(float);
    s->h = p->cell_size;
    s->deterministic = p->deterministic != 0;
    s->max_temp = p->ref_temp;
    int rc = MPMW_OK;
    auto bail = [&](int code) { mpmw_destroy(s); return code; };
    if (pool::stream(p->device, &s->stream) != hipSuccess) return bail(fail(MPMW_E_HIP, "hipStreamCreate failed"));
    void *q;
    if ((rc = dalloc(s, &q, sizeof(double) * kRed))) return bail(rc);
    s->d_stats = (double *)q;
    if ((rc = dalloc(s, &q, sizeof(double)))) return bail(rc);
    s->d_cf = (double *)q;
    if ((rc = dalloc(s, &q, sizeof(int32_t) * (kRed + 1)))) return bail(rc);
    s->d_modes = (int32_t *)q;
    int32_t modes[kRed + 1] = {0, 0, 0, 0, 0, 0, 1, 1, 2, 2};
    if (hipMemcpy(s->d_modes, modes, sizeof(modes), hipMemcpyHostToDevice) != hipSuccess)
        return bail(fail(MPMW_E_HIP, "hipMemcpy modes failed"));
    if (pool::get(p->device, true, sizeof(double) * kRed, (void **)&s->h_stats) != hipSuccess ||
        pool::get(p->device, true, sizeof(double), (void **)&s->h_cf) != hipSuccess)
        return bail(fail(MPMW_E_HIP, "hipHostMalloc failed"));
    std::memset(s->h_stats, 0, sizeof(double) * kRed);
    *s->h_cf = 0.0;
    if ((rc = dalloc(s, &q, sizeof(double) * 4096))) return bail(rc);
    s->d_near = (double *)q;
    if ((rc = dalloc(s, &q, sizeof(int32_t) * 4096))) return bail(rc);
    s->d_near_i = (int32_t *)q;
    if ((rc = dalloc(s, &q, sizeof(double)))) return bail(rc);
    s->d_near_out = (double *)q;
    if ((rc = dalloc(s, &q, sizeof(int32_t)))) return bail(rc);
    s->d_near_out_i = (int32_t *)q;
    *out = s;
    return MPMW_OK;
}

int mpmw_destroy(mpmw_solver *s) {
    if (!s) return MPMW_OK;
    (void)hipSetDevice(s->prm.device);
    if (s->batch) (void)mpmw_batch_remove(s->batch, s);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (auto &e : s->events) { (void)hipEventDestroy(e.a); (void)hipEventDestroy(e.b); }
    for (auto e : s->ev_pool) (void)hipEventDestroy(e);
    free_particles(s);
    dfree(s, s->acc, 0); dfree(s, s->node, 0); dfree(s, s->gm, 0); dfree(s, s->gtm, 0); dfree(s, s->snap, 0);
    void *q;
    q = s->part; dfree(s, q, 0);
    q = s->fpart; dfree(s, q, 0);
    q = s->d_stats; dfree(s, q, 0);
    q = s->d_cf; dfree(s, q, 0);
    q = s->d_modes; dfree(s, q, 0);
    q = s->d_near; dfree(s, q, 0);
    q = s->d_near_i; dfree(s, q, 0);
    q = s->d_near_out; dfree(s, q, 0);
    q = s->d_near_out_i; dfree(s, q, 0);
    dfree(s, s->samples, 0);
    dfree(s, s->sdf, 0);
    dfree(s, s->tpart, 0);
    free_sort_state(s);
    free_staging(s->staging);
    pool::put(s->h_stats);
    pool::put(s->h_cf);
    delete s;
    return MPMW_OK;
}

int mpmw_reset(mpmw_solver *s) {
    CHECK_S(s);
    HIP_TRY(hipStreamSynchronize(s->stream));
    // reset(), :830-841: host state only; grid parameters are kept
    s->n = 0;
    s->tool_kind = MPMW_TOOL_NONE;
    s->nsamples = 0;
    s->time = 0.0;
    s->step = 0;
    s->max_stress = 0.0;
    s->max_temp = s->prm.ref_temp;
    s->ke = 0.0;
    s->tip[0] = s->tip[1] = s->tip[2] = 0.0;
    // the kinematic tool pose advances with every step: back to where it was set
    s->pose = s->pose0;
    s->p2g_ready = false;
    s->stats_fresh = false;
    s->metrics_pending = false;
    s->sort.valid = false;
    return MPMW_OK;
}

int mpmw_set_particles(mpmw_solver *s, const mpmw_particle_arrays *a) {
    CHECK_S(s);
    if (!a || a->n < 0) return fail(MPMW_E_INVAL, "set_particles: bad arrays");
    const int64_t n = a->n;
    if (n > 0 && (!a->pos || !a->vel || !a->mass || !a->volume || !a->temperature || !a->plastic_strain ||
                  !a->strain_rate || !a->damage || !a->residual_stress || !a->stress || !a->id || !a->status))
        return fail(MPMW_E_INVAL, "set_particles: a required array is NULL");
    struct { const double *p; int64_t k; const char *name; } chk[] = {
        {a->pos, 3 * n, "pos"}, {a->vel, 3 * n, "vel"}, {a->mass, n, "mass"}, {a->volume, n, "volume"},
        {a->temperature, n, "temperature"}, {a->plastic_strain, n, "plastic_strain"},
        {a->strain_rate, n, "strain_rate"}, {a->damage, n, "damage"},
        {a->residual_stress, n, "residual_stress"}, {a->stress, 6 * n, "stress"},
        {a->fext, 3 * n, "fext"}, {a->backstress, 6 * n, "backstress"}, {a->density, n, "density"}};
    for (auto &c : chk)
        if (c.p && !finite_array(c.p, c.k)) return fail(MPMW_E_INVAL, std::string("set_particles: non-finite ") + c.name);
    for (int64_t i = 0; i < n; ++i)
        if (a->status[i] < MPMW_ACTIVE || a->status[i] > MPMW_FIXED_BOUNDARY)
            return fail(MPMW_E_INVAL, "set_particles: status out of range");
    return DISPATCH(s, set_particles_impl, s, a);
}

int mpmw_get_particles(mpmw_solver *s, mpmw_particle_arrays *a) {
    CHECK_S(s);
    if (!a || a->n < 0) return fail(MPMW_E_INVAL, "get_particles: bad arrays");
    return DISPATCH(s, get_particles_impl, s, a);
}

int mpmw_particle_count(mpmw_solver *s, int32_t *n) {
    CHECK_S(s);
    if (!n) return fail(MPMW_E_INVAL, "null");
    *n = s->n;
    return MPMW_OK;
}

int mpmw_init_particle_box(mpmw_solver *s, const double mn[3], const double mx[3], double spacing) {
    CHECK_S(s);
    if (!mn || !mx || !(spacing > 0.0)) return fail(MPMW_E_INVAL, "init_particle_box: bad arguments");
    const mpmw_params &p = s->prm;
    const int maxp = p.has_config ? p.max_particles : 100000;
    const double vol = spacing * spacing * spacing;
    Lattice L;
    int32_t id = 0;
    for (double z = mx[2]; z >= mn[2] - 1e-12 && (int)L.mass.size() < maxp; z -= spacing)
        for (double y = mn[1]; y <= mx[1] + 1e-12 && (int)L.mass.size() < maxp; y += spacing)
            for (double x = mn[0]; x <= mx[0] + 1e-12 && (int)L.mass.size() < maxp; x += spacing)
                L.push(p, x, y, z, vol, id++, p.has_config && z <= mn[2] + p.fixture_thickness);
    return L.upload(s);
}

int mpmw_init_cylinder(mpmw_solver *s, const double c[3], double radius, double length, double spacing,
                       int32_t axis) {
    CHECK_S(s);
    (void)axis;  // the reference ignores it too (:618)
    if (!c || !(spacing > 0.0)) return fail(MPMW_E_INVAL, "init_cylinder: bad arguments");
    const mpmw_params &p = s->prm;
    const int maxp = p.has_config ? p.max_particles : 100000;
    const double vol = spacing * spacing * spacing;
    const double minZ = c[2] - length * 0.5, maxZ = c[2] + length * 0.5;
    Lattice L;
    int32_t id = 0;
    for (double z = maxZ; z >= minZ - 1e-12 && (int)L.mass.size() < maxp; z -= spacing)
        for (double y = c[1] - radius; y <= c[1] + radius + 1e-12 && (int)L.mass.size() < maxp; y += spacing)
            for (double x = c[0] - radius; x <= c[0] + radius + 1e-12 && (int)L.mass.size() < maxp; x += spacing) {
                double dx = x - c[0], dy = y - c[1];
                if (dx * dx + dy * dy > radius * radius) continue;
                L.push(p, x, y, z, vol, id++, p.has_config && z <= minZ + p.fixture_thickness);
            }
    return L.upload(s);
}

int mpmw_gen_jittered_slab(uint64_t seed, int32_t cx, int32_t cy, int32_t cz, double h, const double origin[3],
                           double density, double fixture, double *pos, double *mass, double *volume, int32_t *id,
                           int32_t *status) {
#pragma clang fp contract(off)
    if (cx <= 0 || cy <= 0 || cz <= 0 || !(h > 0) || !origin || !pos || !mass || !volume || !id || !status)
        return fail(MPMW_E_INVAL, "gen_jittered_slab: bad arguments");
    const int64_t ncell = (int64_t)cx * cy * cz;
    if (ncell * 8 > INT32_MAX) return fail(MPMW_E_RANGE, "gen_jittered_slab: too many particles");
    const double vol = h * h * h / 8.0;
    for (int64_t c = 0; c < ncell; ++c) {
        const int64_t cell[3] = {c % cx, (c / cx) % cy, c / ((int64_t)cx * cy)};
        for (int k = 0; k < 8; ++k) {
            const int64_t pi = c * 8 + k;
            const int sub[3] = {k & 1, (k >> 1) & 1, (k >> 2) & 1};
            for (int ax = 0; ax < 3; ++ax) {
                const uint64_t ctr = (uint64_t)(pi * 3 + ax);
                const double u = (double)(splitmix64(seed ^ (ctr * 0xD1B54A32D192ED03ull)) >> 11) * 0x1.0p-53;
                const double t = ((double)cell[ax] + 0.5 * sub[ax]) + 0.5 * u;
                const double x = origin[ax] + t * h;  // two roundings, as k_init_slab
                pos[3 * pi + ax] = (double)(float)x;
            }
            mass[pi] = density * vol;
            volume[pi] = vol;
            id[pi] = (int32_t)pi;
            status[pi] = (pos[3 * pi + 2] <= origin[2] + fixture) ? MPMW_FIXED_BOUNDARY : MPMW_ACTIVE;
        }
    }
    return MPMW_OK;
}

int mpmw_init_jittered_slab(mpmw_solver *s, uint64_t seed, int32_t cx, int32_t cy, int32_t cz,
                            const double origin[3], const double vel[3], double temperature) {
    CHECK_S(s);
    if (!origin || !vel) return fail(MPMW_E_INVAL, "init_jittered_slab: null argument");
    const int64_t n = (int64_t)cx * cy * cz * 8;
    if (cx <= 0 || cy <= 0 || cz <= 0 || n > INT32_MAX) return fail(MPMW_E_RANGE, "init_jittered_slab: bad size");
    if (!finite_array(origin, 3) || !finite_array(vel, 3) || !std::isfinite(temperature))
        return fail(MPMW_E_INVAL, "init_jittered_slab: non-finite argument");
    return DISPATCH(s, init_slab_impl, s, seed, cx, cy, cz, origin, vel, temperature);
}

int mpmw_set_tool_samples(mpmw_solver *s, const double *xyz, int32_t n, double contact_radius) {
    CHECK_S(s);
    if (n < 0 || (n > 0 && !xyz)) return fail(MPMW_E_INVAL, "set_tool_samples: bad arguments");
    HIP_TRY(hipStreamSynchronize(s->stream));
    // setToolMesh, :793-810
    s->prm.tool_contact_radius = std::max(contact_radius, s->h);
    const int stride = std::max(1, n / 4000);
    std::vector<double> pts;
    for (int64_t i = 0; i < n; i += stride)
        for (int a = 0; a < 3; ++a) pts.push_back(xyz[3 * i + a]);
    const int32_t m = (int32_t)(pts.size() / 3);
    if (!finite_array(pts.data(), (int64_t)pts.size())) return fail(MPMW_E_INVAL, "set_tool_samples: non-finite");
    dfree(s, s->samples, 0);
    s->nsamples = 0;
    s->tool_kind = MPMW_TOOL_NONE;
    if (m == 0) return MPMW_OK;
    int rc = dalloc(s, &s->samples, 3 * sizeof(double) * (size_t)m);
    if (rc) return rc;
    const double rcr = s->prm.tool_contact_radius;
    for (int a = 0; a < 3; ++a) { s->sample_lo[a] = 1e300; s->sample_hi[a] = -1e300; }
    for (int32_t i = 0; i < m; ++i)
        for (int a = 0; a < 3; ++a) {
            s->sample_lo[a] = std::min(s->sample_lo[a], pts[3 * i + a]);
            s->sample_hi[a] = std::max(s->sample_hi[a], pts[3 * i + a]);
        }
    for (int a = 0; a < 3; ++a) {  // grown by r_c (plus a margin) for the node culling test
        const double marg = rcr * 1.001 + s->h * 1e-3;
        s->sample_lo[a] -= marg;
        s->sample_hi[a] += marg;
    }
    {
        std::vector<double> soa(3 * (size_t)m);
        for (int32_t i = 0; i < m; ++i)
            for (int a = 0; a < 3; ++a) soa[(size_t)a * m + i] = pts[3 * i + a];
        HIP_TRY(hipMemcpy(s->samples, soa.data(), sizeof(double) * soa.size(), hipMemcpyHostToDevice));
    }
    s->nsamples = m;
    s->tool_kind = MPMW_TOOL_POINTS;
    return MPMW_OK;
}

int mpmw_set_tool_analytic(mpmw_solver *s, const mpmw_tool_desc *d) {
    CHECK_S(s);
    if (!d) return fail(MPMW_E_INVAL, "null tool description");
    if (d->kind == MPMW_TOOL_NONE) {
        s->tool_kind = MPMW_TOOL_NONE;
        return MPMW_OK;
    }
    if (d->kind == MPMW_TOOL_VOXEL && !s->sdf)
        return fail(MPMW_E_STATE, "voxel tool without an SDF: build one with mpmw_set_tool_sdf_mesh");
    if (d->kind != MPMW_TOOL_WEDGE && d->kind != MPMW_TOOL_DRILL && d->kind != MPMW_TOOL_VOXEL)
        return fail(MPMW_E_INVAL, "tool kind must be NONE, WEDGE, DRILL or VOXEL here");
    if (!(d->grad_eps > 0.0)) return fail(MPMW_E_INVAL, "tool grad_eps must be > 0");
    if (d->kind == MPMW_TOOL_DRILL && !(d->radius > 0.0)) return fail(MPMW_E_INVAL, "drill radius must be > 0");
    if (d->friction < 0.0) return fail(MPMW_E_INVAL, "friction must be >= 0");
    s->desc = *d;
    s->tool_kind = d->kind;
    return MPMW_OK;
}

}  // extern "C"

namespace {

// ToolSDF::build / rebuild (ToolSDF.cu:212-346) of a triangle mesh into the
// solver's voxel SDF, on the device; the tool becomes the voxel tool `d`.
int build_tool_sdf(mpmw_solver *s, const mpmw_tool_desc *d, const double *verts, int32_t nverts,
                   const int32_t *tris, const double *normals, int32_t ntris, double padding,
                   int32_t voxels_per_axis) {
#pragma clang fp contract(off)
    if (!d || !verts || !tris || !normals) return fail(MPMW_E_INVAL, "set_tool_sdf_mesh: null argument");
    // build, ToolSDF.cu:219-222: an empty mesh builds nothing
    if (nverts <= 0 || ntris <= 0) return fail(MPMW_E_INVAL, "set_tool_sdf_mesh: empty mesh");
    if (voxels_per_axis <= 0) return fail(MPMW_E_INVAL, "set_tool_sdf_mesh: voxels_per_axis must be > 0");
    if (!(d->grad_eps > 0.0)) return fail(MPMW_E_INVAL, "tool grad_eps must be > 0");
    if (d->friction < 0.0) return fail(MPMW_E_INVAL, "friction must be >= 0");
    if (!finite_array(verts, 3 * (int64_t)nverts) || !finite_array(normals, 3 * (int64_t)ntris) ||
        !std::isfinite(padding))
        return fail(MPMW_E_INVAL, "set_tool_sdf_mesh: non-finite input");
    for (int64_t i = 0; i < 3 * (int64_t)ntris; ++i)
        if (tris[i] < 0 || tris[i] >= nverts) return fail(MPMW_E_INVAL, "set_tool_sdf_mesh: vertex index out of range");
    // box, voxel size and dimensions in float (:224-266)
    const float fmax_ = std::numeric_limits<float>::max();
    float lo[3] = {fmax_, fmax_, fmax_}, hi[3] = {-fmax_, -fmax_, -fmax_};
    for (int32_t i = 0; i < nverts; ++i)
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::fmin(lo[k], (float)verts[3 * i + k]);
            hi[k] = std::fmax(hi[k], (float)verts[3 * i + k]);
        }
    const float pad = (float)padding;
    for (int k = 0; k < 3; ++k) { lo[k] -= pad; hi[k] += pad; }
    const float ext[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
    const float vox = std::fmax(ext[0], std::fmax(ext[1], ext[2])) / (float)voxels_per_axis;
    if (!(vox > 0.0f) || !std::isfinite(vox)) return fail(MPMW_E_INVAL, "set_tool_sdf_mesh: degenerate mesh box");
    int dims[3];
    for (int k = 0; k < 3; ++k) dims[k] = std::min(256, (int)(ext[k] / vox) + 2);
    const int nx = dims[0], ny = dims[1], nz = dims[2];
    const int64_t total = (int64_t)nx * ny * nz;
    // triangles with their boxes (buildFromVerts, :348-378) and the voxel
    // index boxes of the scatter (computeSDFCPU, :394-408)
    std::vector<SdfTri> T((size_t)ntris);
    const float inv = 1.0f / vox;
    for (int32_t t = 0; t < ntris; ++t) {
        SdfTri &q = T[(size_t)t];
        for (int k = 0; k < 3; ++k) {
            q.a[k] = (float)verts[3 * tris[3 * t + 0] + k];
            q.b[k] = (float)verts[3 * tris[3 * t + 1] + k];
            q.c[k] = (float)verts[3 * tris[3 * t + 2] + k];
            q.n[k] = (float)normals[3 * t + k];
        }
        for (int k = 0; k < 3; ++k) {
            q.lo[k] = std::fmin(q.a[k], std::fmin(q.b[k], q.c[k]));
            q.hi[k] = std::fmax(q.a[k], std::fmax(q.b[k], q.c[k]));
            q.i0[k] = std::max(0, (int)((q.lo[k] - lo[k]) * inv) - 1);
            q.i1[k] = std::min(dims[k] - 1, (int)((q.hi[k] - lo[k]) * inv) + 2);
        }
    }
    HIP_TRY(hipStreamSynchronize(s->stream));
    void *d_tris = nullptr, *d_bits = nullptr, *d_grid = nullptr;
    auto release = [&]() {
        dfree(s, d_tris, 0);
        dfree(s, d_bits, 0);
        dfree(s, d_grid, 0);
    };
    int rc = dalloc(s, &d_tris, sizeof(SdfTri) * (size_t)ntris);
    if (!rc) rc = dalloc(s, &d_bits, sizeof(uint32_t) * (size_t)total);
    if (!rc) rc = dalloc(s, &d_grid, sizeof(float) * (size_t)total);
    if (rc) {
        release();
        return rc;
    }
    std::vector<float> G((size_t)total);
    uint32_t big_bits;
    std::memcpy(&big_bits, &fmax_, 4);
    hipError_t e = hipMemcpyAsync(d_tris, T.data(), sizeof(SdfTri) * T.size(), hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) {
        const unsigned nb = (unsigned)((total + 255) / 256);
        k_sdf_fill<<<nb, 256, 0, s->stream>>>((uint32_t *)d_bits, total, big_bits);
        k_sdf_scatter<<<(unsigned)ntris, 256, 0, s->stream>>>((const SdfTri *)d_tris, lo[0], lo[1], lo[2], vox, nx,
                                                              ny, (uint32_t *)d_bits);
        k_sdf_sign<<<nb, 256, 0, s->stream>>>((const SdfTri *)d_tris, ntris, lo[0], lo[1], lo[2], vox, nx, ny, total,
                                              vox * 2.0f, (const uint32_t *)d_bits, (float *)d_grid);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(G.data(), d_grid, sizeof(float) * (size_t)total, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) {
        (void)hipStreamSynchronize(s->stream);
        release();
        return fail(MPMW_E_HIP, std::string("set_tool_sdf_mesh: ") + hipGetErrorString(e));
    }
    // voxels no triangle reached: the smallest |value| of the 26 neighbours
    // plus half a voxel, three in-place passes over the interior in index
    // order (:489-520), then the rest at voxel * max(dims) (:523-527)
    const float half = vox * 0.5f;
    for (int pass = 0; pass < 3; ++pass)
        for (int iz = 1; iz < nz - 1; ++iz)
            for (int iy = 1; iy < ny - 1; ++iy)
                for (int ix = 1; ix < nx - 1; ++ix) {
                    const int64_t idx = ix + (int64_t)nx * (iy + (int64_t)ny * iz);
                    if (G[(size_t)idx] < fmax_) continue;
                    float nearest = fmax_;
                    for (int dz = -1; dz <= 1; ++dz)
                        for (int dy = -1; dy <= 1; ++dy)
                            for (int dx = -1; dx <= 1; ++dx) {
                                if (dx == 0 && dy == 0 && dz == 0) continue;
                                const float v = std::fabs(G[(size_t)(idx + dx + (int64_t)nx * (dy + (int64_t)ny * dz))]);
                                if (v < nearest) nearest = v;
                            }
                    if (nearest < fmax_) G[(size_t)idx] = nearest + half;
                }
    const float rest = vox * (float)std::max(nx, std::max(ny, nz));
    for (auto &v : G)
        if (v >= fmax_) v = rest;
    e = hipMemcpyAsync(d_grid, G.data(), sizeof(float) * (size_t)total, hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) {
        (void)hipStreamSynchronize(s->stream);
        release();
        return fail(MPMW_E_HIP, std::string("set_tool_sdf_mesh: ") + hipGetErrorString(e));
    }
    dfree(s, s->sdf, 0);
    s->sdf = d_grid;
    d_grid = nullptr;
    release();
    for (int k = 0; k < 3; ++k) {
        s->sdf_dims[k] = dims[k];
        s->sdf_origin[k] = (double)lo[k];
    }
    s->sdf_voxel = (double)vox;
    s->desc = *d;
    s->desc.kind = MPMW_TOOL_VOXEL;
    s->tool_kind = MPMW_TOOL_VOXEL;
    return MPMW_OK;
}

}  // namespace

extern "C" {

int mpmw_set_tool_sdf_mesh(mpmw_solver *s, const mpmw_tool_desc *d, const double *verts, int32_t nverts,
                           const int32_t *tris, const double *normals, int32_t ntris, double padding,
                           int32_t voxels_per_axis) {
#pragma clang fp contract(off)
    CHECK_S(s);
    int rc = build_tool_sdf(s, d, verts, nverts, tris, normals, ntris, padding, voxels_per_axis);
    if (rc) return rc;
    // the rest mesh and its vertex normals for wear morphing
    // (ContactSolver::initSDF, ContactSolver.cu:1024-1044): each face normal
    // normalised (when its length exceeds 1e-12) and added to its three
    // vertices, each sum normalised the same way
    ToolMesh &m = s->mesh;
    m.verts.assign(verts, verts + 3 * (size_t)nverts);
    m.tris.assign(tris, tris + 3 * (size_t)ntris);
    m.normals.assign(normals, normals + 3 * (size_t)ntris);
    m.vnormals.assign(3 * (size_t)nverts, 0.0);
    for (int32_t t = 0; t < ntris; ++t) {
        double n[3] = {normals[3 * t], normals[3 * t + 1], normals[3 * t + 2]};
        const double len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        if (len > 1e-12)
            for (int k = 0; k < 3; ++k) n[k] /= len;
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k) m.vnormals[3 * (size_t)tris[3 * t + j] + k] += n[k];
    }
    for (int32_t i = 0; i < nverts; ++i) {
        double *v = &m.vnormals[3 * (size_t)i];
        const double len = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        if (len > 1e-12)
            for (int k = 0; k < 3; ++k) v[k] /= len;
    }
    m.padding = padding;
    m.voxels = voxels_per_axis;
    return MPMW_OK;
}

int mpmw_morph_tool_sdf(mpmw_solver *s, const double *wear, int32_t nwear) {
#pragma clang fp contract(off)
    CHECK_S(s);
    ToolMesh &m = s->mesh;
    if (m.verts.empty()) return fail(MPMW_E_STATE, "morph_tool_sdf: no tool mesh was set");
    if (nwear < 0 || (nwear > 0 && !wear)) return fail(MPMW_E_INVAL, "morph_tool_sdf: bad wear array");
    if (!finite_array(wear, nwear)) return fail(MPMW_E_INVAL, "morph_tool_sdf: non-finite wear");
    // ContactSolver::morphToolSDF (:1053-1083): worn = rest - normal * wear
    // per vertex (vertex i takes wear[i], none past nwear), then the build
    // with the rest mesh's triangles and face normals
    const int32_t nv = (int32_t)(m.verts.size() / 3);
    std::vector<double> worn(m.verts.size());
    for (int32_t i = 0; i < nv; ++i) {
        const double w = i < nwear ? wear[i] : 0.0;
        for (int k = 0; k < 3; ++k) worn[3 * (size_t)i + k] = m.verts[3 * (size_t)i + k] - m.vnormals[3 * (size_t)i + k] * w;
    }
    const mpmw_tool_desc d = s->desc;
    return build_tool_sdf(s, &d, worn.data(), nv, m.tris.data(), m.normals.data(), (int32_t)(m.tris.size() / 3),
                          m.padding, m.voxels);
}

int mpmw_get_tool_sdf(mpmw_solver *s, float *grid, int32_t dims[3], double origin[3], double *voxel) {
    CHECK_S(s);
    if (!s->sdf) return fail(MPMW_E_STATE, "no voxel SDF built");
    if (dims)
        for (int k = 0; k < 3; ++k) dims[k] = s->sdf_dims[k];
    if (origin)
        for (int k = 0; k < 3; ++k) origin[k] = s->sdf_origin[k];
    if (voxel) *voxel = s->sdf_voxel;
    if (grid) {
        const size_t n = (size_t)s->sdf_dims[0] * s->sdf_dims[1] * s->sdf_dims[2];
        HIP_TRY(hipMemcpyAsync(grid, s->sdf, sizeof(float) * n, hipMemcpyDeviceToHost, s->stream));
        HIP_TRY(hipStreamSynchronize(s->stream));
    }
    return MPMW_OK;
}

int mpmw_set_tool_pose(mpmw_solver *s, const mpmw_tool_pose *p) {
    CHECK_S(s);
    if (!p) return fail(MPMW_E_INVAL, "null pose");
    s->pose = s->pose0 = *p;
    return MPMW_OK;
}

int mpmw_set_tool_tip(mpmw_solver *s, double x, double y, double z) {
    CHECK_S(s);
    s->tip[0] = x; s->tip[1] = y; s->tip[2] = z;
    return MPMW_OK;
}

int mpmw_apply_external_force(mpmw_solver *s, const double c[3], double radius, const double f[3]) {
    CHECK_S(s);
    if (!c || !f) return fail(MPMW_E_INVAL, "null argument");
    if (s->n <= 0) return MPMW_OK;
    if (!s->has_fext) {
        s->has_fext = true;
        for (int k = F_FX; k <= F_FZ; ++k) {
            int rc = ensure_field(s, k);
            if (rc) return rc;
        }
        if (s->alt_cap) {
            for (int k = F_FX; k <= F_FZ; ++k) dfree(s, s->alt[k], field_bytes(s, k, s->alt_cap));
            s->alt_cap = 0;
        }
    }
    if (s->fp64)
        hipLaunchKernelGGL(k_external_force<double>, dim3(blocks_for(s->n)), dim3(kBlock), 0, s->stream,
                           pview<double>(s), c[0], c[1], c[2], radius * radius, f[0], f[1], f[2]);
    else
        hipLaunchKernelGGL(k_external_force<float>, dim3(blocks_for(s->n)), dim3(kBlock), 0, s->stream,
                           pview<float>(s), c[0], c[1], c[2], radius * radius, (float)f[0], (float)f[1],
                           (float)f[2]);
    HIP_TRY(hipGetLastError());
    s->p2g_ready = false;
    return MPMW_OK;
}

int mpmw_step(mpmw_solver *s, double dt) {
    CHECK_S(s);
    if (!std::isfinite(dt)) return fail(MPMW_E_INVAL, "step: dt must be finite");
    return DISPATCH(s, step_impl, s, dt);
}

int mpmw_run(mpmw_solver *s, int32_t nsteps, double dt) {
    CHECK_S(s);
    for (int32_t i = 0; i < nsteps; ++i) {
        int rc = mpmw_step(s, dt);
        if (rc) return rc;
    }
    return MPMW_OK;
}

int mpmw_synchronize(mpmw_solver *s) {
    CHECK_S(s);
    HIP_TRY(hipStreamSynchronize(s->stream));
    return MPMW_OK;
}

int mpmw_stable_dt(mpmw_solver *s, double *dt) {
    CHECK_S(s);
    if (!dt) return fail(MPMW_E_INVAL, "null");
    // getStableTimeStep, :820-828
    const mpmw_params &p = s->prm;
    double wave = std::sqrt(std::max(p.young / std::max(p.density, 1.0), 1.0));
    double v = 0.35 * s->h / wave;
    if (p.has_config) v = std::max(p.min_time_step, std::min(p.max_time_step, v));
    *dt = v;
    return MPMW_OK;
}

int mpmw_get_metrics(mpmw_solver *s, mpmw_metrics *m) {
    CHECK_S(s);
    if (!m) return fail(MPMW_E_INVAL, "null");
    if (s->metrics_pending || s->cf_pending) {
        if (s->cf_pending) {
            int rc = add_contact_chunks(s);
            if (rc) return rc;
            HIP_TRY(hipMemcpyAsync(s->h_cf, s->d_cf, sizeof(double), hipMemcpyDeviceToHost, s->stream));
        }
        HIP_TRY(hipStreamSynchronize(s->stream));
        if (s->metrics_pending) absorb_stats(s);
        if (s->cf_pending) s->contact_force = *s->h_cf;
        s->metrics_pending = s->cf_pending = false;
    }
    m->max_stress = s->max_stress;
    m->max_temperature = s->max_temp;
    m->kinetic_energy = s->ke;
    m->contact_force = s->contact_force;
    m->time = s->time;
    m->step = s->step;
    return MPMW_OK;
}

int mpmw_sync_metrics(mpmw_solver *s) {
    CHECK_S(s);
    if (s->n <= 0) return MPMW_OK;
    int rc = DISPATCH(s, reduce_particles, s);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(s->stream));
    absorb_stats(s);
    s->metrics_pending = false;
    s->stats_fresh = true;
    return MPMW_OK;
}

int mpmw_compact(mpmw_solver *s) {
    CHECK_S(s);
    return DISPATCH(s, compact_impl, s);
}

int mpmw_get_bounds(mpmw_solver *s, double b[6]) {
    CHECK_S(s);
    if (!b) return fail(MPMW_E_INVAL, "null");
    if (s->metrics_pending) {
        HIP_TRY(hipStreamSynchronize(s->stream));
        absorb_stats(s);
        s->metrics_pending = false;
    }
    if (s->n == 0) {
        for (int k = 0; k < 6; ++k) b[k] = 0.0;
        return MPMW_OK;
    }
    for (int k = 0; k < 6; ++k) b[k] = s->bounds[k];
    return MPMW_OK;
}


int mpmw_temperature_at(mpmw_solver *s, double x, double y, double z, double *t) {
    CHECK_S(s);
    if (!t) return fail(MPMW_E_INVAL, "null");
    *t = s->prm.ref_temp;
    if (s->n <= 0) return MPMW_OK;
    int rc = DISPATCH(s, nearest_impl, s, x, y, z, 0, 0.0);
    if (rc) return rc;
    int32_t bi = -1;
    HIP_TRY(hipMemcpyAsync(&bi, s->d_near_out_i, sizeof(int32_t), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (bi < 0) return MPMW_OK;
    if (s->fp64) {
        double v;
        HIP_TRY(hipMemcpy(&v, (double *)s->f[F_T] + bi, sizeof(double), hipMemcpyDeviceToHost));
        *t = v;
    } else {
        float v;
        HIP_TRY(hipMemcpy(&v, (float *)s->f[F_T] + bi, sizeof(float), hipMemcpyDeviceToHost));
        *t = v;
    }
    return MPMW_OK;
}

int mpmw_apply_heat_flux(mpmw_solver *s, double x, double y, double z, double q) {
    CHECK_S(s);
    if (s->n <= 0) return MPMW_OK;
    int rc = DISPATCH(s, nearest_impl, s, x, y, z, 1, q);
    if (rc) return rc;
    s->p2g_ready = false;
    return MPMW_OK;
}

int mpmw_set_particle_temperatures(mpmw_solver *s, const double *t, int32_t count) {
    CHECK_S(s);
    if (!t || count <= 0 || count > s->n) return MPMW_OK;  // :907 ignores bad counts
    void *tmp = nullptr;
    int rc = dalloc(s, &tmp, sizeof(double) * (size_t)count);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(tmp, t, sizeof(double) * count, hipMemcpyHostToDevice, s->stream));
    if (s->fp64)
        hipLaunchKernelGGL(k_set_temperature<double>, dim3(blocks_for(count)), dim3(kBlock), 0, s->stream,
                           (double *)s->f[F_T], (const double *)tmp, count);
    else
        hipLaunchKernelGGL(k_set_temperature<float>, dim3(blocks_for(count)), dim3(kBlock), 0, s->stream,
                           (float *)s->f[F_T], (const double *)tmp, count);
    hipError_t e = hipGetLastError();
    (void)hipStreamSynchronize(s->stream);
    dfree(s, tmp, sizeof(double) * (size_t)count);
    if (e != hipSuccess) return fail(MPMW_E_HIP, hipGetErrorString(e));
    s->p2g_ready = false;
    return MPMW_OK;
}

int mpmw_set_graph(mpmw_solver *s, int32_t on) {
    CHECK_S(s);
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->staging.graph_on = on != 0;
    if (!on) drop_graphs(s->staging);
    return MPMW_OK;
}

int mpmw_set_autotune(mpmw_solver *s, int32_t trials) {
    CHECK_S(s);
    return set_autotune(s->staging, trials);
}

int mpmw_set_launch_config(mpmw_solver *s, const mpmw_tuning *cfg) {
    CHECK_S(s);
    return set_launch_config(s->staging, cfg);
}

int mpmw_get_tuning(mpmw_solver *s, mpmw_tuning *out) {
    CHECK_S(s);
    if (!out) return fail(MPMW_E_INVAL, "null");
    get_tuning(s->staging, out);
    return MPMW_OK;
}

int mpmw_graph_stats(mpmw_solver *s, int64_t out[2]) {
    CHECK_S(s);
    if (!out) return fail(MPMW_E_INVAL, "null");
    out[0] = s->staging.captures;
    out[1] = s->staging.replays;
    return MPMW_OK;
}

int mpmw_set_deterministic(mpmw_solver *s, int32_t on) {
    CHECK_S(s);
    if (on && s->apic) return fail(MPMW_E_STATE, "the deterministic mode is not available with the APIC transfer");
    s->deterministic = on != 0;
    s->p2g_ready = false;
    s->sort.valid = false;
    return MPMW_OK;
}

int mpmw_set_transfer(mpmw_solver *s, int32_t mode) {
    CHECK_S(s);
    if (mode != MPMW_TRANSFER_PIC && mode != MPMW_TRANSFER_APIC && mode != MPMW_TRANSFER_APIC_TEXTBOOK)
        return fail(MPMW_E_INVAL, "unknown transfer");
    const bool apic = mode != MPMW_TRANSFER_PIC;
    if (apic && s->deterministic)
        return fail(MPMW_E_STATE, "the deterministic mode is not available with the APIC transfer");
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->apic = apic;
    s->apic_textbook = mode == MPMW_TRANSFER_APIC_TEXTBOOK;
    for (int c = F_C0; c <= F_C8; ++c) {
        if (apic) {
            int rc = ensure_field(s, c);  // zero-filled
            if (rc) return rc;
            if (s->f[c] && s->cap > 0)
                hipLaunchKernelGGL(k_fill<float>, dim3(std::min(blocks_for((int64_t)s->cap * s->rsz / 4), 4096)),
                                   dim3(kBlock), 0, s->stream, (float *)s->f[c], (int64_t)s->cap * s->rsz / 4, 0.0f);
        } else {
            dfree(s, s->f[c], field_bytes(s, c, s->cap));
            dfree(s, s->alt[c], field_bytes(s, c, s->alt_cap));
        }
    }
    HIP_TRY(hipGetLastError());
    // the binning's loose rule and the window layout depend on the stencil
    s->p2g_ready = false;
    s->sort.valid = false;
    return MPMW_OK;
}

int mpmw_get_affine(mpmw_solver *s, double *C, int32_t n) {
    CHECK_S(s);
    if (!s->apic) return fail(MPMW_E_STATE, "get_affine: the transfer is not APIC");
    if (!C || n < 0) return fail(MPMW_E_INVAL, "get_affine: bad arguments");
    HIP_TRY(hipStreamSynchronize(s->stream));
    const int32_t m = std::min(n, s->n);
    for (int c = 0; c < 9; ++c) {
        int rc = s->fp64 ? download_field<double>(s, F_C0 + c, C, m, 9, c)
                         : download_field<float>(s, F_C0 + c, C, m, 9, c);
        if (rc) return rc;
    }
    return MPMW_OK;
}

int mpmw_set_affine(mpmw_solver *s, const double *C, int32_t n) {
    CHECK_S(s);
    if (!s->apic) return fail(MPMW_E_STATE, "set_affine: the transfer is not APIC");
    if (!C || n != s->n) return fail(MPMW_E_INVAL, "set_affine: n must equal the particle count");
    if (!finite_array(C, 9 * (int64_t)n)) return fail(MPMW_E_INVAL, "set_affine: non-finite value");
    HIP_TRY(hipStreamSynchronize(s->stream));
    for (int c = 0; c < 9; ++c) {
        int rc = s->fp64 ? upload_field<double>(s, F_C0 + c, C, n, 9, c, 0.0)
                         : upload_field<float>(s, F_C0 + c, C, n, 9, c, 0.0);
        if (rc) return rc;
    }
    s->p2g_ready = false;   // the next P2G uses the new matrices
    return MPMW_OK;
}

int mpmw_get_grid_info(mpmw_solver *s, mpmw_grid_info *g) {
    CHECK_S(s);
    if (!g) return fail(MPMW_E_INVAL, "null");
    g->cell = s->h;
    g->minx = s->minx; g->miny = s->miny; g->minz = s->minz;
    g->nx = s->nx; g->ny = s->ny; g->nz = s->nz;
    g->pad_ = 0;
    return MPMW_OK;
}

int mpmw_set_grid_snapshot(mpmw_solver *s, int32_t on) {
    CHECK_S(s);
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->snap_on = on != 0;
    s->snap_valid = false;
    if (s->snap_on && !s->snap && s->nn_cap > 0) {
        int rc = dalloc(s, &s->snap, 8 * s->rsz * (size_t)s->nn_cap);
        if (rc) { s->snap_on = false; return rc; }
    }
    return MPMW_OK;
}

int mpmw_get_grid_snapshot(mpmw_solver *s, double *mass, double *mom, double *force, double *tm) {
    CHECK_S(s);
    if (!s->snap_valid || !s->snap) return fail(MPMW_E_STATE, "no grid snapshot: enable it and step first");
    HIP_TRY(hipStreamSynchronize(s->stream));
    const int64_t nn = s->nn;
    std::vector<double> buf((size_t)(8 * nn));
    if (s->fp64) {
        HIP_TRY(hipMemcpy(buf.data(), s->snap, sizeof(double) * 8 * nn, hipMemcpyDeviceToHost));
    } else {
        std::vector<float> f((size_t)(8 * nn));
        HIP_TRY(hipMemcpy(f.data(), s->snap, sizeof(float) * 8 * nn, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < f.size(); ++i) buf[i] = f[i];
    }
    for (int64_t i = 0; i < nn; ++i) {
        if (mass) mass[i] = buf[i];
        if (mom) for (int a = 0; a < 3; ++a) mom[3 * i + a] = buf[(1 + a) * nn + i];
        if (force) for (int a = 0; a < 3; ++a) force[3 * i + a] = buf[(4 + a) * nn + i];
        if (tm) tm[i] = buf[7 * nn + i];
    }
    return MPMW_OK;
}

int mpmw_set_kernel_timing(mpmw_solver *s, int32_t mode) {
    CHECK_S(s);
    if (mode != MPMW_TIMING_OFF && mode != MPMW_TIMING_ALL && mode != MPMW_TIMING_PARTICLE_PASS)
        return fail(MPMW_E_INVAL, "set_kernel_timing: unknown mode");
    drain_events(s);
    s->timing = mode != MPMW_TIMING_OFF;
    s->timing_all = mode == MPMW_TIMING_ALL;
    // events for 256 timed steps (three pairs each) made here rather than
    // one by one inside the steps, where creating them held up the host
    constexpr size_t kPool = 256 * 6;
    while (s->timing && s->ev_pool.size() < kPool) {
        hipEvent_t e;
        if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) break;
        s->ev_pool.push_back(e);
    }
    for (int k = 0; k < 8; ++k) { s->t_ms[k] = 0; s->t_n[k] = 0; }
    return MPMW_OK;
}

int mpmw_get_kernel_times(mpmw_solver *s, mpmw_kernel_times *t) {
    CHECK_S(s);
    if (!t) return fail(MPMW_E_INVAL, "null");
    drain_events(s);
    for (int k = 0; k < 8; ++k) { t->ms[k] = s->t_ms[k]; t->launches[k] = s->t_n[k]; }
    return MPMW_OK;
}

int mpmw_get_stream(mpmw_solver *s, void **stream) {
    CHECK_S(s);
    if (!stream) return fail(MPMW_E_INVAL, "null");
    *stream = (void *)s->stream;
    return MPMW_OK;
}

int mpmw_device_bytes(mpmw_solver *s, int64_t *bytes) {
    CHECK_S(s);
    if (!bytes) return fail(MPMW_E_INVAL, "null");
    *bytes = s->device_bytes;
    return MPMW_OK;
}

}  // extern "C"

extern "C" int mpmw_get_diagnostics(mpmw_solver *s, mpmw_diagnostics *d) {
    CHECK_S(s);
    if (!d) return fail(MPMW_E_INVAL, "null");
    HIP_TRY(hipStreamSynchronize(s->stream));
    std::memset(d, 0, sizeof(*d));
    const SortState &S = s->sort;
    d->sorts = S.sorts;
    d->steps_since_rebuild = S.steps_since;
    d->tiles = S.geom.ntiles;
    d->tile_cells = kTileCells;
    if (S.oow) {
        int32_t v = 0;
        HIP_TRY(hipMemcpy(&v, S.oow, sizeof(int32_t), hipMemcpyDeviceToHost));
        d->out_of_window = v;
    }
    d->loose_particles = S.valid ? S.loose_end - S.loose_begin : 0;
    d->deterministic = s->deterministic ? 1 : 0;
    d->det_rebins = S.resorts;
    if (S.valid && S.tile_start && S.geom.ntiles > 0) {
        std::vector<int32_t> ts((size_t)S.geom.ntiles + 1);
        HIP_TRY(hipMemcpy(ts.data(), S.tile_start, sizeof(int32_t) * ts.size(), hipMemcpyDeviceToHost));
        for (int t = 0; t < S.geom.ntiles; ++t) d->max_tile_particles = std::max(d->max_tile_particles, ts[t + 1] - ts[t]);
    }
    return MPMW_OK;
}

// =============================================================== batch ====
//
// Batched stepping: each kernel of the step is launched once for every
// member over a flattened block list (k_*_batch, find_item); the host work
// of each member (rebuilds, binning, standalone P2G, statistics) runs as in
// its own step.  Work items travel in one blob per step, staged through one
// of two pinned host areas (an event guards each area's reuse).

struct mpmw_batch {
    int32_t device = 0;
    bool fp64 = false;
    hipStream_t stream = nullptr;
    std::vector<mpmw_solver *> members;
    StepStaging staging;
};

namespace {

void batch_detach(mpmw_batch *b, mpmw_solver *s) {
    (void)hipStreamSynchronize(b->stream);
    s->stream = s->own_stream;
    s->own_stream = nullptr;
    s->batch = nullptr;
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

template <typename R> void launch_tile_batch(hipStream_t stream, int variant, const TileWork<R> *items, int n,
                                             int blocks, int order) {
    const dim3 grid(blocks), blk(kBlock);
    switch (variant) {
    case 0: hipLaunchKernelGGL((k_tile_batch<R, true, true, false, false>), grid, blk, 0, stream, items, n, order); break;
    case 1: hipLaunchKernelGGL((k_tile_batch<R, true, true, true, false>), grid, blk, 0, stream, items, n, order); break;
    case 2: hipLaunchKernelGGL((k_tile_batch<R, true, true, false, true>), grid, blk, 0, stream, items, n, order); break;
    case 3: hipLaunchKernelGGL((k_tile_batch<R, true, true, true, true>), grid, blk, 0, stream, items, n, order); break;
    case 4: hipLaunchKernelGGL((k_tile_batch<R, true, false, false, false>), grid, blk, 0, stream, items, n, order); break;
    case 5: hipLaunchKernelGGL((k_tile_batch<R, true, false, true, false>), grid, blk, 0, stream, items, n, order); break;
    case 6: hipLaunchKernelGGL((k_apic_tile_batch<R, true, true, false>), grid, blk, 0, stream, items, n, order); break;
    case 7: hipLaunchKernelGGL((k_apic_tile_batch<R, true, true, true>), grid, blk, 0, stream, items, n, order); break;
    case 8: hipLaunchKernelGGL((k_apic_tile_batch<R, true, false, false>), grid, blk, 0, stream, items, n, order); break;
    case 9: hipLaunchKernelGGL((k_apic_tile_batch<R, true, false, true>), grid, blk, 0, stream, items, n, order); break;
    case 10: hipLaunchKernelGGL((k_tile_batch<R, false, true, false, false>), grid, blk, 0, stream, items, n, order); break;
    case 11: hipLaunchKernelGGL((k_tile_batch<R, false, true, true, false>), grid, blk, 0, stream, items, n, order); break;
    case 12: hipLaunchKernelGGL((k_apic_tile_batch<R, false, true, false>), grid, blk, 0, stream, items, n, order); break;
    default: hipLaunchKernelGGL((k_apic_tile_batch<R, false, true, true>), grid, blk, 0, stream, items, n, order); break;
    }
}

// Tile-pass variant of a member: fused (0-3: backstress, deterministic) or
// G2P-only before a rebuild (4-5: backstress); the APIC transfer's fused
// (6-7) and G2P-only (8-9) passes; the standalone P2G of the step's begin
// (10-11 trilinear, 12-13 APIC: backstress).
constexpr int kTileVariants = 14;
int p2g_variant(const mpmw_solver *s) { return (s->apic ? 12 : 10) + (s->has_back ? 1 : 0); }
int tile_variant(const mpmw_solver *s, const StepPlan &sp) {
    const int back = s->has_back ? 1 : 0;
    if (s->apic) return (sp.next_rebuild ? 8 : 6) + back;
    if (sp.next_rebuild) return 4 + back;
    return back + (s->deterministic ? 2 : 0);
}

template <typename R> int loose_after_batch(mpmw_solver *s, double dt, const StepPlan &sp) {
    const bool back = s->has_back, det = s->deterministic;
    if (sp.next_rebuild)
        return back ? loose_pass_t<R, true, false, true, false>(s, dt, sp.zcut, sp.fl)
                    : loose_pass_t<R, true, false, false, false>(s, dt, sp.zcut, sp.fl);
    if (back)
        return det ? loose_pass_t<R, true, true, true, true>(s, dt, sp.zcut, sp.fl)
                   : loose_pass_t<R, true, true, true, false>(s, dt, sp.zcut, sp.fl);
    return det ? loose_pass_t<R, true, true, false, true>(s, dt, sp.zcut, sp.fl)
               : loose_pass_t<R, true, true, false, false>(s, dt, sp.zcut, sp.fl);
}

void drop_graphs(StepStaging &st) {
    for (int k = 0; k < 2; ++k) {
        if (st.exec[k]) (void)hipGraphExecDestroy(st.exec[k]);
        st.exec[k] = nullptr;
        st.sig[k].clear();
    }
}

// Free the staging areas of a solver or a batch.
void free_staging(StepStaging &st) {
    drop_graphs(st);
    for (auto &e : st.tune.ev)
        if (e) (void)hipEventDestroy(e), e = nullptr;
    for (auto &a : st.area) {
        if (a.copied) (void)hipEventDestroy(a.copied);
        pool::put(a.h);
        a = StepStaging::Area();
    }
    pool::put(st.d_blob);
    st.d_blob = nullptr;
    st.d_cap = 0;
}

int set_autotune(StepStaging &st, int32_t trials) {
    if (trials < 0 || trials > 16) return fail(MPMW_E_INVAL, "autotune trials must be in [0, 16]");
    StepStaging::Tuner &t = st.tune;
    t.trials = trials;
    t.phase = t.cand = t.trial = 0;
    std::memset(t.ms, 0, sizeof(t.ms));
    t.state = trials ? 1 : 0;
    return MPMW_OK;
}

int set_launch_config(StepStaging &st, const mpmw_tuning *c) {
    if (!c) return fail(MPMW_E_INVAL, "null");
    if ((c->tile_order | c->grid_order | c->post_order) & ~1)
        return fail(MPMW_E_INVAL, "block orders are 0 (blockIdx) or 1 (per-XCD ranges)");
    if (c->grid_block != 64 && c->grid_block != 128 && c->grid_block != 256)
        return fail(MPMW_E_INVAL, "grid_block must be 64, 128 or 256");
    StepStaging::Tuner &t = st.tune;
    t.tile_order = c->tile_order;
    t.grid_order = c->grid_order;
    t.grid_block = c->grid_block;
    t.post_order = c->post_order;
    t.trials = 0;
    t.state = 0;
    return MPMW_OK;
}

void get_tuning(const StepStaging &st, mpmw_tuning *o) {
    const StepStaging::Tuner &t = st.tune;
    std::memset(o, 0, sizeof(*o));
    o->state = t.state;
    o->tile_order = t.tile_order;
    o->grid_order = t.grid_order;
    o->grid_block = t.grid_block;
    o->post_order = t.post_order;
    o->trials = t.trials;
    for (int k = 0; k < 2; ++k) { o->tile_ms[k] = t.ms[0][k]; o->post_ms[k] = t.ms[2][k]; }
    for (int k = 0; k < 6; ++k) o->grid_ms[k] = t.ms[1][k];
}

// Stages `end` bytes, written by `fill` into one of the two pinned areas,
// into the device blob on `stream` (kernels of the same stream read it);
// returns the blob's device address in *dev.
template <typename Fill> int stage_blob(StepStaging &st, hipStream_t stream, size_t end, Fill fill, char **dev) {
    StepStaging::Area &A = st.area[st.parity];
    st.parity ^= 1;
    // the event only tells the host that the copy has read the area
    if (!A.copied) HIP_TRY(hipEventCreateWithFlags(&A.copied, hipEventDisableTiming | hipEventDisableSystemFence));
    if (A.pending) {
        HIP_TRY(hipEventSynchronize(A.copied));
        A.pending = false;
    }
    if (end > A.cap) {
        pool::put(A.h);
        A.h = nullptr;
        A.cap = 0;
        int dev_id = 0;
        HIP_TRY(hipGetDevice(&dev_id));
        HIP_TRY(pool::get(dev_id, true, end * 2, &A.h));
        A.cap = end * 2;
    }
    if (end > st.d_cap) {
        HIP_TRY(hipStreamSynchronize(stream));
        pool::put(st.d_blob);
        st.d_blob = nullptr;
        st.d_cap = 0;
        int dev_id = 0;
        HIP_TRY(hipGetDevice(&dev_id));
        HIP_TRY(pool::get(dev_id, false, end * 2, &st.d_blob));
        st.d_cap = end * 2;
    }
    fill((char *)A.h);
    HIP_TRY(hipMemcpyAsync(st.d_blob, A.h, end, hipMemcpyHostToDevice, stream));
    HIP_TRY(hipEventRecord(A.copied, stream));
    A.pending = true;
    *dev = (char *)st.d_blob;
    return MPMW_OK;
}

// The step's launches as a HIP graph: `fill` writes the blob into the next
// pinned area, then the graph of that area (its copy into the device blob
// and the kernels `launch` enqueues on the blob) is launched, captured first
// when its signature (blob size and addresses, grid sizes; whatever the
// kernels' arguments hold) differs from the one it was captured with.  The
// per-step values (dt, tool pose, flags) travel in the blob, so a replay
// computes exactly what the eager launches compute.
template <typename Fill, typename Launch>
int graph_step(StepStaging &st, hipStream_t stream, size_t end, std::vector<uint64_t> sig, Fill fill,
               Launch launch) {
    const int k = st.parity;
    StepStaging::Area &A = st.area[k];
    st.parity ^= 1;
    if (!A.copied) HIP_TRY(hipEventCreateWithFlags(&A.copied, hipEventDisableTiming | hipEventDisableSystemFence));
    if (A.pending) {
        HIP_TRY(hipEventSynchronize(A.copied));
        A.pending = false;
    }
    int dev_id = 0;
    HIP_TRY(hipGetDevice(&dev_id));
    if (end > A.cap) {
        if (st.exec[k]) {
            (void)hipGraphExecDestroy(st.exec[k]);
            st.exec[k] = nullptr;
        }
        pool::put(A.h);
        A.h = nullptr;
        A.cap = 0;
        HIP_TRY(pool::get(dev_id, true, end * 2, &A.h));
        A.cap = end * 2;
    }
    if (end > st.d_cap) {
        HIP_TRY(hipStreamSynchronize(stream));
        drop_graphs(st);
        pool::put(st.d_blob);
        st.d_blob = nullptr;
        st.d_cap = 0;
        HIP_TRY(pool::get(dev_id, false, end * 2, &st.d_blob));
        st.d_cap = end * 2;
    }
    fill((char *)A.h);
    sig.push_back((uint64_t)end);
    sig.push_back((uint64_t)(uintptr_t)A.h);
    sig.push_back((uint64_t)(uintptr_t)st.d_blob);
    if (!st.exec[k] || st.sig[k] != sig) {
        if (st.exec[k]) HIP_TRY(hipGraphExecDestroy(st.exec[k]));
        st.exec[k] = nullptr;
        hipGraph_t g = nullptr;
        HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        hipError_t e = hipMemcpyAsync(st.d_blob, A.h, end, hipMemcpyHostToDevice, stream);
        if (e == hipSuccess) {
            launch((char *)st.d_blob);
            e = hipGetLastError();
        }
        const hipError_t e2 = hipStreamEndCapture(stream, &g);
        if (e == hipSuccess) e = e2;
        if (e == hipSuccess) e = hipGraphInstantiate(&st.exec[k], g, nullptr, nullptr, 0);
        if (g) (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            st.exec[k] = nullptr;
            return fail(MPMW_E_HIP, std::string("step graph capture: ") + hipGetErrorString(e));
        }
        st.sig[k] = std::move(sig);
        st.captures++;
    } else {
        st.replays++;
    }
    HIP_TRY(hipGraphLaunch(st.exec[k], stream));
    HIP_TRY(hipEventRecord(A.copied, stream));
    A.pending = true;
    return MPMW_OK;
}

// The tile-pass work item of a solver (blocks = its tiles; block0 is set by
// the caller).
template <typename R> TileWork<R> tile_work(mpmw_solver *s, double dt, double zcut, G2PFlags fl) {
    TileWork<R> t;
    t.P = pview<R>(s);
    t.g = gview<R>(s);
    t.G = grid_arrays<R>(s);
    t.M = material<R>(s);
    t.dt = (R)dt;
    t.dtd = dt;
    t.zcut = zcut;
    t.ref_temp = s->prm.ref_temp;
    t.flags = fl;
    t.part = s->part;
    t.cell_start = s->sort.cell_start;
    t.tg = s->sort.geom;
    t.counters = s->sort.oow;
    t.tpart = (R *)s->tpart;
    t.block0 = 0;
    t.blocks = t.tg.ntiles;
    return t;
}

// Everything before the grid update, for every member (MPMWorkpieceSolver
// ::step :728-745): the rebuild every 25 steps, binning, and the P2G when
// the sums are not ready (first step, after a rebuild or a re-binning).  The
// host reads the members need (AABB, out-of-window counts, loose ranges)
// come from two synchronisations for all of them, and their standalone P2G
// passes are one launch, except for deterministic members, which check
// theirs and may re-bin (det_settle) on their own.
template <typename R> int step_begin_many(mpmw_solver *const *members, size_t nm, const double *dts,
                                         hipStream_t stream, StepStaging &st, std::vector<StepPlan> &plans) {
    int rc;
    std::vector<char> live(nm, 0), sorting(nm, 0);
    bool wait = false;
    for (size_t m = 0; m < nm; ++m) {
        mpmw_solver *s = members[m];
        plans[m] = StepPlan();
        if (s->n <= 0 || s->nn <= 0) continue;  // :728 silent return
        live[m] = 1;
        if (s->step % 25 == 0) {
            if ((rc = rebuild_launch<R>(s))) return rc;
            wait = true;
        }
    }
    if (wait) HIP_TRY(hipStreamSynchronize(stream));
    wait = false;
    for (size_t m = 0; m < nm; ++m) {
        if (!live[m]) continue;
        mpmw_solver *s = members[m];
        bool srt = false;
        if (s->step % 25 == 0 && (rc = rebuild_finish<R>(s, srt))) return rc;
        if (!srt && !s->sort.valid) {
            if ((rc = sort_launch<R>(s))) return rc;
            srt = true;
        }
        sorting[m] = srt ? 1 : 0;
        wait = wait || srt;
    }
    if (wait) {
        HIP_TRY(hipStreamSynchronize(stream));
        for (size_t m = 0; m < nm; ++m)
            if (sorting[m]) sort_finish(members[m]);
    }
    std::vector<TileWork<R>> pw[kTileVariants];
    std::vector<int> pm;
    int pblocks[kTileVariants] = {};
    for (size_t m = 0; m < nm; ++m) {
        if (!live[m]) continue;
        mpmw_solver *s = members[m];
        StepPlan &sp = plans[m];
        sp.zcut = s->tip[2] - s->h * 1.5;
        if ((rc = ensure_part(s, s->sort.geom.ntiles + loose_blocks(s)))) return rc;
        if (s->p2g_ready) continue;
        if (s->deterministic) {
            if ((rc = p2g_alone<R>(s, dts[m], sp.zcut))) return rc;
            if ((rc = det_settle<R>(s, dts[m], sp.zcut))) return rc;
            continue;
        }
        // P2G into empty sums: node accumulators and counters [1], [2] cleared
        if ((rc = zero_acc<R>(s))) return rc;
        hipLaunchKernelGGL(k_fill<int32_t>, dim3(1), dim3(64), 0, stream, s->sort.oow + 1, (int64_t)2, 0);
        const int v = p2g_variant(s);
        TileWork<R> t = tile_work<R>(s, dts[m], sp.zcut, G2PFlags{});
        t.block0 = pblocks[v];
        pblocks[v] += t.blocks;
        pw[v].push_back(t);
        pm.push_back((int)m);
    }
    if (!pm.empty()) {
        size_t off[kTileVariants], end = 0;
        for (int v = 0; v < kTileVariants; ++v) {
            off[v] = end;
            end += align256(sizeof(TileWork<R>) * pw[v].size());
        }
        char *d = nullptr;
        if ((rc = stage_blob(st, stream, end, [&](char *h) {
                 for (int v = 0; v < kTileVariants; ++v)
                     if (!pw[v].empty()) std::memcpy(h + off[v], pw[v].data(), sizeof(TileWork<R>) * pw[v].size());
             }, &d)))
            return rc;
        {
            Timed t(nm == 1 ? members[0] : nullptr, MPMW_K_P2G);
            for (int v = 0; v < kTileVariants; ++v)
                if (!pw[v].empty())
                    launch_tile_batch<R>(stream, v, (const TileWork<R> *)(d + off[v]), (int)pw[v].size(),
                                         pblocks[v], st.tune.tile_order);
        }
        HIP_TRY(hipGetLastError());
        for (int m : pm) {
            mpmw_solver *s = members[m];
            rc = s->has_back ? loose_pass_t<R, false, true, true, false>(s, dts[m], plans[m].zcut, G2PFlags{})
                             : loose_pass_t<R, false, true, false, false>(s, dts[m], plans[m].zcut, G2PFlags{});
            if (rc) return rc;
        }
    }
    for (size_t m = 0; m < nm; ++m) {
        if (!live[m]) continue;
        mpmw_solver *s = members[m];
        StepPlan &sp = plans[m];
        // the binning may have changed in det_settle
        if ((rc = ensure_part(s, s->sort.geom.ntiles + loose_blocks(s)))) return rc;
        sp.diffusion = s->prm.specific_heat > 0.0 &&
                       s->prm.thermal_conductivity / (s->prm.density * s->prm.specific_heat) > 0.0;
        sp.alpha = sp.diffusion ? s->prm.thermal_conductivity / (s->prm.density * s->prm.specific_heat) : 0.0;
        // G2P + chip pass, fused with the next step's P2G unless the next
        // step rebuilds the grid (then the AABB and metrics are reduced
        // instead)
        sp.next_rebuild = ((s->step + 1) % 25 == 0);
        sp.fl.chip = (s->step % 2 == 0);
        sp.fl.stats = sp.next_rebuild ? 1 : 0;
        sp.run = true;
    }
    return MPMW_OK;
}

// The autotuner's candidates, one phase per kernel class: the tile pass's
// block order; the grid update's block order and block size; the post
// pass's block order.
constexpr int kTunePhases = 3;
constexpr int kTuneCands[kTunePhases] = {2, 6, 2};
constexpr int kGridCand[6][2] = {{1, 256}, {1, 128}, {1, 64}, {0, 256}, {0, 128}, {0, 64}};

// The configuration to run a step with: while tuning, the best so far with
// the phase's knob set to the candidate under test.
StepStaging::Tuner step_config(const StepStaging::Tuner &t) {
    StepStaging::Tuner c = t;
    if (t.state != 1) return c;
    if (t.phase == 0) c.tile_order = t.cand;
    if (t.phase == 1) { c.grid_order = kGridCand[t.cand][0]; c.grid_block = kGridCand[t.cand][1]; }
    if (t.phase == 2) c.post_order = t.cand == 0 ? 1 : 0;
    return c;
}

// One timed sample of the tuned phase (ms of its kernel class on a step
// without a rebuild); after `trials` samples the candidate's median is
// kept, after the phase's last candidate the fastest is adopted.
void tune_record(StepStaging::Tuner &t, float ms) {
    t.sample[t.trial++] = ms;
    if (t.trial < t.trials) return;
    std::sort(t.sample, t.sample + t.trial);
    t.ms[t.phase][t.cand] = t.sample[t.trial / 2];
    t.trial = 0;
    if (++t.cand < kTuneCands[t.phase]) return;
    int best = 0;
    for (int k = 1; k < kTuneCands[t.phase]; ++k)
        if (t.ms[t.phase][k] < t.ms[t.phase][best]) best = k;
    if (t.phase == 0) t.tile_order = best;
    if (t.phase == 1) { t.grid_order = kGridCand[best][0]; t.grid_block = kGridCand[best][1]; }
    if (t.phase == 2) t.post_order = best == 0 ? 1 : 0;
    t.cand = 0;
    if (++t.phase == kTunePhases) t.state = 2;
}

// One step of every solver in `members` (all of precision R, on `stream`):
// the host phases of each solver's own step, and each kernel of the step
// launched once for all of them over a flattened block list.
template <typename R> int step_many(mpmw_solver *const *members, size_t nm, const double *dts,
                                   hipStream_t stream, StepStaging &st) {
    std::vector<StepPlan> plans(nm);
    int rc;
    if ((rc = step_begin_many<R>(members, nm, dts, stream, st, plans))) return rc;
    std::vector<int> run;
    for (size_t m = 0; m < nm; ++m)
        if (plans[m].run) run.push_back((int)m);
    if (run.empty()) return MPMW_OK;
    const int nr = (int)run.size();
    // a tuning sample only on a step whose members all run the fused pass
    bool clean = true;
    for (int r = 0; r < nr; ++r) clean = clean && !plans[run[r]].next_rebuild;
    const bool sampling = st.tune.state == 1 && clean;
    const StepStaging::Tuner cfg = sampling ? step_config(st.tune) : st.tune;

    // work items
    std::vector<GridWork<R>> gw(nr);
    std::vector<SumWork> sw(nr);
    std::vector<TileWork<R>> tw[kTileVariants];
    int gblocks = 0, dblocks = 0, sblocks = 0, tblocks[kTileVariants] = {};
    for (int r = 0; r < nr; ++r) {
        mpmw_solver *s = members[run[r]];
        const StepPlan &sp = plans[run[r]];
        const double dt = dts[run[r]];
        GridWork<R> &g = gw[r];
        g.g = gview<R>(s);
        g.G = grid_arrays<R>(s);
        g.tp = tile_parts<R>(s);
        g.M = material<R>(s);
        g.tool = tool_view(s);
        g.dt = (R)dt;
        g.alpha = (R)sp.alpha;
        g.direct_T = sp.diffusion ? 0 : 1;
        g.diffuse = sp.diffusion ? 1 : 0;
        g.force_part = s->fpart;
        g.block0 = gblocks;
        g.blocks = blocks_for(s->nn, cfg.grid_block);
        g.fwaves = (int32_t)((s->nn + kWave - 1) / kWave);
        g.pad_ = 0;
        gblocks += g.blocks;
        g.dblock0 = dblocks;
        g.dblocks = blocks_for(s->nn, kDiffuseBlock);
        dblocks += g.dblocks;
        sw[r].part = s->fpart;
        sw[r].nparts = g.fwaves;
        sw[r].chunks = (g.fwaves + kSumChunk - 1) / kSumChunk;
        sw[r].block0 = sblocks;
        sw[r].pad_ = 0;
        sblocks += sw[r].chunks;
        sw[r].modes = s->d_modes + kRed;
        sw[r].chunk = s->fpart + g.fwaves;
        s->cf_first = g.fwaves;
        s->cf_chunks = sw[r].chunks;
        sw[r].counters = sp.next_rebuild ? nullptr : s->sort.oow;
        const int v = tile_variant(s, sp);
        TileWork<R> t = tile_work<R>(s, dt, sp.zcut, sp.fl);
        t.block0 = tblocks[v];
        tblocks[v] += t.blocks;
        tw[v].push_back(t);
    }
    // a batch of one keeps the per-kernel timing of its solver
    mpmw_solver *one = nr == 1 ? members[run[0]] : nullptr;
    {
        // one blob: grid items, sum items, tile items per variant
        size_t off_sw = align256(sizeof(GridWork<R>) * nr);
        size_t off_tw[kTileVariants];
        size_t end = off_sw + align256(sizeof(SumWork) * nr);
        for (int v = 0; v < kTileVariants; ++v) {
            off_tw[v] = end;
            end += align256(sizeof(TileWork<R>) * tw[v].size());
        }
        auto fill = [&](char *h) {
            std::memcpy(h, gw.data(), sizeof(GridWork<R>) * nr);
            std::memcpy(h + off_sw, sw.data(), sizeof(SumWork) * nr);
            for (int v = 0; v < kTileVariants; ++v)
                if (!tw[v].empty()) std::memcpy(h + off_tw[v], tw[v].data(), sizeof(TileWork<R>) * tw[v].size());
        };
        // the kernels of the step, reading their work items from blob d;
        // with `time_it` each class is bracketed by the solver's timing events
        auto launch = [&](char *d, bool time_it) {
            mpmw_solver *tm = time_it ? one : nullptr;
            hipEvent_t *ev = sampling ? st.tune.ev : nullptr;
            const GridWork<R> *d_gw = (const GridWork<R> *)d;
            if (ev) (void)hipEventRecord(ev[0], stream);
            {
                Timed t(tm, MPMW_K_GRID);
                hipLaunchKernelGGL(k_grid_update_batch<R>, dim3(gblocks), dim3(cfg.grid_block), 0, stream, d_gw, nr,
                                   cfg.grid_order);
            }
            if (ev) (void)hipEventRecord(ev[1], stream);
            {
                // diffusion, contact-force sums and flag resets in one launch
                Timed t(tm, MPMW_K_DIFFUSE);
                hipLaunchKernelGGL(k_post_batch<R>, dim3(sblocks + dblocks), dim3(kBlock), 0, stream, d_gw, nr,
                                   (const SumWork *)(d + off_sw), sblocks, cfg.post_order);
            }
            if (ev) (void)hipEventRecord(ev[2], stream);
            {
                Timed t(tm, MPMW_K_G2P_P2G);
                for (int v = 0; v < kTileVariants; ++v)
                    if (!tw[v].empty())
                        launch_tile_batch<R>(stream, v, (const TileWork<R> *)(d + off_tw[v]), (int)tw[v].size(),
                                             tblocks[v], cfg.tile_order);
            }
            if (ev) (void)hipEventRecord(ev[3], stream);
        };
        if (sampling) {
            for (auto &e : st.tune.ev)
                if (!e && hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) {
                    (void)hipGetLastError();
                    return fail(MPMW_E_HIP, "autotuner: event creation failed");
                }
        }
        // kernel timing and tuning samples need events between the
        // launches: eager then
        if (st.graph_on && !(one && one->timing) && !sampling) {
            std::vector<uint64_t> sig = {(uint64_t)nr, (uint64_t)gblocks, (uint64_t)dblocks, (uint64_t)off_sw,
                                         (uint64_t)cfg.tile_order, (uint64_t)cfg.grid_order,
                                         (uint64_t)cfg.grid_block, (uint64_t)cfg.post_order};
            for (int v = 0; v < kTileVariants; ++v) {
                sig.push_back((uint64_t)tw[v].size());
                sig.push_back((uint64_t)tblocks[v]);
                sig.push_back((uint64_t)off_tw[v]);
            }
            if ((rc = graph_step(st, stream, end, std::move(sig), fill, [&](char *d) { launch(d, false); })))
                return rc;
        } else {
            char *d = nullptr;
            if ((rc = stage_blob(st, stream, end, fill, &d))) return rc;
            launch(d, true);
            if (sampling) {
                HIP_TRY(hipEventSynchronize(st.tune.ev[3]));
                float ms = 0.0f;
                const int a = st.tune.phase == 0 ? 2 : (st.tune.phase == 1 ? 0 : 1);
                HIP_TRY(hipEventElapsedTime(&ms, st.tune.ev[a], st.tune.ev[a + 1]));
                tune_record(st.tune, ms);
            }
        }
    }
    HIP_TRY(hipGetLastError());
    for (int r = 0; r < nr; ++r)
        if ((rc = loose_after_batch<R>(members[run[r]], dts[run[r]], plans[run[r]]))) return rc;
    for (int r = 0; r < nr; ++r)
        if ((rc = step_end<R>(members[run[r]], dts[run[r]], plans[run[r]]))) return rc;
    return MPMW_OK;
}

int check_b(mpmw_batch *b) {
    if (!b) return fail(MPMW_E_INVAL, "null batch handle");
    hipError_t e = hipSetDevice(b->device);
    if (e != hipSuccess) return fail(MPMW_E_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
    return MPMW_OK;
}

}  // namespace

extern "C" {

int mpmw_batch_create(int32_t device, mpmw_batch **out) {
    if (!out) return fail(MPMW_E_INVAL, "null output");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(MPMW_E_HIP, "no HIP device");
    if (device < 0 || device >= ndev) return fail(MPMW_E_INVAL, "device ordinal out of range");
    HIP_TRY(hipSetDevice(device));
    mpmw_batch *b = new mpmw_batch();
    b->device = device;
    // a stream of its own: members being set up on pooled streams do not
    // wait for the batch's queued steps
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        delete b;
        return fail(MPMW_E_HIP, "hipStreamCreate failed");
    }
    *out = b;
    return MPMW_OK;
}

int mpmw_batch_destroy(mpmw_batch *b) {
    if (!b) return MPMW_OK;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (mpmw_solver *s : b->members) batch_detach(b, s);
    b->members.clear();
    free_staging(b->staging);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return MPMW_OK;
}

int mpmw_batch_add(mpmw_batch *b, mpmw_solver *s) {
    int rc = check_b(b);
    if (rc) return rc;
    if (!s) return fail(MPMW_E_INVAL, "null solver handle");
    if (s->batch) return fail(MPMW_E_STATE, "solver is already in a batch");
    if (s->prm.device != b->device) return fail(MPMW_E_INVAL, "solver and batch are on different devices");
    if (!b->members.empty() && s->fp64 != b->fp64)
        return fail(MPMW_E_INVAL, "a batch holds solvers of one precision");
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (b->members.empty()) b->fp64 = s->fp64;
    s->own_stream = s->stream;
    s->stream = b->stream;
    s->batch = b;
    b->members.push_back(s);
    return MPMW_OK;
}

int mpmw_batch_remove(mpmw_batch *b, mpmw_solver *s) {
    int rc = check_b(b);
    if (rc) return rc;
    auto it = std::find(b->members.begin(), b->members.end(), s);
    if (!s || it == b->members.end()) return fail(MPMW_E_INVAL, "solver is not in this batch");
    b->members.erase(it);
    batch_detach(b, s);
    return MPMW_OK;
}

int mpmw_batch_count(mpmw_batch *b, int32_t *n) {
    if (!b || !n) return fail(MPMW_E_INVAL, "null");
    *n = (int32_t)b->members.size();
    return MPMW_OK;
}

int mpmw_batch_step(mpmw_batch *b, const double *dt, int32_t n) {
    int rc = check_b(b);
    if (rc) return rc;
    if (n != (int32_t)b->members.size()) return fail(MPMW_E_INVAL, "batch_step: one dt per member");
    if (n == 0) return MPMW_OK;
    if (!dt) return fail(MPMW_E_INVAL, "null dt");
    for (int32_t i = 0; i < n; ++i)
        if (!std::isfinite(dt[i])) return fail(MPMW_E_INVAL, "batch_step: dt must be finite");
    return b->fp64 ? step_many<double>(b->members.data(), b->members.size(), dt, b->stream, b->staging)
                   : step_many<float>(b->members.data(), b->members.size(), dt, b->stream, b->staging);
}

int mpmw_batch_set_graph(mpmw_batch *b, int32_t on) {
    int rc = check_b(b);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(b->stream));
    b->staging.graph_on = on != 0;
    if (!on) drop_graphs(b->staging);
    return MPMW_OK;
}

int mpmw_batch_set_autotune(mpmw_batch *b, int32_t trials) {
    int rc = check_b(b);
    if (rc) return rc;
    return set_autotune(b->staging, trials);
}

int mpmw_batch_set_launch_config(mpmw_batch *b, const mpmw_tuning *cfg) {
    int rc = check_b(b);
    if (rc) return rc;
    return set_launch_config(b->staging, cfg);
}

int mpmw_batch_get_tuning(mpmw_batch *b, mpmw_tuning *out) {
    if (!b || !out) return fail(MPMW_E_INVAL, "null");
    get_tuning(b->staging, out);
    return MPMW_OK;
}

int mpmw_batch_graph_stats(mpmw_batch *b, int64_t out[2]) {
    if (!b || !out) return fail(MPMW_E_INVAL, "null");
    out[0] = b->staging.captures;
    out[1] = b->staging.replays;
    return MPMW_OK;
}

int mpmw_batch_synchronize(mpmw_batch *b) {
    int rc = check_b(b);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(b->stream));
    return MPMW_OK;
}

}  // extern "C"

#include "mpmw_export.hpp"
