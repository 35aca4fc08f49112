// mpmc_solver.hip -- host orchestration and C ABI (include/mpmc.h) of the
// MI355X tool FEM step (the reference's FEMSolver, src/FEMSolver.cu) and of
// the spatial-hash XPBD contact between the F-tracking particles and the
// tool's surface nodes (ContactSolver's hash path, src/ContactSolver.cu;
// SharedSpatialHash.cu) of Up25-pathan/edegpredict-engine-rotary.
//
// Differences in how, not what: structure-of-arrays tool nodes, the mesh set
// up on the host once (as the reference does) and stepped by one launch per
// reference kernel; the hash as a stable radix sort of (slot, node) pairs
// with the sorted nodes packed into 64-byte contact records; particles
// outside the tool's box (grown by the contact radius) skip the 27-cell
// probe; the node accumulators written at the original node index.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "mpmc.h"
#include "mpmc_kernels.hpp"
#include "mpms_internal.hpp"
#include "mpmw_sort.hpp"

namespace mpmw_internal {
int set_error(int code, const std::string &msg);
}

using namespace mpmc;

namespace {

int fail(int code, const std::string &msg) { return mpmw_internal::set_error(code, msg); }

#define MC_TRY(expr)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail(MPMW_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));      \
    } while (0)

// Node arrays: 30 doubles and 5 int32 per node, carved from one block.
constexpr int kNodeDoubles = 30, kNodeInts = 5;

struct NodeBlock {
    void *block = nullptr;
    Nodes v{};
    int32_t cap = 0;
};

void carve(NodeBlock &b, int32_t cap) {
    double *d = (double *)b.block;
    const size_t c = (size_t)cap;
    double **dp[] = {&b.v.x, &b.v.y, &b.v.z, &b.v.ox, &b.v.oy, &b.v.oz, &b.v.vx, &b.v.vy, &b.v.vz, &b.v.fx,
                     &b.v.fy, &b.v.fz, &b.v.mass, &b.v.T, &b.v.vm, &b.v.wear, &b.v.heat, &b.v.cond, &b.v.fn,
                     &b.v.pen, &b.v.s[0], &b.v.s[1], &b.v.s[2], &b.v.s[3], &b.v.s[4], &b.v.s[5], &b.v.sw,
                     &b.v.loff[0], &b.v.loff[1], &b.v.loff[2]};
    static_assert(sizeof(dp) / sizeof(dp[0]) == kNodeDoubles, "node doubles");
    for (int k = 0; k < kNodeDoubles; ++k) *dp[k] = d + k * c;
    int32_t *ip = (int32_t *)(d + kNodeDoubles * c);
    int32_t **ipp[] = {&b.v.status, &b.v.fixed, &b.v.in_contact, &b.v.driven, &b.v.region};
    static_assert(sizeof(ipp) / sizeof(ipp[0]) == kNodeInts, "node ints");
    for (int k = 0; k < kNodeInts; ++k) *ipp[k] = ip + k * c;
}

int alloc_nodes(NodeBlock &b, int32_t n) {
    if (n <= b.cap && b.block) {
        b.v.n = n;
        return MPMW_OK;
    }
    if (b.block) (void)hipFree(b.block);
    b.block = nullptr;
    b.cap = 0;
    const int32_t cap = std::max(n, 1);
    if (hipMalloc(&b.block, (size_t)cap * (kNodeDoubles * sizeof(double) + kNodeInts * sizeof(int32_t))) !=
        hipSuccess) {
        (void)hipGetLastError();
        return fail(MPMW_E_NOMEM, "hipMalloc of the tool nodes failed");
    }
    b.cap = cap;
    carve(b, cap);
    b.v.n = n;
    return MPMW_OK;
}

// Host image of a node set (the fields set up at initialisation).
struct HostNodes {
    std::vector<double> d[kNodeDoubles];
    std::vector<int32_t> i[kNodeInts];
    void resize(size_t n) {
        for (auto &v : d) v.assign(n, 0.0);
        for (auto &v : i) v.assign(n, 0);
    }
};
enum { HX, HY, HZ, HOX, HOY, HOZ, HVX, HVY, HVZ, HFX, HFY, HFZ, HM, HT, HVM, HWEAR, HHEAT, HCOND, HFN, HPEN,
       HS0, HSW = HS0 + 6, HLOFF };
enum { HSTATUS, HFIXED, HCONTACT, HDRIVEN, HREGION };

int upload_nodes(NodeBlock &b, const HostNodes &h, int32_t n) {
    if (n <= 0) return MPMW_OK;
    const size_t c = (size_t)b.cap;
    for (int k = 0; k < kNodeDoubles; ++k)
        MC_TRY(hipMemcpy((double *)b.block + k * c, h.d[k].data(), sizeof(double) * n, hipMemcpyHostToDevice));
    int32_t *ip = (int32_t *)((double *)b.block + kNodeDoubles * c);
    for (int k = 0; k < kNodeInts; ++k)
        MC_TRY(hipMemcpy(ip + k * c, h.i[k].data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
    return MPMW_OK;
}

}  // namespace

struct mpmt_tool {
    mpmt_params prm{};
    hipStream_t stream = nullptr;
    hipEvent_t ev = nullptr, ev_back = nullptr;
    NodeBlock surf, tet;
    Elem *d_el = nullptr;
    Embed *d_emb = nullptr;
    int32_t ns = 0, nt = 0, ne = 0, nemb = 0;
    double first_volume = 0.0;
    double tool_pos[3] = {0, 0, 0}, tool_axis[3] = {0, 0, 1}, tool_angle = 0, tool_omega = 0;
    double spindle_pos[3] = {0, 0, 0}, spindle_vel[3] = {0, 0, 0};
    double typical_len = 0.001;
    double time = 0.0;
    int64_t step = 0;
};

struct mpmc_contact {
    mpmc_params prm{};
    hipEvent_t ev = nullptr;
    uint32_t table = 0;
    int32_t cap = 0;
    uint32_t *keys = nullptr, *keys2 = nullptr;
    int32_t *idx = nullptr, *idx2 = nullptr;
    NodeRec *rec = nullptr;
    int32_t *start = nullptr, *end = nullptr;
    unsigned long long *d_misc = nullptr;   // box (6), counts (2), totals (2, as doubles), candidates (1)
    int32_t *cand = nullptr;                // particles that can touch the tool (k_contact_cull)
    int32_t cand_cap = 0;
    unsigned long long *h_misc = nullptr;   // pinned read-back of counts and totals
    mpmw::ScratchBuf temp;
    mpmc_results res{};
    int64_t calls = 0, last_contact_call = 0, age = 0;
};

namespace {

int tcheck(mpmt_tool *t) {
    if (!t) return fail(MPMW_E_INVAL, "null tool handle");
    MC_TRY(hipSetDevice(t->prm.device));
    return MPMW_OK;
}

#define MT_CHECK(t)              \
    do {                         \
        int rc_ = tcheck(t);     \
        if (rc_) return rc_;     \
    } while (0)

uint32_t next_prime(uint32_t n) {   // SharedSpatialHash::nextPrime (:117-128)
    for (;; ++n) {
        if (n < 2) continue;
        if (n == 2) return n;
        if (n % 2 == 0) continue;
        bool prime = true;
        for (uint32_t i = 3; (uint64_t)i * i <= n; i += 2)
            if (n % i == 0) { prime = false; break; }
        if (prime) return n;
    }
}

ContactParams contact_params(const mpmc_params &p) {
    ContactParams C;
    C.r = p.contact_radius; C.k = p.contact_stiffness; C.zeta = p.contact_damping_ratio; C.friction = p.friction;
    C.hp = p.heat_partition; C.shear_yield = p.workpiece_shear_yield; C.max_p = p.max_contact_pressure;
    C.A = p.jc_A; C.B = p.jc_B; C.n = p.jc_n; C.C = p.jc_C; C.m = p.jc_m; C.rate_ref = p.jc_rate_ref;
    C.T_ref = p.jc_T_ref; C.T_melt = p.jc_T_melt;
    C.tool_cp = p.tool_specific_heat; C.work_cp = p.work_specific_heat; C.work_rho = p.work_physical_density;
    C.tool_melt = p.tool_melt_temp; C.work_melt = p.work_melt_temp; C.htc = p.contact_htc;
    C.emissivity = p.emissivity; C.sb = p.stefan_boltzmann; C.ambient = p.ambient_temp;
    C.tool_rho = p.tool_physical_density; C.latent = p.latent_heat_fusion; C.solidus = p.solidus_temp;
    C.liquidus = p.liquidus_temp; C.jc_enabled = p.jc_enabled; C.pad_ = 0;
    return C;
}

int download_nodes(mpmt_tool *t, NodeBlock &b, int32_t n, mpmt_node_arrays *a) {
    MC_TRY(hipStreamSynchronize(t->stream));
    if (!a || n <= 0) return MPMW_OK;
    std::vector<double> u((size_t)n);
    auto vec3 = [&](double *dst, const double *x, const double *y, const double *z) -> int {
        if (!dst) return MPMW_OK;
        const double *src[3] = {x, y, z};
        for (int c = 0; c < 3; ++c) {
            MC_TRY(hipMemcpy(u.data(), src[c], sizeof(double) * n, hipMemcpyDeviceToHost));
            for (int32_t i = 0; i < n; ++i) dst[3 * i + c] = u[i];
        }
        return MPMW_OK;
    };
    auto one = [&](double *dst, const double *src) -> int {
        if (dst) MC_TRY(hipMemcpy(dst, src, sizeof(double) * n, hipMemcpyDeviceToHost));
        return MPMW_OK;
    };
    int rc;
    if ((rc = vec3(a->pos, b.v.x, b.v.y, b.v.z)) || (rc = vec3(a->vel, b.v.vx, b.v.vy, b.v.vz)) ||
        (rc = vec3(a->force, b.v.fx, b.v.fy, b.v.fz)) || (rc = one(a->mass, b.v.mass)) ||
        (rc = one(a->temperature, b.v.T)) || (rc = one(a->von_mises, b.v.vm)) || (rc = one(a->wear, b.v.wear)))
        return rc;
    if (a->stress)
        for (int c = 0; c < 6; ++c) {
            MC_TRY(hipMemcpy(u.data(), b.v.s[c], sizeof(double) * n, hipMemcpyDeviceToHost));
            for (int32_t i = 0; i < n; ++i) a->stress[6 * i + c] = u[i];
        }
    if (a->status) MC_TRY(hipMemcpy(a->status, b.v.status, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
    if (a->in_contact) MC_TRY(hipMemcpy(a->in_contact, b.v.in_contact, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
    return MPMW_OK;
}

// The contact pass: the particles that can touch the tool are culled into
// c->cand by a streaming pass over the workpiece, then k_contact runs on
// them (a grid of at most kContactBlocks blocks, striding over the list).
constexpr int kContactBlocks = 2048;
template <typename R> void launch_contact(mpmc_contact *c, const mpms::Particles<R> &P, Nodes S, const ContactArgs &A,
                                          int np, int32_t *ncand, hipStream_t st) {
    const int cull_blocks = (int)((np + (int64_t)kBlock * kCullPer - 1) / ((int64_t)kBlock * kCullPer));
    if (cull_blocks > 0)
        hipLaunchKernelGGL(k_contact_cull<R>, dim3(cull_blocks), dim3(kBlock), 0, st, P, c->prm.contact_radius, A.box,
                           c->cand, ncand);
    hipLaunchKernelGGL(k_contact<R>, dim3(std::max(1, std::min(blocks(np), kContactBlocks))), dim3(kBlock), 0, st, P,
                       S, contact_params(c->prm), A, (const int32_t *)c->cand, (const int32_t *)ncand);
}
template <typename R> mpms::Particles<R> contact_particles(mpms_solver *s);
template <> mpms::Particles<float> contact_particles<float>(mpms_solver *s) { return mpms_internal::particles_f(s); }
template <> mpms::Particles<double> contact_particles<double>(mpms_solver *s) { return mpms_internal::particles_d(s); }

int ensure_hash(mpmc_contact *c, int32_t nn, hipStream_t st) {
    if (nn <= c->cap) return MPMW_OK;   // ensureCapacity (:103-115)
    MC_TRY(hipStreamSynchronize(st));
    for (void *p : {(void *)c->keys, (void *)c->keys2, (void *)c->idx, (void *)c->idx2, (void *)c->rec})
        if (p) (void)hipFree(p);
    c->keys = c->keys2 = nullptr;
    c->idx = c->idx2 = nullptr;
    c->rec = nullptr;
    c->cap = 0;
    if (hipMalloc(&c->keys, sizeof(uint32_t) * nn) != hipSuccess || hipMalloc(&c->keys2, sizeof(uint32_t) * nn) != hipSuccess ||
        hipMalloc(&c->idx, sizeof(int32_t) * nn) != hipSuccess || hipMalloc(&c->idx2, sizeof(int32_t) * nn) != hipSuccess ||
        hipMalloc(&c->rec, sizeof(NodeRec) * nn) != hipSuccess) {
        (void)hipGetLastError();
        return fail(MPMW_E_NOMEM, "hipMalloc of the contact hash failed");
    }
    const uint32_t ts = next_prime((uint32_t)std::max<int64_t>((int64_t)nn * 2, 100003));
    if (ts != c->table) {
        if (c->start) (void)hipFree(c->start);
        if (c->end) (void)hipFree(c->end);
        c->start = c->end = nullptr;
        c->table = 0;
        if (hipMalloc(&c->start, sizeof(int32_t) * ts) != hipSuccess ||
            hipMalloc(&c->end, sizeof(int32_t) * ts) != hipSuccess) {
            (void)hipGetLastError();
            return fail(MPMW_E_NOMEM, "hipMalloc of the hash table failed");
        }
        c->table = ts;
    }
    c->cap = nn;
    return MPMW_OK;
}

template <typename R> int resolve_impl(mpmc_contact *c, mpms_solver *s, mpmt_tool *t, double dt) {
    const int32_t np = mpms_internal::count(s), nn = t->ns;
    const hipStream_t ms = mpms_internal::stream(s);
    const mpmc_params &p = c->prm;
    double scale = 1.0;   // engagement ramp (:1097-1112)
    if (c->age < p.engagement_ramp_steps && p.engagement_ramp_steps > 0) {
        const double u = (double)c->age / (double)p.engagement_ramp_steps;
        scale = 0.1 + 0.9 * (u * u * (3.0 - 2.0 * u));
    }
    // the tool's pending work first
    MC_TRY(hipEventRecord(t->ev, t->stream));
    MC_TRY(hipStreamWaitEvent(ms, t->ev, 0));
    int rc;
    if ((rc = ensure_hash(c, nn, ms))) return rc;
    if (np > c->cand_cap) {
        MC_TRY(hipStreamSynchronize(ms));
        if (c->cand) (void)hipFree(c->cand);
        c->cand = nullptr;
        c->cand_cap = 0;
        const int32_t cap = np + np / 4;
        if (hipMalloc(&c->cand, sizeof(int32_t) * (size_t)cap) != hipSuccess) {
            (void)hipGetLastError();
            return fail(MPMW_E_NOMEM, "hipMalloc of the contact candidates failed");
        }
        c->cand_cap = cap;
    }
    Nodes S = t->surf.v;
    S.n = nn;
    unsigned long long *box = c->d_misc, *counts = c->d_misc + 6;
    double *totals = (double *)(c->d_misc + 8);
    int32_t *ncand = (int32_t *)(c->d_misc + 10);
    // counts, totals and the candidate count start at zero
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(kBlock), 0, ms, (int32_t *)counts, 0, (int64_t)10);
    hipLaunchKernelGGL(k_reset_contact, dim3(blocks(nn)), dim3(kBlock), 0, ms, S, box);
    double o[3];
    mpms_internal::domain_min(s, o);
    hipLaunchKernelGGL(k_node_hash, dim3(blocks(nn)), dim3(kBlock), 0, ms, S, o[0], o[1], o[2], p.contact_radius,
                       c->table, c->keys, c->idx, box);
    MC_TRY(hipGetLastError());
    int bits = 1;
    while (bits < 32 && ((uint64_t)1 << bits) < (uint64_t)c->table) ++bits;
    MC_TRY(mpmw::radix_sort_pairs(c->temp, c->keys, c->keys2, c->idx, c->idx2, nn, bits, ms));
    const int fill_blocks = (int)std::min<int64_t>(blocks(c->table), 1024);
    hipLaunchKernelGGL(k_fill_i32, dim3(fill_blocks), dim3(kBlock), 0, ms, c->start, -1, (int64_t)c->table);
    hipLaunchKernelGGL(k_fill_i32, dim3(fill_blocks), dim3(kBlock), 0, ms, c->end, -1, (int64_t)c->table);
    hipLaunchKernelGGL(k_cell_bounds, dim3(blocks(nn)), dim3(kBlock), 0, ms, c->keys2, nn, c->start, c->end);
    hipLaunchKernelGGL(k_pack_records, dim3(blocks(nn)), dim3(kBlock), 0, ms, S, c->idx2, nn, c->rec);
    ContactArgs A;
    A.ox = o[0]; A.oy = o[1]; A.oz = o[2];
    A.dt = dt;
    A.max_pen = p.contact_radius * p.max_penetration_fraction;
    A.scale = scale;
    A.table = c->table;
    A.nn = nn;
    A.start = c->start; A.end = c->end; A.rec = c->rec; A.box = box;
    A.counts = counts;
    A.totals = totals;
    launch_contact<R>(c, contact_particles<R>(s), S, A, np, ncand, ms);
    hipLaunchKernelGGL(k_node_heat, dim3(blocks(nn)), dim3(kBlock), 0, ms, S, contact_params(p), dt);
    MC_TRY(hipGetLastError());
    MC_TRY(hipMemcpyAsync(c->h_misc, counts, 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost, ms));
    MC_TRY(hipStreamSynchronize(ms));   // transferResults (:1369-1376)
    double tot[2];
    std::memcpy(tot, c->h_misc + 2, sizeof(tot));
    c->res.contacts = (int64_t)c->h_misc[0];
    c->res.probed = (int64_t)c->h_misc[1];
    c->res.heat = tot[0];
    c->res.force = tot[1];
    c->res.cumulative_heat += tot[0];
    c->res.engagement = scale;
    c->res.table_size = (int32_t)c->table;
    c->calls++;   // contact age (:1254-1272)
    if (c->res.contacts > 0) {
        c->age++;
        c->last_contact_call = c->calls;
    } else if (c->calls - c->last_contact_call > 50) {
        c->age = 0;
    }
    if (c->res.contacts > (int64_t)nn * 50) {   // contact-explosion guard (:1276-1296)
        hipLaunchKernelGGL(k_reset_contact, dim3(blocks(nn)), dim3(kBlock), 0, ms, S, box);
        MC_TRY(hipGetLastError());
        c->res.contacts = 0;
        c->res.force = 0.0;
        c->res.heat = 0.0;
    }
    // the tool's later work after this
    MC_TRY(hipEventRecord(c->ev, ms));
    MC_TRY(hipStreamWaitEvent(t->stream, c->ev, 0));
    return MPMW_OK;
}

}  // namespace

extern "C" {

void mpmt_params_default(mpmt_params *p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->young = 600e9;             // FEMMaterialProps (FEMSolver.cuh:42-47)
    p->poisson = 0.22;
    p->density = 14500.0;
    p->thermal_conductivity = 80.0;
    p->specific_heat = 200.0;
    p->damping_ratio = 0.1;       // m_globalDamping (:235)
    p->mass_scaling = 1.0;
    p->coolant_htc = 1.0e3;       // :271
    p->spindle_stiffness = 5e7;   // :265-266
    p->spindle_damping = 1e4;
    p->spindle_enabled = 1;       // :264
    p->thermal_skip = 10;         // :205
    p->polar_textbook = 0;
    p->device = 0;
}

int mpmt_create(const mpmt_params *p, mpmt_tool **out) {
    if (!p || !out) return fail(MPMW_E_INVAL, "null argument");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        (void)hipGetLastError();
        return fail(MPMW_E_HIP, "no HIP device: the tool FEM runs on the GPU only");
    }
    if (p->device < 0 || p->device >= ndev) return fail(MPMW_E_INVAL, "device ordinal out of range");
    if (p->thermal_skip < 1) return fail(MPMW_E_INVAL, "thermal_skip must be at least 1");
    MC_TRY(hipSetDevice(p->device));
    mpmt_tool *t = new mpmt_tool();
    t->prm = *p;
    if (hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&t->ev, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&t->ev_back, hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        delete t;
        return fail(MPMW_E_HIP, "tool FEM: stream set-up failed");
    }
    *out = t;
    return MPMW_OK;
}

int mpmt_destroy(mpmt_tool *t) {
    if (!t) return MPMW_OK;
    (void)hipSetDevice(t->prm.device);
    if (t->stream) (void)hipStreamSynchronize(t->stream);
    for (void *q : {t->surf.block, t->tet.block, (void *)t->d_el, (void *)t->d_emb})
        if (q) (void)hipFree(q);
    if (t->ev) (void)hipEventDestroy(t->ev);
    if (t->ev_back) (void)hipEventDestroy(t->ev_back);
    if (t->stream) (void)hipStreamDestroy(t->stream);
    delete t;
    return MPMW_OK;
}

int mpmt_set_mesh(mpmt_tool *t, const mpmt_mesh *m) {
    MT_CHECK(t);
    if (!m || m->n_surface < 0 || m->n_tet_nodes < 0 || m->n_elements < 0 || m->n_embeds < 0)
        return fail(MPMW_E_INVAL, "bad mesh");
    const int32_t ns = m->n_surface, nt = m->n_tet_nodes, ne = m->n_elements, nb = m->n_embeds;
    if ((ns > 0 && !m->surface_xyz) || (nt > 0 && !m->tet_xyz) || (ne > 0 && !m->elements) ||
        (nb > 0 && (!m->embed_surface || !m->embed_tets || !m->embed_weights)))
        return fail(MPMW_E_INVAL, "mesh arrays missing");
    for (int64_t k = 0; k < 4 * (int64_t)ne; ++k)
        if (m->elements[k] < 0 || m->elements[k] >= nt) return fail(MPMW_E_RANGE, "element node index out of range");
    for (int32_t e = 0; e < nb; ++e) {
        if (m->embed_surface[e] >= ns) return fail(MPMW_E_RANGE, "embed surface index out of range");
        for (int k = 0; k < 4; ++k)
            if (m->embed_tets[4 * e + k] >= nt) return fail(MPMW_E_RANGE, "embed tet index out of range");
    }
    MC_TRY(hipStreamSynchronize(t->stream));
    const mpmt_params &P = t->prm;
    // initializeFromMesh (FEMSolver.cu:900-1010)
    HostNodes S, T;
    S.resize(ns);
    T.resize(nt);
    for (int32_t i = 0; i < ns; ++i) {
        for (int a = 0; a < 3; ++a) S.d[HX + a][i] = S.d[HOX + a][i] = m->surface_xyz[3 * i + a];
        S.d[HM][i] = P.density * 1e-9 * P.mass_scaling;
        S.d[HT][i] = 25.0;
        S.i[HREGION][i] = m->surface_region ? m->surface_region[i] : 0;
    }
    double maxz = -1e10;
    for (int32_t i = 0; i < nt; ++i) {
        for (int a = 0; a < 3; ++a) T.d[HX + a][i] = T.d[HOX + a][i] = m->tet_xyz[3 * i + a];
        T.d[HT][i] = 25.0;
        T.i[HREGION][i] = m->tet_region ? m->tet_region[i] : 0;
        maxz = std::max(maxz, T.d[HOZ][i]);
    }
    for (int32_t i = 0; i < nt; ++i) T.i[HFIXED][i] = T.d[HOZ][i] > maxz - 0.005 ? 1 : 0;
    // createElementsFromMesh (:1012-1078)
    std::vector<Elem> el((size_t)ne);
    for (int32_t e = 0; e < ne; ++e) {
        Elem &E = el[e];
        std::memset(&E, 0, sizeof(E));
        for (int k = 0; k < 4; ++k) E.n[k] = m->elements[4 * e + k];
        double X[4][3];
        for (int k = 0; k < 4; ++k)
            for (int a = 0; a < 3; ++a) X[k][a] = T.d[HOX + a][E.n[k]];
        double D[3][3];
        auto fillD = [&]() {
            for (int i = 0; i < 3; ++i) { D[i][0] = X[1][i] - X[0][i]; D[i][1] = X[2][i] - X[0][i]; D[i][2] = X[3][i] - X[0][i]; }
        };
        fillD();
        double det = D[0][0] * (D[1][1] * D[2][2] - D[1][2] * D[2][1]) - D[0][1] * (D[1][0] * D[2][2] - D[1][2] * D[2][0]) +
                     D[0][2] * (D[1][0] * D[2][1] - D[1][1] * D[2][0]);
        if (det < 0.0) {
            std::swap(E.n[1], E.n[2]);
            for (int a = 0; a < 3; ++a) std::swap(X[1][a], X[2][a]);
            fillD();
            det = -det;
        }
        E.volume = det / 6.0;
        if (E.volume < 1e-18) { E.volume = 1e-18; det = 6e-18; }
        const double id = 1.0 / det;
        double *M = E.inv_dm;
        M[0] = (D[1][1] * D[2][2] - D[1][2] * D[2][1]) * id;
        M[1] = (D[0][2] * D[2][1] - D[0][1] * D[2][2]) * id;
        M[2] = (D[0][1] * D[1][2] - D[0][2] * D[1][1]) * id;
        M[3] = (D[1][2] * D[2][0] - D[1][0] * D[2][2]) * id;
        M[4] = (D[0][0] * D[2][2] - D[0][2] * D[2][0]) * id;
        M[5] = (D[0][2] * D[1][0] - D[0][0] * D[1][2]) * id;
        M[6] = (D[1][0] * D[2][1] - D[1][1] * D[2][0]) * id;
        M[7] = (D[0][1] * D[2][0] - D[0][0] * D[2][1]) * id;
        M[8] = (D[0][0] * D[1][1] - D[0][1] * D[1][0]) * id;
        double mpn = (P.density * E.volume * P.mass_scaling) / 4.0;
        if (mpn < 1e-12) mpn = 1e-12;
        for (int k = 0; k < 4; ++k) T.d[HM][E.n[k]] += mpn;
    }
    std::vector<Embed> emb((size_t)nb);
    for (int32_t e = 0; e < nb; ++e) {
        Embed &E = emb[e];
        std::memset(&E, 0, sizeof(E));
        E.surf = m->embed_surface[e];
        for (int k = 0; k < 4; ++k) { E.tet[k] = m->embed_tets[4 * e + k]; E.w[k] = m->embed_weights[4 * e + k]; }
        if (E.surf < 0) continue;   // surface masses from the embeds (:973-992)
        double total = 0.0;
        for (int k = 0; k < 4; ++k)
            if (E.tet[k] >= 0) total += T.d[HM][E.tet[k]] * E.w[k];
        if (total > 1e-15) S.d[HM][E.surf] = total;
    }
    int rc;
    if ((rc = alloc_nodes(t->surf, ns)) || (rc = alloc_nodes(t->tet, nt))) return rc;
    if ((rc = upload_nodes(t->surf, S, ns)) || (rc = upload_nodes(t->tet, T, nt))) return rc;
    for (void *q : {(void *)t->d_el, (void *)t->d_emb})
        if (q) (void)hipFree(q);
    t->d_el = nullptr;
    t->d_emb = nullptr;
    if (hipMalloc(&t->d_el, sizeof(Elem) * std::max<int32_t>(ne, 1)) != hipSuccess ||
        hipMalloc(&t->d_emb, sizeof(Embed) * std::max<int32_t>(nb, 1)) != hipSuccess) {
        (void)hipGetLastError();
        return fail(MPMW_E_NOMEM, "hipMalloc of the tool elements failed");
    }
    if (ne > 0) MC_TRY(hipMemcpy(t->d_el, el.data(), sizeof(Elem) * ne, hipMemcpyHostToDevice));
    if (nb > 0) MC_TRY(hipMemcpy(t->d_emb, emb.data(), sizeof(Embed) * nb, hipMemcpyHostToDevice));
    t->ns = ns; t->nt = nt; t->ne = ne; t->nemb = nb;
    t->first_volume = ne > 0 ? el[0].volume : 0.0;
    return MPMW_OK;
}

int mpmt_init_driven_nodes(mpmt_tool *t, double top_fraction) {   // :1367-1408
    MT_CHECK(t);
    if (t->ns == 0 || t->nt == 0) return MPMW_OK;
    MC_TRY(hipStreamSynchronize(t->stream));
    const int32_t nt = t->nt;
    std::vector<double> z(nt), o[3] = {std::vector<double>(nt), std::vector<double>(nt), std::vector<double>(nt)};
    std::vector<int32_t> fixed(nt), driven(nt);
    const Nodes &T = t->tet.v;
    MC_TRY(hipMemcpy(z.data(), T.z, sizeof(double) * nt, hipMemcpyDeviceToHost));
    MC_TRY(hipMemcpy(o[0].data(), T.ox, sizeof(double) * nt, hipMemcpyDeviceToHost));
    MC_TRY(hipMemcpy(o[1].data(), T.oy, sizeof(double) * nt, hipMemcpyDeviceToHost));
    MC_TRY(hipMemcpy(o[2].data(), T.oz, sizeof(double) * nt, hipMemcpyDeviceToHost));
    MC_TRY(hipMemcpy(fixed.data(), T.fixed, sizeof(int32_t) * nt, hipMemcpyDeviceToHost));
    double mn = 1e30, mx = -1e30;
    for (double v : z) { mn = std::min(mn, v); mx = std::max(mx, v); }
    const double thr = mx - (mx - mn) * top_fraction;
    std::vector<double> loff[3] = {std::vector<double>(nt, 0.0), std::vector<double>(nt, 0.0),
                                   std::vector<double>(nt, 0.0)};
    MC_TRY(hipMemcpy(loff[0].data(), T.loff[0], sizeof(double) * nt, hipMemcpyDeviceToHost));
    MC_TRY(hipMemcpy(loff[1].data(), T.loff[1], sizeof(double) * nt, hipMemcpyDeviceToHost));
    MC_TRY(hipMemcpy(loff[2].data(), T.loff[2], sizeof(double) * nt, hipMemcpyDeviceToHost));
    for (int32_t i = 0; i < nt; ++i) {
        driven[i] = z[i] >= thr ? 1 : 0;
        if (driven[i])
            for (int a = 0; a < 3; ++a) loff[a][i] = o[a][i];
        if (t->prm.spindle_enabled) fixed[i] = 0;
    }
    MC_TRY(hipMemcpy(T.driven, driven.data(), sizeof(int32_t) * nt, hipMemcpyHostToDevice));
    MC_TRY(hipMemcpy(T.fixed, fixed.data(), sizeof(int32_t) * nt, hipMemcpyHostToDevice));
    for (int a = 0; a < 3; ++a)
        MC_TRY(hipMemcpy(T.loff[a], loff[a].data(), sizeof(double) * nt, hipMemcpyHostToDevice));
    return MPMW_OK;
}

int mpmt_set_spindle_state(mpmt_tool *t, const double pos[3], const double vel[3]) {
    MT_CHECK(t);
    if (!pos || !vel) return fail(MPMW_E_INVAL, "null");
    for (int a = 0; a < 3; ++a) { t->spindle_pos[a] = pos[a]; t->spindle_vel[a] = vel[a]; }
    return MPMW_OK;
}

int mpmt_set_transform(mpmt_tool *t, const double pos[3], const double axis[3], double omega) {
    MT_CHECK(t);
    if (!pos || !axis) return fail(MPMW_E_INVAL, "null");
    const double l = std::sqrt(axis[0] * axis[0] + axis[1] * axis[1] + axis[2] * axis[2]);
    for (int a = 0; a < 3; ++a) {
        t->tool_pos[a] = pos[a];
        t->tool_axis[a] = l > 1e-12 ? axis[a] / l : 0.0;   // Vec3::normalized (Types.h:185-188)
    }
    t->tool_omega = omega;
    return MPMW_OK;
}

int mpmt_set_rigid_velocity(mpmt_tool *t, const double v[3], double omega_z, double cx, double cy) {
    MT_CHECK(t);
    if (!v) return fail(MPMW_E_INVAL, "null");
    if (t->ns > 0)
        hipLaunchKernelGGL(k_rigid_velocity, dim3(blocks(t->ns)), dim3(kBlock), 0, t->stream, t->surf.v, v[0], v[1],
                           v[2], omega_z, cx, cy);
    if (t->nt > 0)
        hipLaunchKernelGGL(k_rigid_velocity, dim3(blocks(t->nt)), dim3(kBlock), 0, t->stream, t->tet.v, v[0], v[1],
                           v[2], omega_z, cx, cy);
    MC_TRY(hipGetLastError());
    return MPMW_OK;
}

int mpmt_apply_contact_force(mpmt_tool *t, int32_t node, const double f[3], double heat) {   // :1524-1536
    MT_CHECK(t);
    if (!f) return fail(MPMW_E_INVAL, "null");
    if (node < 0 || node >= t->ns) return MPMW_OK;
    MC_TRY(hipStreamSynchronize(t->stream));
    const Nodes &S = t->surf.v;
    double fx, fy, fz, T, m;
    MC_TRY(hipMemcpy(&fx, S.fx + node, sizeof(double), hipMemcpyDeviceToHost));
    MC_TRY(hipMemcpy(&fy, S.fy + node, sizeof(double), hipMemcpyDeviceToHost));
    MC_TRY(hipMemcpy(&fz, S.fz + node, sizeof(double), hipMemcpyDeviceToHost));
    MC_TRY(hipMemcpy(&T, S.T + node, sizeof(double), hipMemcpyDeviceToHost));
    MC_TRY(hipMemcpy(&m, S.mass + node, sizeof(double), hipMemcpyDeviceToHost));
    fx += f[0]; fy += f[1]; fz += f[2];
    T += heat / (m * t->prm.specific_heat);
    const int32_t one = 1;
    MC_TRY(hipMemcpy(S.fx + node, &fx, sizeof(double), hipMemcpyHostToDevice));
    MC_TRY(hipMemcpy(S.fy + node, &fy, sizeof(double), hipMemcpyHostToDevice));
    MC_TRY(hipMemcpy(S.fz + node, &fz, sizeof(double), hipMemcpyHostToDevice));
    MC_TRY(hipMemcpy(S.T + node, &T, sizeof(double), hipMemcpyHostToDevice));
    MC_TRY(hipMemcpy(S.in_contact + node, &one, sizeof(int32_t), hipMemcpyHostToDevice));
    return MPMW_OK;
}

int mpmt_step(mpmt_tool *t, double dt) {   // FEMSolver::step (:1112-1300)
    MT_CHECK(t);
    if (!std::isfinite(dt)) return fail(MPMW_E_INVAL, "dt must be finite");
    if (t->ns == 0) return MPMW_OK;
    const mpmt_params &P = t->prm;
    const hipStream_t st = t->stream;
    const Nodes S = t->surf.v, T = t->tet.v;
    const int ns = t->ns, nt = t->nt, ne = t->ne, nb = t->nemb;
    if (std::fabs(t->tool_omega) > 1e-12) {
        const double inc = t->tool_omega * dt;
        t->tool_angle += inc;
        hipLaunchKernelGGL(k_rotate, dim3(blocks(ns)), dim3(kBlock), 0, st, S, t->tool_pos[0], t->tool_pos[1],
                           t->tool_pos[2], t->tool_axis[0], t->tool_axis[1], t->tool_axis[2], inc);
        if (nt > 0)
            hipLaunchKernelGGL(k_rotate, dim3(blocks(nt)), dim3(kBlock), 0, st, T, t->tool_pos[0], t->tool_pos[1],
                               t->tool_pos[2], t->tool_axis[0], t->tool_axis[1], t->tool_axis[2], inc);
    }
    if (P.spindle_enabled) {
        if (nb > 0) hipLaunchKernelGGL(k_transfer_forces, dim3(blocks(nb)), dim3(kBlock), 0, st, S, T, t->d_emb, nb);
        if (nt > 0) {
            hipLaunchKernelGGL(k_spindle, dim3(blocks(nt)), dim3(kBlock), 0, st, T, t->spindle_pos[0],
                               t->spindle_pos[1], t->spindle_pos[2], t->spindle_vel[0], t->spindle_vel[1],
                               t->spindle_vel[2], P.spindle_stiffness, P.spindle_damping);
            hipLaunchKernelGGL(k_reset_stress, dim3(blocks(nt)), dim3(kBlock), 0, st, T);
        }
        if (ne > 0)
            hipLaunchKernelGGL(k_elem_forces, dim3(blocks(ne)), dim3(kBlock), 0, st, T, t->d_el, ne, P.young,
                               P.poisson, P.polar_textbook);
        if (nt > 0) {
            hipLaunchKernelGGL(k_norm_stress, dim3(blocks(nt)), dim3(kBlock), 0, st, T);
            hipLaunchKernelGGL(k_integrate, dim3(blocks(nt)), dim3(kBlock), 0, st, T, dt, P.damping_ratio * 0.01);
        }
        if (nb > 0) hipLaunchKernelGGL(k_update_surface, dim3(blocks(nb)), dim3(kBlock), 0, st, S, T, t->d_emb, nb);
    }
    if (t->step % 100 == 0 && ne > 0) t->typical_len = std::pow(t->first_volume, 1.0 / 3.0);
    if (nb > 0 && nt > 0) {
        hipLaunchKernelGGL(k_surface_heat, dim3(blocks(nb)), dim3(kBlock), 0, st, S, T, t->d_emb, nb,
                           P.specific_heat, 0.35);
        hipLaunchKernelGGL(k_apply_tet_heat, dim3(blocks(nt)), dim3(kBlock), 0, st, T, P.specific_heat);
    }
    if (ne > 0 && t->step % P.thermal_skip == 0) {
        const double alpha = (P.density > 0.0 && P.specific_heat > 0.0)
                                 ? P.thermal_conductivity / (P.density * P.specific_heat) : 0.0;
        const double h = std::fmax(t->typical_len, 0.0001);
        const double dtt = dt * P.thermal_skip;
        int nsub = 1;
        if (alpha > 0.0) {
            const double cfl = alpha * dtt / (h * h);
            if (cfl > 0.25) nsub = (int)std::ceil(cfl / 0.25);
        }
        const double dts = dtt / nsub;
        for (int sub = 0; sub < nsub; ++sub) {
            hipLaunchKernelGGL(k_conduction, dim3(blocks(ne)), dim3(kBlock), 0, st, T, t->d_el, ne,
                               P.thermal_conductivity, P.specific_heat, dts);
            hipLaunchKernelGGL(k_apply_tet_heat, dim3(blocks(nt)), dim3(kBlock), 0, st, T, P.specific_heat);
        }
    }
    if (nt > 0)
        hipLaunchKernelGGL(k_thermal_bc, dim3(blocks(nt)), dim3(kBlock), 0, st, T, P.coolant_htc, P.density,
                           P.specific_heat, dt);
    hipLaunchKernelGGL(k_wear, dim3(blocks(ns)), dim3(kBlock), 0, st, S, dt);
    MC_TRY(hipGetLastError());
    t->time += dt;
    t->step++;
    return MPMW_OK;
}

int mpmt_synchronize(mpmt_tool *t) {
    MT_CHECK(t);
    MC_TRY(hipStreamSynchronize(t->stream));
    return MPMW_OK;
}

int mpmt_counts(mpmt_tool *t, int32_t out[4]) {
    MT_CHECK(t);
    if (!out) return fail(MPMW_E_INVAL, "null");
    out[0] = t->ns; out[1] = t->nt; out[2] = t->ne; out[3] = t->nemb;
    return MPMW_OK;
}

int mpmt_get_surface(mpmt_tool *t, mpmt_node_arrays *a) {
    MT_CHECK(t);
    return download_nodes(t, t->surf, t->ns, a);
}

int mpmt_get_tet_nodes(mpmt_tool *t, mpmt_node_arrays *a) {
    MT_CHECK(t);
    return download_nodes(t, t->tet, t->nt, a);
}

int mpmt_set_surface(mpmt_tool *t, const double *temperature, const double *force) {
    MT_CHECK(t);
    MC_TRY(hipStreamSynchronize(t->stream));
    const int32_t n = t->ns;
    if (n <= 0) return MPMW_OK;
    const Nodes &S = t->surf.v;
    if (temperature) MC_TRY(hipMemcpy(S.T, temperature, sizeof(double) * n, hipMemcpyHostToDevice));
    if (force) {
        std::vector<double> u((size_t)n);
        double *dst[3] = {S.fx, S.fy, S.fz};
        for (int c = 0; c < 3; ++c) {
            for (int32_t i = 0; i < n; ++i) u[i] = force[3 * i + c];
            MC_TRY(hipMemcpy(dst[c], u.data(), sizeof(double) * n, hipMemcpyHostToDevice));
        }
    }
    return MPMW_OK;
}

void mpmc_params_default(mpmc_params *p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->contact_radius = 0.0002;   // ContactConfig (ContactSolver.cuh:29-71)
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
    p->device = 0;
}

int mpmc_create(const mpmc_params *p, mpmc_contact **out) {
    if (!p || !out) return fail(MPMW_E_INVAL, "null argument");
    *out = nullptr;
    if (!(p->contact_radius > 0.0)) return fail(MPMW_E_INVAL, "contact_radius must be positive");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        (void)hipGetLastError();
        return fail(MPMW_E_HIP, "no HIP device: the contact runs on the GPU only");
    }
    if (p->device < 0 || p->device >= ndev) return fail(MPMW_E_INVAL, "device ordinal out of range");
    MC_TRY(hipSetDevice(p->device));
    mpmc_contact *c = new mpmc_contact();
    c->prm = *p;
    if (hipEventCreateWithFlags(&c->ev, hipEventDisableTiming) != hipSuccess ||
        hipMalloc(&c->d_misc, 11 * sizeof(unsigned long long)) != hipSuccess ||
        hipHostMalloc(&c->h_misc, 4 * sizeof(unsigned long long)) != hipSuccess) {
        (void)hipGetLastError();
        if (c->ev) (void)hipEventDestroy(c->ev);
        if (c->d_misc) (void)hipFree(c->d_misc);
        delete c;
        return fail(MPMW_E_NOMEM, "contact set-up failed");
    }
    *out = c;
    return MPMW_OK;
}

int mpmc_destroy(mpmc_contact *c) {
    if (!c) return MPMW_OK;
    (void)hipSetDevice(c->prm.device);
    (void)hipDeviceSynchronize();
    for (void *q : {(void *)c->keys, (void *)c->keys2, (void *)c->idx, (void *)c->idx2, (void *)c->rec,
                    (void *)c->start, (void *)c->end, (void *)c->d_misc, (void *)c->cand})
        if (q) (void)hipFree(q);
    if (c->h_misc) (void)hipHostFree(c->h_misc);
    mpmw::free_scratch(c->temp);
    if (c->ev) (void)hipEventDestroy(c->ev);
    delete c;
    return MPMW_OK;
}

int mpmc_set_params(mpmc_contact *c, const mpmc_params *p) {
    if (!c || !p) return fail(MPMW_E_INVAL, "null argument");
    if (!(p->contact_radius > 0.0)) return fail(MPMW_E_INVAL, "contact_radius must be positive");
    if (p->device != c->prm.device) return fail(MPMW_E_INVAL, "the device of a contact cannot change");
    c->prm = *p;
    return MPMW_OK;
}

int mpmc_resolve(mpmc_contact *c, mpms_solver *s, mpmt_tool *t, double dt) {
    if (!c || !s || !t) return fail(MPMW_E_INVAL, "null handle");
    if (mpms_internal::device(s) != c->prm.device || t->prm.device != c->prm.device)
        return fail(MPMW_E_INVAL, "the contact, the solver and the tool must be on one device");
    MC_TRY(hipSetDevice(c->prm.device));
    if (!(dt > 0.0)) return MPMW_OK;   // :1086-1091 silent return
    if (mpms_internal::count(s) <= 0 || t->ns <= 0) return MPMW_OK;
    return mpms_internal::is_fp64(s) ? resolve_impl<double>(c, s, t, dt) : resolve_impl<float>(c, s, t, dt);
}

int mpmc_get_results(mpmc_contact *c, mpmc_results *r) {
    if (!c || !r) return fail(MPMW_E_INVAL, "null argument");
    *r = c->res;
    return MPMW_OK;
}

int mpmc_link_tool_nodes(mpms_solver *s, mpmt_tool *t) {
    if (!s || !t) return fail(MPMW_E_INVAL, "null handle");
    if (mpms_internal::device(s) != t->prm.device) return fail(MPMW_E_INVAL, "the solver and the tool must be on one device");
    MC_TRY(hipSetDevice(t->prm.device));
    double *xyz = nullptr, *vel = nullptr;
    int rc = mpms_internal::tool_node_buffers(s, t->ns, &xyz, &vel);
    if (rc) return rc;
    if (t->ns <= 0) return MPMW_OK;
    const hipStream_t ms = mpms_internal::stream(s);
    MC_TRY(hipEventRecord(t->ev, t->stream));
    MC_TRY(hipStreamWaitEvent(ms, t->ev, 0));
    hipLaunchKernelGGL(k_link_tool, dim3(blocks(t->ns)), dim3(kBlock), 0, ms, t->surf.v, xyz, vel);
    MC_TRY(hipGetLastError());
    // the tool's next step moves the nodes this copy reads
    MC_TRY(hipEventRecord(t->ev_back, ms));
    MC_TRY(hipStreamWaitEvent(t->stream, t->ev_back, 0));
    return MPMW_OK;
}

}  // extern "C"
