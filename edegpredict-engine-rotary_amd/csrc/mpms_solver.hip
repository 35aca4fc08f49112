// mpms_solver.hip -- host orchestration and C ABI (include/mpms.h) of the
// MI355X F-tracking MPM step, the reference's MPMSolver (src/MPMSolver.cu of
// Up25-pathan/edegpredict-engine-rotary).
//
// Differences in how, not what: particle state in structure-of-arrays (FP32
// bulk with FP64 positions and FP64 islands, or FP64 throughout), the grid
// held only on a sub-box of the reference's lattice (refitted with the
// particle sort every 25 steps), LOD classification folded into the P2G and
// the external-force reset and chip pass folded into the G2P, so a step is
// four launches (five with a tool) and runs ahead of the host between fits.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "mpms.h"
#include "mpmw_sort.hpp"
#include "mpms_kernels.hpp"
#include "mpms_internal.hpp"

namespace mpmw_internal {
int set_error(int code, const std::string &msg);
}

using namespace mpms;

namespace {

int fail(int code, const std::string &msg) { return mpmw_internal::set_error(code, msg); }

#define MS_TRY(expr)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail(MPMW_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));      \
    } while (0)

enum Field {
    X, Y, Z, VX, VY, VZ, EFX, EFY, EFZ, M, V, T,
    F0, FP0 = F0 + 9, C0 = FP0 + 9, S0 = C0 + 9, A0 = S0 + 6, EPS = A0 + 6, RATE, DMG, SRES,
    ID, STATUS, LOD, NFIELD
};
constexpr int kFirstInt = ID;
constexpr int kFitEvery = 25;     // steps between grid fits (and particle sorts)
constexpr int kMargin = 4;        // nodes of margin around the particles' stencils
constexpr double kResortFraction = 0.05;   // particles outside their tile window that trigger a re-sort
constexpr int kTailBlocks = 64;   // G2P blocks over the particles after every tile (INACTIVE at the sort)
constexpr int kStatBlocks = 1024;

inline int blocks(int64_t n) { return (int)std::max<int64_t>(1, (n + kBlock - 1) / kBlock); }

}  // namespace

struct mpms_solver {
    mpms_params prm{};
    bool fp64 = false;
    size_t rsz = 4;
    hipStream_t stream = nullptr;
    int32_t n = 0, cap = 0;
    void *f[NFIELD] = {};
    void *alt[NFIELD] = {};
    // grid
    int32_t dims[3] = {0, 0, 0}, lo[3] = {0, 0, 0}, nb[3] = {1, 1, 1};
    int64_t nn = 0, nn_cap = 0;
    void *gblock = nullptr;           // mass, p xyz, f xyz, v xyz (R), claim (int32)
    bool fitted = false;
    unsigned long long *d_counts = nullptr;   // [0] overflow, [1] island kinematics, [2] island Newton,
                                              // [3] P2G particles outside their tile's window (last step)
    // tool
    double *d_tool_xyz = nullptr, *d_tool_vel = nullptr;
    int32_t ntool = 0, tool_cap = 0;
    double tool_pos[3] = {0, 0, 0};
    // stepping
    double time = 0.0;
    int64_t step = 0, fits = 0, sorts = 0;
    int64_t last_fit = 0;             // step of the last fit
    // reductions and sort scratch
    double *d_part = nullptr;
    uint32_t *keys = nullptr, *keys2 = nullptr;
    int32_t *idx = nullptr, *idx2 = nullptr;
    int32_t sort_cap = 0;
    // the P2G's base-node tiles: first particle of each (ntiles + 1)
    int32_t *tile_start = nullptr;
    int32_t ntiles = 0, tile_cap = 0;
    void *tpart = nullptr;            // the tiles' P2G windows (kWN records of 8 channels each)
    int64_t tpart_cap = 0;            // bytes
    mpmw::ScratchBuf temp;
    // the step's scalars on the device (k_step_args) and its captured
    // launches: one graph per launch configuration, the two most recent kept
    // (the fields swap with their sorted copies at every fit)
    StepArgs *d_sa = nullptr;
    bool graph_on = false;
    struct Captured {
        std::vector<uint64_t> key;
        hipGraphExec_t exec = nullptr;
    } graphs[2];
    int64_t captures = 0, replays = 0;
};

namespace {

size_t field_bytes(const mpms_solver *s, int k, int64_t count) {
    if (k <= Z) return sizeof(double) * (size_t)count;
    if (k >= kFirstInt) return sizeof(int32_t) * (size_t)count;
    return s->rsz * (size_t)count;
}

template <typename R> Particles<R> pview(mpms_solver *s) {
    Particles<R> P;
    P.x = (double *)s->f[X]; P.y = (double *)s->f[Y]; P.z = (double *)s->f[Z];
    for (int a = 0; a < 3; ++a) { P.v[a] = (R *)s->f[VX + a]; P.ef[a] = (R *)s->f[EFX + a]; }
    P.m = (R *)s->f[M]; P.V = (R *)s->f[V]; P.T = (R *)s->f[T];
    for (int c = 0; c < 9; ++c) { P.F[c] = (R *)s->f[F0 + c]; P.Fp[c] = (R *)s->f[FP0 + c]; P.C[c] = (R *)s->f[C0 + c]; }
    for (int c = 0; c < 6; ++c) { P.s[c] = (R *)s->f[S0 + c]; P.a[c] = (R *)s->f[A0 + c]; }
    P.eps = (R *)s->f[EPS]; P.rate = (R *)s->f[RATE]; P.dmg = (R *)s->f[DMG]; P.sres = (R *)s->f[SRES];
    P.id = (int32_t *)s->f[ID]; P.status = (int32_t *)s->f[STATUS]; P.lod = (int32_t *)s->f[LOD];
    P.n = s->n;
    return P;
}

template <typename R> Grid<R> gview(mpms_solver *s) {
    Grid<R> g;
    g.ox = s->prm.domain_min[0]; g.oy = s->prm.domain_min[1]; g.oz = s->prm.domain_min[2];
    g.dx = s->prm.dx;
    g.inv_dx = 1.0 / s->prm.dx;
    for (int a = 0; a < 3; ++a) { g.dims[a] = s->dims[a]; g.lo[a] = s->lo[a]; g.nb[a] = s->nb[a]; }
    g.nn = s->nn;
    R *b = (R *)s->gblock;
    const int64_t c = s->nn_cap;
    g.mass = b;
    for (int a = 0; a < 3; ++a) { g.p[a] = b + (1 + a) * c; g.f[a] = b + (4 + a) * c; g.v[a] = b + (7 + a) * c; }
    g.claim = (int32_t *)(b + 10 * c);
    g.overflow = s->d_counts;
    return g;
}

template <typename R> Mat<R> material(const mpms_solver *s) {
    const mpms_params &p = s->prm;
    Mat<R> m;
    const double mu = p.young / (2.0 * (1.0 + p.poisson));
    const double lambda = p.young * p.poisson / ((1.0 + p.poisson) * (1.0 - 2.0 * p.poisson));
    m.mu = (R)mu; m.lambda = (R)lambda;
    m.rho_phys = (R)p.density; m.cp = (R)p.specific_heat; m.tq = (R)p.taylor_quinney;
    m.melt = (R)p.melt_temp; m.ambient = (R)p.ambient_temp;
    m.A = (R)p.jc_A; m.B = (R)p.jc_B; m.n = (R)p.jc_n; m.C = (R)p.jc_C; m.m = (R)p.jc_m;
    m.H = (R)p.kinematic_hardening;
    m.D1 = (R)p.jc_D1; m.D2 = (R)p.jc_D2; m.D3 = (R)p.jc_D3; m.D4 = (R)p.jc_D4; m.D5 = (R)p.jc_D5;
    m.dth = (R)p.damage_threshold; m.ref_rate = (R)p.ref_strain_rate;
    m.mu_d = mu; m.lambda_d = lambda;
    m.A_d = p.jc_A; m.B_d = p.jc_B; m.n_d = p.jc_n; m.C_d = p.jc_C; m.m_d = p.jc_m;
    m.corotated = p.elastic_model == MPMS_ELASTIC_COROTATED ? 1 : 0;
    m.newton = p.return_map == MPMS_RETURN_NEWTON ? 1 : 0;
    m.newton_fp64 = p.newton_fp64;
    m.polar_fp64 = p.polar_fp64;
    m.island_det = p.island_det;
    return m;
}

int check(mpms_solver *s) {
    if (!s) return fail(MPMW_E_INVAL, "null solver handle");
    MS_TRY(hipSetDevice(s->prm.device));
    return MPMW_OK;
}

#define MS_CHECK(s)              \
    do {                         \
        int rc_ = check(s);      \
        if (rc_) return rc_;     \
    } while (0)

void free_fields(void **f, const mpms_solver *s) {
    (void)s;
    for (int k = 0; k < NFIELD; ++k) {
        if (f[k]) (void)hipFree(f[k]);
        f[k] = nullptr;
    }
}

int alloc_fields(mpms_solver *s, void **f, int32_t count) {
    for (int k = 0; k < NFIELD; ++k) {
        if (hipMalloc(&f[k], std::max<size_t>(field_bytes(s, k, count), 256)) != hipSuccess) {
            (void)hipGetLastError();
            free_fields(f, s);
            return fail(MPMW_E_NOMEM, "hipMalloc of the particle fields failed");
        }
    }
    return MPMW_OK;
}

// Upload `count` values of field k from FP64 (or int32) host data; `shift`
// is subtracted first (the identity's entry for F and Fp, held as F - I).
int upload(mpms_solver *s, int k, const void *src, int32_t count, int comp = 1, int c = 0, double shift = 0.0) {
    if (count <= 0) return MPMW_OK;
    if (shift != 0.0) {
        std::vector<double> t(count);
        for (int32_t i = 0; i < count; ++i) t[i] = ((const double *)src)[(size_t)i * comp + c] - shift;
        return upload(s, k, t.data(), count);
    }
    if (k >= kFirstInt || k <= Z || s->fp64) {
        if (comp == 1) {
            MS_TRY(hipMemcpy(s->f[k], src, field_bytes(s, k, count), hipMemcpyHostToDevice));
            return MPMW_OK;
        }
        std::vector<double> t(count);
        for (int32_t i = 0; i < count; ++i) t[i] = ((const double *)src)[(size_t)i * comp + c];
        MS_TRY(hipMemcpy(s->f[k], t.data(), sizeof(double) * count, hipMemcpyHostToDevice));
        return MPMW_OK;
    }
    std::vector<float> t(count);
    for (int32_t i = 0; i < count; ++i) t[i] = (float)((const double *)src)[(size_t)i * comp + c];
    MS_TRY(hipMemcpy(s->f[k], t.data(), sizeof(float) * count, hipMemcpyHostToDevice));
    return MPMW_OK;
}

int download(mpms_solver *s, int k, void *dst, int32_t count, int comp = 1, int c = 0, double shift = 0.0) {
    if (count <= 0 || !dst) return MPMW_OK;
    if (k >= kFirstInt) {
        MS_TRY(hipMemcpy(dst, s->f[k], sizeof(int32_t) * count, hipMemcpyDeviceToHost));
        return MPMW_OK;
    }
    std::vector<double> t(count);
    if (k <= Z || s->fp64) {
        MS_TRY(hipMemcpy(t.data(), s->f[k], sizeof(double) * count, hipMemcpyDeviceToHost));
    } else {
        std::vector<float> u(count);
        MS_TRY(hipMemcpy(u.data(), s->f[k], sizeof(float) * count, hipMemcpyDeviceToHost));
        for (int32_t i = 0; i < count; ++i) t[i] = u[i];
    }
    for (int32_t i = 0; i < count; ++i) ((double *)dst)[(size_t)i * comp + c] = t[i] + shift;
    return MPMW_OK;
}

int set_particles_impl(mpms_solver *s, const mpms_particle_arrays *a) {
    const int32_t n = a->n;
    MS_TRY(hipStreamSynchronize(s->stream));
    if (n > s->cap) {
        free_fields(s->f, s);
        free_fields(s->alt, s);
        s->cap = 0;
        int rc = alloc_fields(s, s->f, n);
        if (rc) return rc;
        s->cap = n;
    }
    s->n = n;
    int rc = 0;
    const std::vector<double> zero(std::max<int32_t>(n, 1) * 9, 0.0);
    std::vector<double> eye(std::max<int32_t>(n, 1) * 9, 0.0);
    for (int32_t i = 0; i < n; ++i) eye[9 * i] = eye[9 * i + 4] = eye[9 * i + 8] = 1.0;
    std::vector<int32_t> izero(std::max<int32_t>(n, 1), 0);
    for (int c = 0; c < 3 && !rc; ++c) {
        rc = upload(s, X + c, a->pos, n, 3, c);
        if (!rc) rc = upload(s, VX + c, a->vel ? a->vel : zero.data(), n, 3, c);
        if (!rc) rc = upload(s, EFX + c, a->fext ? a->fext : zero.data(), n, 3, c);
    }
    if (!rc) rc = upload(s, M, a->mass, n);
    if (!rc) rc = upload(s, V, a->volume, n);
    if (!rc) rc = upload(s, T, a->temperature, n);
    for (int c = 0; c < 9 && !rc; ++c) {
        const double id = (c % 4 == 0) ? 1.0 : 0.0;
        rc = upload(s, F0 + c, a->F ? a->F : eye.data(), n, 9, c, id);
        if (!rc) rc = upload(s, FP0 + c, a->Fp ? a->Fp : eye.data(), n, 9, c, id);
        if (!rc) rc = upload(s, C0 + c, a->C ? a->C : zero.data(), n, 9, c);
    }
    for (int c = 0; c < 6 && !rc; ++c) {
        rc = upload(s, S0 + c, a->stress ? a->stress : zero.data(), n, 6, c);
        if (!rc) rc = upload(s, A0 + c, a->backstress ? a->backstress : zero.data(), n, 6, c);
    }
    if (!rc) rc = upload(s, EPS, a->plastic_strain ? a->plastic_strain : zero.data(), n);
    if (!rc) rc = upload(s, RATE, a->strain_rate ? a->strain_rate : zero.data(), n);
    if (!rc) rc = upload(s, DMG, a->damage ? a->damage : zero.data(), n);
    if (!rc) rc = upload(s, SRES, a->residual_stress ? a->residual_stress : zero.data(), n);
    if (!rc) rc = upload(s, ID, a->id, n);
    if (!rc) rc = upload(s, STATUS, a->status ? a->status : izero.data(), n);
    if (!rc) rc = upload(s, LOD, a->lod ? a->lod : izero.data(), n);
    s->fitted = false;
    return rc;
}

int get_particles_impl(mpms_solver *s, mpms_particle_arrays *a) {
    MS_TRY(hipStreamSynchronize(s->stream));
    const int32_t n = std::min(a->n, s->n);
    int rc = 0;
    for (int c = 0; c < 3 && !rc; ++c) {
        rc = download(s, X + c, a->pos, n, 3, c);
        if (!rc) rc = download(s, VX + c, a->vel, n, 3, c);
        if (!rc) rc = download(s, EFX + c, a->fext, n, 3, c);
    }
    if (!rc) rc = download(s, M, a->mass, n);
    if (!rc) rc = download(s, V, a->volume, n);
    if (!rc) rc = download(s, T, a->temperature, n);
    for (int c = 0; c < 9 && !rc; ++c) {
        const double id = (c % 4 == 0) ? 1.0 : 0.0;
        rc = download(s, F0 + c, a->F, n, 9, c, id);
        if (!rc) rc = download(s, FP0 + c, a->Fp, n, 9, c, id);
        if (!rc) rc = download(s, C0 + c, a->C, n, 9, c);
    }
    for (int c = 0; c < 6 && !rc; ++c) {
        rc = download(s, S0 + c, a->stress, n, 6, c);
        if (!rc) rc = download(s, A0 + c, a->backstress, n, 6, c);
    }
    if (!rc) rc = download(s, EPS, a->plastic_strain, n);
    if (!rc) rc = download(s, RATE, a->strain_rate, n);
    if (!rc) rc = download(s, DMG, a->damage, n);
    if (!rc) rc = download(s, SRES, a->residual_stress, n);
    if (!rc) rc = download(s, ID, a->id, n);
    if (!rc) rc = download(s, STATUS, a->status, n);
    if (!rc) rc = download(s, LOD, a->lod, n);
    a->n = n;
    return rc;
}

// The 16 statistics of k_stats reduced on the host (min as -max): [0, 6)
// the AABB of the non-INACTIVE particles, [6, 12) that of ACTIVE and CHIP,
// then max von Mises, max T, KE, internal energy.
template <typename R> int stats(mpms_solver *s, double out[kStats]) {
    const int nb = std::min(blocks(s->n), kStatBlocks);
    hipLaunchKernelGGL(k_stats<R>, dim3(nb), dim3(kBlock), 0, s->stream, pview<R>(s), s->d_part);
    MS_TRY(hipGetLastError());
    std::vector<double> h((size_t)nb * kStats);
    MS_TRY(hipMemcpyAsync(h.data(), s->d_part, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s->stream));
    MS_TRY(hipStreamSynchronize(s->stream));
    for (int c = 0; c < kStats; ++c) out[c] = c >= 14 ? 0.0 : -INFINITY;
    for (int b = 0; b < nb; ++b)
        for (int c = 0; c < kStats; ++c)
            out[c] = c >= 14 ? out[c] + h[(size_t)b * kStats + c] : std::max(out[c], h[(size_t)b * kStats + c]);
    if (!(out[12] > 0.0)) out[12] = 0.0;
    if (!(out[13] > 0.0)) out[13] = 0.0;
    return MPMW_OK;
}

// The grid fit (every kFitEvery steps): the overflow count of the last
// interval must be zero; the sub-box covers every non-INACTIVE particle's
// stencil plus kMargin nodes; the particles are sorted by base-node tile and
// base node, and the tiles' particle ranges found.
template <typename R> int fit(mpms_solver *s) {
    unsigned long long cnt[3] = {0, 0, 0};
    MS_TRY(hipMemcpyAsync(cnt, s->d_counts, sizeof(cnt), hipMemcpyDeviceToHost, s->stream));
    double st[kStats];
    int rc = stats<R>(s, st);
    if (rc) return rc;
    if (s->fitted && cnt[0])
        return fail(MPMW_E_STATE, "F-tracking step: " + std::to_string(cnt[0]) +
                                      " particle stencil visits fell outside the grid sub-box since the last fit "
                                      "(a particle moved more than the margin in 25 steps)");
    const double inv = 1.0 / s->prm.dx;
    int64_t nn = 1;
    for (int a = 0; a < 3; ++a) {
        const double mn = -st[a], mx = st[3 + a];
        int lo = 0, hi = 1;
        if (mn <= mx) {
            lo = (int)std::floor((mn - s->prm.domain_min[a]) * inv - 0.5) - kMargin;
            hi = (int)std::floor((mx - s->prm.domain_min[a]) * inv - 0.5) + 3 + kMargin;
        }
        lo = std::max(lo, 0);
        hi = std::min(hi, s->dims[a]);
        if (hi <= lo) hi = std::min(lo + 1, s->dims[a]), lo = hi - 1;
        s->lo[a] = lo;
        s->nb[a] = hi - lo;
        nn *= s->nb[a];
    }
    if (nn > s->nn_cap) {
        if (s->gblock) (void)hipFree(s->gblock);
        s->gblock = nullptr;
        s->nn_cap = 0;
        const int64_t c = nn + nn / 4;
        if (hipMalloc(&s->gblock, (size_t)c * (10 * s->rsz + sizeof(int32_t))) != hipSuccess) {
            (void)hipGetLastError();
            return fail(MPMW_E_NOMEM, "hipMalloc of the grid sub-box failed");
        }
        s->nn_cap = c;
    }
    s->nn = nn;
    // sort by base node
    const int n = s->n;
    if (n > s->sort_cap) {
        for (void *p : {(void *)s->keys, (void *)s->keys2, (void *)s->idx, (void *)s->idx2})
            if (p) (void)hipFree(p);
        s->keys = s->keys2 = nullptr;
        s->idx = s->idx2 = nullptr;
        s->sort_cap = 0;
        if (hipMalloc(&s->keys, sizeof(uint32_t) * n) != hipSuccess || hipMalloc(&s->keys2, sizeof(uint32_t) * n) != hipSuccess ||
            hipMalloc(&s->idx, sizeof(int32_t) * n) != hipSuccess || hipMalloc(&s->idx2, sizeof(int32_t) * n) != hipSuccess) {
            (void)hipGetLastError();
            return fail(MPMW_E_NOMEM, "hipMalloc of the sort arrays failed");
        }
        s->sort_cap = n;
    }
    if (!s->alt[0]) {
        if ((rc = alloc_fields(s, s->alt, s->cap))) return rc;
    }
    hipLaunchKernelGGL(k_sort_keys<R>, dim3(blocks(n)), dim3(kBlock), 0, s->stream, pview<R>(s), gview<R>(s), s->keys,
                       s->idx);
    MS_TRY(hipGetLastError());
    MS_TRY(mpmw::radix_sort_pairs(s->temp, s->keys, s->keys2, s->idx, s->idx2, n, 32, s->stream));
    {
        int64_t nt = 1;
        const int tdim[3] = {kTX, kTY, kTZ};
        for (int a = 0; a < 3; ++a) nt *= (s->nb[a] + tdim[a] - 1) / tdim[a];
        if (nt * (kTX * kTY * kTZ) >= (int64_t)0xFFFFFFFFu) return fail(MPMW_E_RANGE, "grid sub-box too large for the sort keys");
        if (nt + 1 > s->tile_cap) {
            if (s->tile_start) (void)hipFree(s->tile_start);
            s->tile_start = nullptr;
            s->tile_cap = 0;
            const int64_t c = nt + 1 + nt / 4;
            if (hipMalloc(&s->tile_start, sizeof(int32_t) * (size_t)c) != hipSuccess) {
                (void)hipGetLastError();
                return fail(MPMW_E_NOMEM, "hipMalloc of the tile ranges failed");
            }
            s->tile_cap = (int32_t)c;
        }
        s->ntiles = (int32_t)nt;
        const int64_t tb = nt * (int64_t)kWN * 8 * (int64_t)s->rsz;
        if (tb > s->tpart_cap) {
            if (s->tpart) (void)hipFree(s->tpart);
            s->tpart = nullptr;
            s->tpart_cap = 0;
            const int64_t c = tb + tb / 4;
            if (hipMalloc(&s->tpart, (size_t)c) != hipSuccess) {
                (void)hipGetLastError();
                return fail(MPMW_E_NOMEM, "hipMalloc of the tile windows failed");
            }
            s->tpart_cap = c;
        }
        hipLaunchKernelGGL(k_tile_starts, dim3(blocks(n + 1)), dim3(kBlock), 0, s->stream, (const uint32_t *)s->keys2,
                           n, s->ntiles, s->tile_start);
    }
    {
        // positions (and in FP64 every real field) 8 bytes wide, the rest 4
        FieldPtrs<NFIELD> G;
        for (int k = 0; k < NFIELD; ++k) {
            G.src[k] = s->f[k];
            G.dst[k] = s->alt[k];
        }
        if (s->fp64)
            hipLaunchKernelGGL((k_gather_fields<NFIELD, kFirstInt>), dim3(blocks(n)), dim3(kBlock), 0, s->stream, G,
                               (const int32_t *)s->idx2, n);
        else
            hipLaunchKernelGGL((k_gather_fields<NFIELD, Z + 1>), dim3(blocks(n)), dim3(kBlock), 0, s->stream, G,
                               (const int32_t *)s->idx2, n);
        for (int k = 0; k < NFIELD; ++k) std::swap(s->f[k], s->alt[k]);
    }
    MS_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(kBlock), 0, s->stream, (uint32_t *)s->d_counts, (int64_t)6);
    s->fitted = true;
    s->fits++;
    s->sorts++;
    s->last_fit = s->step;
    return MPMW_OK;
}

void drop_graphs(mpms_solver *s) {
    for (auto &c : s->graphs) {
        if (c.exec) (void)hipGraphExecDestroy(c.exec);
        c.exec = nullptr;
        c.key.clear();
    }
}

// The launch configuration a captured step bakes in: sizes, grid box and
// every device address its kernels take by value.
std::vector<uint64_t> graph_key(const mpms_solver *s) {
    std::vector<uint64_t> k = {(uint64_t)s->n, (uint64_t)s->nn, (uint64_t)s->nn_cap, (uint64_t)s->ntool,
                               (uint64_t)(uintptr_t)s->gblock, (uint64_t)(uintptr_t)s->d_tool_xyz,
                               (uint64_t)(uintptr_t)s->d_tool_vel, (uint64_t)(uintptr_t)s->d_sa,
                               (uint64_t)s->ntiles, (uint64_t)(uintptr_t)s->tile_start,
                               (uint64_t)(uintptr_t)s->tpart};
    for (int a = 0; a < 3; ++a) k.push_back(((uint64_t)(uint32_t)s->lo[a] << 32) | (uint32_t)s->nb[a]);
    for (int f = 0; f < NFIELD; ++f) k.push_back((uint64_t)(uintptr_t)s->f[f]);
    return k;
}

// Replay the step's launches from a graph captured for this configuration,
// capturing it first when there is none (the older of two is replaced).
template <typename Issue> int graph_step(mpms_solver *s, Issue &&issue) {
    std::vector<uint64_t> key = graph_key(s);
    mpms_solver::Captured *hit = nullptr;
    for (auto &c : s->graphs)
        if (c.exec && c.key == key) hit = &c;
    if (!hit) {
        hit = &s->graphs[s->graphs[0].exec && !s->graphs[1].exec ? 1 : 0];
        if (s->graphs[0].exec && s->graphs[1].exec) {
            // keep the more recent one in slot 1
            std::swap(s->graphs[0], s->graphs[1]);
            hit = &s->graphs[1];
        }
        if (hit->exec) (void)hipGraphExecDestroy(hit->exec);
        hit->exec = nullptr;
        hit->key.clear();
        MS_TRY(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
        issue();
        hipGraph_t graph = nullptr;
        const hipError_t launch_err = hipGetLastError();
        const hipError_t end_err = hipStreamEndCapture(s->stream, &graph);
        if (launch_err != hipSuccess || end_err != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            (void)hipGetLastError();
            return fail(MPMW_E_HIP, std::string("F-tracking step capture failed: ") +
                                        hipGetErrorString(launch_err != hipSuccess ? launch_err : end_err));
        }
        const hipError_t inst = hipGraphInstantiate(&hit->exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (inst != hipSuccess) {
            hit->exec = nullptr;
            (void)hipGetLastError();
            return fail(MPMW_E_HIP, std::string("F-tracking step graph instantiation failed: ") + hipGetErrorString(inst));
        }
        hit->key = std::move(key);
        s->captures++;
    }
    MS_TRY(hipGraphLaunch(hit->exec, s->stream));
    s->replays++;
    return MPMW_OK;
}

template <typename R> int step_impl(mpms_solver *s, double dt) {
    if (s->n <= 0) return MPMW_OK;   // MPMSolver::step :1063 silent return
    int rc;
    if (!s->fitted || s->step % kFitEvery == 0) {
        if ((rc = fit<R>(s))) return rc;
    } else if (s->step - s->last_fit == 2 || s->step - s->last_fit == 6 || s->step - s->last_fit == 12) {
        // The sort order decays as particles cross base-node planes: their
        // runs break up, or leave the tile, and take the device atomics.
        // A lattice start decays at once (half the particles sit on the
        // planes and the first motion moves them across).  Three checks
        // per fit interval read the last step's count (one synchronisation
        // each) and fit again when more than 5 % of the particles left.
        unsigned long long out = 0;
        MS_TRY(hipMemcpyAsync(&out, s->d_counts + 3, sizeof(out), hipMemcpyDeviceToHost, s->stream));
        MS_TRY(hipStreamSynchronize(s->stream));
        if ((double)out > kResortFraction * (double)s->n)
            if ((rc = fit<R>(s))) return rc;
    }
    const Particles<R> P = pview<R>(s);
    const Grid<R> g = gview<R>(s);
    StepArgs sa;
    sa.dt = dt;
    for (int a = 0; a < 3; ++a) sa.tool[a] = s->tool_pos[a];
    sa.lod_active = s->prm.lod_active_radius;
    sa.lod_near = s->prm.lod_near_radius;
    sa.lod_enabled = s->prm.lod_enabled;
    sa.chip_step = s->step % 2 == 0 ? 1 : 0;
    sa.fixture = s->prm.fixture_thickness;
    sa.fixture_damping = s->prm.fixture_damping;
    sa.gravity = -9.81;
    sa.tool_vel = s->d_tool_vel;
    sa.islands = s->d_counts + 1;
    hipLaunchKernelGGL(k_step_args, dim3(1), dim3(64), 0, s->stream, s->d_sa, sa);
    const StepArgs *dsa = s->d_sa;
    // resetGridKernel (:113-122): the node channels and tool claims, and the
    // island counters of this step (fill kernels, no memset nodes, so the
    // sequence can be captured)
    const int64_t words = s->nn_cap * 7 * (int64_t)(s->rsz / 4);
    auto issue = [&]() {
        hipLaunchKernelGGL(k_zero_words, dim3((unsigned)std::min<int64_t>(blocks(words), 4096)), dim3(kBlock), 0,
                           s->stream, (uint32_t *)s->gblock, words);
        hipLaunchKernelGGL(k_zero_words, dim3((unsigned)std::min<int64_t>(blocks(s->nn), 2048)), dim3(kBlock), 0,
                           s->stream, (uint32_t *)g.claim, s->nn);
        hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(kBlock), 0, s->stream, (uint32_t *)(s->d_counts + 1),
                           (int64_t)6);
        hipLaunchKernelGGL(k_p2g<R>, dim3((unsigned)s->ntiles), dim3(kBlock), 0, s->stream, P, g, material<R>(s), dsa,
                           (const int32_t *)s->tile_start, (R *)s->tpart);
        if (s->ntool > 0)
            hipLaunchKernelGGL(k_raster<R>, dim3(blocks(s->ntool)), dim3(kBlock), 0, s->stream, g,
                               (const double *)s->d_tool_xyz, s->ntool);
        hipLaunchKernelGGL(k_grid<R>, dim3(blocks(s->nn)), dim3(kBlock), 0, s->stream, g, dsa, (const R *)s->tpart,
                           (const int32_t *)s->tile_start);
        hipLaunchKernelGGL(k_g2p<R>, dim3((unsigned)(s->ntiles + kTailBlocks)), dim3(kBlock), 0, s->stream, P, g, dsa,
                           (R)s->prm.damage_threshold, (const int32_t *)s->tile_start, s->ntiles, kTailBlocks);
    };
    if (!s->graph_on) {
        issue();
        MS_TRY(hipGetLastError());
    } else if ((rc = graph_step(s, issue))) {
        return rc;
    }
    s->time += dt;
    s->step++;
    return MPMW_OK;
}

// The lattices of initializeParticleBox / initializeCylindricalWorkpiece.
struct Lat {
    std::vector<double> pos, mass, vol, T;
    std::vector<int32_t> id, status;
    void push(double x, double y, double z, double m, double v, double t, bool fixed) {
        pos.push_back(x); pos.push_back(y); pos.push_back(z);
        mass.push_back(m); vol.push_back(v); T.push_back(t);
        id.push_back((int32_t)id.size());
        status.push_back(fixed ? kFixed : kActive);
    }
    int upload(mpms_solver *s) {
        mpms_particle_arrays a{};
        a.n = (int32_t)mass.size();
        a.pos = pos.data(); a.mass = mass.data(); a.volume = vol.data(); a.temperature = T.data();
        a.id = id.data(); a.status = status.data();
        return set_particles_impl(s, &a);
    }
};

}  // namespace

extern "C" {

void mpms_params_default(mpms_params *p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->young = 113.8e9; p->poisson = 0.34; p->density = 4430.0; p->mass_scaling = 1.0;
    p->specific_heat = 526.3; p->taylor_quinney = 0.9; p->melt_temp = 1660.0; p->ambient_temp = 25.0;
    p->jc_A = 880e6; p->jc_B = 290e6; p->jc_n = 0.47; p->jc_C = 0.015; p->jc_m = 1.0;
    p->kinematic_hardening = 0.0;
    p->jc_D1 = 0.06; p->jc_D2 = 3.31; p->jc_D3 = -1.96; p->jc_D4 = 0.018; p->jc_D5 = 0.58;
    p->damage_threshold = 1.0; p->ref_strain_rate = 1.0;
    p->lod_active_radius = 0.002; p->lod_near_radius = 0.01; p->lod_enabled = 1;
    p->fixture_thickness = 0.002; p->fixture_damping = 1e7;
    p->dx = std::max(0.0001 * 2.0, 0.0005);
    for (int a = 0; a < 3; ++a) { p->domain_min[a] = -0.1; p->domain_max[a] = 0.1; }
    p->cfl = 0.2;
    p->max_particles = 5000000;
    p->elastic_model = MPMS_ELASTIC_NEOHOOKEAN;
    p->return_map = MPMS_RETURN_RADIAL;
    p->precision = MPMS_FP32;
    p->device = 0;
    p->island_det = 0.5;
    p->newton_fp64 = 1;
    p->polar_fp64 = 1;
}

int mpms_create(const mpms_params *p, mpms_solver **out) {
    if (!p || !out) return fail(MPMW_E_INVAL, "null argument");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        (void)hipGetLastError();
        return fail(MPMW_E_HIP, "no HIP device: the F-tracking solver runs on the GPU only");
    }
    if (p->device < 0 || p->device >= ndev) return fail(MPMW_E_INVAL, "device ordinal out of range");
    if (!(p->dx > 0.0) || !std::isfinite(p->dx)) return fail(MPMW_E_INVAL, "dx must be positive");
    for (int a = 0; a < 3; ++a)
        if (!(p->domain_max[a] > p->domain_min[a])) return fail(MPMW_E_INVAL, "empty domain");
    MS_TRY(hipSetDevice(p->device));
    mpms_solver *s = new mpms_solver();
    s->prm = *p;
    s->fp64 = p->precision == MPMS_FP64;
    s->rsz = s->fp64 ? 8 : 4;
    for (int a = 0; a < 3; ++a) {
        const double d = (double)(int64_t)((p->domain_max[a] - p->domain_min[a]) * (1.0 / p->dx)) + 1.0;
        if (d >= 2147483647.0) { delete s; return fail(MPMW_E_RANGE, "lattice beyond 2^31 nodes per axis"); }
        s->dims[a] = (int32_t)d;
    }
    // node indices of the whole lattice are 32-bit (the P2G's run keys)
    if ((int64_t)s->dims[0] * s->dims[1] * s->dims[2] >= ((int64_t)1 << 31)) {
        delete s;
        return fail(MPMW_E_RANGE, "lattice of 2^31 nodes or more: shrink the domain or grow dx");
    }
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc(&s->d_counts, 8 * sizeof(unsigned long long)) != hipSuccess ||
        hipMalloc(&s->d_part, sizeof(double) * kStatBlocks * kStats) != hipSuccess ||
        hipMalloc(&s->d_sa, sizeof(StepArgs)) != hipSuccess) {
        (void)hipGetLastError();
        delete s;
        return fail(MPMW_E_NOMEM, "F-tracking solver: device set-up failed");
    }
    hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(kBlock), 0, s->stream, (uint32_t *)s->d_counts, (int64_t)16);
    *out = s;
    return MPMW_OK;
}

int mpms_destroy(mpms_solver *s) {
    if (!s) return MPMW_OK;
    (void)hipSetDevice(s->prm.device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    drop_graphs(s);
    free_fields(s->f, s);
    free_fields(s->alt, s);
    for (void *p : {(void *)s->gblock, (void *)s->d_counts, (void *)s->d_part, (void *)s->d_sa, (void *)s->d_tool_xyz,
                    (void *)s->d_tool_vel, (void *)s->keys, (void *)s->keys2, (void *)s->idx, (void *)s->idx2,
                    (void *)s->tile_start, (void *)s->tpart})
        if (p) (void)hipFree(p);
    mpmw::free_scratch(s->temp);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return MPMW_OK;
}

int mpms_set_particles(mpms_solver *s, const mpms_particle_arrays *a) {
    MS_CHECK(s);
    if (!a || a->n < 0) return fail(MPMW_E_INVAL, "bad particle arrays");
    if (a->n > 0 && (!a->pos || !a->mass || !a->volume || !a->temperature || !a->id))
        return fail(MPMW_E_INVAL, "pos, mass, volume, temperature and id are required");
    return set_particles_impl(s, a);
}

int mpms_get_particles(mpms_solver *s, mpms_particle_arrays *a) {
    MS_CHECK(s);
    if (!a) return fail(MPMW_E_INVAL, "null");
    return get_particles_impl(s, a);
}

int mpms_particle_count(mpms_solver *s, int32_t *n) {
    MS_CHECK(s);
    if (!n) return fail(MPMW_E_INVAL, "null");
    *n = s->n;
    return MPMW_OK;
}

int mpms_init_particle_box(mpms_solver *s, const double mn[3], const double mx[3], double spacing) {
    MS_CHECK(s);
    if (!mn || !mx || !(spacing > 0)) return fail(MPMW_E_INVAL, "bad box");
    Lat L;
    const double density = s->prm.density * s->prm.mass_scaling;
    const double mass = density * spacing * spacing * spacing, vol = spacing * spacing * spacing;
    const double fixed_z = mn[2] + (mx[2] - mn[2]) * 0.1;
    for (double z = mn[2]; z <= mx[2]; z += spacing)
        for (double y = mn[1]; y <= mx[1]; y += spacing)
            for (double x = mn[0]; x <= mx[0]; x += spacing) {
                if ((int32_t)L.mass.size() >= s->prm.max_particles) break;
                L.push(x, y, z, mass, vol, s->prm.ambient_temp, z <= fixed_z);
            }
    return L.upload(s);
}

int mpms_init_cylinder(mpms_solver *s, const double c[3], double radius, double length, double spacing) {
    MS_CHECK(s);
    if (!c || !(spacing > 0)) return fail(MPMW_E_INVAL, "bad cylinder");
    Lat L;
    const double density = s->prm.density * s->prm.mass_scaling;
    const double mass = density * spacing * spacing * spacing, vol = spacing * spacing * spacing;
    const double fixed_z = c[2] - length * 0.1, r2 = radius * radius;
    for (double dz = 0; dz <= length; dz += spacing)
        for (double dy = -radius; dy <= radius; dy += spacing)
            for (double dx = -radius; dx <= radius; dx += spacing) {
                if (dx * dx + dy * dy > r2) continue;
                if ((int32_t)L.mass.size() >= s->prm.max_particles) break;
                L.push(c[0] + dx, c[1] + dy, c[2] - dz, mass, vol, s->prm.ambient_temp, c[2] - dz <= fixed_z);
            }
    return L.upload(s);
}

int mpms_set_tool_nodes(mpms_solver *s, const double *xyz, const double *vel, int32_t n) {
    MS_CHECK(s);
    if (n < 0 || (n > 0 && (!xyz || !vel))) return fail(MPMW_E_INVAL, "bad tool nodes");
    MS_TRY(hipStreamSynchronize(s->stream));
    if (n > s->tool_cap) {
        if (s->d_tool_xyz) (void)hipFree(s->d_tool_xyz);
        if (s->d_tool_vel) (void)hipFree(s->d_tool_vel);
        s->d_tool_xyz = s->d_tool_vel = nullptr;
        s->tool_cap = 0;
        if (hipMalloc(&s->d_tool_xyz, sizeof(double) * 3 * n) != hipSuccess ||
            hipMalloc(&s->d_tool_vel, sizeof(double) * 3 * n) != hipSuccess) {
            (void)hipGetLastError();
            return fail(MPMW_E_NOMEM, "hipMalloc of the tool nodes failed");
        }
        s->tool_cap = n;
    }
    if (n > 0) {
        MS_TRY(hipMemcpy(s->d_tool_xyz, xyz, sizeof(double) * 3 * n, hipMemcpyHostToDevice));
        MS_TRY(hipMemcpy(s->d_tool_vel, vel, sizeof(double) * 3 * n, hipMemcpyHostToDevice));
    }
    s->ntool = n;
    return MPMW_OK;
}

int mpms_set_tool_position(mpms_solver *s, double x, double y, double z) {
    MS_CHECK(s);
    s->tool_pos[0] = x; s->tool_pos[1] = y; s->tool_pos[2] = z;
    return MPMW_OK;
}

int mpms_apply_external_force(mpms_solver *s, const double c[3], double radius, const double f[3]) {
    MS_CHECK(s);
    if (!c || !f) return fail(MPMW_E_INVAL, "null");
    if (s->n == 0 || std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]) < 1e-12) return MPMW_OK;   // :1441
    if (s->fp64)
        hipLaunchKernelGGL(k_ext_force<double>, dim3(blocks(s->n)), dim3(kBlock), 0, s->stream, pview<double>(s), c[0],
                           c[1], c[2], radius, f[0], f[1], f[2]);
    else
        hipLaunchKernelGGL(k_ext_force<float>, dim3(blocks(s->n)), dim3(kBlock), 0, s->stream, pview<float>(s), c[0],
                           c[1], c[2], radius, f[0], f[1], f[2]);
    MS_TRY(hipGetLastError());
    return MPMW_OK;
}

int mpms_step(mpms_solver *s, double dt) {
    MS_CHECK(s);
    if (!std::isfinite(dt)) return fail(MPMW_E_INVAL, "dt must be finite");
    return s->fp64 ? step_impl<double>(s, dt) : step_impl<float>(s, dt);
}

int mpms_run(mpms_solver *s, int32_t nsteps, double dt) {
    for (int32_t k = 0; k < nsteps; ++k) {
        int rc = mpms_step(s, dt);
        if (rc) return rc;
    }
    return MPMW_OK;
}

int mpms_synchronize(mpms_solver *s) {
    MS_CHECK(s);
    MS_TRY(hipStreamSynchronize(s->stream));
    return MPMW_OK;
}

int mpms_stable_dt(mpms_solver *s, double *dt) {
    MS_CHECK(s);
    if (!dt) return fail(MPMW_E_INVAL, "null");
    const double c = std::sqrt(s->prm.young / (s->prm.density * s->prm.mass_scaling));
    *dt = std::max(s->prm.cfl * s->prm.dx / std::max(c, 1.0), 1e-10);
    return MPMW_OK;
}

int mpms_metrics(mpms_solver *s, double out[4]) {
    MS_CHECK(s);
    if (!out) return fail(MPMW_E_INVAL, "null");
    out[0] = out[2] = out[3] = 0.0;
    out[1] = 0.0;
    if (s->n == 0) return MPMW_OK;
    double st[kStats];
    int rc = s->fp64 ? stats<double>(s, st) : stats<float>(s, st);
    if (rc) return rc;
    for (int k = 0; k < 4; ++k) out[k] = st[12 + k];
    return MPMW_OK;
}

int mpms_get_bounds(mpms_solver *s, double b[6]) {
    MS_CHECK(s);
    if (!b) return fail(MPMW_E_INVAL, "null");
    if (s->n == 0) {
        for (int k = 0; k < 6; ++k) b[k] = k < 3 ? INFINITY : -INFINITY;
        return MPMW_OK;
    }
    double st[kStats];
    int rc = s->fp64 ? stats<double>(s, st) : stats<float>(s, st);
    if (rc) return rc;
    for (int a = 0; a < 3; ++a) { b[a] = -st[6 + a]; b[3 + a] = st[9 + a]; }
    return MPMW_OK;
}

int mpms_get_grid(mpms_solver *s, int32_t lo[3], int32_t n[3], double *mass, double *mom, double *force) {
    MS_CHECK(s);
    if (!lo || !n) return fail(MPMW_E_INVAL, "null");
    for (int a = 0; a < 3; ++a) { lo[a] = s->lo[a]; n[a] = s->nb[a]; }
    if (!s->gblock || s->nn == 0) return MPMW_OK;
    MS_TRY(hipStreamSynchronize(s->stream));
    const int64_t nn = s->nn, c = s->nn_cap;
    std::vector<double> t((size_t)nn);
    auto fetch = [&](int ch, double *dst, int comp, int k) -> int {
        if (s->fp64) {
            MS_TRY(hipMemcpy(t.data(), (double *)s->gblock + ch * c, sizeof(double) * nn, hipMemcpyDeviceToHost));
        } else {
            std::vector<float> u((size_t)nn);
            MS_TRY(hipMemcpy(u.data(), (float *)s->gblock + ch * c, sizeof(float) * nn, hipMemcpyDeviceToHost));
            for (int64_t i = 0; i < nn; ++i) t[i] = u[i];
        }
        for (int64_t i = 0; i < nn; ++i) dst[i * comp + k] = t[i];
        return MPMW_OK;
    };
    int rc = 0;
    if (mass) rc = fetch(0, mass, 1, 0);
    for (int a = 0; a < 3 && !rc; ++a) {
        if (mom) rc = fetch(1 + a, mom, 3, a);
        if (!rc && force) rc = fetch(4 + a, force, 3, a);
    }
    return rc;
}

int mpms_get_diagnostics(mpms_solver *s, mpms_diagnostics *d) {
    MS_CHECK(s);
    if (!d) return fail(MPMW_E_INVAL, "null");
    unsigned long long cnt[4] = {0, 0, 0, 0};
    MS_TRY(hipMemcpyAsync(cnt, s->d_counts, sizeof(cnt), hipMemcpyDeviceToHost, s->stream));
    MS_TRY(hipStreamSynchronize(s->stream));
    for (int a = 0; a < 3; ++a) { d->box_lo[a] = s->lo[a]; d->box_n[a] = s->nb[a]; }
    d->out_of_tile = (int64_t)cnt[3];
    d->fits = s->fits;
    d->sorts = s->sorts;
    d->overflow = (int64_t)cnt[0];
    d->island_kinematics = (int64_t)cnt[1];
    d->island_newton = (int64_t)cnt[2];
    return MPMW_OK;
}

int mpms_set_graph(mpms_solver *s, int32_t on) {
    MS_CHECK(s);
    if (!on) {
        MS_TRY(hipStreamSynchronize(s->stream));
        drop_graphs(s);
    }
    s->graph_on = on != 0;
    return MPMW_OK;
}

int mpms_graph_stats(mpms_solver *s, int64_t out[2]) {
    MS_CHECK(s);
    if (!out) return fail(MPMW_E_INVAL, "null");
    out[0] = s->captures;
    out[1] = s->replays;
    return MPMW_OK;
}

int mpms_get_stream(mpms_solver *s, void **stream) {
    MS_CHECK(s);
    if (!stream) return fail(MPMW_E_INVAL, "null");
    *stream = (void *)s->stream;
    return MPMW_OK;
}

}  // extern "C"

namespace mpms_internal {
bool is_fp64(const mpms_solver *s) { return s->fp64; }
int32_t count(const mpms_solver *s) { return s->n; }
int device(const mpms_solver *s) { return s->prm.device; }
hipStream_t stream(const mpms_solver *s) { return s->stream; }
void domain_min(const mpms_solver *s, double o[3]) {
    for (int a = 0; a < 3; ++a) o[a] = s->prm.domain_min[a];
}
mpms::Particles<float> particles_f(mpms_solver *s) { return pview<float>(s); }
mpms::Particles<double> particles_d(mpms_solver *s) { return pview<double>(s); }
int tool_node_buffers(mpms_solver *s, int32_t n, double **xyz, double **vel) {
    if (n > s->tool_cap) {
        MS_TRY(hipStreamSynchronize(s->stream));
        if (s->d_tool_xyz) (void)hipFree(s->d_tool_xyz);
        if (s->d_tool_vel) (void)hipFree(s->d_tool_vel);
        s->d_tool_xyz = s->d_tool_vel = nullptr;
        s->tool_cap = 0;
        if (hipMalloc(&s->d_tool_xyz, sizeof(double) * 3 * n) != hipSuccess ||
            hipMalloc(&s->d_tool_vel, sizeof(double) * 3 * n) != hipSuccess) {
            (void)hipGetLastError();
            return fail(MPMW_E_NOMEM, "hipMalloc of the tool nodes failed");
        }
        s->tool_cap = n;
    }
    s->ntool = n;
    *xyz = s->d_tool_xyz;
    *vel = s->d_tool_vel;
    return MPMW_OK;
}
}  // namespace mpms_internal
