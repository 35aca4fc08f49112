// mpmw_export.hpp -- the result formats of the serving loop.
//
// * NDJSON particle message (StdoutIPCExporter::exportParticles,
//   StdoutIPCExporter.cpp:70-113): the particles that are not INACTIVE, in
//   index order; positions as FP32 x, y, z and temperatures as FP32, each
//   array little-endian bytes in standard base64 with '=' padding
//   (Base64.h:13-52), in one JSON object with the keys in sorted order and
//   no spaces, {"count":N,"positions_b64":"..","step":S,"temps_b64":"..",
//   "type":"particles"}, and a newline.  No message when no particle is
//   exported (:79).  The device builds both base64 strings: a particle's 12
//   position bytes are exactly 16 characters, so its position encodes on
//   its own at 16 times its rank among the exported particles; the packed
//   temperatures encode three per thread.
// * Legacy VTK particle file (VTKExporter::exportParticles,
//   VTKExporter.cpp:252-336): POLYDATA points (%.8e), one vertex per point,
//   point data Temperature_C (%.2f), Pressure_MPa (-(sxx + syy + szz) / 3 /
//   1e6, %.6e; the pressure of getParticlesForExport,
//   MPMWorkpieceSolver.cu:870-894), Velocity_m_s (|v|, %.6e),
//   ParticleStatus (int) and the Velocity vectors (%.6e), written on the
//   host from the downloaded fields.
#pragma once

namespace mpmw {

__device__ __forceinline__ uint32_t b64_char(uint32_t v) {
    // A-Z a-z 0-9 + /
    return v < 26 ? 'A' + v : (v < 52 ? 'a' + (v - 26) : (v < 62 ? '0' + (v - 52) : (v == 62 ? '+' : '/')));
}

// Four characters of three bytes, packed little-endian into one word (the
// first character in the lowest byte, as they lie in memory).
__device__ __forceinline__ uint32_t b64_quad(uint32_t b0, uint32_t b1, uint32_t b2) {
    const uint32_t t = (b0 << 16) | (b1 << 8) | b2;
    return b64_char(t >> 18) | (b64_char((t >> 12) & 63) << 8) | (b64_char((t >> 6) & 63) << 16) |
           (b64_char(t & 63) << 24);
}

__global__ void __launch_bounds__(kBlock) k_export_flags(const int32_t *status, int32_t *flag, int n) {
    const int i = blockIdx.x * kBlock + (int)threadIdx.x;
    if (i < n) flag[i] = status[i] != kInactive ? 1 : 0;
    else if (i == n) flag[i] = 0;
}

// Exported particle i, of rank off[i]: the base64 of its FP32 position at
// pos64 + 16 off[i], its FP32 temperature at temps[off[i]].
template <typename R>
__global__ void __launch_bounds__(kBlock)
    k_export_pack(const double *x, const double *y, const double *z, const R *T, const int32_t *status,
                  const int32_t *off, uint4 *pos64, float *temps, int n) {
    const int i = blockIdx.x * kBlock + (int)threadIdx.x;
    if (i >= n || status[i] == kInactive) return;
    const int o = off[i];
    const uint32_t w[3] = {__float_as_uint((float)x[i]), __float_as_uint((float)y[i]), __float_as_uint((float)z[i])};
    uint8_t b[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) b[k] = (uint8_t)(w[k / 4] >> (8 * (k % 4)));
    uint4 q;
    q.x = b64_quad(b[0], b[1], b[2]);
    q.y = b64_quad(b[3], b[4], b[5]);
    q.z = b64_quad(b[6], b[7], b[8]);
    q.w = b64_quad(b[9], b[10], b[11]);
    pos64[o] = q;
    temps[o] = (float)T[i];
}

// Standard base64 of nbytes bytes, group g (bytes 3g .. 3g + 2) to
// characters 4g .. 4g + 3; a last group of one or two bytes ends in "==" or
// "=".
__global__ void __launch_bounds__(kBlock) k_b64(const uint8_t *src, int64_t nbytes, uint32_t *dst) {
    const int64_t g = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t b = 3 * g;
    if (b >= nbytes) return;
    const int64_t left = nbytes - b;
    const uint32_t b0 = src[b], b1 = left > 1 ? src[b + 1] : 0u, b2 = left > 2 ? src[b + 2] : 0u;
    uint32_t q = b64_quad(b0, b1, b2);
    if (left == 1) q = (q & 0x0000FFFFu) | ((uint32_t)'=' << 16) | ((uint32_t)'=' << 24);
    else if (left == 2) q = (q & 0x00FFFFFFu) | ((uint32_t)'=' << 24);
    dst[g] = q;
}

}  // namespace mpmw

namespace {

// Exported-particle ranks (off[0..n], off[n] = the count) on the device;
// returns the count.
int export_ranks(mpmw_solver *s, int32_t *flag, int32_t *off, int64_t *count) {
    const int n = s->n;
    hipLaunchKernelGGL(k_export_flags, dim3(blocks_for(n + 1)), dim3(kBlock), 0, s->stream,
                       (const int32_t *)s->f[F_STATUS], flag, n);
    HIP_TRY(hipGetLastError());
    HIP_TRY(exclusive_scan(s->sort.temp2, flag, off, n + 1, s->stream));
    int32_t c = 0;
    HIP_TRY(hipMemcpyAsync(&c, off + n, sizeof(int32_t), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    *count = c;
    return MPMW_OK;
}

// Device-to-host copies of (src, dst, bytes) pieces into caller memory,
// staged through two pinned chunks: the copy of one chunk runs while the
// host moves the previous one out (a copy straight into pageable memory ran
// at ~2.4 GB/s).
struct CopyPiece {
    const char *src;
    char *dst;
    size_t bytes;
};

int staged_download(mpmw_solver *s, const std::vector<CopyPiece> &pieces) {
    constexpr size_t kChunk = 32u << 20;
    std::vector<CopyPiece> chunks;
    for (const CopyPiece &p : pieces)
        for (size_t o = 0; o < p.bytes; o += kChunk)
            chunks.push_back({p.src + o, p.dst + o, std::min(kChunk, p.bytes - o)});
    if (chunks.empty()) return MPMW_OK;
    void *stage = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipError_t e = pool::get(s->prm.device, true, 2 * kChunk, &stage);
    for (int k = 0; k < 2 && e == hipSuccess; ++k) e = hipEventCreateWithFlags(&ev[k], hipEventDisableTiming);
    for (size_t k = 0; k < chunks.size() && e == hipSuccess; ++k) {
        char *slot = (char *)stage + (k % 2) * kChunk;
        e = hipMemcpyAsync(slot, chunks[k].src, chunks[k].bytes, hipMemcpyDeviceToHost, s->stream);
        if (e == hipSuccess) e = hipEventRecord(ev[k % 2], s->stream);
        if (e == hipSuccess && k > 0) {
            e = hipEventSynchronize(ev[(k - 1) % 2]);
            if (e == hipSuccess)
                std::memcpy(chunks[k - 1].dst, (char *)stage + ((k - 1) % 2) * kChunk, chunks[k - 1].bytes);
        }
    }
    if (e == hipSuccess) {
        const size_t k = chunks.size() - 1;
        e = hipEventSynchronize(ev[k % 2]);
        if (e == hipSuccess) std::memcpy(chunks[k].dst, (char *)stage + (k % 2) * kChunk, chunks[k].bytes);
    }
    (void)hipStreamSynchronize(s->stream);
    for (hipEvent_t x : ev)
        if (x) (void)hipEventDestroy(x);
    pool::put(stage);
    if (e != hipSuccess) return fail(MPMW_E_HIP, std::string("download: ") + hipGetErrorString(e));
    return MPMW_OK;
}

template <typename R> int export_ndjson_impl(mpmw_solver *s, int32_t step, char *buf, size_t cap, size_t *len) {
    *len = 0;
    if (s->n <= 0) return MPMW_OK;
    const int n = s->n;
    void *flag = nullptr, *off = nullptr;
    int rc = dalloc(s, &flag, sizeof(int32_t) * ((size_t)n + 1));
    if (!rc) rc = dalloc(s, &off, sizeof(int32_t) * ((size_t)n + 1));
    int64_t count = 0;
    if (!rc) rc = export_ranks(s, (int32_t *)flag, (int32_t *)off, &count);
    if (rc || count == 0) {
        dfree(s, flag, 0);
        dfree(s, off, 0);
        return rc;
    }
    const size_t npos = 16 * (size_t)count;                  // 12 bytes -> 16 characters
    const size_t ntmp = 4 * (((size_t)count * 4 + 2) / 3);    // 4 bytes each, padded groups
    const std::string h0 = "{\"count\":" + std::to_string(count) + ",\"positions_b64\":\"";
    const std::string h1 = "\",\"step\":" + std::to_string(step) + ",\"temps_b64\":\"";
    const std::string h2 = "\",\"type\":\"particles\"}\n";
    const size_t total = h0.size() + npos + h1.size() + ntmp + h2.size();
    *len = total;
    if (!buf || cap < total) {
        dfree(s, flag, 0);
        dfree(s, off, 0);
        return buf ? fail(MPMW_E_RANGE, "export_ndjson: buffer of " + std::to_string(cap) + " bytes, " +
                                            std::to_string(total) + " needed")
                   : MPMW_OK;
    }
    void *pos64 = nullptr, *temps = nullptr, *t64 = nullptr;
    rc = dalloc(s, &pos64, npos);
    if (!rc) rc = dalloc(s, &temps, sizeof(float) * (size_t)count);
    if (!rc) rc = dalloc(s, &t64, ntmp);
    if (!rc) {
        hipLaunchKernelGGL(k_export_pack<R>, dim3(blocks_for(n)), dim3(kBlock), 0, s->stream,
                           (const double *)s->f[F_X], (const double *)s->f[F_Y], (const double *)s->f[F_Z],
                           (const R *)s->f[F_T], (const int32_t *)s->f[F_STATUS], (const int32_t *)off,
                           (uint4 *)pos64, (float *)temps, n);
        const int64_t groups = (4 * count + 2) / 3;
        hipLaunchKernelGGL(k_b64, dim3((unsigned)((groups + kBlock - 1) / kBlock)), dim3(kBlock), 0, s->stream,
                           (const uint8_t *)temps, 4 * count, (uint32_t *)t64);
        hipError_t e = hipGetLastError();
        char *p = buf;
        std::memcpy(p, h0.data(), h0.size());
        char *p_pos = p + h0.size();
        p = p_pos + npos;
        std::memcpy(p, h1.data(), h1.size());
        char *p_tmp = p + h1.size();
        p = p_tmp + ntmp;
        std::memcpy(p, h2.data(), h2.size());
        if (e != hipSuccess) rc = fail(MPMW_E_HIP, std::string("export_ndjson: ") + hipGetErrorString(e));
        else rc = staged_download(s, {{(const char *)pos64, p_pos, npos}, {(const char *)t64, p_tmp, ntmp}});
    }
    dfree(s, pos64, 0);
    dfree(s, temps, 0);
    dfree(s, t64, 0);
    dfree(s, flag, 0);
    dfree(s, off, 0);
    return rc;
}

// ---- the allocation-free NDJSON export ---------------------------------
//
// Every device buffer of mpmw_export_ndjson_noalloc (flags, ranks, the
// scan's temporary storage, packed positions, packed temperatures and their
// base64) lies in `scratch`, one static array per device that the library
// holds until the process ends.  Host code never passes `scratch` itself to
// the device (its name there is a host-side stand-in): the address of each
// device's copy comes from hipGetSymbolAddress on the first call on that
// device and is kept per device.  Calls on one device are serialised by a
// lock, so they do not overlap in the array; the array keeps the last
// call's contents, and no part of it needs to start at zero (every word a
// call reads was written earlier in the same call).
constexpr size_t kExportScratchBytes = MPMW_EXPORT_SCRATCH_BYTES;

}  // namespace

static __device__ __align__(256) unsigned char scratch[kExportScratchBytes];

namespace {

std::mutex g_scratch_mu;
std::map<int, char *> g_scratch_addr;       // device -> address of its copy of `scratch`
std::map<int, std::mutex> g_scratch_lock;   // device -> lock of its copy

int scratch_on_device(int dev, char **out) {
    std::lock_guard<std::mutex> g(g_scratch_mu);
    auto it = g_scratch_addr.find(dev);
    if (it != g_scratch_addr.end()) {
        *out = it->second;
        return MPMW_OK;
    }
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(MPMW_E_HIP, std::string("export_ndjson_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
    g_scratch_addr[dev] = (char *)p;
    *out = (char *)p;
    return MPMW_OK;
}

std::mutex &scratch_lock(int dev) {
    std::lock_guard<std::mutex> g(g_scratch_mu);
    return g_scratch_lock[dev];
}

// Offsets of the buffers of an export of n particles inside the scratch
// array; `end` is the bytes it needs.
struct ExportLayout {
    size_t flag, off, scan, scan_bytes, pos, tmp, t64, end;
};

int export_layout(int n, ExportLayout *L) {
    HIP_TRY(exclusive_scan_temp_bytes(n + 1, &L->scan_bytes));
    L->flag = 0;
    L->off = align256(L->flag + sizeof(int32_t) * ((size_t)n + 1));
    L->scan = align256(L->off + sizeof(int32_t) * ((size_t)n + 1));
    L->pos = align256(L->scan + L->scan_bytes);
    L->tmp = align256(L->pos + 16 * (size_t)n);
    L->t64 = align256(L->tmp + sizeof(float) * (size_t)n);
    L->end = L->t64 + 4 * ((4 * (size_t)n + 2) / 3);
    return MPMW_OK;
}

template <typename R>
int export_ndjson_noalloc_impl(mpmw_solver *s, int32_t step, char *buf, size_t cap, size_t *len) {
    *len = 0;
    if (s->n <= 0) return MPMW_OK;
    const int n = s->n;
    ExportLayout L;
    int rc = export_layout(n, &L);
    if (rc) return rc;
    if (L.end > kExportScratchBytes) {
        // bytes per particle: flag + rank 8, position 16, temperature 4, its base64 16/3
        const size_t most = (kExportScratchBytes - 4096 - L.scan_bytes) * 3 / 100;
        return fail(MPMW_E_RANGE, "export_ndjson_noalloc: " + std::to_string(n) + " particles need " +
                                      std::to_string(L.end) + " bytes of scratch, the limit is " +
                                      std::to_string(kExportScratchBytes) + " (about " + std::to_string(most) +
                                      " particles); use mpmw_export_ndjson");
    }
    const int dev = s->prm.device;
    char *base = nullptr;
    if ((rc = scratch_on_device(dev, &base))) return rc;
    std::lock_guard<std::mutex> lk(scratch_lock(dev));
    int32_t *flag = (int32_t *)(base + L.flag), *off = (int32_t *)(base + L.off);
    hipLaunchKernelGGL(k_export_flags, dim3(blocks_for(n + 1)), dim3(kBlock), 0, s->stream,
                       (const int32_t *)s->f[F_STATUS], flag, n);
    HIP_TRY(hipGetLastError());
    HIP_TRY(exclusive_scan_buf(base + L.scan, L.scan_bytes, flag, off, n + 1, s->stream));
    int32_t count = 0;
    HIP_TRY(hipMemcpyAsync(&count, off + n, sizeof(int32_t), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (count == 0) return MPMW_OK;
    const size_t npos = 16 * (size_t)count;
    const size_t ntmp = 4 * (((size_t)count * 4 + 2) / 3);
    const std::string h0 = "{\"count\":" + std::to_string(count) + ",\"positions_b64\":\"";
    const std::string h1 = "\",\"step\":" + std::to_string(step) + ",\"temps_b64\":\"";
    const std::string h2 = "\",\"type\":\"particles\"}\n";
    const size_t total = h0.size() + npos + h1.size() + ntmp + h2.size();
    *len = total;
    if (!buf) return MPMW_OK;
    if (cap < total)
        return fail(MPMW_E_RANGE, "export_ndjson_noalloc: buffer of " + std::to_string(cap) + " bytes, " +
                                      std::to_string(total) + " needed");
    hipLaunchKernelGGL(k_export_pack<R>, dim3(blocks_for(n)), dim3(kBlock), 0, s->stream,
                       (const double *)s->f[F_X], (const double *)s->f[F_Y], (const double *)s->f[F_Z],
                       (const R *)s->f[F_T], (const int32_t *)s->f[F_STATUS], (const int32_t *)off,
                       (uint4 *)(base + L.pos), (float *)(base + L.tmp), n);
    const int64_t groups = (4 * (int64_t)count + 2) / 3;
    hipLaunchKernelGGL(k_b64, dim3((unsigned)((groups + kBlock - 1) / kBlock)), dim3(kBlock), 0, s->stream,
                       (const uint8_t *)(base + L.tmp), 4 * (int64_t)count, (uint32_t *)(base + L.t64));
    HIP_TRY(hipGetLastError());
    char *p = buf;
    std::memcpy(p, h0.data(), h0.size());
    char *p_pos = p + h0.size();
    p = p_pos + npos;
    std::memcpy(p, h1.data(), h1.size());
    char *p_tmp = p + h1.size();
    p = p_tmp + ntmp;
    std::memcpy(p, h2.data(), h2.size());
    // straight into the caller's memory: no pinned staging is allocated
    HIP_TRY(hipMemcpyAsync(p_pos, base + L.pos, npos, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipMemcpyAsync(p_tmp, base + L.t64, ntmp, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return MPMW_OK;
}

template <typename R> int export_vtk_impl(mpmw_solver *s, const char *path) {
    const int n = s->n;
    std::vector<double> pos(3 * (size_t)n), vel(3 * (size_t)n), T(n), stress(6 * (size_t)n);
    std::vector<int32_t> status(n);
    mpmw_particle_arrays a{};
    a.n = n;
    a.pos = pos.data(); a.vel = vel.data(); a.temperature = T.data(); a.stress = stress.data();
    a.status = status.data();
    int rc = get_particles_impl<R>(s, &a);
    if (rc) return rc;
    std::FILE *f = std::fopen(path, "w");
    if (!f) return fail(MPMW_E_INVAL, std::string("export_vtk: cannot open ") + path);
    std::vector<int> live;
    live.reserve(n);
    for (int i = 0; i < n; ++i)
        if (status[i] != MPMW_INACTIVE) live.push_back(i);
    const int m = (int)live.size();
    std::fprintf(f, "# vtk DataFile Version 3.0\nEdgePredict SPH Particles\nASCII\nDATASET POLYDATA\n");
    std::fprintf(f, "POINTS %d double\n", m);
    for (int i : live) std::fprintf(f, "%.8e %.8e %.8e\n", pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
    std::fprintf(f, "\nVERTICES %d %d\n", m, m * 2);
    for (int k = 0; k < m; ++k) std::fprintf(f, "1 %d\n", k);
    std::fprintf(f, "\nPOINT_DATA %d\n", m);
    std::fprintf(f, "SCALARS Temperature_C double 1\nLOOKUP_TABLE default\n");
    for (int i : live) std::fprintf(f, "%.2f\n", T[i]);
    std::fprintf(f, "\nSCALARS Pressure_MPa double 1\nLOOKUP_TABLE default\n");
    for (int i : live) {
        const double p = -(stress[6 * i] + stress[6 * i + 1] + stress[6 * i + 2]) / 3.0;
        std::fprintf(f, "%.6e\n", p / 1e6);
    }
    std::fprintf(f, "\nSCALARS Velocity_m_s double 1\nLOOKUP_TABLE default\n");
    for (int i : live) {
        const double vx = vel[3 * i], vy = vel[3 * i + 1], vz = vel[3 * i + 2];
        std::fprintf(f, "%.6e\n", std::sqrt(vx * vx + vy * vy + vz * vz));
    }
    std::fprintf(f, "\nSCALARS ParticleStatus int 1\nLOOKUP_TABLE default\n");
    for (int i : live) std::fprintf(f, "%d\n", status[i]);
    std::fprintf(f, "\nVECTORS Velocity double\n");
    for (int i : live) std::fprintf(f, "%.6e %.6e %.6e\n", vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]);
    const bool bad = std::ferror(f) != 0;
    if (std::fclose(f) != 0 || bad) return fail(MPMW_E_INVAL, std::string("export_vtk: write failed: ") + path);
    return MPMW_OK;
}

}  // namespace

extern "C" {

int mpmw_export_ndjson(mpmw_solver *s, int32_t step, char *buf, size_t cap, size_t *len) {
    CHECK_S(s);
    if (!len) return fail(MPMW_E_INVAL, "export_ndjson: null length");
    return DISPATCH(s, export_ndjson_impl, s, step, buf, cap, len);
}

int mpmw_export_ndjson_noalloc(mpmw_solver *s, int32_t step, char *buf, size_t cap, size_t *len) {
    CHECK_S(s);
    if (!len) return fail(MPMW_E_INVAL, "export_ndjson_noalloc: null length");
    return DISPATCH(s, export_ndjson_noalloc_impl, s, step, buf, cap, len);
}

int mpmw_export_vtk(mpmw_solver *s, const char *path) {
    CHECK_S(s);
    if (!path) return fail(MPMW_E_INVAL, "export_vtk: null path");
    return DISPATCH(s, export_vtk_impl, s, path);
}

}  // extern "C"
