// hbm_probe.hip -- what the HBM of one MI355X sustains for plain streams:
// a read-only, a write-only and two copy kernels (16-byte and 4-byte
// elements) over 2 GiB buffers, grid-stride loops at three grid sizes.  The
// rates are the practical ceilings the particle pass is compared with (it
// reads and writes about equal byte counts, like a copy).
#include <hip/hip_runtime.h>
#include <cstdio>

__global__ void __launch_bounds__(256) k_read(const float4 *__restrict__ a, size_t n, float *out) {
    float s = 0.f;
    for (size_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        float4 v = a[i];
        s += v.x + v.y + v.z + v.w;
    }
    if (s == 12345.f) out[blockIdx.x] = s;
}
__global__ void __launch_bounds__(256) k_write(float4 *__restrict__ b, size_t n) {
    for (size_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        b[i] = make_float4(1.f, 2.f, 3.f, 4.f);
}
__global__ void __launch_bounds__(256) k_copy(const float4 *__restrict__ a, float4 *__restrict__ b, size_t n) {
    for (size_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) b[i] = a[i];
}
// scalar (4-byte) copy, the access width of the solver's SoA fields
__global__ void __launch_bounds__(256) k_copy1(const float *__restrict__ a, float *__restrict__ b, size_t n) {
    for (size_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) b[i] = a[i];
}

int main() {
    const size_t bytes = (size_t)2 << 30;
    const size_t n4 = bytes / 16;
    float4 *a, *b;
    float *out;
    if (hipMalloc(&a, bytes) || hipMalloc(&b, bytes) || hipMalloc(&out, 1 << 20)) return 1;
    (void)hipMemset(a, 0, bytes);
    (void)hipMemset(b, 0, bytes);
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    const int grids[] = {256 * 8, 256 * 32, 256 * 128};
    for (int g : grids) {
        for (int kind = 0; kind < 4; ++kind) {
            auto launch = [&]() {
                if (kind == 0) hipLaunchKernelGGL(k_read, dim3(g), dim3(256), 0, 0, a, n4, out);
                else if (kind == 1) hipLaunchKernelGGL(k_write, dim3(g), dim3(256), 0, 0, b, n4);
                else if (kind == 2) hipLaunchKernelGGL(k_copy, dim3(g), dim3(256), 0, 0, a, b, n4);
                else hipLaunchKernelGGL(k_copy1, dim3(g), dim3(256), 0, 0, (const float *)a, (float *)b, n4 * 4);
            };
            launch();
            (void)hipDeviceSynchronize();
            (void)hipEventRecord(e0);
            for (int r = 0; r < 10; ++r) launch();
            (void)hipEventRecord(e1);
            (void)hipEventSynchronize(e1);
            float ms = 0;
            (void)hipEventElapsedTime(&ms, e0, e1);
            const double moved = (kind >= 2 ? 2.0 : 1.0) * bytes * 10;
            const char *nm[] = {"read", "write", "copy16", "copy4"};
            printf("grid %6d %-7s %.3f ms/launch  %.0f GB/s\n", g, nm[kind], ms / 10, moved / (ms * 1e-3) / 1e9);
        }
    }
    return 0;
}
