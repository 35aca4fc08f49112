// mpmw_sort.hip -- device-wide primitives for particle binning (hipCUB):
// a stable radix sort of (cell key, particle index) pairs and an exclusive
// scan of per-cell counts.  Kept in its own translation unit because the
// rocPRIM templates dominate compile time.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "mpmw_sort.hpp"

namespace mpmw {

static hipError_t ensure_temp(ScratchBuf &t, size_t need) {
    if (need <= t.bytes) return hipSuccess;
    if (t.ptr) (void)hipFree(t.ptr);
    t.ptr = nullptr;
    t.bytes = 0;
    hipError_t e = hipMalloc(&t.ptr, need);
    if (e == hipSuccess) t.bytes = need;
    return e;
}

hipError_t radix_sort_pairs(ScratchBuf &temp, const uint32_t *kin, uint32_t *kout, const int32_t *vin,
                            int32_t *vout, int n, int nbits, hipStream_t st) {
    size_t need = 0;
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(nullptr, need, kin, kout, vin, vout, n, 0, nbits, st);
    if (e != hipSuccess) return e;
    if ((e = ensure_temp(temp, need)) != hipSuccess) return e;
    size_t have = temp.bytes;
    return hipcub::DeviceRadixSort::SortPairs(temp.ptr, have, kin, kout, vin, vout, n, 0, nbits, st);
}

hipError_t exclusive_scan(ScratchBuf &temp, const int32_t *in, int32_t *out, int n, hipStream_t st) {
    size_t need = 0;
    hipError_t e = hipcub::DeviceScan::ExclusiveSum(nullptr, need, in, out, n, st);
    if (e != hipSuccess) return e;
    if ((e = ensure_temp(temp, need)) != hipSuccess) return e;
    size_t have = temp.bytes;
    return hipcub::DeviceScan::ExclusiveSum(temp.ptr, have, in, out, n, st);
}

hipError_t exclusive_scan_temp_bytes(int n, size_t *bytes) {
    *bytes = 0;
    return hipcub::DeviceScan::ExclusiveSum(nullptr, *bytes, (const int32_t *)nullptr, (int32_t *)nullptr, n);
}

hipError_t exclusive_scan_buf(void *temp, size_t bytes, const int32_t *in, int32_t *out, int n, hipStream_t st) {
    return hipcub::DeviceScan::ExclusiveSum(temp, bytes, in, out, n, st);
}

void free_scratch(ScratchBuf &t) {
    if (t.ptr) (void)hipFree(t.ptr);
    t.ptr = nullptr;
    t.bytes = 0;
}

}  // namespace mpmw
