// bf_probe.hip — device copy-bandwidth probe.
//
// Streams `bytes` from one device buffer to another with 16-byte accesses and
// reports the sustained rate (read + write bytes per second).  It gives the
// achievable HBM bandwidth the tick kernel is compared against, and a kernel of
// known traffic for calibrating the FETCH_SIZE / WRITE_SIZE counters.
#include <hip/hip_runtime.h>


#include "../../include/bf/bf.h"

namespace {

// Each block copies 4 x 256 consecutive float4 (16 KiB); all four loads are
// issued before the stores, with streaming (nontemporal) hints.
typedef float v4f __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(256) k_copy(const v4f *__restrict__ src, v4f *__restrict__ dst, size_t n) {
    const size_t base = (size_t)blockIdx.x * 1024 + threadIdx.x;
    v4f v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const size_t i = base + (size_t)k * 256;
        if (i < n) v[k] = __builtin_nontemporal_load(src + i);
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const size_t i = base + (size_t)k * 256;
        if (i < n) __builtin_nontemporal_store(v[k], dst + i);
    }
}

}  // namespace

extern "C" bf_status bf_probe_copy_bandwidth(int32_t device, int64_t bytes, int32_t iters, double *gbps) {
    if (!gbps || bytes < 4096 || iters < 1) return BF_ERR_INVALID;
    if (hipSetDevice(device) != hipSuccess) return BF_ERR_DEVICE;
    const size_t n = (size_t)bytes / sizeof(float4);
    float4 *a = nullptr, *b = nullptr;
    if (hipMalloc((void **)&a, n * sizeof(float4)) != hipSuccess) return BF_ERR_OOM;
    if (hipMalloc((void **)&b, n * sizeof(float4)) != hipSuccess) { (void)hipFree(a); return BF_ERR_OOM; }
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    bf_status st = BF_OK;
    float ms = 0.0f;
    const unsigned grid = (unsigned)((n + 1023) / 1024);
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess) {
        st = BF_ERR_DEVICE;
    } else {
        (void)hipMemsetAsync(a, 0, n * sizeof(float4), s);
        hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, s, (const v4f *)a, (v4f *)b, n);  // warm-up
        (void)hipEventRecord(e0, s);
        for (int k = 0; k < iters; ++k)
            hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, s, (const v4f *)a, (v4f *)b, n);
        (void)hipEventRecord(e1, s);
        if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) st = BF_ERR_DEVICE;
    }
    if (st == BF_OK) *gbps = 2.0 * (double)(n * sizeof(float4)) * iters / (ms * 1e-3) / 1e9;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    (void)hipFree(a);
    (void)hipFree(b);
    return st;
}
