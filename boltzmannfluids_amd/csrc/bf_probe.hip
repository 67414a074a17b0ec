// bf_probe.hip — device bandwidth probes (measurement infrastructure).
//
// Built into its own library, libbfprobe.so, which only the benchmark loads;
// the solver library libbfamd.so does not contain these kernels.  Entry points
// are declared in bf_probe.h and return the bf_status codes of bf.h.
//
// bfp_copy_bandwidth streams `bytes` from one device buffer to another
// with 16-byte accesses and reports the sustained rate (read + write bytes per
// second): the achievable HBM bandwidth of a plain copy, and a kernel of known
// traffic for calibrating the FETCH_SIZE / WRITE_SIZE counters.
//
// bfp_inplace_bandwidth reproduces the memory traffic of one lattice tick
// without its arithmetic: every thread reads one 4-byte element of each of
// `slots` slots of its cell, in the solver's slab layout (slabs of 2^kSlabLog
// cells holding the slots one after the other), and writes a value back to the
// same places.  Its rate is the ceiling of the in-place tick's access pattern.
#include <hip/hip_runtime.h>

#include "bf_probe.h"
#include "bf_kernels.h"
#include "bf_lattice.h"

namespace {

typedef float v4f __attribute__((ext_vector_type(4)));

// one float4 per thread, streaming (nontemporal) hints
__global__ void __launch_bounds__(256) k_copy(const v4f *__restrict__ src, v4f *__restrict__ dst, size_t n) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

constexpr int kMaxSlots = 27;

// all loads of a thread are issued before its stores, as in the tick kernel
template <int NS>
__global__ void __launch_bounds__(256) k_inplace(float *__restrict__ f, int64_t cells) {
    const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (c >= cells) return;
    constexpr int64_t P = int64_t(1) << bf::kSlabLog;
    float *p = f + (c >> bf::kSlabLog) * (NS * P) + (c & (P - 1));
    float v[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) v[s] = __builtin_nontemporal_load(p + s * P);
#pragma unroll
    for (int s = 0; s < NS; ++s) __builtin_nontemporal_store(0.5f * v[s] + v[(s + 1) % NS], p + s * P);
}

template <int NS>
void launch_inplace(float *f, int64_t cells, hipStream_t s) {
    hipLaunchKernelGGL((k_inplace<NS>), dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, s, f, cells);
}

void launch_inplace_n(int slots, float *f, int64_t cells, hipStream_t s) {
    switch (slots) {
    case 5: launch_inplace<5>(f, cells, s); break;
    case 7: launch_inplace<7>(f, cells, s); break;
    case 9: launch_inplace<9>(f, cells, s); break;
    case 15: launch_inplace<15>(f, cells, s); break;
    case 19: launch_inplace<19>(f, cells, s); break;
    case 27: launch_inplace<27>(f, cells, s); break;
    default: break;
    }
}

// The D3Q19 AA pattern on a periodic X x Y x Z grid in the slab layout, one
// block per x-row: on the odd tick every cell reads and rewrites its 19 slots
// at the cell itself; on the even tick it reads slot i in row (y - c_iy,
// z - c_iz) and writes slot opp(i) in row (y + c_iy, z + c_iz), each at its
// own x (the tick kernel moves the x shift through LDS, so memory sees these
// rows).  The read set of every cell is its write set, as in the tick.
template <bool EVEN>
__global__ void __launch_bounds__(1024) k_aa_pattern(float *__restrict__ f, int X, int Y, int Z) {
    using L = bf::Lattice<bf::D3Q19>;
    const int x = threadIdx.x, y = blockIdx.x, z = blockIdx.y;
    if (x >= X) return;
    const int ys[3] = {y == 0 ? Y - 1 : y - 1, y, y == Y - 1 ? 0 : y + 1};
    const int zs[3] = {z == 0 ? Z - 1 : z - 1, z, z == Z - 1 ? 0 : z + 1};
    constexpr int64_t P = int64_t(1) << bf::kSlabLog;
    auto el = [&](int i, int yy, int zz) {
        const int64_t c = x + (int64_t)X * (yy + (int64_t)Y * zz);
        return ((c >> bf::kSlabLog) * L::Q + i) * P + (c & (P - 1));
    };
    float v[L::Q];
#pragma unroll
    for (int i = 0; i < L::Q; ++i)
        v[i] = __builtin_nontemporal_load(f + (EVEN ? el(i, ys[1 - L::c[i][1]], zs[1 - L::c[i][2]]) : el(i, y, z)));
#pragma unroll
    for (int i = 0; i < L::Q; ++i) {
        const int o = bf::opposite(i);
        const int64_t e = EVEN ? el(o, ys[1 + L::c[i][1]], zs[1 + L::c[i][2]]) : el(o, y, z);
        __builtin_nontemporal_store(0.5f * v[i] + v[(i + 1) % L::Q], f + e);
    }
}

// The D3Q19 two-lattice pull pattern (for comparison with the AA pattern
// above): slot i is read from lattice `a` in row (y - c_iy, z - c_iz) and
// written to lattice `b` at the cell itself, each at its own x.  Launches
// alternate the two lattices, so every launch reads what the last one wrote.
__global__ void __launch_bounds__(1024) k_pull_pattern(const float *__restrict__ a, float *__restrict__ b, int X,
                                                       int Y, int Z) {
    using L = bf::Lattice<bf::D3Q19>;
    const int x = threadIdx.x, y = blockIdx.x, z = blockIdx.y;
    if (x >= X) return;
    const int ys[3] = {y == 0 ? Y - 1 : y - 1, y, y == Y - 1 ? 0 : y + 1};
    const int zs[3] = {z == 0 ? Z - 1 : z - 1, z, z == Z - 1 ? 0 : z + 1};
    constexpr int64_t P = int64_t(1) << bf::kSlabLog;
    auto el = [&](int i, int yy, int zz) {
        const int64_t c = x + (int64_t)X * (yy + (int64_t)Y * zz);
        return ((c >> bf::kSlabLog) * L::Q + i) * P + (c & (P - 1));
    };
    float v[L::Q];
#pragma unroll
    for (int i = 0; i < L::Q; ++i)
        v[i] = __builtin_nontemporal_load(a + el(i, ys[1 - L::c[i][1]], zs[1 - L::c[i][2]]));
#pragma unroll
    for (int i = 0; i < L::Q; ++i)
        __builtin_nontemporal_store(0.5f * v[i] + v[(i + 1) % L::Q], b + el(i, y, z));
}

// Times `iters` launches of `launch` after one warm-up launch on a private stream.
template <class F>
bf_status time_launches(F launch, int32_t iters, float *ms) {
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    bf_status st = BF_OK;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess) {
        st = BF_ERR_DEVICE;
    } else {
        launch(s);
        (void)hipEventRecord(e0, s);
        for (int k = 0; k < iters; ++k) launch(s);
        (void)hipEventRecord(e1, s);
        if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(ms, e0, e1) != hipSuccess)
            st = BF_ERR_DEVICE;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    return st;
}

}  // namespace

extern "C" bf_status bfp_copy_bandwidth(int32_t device, int64_t bytes, int32_t iters, double *gbps) {
    if (!gbps || bytes < 4096 || iters < 1) return BF_ERR_INVALID;
    if (hipSetDevice(device) != hipSuccess) return BF_ERR_DEVICE;
    const size_t n = (size_t)bytes / sizeof(float4);
    float4 *a = nullptr, *b = nullptr;
    if (hipMalloc((void **)&a, n * sizeof(float4)) != hipSuccess) return BF_ERR_OOM;
    if (hipMalloc((void **)&b, n * sizeof(float4)) != hipSuccess) { (void)hipFree(a); return BF_ERR_OOM; }
    float ms = 0.0f;
    bf_status st = BF_OK;
    if (hipMemset(a, 0, n * sizeof(float4)) != hipSuccess) st = BF_ERR_DEVICE;
    if (st == BF_OK) {
        const unsigned grid = (unsigned)((n + 255) / 256);
        st = time_launches([&](hipStream_t s) {
            hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, s, (const v4f *)a, (v4f *)b, n);
        }, iters, &ms);
    }
    if (st == BF_OK) *gbps = 2.0 * (double)(n * sizeof(float4)) * iters / (ms * 1e-3) / 1e9;
    (void)hipFree(a);
    (void)hipFree(b);
    return st;
}

extern "C" bf_status bfp_inplace_bandwidth(int32_t device, int64_t cells, int32_t slots, int32_t iters,
                                                double *gbps) {
    if (!gbps || cells < 64 || iters < 1) return BF_ERR_INVALID;
    if (!(slots == 5 || slots == 7 || slots == 9 || slots == 15 || slots == 19 || slots == kMaxSlots))
        return BF_ERR_INVALID;
    if (hipSetDevice(device) != hipSuccess) return BF_ERR_DEVICE;
    // whole slabs: the last one is partly used
    const int64_t P = int64_t(1) << bf::kSlabLog;
    const size_t bytes = sizeof(float) * (size_t)((cells + P - 1) / P * P) * slots;
    float *f = nullptr;
    if (hipMalloc((void **)&f, bytes) != hipSuccess) return BF_ERR_OOM;
    float ms = 0.0f;
    bf_status st = BF_OK;
    if (hipMemset(f, 0, bytes) != hipSuccess) st = BF_ERR_DEVICE;
    if (st == BF_OK)
        st = time_launches([&](hipStream_t s) { launch_inplace_n(slots, f, cells, s); }, iters, &ms);
    if (st == BF_OK) *gbps = 2.0 * sizeof(float) * (double)cells * slots * iters / (ms * 1e-3) / 1e9;
    (void)hipFree(f);
    return st;
}

extern "C" bf_status bfp_aa_bandwidth(int32_t device, int32_t X, int32_t Y, int32_t Z, int32_t even,
                                           int32_t iters, double *gbps) {
    if (!gbps || iters < 1 || X < 1 || X > 1024 || Y < 1 || Y > 65535 || Z < 1 || Z > 65535) return BF_ERR_INVALID;
    if (hipSetDevice(device) != hipSuccess) return BF_ERR_DEVICE;
    const int64_t cells = (int64_t)X * Y * Z;
    const int64_t P = int64_t(1) << bf::kSlabLog;
    const size_t bytes = sizeof(float) * (size_t)((cells + P - 1) / P * P) * 19;
    float *f = nullptr;
    if (hipMalloc((void **)&f, bytes) != hipSuccess) return BF_ERR_OOM;
    float ms = 0.0f;
    bf_status st = BF_OK;
    if (hipMemset(f, 0, bytes) != hipSuccess) st = BF_ERR_DEVICE;
    const dim3 grid((unsigned)Y, (unsigned)Z), block((unsigned)((X + 63) / 64 * 64));
    if (st == BF_OK)
        st = time_launches([&](hipStream_t s) {
            if (even) hipLaunchKernelGGL(k_aa_pattern<true>, grid, block, 0, s, f, X, Y, Z);
            else hipLaunchKernelGGL(k_aa_pattern<false>, grid, block, 0, s, f, X, Y, Z);
        }, iters, &ms);
    if (st == BF_OK) *gbps = 2.0 * sizeof(float) * 19.0 * (double)cells * iters / (ms * 1e-3) / 1e9;
    (void)hipFree(f);
    return st;
}

extern "C" bf_status bfp_pull_bandwidth(int32_t device, int32_t X, int32_t Y, int32_t Z, int32_t iters,
                                             double *gbps) {
    if (!gbps || iters < 1 || X < 1 || X > 1024 || Y < 1 || Y > 65535 || Z < 1 || Z > 65535) return BF_ERR_INVALID;
    if (hipSetDevice(device) != hipSuccess) return BF_ERR_DEVICE;
    const int64_t cells = (int64_t)X * Y * Z;
    const int64_t P = int64_t(1) << bf::kSlabLog;
    const size_t bytes = sizeof(float) * (size_t)((cells + P - 1) / P * P) * 19;
    float *a = nullptr, *b = nullptr;
    if (hipMalloc((void **)&a, bytes) != hipSuccess) return BF_ERR_OOM;
    if (hipMalloc((void **)&b, bytes) != hipSuccess) { (void)hipFree(a); return BF_ERR_OOM; }
    float ms = 0.0f;
    bf_status st = BF_OK;
    if (hipMemset(a, 0, bytes) != hipSuccess || hipMemset(b, 0, bytes) != hipSuccess) st = BF_ERR_DEVICE;
    const dim3 grid((unsigned)Y, (unsigned)Z), block((unsigned)((X + 63) / 64 * 64));
    int k = 0;
    if (st == BF_OK)
        st = time_launches([&](hipStream_t s) {
            float *src = (k & 1) ? b : a, *dst = (k & 1) ? a : b;
            hipLaunchKernelGGL(k_pull_pattern, grid, block, 0, s, src, dst, X, Y, Z);
            ++k;
        }, iters, &ms);
    if (st == BF_OK) *gbps = 2.0 * sizeof(float) * 19.0 * (double)cells * iters / (ms * 1e-3) / 1e9;
    (void)hipFree(a);
    (void)hipFree(b);
    return st;
}
