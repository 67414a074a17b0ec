// bf_voxel.hip — solid voxels of a closed triangle mesh (ray parity along z),
// for the boundary ids that bf_init_fields takes.
//
// Replaces MeshRasterizer3D::rasterize (BoltzmannFluids/Source/MeshRasterizer3D/
// MeshRasterizer3D.cpp:7-100, GLSL/MeshRasterizer3D/raster_unifier.comp): the
// reference draws the mesh in 64 thin depth slices per voxel with an
// orthographic camera at z = +1 looking down -z over the cube [-1, 1]^3, flips
// a per-pixel inside flag wherever a slice holds a surface fragment, and marks
// a voxel when the flag was set in any of its slices.  Here every voxel column
// intersects its centre line with all triangles directly:
//   voxel (i, j, k) covers model x in [-1 + 2i/X, -1 + 2(i+1)/X], y likewise,
//   and z in [1 - 2(k+1)/Z, 1 - 2k/Z] (slice 0 nearest the camera at +z);
//   column (i, j) is the line x = (2i+1)/X - 1, y = (2j+1)/Y - 1;
//   the crossings of that line with the mesh, sorted by z, pair up into inside
//   intervals, and a voxel is solid when its z range overlaps one of them with
//   positive length (the limit of the reference's sub-slice union).
// A line through an edge or vertex of the projected mesh is decided by a fixed
// symbolic perturbation of the line (x + e, y + e^2): each edge function is
// evaluated with the edge's endpoints in index order, so the two triangles
// sharing an edge always disagree about the line and it crosses exactly one.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <string>

#include "../../include/bf/bf.h"

namespace {

constexpr int kMaxCrossings = 64;  // per column
constexpr int kBatch = 256;         // triangles staged in LDS per pass

thread_local std::string g_vox_error;

bf_status vfail(bf_status st, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_vox_error = buf;
    return st;
}

// p' = M p with M column-major (glm::mat4 layout), summed in a fixed order
__global__ void __launch_bounds__(256) k_transform(const float *__restrict__ v, int64_t n, const float *__restrict__ m,
                                                   float *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float x = v[3 * i], y = v[3 * i + 1], z = v[3 * i + 2];
#pragma unroll
    for (int r = 0; r < 3; ++r) out[3 * i + r] = ((m[r] * x + m[4 + r] * y) + m[8 + r] * z) + m[12 + r];
}

// Edge function of the line point p against edge (a -> b) taken in index order
// (ia < ib), with the perturbation deciding exact zeros; the sign is returned
// for the direction the triangle walks the edge.
__device__ inline float edge_fn(const float *a, const float *b, bool reversed, float px, float py, bool &positive) {
    const float e = (b[0] - a[0]) * (py - a[1]) - (b[1] - a[1]) * (px - a[0]);
    bool pos = e > 0.0f;
    if (e == 0.0f) pos = (b[1] != a[1]) ? (b[1] < a[1]) : (b[0] > a[0]);
    positive = reversed ? !pos : pos;
    return reversed ? -e : e;
}

__global__ void __launch_bounds__(256) k_columns(const float *__restrict__ tv, const uint32_t *__restrict__ tri,
                                                 int64_t ntri, int X, int Y, int Z, uint8_t value,
                                                 uint8_t *__restrict__ vol, int *__restrict__ overflow) {
    __shared__ float sv[kBatch][9];
    __shared__ uint32_t sidx[kBatch][3];
    const int64_t col = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const bool live = col < (int64_t)X * Y;
    const int i = live ? (int)(col % X) : 0, j = live ? (int)(col / X) : 0;
    const float px = (float)(2 * i + 1) / (float)X - 1.0f;
    const float py = (float)(2 * j + 1) / (float)Y - 1.0f;
    float zc[kMaxCrossings];
    int nc = 0;
    bool over = false;
    for (int64_t t0 = 0; t0 < ntri; t0 += kBatch) {
        __syncthreads();
        for (int k = threadIdx.x; k < kBatch; k += 256) {
            const int64_t t = t0 + k;
            if (t < ntri) {
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const uint32_t vi = tri[3 * t + c];
                    sidx[k][c] = vi;
                    sv[k][3 * c] = tv[3 * (size_t)vi];
                    sv[k][3 * c + 1] = tv[3 * (size_t)vi + 1];
                    sv[k][3 * c + 2] = tv[3 * (size_t)vi + 2];
                }
            }
        }
        __syncthreads();
        if (!live) continue;
        const int nb = (int)std::min<int64_t>(kBatch, ntri - t0);
        for (int k = 0; k < nb; ++k) {
            bool p[3];
            float e[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const int c1 = (c + 1) % 3;
                const bool rev = sidx[k][c] > sidx[k][c1];
                const float *a = rev ? sv[k] + 3 * c1 : sv[k] + 3 * c;
                const float *b = rev ? sv[k] + 3 * c : sv[k] + 3 * c1;
                // edge opposite vertex (c + 2) % 3 carries that vertex's weight
                e[(c + 2) % 3] = edge_fn(a, b, rev, px, py, p[(c + 2) % 3]);
            }
            if (!((p[0] && p[1] && p[2]) || (!p[0] && !p[1] && !p[2]))) continue;
            const float s = (e[0] + e[1]) + e[2];
            if (s == 0.0f) continue;  // triangle seen edge-on
            const float z = ((e[0] * sv[k][2] + e[1] * sv[k][5]) + e[2] * sv[k][8]) / s;
            if (nc < kMaxCrossings) zc[nc++] = z;
            else over = true;
        }
    }
    if (!live) return;
    if (over) { atomicAdd(overflow, 1); return; }
    // insertion sort (few crossings per column)
    for (int a = 1; a < nc; ++a) {
        const float v = zc[a];
        int b = a - 1;
        while (b >= 0 && zc[b] > v) { zc[b + 1] = zc[b]; --b; }
        zc[b + 1] = v;
    }
    for (int k = 0; k < Z; ++k) {
        const float hi = 1.0f - (float)(2 * k) / (float)Z, lo = 1.0f - (float)(2 * k + 2) / (float)Z;
        bool in = false;
        for (int a = 0; a + 1 < nc && !in; a += 2) in = zc[a] < hi && zc[a + 1] > lo;
        if (in) vol[(size_t)k * X * Y + col] = value;
    }
}

struct Mem {
    void *p[4] = {nullptr, nullptr, nullptr, nullptr};
    ~Mem() {
        for (void *q : p)
            if (q) (void)hipFree(q);
    }
};

}  // namespace

extern "C" {

const char *bf_voxelize_last_error(void) { return g_vox_error.c_str(); }

bf_status bf_voxelize_mesh(int32_t device, const float *vertices, int64_t vertex_count, const uint32_t *indices,
                           int64_t triangle_count, const float model[16], const int32_t resolution[3], uint8_t value,
                           uint8_t *volume, int64_t *solid_count) {
    if (!vertices || !indices || !model || !resolution || !volume || vertex_count < 0 || triangle_count < 0)
        return vfail(BF_ERR_INVALID, "[MeshRasterizer3D Error] null or negative argument");
    const int X = resolution[0], Y = resolution[1], Z = resolution[2];
    if (X <= 0 || Y <= 0 || Z <= 0) return vfail(BF_ERR_INVALID, "[MeshRasterizer3D Error] resolution must be positive");
    for (int64_t k = 0; k < 3 * triangle_count; ++k)
        if ((int64_t)indices[k] >= vertex_count)
            return vfail(BF_ERR_INVALID, "[MeshRasterizer3D Error] index %lld of triangle %lld is out of range",
                         (long long)indices[k], (long long)(k / 3));
    const size_t nvox = (size_t)X * Y * Z;
    if (hipSetDevice(device) != hipSuccess) return vfail(BF_ERR_DEVICE, "[MeshRasterizer3D Error] no device %d", device);
    Mem m;
    float *dv = nullptr, *dt = nullptr, *dm = nullptr;
    uint32_t *di = nullptr;
    uint8_t *dvol = nullptr;
    int *dover = nullptr;
    auto alloc = [&](void **p, size_t bytes) { return hipMalloc(p, std::max<size_t>(bytes, 16)) == hipSuccess; };
    if (!alloc((void **)&dv, sizeof(float) * 3 * vertex_count)) return vfail(BF_ERR_OOM, "[MeshRasterizer3D Error] out of memory");
    m.p[0] = dv;
    // transformed vertices, then the matrix (16 floats) and the overflow counter
    if (!alloc((void **)&dt, sizeof(float) * (3 * vertex_count + 16) + sizeof(int))) return vfail(BF_ERR_OOM, "[MeshRasterizer3D Error] out of memory");
    m.p[1] = dt;
    dm = dt + 3 * vertex_count;
    dover = reinterpret_cast<int *>(dm + 16);
    if (!alloc((void **)&di, sizeof(uint32_t) * 3 * triangle_count)) return vfail(BF_ERR_OOM, "[MeshRasterizer3D Error] out of memory");
    m.p[2] = di;
    if (!alloc((void **)&dvol, nvox)) return vfail(BF_ERR_OOM, "[MeshRasterizer3D Error] out of memory");
    m.p[3] = dvol;
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess)
        return vfail(BF_ERR_DEVICE, "[MeshRasterizer3D Error] stream creation failed");
    bf_status st = BF_OK;
    int over = 0;
    do {
        if (hipMemcpyAsync(dv, vertices, sizeof(float) * 3 * vertex_count, hipMemcpyHostToDevice, s) != hipSuccess ||
            hipMemcpyAsync(di, indices, sizeof(uint32_t) * 3 * triangle_count, hipMemcpyHostToDevice, s) != hipSuccess ||
            hipMemcpyAsync(dm, model, sizeof(float) * 16, hipMemcpyHostToDevice, s) != hipSuccess ||
            hipMemcpyAsync(dvol, volume, nvox, hipMemcpyHostToDevice, s) != hipSuccess ||
            hipMemsetAsync(dover, 0, sizeof(int), s) != hipSuccess) {
            st = vfail(BF_ERR_DEVICE, "[MeshRasterizer3D Error] upload failed");
            break;
        }
        if (vertex_count > 0)
            hipLaunchKernelGGL(k_transform, dim3((unsigned)((vertex_count + 255) / 256)), dim3(256), 0, s, dv,
                               vertex_count, dm, dt);
        const int64_t cols = (int64_t)X * Y;
        hipLaunchKernelGGL(k_columns, dim3((unsigned)((cols + 255) / 256)), dim3(256), 0, s, dt, di, triangle_count, X,
                           Y, Z, value, dvol, dover);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(volume, dvol, nvox, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipMemcpyAsync(&over, dover, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess) {
            st = vfail(BF_ERR_DEVICE, "[MeshRasterizer3D Error] voxelisation failed");
            break;
        }
        if (over)
            st = vfail(BF_ERR_UNSUPPORTED, "[MeshRasterizer3D Error] %d columns cross the mesh more than %d times", over,
                       kMaxCrossings);
    } while (false);
    (void)hipStreamDestroy(s);
    if (st == BF_OK && solid_count) {
        int64_t n = 0;
        for (size_t c = 0; c < nvox; ++c) n += volume[c] == value;
        *solid_count = n;
    }
    return st;
}

}  // extern "C"
