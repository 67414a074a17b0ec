// bf_solver.hip — host runtime behind the C ABI of include/bf/bf.h.
//
// Owns the device lattice, the boundary / force / thermal / pseudopotential
// fields and one HIP stream per handle, and sequences the fused tick kernel.
// It takes the place of the reference's LBM host class
// (BoltzmannFluids/Source/LBM/LBM.cpp): field initialisation with the same
// detection rules (LBM.cpp:508-591), the object table (LBM.cpp:1015-1070),
// the tick loop with parity (LBM.cpp:220-240,1557-1653) and field readback.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include <sys/stat.h>

#include "../../include/bf/bf.h"
#include "bf_kernels.h"
#include "bf_lattice.h"

using namespace bf;

struct bf_solver {
    bf_config cfg{};
    int vs = 0, q = 0, qt = 0;
    int X = 0, Y = 0, Z = 0;
    int64_t N = 0, stride = 0;  // stride: thermal slots (and lattice slots without slabs)
    // lattice layout (described with StepArgs in bf_kernels.h)
    uint32_t slab_log = 31, slab_skip = 0, pair_step = 0;
    int group_width = 0, groups = 0;  // 0 groups: plain structure of arrays
    size_t slot_off[kMaxQ] = {};      // byte offset of slot i from f
    size_t lattice_bytes = 0;
    int storage = ST_FP32;
    size_t elem = 4;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;

    // device memory
    void *f = nullptr;
    float *g = nullptr;
    uint8_t *ids = nullptr;
    uint32_t *nmask = nullptr;  // solid-neighbour masks (with ids)
    ObjectEntry *objects_d = nullptr;
    float *rho_eff_d = nullptr;
    float4 *force_d = nullptr;
    float *psi[2] = {nullptr, nullptr};
    float4 *vd = nullptr, *ft = nullptr;
    double *partial = nullptr;
    size_t partial_blocks = 0, partial_capacity = 0;
    double *sums_d = nullptr;
    int64_t device_bytes = 0;

    // host state
    ObjectEntry objects_h[256];
    float rho_eff_h[256];
    int object_count = 1;  // objects_cpu.size() in the reference (slot 0 = fluid)
    float tau = 0.53f, G = -6.0f, beta = 0.5f, tau_t = 2.0f, t_ref = 1.0f, cs2_t = 1.0f / 3.0f;
    float ulim = 0.25f, uext = 0.30f;
    int sc_sync = 0;  // synchronous Shan-Chen pseudopotential (density pre-pass per tick)
    int pair_planes = 0;  // chunked tick pairs: 0 automatic, -1 off, k > 0 chunks of k planes
    int sweep_ticks = 0;  // ticks per chunked sweep: 0 automatic, 2 (one pair) or 4 (two pairs)
    float fconst[3] = {0, 0, 0};
    int forcing = 0, thermal = 0, has_ids = 0, initialized = 0;
    int closed = 0;  // see StepArgs::closed
    int cur = 0;
    int64_t ticks = 0;
    int64_t tick_base = 0;  // tick count at the last equilibrium initialisation: AA parity = ticks - tick_base
    LaunchShape shape{};
    std::string err;
};

static std::string g_create_error;

// Slot stride of the thermal lattice and of a lattice without slabs: the cell
// count rounded up to 64 elements (256-byte aligned slots) plus a pad of 16384
// elements, which places the slot bases at different offsets modulo the
// memory interleave (measured 0.6-2 % faster than no pad for 19 plain slots).
constexpr int64_t kSlotPad = 16384;
// Lattice slabs of 2^kSlabLog cells.  With all 19 slots of a slab in one
// region (608 KB in FP32) instead of 19 arrays a lattice apart, an in-place
// pass over 19 slots moves 6.4-6.5 instead of 6.0 TB/s at 256^3 and 6.3-6.4
// instead of 5.5 TB/s at 384^3 (measured on the MI355X for 2^12 to 2^14
// cells; 2^18 and more approach the plain arrays again).  2^13 (kSlabLog)
// gave the fastest tick.
// Largest grid: a slot group of one opposite pair of 16-bit slots (slabs of
// 2 x 8192 elements) and, for FP32 lattices too large for slabs, one plain
// slot array (N + 63 + kSlotPad elements) stay within 4 GiB of 32-bit byte
// offsets.
constexpr int64_t kMaxCells = (int64_t(1) << 30) - (int64_t(1) << 14);

namespace {

bf_status fail(bf_solver *s, bf_status st, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (s) s->err = buf; else g_create_error = buf;
    return st;
}

#define BF_HIP(s, call)                                                                        \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return fail((s), e_ == hipErrorOutOfMemory ? BF_ERR_OOM : BF_ERR_DEVICE,          \
                        "[LBM Error] %s failed: %s", #call, hipGetErrorString(e_));            \
    } while (0)

template <class T>
bf_status dalloc(bf_solver *s, T **p, size_t bytes) {
    BF_HIP(s, hipMalloc((void **)p, bytes));
    s->device_bytes += (int64_t)bytes;
    return BF_OK;
}

void dfree(void *p) {
    if (p) (void)hipFree(p);
}

// Lattice layout: slabs of 2^kSlabLog cells, the slots in the fewest groups
// of equal width whose arrays stay below 4 GiB (the kernels address a slot
// with 32-bit byte offsets from its base).  Group 0 holds the rest slot and
// the first opposite pairs, each further group the next pairs; a group's
// unused positions are padding.  Grids that need more than 4 GiB even for one
// pair per group keep plain slot arrays.
void choose_layout(bf_solver *s) {
    const int64_t P = int64_t(1) << kSlabLog;
    const int64_t nslabs = (s->N + P - 1) / P;
    const int pairs = (s->q - 1) / 2;
    for (int G = 1; G <= pairs + 1; ++G) {
        int W = 2;
        while ((W - 1) / 2 + (G - 1) * (W / 2) < pairs) ++W;
        const uint64_t region = (uint64_t)nslabs * P * W * s->elem;
        if (region > (uint64_t(1) << 32)) continue;
        s->groups = G;
        s->group_width = W;
        s->slab_log = kSlabLog;
        s->slab_skip = (uint32_t)((W - 1) * P);
        s->pair_step = (uint32_t)P;
        s->lattice_bytes = (size_t)G * region;
        int g = 0, pos = 1;
        s->slot_off[0] = 0;
        for (int j = 0; j < pairs; ++j) {
            if (pos + 2 > W) { ++g; pos = 0; }
            for (int h = 0; h < 2; ++h, ++pos)
                s->slot_off[1 + 2 * j + h] = (size_t)g * region + (size_t)pos * P * s->elem;
        }
        return;
    }
    s->groups = 0;
    s->group_width = 0;
    s->slab_log = 31;
    s->slab_skip = 0;
    s->pair_step = (uint32_t)s->stride;
    s->lattice_bytes = (size_t)s->q * s->stride * s->elem;
    for (int i = 0; i < s->q; ++i) s->slot_off[i] = (size_t)i * s->stride * s->elem;
}

// Copies slot i between the device layout and N contiguous host elements.
hipError_t copy_slot(const bf_solver *s, int i, void *host, bool to_host) {
    char *dev = static_cast<char *>(s->f) + s->slot_off[i];
    const size_t e = s->elem, N = (size_t)s->N;
    if (s->groups == 0)
        return to_host ? hipMemcpy(host, dev, N * e, hipMemcpyDeviceToHost)
                       : hipMemcpy(dev, host, N * e, hipMemcpyHostToDevice);
    const size_t P = size_t(1) << s->slab_log, pitch = (size_t)s->group_width * P * e;
    const size_t full = N / P, tail = N % P;
    char *h = static_cast<char *>(host);
    hipError_t r = hipSuccess;
    if (full)
        r = to_host ? hipMemcpy2D(h, P * e, dev, pitch, P * e, full, hipMemcpyDeviceToHost)
                    : hipMemcpy2D(dev, pitch, h, P * e, P * e, full, hipMemcpyHostToDevice);
    if (r == hipSuccess && tail)
        r = to_host ? hipMemcpy(h + full * P * e, dev + full * pitch, tail * e, hipMemcpyDeviceToHost)
                    : hipMemcpy(dev + full * pitch, h + full * P * e, tail * e, hipMemcpyHostToDevice);
    return r;
}

template <int VS>
void table_of(float *out) {
    using L = Lattice<VS>;
    for (int i = 0; i < L::Q; ++i) {
        out[4 * i + 0] = L::c[i][0];
        out[4 * i + 1] = L::c[i][1];
        out[4 * i + 2] = L::c[i][2];
        out[4 * i + 3] = L::w[i];
    }
}

float weight_of(int vs, int i) {
    switch (vs) {
    case D2Q9: return Lattice<D2Q9>::w[i];
    case D3Q15: return Lattice<D3Q15>::w[i];
    case D3Q19: return Lattice<D3Q19>::w[i];
    case D3Q27: return Lattice<D3Q27>::w[i];
    }
    return 0.0f;
}

hipError_t dispatch_step(const bf_solver *s, const StepArgs &a, bool odd, double *partial) {
    const bool walls = s->has_ids != 0, mp = s->cfg.multiphase != 0, th = s->thermal != 0;
    switch (s->vs) {
    case D2Q9: return launch_step<D2Q9>(a, s->storage, odd, walls, mp, th, s->shape, partial, s->stream);
    case D3Q15: return launch_step<D3Q15>(a, s->storage, odd, walls, mp, th, s->shape, partial, s->stream);
    case D3Q19: return launch_step<D3Q19>(a, s->storage, odd, walls, mp, th, s->shape, partial, s->stream);
    case D3Q27: return launch_step<D3Q27>(a, s->storage, odd, walls, mp, th, s->shape, partial, s->stream);
    }
    return hipErrorInvalidValue;
}

// StepArgs::closed from the host ids (cell order, or null for no boundaries):
// every axis periodic, or each non-periodic axis solid on both of its boundary
// planes, so that no cell that updates has a neighbour outside the grid.  The
// z axis of a 2-D velocity set has no neighbours and is ignored.
int closed_domain(const bf_solver *s, const uint8_t *ids) {
    const int64_t n[3] = {s->X, s->Y, s->Z};
    const int dims = (s->vs == D2Q9) ? 2 : 3;
    bool all_periodic = true;
    for (int d = 0; d < dims; ++d) all_periodic = all_periodic && s->cfg.periodic[d] != 0;
    if (all_periodic) return 1;
    if (!ids) return 0;
    for (int d = 0; d < dims; ++d) {
        if (s->cfg.periodic[d]) continue;
        // the two planes coord_d = 0 and coord_d = n[d] - 1
        const int a1 = (d + 1) % 3, a2 = (d + 2) % 3;
        for (int64_t side = 0; side < 2; ++side) {
            const int64_t cd = side ? n[d] - 1 : 0;
            for (int64_t j = 0; j < n[a2]; ++j)
                for (int64_t i = 0; i < n[a1]; ++i) {
                    int64_t co[3];
                    co[d] = cd; co[a1] = i; co[a2] = j;
                    if (ids[co[0] + n[0] * (co[1] + n[1] * co[2])] == 0) return 0;
                }
        }
    }
    return 1;
}

hipError_t dispatch_density(const bf_solver *s, const StepArgs &a, bool odd) {
    const bool walls = s->has_ids != 0;
    switch (s->vs) {
    case D2Q9: return launch_density<D2Q9>(a, s->storage, odd, walls, s->shape, s->stream);
    case D3Q15: return launch_density<D3Q15>(a, s->storage, odd, walls, s->shape, s->stream);
    case D3Q19: return launch_density<D3Q19>(a, s->storage, odd, walls, s->shape, s->stream);
    case D3Q27: return launch_density<D3Q27>(a, s->storage, odd, walls, s->shape, s->stream);
    }
    return hipErrorInvalidValue;
}

hipError_t dispatch_mask(const bf_solver *s) {
    const int px = s->cfg.periodic[0] != 0, py = s->cfg.periodic[1] != 0, pz = s->cfg.periodic[2] != 0;
    switch (s->vs) {
    case D2Q9: return launch_mask<D2Q9>(s->ids, s->nmask, s->X, s->Y, s->Z, px, py, pz, s->stream);
    case D3Q15: return launch_mask<D3Q15>(s->ids, s->nmask, s->X, s->Y, s->Z, px, py, pz, s->stream);
    case D3Q19: return launch_mask<D3Q19>(s->ids, s->nmask, s->X, s->Y, s->Z, px, py, pz, s->stream);
    case D3Q27: return launch_mask<D3Q27>(s->ids, s->nmask, s->X, s->Y, s->Z, px, py, pz, s->stream);
    }
    return hipErrorInvalidValue;
}

hipError_t dispatch_init(const bf_solver *s, const InitArgs &a) {
    const bool th = s->thermal != 0;
    switch (s->vs) {
    case D2Q9: return launch_init<D2Q9>(a, s->storage, th, s->stream);
    case D3Q15: return launch_init<D3Q15>(a, s->storage, th, s->stream);
    case D3Q19: return launch_init<D3Q19>(a, s->storage, th, s->stream);
    case D3Q27: return launch_init<D3Q27>(a, s->storage, th, s->stream);
    }
    return hipErrorInvalidValue;
}

// Deterministic second stage of the FP64 conservation sums: one block adds the
// per-block partials in a fixed order.
__global__ void __launch_bounds__(256) k_reduce_partials(const double *__restrict__ partial, size_t nblocks,
                                                         double *__restrict__ out) {
    __shared__ double red[256][5];
    const int t = threadIdx.x;
    double acc[5] = {0, 0, 0, 0, 0};
    for (size_t b = t; b < nblocks; b += 256)
        for (int k = 0; k < 5; ++k) acc[k] += partial[b * 5 + k];
    for (int k = 0; k < 5; ++k) red[t][k] = acc[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w)
            for (int k = 0; k < 5; ++k) red[t][k] += red[t + w][k];
        __syncthreads();
    }
    if (t < 5) out[t] = red[0][t];
}

StepArgs make_args(bf_solver *s, bool update) {
    StepArgs a{};
    a.f = s->f;
    a.g = s->g;
    a.stride = s->stride;
    for (int i = 0; i < s->q; ++i) a.slot[i] = static_cast<char *>(s->f) + s->slot_off[i];
    a.slab_log = s->slab_log;
    a.slab_skip = s->slab_skip;
    a.pair_step = s->pair_step;
    if (s->g)
        for (int q = 0; q < s->qt; ++q) a.gslot[q] = s->g + (size_t)q * s->stride;
    a.X = s->X; a.Y = s->Y; a.Z = s->Z;
    a.periodic_x = s->cfg.periodic[0] != 0;
    a.periodic_y = s->cfg.periodic[1] != 0;
    a.periodic_z = s->cfg.periodic[2] != 0;
    a.closed = s->closed;
    a.ids = s->ids;
    a.nmask = s->nmask;
    a.objects = s->objects_d;
    a.halfway = s->cfg.boundary_mode == BF_BB_HALFWAY;
    // wall terms read the solid neighbour's object only when one can be non-zero
    a.wall_ids = s->thermal ? 1 : 0;
    for (int i = 1; i < 256 && a.halfway && !a.wall_ids; ++i)
        if (s->objects_h[i].ux != 0.0f || s->objects_h[i].uy != 0.0f || s->objects_h[i].uz != 0.0f) a.wall_ids = 1;
    a.forcing = s->forcing;
    a.fx = s->fconst[0]; a.fy = s->fconst[1]; a.fz = s->fconst[2];
    a.force = s->force_d;
    a.tau = s->tau;
    a.inv_tau = 1.0f / s->tau;
    a.guo_pre = 1.0f - 1.0f / (2.0f * s->tau);
    a.ulim = s->ulim; a.uext = s->uext;
    a.source_on = (s->forcing != 0 || s->cfg.multiphase || s->thermal) ? 1 : 0;
    a.G = s->G;
    if (s->cfg.multiphase) {
        a.psi_prev = s->psi[s->cur ^ 1];
        a.psi_cur = s->psi[s->cur];
    }
    a.write_psi = 1;
    a.tau_t = s->tau_t; a.beta = s->beta; a.t_ref = s->t_ref; a.cs2_t = s->cs2_t;
    a.inv_tau_t = 1.0f / s->tau_t;
    a.k1_t = 1.0f / s->cs2_t;
    a.k2_t = 0.5f * a.k1_t * a.k1_t;
    a.update = update ? 1 : 0;
    return a;
}

bf_status upload_objects(bf_solver *s) {
    if (!s->objects_d) return BF_OK;
    BF_HIP(s, hipMemcpyAsync(s->objects_d, s->objects_h, sizeof(ObjectEntry) * 256, hipMemcpyHostToDevice, s->stream));
    BF_HIP(s, hipMemcpyAsync(s->rho_eff_d, s->rho_eff_h, sizeof(float) * 256, hipMemcpyHostToDevice, s->stream));
    BF_HIP(s, hipStreamSynchronize(s->stream));
    return BF_OK;
}

float half_to_float(uint16_t h) {
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    const uint32_t e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
    float v;
    if (e == 0) v = std::ldexp((float)m, -24);
    else if (e == 31) v = m ? NAN : INFINITY;
    else v = std::ldexp((float)(m | 0x400u), (int)e - 25);
    uint32_t b;
    std::memcpy(&b, &v, 4);
    b |= sign;
    std::memcpy(&v, &b, 4);
    return v;
}

bf_status check_solver(const bf_solver *s) {
    return s ? BF_OK : BF_ERR_INVALID;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI

extern "C" {

int32_t bf_abi_version(void) { return BF_ABI_VERSION; }

int32_t bf_velocity_count(int32_t vs) { return velocity_count(vs); }

bf_status bf_velocity_table(int32_t set, int32_t thermal, float *out, int32_t capacity) {
    if (!out) return BF_ERR_INVALID;
    if (thermal) {
        if (set == BF_D2Q5) {
            if (capacity < 20) return BF_ERR_INVALID;
            float t[36];
            table_of<D2Q9>(t);
            for (int k = 0; k < 5; ++k) {
                out[4 * k] = t[4 * k]; out[4 * k + 1] = t[4 * k + 1]; out[4 * k + 2] = t[4 * k + 2];
                out[4 * k + 3] = Lattice<D2Q9>::wt[k];
            }
            return BF_OK;
        }
        if (set == BF_D3Q7) {
            if (capacity < 28) return BF_ERR_INVALID;
            float t[76];
            table_of<D3Q19>(t);
            for (int k = 0; k < 7; ++k) {
                out[4 * k] = t[4 * k]; out[4 * k + 1] = t[4 * k + 1]; out[4 * k + 2] = t[4 * k + 2];
                out[4 * k + 3] = Lattice<D3Q19>::wt[k];
            }
            return BF_OK;
        }
        return BF_ERR_INVALID;
    }
    const int q = velocity_count(set);
    if (!q || capacity < 4 * q) return BF_ERR_INVALID;
    switch (set) {
    case D2Q9: table_of<D2Q9>(out); break;
    case D3Q15: table_of<D3Q15>(out); break;
    case D3Q19: table_of<D3Q19>(out); break;
    case D3Q27: table_of<D3Q27>(out); break;
    }
    return BF_OK;
}

const char *bf_last_error(const bf_solver *s) { return s ? s->err.c_str() : g_create_error.c_str(); }

bf_status bf_create(const bf_config *cfg, bf_solver **out) {
    if (!cfg || !out) return fail(nullptr, BF_ERR_INVALID, "[LBM Error] bf_create: null argument");
    *out = nullptr;
    const int q = velocity_count(cfg->velocity_set);
    if (!q) return fail(nullptr, BF_ERR_INVALID, "[LBM Error] bf_create: velocity set %d is not supported", cfg->velocity_set);
    const int X = cfg->resolution[0], Y = cfg->resolution[1], Z = cfg->resolution[2];
    // LBM.cpp:429-432
    if (X <= 0 || Y <= 0 || Z <= 0)
        return fail(nullptr, BF_ERR_INVALID,
                    "[LBM Error] bf_create: resolution (%d, %d, %d) is not valid, all components must be greater than zero", X, Y, Z);
    // LBM.cpp:434-437
    if (cfg->velocity_set == BF_D2Q9 && Z != 1)
        return fail(nullptr, BF_ERR_INVALID,
                    "[LBM Error] bf_create: velocity set dimension doesn't match the resolution dimension (2-D set needs z == 1)");
    if (cfg->storage != BF_STORE_FP32 && cfg->storage != BF_STORE_FP16 && cfg->storage != BF_STORE_BF16)
        return fail(nullptr, BF_ERR_INVALID, "[LBM Error] bf_create: storage %d is not supported", cfg->storage);
    if (cfg->boundary_mode != BF_BB_HALFWAY && cfg->boundary_mode != BF_BB_FULLWAY_COMPAT)
        return fail(nullptr, BF_ERR_INVALID, "[LBM Error] bf_create: boundary mode %d is not supported", cfg->boundary_mode);
    if (!(cfg->relaxation_time > 0.5f))
        return fail(nullptr, BF_ERR_INVALID, "[LBM Error] bf_create: relaxation time %g must exceed 0.5", (double)cfg->relaxation_time);
    const int64_t N = (int64_t)X * Y * Z;
    // slots are addressed with 32-bit byte offsets from their base (see
    // kMaxCells); the reference's uint32 element addressing
    // (collide.comp:196-200) allows Q * N < 2^32 only
    if (N > kMaxCells)
        return fail(nullptr, BF_ERR_INVALID, "[LBM Error] bf_create: %lld cells exceed the limit of %lld",
                    (long long)N, (long long)kMaxCells);

    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return fail(nullptr, BF_ERR_DEVICE, "[LBM Error] bf_create: no HIP device available");
    if (cfg->device < 0 || cfg->device >= ndev)
        return fail(nullptr, BF_ERR_INVALID, "[LBM Error] bf_create: device %d out of range (%d devices)", cfg->device, ndev);

    bf_solver *s = new (std::nothrow) bf_solver();
    if (!s) return fail(nullptr, BF_ERR_OOM, "[LBM Error] bf_create: host allocation failed");
    s->cfg = *cfg;
    s->vs = cfg->velocity_set;
    s->q = q;
    s->qt = (cfg->velocity_set == BF_D2Q9) ? 5 : 7;
    s->X = X; s->Y = Y; s->Z = Z;
    s->N = N;
    s->stride = (N + 63) / 64 * 64 + kSlotPad;
    s->storage = cfg->storage;
    s->elem = cfg->storage == BF_STORE_FP32 ? 4 : 2;
    choose_layout(s);
    s->tau = cfg->relaxation_time;
    for (int i = 0; i < 256; ++i) {
        s->objects_h[i] = ObjectEntry{0.0f, 0.0f, 0.0f, 1.0f};  // referance_temperature (LBM.h:30)
        s->rho_eff_h[i] = 1.0f;                                 // referance_boundry_density (LBM.h:31)
    }
    auto bail = [&](bf_status st) {
        g_create_error = s->err;
        bf_destroy(s);
        return st;
    };
    if (hipSetDevice(cfg->device) != hipSuccess) { fail(s, BF_ERR_DEVICE, "[LBM Error] hipSetDevice failed"); return bail(BF_ERR_DEVICE); }
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&s->ev0) != hipSuccess || hipEventCreate(&s->ev1) != hipSuccess) {
        fail(s, BF_ERR_DEVICE, "[LBM Error] stream/event creation failed");
        return bail(BF_ERR_DEVICE);
    }
    bf_status st;
    if ((st = dalloc(s, &s->f, s->lattice_bytes)) != BF_OK) return bail(st);
    if ((st = dalloc(s, &s->vd, sizeof(float4) * N)) != BF_OK) return bail(st);
    if ((st = dalloc(s, &s->ft, sizeof(float4) * N)) != BF_OK) return bail(st);
    if (hipMemsetAsync(s->vd, 0, sizeof(float4) * N, s->stream) != hipSuccess ||
        hipMemsetAsync(s->ft, 0, sizeof(float4) * N, s->stream) != hipSuccess) {
        fail(s, BF_ERR_DEVICE, "[LBM Error] clearing macroscopic fields failed");
        return bail(BF_ERR_DEVICE);
    }
    if ((st = dalloc(s, &s->objects_d, sizeof(ObjectEntry) * 256)) != BF_OK) return bail(st);
    if ((st = dalloc(s, &s->rho_eff_d, sizeof(float) * 256)) != BF_OK) return bail(st);
    if ((st = dalloc(s, &s->sums_d, sizeof(double) * 8)) != BF_OK) return bail(st);
    // the launch shape depends on whether the flow is thermal, known after
    // bf_init_fields; size the partial-sum buffer for the larger grid
    s->shape = step_shape(X, Y, Z, cfg->multiphase != 0, 1, cfg->velocity_set);
    s->partial_blocks = (size_t)s->shape.grid.x * s->shape.grid.y * s->shape.grid.z;
    s->partial_capacity = s->partial_blocks;
    if ((st = dalloc(s, &s->partial, sizeof(double) * 5 * s->partial_capacity)) != BF_OK) return bail(st);
    if (cfg->multiphase) {
        if ((st = dalloc(s, &s->psi[0], sizeof(float) * N)) != BF_OK) return bail(st);
        if ((st = dalloc(s, &s->psi[1], sizeof(float) * N)) != BF_OK) return bail(st);
    }
    if ((st = upload_objects(s)) != BF_OK) return bail(st);
    *out = s;
    return BF_OK;
}

void bf_destroy(bf_solver *s) {
    if (!s) return;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    dfree(s->f); dfree(s->g); dfree(s->ids); dfree(s->nmask); dfree(s->objects_d); dfree(s->rho_eff_d);
    dfree(s->force_d); dfree(s->psi[0]); dfree(s->psi[1]); dfree(s->vd); dfree(s->ft);
    dfree(s->partial); dfree(s->sums_d);
    if (s->ev0) (void)hipEventDestroy(s->ev0);
    if (s->ev1) (void)hipEventDestroy(s->ev1);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

bf_status bf_set_object(bf_solver *s, uint32_t id, const float velocity[3], const float angular_velocity[3],
                        const float center_of_mass[3], float temperature, float effective_density) {
    if (!s) return BF_ERR_INVALID;
    // LBM.cpp:1024-1032
    if (id > 255)
        return fail(s, BF_ERR_INVALID, "[LBM Error] bf_set_object: boundry_id(%u) is greater than maximum(255)", id);
    if (id == 0)
        return fail(s, BF_ERR_INVALID,
                    "[LBM Error] bf_set_object: boundry_id(0) is defined to be fluid, it cannot be treated as an object");
    (void)angular_velocity;  // stored by the reference but not used by its update (stream.comp:563-565)
    (void)center_of_mass;
    const float zero[3] = {0, 0, 0};
    const float *v = velocity ? velocity : zero;
    s->objects_h[id] = ObjectEntry{v[0], v[1], v[2], temperature};
    s->rho_eff_h[id] = effective_density;
    if ((int)id + 1 > s->object_count) s->object_count = (int)id + 1;
    return upload_objects(s);
}

bf_status bf_clear_objects(bf_solver *s) {
    if (!s) return BF_ERR_INVALID;
    for (int i = 0; i < 256; ++i) {
        s->objects_h[i] = ObjectEntry{0.0f, 0.0f, 0.0f, 1.0f};
        s->rho_eff_h[i] = 1.0f;
    }
    s->object_count = 1;
    return upload_objects(s);
}

bf_status bf_init_fields(bf_solver *s, const float *rho, const float *u, const float *force,
                         const float *temperature, const uint8_t *boundary_id) {
    if (!s) return BF_ERR_INVALID;
    if (!rho || !u) return fail(s, BF_ERR_INVALID, "[LBM Error] bf_init_fields: rho and u are required");
    const int64_t N = s->N;
    BF_HIP(s, hipSetDevice(s->cfg.device));
    BF_HIP(s, hipStreamSynchronize(s->stream));

    // boundaries (LBM.cpp:593-691): ids grow the object table like objects_cpu.resize
    dfree(s->ids); s->ids = nullptr;
    dfree(s->nmask); s->nmask = nullptr;
    s->has_ids = 0;
    if (boundary_id) {
        int maxid = 0;
        for (int64_t c = 0; c < N; ++c) maxid = std::max(maxid, (int)boundary_id[c]);
        if (maxid > 0) {
            s->has_ids = 1;
            if (maxid + 1 > s->object_count) s->object_count = maxid + 1;
            bf_status st;
            if ((st = dalloc(s, &s->ids, (size_t)N)) != BF_OK) return st;
            BF_HIP(s, hipMemcpy(s->ids, boundary_id, (size_t)N, hipMemcpyHostToDevice));
            if ((st = dalloc(s, &s->nmask, sizeof(uint32_t) * (size_t)N)) != BF_OK) return st;
            BF_HIP(s, dispatch_mask(s));
        }
    }
    s->closed = closed_domain(s, s->has_ids ? boundary_id : nullptr);
    // force detection (LBM.cpp:531-541,562)
    dfree(s->force_d); s->force_d = nullptr;
    s->forcing = 0;
    if (force) {
        bool constant = true;
        for (int64_t c = 1; c < N && constant; ++c)
            if (force[3 * c] != force[0] || force[3 * c + 1] != force[1] || force[3 * c + 2] != force[2]) constant = false;
        if (constant) {
            s->fconst[0] = force[0]; s->fconst[1] = force[1]; s->fconst[2] = force[2];
            s->forcing = (force[0] != 0.0f || force[1] != 0.0f || force[2] != 0.0f) ? 1 : 0;
        } else {
            s->forcing = 2;
            std::vector<float4> tmp((size_t)N);
            for (int64_t c = 0; c < N; ++c) tmp[c] = make_float4(force[3 * c], force[3 * c + 1], force[3 * c + 2], 0.0f);
            bf_status st;
            if ((st = dalloc(s, &s->force_d, sizeof(float4) * N)) != BF_OK) return st;
            BF_HIP(s, hipMemcpy(s->force_d, tmp.data(), sizeof(float4) * N, hipMemcpyHostToDevice));
        }
    } else {
        s->fconst[0] = s->fconst[1] = s->fconst[2] = 0.0f;  // lambda default force (LBM.h:43)
    }
    // thermal detection (LBM.cpp:542-545)
    s->thermal = 0;
    if (temperature) {
        for (int64_t c = 0; c < N && !s->thermal; ++c) {
            if (temperature[c] != temperature[0]) s->thermal = 1;
            if (s->has_ids && boundary_id[c] && s->objects_h[boundary_id[c]].temperature != temperature[0]) s->thermal = 1;
        }
    } else if (s->has_ids) {
        for (int64_t c = 0; c < N && !s->thermal; ++c)
            if (boundary_id[c] && s->objects_h[boundary_id[c]].temperature != 1.0f) s->thermal = 1;
    }
    dfree(s->g); s->g = nullptr;
    if (s->thermal) {
        bf_status st;
        if ((st = dalloc(s, &s->g, sizeof(float) * (size_t)s->qt * s->stride)) != BF_OK) return st;
    }
    bf_status st = upload_objects(s);
    if (st != BF_OK) return st;
    s->shape = step_shape(s->X, s->Y, s->Z, s->cfg.multiphase != 0,
                          cells_per_thread(s->storage, s->cfg.multiphase != 0, s->thermal != 0), s->vs);
    s->partial_blocks = (size_t)s->shape.grid.x * s->shape.grid.y * s->shape.grid.z;
    if (s->partial_blocks > s->partial_capacity) {
        dfree(s->partial);
        s->partial = nullptr;
        s->partial_capacity = s->partial_blocks;
        if ((st = dalloc(s, &s->partial, sizeof(double) * 5 * s->partial_capacity)) != BF_OK) return st;
    }

    // upload rho, u, T and run the equilibrium initialisation
    float *rho_d = nullptr, *u_d = nullptr, *t_d = nullptr;
    if ((st = dalloc(s, &rho_d, sizeof(float) * N)) != BF_OK) return st;
    if ((st = dalloc(s, &u_d, sizeof(float) * 3 * N)) != BF_OK) { dfree(rho_d); return st; }
    if (temperature && s->thermal) {
        if ((st = dalloc(s, &t_d, sizeof(float) * N)) != BF_OK) { dfree(rho_d); dfree(u_d); return st; }
    }
    auto cleanup = [&]() {
        (void)hipStreamSynchronize(s->stream);
        dfree(rho_d); dfree(u_d); dfree(t_d);
        s->device_bytes -= (int64_t)(sizeof(float) * N * 4 + (t_d ? sizeof(float) * N : 0));
    };
    hipError_t e = hipMemcpy(rho_d, rho, sizeof(float) * N, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(u_d, u, sizeof(float) * 3 * N, hipMemcpyHostToDevice);
    if (e == hipSuccess && t_d) e = hipMemcpy(t_d, temperature, sizeof(float) * N, hipMemcpyHostToDevice);
    if (e != hipSuccess) { cleanup(); return fail(s, BF_ERR_DEVICE, "[LBM Error] upload failed: %s", hipGetErrorString(e)); }

    InitArgs ia{};
    for (int i = 0; i < s->q; ++i) ia.slot[i] = static_cast<char *>(s->f) + s->slot_off[i];
    ia.slab_log = s->slab_log; ia.slab_skip = s->slab_skip;
    ia.g = s->g; ia.stride = s->stride; ia.n = N;
    ia.rho = rho_d; ia.u = u_d; ia.temperature = t_d; ia.ids = s->ids;
    ia.objects = s->objects_d; ia.rho_eff = s->rho_eff_d;
    ia.psi0 = s->psi[0]; ia.psi1 = s->psi[1];
    ia.cs2_t = s->cs2_t;
    e = dispatch_init(s, ia);
    if (e != hipSuccess) { cleanup(); return fail(s, BF_ERR_DEVICE, "[LBM Error] init kernel failed: %s", hipGetErrorString(e)); }

    // save the macroscopic fields once without updating (LBM.cpp:502)
    s->cur = 0;
    s->ticks = 0;
    s->tick_base = 0;
    StepArgs a = make_args(s, false);
    a.vd = s->vd; a.ft = s->ft;
    e = dispatch_step(s, a, false, nullptr);
    if (e != hipSuccess) { cleanup(); return fail(s, BF_ERR_DEVICE, "[LBM Error] init pass failed: %s", hipGetErrorString(e)); }
    if (s->cfg.multiphase) s->cur ^= 1;
    e = hipStreamSynchronize(s->stream);
    cleanup();
    if (e != hipSuccess) return fail(s, BF_ERR_DEVICE, "[LBM Error] init failed: %s", hipGetErrorString(e));
    s->initialized = 1;
    return BF_OK;
}

bf_status bf_set_param(bf_solver *s, int32_t key, const float *v, int32_t n) {
    if (!s || !v || n < 1) return s ? fail(s, BF_ERR_INVALID, "[LBM Error] bf_set_param: bad value") : BF_ERR_INVALID;
    switch (key) {
    case BF_PARAM_RELAXATION_TIME:
        if (!(v[0] > 0.5f)) return fail(s, BF_ERR_INVALID, "[LBM Error] relaxation time must exceed 0.5");
        s->tau = v[0]; return BF_OK;
    case BF_PARAM_CONSTANT_FORCE:
        if (n < 3) return fail(s, BF_ERR_INVALID, "[LBM Error] constant force needs 3 values");
        s->fconst[0] = v[0]; s->fconst[1] = v[1]; s->fconst[2] = v[2];
        if (s->forcing != 2) s->forcing = (v[0] != 0.0f || v[1] != 0.0f || v[2] != 0.0f) ? 1 : 0;
        return BF_OK;
    case BF_PARAM_INTERACTION_STRENGTH: s->G = v[0]; return BF_OK;
    case BF_PARAM_THERMAL_EXPANSION: s->beta = v[0]; return BF_OK;
    case BF_PARAM_THERMAL_RELAXATION_TIME:
        if (!(v[0] > 0.5f)) return fail(s, BF_ERR_INVALID, "[LBM Error] thermal relaxation time must exceed 0.5");
        s->tau_t = v[0]; return BF_OK;
    case BF_PARAM_REFERENCE_TEMPERATURE: s->t_ref = v[0]; return BF_OK;
    case BF_PARAM_THERMAL_CS2:
        if (!(v[0] > 0.0f)) return fail(s, BF_ERR_INVALID, "[LBM Error] thermal c_s^2 must be positive");
        s->cs2_t = v[0]; return BF_OK;
    case BF_PARAM_VELOCITY_LIMIT: s->ulim = v[0]; return BF_OK;
    case BF_PARAM_VELOCITY_LIMIT_EXTREME: s->uext = v[0]; return BF_OK;
    case BF_PARAM_SHAN_CHEN_SYNCHRONOUS: s->sc_sync = v[0] != 0.0f; return BF_OK;
    case BF_PARAM_TICK_PAIR_PLANES:
        if (!(v[0] >= -1.0f) || v[0] != std::floor(v[0]) || v[0] > 65535.0f)
            return fail(s, BF_ERR_INVALID, "[LBM Error] bf_set_param: tick-pair chunk must be -1, 0 or a plane count");
        s->pair_planes = (int)v[0];
        return BF_OK;
    case BF_PARAM_TICK_SWEEP_TICKS:
        if (!(v[0] == 0.0f || v[0] == 2.0f || v[0] == 4.0f))
            return fail(s, BF_ERR_INVALID, "[LBM Error] bf_set_param: ticks per chunked sweep must be 0, 2 or 4");
        s->sweep_ticks = (int)v[0];
        return BF_OK;
    }
    return fail(s, BF_ERR_INVALID, "[LBM Error] bf_set_param: unknown key %d", key);
}

bf_status bf_get_param(const bf_solver *s, int32_t key, float *v, int32_t n) {
    if (!s || !v || n < 1) return BF_ERR_INVALID;
    switch (key) {
    case BF_PARAM_RELAXATION_TIME: v[0] = s->tau; return BF_OK;
    case BF_PARAM_CONSTANT_FORCE:
        if (n < 3) return BF_ERR_INVALID;
        v[0] = s->fconst[0]; v[1] = s->fconst[1]; v[2] = s->fconst[2]; return BF_OK;
    case BF_PARAM_INTERACTION_STRENGTH: v[0] = s->G; return BF_OK;
    case BF_PARAM_THERMAL_EXPANSION: v[0] = s->beta; return BF_OK;
    case BF_PARAM_THERMAL_RELAXATION_TIME: v[0] = s->tau_t; return BF_OK;
    case BF_PARAM_REFERENCE_TEMPERATURE: v[0] = s->t_ref; return BF_OK;
    case BF_PARAM_THERMAL_CS2: v[0] = s->cs2_t; return BF_OK;
    case BF_PARAM_VELOCITY_LIMIT: v[0] = s->ulim; return BF_OK;
    case BF_PARAM_VELOCITY_LIMIT_EXTREME: v[0] = s->uext; return BF_OK;
    case BF_PARAM_SHAN_CHEN_SYNCHRONOUS: v[0] = (float)s->sc_sync; return BF_OK;
    case BF_PARAM_TICK_PAIR_PLANES: v[0] = (float)s->pair_planes; return BF_OK;
    case BF_PARAM_TICK_SWEEP_TICKS: v[0] = (float)s->sweep_ticks; return BF_OK;
    }
    return BF_ERR_INVALID;
}

namespace {

// Chunked tick pairs.  An even tick followed by an odd tick can run chunk by
// chunk of z-planes: the odd tick of a plane needs the even tick of that plane
// and of its two neighbours (the even tick writes into the planes next to its
// own), so the odd tick of the planes up to two before chunk j can share a
// launch with chunk j's even tick (run_pair).  The lines the even tick wrote
// are then still in the
// 256 MB last-level cache when the odd tick reads them.  Measured on the
// MI355X (D3Q19, device time per tick, interleaved on one solver): with the
// odd planes one plane behind the even chunk, the 256^3 periodic box in FP32
// 378-381 us per tick with chunks of 24 planes (119 MB; 20 planes 381.5, 28
// planes 376.5-379.8, whole-grid ticks 407-413), the 384^3 periodic box
// 1294-1297 us (10-12 planes; 16 planes 1415), the 384^3 porous duct
// 1472-1474 us in FP32 (10 planes; 12 planes 1482, 14 planes 1546; whole-grid
// ticks 1733-1742) and 862-871 us in FP16 (20-24 planes; 28 planes 877-880,
// whole-grid ticks 1036-1041).  With the odd chunk two chunks behind the
// even one the best chunks held 80-90 MB (256^3: 384 us per tick).  The
// periodic 16-bit boxes gain once their even tick issues few enough
// instructions: 256^3 FP16 193.6-196.6 us per tick with 48-plane chunks
// (120 MB; 32 planes 197.6, 64 planes 194.8, whole-grid ticks 199.4-200.0),
// BF16 195.6-196.2 against 199.8; 384^3 FP16 679.1-679.8 us with 21 planes
// (28 planes 675.1-675.3, whole-grid ticks 703.9-706.2).  The automatic mode
// pairs the ticks of every lattice above kPairMinBytes, in chunks of about
// kPairChunkBytes.
//
// Sweeps of two pairs (BF_PARAM_TICK_SWEEP_TICKS 4) hold four ticks' lines in
// the cache at once and want smaller chunks (D3Q19, device time per tick,
// interleaved on one solver, three rounds each): 256^3 FP32 370.1-370.2 us
// with 16-plane chunks (80 MB; 12 planes 377, 20 planes 378, 24 planes 399)
// against 371.7-372.0 for pairs; 256^3 FP16 194.0-194.2 us with 32 planes
// (80 MB; 40-64 planes 195.3-196.5) against 195.8-199.2; the 384^3 porous
// duct no faster (FP16 829.6-850.2 us with 14 planes against 834.1-842.2,
// FP32 1472.9 with 5 planes against 1470.1-1472.8).  The automatic mode runs
// two-pair sweeps in chunks of about kQuadChunkBytes for the 16-bit lattices
// without walls or a thermal lattice, where the gain stands clear of the
// noise; pairs otherwise.
constexpr double kPairChunkBytes = 120e6;
constexpr double kQuadChunkBytes = 80e6;
constexpr double kPairMinBytes = 384e6;

// ticks per chunked sweep (2 or 4)
int sweep_depth(const bf_solver *s) {
    if (s->sweep_ticks != 0) return s->sweep_ticks;
    return (s->pair_planes == 0 && s->storage != ST_FP32 && !s->has_ids && !s->thermal && !s->cfg.multiphase) ? 4 : 2;
}

// planes per chunk, or 0 for whole-grid ticks
int pair_chunk(const bf_solver *s) {
    // The synchronous Shan-Chen mode runs a density pre-pass before each
    // tick.  The lagged mode pairs only on request: its even tick is bound by
    // instruction issue, and pairs measured slower (D3Q27 + D3Q7 thermal
    // walls, 192^3 FP32, device time per tick, three rounds: whole-grid ticks
    // 369.8-371.1 us; pairs of 12, 16, 24 and 32 planes 425.6-426.1, 409.3-409.5,
    // 391.0-391.1 and 384.9-385.0; two-pair sweeps of 8, 12 and 16 planes
    // 401.7-401.9, 391.9-392.2 and 381.9-382.0).
    if ((s->cfg.multiphase && (s->sc_sync || s->pair_planes == 0)) || s->pair_planes < 0) return 0;
    int zc = s->pair_planes;
    if (zc == 0) {
        const double plane = (double)s->X * s->Y * ((double)s->q * s->elem + (s->thermal ? 4.0 * s->qt : 0.0));
        if (plane * s->Z < kPairMinBytes) return 0;
        const double chunk = sweep_depth(s) == 4 ? kQuadChunkBytes : kPairChunkBytes;
        zc = std::max(1, (int)std::lround(chunk / plane));
    }
    // the schedule needs at least four chunks (chunk 0's odd tick waits for
    // the last chunk's even tick, which writes into plane 0 on a periodic z axis)
    return (s->Z + zc - 1) / zc >= 4 ? zc : 0;
}

hipError_t dispatch_pair(const bf_solver *s, const StepArgs &ae, const StepArgs &ao, const PairStages &ps,
                         double *partial) {
    hipStream_t stream = s->stream;
    const bool walls = s->has_ids != 0, mp = s->cfg.multiphase != 0, th = s->thermal != 0;
    switch (s->vs) {
    case D2Q9: return launch_pair<D2Q9>(ae, ao, ps, s->storage, walls, mp, th, s->shape, partial, stream);
    case D3Q15: return launch_pair<D3Q15>(ae, ao, ps, s->storage, walls, mp, th, s->shape, partial, stream);
    case D3Q19: return launch_pair<D3Q19>(ae, ao, ps, s->storage, walls, mp, th, s->shape, partial, stream);
    case D3Q27: return launch_pair<D3Q27>(ae, ao, ps, s->storage, walls, mp, th, s->shape, partial, stream);
    }
    return hipErrorInvalidValue;
}

// An even tick and the odd tick after it (depth 2), or two such pairs
// (depth 4), in chunks of zc planes (m >= 4 chunks), on the handle's stream.
// Depth 2, m + 1 launches:
//   E0 | E1 + O[1, lo(1) - 1) | ... | E(m-1) + O[.., lo(m-1) - 1) | O[lo(m-1) - 1, Z) O0
// Within a launch the odd planes end two planes before the even chunk: even
// ticks of different cells, or odd ticks, never touch one another's slots.
// The odd planes read what the even chunk of the previous launch wrote, which
// the last-level cache still holds.  Depth 4 adds the second pair's even tick
// E' and odd tick O' to the same launches (m + 3 launches), each two planes
// behind the tick before it, so each tick of a plane again reads what the
// previous launch wrote.  `flags` apply to the last odd tick; depth 4 is run
// without flags (run_ticks ends a call that saves or sums with a pair).
bf_status run_pair(bf_solver *s, int zc, int depth, uint32_t flags) {
    const int Z = s->Z, m = (Z + zc - 1) / zc;
    StepArgs ae = make_args(s, true), ao = make_args(s, true);
    if (s->cfg.multiphase) {
        // lagged psi: the odd tick reads what the even tick writes, and writes
        // the buffer the even tick read (the next even tick reads it)
        ao.psi_prev = s->psi[s->cur];
        ao.psi_cur = s->psi[s->cur ^ 1];
    }
    if (flags & BF_STEP_SAVE_MACRO) { ao.vd = s->vd; ao.ft = s->ft; }
    double *partial = (flags & BF_STEP_CONSERVATION) ? s->partial : nullptr;
    auto lo = [&](int j) { return j * zc; };
    auto hi = [&](int j) { return std::min(Z, (j + 1) * zc); };
    // ranges in order E, O, E', O': first plane and plane count (modulo Z)
    auto launch = [&](int e0, int ne, int o0, int no, int e1, int ne1, int o1, int no1) -> bf_status {
        const PairStages ps{{e0, o0, e1, o1}, {ne, no, ne1, no1}};
        const hipError_t e = dispatch_pair(s, ae, ao, ps, partial);
        if (e != hipSuccess) return fail(s, BF_ERR_DEVICE, "[LBM Error] tick-pair launch failed: %s", hipGetErrorString(e));
        return BF_OK;
    };
    bf_status st = BF_OK;
    if (depth == 2) {
        // The odd tick one plane behind the even tick: launch j runs the even
        // tick on chunk j and the odd tick on every plane p in 1 .. lo(j) - 2
        // not yet done (plane p needs the even ticks of p - 1, p and p + 1, all
        // in earlier launches; this launch's even chunk reads and writes from
        // plane lo(j) - 1 on).  Plane 0 and the planes from lo(m - 1) - 1 come
        // last: on a periodic z axis the last chunk's even tick writes into
        // plane 0.
        st = launch(0, hi(0), 0, 0, 0, 0, 0, 0);
        int zo = 1;
        for (int j = 1; j < m && st == BF_OK; ++j) {
            const int nzo = std::max(0, lo(j) - 1 - zo);
            st = launch(lo(j), hi(j) - lo(j), zo, nzo, 0, 0, 0, 0);
            zo += nzo;
        }
        if (st == BF_OK) st = launch(0, 0, zo, Z - zo + 1, 0, 0, 0, 0);
    } else {
        // Each tick after the first covers, per launch, the planes that the
        // tick before it finished in earlier launches, up to two planes short
        // of that tick's first plane in this launch (a plane's tick needs the
        // tick before it on the plane and both neighbours; an even tick reads
        // and writes the neighbour planes).  O starts at plane 1, E' at 2, O'
        // at 3: plane 0 waits for the last chunk's even tick, which writes into
        // it on a periodic z axis.  The last three launches finish O, E' and O'
        // over the planes from their fronts to Z - 1 and then 0, 0 .. 1 and
        // 0 .. 2.
        int zo = 1, ze1 = 2, zo1 = 3;
        for (int j = 0; j < m && st == BF_OK; ++j) {
            const int no = j == 0 ? 0 : std::max(0, lo(j) - 1 - zo);
            const int ne1 = std::max(0, zo - 1 - ze1);
            const int no1 = std::max(0, ze1 - 1 - zo1);
            st = launch(lo(j), hi(j) - lo(j), zo, no, ze1, ne1, zo1, no1);
            zo += no; ze1 += ne1; zo1 += no1;
        }
        if (st == BF_OK) {
            const int ne1 = std::max(0, zo - 1 - ze1), no1 = std::max(0, ze1 - 1 - zo1);
            st = launch(0, 0, zo, Z - zo + 1, ze1, ne1, zo1, no1);
            ze1 += ne1; zo1 += no1;
        }
        if (st == BF_OK) {
            const int no1 = std::max(0, ze1 - 1 - zo1);
            st = launch(0, 0, 0, 0, ze1, Z - ze1 + 2, zo1, no1);
            zo1 += no1;
        }
        if (st == BF_OK) st = launch(0, 0, 0, 0, 0, 0, zo1, Z - zo1 + 3);
    }
    if (st != BF_OK) return st;
    if (partial) {
        hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(256), 0, s->stream, s->partial, s->partial_blocks, s->sums_d);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return fail(s, BF_ERR_DEVICE, "[LBM Error] reduction launch failed: %s", hipGetErrorString(e));
    }
    s->ticks += depth;
    return BF_OK;
}

}  // namespace

static bf_status run_ticks(bf_solver *s, int64_t n, uint32_t flags) {
    const int zc = pair_chunk(s), depth = sweep_depth(s);
    for (int64_t k = 0; k < n; ++k) {
        const bool last = k == n - 1;
        const bool odd = ((s->ticks - s->tick_base) & 1) != 0;
        if (zc > 0 && !odd && depth == 4 && (k + 4 < n || (k + 4 == n && flags == 0u))) {
            // two pairs of ticks, chunk by chunk (a call that saves or sums
            // ends with a single pair, which takes the flags)
            const bf_status st = run_pair(s, zc, 4, 0u);
            if (st != BF_OK) return st;
            k += 3;
            continue;
        }
        if (zc > 0 && !odd && k + 1 < n) {
            // an even tick and the odd tick after it, chunk by chunk
            const bf_status st = run_pair(s, zc, 2, k + 2 == n ? flags : 0u);
            if (st != BF_OK) return st;
            ++k;
            continue;
        }
        StepArgs a = make_args(s, true);
        if (last && (flags & BF_STEP_SAVE_MACRO)) { a.vd = s->vd; a.ft = s->ft; }
        double *partial = (last && (flags & BF_STEP_CONSERVATION)) ? s->partial : nullptr;
        if (s->cfg.multiphase && s->sc_sync) {
            // density pre-pass: the tick's reads without writes of populations;
            // psi of this tick's density goes to psi[cur], which the tick then
            // reads as its neighbours' psi (and it writes its own to the other)
            StepArgs p = make_args(s, false);
            p.psi_cur = s->psi[s->cur];
            const hipError_t e = dispatch_density(s, p, odd);
            if (e != hipSuccess) return fail(s, BF_ERR_DEVICE, "[LBM Error] density pre-pass launch failed: %s", hipGetErrorString(e));
            a.psi_prev = s->psi[s->cur];
            a.psi_cur = s->psi[s->cur ^ 1];
        }
        hipError_t e = dispatch_step(s, a, odd, partial);
        if (e != hipSuccess) return fail(s, BF_ERR_DEVICE, "[LBM Error] tick kernel launch failed: %s", hipGetErrorString(e));
        if (partial) {
            hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(256), 0, s->stream, s->partial, s->partial_blocks, s->sums_d);
            e = hipGetLastError();
            if (e != hipSuccess) return fail(s, BF_ERR_DEVICE, "[LBM Error] reduction launch failed: %s", hipGetErrorString(e));
        }
        if (s->cfg.multiphase) s->cur ^= 1;
        s->ticks++;
    }
    return BF_OK;
}

bf_status bf_step(bf_solver *s, int64_t n, uint32_t flags) {
    if (!s) return BF_ERR_INVALID;
    if (!s->initialized) return fail(s, BF_ERR_STATE, "[LBM Error] bf_step called before bf_init_fields");
    if (n < 0) return fail(s, BF_ERR_INVALID, "[LBM Error] bf_step: negative tick count");
    if (hipSetDevice(s->cfg.device) != hipSuccess) return fail(s, BF_ERR_DEVICE, "[LBM Error] hipSetDevice failed");
    return run_ticks(s, n, flags);
}

bf_status bf_step_timed(bf_solver *s, int64_t n, uint32_t flags, double *elapsed_ms) {
    if (!s || !elapsed_ms) return BF_ERR_INVALID;
    if (!s->initialized) return fail(s, BF_ERR_STATE, "[LBM Error] bf_step_timed called before bf_init_fields");
    if (n < 0) return fail(s, BF_ERR_INVALID, "[LBM Error] bf_step_timed: negative tick count");
    BF_HIP(s, hipSetDevice(s->cfg.device));
    BF_HIP(s, hipEventRecord(s->ev0, s->stream));
    bf_status st = run_ticks(s, n, flags);
    if (st != BF_OK) return st;
    BF_HIP(s, hipEventRecord(s->ev1, s->stream));
    // poll instead of a blocking wait: the call returns within microseconds of
    // the last tick rather than after the interrupt-driven wake-up
    hipError_t hipEventQuery_ev1;
    while ((hipEventQuery_ev1 = hipEventQuery(s->ev1)) == hipErrorNotReady) {}
    BF_HIP(s, hipEventQuery_ev1);
    float ms = 0.0f;
    BF_HIP(s, hipEventElapsedTime(&ms, s->ev0, s->ev1));
    *elapsed_ms = ms;
    return BF_OK;
}

int64_t bf_ticks(const bf_solver *s) { return s ? s->ticks : -1; }

bf_status bf_synchronize(bf_solver *s) {
    if (!s) return BF_ERR_INVALID;
    BF_HIP(s, hipStreamSynchronize(s->stream));
    return BF_OK;
}

size_t bf_field_bytes(const bf_solver *s, int32_t field) {
    if (!s) return 0;
    switch (field) {
    case BF_FIELD_VELOCITY_DENSITY: return sizeof(float) * 4 * (size_t)s->N;
    case BF_FIELD_BOUNDARY: return (size_t)s->N;
    case BF_FIELD_FORCE_TEMPERATURE: return sizeof(float) * 4 * (size_t)s->N;
    case BF_FIELD_POPULATIONS_RAW: return sizeof(float) * (size_t)s->q * s->N;
    case BF_FIELD_THERMAL_RAW: return s->thermal ? sizeof(float) * (size_t)s->qt * s->N : 0;
    case BF_FIELD_PSI: return s->cfg.multiphase ? sizeof(float) * (size_t)s->N : 0;
    case BF_FIELD_POPULATIONS: return sizeof(float) * (size_t)s->q * s->N;
    }
    return 0;
}

namespace {

// The populations each cell's next collision uses, from the slots (unshifted
// FP32, population-major) by the AA read rules of that tick, as the tick
// kernel applies them: odd tick slot opp(i) at the cell; even tick slot i at
// x - c_i, 0.01 outside the domain (collide.comp:297-298), and with halfway
// bounce-back the own slot opp(i) plus 6 w_i (c_i . u_w) where x - c_i is
// solid.  Solid cells keep their slots as stored.
void logical_populations(const bf_solver *s, const std::vector<float> &raw, const std::vector<uint8_t> &ids,
                         bool odd, float *out) {
    float tab[4 * kMaxQ];
    bf_velocity_table(s->vs, 0, tab, 4 * kMaxQ);
    const int64_t N = s->N;
    const int X = s->X, Y = s->Y, Z = s->Z;
    const bool halfway = s->cfg.boundary_mode == BF_BB_HALFWAY;
    const int per[3] = {s->cfg.periodic[0] != 0, s->cfg.periodic[1] != 0, s->cfg.periodic[2] != 0};
    for (int64_t c = 0; c < N; ++c) {
        const int x = (int)(c % X), y = (int)((c / X) % Y), z = (int)(c / ((int64_t)X * Y));
        for (int i = 0; i < s->q; ++i) {
            const int o = opposite(i);
            float v;
            if (ids[c] != 0) {
                v = raw[(size_t)i * N + c];
            } else if (odd) {
                v = raw[(size_t)o * N + c];
            } else if (i == 0) {
                v = raw[c];
            } else {
                int n3[3] = {x - (int)tab[4 * i], y - (int)tab[4 * i + 1], z - (int)tab[4 * i + 2]};
                const int ext[3] = {X, Y, Z};
                bool inside = true;
                for (int d = 0; d < 3; ++d) {
                    if (n3[d] >= 0 && n3[d] < ext[d]) continue;
                    if (per[d]) n3[d] = (n3[d] + ext[d]) % ext[d];
                    else inside = false;
                }
                if (!inside) {
                    v = 0.01f;
                } else {
                    const int64_t n = n3[0] + (int64_t)X * (n3[1] + (int64_t)Y * n3[2]);
                    if (halfway && ids[n] != 0) {
                        const ObjectEntry &ob = s->objects_h[ids[n]];
                        const float cu = tab[4 * i] * ob.ux + tab[4 * i + 1] * ob.uy + tab[4 * i + 2] * ob.uz;
                        v = raw[(size_t)o * N + c] + 6.0f * tab[4 * i + 3] * cu;
                    } else {
                        v = raw[(size_t)i * N + n];
                    }
                }
            }
            out[(size_t)i * N + c] = v;
        }
    }
}

}  // namespace

// Device view of the saved density (the w channel of the velocity_density
// field) for the iso-surface pass (bf_mc.hip).
bf_status bf_solver_density_field(bf_solver *s, const float **v, int *step, int32_t res[3], hipStream_t *stream) {
    if (!s || !s->initialized || !s->vd) return BF_ERR_STATE;
    if (hipSetDevice(s->cfg.device) != hipSuccess) return BF_ERR_DEVICE;
    *v = reinterpret_cast<const float *>(s->vd) + 3;
    *step = 4;
    res[0] = s->X; res[1] = s->Y; res[2] = s->Z;
    *stream = s->stream;
    return BF_OK;
}

bf_status bf_read_field(bf_solver *s, int32_t field, void *dst, size_t bytes) {
    if (!s || !dst) return BF_ERR_INVALID;
    const size_t need = bf_field_bytes(s, field);
    if (need == 0) return fail(s, BF_ERR_INVALID, "[LBM Error] bf_read_field: field %d does not exist for this solver", field);
    if (bytes != need) return fail(s, BF_ERR_INVALID, "[LBM Error] bf_read_field: %zu bytes given, %zu needed", bytes, need);
    if (!s->initialized) return fail(s, BF_ERR_STATE, "[LBM Error] bf_read_field called before bf_init_fields");
    BF_HIP(s, hipSetDevice(s->cfg.device));
    BF_HIP(s, hipStreamSynchronize(s->stream));
    const size_t N = (size_t)s->N;
    switch (field) {
    case BF_FIELD_VELOCITY_DENSITY:
        BF_HIP(s, hipMemcpy(dst, s->vd, need, hipMemcpyDeviceToHost));
        return BF_OK;
    case BF_FIELD_FORCE_TEMPERATURE:
        BF_HIP(s, hipMemcpy(dst, s->ft, need, hipMemcpyDeviceToHost));
        return BF_OK;
    case BF_FIELD_BOUNDARY:
        if (s->ids) BF_HIP(s, hipMemcpy(dst, s->ids, N, hipMemcpyDeviceToHost));
        else std::memset(dst, 0, N);
        return BF_OK;
    case BF_FIELD_POPULATIONS_RAW: {
        // the lattice holds h = f - w: add the weight back on the host
        float *out = static_cast<float *>(dst);
        if (s->storage == ST_FP32) {
            for (int i = 0; i < s->q; ++i) {
                BF_HIP(s, copy_slot(s, i, out + i * N, true));
                const float w = weight_of(s->vs, i);
                for (size_t c = 0; c < N; ++c) out[i * N + c] += w;
            }
            return BF_OK;
        }
        std::vector<uint16_t> tmp(N);
        for (int i = 0; i < s->q; ++i) {
            BF_HIP(s, copy_slot(s, i, tmp.data(), true));
            const float w = weight_of(s->vs, i);
            for (size_t c = 0; c < N; ++c) {
                float v;
                if (s->storage == ST_FP16) v = half_to_float(tmp[c]);
                else { uint32_t b = (uint32_t)tmp[c] << 16; std::memcpy(&v, &b, 4); }
                out[i * N + c] = v + w;
            }
        }
        return BF_OK;
    }
    case BF_FIELD_THERMAL_RAW: {
        float *out = static_cast<float *>(dst);
        for (int k = 0; k < s->qt; ++k)
            BF_HIP(s, hipMemcpy(out + k * N, s->g + (size_t)k * s->stride, sizeof(float) * N, hipMemcpyDeviceToHost));
        return BF_OK;
    }
    case BF_FIELD_PSI:
        BF_HIP(s, hipMemcpy(dst, s->psi[s->cur ^ 1], need, hipMemcpyDeviceToHost));
        return BF_OK;
    case BF_FIELD_POPULATIONS: {
        std::vector<float> raw((size_t)s->q * N);
        const bf_status st = bf_read_field(s, BF_FIELD_POPULATIONS_RAW, raw.data(), raw.size() * sizeof(float));
        if (st != BF_OK) return st;
        std::vector<uint8_t> ids(N, 0);
        if (s->ids) BF_HIP(s, hipMemcpy(ids.data(), s->ids, N, hipMemcpyDeviceToHost));
        logical_populations(s, raw, ids, ((s->ticks - s->tick_base) & 1) != 0, static_cast<float *>(dst));
        return BF_OK;
    }
    }
    return BF_ERR_INVALID;
}

bf_status bf_conservation(bf_solver *s, double out[5]) {
    if (!s || !out) return BF_ERR_INVALID;
    if (!s->initialized) return fail(s, BF_ERR_STATE, "[LBM Error] bf_conservation called before bf_init_fields");
    BF_HIP(s, hipSetDevice(s->cfg.device));
    StepArgs a = make_args(s, false);
    a.write_psi = 0;
    hipError_t e = dispatch_step(s, a, ((s->ticks - s->tick_base) & 1) != 0, s->partial);
    if (e != hipSuccess) return fail(s, BF_ERR_DEVICE, "[LBM Error] conservation kernel failed: %s", hipGetErrorString(e));
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(256), 0, s->stream, s->partial, s->partial_blocks, s->sums_d);
    BF_HIP(s, hipGetLastError());
    double tmp[5];
    BF_HIP(s, hipMemcpyAsync(tmp, s->sums_d, sizeof tmp, hipMemcpyDeviceToHost, s->stream));
    BF_HIP(s, hipStreamSynchronize(s->stream));
    std::memcpy(out, tmp, sizeof tmp);
    if (!s->thermal) out[4] = 0.0;
    return BF_OK;
}

bf_status bf_last_step_sums(bf_solver *s, double out[5]) {
    if (!s || !out) return BF_ERR_INVALID;
    double tmp[5];
    BF_HIP(s, hipMemcpyAsync(tmp, s->sums_d, sizeof tmp, hipMemcpyDeviceToHost, s->stream));
    BF_HIP(s, hipStreamSynchronize(s->stream));
    std::memcpy(out, tmp, sizeof tmp);
    if (!s->thermal) out[4] = 0.0;
    return BF_OK;
}

bf_status bf_get_info(const bf_solver *s, bf_info *o) {
    if (!s || !o) return BF_ERR_INVALID;
    o->resolution[0] = s->X; o->resolution[1] = s->Y; o->resolution[2] = s->Z;
    o->q = s->q;
    o->q_thermal = s->thermal ? s->qt : 0;
    o->storage = s->storage;
    o->forcing = s->forcing;
    o->thermal = s->thermal;
    o->multiphase = s->cfg.multiphase;
    o->has_boundaries = s->has_ids;
    o->cells = s->N;
    o->stride = s->pair_step;
    o->device_bytes = s->device_bytes;
    return BF_OK;
}

// bits = 2^ceil(log2(ceil(log2(object_count)))), 0 when only the fluid slot exists (LBM.cpp:566-573)
int32_t bf_bits_per_boundary(int32_t object_count) {
    if (object_count <= 1) return 0;
    return (int32_t)std::exp2(std::ceil(std::log2f(std::ceil(std::log2f((float)object_count)))));
}

bf_status bf_pack_boundaries(const uint8_t *ids, int64_t n, int32_t bits, uint32_t *words) {
    if (!ids || !words || n < 0 || !(bits == 1 || bits == 2 || bits == 4 || bits == 8)) return BF_ERR_INVALID;
    const int64_t nw = (n * bits + 31) / 32;
    std::memset(words, 0, sizeof(uint32_t) * nw);
    const uint32_t mask = (1u << bits) - 1u;
    for (int64_t c = 0; c < n; ++c) {
        if (ids[c] > mask) return BF_ERR_INVALID;
        const int64_t b = c * bits;
        words[b / 32] |= (uint32_t)ids[c] << (b % 32);
    }
    return BF_OK;
}

bf_status bf_unpack_boundaries(const uint32_t *words, int64_t n, int32_t bits, uint8_t *ids) {
    if (!ids || !words || n < 0 || !(bits == 1 || bits == 2 || bits == 4 || bits == 8)) return BF_ERR_INVALID;
    const uint32_t mask = (1u << bits) - 1u;
    for (int64_t c = 0; c < n; ++c) {
        const int64_t b = c * bits;
        ids[c] = (uint8_t)((words[b / 32] >> (b % 32)) & mask);
    }
    return BF_OK;
}

}  // extern "C"

// ---- checkpoint files (LBM.h:123-126 declares them; the reference has no body) ----
//
// One file per save: <directory>/tick_<ticks>.bfmeso or .bfmacro.  Little
// endian.  A fixed header (kind, grid, configuration, parameters, object table)
// is followed by tagged sections (uint32 tag, uint64 byte count, bytes).
// Mesoscopic files hold the raw state (populations as stored, in slot order and
// the AA parity of the tick count; thermal populations; both pseudopotential buffers;
// the force field; boundary ids; the saved macroscopic fields), so a loaded
// solver continues bit for bit.  Macroscopic files hold the saved fields
// (u, rho), (F, T), the ids and the external force (the constant force in the
// header, a per-cell force field as a section); loading one re-initialises the
// populations at equilibrium like initialize_fields, from the saved (u, rho),
// the saved T and the external force.  The F channel of the saved fields is
// not a force to restore: it holds the total force of the save tick, with the
// buoyancy and the Shan-Chen force added (collide.comp:707-720).

namespace {

constexpr char kMagic[8] = {'B', 'F', 'L', 'B', 'M', 'S', 'T', '1'};
// version 2: the population section holds h = f - w for every storage type;
// version 3: macroscopic files carry the per-cell external force field
constexpr uint32_t kVersion = 3;
enum : uint32_t { kKindMacro = 0, kKindMeso = 1 };
enum : uint32_t {
    kTagPopulations = 1, kTagThermal = 2, kTagPsi = 3, kTagForce = 4, kTagIds = 5, kTagVelocityDensity = 6,
    kTagForceTemperature = 7
};

struct FileHeader {
    char magic[8];
    uint32_t version, kind;
    int32_t X, Y, Z, velocity_set, storage, multiphase, boundary_mode, periodic[3];
    int32_t thermal, forcing, has_ids, object_count, cur, sc_sync;
    int64_t ticks, tick_base;
    float tau, G, beta, tau_t, t_ref, cs2_t, ulim, uext, fconst[3];
    ObjectEntry objects[256];
    float rho_eff[256];
};

std::string state_path(const char *dir, int64_t tick, uint32_t kind) {
    return std::string(dir) + "/tick_" + std::to_string((long long)tick) + (kind == kKindMeso ? ".bfmeso" : ".bfmacro");
}

struct File {
    FILE *f = nullptr;
    ~File() {
        if (f) std::fclose(f);
    }
};

bool put(FILE *f, uint32_t tag, const void *p, uint64_t n) {
    return std::fwrite(&tag, 4, 1, f) == 1 && std::fwrite(&n, 8, 1, f) == 1 && (n == 0 || std::fwrite(p, 1, n, f) == n);
}

// reads the next section; it must carry `tag` and exactly n bytes
bool get(FILE *f, uint32_t tag, void *p, uint64_t n) {
    uint32_t t = 0;
    uint64_t m = 0;
    return std::fread(&t, 4, 1, f) == 1 && std::fread(&m, 8, 1, f) == 1 && t == tag && m == n &&
           (n == 0 || std::fread(p, 1, n, f) == n);
}

FileHeader header_of(const bf_solver *s, uint32_t kind) {
    FileHeader h{};
    std::memcpy(h.magic, kMagic, 8);
    h.version = kVersion;
    h.kind = kind;
    h.X = s->X; h.Y = s->Y; h.Z = s->Z;
    h.velocity_set = s->vs; h.storage = s->storage; h.multiphase = s->cfg.multiphase;
    h.boundary_mode = s->cfg.boundary_mode;
    for (int k = 0; k < 3; ++k) h.periodic[k] = s->cfg.periodic[k];
    h.thermal = s->thermal; h.forcing = s->forcing; h.has_ids = s->has_ids; h.object_count = s->object_count;
    h.cur = s->cur; h.sc_sync = s->sc_sync;
    h.ticks = s->ticks;
    h.tick_base = s->tick_base;
    h.tau = s->tau; h.G = s->G; h.beta = s->beta; h.tau_t = s->tau_t; h.t_ref = s->t_ref; h.cs2_t = s->cs2_t;
    h.ulim = s->ulim; h.uext = s->uext;
    for (int k = 0; k < 3; ++k) h.fconst[k] = s->fconst[k];
    std::memcpy(h.objects, s->objects_h, sizeof h.objects);
    std::memcpy(h.rho_eff, s->rho_eff_h, sizeof h.rho_eff);
    return h;
}

void apply_params(bf_solver *s, const FileHeader &h) {
    s->tau = h.tau; s->G = h.G; s->beta = h.beta; s->tau_t = h.tau_t; s->t_ref = h.t_ref; s->cs2_t = h.cs2_t;
    s->ulim = h.ulim; s->uext = h.uext; s->sc_sync = h.sc_sync;
    for (int k = 0; k < 3; ++k) s->fconst[k] = h.fconst[k];
    std::memcpy(s->objects_h, h.objects, sizeof h.objects);
    std::memcpy(s->rho_eff_h, h.rho_eff, sizeof h.rho_eff);
    s->object_count = h.object_count;
}

template <class T>
bf_status download(bf_solver *s, std::vector<T> &out, const void *src, size_t bytes) {
    out.resize((bytes + sizeof(T) - 1) / sizeof(T));
    BF_HIP(s, hipMemcpy(out.data(), src, bytes, hipMemcpyDeviceToHost));
    return BF_OK;
}

bf_status write_state(bf_solver *s, const char *dir, uint32_t kind) {
    if (!s || !dir) return BF_ERR_INVALID;
    if (!s->initialized) return fail(s, BF_ERR_STATE, "[LBM Error] save called before initialize_fields");
    BF_HIP(s, hipSetDevice(s->cfg.device));
    BF_HIP(s, hipStreamSynchronize(s->stream));
    (void)mkdir(dir, 0755);
    const std::string path = state_path(dir, s->ticks, kind);
    File file;
    file.f = std::fopen(path.c_str(), "wb");
    if (!file.f) return fail(s, BF_ERR_INVALID, "[LBM Error] cannot write %s", path.c_str());
    const FileHeader h = header_of(s, kind);
    bool ok = std::fwrite(&h, sizeof h, 1, file.f) == 1;
    const size_t N = (size_t)s->N;
    std::vector<uint8_t> buf;
    bf_status st;
    auto section = [&](uint32_t tag, const void *dev, size_t bytes) -> bf_status {
        if ((st = download(s, buf, dev, bytes)) != BF_OK) return st;
        ok = ok && put(file.f, tag, buf.data(), bytes);
        return BF_OK;
    };
    if (kind == kKindMeso) {
        // populations slot by slot, N cells each (the layout's padding is dropped)
        std::vector<uint8_t> pops(s->q * N * s->elem);
        for (int i = 0; i < s->q; ++i) BF_HIP(s, copy_slot(s, i, pops.data() + i * N * s->elem, true));
        ok = ok && put(file.f, kTagPopulations, pops.data(), pops.size());
        if (s->thermal) {
            std::vector<uint8_t> th(s->qt * N * sizeof(float));
            for (int k = 0; k < s->qt; ++k)
                BF_HIP(s, hipMemcpy(th.data() + k * N * sizeof(float), s->g + (size_t)k * s->stride, N * sizeof(float),
                                    hipMemcpyDeviceToHost));
            ok = ok && put(file.f, kTagThermal, th.data(), th.size());
        }
        if (s->cfg.multiphase) {
            std::vector<uint8_t> ps(2 * N * sizeof(float));
            BF_HIP(s, hipMemcpy(ps.data(), s->psi[0], N * sizeof(float), hipMemcpyDeviceToHost));
            BF_HIP(s, hipMemcpy(ps.data() + N * sizeof(float), s->psi[1], N * sizeof(float), hipMemcpyDeviceToHost));
            ok = ok && put(file.f, kTagPsi, ps.data(), ps.size());
        }
    }
    if (s->forcing == 2 && (st = section(kTagForce, s->force_d, N * sizeof(float4))) != BF_OK) return st;
    if (s->has_ids && (st = section(kTagIds, s->ids, N)) != BF_OK) return st;
    if ((st = section(kTagVelocityDensity, s->vd, N * sizeof(float4))) != BF_OK) return st;
    if ((st = section(kTagForceTemperature, s->ft, N * sizeof(float4))) != BF_OK) return st;
    if (!ok || std::fflush(file.f) != 0) return fail(s, BF_ERR_INVALID, "[LBM Error] writing %s failed", path.c_str());
    return BF_OK;
}

bf_status open_state(bf_solver *s, const char *dir, int64_t tick, uint32_t kind, File &file, FileHeader &h) {
    const std::string path = state_path(dir, tick, kind);
    file.f = std::fopen(path.c_str(), "rb");
    if (!file.f) return fail(s, BF_ERR_INVALID, "[LBM Error] no checkpoint %s", path.c_str());
    if (std::fread(&h, sizeof h, 1, file.f) != 1 || std::memcmp(h.magic, kMagic, 8) != 0 || h.version != kVersion ||
        h.kind != kind)
        return fail(s, BF_ERR_INVALID, "[LBM Error] %s is not a checkpoint of this kind", path.c_str());
    if (h.X != s->X || h.Y != s->Y || h.Z != s->Z || h.velocity_set != s->vs || h.multiphase != s->cfg.multiphase ||
        h.boundary_mode != s->cfg.boundary_mode || h.periodic[0] != s->cfg.periodic[0] ||
        h.periodic[1] != s->cfg.periodic[1] || h.periodic[2] != s->cfg.periodic[2] ||
        (kind == kKindMeso && h.storage != s->storage))
        return fail(s, BF_ERR_INVALID, "[LBM Error] %s was saved by a solver of another configuration", path.c_str());
    return BF_OK;
}

}  // namespace

extern "C" {

bf_status bf_save_mesoscopic(bf_solver *s, const char *directory) { return write_state(s, directory, kKindMeso); }
bf_status bf_save_macroscopic(bf_solver *s, const char *directory) { return write_state(s, directory, kKindMacro); }

bf_status bf_load_macroscopic(bf_solver *s, const char *directory, int64_t tick) {
    if (!s || !directory) return BF_ERR_INVALID;
    File file;
    FileHeader h;
    bf_status st = open_state(s, directory, tick, kKindMacro, file, h);
    if (st != BF_OK) return st;
    const size_t N = (size_t)s->N;
    auto truncated = [&]() { return fail(s, BF_ERR_INVALID, "[LBM Error] truncated checkpoint"); };
    // the external force: the header's constant force, or the per-cell field
    std::vector<float> field;
    if (h.forcing == 2) {
        field.resize(4 * N);
        if (!get(file.f, kTagForce, field.data(), 16 * N)) return truncated();
    }
    std::vector<uint8_t> ids;
    if (h.has_ids) {
        ids.resize(N);
        if (!get(file.f, kTagIds, ids.data(), N)) return truncated();
    }
    std::vector<float> vd(4 * N), ft(4 * N);
    if (!get(file.f, kTagVelocityDensity, vd.data(), 16 * N) || !get(file.f, kTagForceTemperature, ft.data(), 16 * N))
        return truncated();
    std::vector<float> rho(N), u(3 * N), force(3 * N), T(N);
    for (size_t c = 0; c < N; ++c) {
        rho[c] = vd[4 * c + 3];
        for (int k = 0; k < 3; ++k) {
            u[3 * c + k] = vd[4 * c + k];
            force[3 * c + k] = h.forcing == 2 ? field[4 * c + k] : h.fconst[k];
        }
        T[c] = ft[4 * c + 3];
    }
    apply_params(s, h);
    st = bf_init_fields(s, rho.data(), u.data(), h.forcing ? force.data() : nullptr, h.thermal ? T.data() : nullptr,
                        h.has_ids ? ids.data() : nullptr);
    if (st != BF_OK) return st;
    // the populations restart at equilibrium in the layout of an even tick
    s->ticks = h.ticks;
    s->tick_base = h.ticks;
    return BF_OK;
}

bf_status bf_load_mesoscopic(bf_solver *s, const char *directory, int64_t tick) {
    if (!s || !directory) return BF_ERR_INVALID;
    File file;
    FileHeader h;
    bf_status st = open_state(s, directory, tick, kKindMeso, file, h);
    if (st != BF_OK) return st;
    BF_HIP(s, hipSetDevice(s->cfg.device));
    BF_HIP(s, hipStreamSynchronize(s->stream));
    const size_t N = (size_t)s->N;
    auto bad = [&]() { return fail(s, BF_ERR_INVALID, "[LBM Error] truncated or inconsistent checkpoint"); };
    std::vector<uint8_t> pops(s->q * N * s->elem), th, ps, force, ids, vd(16 * N), ft(16 * N);
    if (!get(file.f, kTagPopulations, pops.data(), pops.size())) return bad();
    if (h.thermal) {
        th.resize(s->qt * N * sizeof(float));
        if (!get(file.f, kTagThermal, th.data(), th.size())) return bad();
    }
    if (h.multiphase) {
        ps.resize(2 * N * sizeof(float));
        if (!get(file.f, kTagPsi, ps.data(), ps.size())) return bad();
    }
    if (h.forcing == 2) {
        force.resize(N * sizeof(float4));
        if (!get(file.f, kTagForce, force.data(), force.size())) return bad();
    }
    if (h.has_ids) {
        ids.resize(N);
        if (!get(file.f, kTagIds, ids.data(), N)) return bad();
    }
    if (!get(file.f, kTagVelocityDensity, vd.data(), vd.size()) || !get(file.f, kTagForceTemperature, ft.data(), ft.size()))
        return bad();
    // device buffers for what the state uses
    dfree(s->ids); s->ids = nullptr;
    dfree(s->nmask); s->nmask = nullptr;
    if (h.has_ids) {
        if ((st = dalloc(s, &s->ids, N)) != BF_OK) return st;
        if ((st = dalloc(s, &s->nmask, sizeof(uint32_t) * N)) != BF_OK) return st;
        BF_HIP(s, hipMemcpy(s->ids, ids.data(), N, hipMemcpyHostToDevice));
    }
    dfree(s->force_d); s->force_d = nullptr;
    if (h.forcing == 2) {
        if ((st = dalloc(s, &s->force_d, force.size())) != BF_OK) return st;
        BF_HIP(s, hipMemcpy(s->force_d, force.data(), force.size(), hipMemcpyHostToDevice));
    }
    dfree(s->g); s->g = nullptr;
    if (h.thermal) {
        if ((st = dalloc(s, &s->g, sizeof(float) * (size_t)s->qt * s->stride)) != BF_OK) return st;
        for (int k = 0; k < s->qt; ++k)
            BF_HIP(s, hipMemcpy(s->g + (size_t)k * s->stride, th.data() + k * N * sizeof(float), N * sizeof(float),
                                hipMemcpyHostToDevice));
    }
    for (int i = 0; i < s->q; ++i) BF_HIP(s, copy_slot(s, i, pops.data() + i * N * s->elem, false));
    if (h.multiphase) {
        BF_HIP(s, hipMemcpy(s->psi[0], ps.data(), N * sizeof(float), hipMemcpyHostToDevice));
        BF_HIP(s, hipMemcpy(s->psi[1], ps.data() + N * sizeof(float), N * sizeof(float), hipMemcpyHostToDevice));
    }
    BF_HIP(s, hipMemcpy(s->vd, vd.data(), vd.size(), hipMemcpyHostToDevice));
    BF_HIP(s, hipMemcpy(s->ft, ft.data(), ft.size(), hipMemcpyHostToDevice));
    apply_params(s, h);
    s->has_ids = h.has_ids;
    s->closed = closed_domain(s, h.has_ids ? ids.data() : nullptr);
    s->thermal = h.thermal;
    s->forcing = h.forcing;
    s->cur = h.cur;
    s->ticks = h.ticks;
    s->tick_base = h.tick_base;
    if ((st = upload_objects(s)) != BF_OK) return st;
    if (s->has_ids) BF_HIP(s, dispatch_mask(s));
    s->shape = step_shape(s->X, s->Y, s->Z, s->cfg.multiphase != 0,
                          cells_per_thread(s->storage, s->cfg.multiphase != 0, s->thermal != 0), s->vs);
    s->partial_blocks = (size_t)s->shape.grid.x * s->shape.grid.y * s->shape.grid.z;
    if (s->partial_blocks > s->partial_capacity) {
        dfree(s->partial);
        s->partial = nullptr;
        s->partial_capacity = s->partial_blocks;
        if ((st = dalloc(s, &s->partial, sizeof(double) * 5 * s->partial_capacity)) != BF_OK) return st;
    }
    BF_HIP(s, hipStreamSynchronize(s->stream));
    s->initialized = 1;
    return BF_OK;
}

}  // extern "C"

namespace bf {

#ifndef BF_ALIGN_ROWS
#define BF_ALIGN_ROWS 1  // build-time switch of the aligned even tick
#endif
#ifndef BF_BLOCK_CELLS
#define BF_BLOCK_CELLS 256
#endif

int x_direction_count(int vs) {
    switch (vs) {
    case D2Q9: return 6;
    case D3Q15: return 10;
    case D3Q19: return 10;
    case D3Q27: return 18;
    }
    return 0;
}

LaunchShape step_shape(int X, int Y, int Z, bool multiphase, int cpt, int vs) {
    LaunchShape sh{};
    sh.lds_bytes = 0;
    if (multiphase) {
        sh.block = dim3(64, 4, 1);
        cpt = 1;
    } else {
        const int per_thread_x = (X + cpt - 1) / cpt;
        const int row_threads = (per_thread_x + 63) / 64 * 64;
        if (row_threads <= 512) {
            // whole rows per block: aligned even tick through an LDS row buffer
            const int bx = row_threads;
            const int by = std::max(1, BF_BLOCK_CELLS / bx);
            sh.block = dim3(bx, by, 1);
            // [row][x][x-direction], the directions at an odd stride
            sh.lds_bytes = BF_ALIGN_ROWS ? sizeof(float) * (size_t)(x_direction_count(vs) | 1) * by * X : 0;
        } else {
            const int bx = std::min(BF_BLOCK_CELLS, row_threads);
            const int by = std::max(1, BF_BLOCK_CELLS / bx);
            sh.block = dim3(bx, by, 1);
        }
    }
    const int cells_x = (int)sh.block.x * cpt;
    sh.grid = dim3((X + cells_x - 1) / cells_x, (Y + sh.block.y - 1) / sh.block.y, Z);
    return sh;
}

}  // namespace bf

// ---- field textures: device-to-device copies ------------------------------
//
// The reference keeps its macroscopic fields in GL 3-D textures that the
// collide shader fills with imageStore (collide.comp:158-163,784-792; formats
// LBM.h:299-301, created in LBM.cpp:957-1003) and that its renderers sample.
// The MI355X has no image support (HIP arrays are not supported on gfx950)
// and no graphics pipeline, so no GL texture can live in or share its memory;
// a device consumer of the fields gets them in the texture encoding by a
// device-to-device copy instead.

namespace {

// R8 (unsigned normalised) value of imageStore(boundry_texture, vec4(id)):
// the store clamps to [0, 1], so fluid is 0 and every object 255.
__global__ void __launch_bounds__(256) k_boundary_r8(const uint8_t *__restrict__ ids, uint8_t *__restrict__ out,
                                                     int64_t n) {
    const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (c < n) out[c] = ids && ids[c] != 0 ? 255 : 0;
}

}  // namespace

extern "C" {

bf_status bf_copy_field_to_device(bf_solver *s, int32_t field, void *dst, size_t bytes) {
    if (!s || !dst) return BF_ERR_INVALID;
    if (!s->initialized) return fail(s, BF_ERR_STATE, "[LBM Error] field copy before bf_init_fields");
    size_t need = 0;
    if (field == BF_FIELD_VELOCITY_DENSITY || field == BF_FIELD_FORCE_TEMPERATURE) need = sizeof(float4) * (size_t)s->N;
    else if (field == BF_FIELD_BOUNDARY) need = (size_t)s->N;
    else return fail(s, BF_ERR_INVALID, "[LBM Error] field %d has no texture", field);
    if (bytes != need)
        return fail(s, BF_ERR_INVALID, "[LBM Error] field copy: %zu bytes given, the field has %zu", bytes, need);
    BF_HIP(s, hipSetDevice(s->cfg.device));
    if (field == BF_FIELD_BOUNDARY) {
        hipLaunchKernelGGL(k_boundary_r8, dim3((unsigned)((s->N + 255) / 256)), dim3(256), 0, s->stream, s->ids,
                           static_cast<uint8_t *>(dst), s->N);
        BF_HIP(s, hipGetLastError());
        return BF_OK;
    }
    const void *src = field == BF_FIELD_VELOCITY_DENSITY ? (const void *)s->vd : (const void *)s->ft;
    BF_HIP(s, hipMemcpyAsync(dst, src, need, hipMemcpyDeviceToDevice, s->stream));
    return BF_OK;
}

}  // extern "C"
