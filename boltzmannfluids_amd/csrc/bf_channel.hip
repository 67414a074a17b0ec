// bf_channel.hip — differentiable channel: forward ticks and the hand-written
// discrete adjoint, for D2Q9 and D3Q19 on gfx950.
//
// Geometry: walls (halfway bounce-back) at y = 0 and y = Y-1, z periodic,
// Zou-He velocity inlet at x = 0 with a Poiseuille profile of mean u_in and
// Zou-He pressure outlet at x = X-1 (Zou & He, Phys. Fluids 9:1591, 1997;
// D3Q19 transverse terms after Hecht & Harting, J. Stat. Mech. 2010 P01018).
// Porosity gamma in [0, 1] per cell enters through partial bounce-back
// (Walsh, Burwinkle & Saar, Comput. Geosci. 35:1186, 2009):
//   f**_i = (1 - gamma) f*_i + gamma f_opp(i),   f* = BGK post-collision.
// The equilibrium and BGK step follow collide.comp:589-629 of the reference.
//
// State: s^t = post-collision populations of tick t, population-major, held
// shifted (h = f - w, rho = 1 + sum h) so that FP32 sums carry no O(1) part.
// The closures and partial bounce-back keep their form in h because
// w_i = w_opp(i) and sum_{c_x=0} w + 2 sum_{c_x=-s} w = 1.  A tick
// pulls the pre-collision populations of a cell from s^{t-1} (one load per
// direction), closes the boundaries, collides and writes s^t locally.  The
// adjoint of a tick runs per cell: it reads dJ/ds^t at the cell, re-derives the
// cell's forward state from s^{t-1}, applies the transpose of collision +
// partial bounce-back (O(Q) through three moment sums) and the transpose of the
// boundary closure, giving m_i(x) = dJ/d(pulled f_i at x).  m is stored at the
// puller (slot i of cell x: aligned, local stores); the next adjoint tick
// gathers dJ/ds^{t-1}_i(y) = m_i(y + c_i) with the mirror image of the forward
// pull (zero where no cell pulls the slot, m_opp(i)(y) where y pulled its own
// slot back from a wall).  Every slot has at most one puller, so the gather is
// exact, and every tick moves its misaligned accesses into loads.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/bf/bf.h"
#include "bf_lattice.h"

using namespace bf;

namespace {

struct ChArgs {
    int X, Y, Z, q;
    int64_t stride;
    float tau, u_in, rho_out;
    const float *gamma;   // N
    const float *prof;    // Y
};

template <int VS>
__device__ inline float dotc3(int i, float vx, float vy, float vz) {
    using L = Lattice<VS>;
    float r = 0.0f;
    bool first = true;
    if (L::c[i][0] != 0) { r = L::c[i][0] > 0 ? vx : -vx; first = false; }
    if (L::c[i][1] != 0) { float t = L::c[i][1] > 0 ? vy : -vy; r = first ? t : r + t; first = false; }
    if (L::c[i][2] != 0) { float t = L::c[i][2] > 0 ? vz : -vz; r = first ? t : r + t; }
    return r;
}

// Unknown-population weight of the closure: axis 2/3 (D2Q9) or 1/3 (D3Q19), diagonals 1/6.
template <int VS>
__device__ constexpr float zh_wt(int i) {
    using L = Lattice<VS>;
    const int nz = (L::c[i][0] != 0) + (L::c[i][1] != 0) + (L::c[i][2] != 0);
    return nz == 1 ? (L::Q == 9 ? 2.0f / 3.0f : 1.0f / 3.0f) : 1.0f / 6.0f;
}

// Every state and adjoint slot is read once and written once per tick: streaming hints.
__device__ inline float ldnt(const float *p) { return __builtin_nontemporal_load(p); }
__device__ inline void stnt(float *p, float v) { __builtin_nontemporal_store(v, p); }

struct CellPos {
    int x, y, z, self;
    int zm, zp;   // periodic z neighbours
};

template <int VS>
__device__ inline void pull(const ChArgs &a, const float *__restrict__ src, const CellPos &p, float *f) {
    using L = Lattice<VS>;
    const int XY = a.X * a.Y;
#pragma unroll
    for (int i = 0; i < L::Q; ++i) {
        const int sx = p.x - L::c[i][0], sy = p.y - L::c[i][1];
        const int sz = L::c[i][2] > 0 ? p.zm : (L::c[i][2] < 0 ? p.zp : p.z);
        if (L::c[i][0] != 0 && (sx < 0 || sx >= a.X)) { f[i] = 0.0f; continue; }
        if (L::c[i][1] != 0 && (sy == 0 || sy == a.Y - 1)) {
            f[i] = ldnt(src + (size_t)opposite(i) * a.stride + p.self);
            continue;
        }
        f[i] = ldnt(src + (size_t)i * a.stride + sx + sy * a.X + sz * XY);
    }
}

// Boundary closure; returns rho u_x and the partial sums used by its transpose.
template <int VS>
__device__ inline float closure(const ChArgs &a, float *f, int sign, float ux_in, float &S_out, float &ny_out,
                                float &nz_out) {
    using L = Lattice<VS>;
    float S = 0.0f, ny = 0.0f, nz = 0.0f;
    bool fs = true, fy = true, fz = true;
#pragma unroll
    for (int i = 0; i < L::Q; ++i) {
        if (L::c[i][0] == 0) {
            S = fs ? f[i] : S + f[i];
            fs = false;
            if (L::c[i][1] != 0) { float t = L::c[i][1] > 0 ? f[i] : -f[i]; ny = fy ? t : ny + t; fy = false; }
            if (L::c[i][2] != 0) { float t = L::c[i][2] > 0 ? f[i] : -f[i]; nz = fz ? t : nz + t; fz = false; }
        } else if (L::c[i][0] == -sign) {
            S = S + 2.0f * f[i];
        }
    }
    ny *= 0.5f;
    nz *= 0.5f;
    const float rux = sign > 0 ? (1.0f + S) * ux_in / (1.0f - ux_in) : S + (1.0f - a.rho_out);
#pragma unroll
    for (int i = 0; i < L::Q; ++i) {
        if (L::c[i][0] != sign) continue;
        float tr = 0.0f;
        bool first = true;
        if (L::c[i][1] != 0) { tr = L::c[i][1] > 0 ? ny : -ny; first = false; }
        if (L::c[i][2] != 0) { float t = L::c[i][2] > 0 ? nz : -nz; tr = first ? t : tr + t; }
        f[i] = f[opposite(i)] + (float)sign * rux * zh_wt<VS>(i) - tr;
    }
    S_out = 1.0f + S; ny_out = ny; nz_out = nz;
    return rux;
}

template <int VS>
__device__ inline bool cell_of(const ChArgs &a, CellPos &p) {
    p.x = blockIdx.x * blockDim.x + threadIdx.x;
    p.y = blockIdx.y * blockDim.y + threadIdx.y + 1;  // fluid rows 1..Y-2
    p.z = blockIdx.z;
    if (p.x >= a.X || p.y >= a.Y - 1) return false;
    p.zm = p.z == 0 ? a.Z - 1 : p.z - 1;
    p.zp = p.z == a.Z - 1 ? 0 : p.z + 1;
    p.self = p.x + p.y * a.X + p.z * a.X * a.Y;
    return true;
}

template <int VS>
__device__ inline void precollision(const ChArgs &a, const float *__restrict__ src, const CellPos &p, float *f,
                                    int &sign, float &S, float &ny, float &nz, float &uin) {
    pull<VS>(a, src, p, f);
    sign = p.x == 0 ? 1 : (p.x == a.X - 1 ? -1 : 0);
    uin = a.u_in * a.prof[p.y];
    S = ny = nz = 0.0f;
    if (sign) closure<VS>(a, f, sign, uin, S, ny, nz);
}

template <int VS>
__global__ void __launch_bounds__(256) k_ch_step(const ChArgs a, const float *__restrict__ src, float *__restrict__ dst) {
    using L = Lattice<VS>;
    CellPos p;
    if (!cell_of<VS>(a, p)) return;
    float f[L::Q];
    int sign;
    float S, ny, nz, uin;
    precollision<VS>(a, src, p, f, sign, S, ny, nz, uin);
    float delta = f[0];
#pragma unroll
    for (int i = 1; i < L::Q; ++i) delta += f[i];
    const float rho = 1.0f + delta;
    float mx = 0.0f, my = 0.0f, mz = 0.0f;
    bool bx = true, by = true, bz = true;
#pragma unroll
    for (int i = 0; i < L::Q; ++i) {
        if (L::c[i][0] != 0) { float t = L::c[i][0] > 0 ? f[i] : -f[i]; mx = bx ? t : mx + t; bx = false; }
        if (L::c[i][1] != 0) { float t = L::c[i][1] > 0 ? f[i] : -f[i]; my = by ? t : my + t; by = false; }
        if (L::c[i][2] != 0) { float t = L::c[i][2] > 0 ? f[i] : -f[i]; mz = bz ? t : mz + t; bz = false; }
    }
    const float ux = mx / rho, uy = my / rho, uz = mz / rho;
    const float uu = ux * ux + uy * uy + uz * uz;
    const float g = a.gamma[p.self];
    float fs[L::Q];
#pragma unroll
    for (int i = 0; i < L::Q; ++i) {
        const float cu = i == 0 ? 0.0f : dotc3<VS>(i, ux, uy, uz);
        const float e = L::w[i] * (delta + rho * (3.0f * cu + 4.5f * cu * cu - 1.5f * uu));
        fs[i] = f[i] - (f[i] - e) / a.tau;
    }
#pragma unroll
    for (int i = 0; i < L::Q; ++i)
        stnt(dst + (size_t)i * a.stride + p.self, (1.0f - g) * fs[i] + g * f[opposite(i)]);
}

// dJ/ds_i(y) from the puller-indexed adjoint m of the following tick: the
// forward pull (pull above) read slot i of cell y from cell y + c_i, or from y
// itself through direction opp(i) when y + c_i lies in a wall row, or not at
// all when y + c_i leaves the channel in x.
template <int VS>
__device__ inline void gather(const ChArgs &a, const float *__restrict__ m, const CellPos &p, float *mu) {
    using L = Lattice<VS>;
    const int XY = a.X * a.Y;
#pragma unroll
    for (int i = 0; i < L::Q; ++i) {
        const int tx = p.x + L::c[i][0], ty = p.y + L::c[i][1];
        const int tz = L::c[i][2] > 0 ? p.zp : (L::c[i][2] < 0 ? p.zm : p.z);
        if (L::c[i][0] != 0 && (tx < 0 || tx >= a.X)) { mu[i] = 0.0f; continue; }
        if (L::c[i][1] != 0 && (ty == 0 || ty == a.Y - 1)) {
            mu[i] = ldnt(m + (size_t)opposite(i) * a.stride + p.self);
            continue;
        }
        mu[i] = ldnt(m + (size_t)i * a.stride + tx + ty * a.X + tz * XY);
    }
}

// One adjoint tick.  PLAIN: `nu` holds dJ/ds^t slot by slot (the seed);
// otherwise `nu` is the puller-indexed m of tick t + 1.  Writes m of tick t.
template <int VS, bool PLAIN>
__global__ void __launch_bounds__(256) k_ch_adjoint(const ChArgs a, const float *__restrict__ src,
                                                    const float *__restrict__ nu, float *__restrict__ m_out,
                                                    float *__restrict__ dg, double *__restrict__ du_rows) {
    using L = Lattice<VS>;
    constexpr int Q = L::Q;
    CellPos p;
    if (!cell_of<VS>(a, p)) return;
    // the adjoint slots are loaded before the state is pulled: both sets of
    // loads are in flight together
    float mu[Q];
    if constexpr (PLAIN) {
#pragma unroll
        for (int i = 0; i < Q; ++i) mu[i] = ldnt(nu + (size_t)i * a.stride + p.self);
    } else {
        gather<VS>(a, nu, p, mu);
    }
    const float g = a.gamma[p.self];
    float f[Q];
    int sign;
    float S, ny, nz, uin;
    precollision<VS>(a, src, p, f, sign, S, ny, nz, uin);
    float delta = f[0];
#pragma unroll
    for (int i = 1; i < Q; ++i) delta += f[i];
    const float rho = 1.0f + delta;
    float mx = 0.0f, my = 0.0f, mz = 0.0f;
    bool bx = true, by = true, bz = true;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        if (L::c[i][0] != 0) { float t = L::c[i][0] > 0 ? f[i] : -f[i]; mx = bx ? t : mx + t; bx = false; }
        if (L::c[i][1] != 0) { float t = L::c[i][1] > 0 ? f[i] : -f[i]; my = by ? t : my + t; by = false; }
        if (L::c[i][2] != 0) { float t = L::c[i][2] > 0 ? f[i] : -f[i]; mz = bz ? t : mz + t; bz = false; }
    }
    const float ux = mx / rho, uy = my / rho, uz = mz / rho;
    const float uu = ux * ux + uy * uy + uz * uz;

    // d post_i / d gamma = f_opp(i) - f*_i; transpose of collision + partial bounce-back
    float acc = 0.0f, A = 0.0f, C = 0.0f, Bx = 0.0f, By = 0.0f, Bz = 0.0f;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        const float cu = i == 0 ? 0.0f : dotc3<VS>(i, ux, uy, uz);
        const float e = L::w[i] * (delta + rho * (3.0f * cu + 4.5f * cu * cu - 1.5f * uu));
        const float fsi = f[i] - (f[i] - e) / a.tau;
        acc += mu[i] * (f[opposite(i)] - fsi);
        const float wm = L::w[i] * mu[i];
        A += wm * (1.0f - 4.5f * cu * cu + 1.5f * uu);
        C += wm;
        const float b = wm * (3.0f + 9.0f * cu);
        if (L::c[i][0] != 0) Bx += L::c[i][0] > 0 ? b : -b;
        if (L::c[i][1] != 0) By += L::c[i][1] > 0 ? b : -b;
        if (L::c[i][2] != 0) Bz += L::c[i][2] > 0 ? b : -b;
    }
    dg[p.self] += acc;
    const float Dx = Bx - 3.0f * C * ux, Dy = By - 3.0f * C * uy, Dz = Bz - 3.0f * C * uz;
    const float a0 = (1.0f - g) * (1.0f - 1.0f / a.tau), a1 = (1.0f - g) / a.tau;
    float mh[Q];
#pragma unroll
    for (int j = 0; j < Q; ++j)
        mh[j] = a0 * mu[j] + g * mu[opposite(j)] + a1 * (A + dotc3<VS>(j, Dx, Dy, Dz));

    // transpose of the closure: the unknown slots feed the known ones
    if (sign) {
        float gr = 0.0f, gny = 0.0f, gnz = 0.0f;
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            if (L::c[i][0] != sign) continue;
            mh[opposite(i)] += mh[i];
            gr += mh[i] * (float)sign * zh_wt<VS>(i);
            if (L::c[i][1] != 0) gny -= L::c[i][1] > 0 ? mh[i] : -mh[i];
            if (L::c[i][2] != 0) gnz -= L::c[i][2] > 0 ? mh[i] : -mh[i];
            mh[i] = 0.0f;
        }
        const float dS = sign > 0 ? gr * uin / (1.0f - uin) : gr;
        if (sign > 0)
            du_rows[p.y + p.z * a.Y] += (double)(gr * S / ((1.0f - uin) * (1.0f - uin)) * a.prof[p.y]);
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            if (L::c[i][0] == 0) {
                float t = 0.0f;
                if (L::c[i][1] != 0) t += L::c[i][1] > 0 ? gny : -gny;
                if (L::c[i][2] != 0) t += L::c[i][2] > 0 ? gnz : -gnz;
                mh[i] += dS + 0.5f * t;
            } else if (L::c[i][0] == -sign) {
                mh[i] += 2.0f * dS;
            }
        }
    }

    // m at the puller; the transpose of the pull is the next tick's gather
#pragma unroll
    for (int i = 0; i < Q; ++i) stnt(m_out + (size_t)i * a.stride + p.self, mh[i]);
}

// J = mean over the outlet cells of u_x = (S / rho_out - 1); one block, fixed order.
template <int VS>
__global__ void __launch_bounds__(256) k_ch_objective(const ChArgs a, const float *__restrict__ src, double *out) {
    using L = Lattice<VS>;
    __shared__ double red[256];
    double acc = 0.0;
    const int nrows = (a.Y - 2) * a.Z;
    for (int r = threadIdx.x; r < nrows; r += 256) {
        CellPos p;
        p.x = a.X - 1;
        p.y = r % (a.Y - 2) + 1;
        p.z = r / (a.Y - 2);
        p.zm = p.z == 0 ? a.Z - 1 : p.z - 1;
        p.zp = p.z == a.Z - 1 ? 0 : p.z + 1;
        p.self = p.x + p.y * a.X + p.z * a.X * a.Y;
        float f[L::Q];
        pull<VS>(a, src, p, f);
        float S, ny, nz;
        const float rux = closure<VS>(a, f, -1, 0.0f, S, ny, nz);
        acc += (double)(rux / a.rho_out);
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0] / (double)nrows;
}

// dJ/ds^{T-1}: the objective's sensitivities scattered to the slots the outlet
// cells pull (nu must be zero).
template <int VS>
__global__ void __launch_bounds__(256) k_ch_seed(const ChArgs a, float *__restrict__ nu) {
    using L = Lattice<VS>;
    const int nrows = (a.Y - 2) * a.Z;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    const float k = 1.0f / ((float)nrows * a.rho_out);
    CellPos p;
    p.x = a.X - 1;
    p.y = r % (a.Y - 2) + 1;
    p.z = r / (a.Y - 2);
    p.zm = p.z == 0 ? a.Z - 1 : p.z - 1;
    p.zp = p.z == a.Z - 1 ? 0 : p.z + 1;
    p.self = p.x + p.y * a.X + p.z * a.X * a.Y;
    const int XY = a.X * a.Y;
#pragma unroll
    for (int i = 0; i < L::Q; ++i) {
        const float coef = L::c[i][0] == 0 ? k : (L::c[i][0] == 1 ? 2.0f * k : 0.0f);
        if (coef == 0.0f) continue;
        const int sx = p.x - L::c[i][0], sy = p.y - L::c[i][1];
        const int sz = L::c[i][2] > 0 ? p.zm : (L::c[i][2] < 0 ? p.zp : p.z);
        if (L::c[i][1] != 0 && (sy == 0 || sy == a.Y - 1)) nu[(size_t)opposite(i) * a.stride + p.self] = coef;
        else nu[(size_t)i * a.stride + sx + sy * a.X + sz * XY] = coef;
    }
}

// Pre-collision moments of the state after the last run: (u_x, u_y, u_z, rho).
template <int VS>
__global__ void __launch_bounds__(256) k_ch_moments(const ChArgs a, const float *__restrict__ src, float4 *out) {
    using L = Lattice<VS>;
    CellPos p;
    if (!cell_of<VS>(a, p)) return;
    float f[L::Q];
    int sign;
    float S, ny, nz, uin;
    precollision<VS>(a, src, p, f, sign, S, ny, nz, uin);
    float rho = 1.0f, mx = 0.0f, my = 0.0f, mz = 0.0f;
#pragma unroll
    for (int i = 0; i < L::Q; ++i) {
        rho += f[i];
        mx += f[i] * L::c[i][0];
        my += f[i] * L::c[i][1];
        mz += f[i] * L::c[i][2];
    }
    out[p.self] = make_float4(mx / rho, my / rho, mz / rho, rho);
}

// Fluid at rest: f = w, i.e. h = 0.
template <int VS>
__global__ void __launch_bounds__(256) k_ch_rest(float *__restrict__ s, int64_t stride, int64_t n) {
    using L = Lattice<VS>;
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
#pragma unroll
    for (int i = 0; i < L::Q; ++i) s[(size_t)i * stride + c] = 0.0f;
}

__global__ void __launch_bounds__(256) k_zero_f32(float *__restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

__global__ void __launch_bounds__(256) k_zero_f64(double *__restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// total += segment; segment = 0
__global__ void __launch_bounds__(256) k_fold(float *__restrict__ seg, double *__restrict__ total, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        total[i] += (double)seg[i];
        seg[i] = 0.0f;
    }
}

}  // namespace

struct bf_channel {
    bf_channel_config cfg{};
    int q = 0;
    int64_t N = 0, stride = 0;
    hipStream_t stream = nullptr;
    hipStream_t stream2 = nullptr;         // segment recomputations, beside the adjoint sweep
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    float *gamma = nullptr, *prof = nullptr;
    float *pp[2] = {nullptr, nullptr};     // forward ping-pong
    float *rest = nullptr;                 // s^{-1}
    float *last = nullptr;                 // state after the last run (points into another buffer)
    double *scalar = nullptr;              // J
    float4 *moments = nullptr;
    double timing_ms[3] = {0, 0, 0};
    std::string err;
};

static std::string g_channel_error;

namespace {

bf_status cfail(bf_channel *c, bf_status st, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf; else g_channel_error = buf;
    return st;
}

#define CH_HIP(c, call)                                                                          \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return cfail((c), e_ == hipErrorOutOfMemory ? BF_ERR_OOM : BF_ERR_DEVICE,           \
                         "[LBM Error] %s failed: %s", #call, hipGetErrorString(e_));             \
    } while (0)

ChArgs args_of(const bf_channel *c) {
    ChArgs a{};
    a.X = c->cfg.resolution[0]; a.Y = c->cfg.resolution[1]; a.Z = c->cfg.resolution[2];
    a.q = c->q; a.stride = c->stride;
    a.tau = c->cfg.relaxation_time; a.u_in = c->cfg.inlet_velocity; a.rho_out = c->cfg.outlet_density;
    a.gamma = c->gamma; a.prof = c->prof;
    return a;
}

void shape(const bf_channel *c, dim3 &grid, dim3 &block) {
    const int X = c->cfg.resolution[0], Y = c->cfg.resolution[1], Z = c->cfg.resolution[2];
    const int bx = std::min(256, (X + 63) / 64 * 64);
    const int by = std::max(1, 256 / bx);
    block = dim3(bx, by, 1);
    grid = dim3((X + bx - 1) / bx, (Y - 2 + by - 1) / by, Z);
}

hipError_t launch_tick(const bf_channel *c, const float *src, float *dst, hipStream_t stream) {
    dim3 g, b;
    shape(c, g, b);
    const ChArgs a = args_of(c);
    if (c->q == 9) hipLaunchKernelGGL(k_ch_step<D2Q9>, g, b, 0, stream, a, src, dst);
    else hipLaunchKernelGGL(k_ch_step<D3Q19>, g, b, 0, stream, a, src, dst);
    return hipGetLastError();
}
hipError_t launch_tick(const bf_channel *c, const float *src, float *dst) { return launch_tick(c, src, dst, c->stream); }

hipError_t launch_adjoint(const bf_channel *c, bool plain, const float *src, const float *nu, float *m_out, float *dg,
                          double *du_rows) {
    dim3 g, b;
    shape(c, g, b);
    const ChArgs a = args_of(c);
    if (c->q == 9) {
        if (plain) hipLaunchKernelGGL((k_ch_adjoint<D2Q9, true>), g, b, 0, c->stream, a, src, nu, m_out, dg, du_rows);
        else hipLaunchKernelGGL((k_ch_adjoint<D2Q9, false>), g, b, 0, c->stream, a, src, nu, m_out, dg, du_rows);
    } else {
        if (plain) hipLaunchKernelGGL((k_ch_adjoint<D3Q19, true>), g, b, 0, c->stream, a, src, nu, m_out, dg, du_rows);
        else hipLaunchKernelGGL((k_ch_adjoint<D3Q19, false>), g, b, 0, c->stream, a, src, nu, m_out, dg, du_rows);
    }
    return hipGetLastError();
}

hipError_t launch_objective(const bf_channel *c, const float *src) {
    const ChArgs a = args_of(c);
    if (c->q == 9) hipLaunchKernelGGL(k_ch_objective<D2Q9>, dim3(1), dim3(256), 0, c->stream, a, src, c->scalar);
    else hipLaunchKernelGGL(k_ch_objective<D3Q19>, dim3(1), dim3(256), 0, c->stream, a, src, c->scalar);
    return hipGetLastError();
}

hipError_t launch_rest(const bf_channel *c, float *s) {
    const dim3 g((unsigned)((c->N + 255) / 256)), b(256);
    if (c->q == 9) hipLaunchKernelGGL(k_ch_rest<D2Q9>, g, b, 0, c->stream, s, c->stride, c->N);
    else hipLaunchKernelGGL(k_ch_rest<D3Q19>, g, b, 0, c->stream, s, c->stride, c->N);
    return hipGetLastError();
}

hipError_t zero_f32(const bf_channel *c, float *p, int64_t n) {
    hipLaunchKernelGGL(k_zero_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, p, n);
    return hipGetLastError();
}

hipError_t zero_f64(const bf_channel *c, double *p, int64_t n) {
    hipLaunchKernelGGL(k_zero_f64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, p, n);
    return hipGetLastError();
}

// Device buffer list freed together on every exit path.
struct Scratch {
    std::vector<void *> ptrs;
    std::vector<hipEvent_t> events;
    ~Scratch() {
        for (void *p : ptrs) (void)hipFree(p);
        for (hipEvent_t e : events) (void)hipEventDestroy(e);
    }
    hipError_t event(hipEvent_t *e) {
        hipError_t r = hipEventCreate(e);
        if (r == hipSuccess) events.push_back(*e);
        return r;
    }
    template <class T>
    hipError_t alloc(T **p, size_t bytes) {
        hipError_t e = hipMalloc((void **)p, bytes);
        if (e == hipSuccess) ptrs.push_back(*p);
        return e;
    }
};

}  // namespace

extern "C" {

bf_status bf_channel_create(const bf_channel_config *cfg, bf_channel **out) {
    if (!cfg || !out) return cfail(nullptr, BF_ERR_INVALID, "[LBM Error] bf_channel_create: null argument");
    *out = nullptr;
    const int X = cfg->resolution[0], Y = cfg->resolution[1], Z = cfg->resolution[2];
    if (cfg->velocity_set != BF_D2Q9 && cfg->velocity_set != BF_D3Q19)
        return cfail(nullptr, BF_ERR_INVALID, "[LBM Error] bf_channel_create: velocity set must be D2Q9 or D3Q19");
    if (X < 2 || Y < 3 || Z < 1 || (cfg->velocity_set == BF_D2Q9 && Z != 1))
        return cfail(nullptr, BF_ERR_INVALID, "[LBM Error] bf_channel_create: resolution (%d, %d, %d) is not valid", X, Y, Z);
    if ((int64_t)X * Y * Z >= (int64_t(1) << 31))
        return cfail(nullptr, BF_ERR_INVALID, "[LBM Error] bf_channel_create: too many cells");
    if (!(cfg->relaxation_time > 0.5f))
        return cfail(nullptr, BF_ERR_INVALID, "[LBM Error] bf_channel_create: relaxation time must exceed 0.5");
    if (!(cfg->inlet_velocity < 0.5f) || !(cfg->inlet_velocity > -0.5f) || !(cfg->outlet_density > 0.0f))
        return cfail(nullptr, BF_ERR_INVALID, "[LBM Error] bf_channel_create: inlet velocity or outlet density out of range");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return cfail(nullptr, BF_ERR_DEVICE, "[LBM Error] bf_channel_create: no HIP device available");
    if (cfg->device < 0 || cfg->device >= ndev)
        return cfail(nullptr, BF_ERR_INVALID, "[LBM Error] bf_channel_create: device out of range");
    bf_channel *c = new (std::nothrow) bf_channel();
    if (!c) return cfail(nullptr, BF_ERR_OOM, "[LBM Error] host allocation failed");
    c->cfg = *cfg;
    c->q = cfg->velocity_set == BF_D2Q9 ? 9 : 19;
    c->N = (int64_t)X * Y * Z;
    c->stride = (c->N + 63) / 64 * 64;
    auto bail = [&](const char *what) {
        g_channel_error = std::string("[LBM Error] bf_channel_create: ") + what;
        bf_channel_destroy(c);
        return BF_ERR_DEVICE;
    };
    if (hipSetDevice(cfg->device) != hipSuccess) return bail("hipSetDevice");
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking) != hipSuccess)
        return bail("stream");
    for (auto &e : c->ev)
        if (hipEventCreate(&e) != hipSuccess) return bail("event");
    const size_t sb = sizeof(float) * c->q * c->stride;
    if (hipMalloc((void **)&c->gamma, sizeof(float) * c->N) != hipSuccess ||
        hipMalloc((void **)&c->prof, sizeof(float) * Y) != hipSuccess ||
        hipMalloc((void **)&c->pp[0], sb) != hipSuccess || hipMalloc((void **)&c->pp[1], sb) != hipSuccess ||
        hipMalloc((void **)&c->rest, sb) != hipSuccess || hipMalloc((void **)&c->scalar, sizeof(double) * 4) != hipSuccess ||
        hipMalloc((void **)&c->moments, sizeof(float4) * c->N) != hipSuccess)
        return bail("device allocation");
    // Poiseuille profile 6 s (1 - s), s = (y - 1/2)/(Y - 2): unit mean over the channel between the halfway walls
    std::vector<float> prof(Y, 0.0f);
    for (int y = 1; y < Y - 1; ++y) {
        const double e = (y - 0.5) / (Y - 2.0);
        prof[y] = (float)(6.0 * e * (1.0 - e));
    }
    if (hipMemcpy(c->prof, prof.data(), sizeof(float) * Y, hipMemcpyHostToDevice) != hipSuccess) return bail("upload");
    if (zero_f32(c, c->gamma, c->N) != hipSuccess || launch_rest(c, c->rest) != hipSuccess ||
        launch_rest(c, c->pp[0]) != hipSuccess || launch_rest(c, c->pp[1]) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return bail("initialisation");
    c->last = c->rest;
    *out = c;
    return BF_OK;
}

void bf_channel_destroy(bf_channel *c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->stream2) (void)hipStreamSynchronize(c->stream2);
    for (void *p : {(void *)c->gamma, (void *)c->prof, (void *)c->pp[0], (void *)c->pp[1], (void *)c->rest,
                    (void *)c->scalar, (void *)c->moments})
        if (p) (void)hipFree(p);
    for (auto &e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    if (c->stream2) (void)hipStreamDestroy(c->stream2);
    delete c;
}

const char *bf_channel_last_error(const bf_channel *c) { return c ? c->err.c_str() : g_channel_error.c_str(); }

bf_status bf_channel_set_porosity(bf_channel *c, const float *gamma) {
    if (!c || !gamma) return BF_ERR_INVALID;
    for (int64_t k = 0; k < c->N; ++k)
        if (!(gamma[k] >= 0.0f && gamma[k] <= 1.0f))
            return cfail(c, BF_ERR_INVALID, "[LBM Error] porosity must lie in [0, 1] (cell %lld)", (long long)k);
    CH_HIP(c, hipSetDevice(c->cfg.device));
    CH_HIP(c, hipMemcpyAsync(c->gamma, gamma, sizeof(float) * c->N, hipMemcpyHostToDevice, c->stream));
    CH_HIP(c, hipStreamSynchronize(c->stream));
    return BF_OK;
}

bf_status bf_channel_set_inlet_velocity(bf_channel *c, float u) {
    if (!c || !(u < 0.5f) || !(u > -0.5f)) return c ? cfail(c, BF_ERR_INVALID, "[LBM Error] inlet velocity out of range") : BF_ERR_INVALID;
    c->cfg.inlet_velocity = u;
    return BF_OK;
}

bf_status bf_channel_forward(bf_channel *c, int64_t steps, double *J) {
    if (!c || !J || steps < 1) return c ? cfail(c, BF_ERR_INVALID, "[LBM Error] bf_channel_forward: bad arguments") : BF_ERR_INVALID;
    CH_HIP(c, hipSetDevice(c->cfg.device));
    CH_HIP(c, hipEventRecord(c->ev[0], c->stream));
    const float *src = c->rest;
    for (int64_t t = 0; t < steps; ++t) {
        float *dst = c->pp[t & 1];
        CH_HIP(c, launch_tick(c, src, dst));
        src = dst;
    }
    CH_HIP(c, hipEventRecord(c->ev[1], c->stream));
    CH_HIP(c, launch_objective(c, src));
    double j = 0.0;
    CH_HIP(c, hipMemcpyAsync(&j, c->scalar, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CH_HIP(c, hipStreamSynchronize(c->stream));
    float ms = 0.0f;
    CH_HIP(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
    c->timing_ms[0] = ms;
    c->last = const_cast<float *>(src);
    *J = j;
    return BF_OK;
}

bf_status bf_channel_gradient(bf_channel *c, const bf_adjoint_config *acfg, double *J, float *dJ_dgamma,
                              double *dJ_duin) {
    if (!c || !acfg || !J || !dJ_dgamma || !dJ_duin) return BF_ERR_INVALID;
    const int64_t T = acfg->steps;
    const int64_t K = acfg->checkpoint_every;
    if (T < 1 || K < 1) return cfail(c, BF_ERR_INVALID, "[LBM Error] bf_channel_gradient: steps and checkpoint_every must be >= 1");
    CH_HIP(c, hipSetDevice(c->cfg.device));
    const int64_t nseg = (T + K - 1) / K;
    const size_t sb = sizeof(float) * c->q * c->stride;
    // segments the trajectory budget keeps resident (at least one)
    int64_t resident = std::max<int64_t>(1, acfg->trajectory_budget_bytes / (int64_t)(sb * K));
    resident = std::min(resident, nseg);
    // Trajectory slots: one per resident segment, plus one when segments are
    // recomputed, so that the recomputation of the next segment (second
    // stream) runs beside the adjoint sweep of the current one.  The forward
    // pass fills that extra slot too, with the last segment before the
    // resident ones, which is then not recomputed.
    const int64_t nrec = nseg - resident;
    const int64_t nslots = resident + (nrec > 0 ? 1 : 0);
    const int64_t first_kept = nseg - nslots;   // segments from here keep the forward pass's states
    Scratch sc;
    float *ckpt = nullptr, *traj = nullptr, *nu[2] = {nullptr, nullptr}, *dgseg = nullptr;
    double *dg64 = nullptr, *du_rows = nullptr;
    const int Y = c->cfg.resolution[1], Z = c->cfg.resolution[2];
    const int64_t per_seg = std::max<int64_t>(1, K - 1);
    if (sc.alloc(&ckpt, sb * (size_t)nseg) != hipSuccess || sc.alloc(&traj, sb * (size_t)(per_seg * nslots)) != hipSuccess ||
        sc.alloc(&nu[0], sb) != hipSuccess || sc.alloc(&nu[1], sb) != hipSuccess ||
        sc.alloc(&dgseg, sizeof(float) * c->N) != hipSuccess || sc.alloc(&dg64, sizeof(double) * c->N) != hipSuccess ||
        sc.alloc(&du_rows, sizeof(double) * Y * Z) != hipSuccess)
        return cfail(c, BF_ERR_OOM, "[LBM Error] bf_channel_gradient: cannot allocate %lld checkpoints + %lld trajectory segments",
                     (long long)nseg, (long long)nslots);
    const int64_t qs = (int64_t)c->q * c->stride;
    // Segment k covers ticks [kK, min(T, (k+1)K)).  State j of segment k is
    // s^{kK+j-1}, the state tick kK+j reads: j = 0 is checkpoint k, j >= 1 live
    // in the segment's trajectory slot (resident segments keep their own, the
    // segment before them the extra slot; a recomputed segment takes the
    // extra slot or slot 0, whichever holds no segment still to be swept).
    auto resident_slot = [&](int64_t k) {
        if (k >= nseg - resident) return k - (nseg - resident);
        return k == first_kept ? resident : (int64_t)-1;   // the extra slot
    };
    std::vector<int64_t> seg_slot((size_t)nseg, -1);
    for (int64_t k = first_kept; k < nseg; ++k) seg_slot[(size_t)k] = resident_slot(k);
    auto seg_state = [&](int64_t k, int64_t j) -> float * {
        if (j == 0) return ckpt + k * qs;
        return traj + (seg_slot[(size_t)k] * per_seg + (j - 1)) * qs;
    };

    // ---- forward with checkpoints (and the resident segments' trajectories) ----
    CH_HIP(c, hipEventRecord(c->ev[0], c->stream));
    CH_HIP(c, hipMemcpyAsync(ckpt, c->rest, sb, hipMemcpyDeviceToDevice, c->stream));
    const float *src = ckpt;
    for (int64_t t = 0; t < T; ++t) {
        const int64_t k = t / K, j = t - k * K;   // tick t reads state j of segment k
        float *dst;
        if (t + 1 < T && (t + 1) % K == 0) dst = seg_state(k + 1, 0);                         // next checkpoint
        else if (t + 1 < T && resident_slot(k) >= 0) dst = seg_state(k, j + 1);               // resident state
        else dst = c->pp[t & 1] == src ? c->pp[(t + 1) & 1] : c->pp[t & 1];                  // not kept
        CH_HIP(c, launch_tick(c, src, dst));
        src = dst;
    }
    float *final_state = const_cast<float *>(src);
    CH_HIP(c, launch_objective(c, src));
    CH_HIP(c, hipEventRecord(c->ev[1], c->stream));

    // ---- adjoint sweep ----
    CH_HIP(c, zero_f32(c, nu[0], qs));
    CH_HIP(c, zero_f32(c, nu[1], qs));
    CH_HIP(c, zero_f32(c, dgseg, c->N));
    CH_HIP(c, zero_f64(c, dg64, c->N));
    CH_HIP(c, zero_f64(c, du_rows, (int64_t)Y * Z));
    {
        const ChArgs a = args_of(c);
        const int nrows = (Y - 2) * Z;
        const dim3 g((unsigned)((nrows + 255) / 256)), b(256);
        if (c->q == 9) hipLaunchKernelGGL(k_ch_seed<D2Q9>, g, b, 0, c->stream, a, nu[0]);
        else hipLaunchKernelGGL(k_ch_seed<D3Q19>, g, b, 0, c->stream, a, nu[0]);
        CH_HIP(c, hipGetLastError());
    }
    int cur = 0;          // nu[cur]: dJ/ds^T (the seed), then m of the tick after
    bool plain = true;    // only the first adjoint tick reads the seed slot by slot
    // Recomputations run on the second stream, each timed by an event pair
    // (read after the sweep: the host never waits between segments).  A
    // recomputation into a slot waits for the sweep that last used the slot;
    // the sweep of a recomputed segment waits for its recomputation.
    std::vector<hipEvent_t> rec_ev((size_t)(2 * nrec)), slot_free((size_t)nslots);
    for (auto &e : rec_ev)
        if (sc.event(&e) != hipSuccess) return cfail(c, BF_ERR_DEVICE, "[LBM Error] bf_channel_gradient: cannot create events");
    for (auto &e : slot_free)
        if (sc.event(&e) != hipSuccess) return cfail(c, BF_ERR_DEVICE, "[LBM Error] bf_channel_gradient: cannot create events");
    hipEvent_t fwd_done;
    if (sc.event(&fwd_done) != hipSuccess) return cfail(c, BF_ERR_DEVICE, "[LBM Error] bf_channel_gradient: cannot create events");
    std::vector<int64_t> owner((size_t)nslots, -1);   // segment held by the slot and not swept yet
    for (int64_t k = first_kept; k < nseg; ++k) owner[(size_t)resident_slot(k)] = k;
    std::vector<char> released((size_t)nslots, 0);     // slot_free recorded at least once
    std::vector<int64_t> rec_of((size_t)nseg, -1);     // index of the segment's event pair
    int64_t k_rec = first_kept - 1, n_rec = 0;         // next segment to recompute (descending)
    CH_HIP(c, hipEventRecord(fwd_done, c->stream));
    CH_HIP(c, hipStreamWaitEvent(c->stream2, fwd_done, 0));
    auto try_recompute = [&]() -> hipError_t {
        if (k_rec < 0) return hipSuccess;
        const int64_t extra = resident;
        int64_t sl = -1;
        if (owner[(size_t)extra] < 0) sl = extra;
        else if (owner[0] < 0) sl = 0;
        if (sl < 0) return hipSuccess;                  // both slots hold segments still to be swept
        hipError_t e;
        if (released[(size_t)sl] && (e = hipStreamWaitEvent(c->stream2, slot_free[(size_t)sl], 0)) != hipSuccess) return e;
        const int64_t k = k_rec--;
        seg_slot[(size_t)k] = sl;
        owner[(size_t)sl] = k;
        rec_of[(size_t)k] = n_rec;
        const int64_t len = std::min(K, T - k * K);
        if ((e = hipEventRecord(rec_ev[(size_t)(2 * n_rec)], c->stream2)) != hipSuccess) return e;
        for (int64_t j = 1; j < len; ++j)
            if ((e = launch_tick(c, seg_state(k, j - 1), seg_state(k, j), c->stream2)) != hipSuccess) return e;
        if ((e = hipEventRecord(rec_ev[(size_t)(2 * n_rec + 1)], c->stream2)) != hipSuccess) return e;
        ++n_rec;
        return hipSuccess;
    };
    CH_HIP(c, try_recompute());
    for (int64_t k = nseg - 1; k >= 0; --k) {
        const int64_t t0 = k * K, len = std::min(K, T - t0);
        if (resident_slot(k) < 0) {
            // every later segment is swept (enqueued), so a slot is free for k
            while (rec_of[(size_t)k] < 0) {
                const int64_t before = k_rec;
                CH_HIP(c, try_recompute());
                if (k_rec == before)
                    return cfail(c, BF_ERR_STATE, "[LBM Error] bf_channel_gradient: no free trajectory slot (internal)");
            }
            CH_HIP(c, hipStreamWaitEvent(c->stream, rec_ev[(size_t)(2 * rec_of[(size_t)k] + 1)], 0));
        }
        for (int64_t j = len - 1; j >= 0; --j) {
            CH_HIP(c, launch_adjoint(c, plain, seg_state(k, j), nu[cur], nu[cur ^ 1], dgseg, du_rows));
            cur ^= 1;
            plain = false;
        }
        hipLaunchKernelGGL(k_fold, dim3((unsigned)((c->N + 255) / 256)), dim3(256), 0, c->stream, dgseg, dg64, c->N);
        CH_HIP(c, hipGetLastError());
        const int64_t sl = seg_slot[(size_t)k];
        if (sl >= 0) {
            CH_HIP(c, hipEventRecord(slot_free[(size_t)sl], c->stream));
            released[(size_t)sl] = 1;
            owner[(size_t)sl] = -1;
        }
        CH_HIP(c, try_recompute());
    }
    CH_HIP(c, hipEventRecord(c->ev[3], c->stream));
    std::vector<double> dg(c->N), du((size_t)Y * Z);
    double j = 0.0;
    CH_HIP(c, hipMemcpyAsync(&j, c->scalar, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CH_HIP(c, hipMemcpyAsync(dg.data(), dg64, sizeof(double) * c->N, hipMemcpyDeviceToHost, c->stream));
    CH_HIP(c, hipMemcpyAsync(du.data(), du_rows, sizeof(double) * Y * Z, hipMemcpyDeviceToHost, c->stream));
    CH_HIP(c, hipStreamSynchronize(c->stream));
    double recompute_ms = 0.0;
    for (int64_t e = 0; e < n_rec; ++e) {
        float ms = 0.0f;
        CH_HIP(c, hipEventElapsedTime(&ms, rec_ev[(size_t)(2 * e)], rec_ev[(size_t)(2 * e + 1)]));
        recompute_ms += ms;
    }
    float fwd_ms = 0.0f, total_ms = 0.0f;
    CH_HIP(c, hipEventElapsedTime(&fwd_ms, c->ev[0], c->ev[1]));
    CH_HIP(c, hipEventElapsedTime(&total_ms, c->ev[0], c->ev[3]));
    c->timing_ms[0] = fwd_ms;
    c->timing_ms[1] = recompute_ms;
    c->timing_ms[2] = total_ms;
    for (int64_t k = 0; k < c->N; ++k) dJ_dgamma[k] = (float)dg[k];
    double s = 0.0;
    for (double v : du) s += v;
    *dJ_duin = s;
    *J = j;
    c->last = final_state;   // a ping-pong buffer, untouched by the adjoint sweep
    return BF_OK;
}

bf_status bf_channel_read_moments(bf_channel *c, float *out, size_t bytes) {
    if (!c || !out) return BF_ERR_INVALID;
    if (bytes != sizeof(float) * 4 * (size_t)c->N)
        return cfail(c, BF_ERR_INVALID, "[LBM Error] bf_channel_read_moments: %zu bytes given", bytes);
    if (!c->last) return cfail(c, BF_ERR_STATE, "[LBM Error] bf_channel_read_moments: run bf_channel_forward first");
    CH_HIP(c, hipSetDevice(c->cfg.device));
    dim3 g, b;
    shape(c, g, b);
    const ChArgs a = args_of(c);
    CH_HIP(c, hipMemsetAsync(c->moments, 0, sizeof(float4) * c->N, c->stream));
    if (c->q == 9) hipLaunchKernelGGL(k_ch_moments<D2Q9>, g, b, 0, c->stream, a, c->last, c->moments);
    else hipLaunchKernelGGL(k_ch_moments<D3Q19>, g, b, 0, c->stream, a, c->last, c->moments);
    CH_HIP(c, hipGetLastError());
    CH_HIP(c, hipMemcpyAsync(out, c->moments, bytes, hipMemcpyDeviceToHost, c->stream));
    CH_HIP(c, hipStreamSynchronize(c->stream));
    return BF_OK;
}

bf_status bf_channel_timing(const bf_channel *c, double out[3]) {
    if (!c || !out) return BF_ERR_INVALID;
    out[0] = c->timing_ms[0];
    out[1] = c->timing_ms[1];
    out[2] = c->timing_ms[2];
    return BF_OK;
}

}  // extern "C"
