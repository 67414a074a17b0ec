// bf_step.hip — fused in-place lattice Boltzmann tick for gfx950.
//
// One thread per cell, populations in slot arrays (slot i of cell c at element
// pop_index(c) of slot i: slabs of 8192 cells hold a group of slots one after
// the other, see bf_kernels.h), updated in place with the AA ("esoteric
// pull") pattern of BoltzmannFluids/Source/GLSL/LBM/collide.comp:244-369:
//   even tick: read slot i at x - c_i, write slot opp(i) at x + c_i;
//   odd tick:  read slot opp(i) at x,   write slot i at x.
// Each cell reads exactly the slots it writes, so the update is race free
// without a second lattice.  The per-cell work restates collide.comp:676-793:
// density floor, constant / per-cell force, Boussinesq buoyancy, Shan-Chen
// pseudopotential force, half-force velocity shift with clamps, BGK with the
// Guo source, thermal BGK, macroscopic save.  Halfway bounce-back with the
// moving-wall term (stream.comp:487-496,576-580) and thermal anti-bounce-back
// (stream_thermal.comp:520-524) are applied by redirecting the even-tick
// accesses of directions whose neighbour is solid to the cell's own opposite
// slot; odd ticks stay local.
//
// This file is compiled once per velocity set with -DBF_VS=<VelocitySetId>.
// Floating-point contraction is disabled at build time so that the FP32 path
// reproduces the CPU restatement's operation order bit for bit.
#include "bf_kernels.h"
#include "bf_lattice.h"

#ifndef BF_VS
#error "compile with -DBF_VS=<velocity set id>"
#endif

namespace bf {
namespace {

constexpr int kSrcOff = 0, kSrcOn = 1, kSrcRuntime = 2;  // Guo source policy of a kernel (SRC)

constexpr int kMpBX = 64;   // multiphase tile: 64 x 4 cells, one z-plane
constexpr int kMpBY = 4;
constexpr int kTileX = kMpBX + 2, kTileY = kMpBY + 2;

// FP32 -> BF16, round to nearest even: one v_cvt_pk_bf16_f32 on gfx950
// (the integer form, add 0x7fff + lsb and keep the high half, is two
// instructions per store)
__device__ inline uint16_t bf16_rne(float v) { return __builtin_bit_cast(uint16_t, (__bf16)v); }

// Streaming hints for the population traffic: every slot is read once and
// written once per tick, so the accesses may bypass cache allocation.
#ifndef BF_NT_LOAD
#define BF_NT_LOAD 1
#endif
#ifndef BF_NT_STORE
#define BF_NT_STORE 1
#endif

// NT selects the streaming form; the even tick of a chunked tick pair uses
// ordinary accesses instead, so that the lines it writes stay in the last-level
// cache for the odd tick of the same chunk (see k_pair).
template <bool NT, class T>
__device__ inline T ld(const T *p) {
    if constexpr (NT && BF_NT_LOAD) return __builtin_nontemporal_load(p);
    else return *p;
}
template <bool NT, class T>
__device__ inline void st(T *p, T v) {
    if constexpr (NT && BF_NT_STORE) __builtin_nontemporal_store(v, p);
    else *p = v;
}

// Element `idx` of a slot whose base is wave-uniform, addressed as base + a
// 32-bit byte offset: the loads and stores take the scalar-base form and cells
// that share an index share one offset register.  The host keeps every slot
// group below 4 GiB, so every element's byte offset from its slot base is 32-bit.
// The byte offset passes through an empty asm: the compiler otherwise
// rewrites zext(idx * size) as zext(idx) << log2(size) where idx comes from a
// select, a 64-bit offset that the scalar-base form cannot take.
template <class T>
__device__ inline const T *at(const T *p, uint32_t idx) {
    uint32_t off = idx * (uint32_t)sizeof(T);
    asm("" : "+v"(off));
    return reinterpret_cast<const T *>(reinterpret_cast<const char *>(p) + off);
}
template <class T>
__device__ inline T *at(T *p, uint32_t idx) {
    uint32_t off = idx * (uint32_t)sizeof(T);
    asm("" : "+v"(off));
    return reinterpret_cast<T *>(reinterpret_cast<char *>(p) + off);
}
// Byte offset of element idx with its known bits hidden.  Where one index
// serves every slot (the odd tick), the compiler otherwise rewrites
// zext(idx * size) as zext(idx) << log2(size), a 64-bit offset that the
// scalar-base form cannot take, and builds a 64-bit address per access.
__device__ inline uint32_t opaque_offset(uint32_t idx, uint32_t size) {
    uint32_t off = idx * size;
    asm("" : "+v"(off));
    return off;
}
// Element of cell c in every slot array: c + (c >> slab_log) * slab_skip.
// (c >> slab_log) < 2^18 and slab_skip < 2^24, so one 24-bit multiply-add.
__device__ inline uint32_t pop_index(uint32_t c, uint32_t slab_log, uint32_t slab_skip) {
    return c + __umul24(c >> slab_log, slab_skip);
}
template <class T>
__device__ inline T *at_bytes(T *p, uint32_t off) {
    return reinterpret_cast<T *>(reinterpret_cast<char *>(p) + off);
}
template <class T>
__device__ inline const T *at_bytes(const T *p, uint32_t off) {
    return reinterpret_cast<const T *>(reinterpret_cast<const char *>(p) + off);
}

// Storage policies.  Every storage type holds h = f - w (the population minus
// its rest-state weight), and the tick computes on h: at low Mach number f is
// w plus a small deviation, so the deviation keeps all 24 bits of an FP32
// mantissa (and all 11 or 8 bits of FP16 or BF16) instead of losing the bits
// that w occupies.  `raw` issues the load, `cvt` widens it to FP32; the tick
// issues all raw loads of a cell before it converts any, so the loads stay in
// flight together.
//
// Opposite slots i and opp(i) (adjacent: odd i pairs with i + 1, always in
// one slot group) are addressed as a pair from the lower slot's base: the
// higher slot's element of a cell lies pair_step elements after the lower
// one's.  A per-lane select of the element then replaces a per-lane select
// between two base pointers.  Two 16-bit slots stay below 4 GiB for every grid
// the solver accepts, so their byte offset is 32-bit (scalar-base form); FP32
// pairs use a 64-bit address (an FP32 lattice too large for slabs keeps plain
// slot arrays, whose pairs span more than 4 GiB).
template <int ST, bool NTL, bool NTS> struct Store;
template <bool NTL, bool NTS> struct Store<ST_FP32, NTL, NTS> {
    using T = float;
    __device__ static T raw_pair(const T *p, uint32_t e) { return ld<NTL>(p + (size_t)e); }
    __device__ static void store_pair(T *p, uint32_t e, float h) { st<NTS>(p + (size_t)e, h); }
    __device__ static T raw(const T *p, uint32_t idx) { return ld<NTL>(at(p, idx)); }
    __device__ static float cvt(T v) { return v; }
    __device__ static void store(T *p, uint32_t idx, float h) { st<NTS>(at(p, idx), h); }
    __device__ static T raw_b(const T *p, uint32_t off) { return ld<NTL>(at_bytes(p, off)); }
    __device__ static void store_b(T *p, uint32_t off, float h) { st<NTS>(at_bytes(p, off), h); }
};
template <bool NTL, bool NTS> struct Store<ST_FP16, NTL, NTS> {
    using T = _Float16;
    __device__ static T raw_pair(const T *p, uint32_t e) { return ld<NTL>(at(p, e)); }
    __device__ static void store_pair(T *p, uint32_t e, float h) { st<NTS>(at(p, e), (_Float16)h); }
    __device__ static T raw(const T *p, uint32_t idx) { return ld<NTL>(at(p, idx)); }
    __device__ static float cvt(T v) { return (float)v; }
    __device__ static void store(T *p, uint32_t idx, float h) { st<NTS>(at(p, idx), (_Float16)h); }
    __device__ static T raw_b(const T *p, uint32_t off) { return ld<NTL>(at_bytes(p, off)); }
    __device__ static void store_b(T *p, uint32_t off, float h) { st<NTS>(at_bytes(p, off), (_Float16)h); }
};
template <bool NTL, bool NTS> struct Store<ST_BF16, NTL, NTS> {
    using T = uint16_t;
    __device__ static T raw_pair(const T *p, uint32_t e) { return ld<NTL>(at(p, e)); }
    __device__ static void store_pair(T *p, uint32_t e, float h) { st<NTS>(at(p, e), bf16_rne(h)); }
    __device__ static T raw(const T *p, uint32_t idx) { return ld<NTL>(at(p, idx)); }
    __device__ static float cvt(T v) { return __uint_as_float((uint32_t)v << 16); }
    __device__ static void store(T *p, uint32_t idx, float h) { st<NTS>(at(p, idx), bf16_rne(h)); }
    __device__ static T raw_b(const T *p, uint32_t off) { return ld<NTL>(at_bytes(p, off)); }
    __device__ static void store_b(T *p, uint32_t off, float h) { st<NTS>(at_bytes(p, off), bf16_rne(h)); }
};

// c_i . v summed over the non-zero components of c_i in x, y, z order.  Equal
// to the full three-term dot product (the skipped terms are signed zeros).
template <int VS>
__device__ inline float dotc(int i, float vx, float vy, float vz) {
    using L = Lattice<VS>;
    float r = 0.0f;
    bool first = true;
    if (L::c[i][0] != 0) { r = L::c[i][0] > 0 ? vx : -vx; first = false; }
    if (L::c[i][1] != 0) {
        float t = L::c[i][1] > 0 ? vy : -vy;
        r = first ? t : r + t;
        first = false;
    }
    if (L::c[i][2] != 0) {
        float t = L::c[i][2] > 0 ? vz : -vz;
        r = first ? t : r + t;
    }
    return r;
}

// Directions with a non-zero x component: their even-tick accesses are shifted
// by one cell along x and are realigned through LDS (see k_step).
template <int VS>
__host__ __device__ constexpr int x_dir_count() {
    int n = 0;
    for (int i = 0; i < Lattice<VS>::Q; ++i) n += Lattice<VS>::c[i][0] != 0;
    return n;
}
template <int VS>
__host__ __device__ constexpr int x_dir_index(int i) {
    int n = 0;
    for (int j = 0; j < i; ++j) n += Lattice<VS>::c[j][0] != 0;
    return n;
}
// Marks an LDS entry that no cell wrote (a signalling-NaN pattern that arithmetic never produces).
constexpr uint32_t kNoValue = 0x7fa5a5a5u;

// Occupancy: the plain and wall kernels are held to 8 waves per SIMD (the
// register budget then covers 8 resident waves; the compiler spills a few
// scalar registers to VGPR lanes instead of lowering occupancy): porous duct
// FP16 +4 %, FP32 +2.4 %, TGV +1-2 %.  The multiphase and thermal kernels are
// left to the compiler: held to 7 or 8 waves they spill and run at half speed.
#ifndef BF_WAVES_PER_EU
#define BF_WAVES_PER_EU 8
#endif
#ifndef BF_WAVES_PER_EU_MPTH
#define BF_WAVES_PER_EU_MPTH 1  // 1: the compiler's choice
#endif
#define BF_OCCUPANCY __attribute__((amdgpu_waves_per_eu((MP || TH) ? BF_WAVES_PER_EU_MPTH : BF_WAVES_PER_EU)))

// v modulo n in [0, n).  Tile coordinates are inside the grid except in the
// halo at the grid's edges, so the integer division only runs on those lanes.
__device__ inline int wrap_index(int v, int n) {
    return (unsigned)v < (unsigned)n ? v : ((v % n) + n) % n;
}

// Where a block works: its first cell is (xb, yb, z) (cells xb + k a.bdx
// + threadIdx.x along x, rows yb + threadIdx.y), and its FP64 partial sums go
// to entry pb.
struct BlockAt {
    int xb, yb, z;
    size_t pb;
};
__device__ inline BlockAt grid_block(const StepArgs &a, int cpt, int z) {
    return BlockAt{(int)blockIdx.x * cpt * a.bdx, (int)blockIdx.y * a.bdy, z,
                   blockIdx.x + (size_t)gridDim.x * (blockIdx.y + (size_t)gridDim.y * (size_t)z)};
}

// One tick of the cells of the block at `blk` (the body of k_step, of both
// halves of k_pair and of k_list).  NTL / NTS: streaming (nontemporal)
// population loads / stores.
template <int VS, int ST, bool ODD, bool WALLS, bool MP, bool TH, int CPT, bool AL, bool PER, int SRC, bool NTL,
          bool NTS, bool DEN = false>
__device__ __forceinline__ void step_body(const StepArgs &a, double *__restrict__ partial, const BlockAt blk) {
    const int z = blk.z;
    // PER: the grid is closed (every axis periodic, or solid boundary planes
    // on the others): every neighbour of a cell that updates exists, so the
    // open-boundary branches vanish at compile time.  Without walls every
    // cell of the grid updates (kEvery).
    // CPT cells per thread along x, blockDim.x apart: all their loads are issued
    // before any of them is updated.
    // AL (even ticks, blocks spanning whole x-rows): the accesses of directions
    // with c_x != 0 are made aligned.  Each cell loads slot i of the row it pulls
    // from at its own x into LDS and reads back the entry at x - c_x; on the way
    // out it writes its value to the LDS entry at x + c_x and the cell owning
    // that column stores it.  Every global access of the tick then covers whole
    // cache lines (the periodic x wrap stays inside the block).
    static_assert(CPT == 1 || (!MP && !TH), "multi-cell threads only for the plain and wall kernels");
    static_assert(!(AL && (ODD || MP)), "aligned rows only for the even tick of non-multiphase kernels");
    constexpr bool kEvery = PER && !WALLS;
    // SRC: the Guo source (a force, Shan-Chen or a thermal lattice) is off
    // (kSrcOff), on (kSrcOn), or decided per launch (kSrcRuntime).  The
    // launch's flag tested per direction is compiled as the source of every
    // direction computed and selected away: 3-4 vector instructions per
    // direction without a force.
    static_assert(SRC == kSrcOn || !(MP || TH), "the Shan-Chen and thermal kernels always have the source on");
    // DEN: the synchronous Shan-Chen density pre-pass.  The tick's reads (the
    // same addresses, bounce-back, moving-wall and open-boundary rules, so the
    // same density bit for bit), then psi of the floored density is written
    // and the cell is done: no tile, force, collision or thermal lattice.
    static_assert(!DEN || (MP && !TH && !AL && CPT == 1), "the density pre-pass is a multiphase read pass");
    extern __shared__ float lds_rows[];
    using L = Lattice<VS>;
    using S = Store<ST, NTL, NTS>;
    using T = typename S::T;
    constexpr int Q = L::Q;
    constexpr int QT = L::QT;
    constexpr bool D3 = L::DIM == 3;

    // slot bases come from the argument block (wave-uniform, scalar registers)
#define SLOT(i) static_cast<T *>(a.slot[i])
    // element of cell c of slot j in the pair (j, opp(j)) based at the lower
    // slot, for the lattice (c from POP) and the thermal lattice (plain cells)
    const uint32_t pair_step = a.pair_step, gstride32 = (uint32_t)a.stride;
#define PAIR_E(j, c) ((uint32_t)(c) + ((j) > opposite(j) ? pair_step : 0u))
#define GPAIR_E(j, c) ((uint32_t)(c) + ((j) > opposite(j) ? gstride32 : 0u))
#define POP(c) pop_index((uint32_t)(c), a.slab_log, a.slab_skip)
#define PAIR_BASE(j) SLOT((j) < opposite(j) ? (j) : opposite(j))
    const int X = a.X, Y = a.Y, Z = a.Z;
    const int tx = threadIdx.x, ty = threadIdx.y;
    // PER: lanes past the grid address the grid's last row / column (they
    // load in range and store nothing), so no access needs a per-lane guard
    const int y_in = blk.yb + ty;
    const int y = PER ? min(y_in, Y - 1) : y_in;

    // Neighbour coordinates: index 0 = -1, 1 = 0, 2 = +1 along each axis.
    // Written as selects: a branch here would split the block that issues the
    // loads and cost the scalar-base addressing of every access.
    const bool wy0 = y - 1 < 0, wy2 = y + 1 >= Y;
    const int ys[3] = {wy0 ? y - 1 + Y : y - 1, y, wy2 ? y + 1 - Y : y + 1};
    // (the fully periodic kernel runs only on grids periodic on every axis)
    const bool vy[3] = {PER || !wy0 || a.periodic_y != 0, true, PER || !wy2 || a.periodic_y != 0};
    const bool wz0 = D3 && z - 1 < 0, wz2 = D3 && z + 1 >= Z;
    const int zs[3] = {D3 ? (wz0 ? z - 1 + Z : z - 1) : z, z, D3 ? (wz2 ? z + 1 - Z : z + 1) : z};
    const bool vz[3] = {PER || !wz0 || a.periodic_z != 0, true, PER || !wz2 || a.periodic_z != 0};
    int ry[3] = {ys[0] * X, ys[1] * X, ys[2] * X};
    const int XY = X * Y;
    int pz[3] = {zs[0] * XY, zs[1] * XY, zs[2] * XY};

    // neighbour (dx, dy, dz) of cell k inside the domain (indices 0/1/2 = -1/0/+1)
#define OK(k, dx, dy, dz) (PER || (vx[k][dx] && vy[dy] && vz[dz]))
    int xs[CPT][3];
    bool vx[CPT][3];
    int self[CPT];
    uint32_t pself[CPT];  // the cell's element in the slot arrays
    bool active[CPT];
    bool in_bounds[CPT];
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        const int x_in = blk.xb + k * a.bdx + tx;
        const int x = PER ? min(x_in, X - 1) : x_in;
        const bool w0 = x - 1 < 0, w2 = x + 1 >= X;
        xs[k][0] = w0 ? x - 1 + X : x - 1; xs[k][1] = x; xs[k][2] = w2 ? x + 1 - X : x + 1;
        vx[k][0] = PER || !w0 || a.periodic_x != 0; vx[k][1] = true; vx[k][2] = PER || !w2 || a.periodic_x != 0;
        self[k] = x + ry[1] + pz[1];
        active[k] = x_in < X && y_in < Y;
        in_bounds[k] = active[k];   // in the grid (active drops solid cells below)
    }
    // AL: every population access of the tick is at the cell's own x in one of
    // the 9 neighbouring rows.  With 16-bit storage their elements are computed
    // once and made opaque, so that the compiler keeps them instead of
    // recomputing the slab index at each of the ~4 accesses per row (-130 VALU
    // instructions, FP16 7 % faster).  The FP32 kernels compute them at each
    // access: kept, the FP32 wall kernel runs 25 % slower and the periodic
    // one 0.7 % slower.
    constexpr bool kKeepRows = AL && ST != ST_FP32;
    uint32_t prow[CPT][3][3];
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        if constexpr (kKeepRows) {
#pragma unroll
            for (int dy = 0; dy < 3; ++dy)
#pragma unroll
                for (int dz = 0; dz < 3; ++dz) {
                    if (!D3 && dz != 1) continue;
                    uint32_t e = POP(xs[k][1] + ry[dy] + pz[dz]);
                    asm("" : "+v"(e));
                    prow[k][dy][dz] = e;
                }
            pself[k] = prow[k][1][1];
        } else {
            pself[k] = POP(self[k]);
        }
    }
    // element of the cell at the own x of row (dy, dz), and of the neighbour
    // (dx, dy, dz) of cell k (indices 0/1/2 = -1/0/+1)
#define ROW(k, dy, dz) (kKeepRows ? prow[k][dy][dz] : POP(xs[k][1] + ry[dy] + pz[dz]))
#define NB(k, dx, dy, dz) ((dx) == 1 ? ROW(k, dy, dz) : POP(xs[k][dx] + ry[dy] + pz[dz]))

    // Shan-Chen: stage psi of the previous tick for the block tile + halo.
    __shared__ float tile[(MP && !DEN) ? kTileX * kTileY * 3 : 1];
    if constexpr (MP && !DEN) {
        const int tid = ty * kMpBX + tx;
        const int nplanes = D3 ? 3 : 1;
        const int x0 = blk.xb - 1, y0 = blk.yb - 1;
        for (int k = tid; k < kTileX * kTileY * nplanes; k += kMpBX * kMpBY) {
            const int lx = k % kTileX, ly = (k / kTileX) % kTileY, lz = k / (kTileX * kTileY);
            const int gx = wrap_index(x0 + lx, X);
            const int gy = wrap_index(y0 + ly, Y);
            const int gz = D3 ? wrap_index(z - 1 + lz, Z) : z;
            tile[k] = a.psi_prev[gx + gy * X + gz * XY];
        }
        __syncthreads();
    }
    // D3Q27: the weights factor as w1(c_x) w1(c_y) w1(c_z), w1(0) = 2/3,
    // w1(+-1) = 1/6, so sum_i w_i psi(x + c_i) c_i separates into smoothings S
    // and differences D along the axes: x (1/6) D_x S_y S_z psi, y (1/6) S_x
    // D_y S_z psi, z (1/6) S_x S_y D_z psi.  S_y S_z psi, D_y S_z psi and S_y
    // D_z psi are formed once per block at the 66 x 4 tile columns of its rows;
    // a cell then reads 8 of them instead of 26 tile entries.
#ifndef BF_SC_SEPARABLE
#define BF_SC_SEPARABLE 1  // build-time switch of the separable D3Q27 pseudopotential gradient
#endif
    constexpr bool kSep = BF_SC_SEPARABLE && MP && !DEN && VS == D3Q27;
    constexpr int kSepN = kTileX * kMpBY;
    constexpr float k23 = float(2.0 / 3.0), k16 = float(1.0 / 6.0);
    __shared__ float sep[kSep ? 3 * kSepN : 1];
    if constexpr (kSep) {
        constexpr int plane = kTileX * kTileY;
        for (int k = ty * kMpBX + tx; k < kSepN; k += kMpBX * kMpBY) {
            const int lx = k % kTileX, ly = k / kTileX + 1;
            float Pz[3], Dz[3];
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const int t = lx + (ly + j - 1) * kTileX;
                const float pm = tile[t], p0 = tile[t + plane], pp = tile[t + 2 * plane];
                Pz[j] = __builtin_fmaf(p0, k23, (pm + pp) * k16);
                Dz[j] = pp - pm;
            }
            sep[k] = __builtin_fmaf(Pz[1], k23, (Pz[0] + Pz[2]) * k16);
            sep[kSepN + k] = Pz[2] - Pz[0];
            sep[2 * kSepN + k] = __builtin_fmaf(Dz[1], k23, (Dz[0] + Dz[2]) * k16);
        }
        __syncthreads();
    }

    // Solid cells skip (collide.comp:696-700).  Even ticks take the solid-neighbour
    // mask (bit 0 = self solid, bit j = neighbour at x + c_j solid), odd ticks
    // only need the cell's own id.
    uint32_t nm[CPT];
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        nm[k] = 0u;
        if (WALLS && active[k]) {
            if constexpr (ODD) {
                active[k] = a.ids[self[k]] == 0;
            } else {
                nm[k] = a.nmask[self[k]];
                active[k] = (nm[k] & 1u) == 0u;
            }
        }
    }
#define SOLID(k, j) (WALLS && !ODD && ((nm[k] >> (j)) & 1u))

    double s_rho = 0.0, s_mx = 0.0, s_my = 0.0, s_mz = 0.0, s_t = 0.0;

    // Plain periodic kernels: every lane loads in range (lanes past the grid
    // take its last row or column), so the reads, the row-buffer exchange and
    // the conversions run on every lane and only the stores test `active`.
    // Behind a branch the skipped path would need its registers set (zeros,
    // the open-boundary constants, copies): about 40 instructions per tick.
    // A lane past the grid in x writes the row-buffer entry of the grid's
    // last column, the value that column's own cell writes there.
#ifndef BF_ALL_LANES
#define BF_ALL_LANES 1
#endif
    constexpr bool kAllLanes = BF_ALL_LANES && PER && !WALLS && !MP && !TH;

    // ---- phase 1: AA reads (collide.comp:244-305) ----
    // 1b issues every population load of the tick (addresses from the
    // neighbour mask), then the ids of solid neighbours where wall terms need
    // them; 1c converts and applies the open-boundary and bounce-back rules.
    // No loaded value is used before 1c.
    T raw[CPT][Q];    // x-directions under AL: the value staged for the row buffer
    float graw[QT];
    float fi[CPT][Q];
    float gi[QT];
    // row buffer: [row of the block][x][x-direction], the directions of one
    // column at an odd stride (lanes of consecutive columns hit different
    // banks), so that the directions of a column are immediate offsets from
    // one address
    constexpr int kXdStride = x_dir_count<VS>() | 1;
    const int lds_base = ty * X;
#define LDS(i, col) lds_rows[(lds_base + (col)) * kXdStride + x_dir_index<VS>(i)]

    // 1b.  Cells that do not update (solid, or past the grid) load cell 0 of
    // the same slots: one shared line per wave, and no branch splits the
    // batch of loads (a branch would let the compiler consume the first loads
    // before issuing the rest).
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        if constexpr (AL) {
            // every cell of the block's rows loads, at its own x, slot i of the
            // row the cell at x + c_x pulls from
            const bool in_grid = PER || in_bounds[k];
#pragma unroll
            for (int i = 1; i < Q; ++i) {
                if (L::c[i][0] == 0) continue;
                const int dy = 1 - L::c[i][1], dz = 1 - L::c[i][2];
                const uint32_t idx = !in_grid ? 0u : (PER || (vy[dy] && vz[dz])) ? ROW(k, dy, dz) : pself[k];
                raw[k][i] = S::raw(SLOT(i), idx);
            }
        }
        // Closed 16-bit wall kernels: every lane's addresses lie in the grid
        // (a solid cell's neighbours included), so solid cells load their own
        // neighbourhood, lines their fluid neighbours read as well, and no
        // select redirects them to cell 0.
#ifndef BF_SOLID_LOADS
#define BF_SOLID_LOADS 1
#endif
        constexpr bool kAnyAddr = BF_SOLID_LOADS && PER && WALLS && ST != ST_FP32 && !MP && !TH;
        const bool act = kEvery || kAnyAddr || active[k];
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            const int o = opposite(i);
            if constexpr (ODD) {
                // every cell of the grid loads its own slots: the addresses do
                // not wait for the boundary id (a solid cell shares its lines
                // with the fluid cells of its wave)
                raw[k][i] = S::raw_b(SLOT(o), opaque_offset((PER || in_bounds[k]) ? pself[k] : 0u, sizeof(T)));
                continue;
            }
            if (i == 0) {
                raw[k][i] = S::raw(SLOT(0), act ? pself[k] : 0u);
                continue;
            }
            const int dx = 1 - L::c[i][0], dy = 1 - L::c[i][1], dz = 1 - L::c[i][2];
            const bool valid = vx[k][dx] && vy[dy] && vz[dz];
            const bool redir = a.halfway && SOLID(k, o);
            if (AL && L::c[i][0] != 0) {
                // own opposite slot for a wall redirect: loaded below
            } else if constexpr (WALLS) {
                // bounce-back reads slot o at the cell itself, otherwise slot i
                // at the source: one pair, one base.  (Both candidates are
                // computed first so that the choice compiles to selects.)
                const uint32_t e_back = PAIR_E(o, pself[k]);
                const uint32_t e_src = PAIR_E(i, valid ? NB(k, dx, dy, dz) : pself[k]);
                const uint32_t e_act = redir ? e_back : e_src;
                raw[k][i] = S::raw_pair(PAIR_BASE(i), act ? e_act : 0u);
            } else {
                const uint32_t idx = valid ? NB(k, dx, dy, dz) : pself[k];
                raw[k][i] = S::raw(SLOT(i), act ? idx : 0u);
            }
        }
        if constexpr (TH) {
#pragma unroll
            for (int q = 0; q < QT; ++q) {
                const int o = opposite(q);
                if (ODD || q == 0) {
                    if constexpr (ODD) graw[q] = *at_bytes(a.gslot[o], opaque_offset(in_bounds[k] ? self[k] : 0, 4));
                    else graw[q] = *at(a.gslot[0], act ? self[k] : 0);
                    continue;
                }
                const int dx = 1 - L::c[q][0], dy = 1 - L::c[q][1], dz = 1 - L::c[q][2];
                const bool pull = OK(k, dx, dy, dz) && !SOLID(k, o);
                const uint32_t e_src = GPAIR_E(q, xs[k][dx] + ry[dy] + pz[dz]);
                const uint32_t e_back = GPAIR_E(o, self[k]);
                const uint32_t e_act = pull ? e_src : e_back;
                graw[q] = a.gslot[q < o ? q : o][act ? e_act : 0u];
            }
        }
    }
    // object of the solid neighbour at x + c_j (looked up only where a wall
    // term can be non-zero: a moving wall or a thermal wall)
#define WALL(k, j) a.objects[a.ids[xs[k][1 + L::c[j][0]] + ry[1 + L::c[j][1]] + pz[1 + L::c[j][2]]]]
    // 1c
    // Every load of 1b stays ahead of every use in 1c where that measured
    // faster (device time per tick, interleaved builds): the FP32 even tick
    // without walls (left to itself the scheduler splits its loads into
    // batches with waits between them: tick pairs 5 % slower), and the 16-bit
    // wall kernels (porous duct 1-1.5 % faster).  Not for the FP32 wall
    // kernels (+2 %) or the Shan-Chen kernels (+1 %).
#ifndef BF_SCHED_SPLIT
#define BF_SCHED_SPLIT 1  // build-time switch of the scheduling barrier
#endif
    constexpr bool kSchedSplit = BF_SCHED_SPLIT && !MP && !TH &&
                                 ((ST == ST_FP32 && AL && !WALLS) || (ST != ST_FP32 && WALLS));
    if constexpr (kSchedSplit) __builtin_amdgcn_sched_barrier(0);
    if constexpr (AL) {
#pragma unroll
        for (int k = 0; k < CPT; ++k) {
            if (!kAllLanes && !in_bounds[k]) continue;
#pragma unroll
            for (int i = 1; i < Q; ++i) {
                if (L::c[i][0] == 0) continue;
                LDS(i, xs[k][1]) = S::cvt(raw[k][i]);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        if (!kAllLanes && !active[k]) continue;
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            const float w = L::w[i];
            const int o = opposite(i);
            if (ODD || i == 0) {
                fi[k][i] = S::cvt(raw[k][i]);
                continue;
            }
            const int dx = 1 - L::c[i][0], dy = 1 - L::c[i][1], dz = 1 - L::c[i][2];
            const bool valid = vx[k][dx] && vy[dy] && vz[dz];
            const bool redir = a.halfway && SOLID(k, o);
            if (AL && L::c[i][0] != 0) {
                // the open-boundary value, replaced after the row exchange
                // unless the source is outside the domain
                fi[k][i] = 0.01f - w;
                continue;
            }
            float v = S::cvt(raw[k][i]);
            if (a.wall_ids && redir) {
                // moving wall (Ladd): + 6 w rho_w (c_i . u_w), rho_w = 1; a wall at
                // rest adds exactly zero
                const ObjectEntry ob = WALL(k, o);
                v = __builtin_fmaf(6.0f * w, dotc<BF_VS>(i, ob.ux, ob.uy, ob.uz), v);
            }
            fi[k][i] = valid ? v : 0.01f - w;  // open boundary read f = 0.01 (collide.comp:297-298)
        }
        if constexpr (TH) {
#pragma unroll
            for (int q = 0; q < QT; ++q) {
                gi[q] = graw[q];
                if (ODD || q == 0) continue;
                const int o = opposite(q);
                if (SOLID(k, o)) gi[q] = __builtin_fmaf(2.0f * L::wt[q], WALL(k, o).temperature, -graw[q]);
            }
        }
    }
    if constexpr (AL) {
        __syncthreads();
        // pull the x-shifted values from the row buffer
#pragma unroll
        for (int k = 0; k < CPT; ++k) {
            if (!kAllLanes && !active[k]) continue;
#pragma unroll
            for (int i = 1; i < Q; ++i) {
                if (L::c[i][0] == 0) continue;
                const int dx = 1 - L::c[i][0], dy = 1 - L::c[i][1], dz = 1 - L::c[i][2];
                const bool take = OK(k, dx, dy, dz) && !(a.halfway && SOLID(k, opposite(i)));
                const float lv = LDS(i, xs[k][dx]);
                fi[k][i] = take ? lv : fi[k][i];
            }
        }
        __syncthreads();
        // Each row-buffer entry (i, col) has one writer, the cell at col - c_x
        // (periodic in x within the row): cells that update write theirs in
        // phase 2 (empty where the write is dropped or bounced back), the
        // other cells of the grid mark theirs empty here.
#pragma unroll
        for (int k = 0; k < CPT; ++k) {
            if (!in_bounds[k] || active[k]) continue;
#pragma unroll
            for (int i = 1; i < Q; ++i) {
                if (L::c[i][0] == 0) continue;
                LDS(i, xs[k][1 + L::c[i][0]]) = __uint_as_float(kNoValue);
            }
        }
        if constexpr (WALLS) {
            // x-directions whose source is solid read the cell's own opposite
            // slot.  A wave with such a cell issues all these loads together
            // (cells without that redirect load cell 0), then converts them.
#pragma unroll
            for (int k = 0; k < CPT; ++k) {
                if (__builtin_amdgcn_ballot_w64(active[k] && a.halfway && (nm[k] >> 1) != 0u) == 0) continue;
                T rr[Q];
#pragma unroll
                for (int i = 1; i < Q; ++i) {
                    if (L::c[i][0] == 0) continue;
                    const int o = opposite(i);
                    rr[i] = S::raw(SLOT(o), (active[k] && a.halfway && SOLID(k, o)) ? pself[k] : 0u);
                }
#pragma unroll
                for (int i = 1; i < Q; ++i) {
                    if (L::c[i][0] == 0) continue;
                    const int o = opposite(i);
                    const bool red = active[k] && a.halfway && SOLID(k, o);
                    float v = S::cvt(rr[i]);
                    if (a.wall_ids && red) {
                        const ObjectEntry ob = WALL(k, o);
                        v = __builtin_fmaf(6.0f * L::w[i], dotc<BF_VS>(i, ob.ux, ob.uy, ob.uz), v);
                    }
                    fi[k][i] = red ? v : fi[k][i];
                }
            }
        }
    }

    // ---- phase 2: per-cell update and AA writes ----
    // 16-bit wall kernels: the solid-neighbour mask passes through an empty asm
    // here, so the bounce-back tests of the stores are made afresh from it
    // instead of keeping the load phase's 64-bit lane masks across the
    // collision (the kernel holds more wave-uniform values than the scalar
    // registers; 30-50 fewer of them spill to vector lanes).  Measured on the
    // 384^3 porous duct: FP16 tick pairs 1056 -> 1040 us; FP32 1.0 % slower,
    // so FP32 keeps the masks.
    if constexpr (WALLS && !ODD && ST != ST_FP32) {
#pragma unroll
        for (int k = 0; k < CPT; ++k) asm volatile("" : "+v"(nm[k]));
    }
    // Closed wall kernels: the neighbour coordinates pass through an empty asm
    // too, so that the stores compute their neighbours' elements afresh.
    // Without the neighbour checks the compiler otherwise keeps the load
    // phase's 18 elements live across the collision and spills to memory.
    if constexpr (WALLS && PER && !ODD) {
#pragma unroll
        for (int k = 0; k < CPT; ++k) asm volatile("" : "+v"(xs[k][0]), "+v"(xs[k][1]), "+v"(xs[k][2]));
        asm volatile("" : "+v"(ry[0]), "+v"(ry[1]), "+v"(ry[2]), "+v"(pz[0]), "+v"(pz[1]), "+v"(pz[2]));
    }
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        if (!active[k]) continue;
        // moments (collide.comp:383-396, 704-705): d = rho - 1 = sum of h_i;
        // sum c_i w_i = 0, so the momentum is the same sum over h as over f
        float d = fi[k][0];
#pragma unroll
        for (int i = 1; i < Q; ++i) d += fi[k][i];
        float rho = 1.0f + d;
        float mx = 0.0f, my = 0.0f, mz = 0.0f;
        {
            bool fx0 = true, fy0 = true, fz0 = true;
#pragma unroll
            for (int i = 0; i < Q; ++i) {
                const float v = fi[k][i];
                if (L::c[i][0] != 0) { float t = L::c[i][0] > 0 ? v : -v; mx = fx0 ? t : mx + t; fx0 = false; }
                if (L::c[i][1] != 0) { float t = L::c[i][1] > 0 ? v : -v; my = fy0 ? t : my + t; fy0 = false; }
                if (L::c[i][2] != 0) { float t = L::c[i][2] > 0 ? v : -v; mz = fz0 ? t : mz + t; fz0 = false; }
            }
        }
        if (partial) { s_rho += 1.0 + (double)d; s_mx += mx; s_my += my; s_mz += mz; }
        {
            const bool low = !(rho >= 0.01f);   // density floor 0.01 (also taken by NaN)
            rho = low ? 0.01f : rho;
            d = low ? -0.99f : d;
        }
        if constexpr (DEN) {
            if (a.write_psi) a.psi_cur[self[k]] = 1.0f - expf(-rho);  // as psi_self below
            continue;
        }

        // force (collide.comp:707-720)
        float Fx = 0.0f, Fy = 0.0f, Fz = 0.0f;
        if (a.forcing == 1) { Fx = a.fx; Fy = a.fy; Fz = a.fz; }
        else if (a.forcing == 2) { const float4 fv = a.force[self[k]]; Fx = fv.x; Fy = fv.y; Fz = fv.z; }

        float temp = 0.0f;
        if constexpr (TH) {
            temp = gi[0];
#pragma unroll
            for (int q = 1; q < QT; ++q) temp += gi[q];
            if (partial) s_t += temp;
            const float b = -a.beta * (temp - a.t_ref);
            const float bx = b * Fx, by = b * Fy, bz = b * Fz;
            Fx += bx; Fy += by; Fz += bz;
        }

        float psi_self = 0.0f;
        if constexpr (MP) {
            psi_self = 1.0f - expf(-rho);
            float ax = 0.0f, ay = 0.0f, az = 0.0f;
            // a neighbour outside the domain on an open axis takes psi_self
            // (collide.comp:552-556): such cells keep the direct sum
            const bool separable = kSep && vx[k][0] && vx[k][2] && vy[0] && vy[2] && vz[0] && vz[2];
            if (separable) {
                const int b = ty * kTileX + tx;  // tile column x - 1 of this cell's row
                ax = (sep[b + 2] - sep[b]) * k16;
                ay = __builtin_fmaf(sep[kSepN + b + 1], k23, (sep[kSepN + b] + sep[kSepN + b + 2]) * k16) * k16;
                az = __builtin_fmaf(sep[2 * kSepN + b + 1], k23,
                                    (sep[2 * kSepN + b] + sep[2 * kSepN + b + 2]) * k16) * k16;
            } else {
                bool ax0 = true, ay0 = true, az0 = true;
#pragma unroll
                for (int i = 1; i < Q; ++i) {
                    const int dx = 1 + L::c[i][0], dy = 1 + L::c[i][1], dz = 1 + L::c[i][2];
                    const int li = (tx + dx) + (ty + dy) * kTileX + (D3 ? dz : 0) * kTileX * kTileY;
                    const float pn = OK(k, dx, dy, dz) ? tile[li] : psi_self;
                    const float wp = L::w[i] * pn;
                    if (L::c[i][0] != 0) { float t = L::c[i][0] > 0 ? wp : -wp; ax = ax0 ? t : ax + t; ax0 = false; }
                    if (L::c[i][1] != 0) { float t = L::c[i][1] > 0 ? wp : -wp; ay = ay0 ? t : ay + t; ay0 = false; }
                    if (L::c[i][2] != 0) { float t = L::c[i][2] > 0 ? wp : -wp; az = az0 ? t : az + t; az0 = false; }
                }
            }
            const float kk = -psi_self * a.G;
            Fx += ax * kk; Fy += ay * kk; Fz += az * kk;
        }

        // velocity with half-force shift and clamps (collide.comp:722-725)
        const float inv_rho = 1.0f / rho, half_inv = 0.5f * inv_rho;
        float ux = __builtin_fmaf(Fx, half_inv, mx * inv_rho);
        float uy = __builtin_fmaf(Fy, half_inv, my * inv_rho);
        float uz = __builtin_fmaf(Fz, half_inv, mz * inv_rho);
        const float un = sqrtf(__builtin_fmaf(uz, uz, __builtin_fmaf(uy, uy, ux * ux)));
        const bool ext = un > a.uext;
        if (un > a.ulim) {
            const float sc = a.ulim / un;
            ux *= sc; uy *= sc; uz *= sc;
        }

        // BGK + Guo source, AA write (collide.comp:727-745, 307-369)
        if (a.update) {
            const float uu = __builtin_fmaf(uz, uz, __builtin_fmaf(uy, uy, ux * ux));
            const float fu = __builtin_fmaf(Fz, uz, __builtin_fmaf(Fy, uy, Fx * ux));
            const float c1 = -1.5f * uu;
            const float inv_tau = a.inv_tau;
            const float pre3 = 3.0f * a.guo_pre;
            // c_i . u and c_i . F of the last odd direction: the even direction
            // after it is its opposite and takes the negated products
            float cu_odd = 0.0f, cf_odd = 0.0f;
#pragma unroll
            for (int i = 0; i < Q; ++i) {
                const float w = L::w[i];
                const int o = opposite(i);
                static_assert(opposite(2) == 1 && opposite(Q - 1) == Q - 2, "opposite pairs are (2j - 1, 2j)");
                float cu = 0.0f;
                if (i != 0 && (i & 1)) { cu = dotc<BF_VS>(i, ux, uy, uz); cu_odd = cu; }
                else if (i != 0) cu = -cu_odd;
                // h_eq = w rho (1 + 3 cu + 9/2 cu^2 - 3/2 u^2) - w
                //      = (w rho)(cu (3 + 9/2 cu) - 3/2 u^2) + w d,
                // the rest population's mass part with w_0 + rest_fix, rounded
                // once (rest_fix * d is exact, w_0 d is not rounded before the sum)
                const float md = (i == 0) ? __builtin_fmaf(w, d, L::rest_fix * d) : w * d;
                const float e = __builtin_fmaf(w * rho, __builtin_fmaf(cu, __builtin_fmaf(4.5f, cu, 3.0f), c1), md);
                float v;
                // h - (h - h_eq)/tau + Guo source, or h_eq past the reset
                // threshold (a select, not a per-direction branch)
                v = __builtin_fmaf(e - fi[k][i], inv_tau, fi[k][i]);
                if (SRC == kSrcOn || (SRC == kSrcRuntime && a.source_on)) {
                    float cf = 0.0f;
                    if (i != 0 && (i & 1)) { cf = dotc<BF_VS>(i, Fx, Fy, Fz); cf_odd = cf; }
                    else if (i != 0) cf = -cf_odd;
                    // (1 - 1/(2 tau)) w (3 cF + 9 cF cu - 3 F.u), as
                    // 3 pre w (cF (3 cu + 1) - F.u): three fused multiply-adds
                    v = __builtin_fmaf(pre3 * w, __builtin_fmaf(cf, __builtin_fmaf(3.0f, cu, 1.0f), -fu), v);
                }
                v = ext ? e : v;
                if (ODD) {
                    S::store_b(SLOT(i), opaque_offset(pself[k], sizeof(T)), v);
                } else if (i == 0) {
                    S::store(SLOT(i), pself[k], v);
                } else if constexpr (WALLS) {
                    // the destination is slot o at x + c_i, or slot i at the cell
                    // itself when x + c_i is solid (the address this cell read
                    // f_o from); writes leaving the domain are dropped
                    // (collide.comp:358-359)
                    const int dx = 1 + L::c[i][0], dy = 1 + L::c[i][1], dz = 1 + L::c[i][2];
                    const bool ok = OK(k, dx, dy, dz);
                    const bool red = a.halfway && SOLID(k, i);
                    float vb = v;
                    if (a.wall_ids && red) {
                        const ObjectEntry ob = WALL(k, i);
                        vb = __builtin_fmaf(-6.0f * w, dotc<BF_VS>(i, ob.ux, ob.uy, ob.uz), v);
                    }
                    if (AL && L::c[i][0] != 0) {
                        // the row buffer entry has this cell as its only writer
                        LDS(i, xs[k][dx]) = ok && !red ? v : __uint_as_float(kNoValue);
                        if (ok && red) S::store(SLOT(i), pself[k], vb);
                    } else {
                        const uint32_t e_back = PAIR_E(i, pself[k]);
                        const uint32_t e_dst = PAIR_E(o, NB(k, dx, dy, dz));
                        const uint32_t e = red ? e_back : e_dst;
                        const float out = red ? vb : v;
                        if (ok) S::store_pair(PAIR_BASE(i), e, out);
                    }
                } else {
                    // no walls: slot o at x + c_i; writes leaving the domain are
                    // dropped (collide.comp:358-359)
                    const int dx = 1 + L::c[i][0], dy = 1 + L::c[i][1], dz = 1 + L::c[i][2];
                    const bool ok = OK(k, dx, dy, dz);
                    if (AL && L::c[i][0] != 0) LDS(i, xs[k][dx]) = ok ? v : __uint_as_float(kNoValue);
                    else if (ok) S::store(SLOT(o), NB(k, dx, dy, dz), v);
                }
            }
        }

        if constexpr (MP) {
            if (a.write_psi) a.psi_cur[self[k]] = psi_self;  // collide.comp:763-769 (psi of rho)
        }

        // thermal BGK (collide.comp:771-782)
        if constexpr (TH) {
            if (a.update) {
                const float uu = __builtin_fmaf(uz, uz, __builtin_fmaf(uy, uy, ux * ux));
                const float k1 = a.k1_t, k2 = a.k2_t;
                const float t1 = __builtin_fmaf(-0.5f * k1, uu, 1.0f);
                const float inv_tau_t = a.inv_tau_t;
#pragma unroll
                for (int q = 0; q < QT; ++q) {
                    const int o = opposite(q);
                    const float cu = (q == 0) ? 0.0f : dotc<BF_VS>(q, ux, uy, uz);
                    // wt T (1 + cu/cs2 + cu^2/(2 cs2^2) - u^2/(2 cs2))
                    const float wT = (q == 0) ? __builtin_fmaf(L::wt[0], temp, L::rest_fix_t * temp) : L::wt[q] * temp;
                    const float e = wT * __builtin_fmaf(cu, __builtin_fmaf(cu, k2, k1), t1);
                    const float v = __builtin_fmaf(e - gi[q], inv_tau_t, gi[q]);
                    if (ODD) {
                        *at_bytes(a.gslot[q], opaque_offset(self[k], 4)) = v;
                    } else if (q == 0) {
                        *at(a.gslot[q], self[k]) = v;
                    } else {
                        // zero-flux bounce-back at an open edge and anti-bounce-back
                        // at a wall stay at the cell; otherwise slot o at x + c_q
                        const int dx = 1 + L::c[q][0], dy = 1 + L::c[q][1], dz = 1 + L::c[q][2];
                        const bool ok = OK(k, dx, dy, dz), sol = SOLID(k, q);
                        float vv = v;
                        if (ok && sol) vv = __builtin_fmaf(2.0f * L::wt[q], WALL(k, q).temperature, -v);
                        const uint32_t e_here = GPAIR_E(q, self[k]);
                        const uint32_t e_dst = GPAIR_E(o, xs[k][dx] + ry[dy] + pz[dz]);
                        a.gslot[q < o ? q : o][(!ok || sol) ? e_here : e_dst] = vv;
                    }
                }
            }
        }

        // macroscopic save (collide.comp:784-792)
        if (a.vd) a.vd[self[k]] = make_float4(ux, uy, uz, rho);
        if (a.ft) a.ft[self[k]] = make_float4(Fx, Fy, Fz, temp);
    }

    if constexpr (AL) {
        __syncthreads();
        // store the values the row buffer received, each at its own column
        // (a tick that does not update wrote no entries)
#pragma unroll
        for (int k = 0; k < CPT; ++k) {
            const int x = xs[k][1];
            if (!in_bounds[k] || !a.update) continue;
#pragma unroll
            for (int i = 1; i < Q; ++i) {
                if (L::c[i][0] == 0) continue;
                const int sx = 1 - L::c[i][0];             // the writer sits at x - c_x
                if (!(PER || vx[k][sx])) continue;
                const float v = LDS(i, x);
                // (the periodic plain kernel has no empty entries: every cell
                // of the grid updates and no write leaves the domain)
                if (!kEvery && __float_as_uint(v) == kNoValue) continue;
                const int dy = 1 + L::c[i][1], dz = 1 + L::c[i][2];
                S::store(SLOT(opposite(i)), ROW(k, dy, dz), v);
            }
        }
    }

#undef LDS
#undef PAIR_E
#undef GPAIR_E
#undef NB
#undef ROW
#undef POP
#undef PAIR_BASE
#undef SLOT
#undef OK
#undef SOLID
#undef WALL
    // ---- FP64 block partial sums (deterministic two-stage reduction) ----
    if (partial) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            s_rho += __shfl_xor(s_rho, off);
            s_mx += __shfl_xor(s_mx, off);
            s_my += __shfl_xor(s_my, off);
            s_mz += __shfl_xor(s_mz, off);
            s_t += __shfl_xor(s_t, off);
        }
        __shared__ double red[8][5];
        const int tid = ty * a.bdx + tx;
        const int wave = tid >> 6, lane = tid & 63;
        if (lane == 0) {
            red[wave][0] = s_rho; red[wave][1] = s_mx; red[wave][2] = s_my;
            red[wave][3] = s_mz; red[wave][4] = s_t;
        }
        __syncthreads();
        if (tid < 5) {
            const int nw = (a.bdx * a.bdy + 63) >> 6;
            double acc = 0.0;
            for (int wv = 0; wv < nw; ++wv) acc += red[wv][tid];
            partial[blk.pb * 5 + tid] = acc;
        }
    }
}

template <int VS, int ST, bool ODD, bool WALLS, bool MP, bool TH, int CPT, bool AL, bool PER, int SRC>
__global__ void __launch_bounds__(512) BF_OCCUPANCY k_step(const StepArgs a, double *__restrict__ partial) {
    step_body<VS, ST, ODD, WALLS, MP, TH, CPT, AL, PER, SRC, true, true>(a, partial,
                                                                      grid_block(a, CPT, a.z0 + (int)blockIdx.z));
}

// Synchronous Shan-Chen: psi of this tick's densities, before the tick.
template <int VS, int ST, bool ODD, bool WALLS, bool PER>
__global__ void __launch_bounds__(512) k_density(const StepArgs a) {
    step_body<VS, ST, ODD, WALLS, true, false, 1, false, PER, kSrcOn, true, true, true>(
        a, nullptr, grid_block(a, 1, a.z0 + (int)blockIdx.z));
}

// A launch of a chunked sweep: the blocks run the plane ranges of `ps` in
// order, even tick, odd tick, and for a sweep of two pairs the next even and
// odd ticks.  The host schedules chunk j's even tick with the odd tick of the
// planes that end two planes before the chunk (and the next pair's ticks two
// planes behind each other), so the ranges touch disjoint planes (the even
// tick of a plane reads and writes its neighbour planes z +- 1; the odd tick
// stays at its cells).  The even tick's
// accesses are ordinary loads and stores, whose lines the 256 MB last-level
// cache keeps; the odd tick of those planes, one launch later, finds most of
// them there and streams its own traffic with nontemporal hints.  (Which hints the two halves use measured within noise
// of each other on the tick kernels; they decide the rate of the bare access
// pattern.)
#ifndef BF_PAIR_EVEN_NT_LOAD
#define BF_PAIR_EVEN_NT_LOAD 0
#endif
#ifndef BF_PAIR_EVEN_NT_STORE
#define BF_PAIR_EVEN_NT_STORE 0
#endif
#ifndef BF_PAIR_ODD_NT_LOAD
#define BF_PAIR_ODD_NT_LOAD 1
#endif
#ifndef BF_PAIR_ODD_NT_STORE
#define BF_PAIR_ODD_NT_STORE 1
#endif
#define BF_OCCUPANCY_PAIR __attribute__((amdgpu_waves_per_eu((MP || TH) ? 1 : BF_WAVES_PER_EU)))
// MP: the lagged Shan-Chen kernels.  Tick t reads its neighbours' psi from
// the buffer tick t - 1 wrote (radius one, like the populations) and writes
// its own psi to the other buffer, so the planes that the schedule keeps
// apart for the populations keep the psi reads and writes apart as well.
template <int VS, int ST, bool WALLS, bool MP, bool TH, int CPT, bool AL, bool PER, int SRC>
__global__ void __launch_bounds__(512) BF_OCCUPANCY_PAIR k_pair(const StepArgs ae, const StepArgs ao, const PairStages ps,
                                                                double *__restrict__ partial) {
    // the even chunk's planes come first in dispatch order (alternating the
    // planes of the two chunks measured 2-8 % slower)
    const int bz = (int)blockIdx.z;
    const int b1 = ps.n[0], b2 = b1 + ps.n[1], b3 = b2 + ps.n[2];
    const bool even = bz < b1 || (bz >= b2 && bz < b3);
    const int z = bz < b1 ? ps.z0[0] + bz
                : bz < b2 ? ps.z0[1] + (bz - b1)
                : bz < b3 ? ps.z0[2] + (bz - b2)
                          : ps.z0[3] + (bz - b3);
    // the last launches of a sweep wrap from plane Z - 1 to 0
    const int zw = z < ae.Z ? z : z - ae.Z;
    if (even) {
        step_body<VS, ST, false, WALLS, MP, TH, CPT, AL, PER, SRC, BF_PAIR_EVEN_NT_LOAD, BF_PAIR_EVEN_NT_STORE>(
            ae, nullptr, grid_block(ae, CPT, zw));
    } else {
        step_body<VS, ST, true, WALLS, MP, TH, CPT, false, PER, SRC, BF_PAIR_ODD_NT_LOAD, BF_PAIR_ODD_NT_STORE>(
            ao, partial, grid_block(ao, CPT, zw));
    }
}

// Equilibrium initialisation (set_equilibrium_populations.comp:198-220; the
// thermal lattice likewise, set_equilibrium_populations_thermal.comp:158-180,
// indexed by population).  Solid cells start at rest (LBM.cpp:534-535); psi of
// solid cells is psi(effective density) and never changes.
template <int VS, int ST, bool TH>
__global__ void __launch_bounds__(256) k_init(const InitArgs a) {
    using L = Lattice<VS>;
    using S = Store<ST, true, true>;
    using T = typename S::T;
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.n) return;
    const uint32_t pc = pop_index((uint32_t)c, a.slab_log, a.slab_skip);
    const float r = a.rho[c];
    float ux = a.u[3 * c], uy = a.u[3 * c + 1], uz = a.u[3 * c + 2];
    const uint8_t id = a.ids ? a.ids[c] : 0;
    if (id) { ux = 0.0f; uy = 0.0f; uz = 0.0f; }
    const float uu = __builtin_fmaf(uz, uz, __builtin_fmaf(uy, uy, ux * ux));
    const float c1 = -1.5f * uu, d = r - 1.0f;
#pragma unroll
    for (int i = 0; i < L::Q; ++i) {
        const float cu = i == 0 ? 0.0f : dotc<VS>(i, ux, uy, uz);
        const float w = L::w[i];
        const float md = (i == 0) ? __builtin_fmaf(w, d, L::rest_fix * d) : w * d;
        float e = __builtin_fmaf(w * r, __builtin_fmaf(cu, __builtin_fmaf(4.5f, cu, 3.0f), c1), md);
        // round to FP32 before the storage conversion: without the empty asm the
        // FMA and an FP16 conversion fuse into v_fma_mixlo_f16, which rounds once
        if constexpr (ST == ST_FP16) asm("" : "+v"(e));
        S::store(static_cast<T *>(a.slot[i]), pc, e);
    }
    if constexpr (TH) {
        const float T0 = a.temperature ? a.temperature[c] : 1.0f;
        const float k1 = 1.0f / a.cs2_t, k2 = 0.5f * k1 * k1;
        const float t1 = __builtin_fmaf(-0.5f * k1, uu, 1.0f);
#pragma unroll
        for (int k = 0; k < L::QT; ++k) {
            const float cu = k == 0 ? 0.0f : dotc<VS>(k, ux, uy, uz);
            const float wT = (k == 0) ? __builtin_fmaf(L::wt[0], T0, L::rest_fix_t * T0) : L::wt[k] * T0;
            a.g[(size_t)k * a.stride + c] = wT * __builtin_fmaf(cu, __builtin_fmaf(cu, k2, k1), t1);
        }
    }
    if (a.psi0) {
        const float rr = id ? a.rho_eff[id] : r;
        const float p = 1.0f - expf(-rr);
        a.psi0[c] = p;
        a.psi1[c] = p;
    }
}

template <int VS>
__global__ void __launch_bounds__(256) k_mask(const uint8_t *__restrict__ ids, uint32_t *__restrict__ nmask, int X,
                                              int Y, int Z, int px, int py, int pz) {
    using L = Lattice<VS>;
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (int64_t)X * Y * Z) return;
    const int x = (int)(c % X), y = (int)((c / X) % Y), z = (int)(c / ((int64_t)X * Y));
    uint32_t m = ids[c] != 0 ? 1u : 0u;
#pragma unroll
    for (int j = 1; j < L::Q; ++j) {
        int nx = x + L::c[j][0], ny = y + L::c[j][1], nz = z + L::c[j][2];
        if (nx < 0 || nx >= X) { if (!px) continue; nx = (nx + X) % X; }
        if (ny < 0 || ny >= Y) { if (!py) continue; ny = (ny + Y) % Y; }
        if (nz < 0 || nz >= Z) { if (!pz) continue; nz = (nz + Z) % Z; }
        if (ids[nx + (int64_t)X * (ny + (int64_t)Y * nz)] != 0) m |= 1u << j;
    }
    nmask[c] = m;
}

// The FP32 kernels keep the per-launch source flag (measured: the TGV tick
// pairs ran 2.8 % slower with the source compiled out, although they issue a
// quarter fewer vector instructions; they are bound by memory, not issue).
#ifndef BF_FP32_SRC_RUNTIME
#define BF_FP32_SRC_RUNTIME 1
#endif
#ifndef BF_ROWBUF_16BIT_WALLS
#define BF_ROWBUF_16BIT_WALLS 0  // build-time switch: row buffer also for 16-bit storage with walls
#endif

template <int ST, bool ODD>
hipError_t launch_step_st(const StepArgs &a_in, bool walls, bool mp, bool th, const LaunchShape &sh,
                          double *partial, hipStream_t stream) {
    constexpr int VS = BF_VS;
    StepArgs a = a_in;
    a.bdx = (int)sh.block.x;
    a.bdy = (int)sh.block.y;
    constexpr int CPT = cells_per_thread(ST, false, false);
    // The row buffer pays off for the plain kernels and for FP32 walls; with
    // walls and 16-bit storage the direct accesses are faster (the barriers
    // stall blocks whose loads wait on the solid-neighbour mask).
    const bool al = !ODD && sh.lds_bytes > 0 && !(walls && ST != ST_FP32 && !BF_ROWBUF_16BIT_WALLS);
    // Closed grids run the kernels without neighbour checks: the plain kernel
    // on a periodic grid, and the wall kernels (plain walls, thermal walls,
    // Shan-Chen with thermal walls) where solid planes close the grid.
#define BF_L2(W, M, T, P, S)                                                                               \
    do {                                                                                                   \
        if constexpr (!ODD && !M) {                                                                        \
            if (al) {                                                                                      \
                hipLaunchKernelGGL((k_step<VS, ST, ODD, W, M, T, (M || T) ? 1 : CPT, true, P, S>), sh.grid,    \
                                   sh.block, sh.lds_bytes, stream, a, partial);                            \
                return hipGetLastError();                                                                  \
            }                                                                                              \
        }                                                                                                  \
        hipLaunchKernelGGL((k_step<VS, ST, ODD, W, M, T, (M || T) ? 1 : CPT, false, P, S>), sh.grid, sh.block, \
                           0, stream, a, partial);                                                         \
        return hipGetLastError();                                                                          \
    } while (0)
#define BF_L(W, M, T, P)                                                                                   \
    do {                                                                                                   \
        if constexpr (M || T) BF_L2(W, M, T, P, kSrcOn);                                                   \
        else if constexpr (ST == ST_FP32 && BF_FP32_SRC_RUNTIME) BF_L2(W, M, T, P, kSrcRuntime);           \
        else if (a.source_on) BF_L2(W, M, T, P, kSrcOn);                                                   \
        else BF_L2(W, M, T, P, kSrcOff);                                                                   \
    } while (0)
    const bool closed = a.closed != 0;
    if (walls) {
        if (mp) {
            if (th) { if (closed) BF_L(true, true, true, true); BF_L(true, true, true, false); }
            BF_L(true, true, false, false);
        }
        if (th) { if (closed) BF_L(true, false, true, true); BF_L(true, false, true, false); }
        if (closed) BF_L(true, false, false, true);
        BF_L(true, false, false, false);
    }
    if (mp) { if (th) BF_L(false, true, true, false); else BF_L(false, true, false, false); }
    if (th) BF_L(false, false, true, false);
    if (closed) BF_L(false, false, false, true);
    BF_L(false, false, false, false);
#undef BF_L
#undef BF_L2
}

template <int ST>
hipError_t launch_pair_st(const StepArgs &ae_in, const StepArgs &ao_in, const PairStages &ps, bool walls, bool mp,
                          bool th, const LaunchShape &sh, double *partial, hipStream_t stream) {
    constexpr int VS = BF_VS;
    StepArgs ae = ae_in, ao = ao_in;
    ae.bdx = ao.bdx = (int)sh.block.x;
    ae.bdy = ao.bdy = (int)sh.block.y;
    constexpr int CPT = cells_per_thread(ST, false, false);
    // as for single ticks: the row buffer for the plain kernels and FP32 walls
    const bool al = !mp && sh.lds_bytes > 0 && !(walls && ST != ST_FP32 && !BF_ROWBUF_16BIT_WALLS);
    const dim3 grid(sh.grid.x, sh.grid.y, (unsigned)(ps.n[0] + ps.n[1] + ps.n[2] + ps.n[3]));
    const size_t lds = al ? sh.lds_bytes : 0;
#define BF_P2(W, T, A, P, S)                                                                               \
    do {                                                                                                   \
        hipLaunchKernelGGL((k_pair<VS, ST, W, false, T, T ? 1 : CPT, A, P, S>), grid, sh.block, lds, stream, ae, \
                           ao, ps, partial);                                                               \
        return hipGetLastError();                                                                          \
    } while (0)
    // the lagged Shan-Chen kernels (the combinations the single ticks use)
#define BF_PM(W, T, P)                                                                                     \
    do {                                                                                                   \
        hipLaunchKernelGGL((k_pair<VS, ST, W, true, T, 1, false, P, kSrcOn>), grid, sh.block, 0, stream, ae, ao, \
                           ps, partial);                                                                   \
        return hipGetLastError();                                                                          \
    } while (0)
#define BF_P(W, T, A, P)                                                                                   \
    do {                                                                                                   \
        if constexpr (T) BF_P2(W, T, A, P, kSrcOn);                                                        \
        else if constexpr (ST == ST_FP32 && BF_FP32_SRC_RUNTIME) BF_P2(W, T, A, P, kSrcRuntime);           \
        else if (ae.source_on) BF_P2(W, T, A, P, kSrcOn);                                                  \
        else BF_P2(W, T, A, P, kSrcOff);                                                                   \
    } while (0)
    const bool closed = ae.closed != 0;
    if (mp) {
        if (walls) {
            if (th) { if (closed) BF_PM(true, true, true); BF_PM(true, true, false); }
            BF_PM(true, false, false);
        }
        if (th) BF_PM(false, true, false);
        BF_PM(false, false, false);
    }
    if (walls) {
        if (th) {
            if (closed) { if (al) BF_P(true, true, true, true); BF_P(true, true, false, true); }
            if (al) BF_P(true, true, true, false);
            BF_P(true, true, false, false);
        }
        if (closed) { if (al) BF_P(true, false, true, true); BF_P(true, false, false, true); }
        if (al) BF_P(true, false, true, false);
        BF_P(true, false, false, false);
    }
    if (th) { if (al) BF_P(false, true, true, false); BF_P(false, true, false, false); }
    if (closed) { if (al) BF_P(false, false, true, true); BF_P(false, false, false, true); }
    if (al) BF_P(false, false, true, false);
    BF_P(false, false, false, false);
#undef BF_P
#undef BF_P2
#undef BF_PM
}

template <int ST, bool ODD>
hipError_t launch_density_st(const StepArgs &a_in, bool walls, const LaunchShape &sh, hipStream_t stream) {
    constexpr int VS = BF_VS;
    StepArgs a = a_in;
    a.bdx = (int)sh.block.x;
    a.bdy = (int)sh.block.y;
    const bool closed = a.closed != 0;
#define BF_D(W, P)                                                                                         \
    do {                                                                                                   \
        hipLaunchKernelGGL((k_density<VS, ST, ODD, W, P>), sh.grid, sh.block, 0, stream, a);               \
        return hipGetLastError();                                                                          \
    } while (0)
    if (walls) { if (closed) BF_D(true, true); BF_D(true, false); }
    BF_D(false, false);
#undef BF_D
}

template <int ST>
hipError_t launch_init_st(const InitArgs &a, bool th, hipStream_t stream) {
    const dim3 block(256), grid((unsigned)((a.n + 255) / 256));
    if (th) hipLaunchKernelGGL((k_init<BF_VS, ST, true>), grid, block, 0, stream, a);
    else hipLaunchKernelGGL((k_init<BF_VS, ST, false>), grid, block, 0, stream, a);
    return hipGetLastError();
}

}  // namespace

template <>
hipError_t launch_step<BF_VS>(const StepArgs &a, int storage, bool odd, bool walls, bool mp, bool th,
                              const LaunchShape &sh, double *partial, hipStream_t stream) {
    switch (storage) {
    case ST_FP32: return odd ? launch_step_st<ST_FP32, true>(a, walls, mp, th, sh, partial, stream)
                             : launch_step_st<ST_FP32, false>(a, walls, mp, th, sh, partial, stream);
    case ST_FP16: return odd ? launch_step_st<ST_FP16, true>(a, walls, mp, th, sh, partial, stream)
                             : launch_step_st<ST_FP16, false>(a, walls, mp, th, sh, partial, stream);
    case ST_BF16: return odd ? launch_step_st<ST_BF16, true>(a, walls, mp, th, sh, partial, stream)
                             : launch_step_st<ST_BF16, false>(a, walls, mp, th, sh, partial, stream);
    }
    return hipErrorInvalidValue;
}

template <>
hipError_t launch_density<BF_VS>(const StepArgs &a, int storage, bool odd, bool walls, const LaunchShape &sh,
                                 hipStream_t stream) {
    switch (storage) {
    case ST_FP32: return odd ? launch_density_st<ST_FP32, true>(a, walls, sh, stream)
                             : launch_density_st<ST_FP32, false>(a, walls, sh, stream);
    case ST_FP16: return odd ? launch_density_st<ST_FP16, true>(a, walls, sh, stream)
                             : launch_density_st<ST_FP16, false>(a, walls, sh, stream);
    case ST_BF16: return odd ? launch_density_st<ST_BF16, true>(a, walls, sh, stream)
                             : launch_density_st<ST_BF16, false>(a, walls, sh, stream);
    }
    return hipErrorInvalidValue;
}

template <>
hipError_t launch_pair<BF_VS>(const StepArgs &ae, const StepArgs &ao, const PairStages &ps, int storage, bool walls,
                              bool mp, bool th, const LaunchShape &sh, double *partial, hipStream_t stream) {
    switch (storage) {
    case ST_FP32: return launch_pair_st<ST_FP32>(ae, ao, ps, walls, mp, th, sh, partial, stream);
    case ST_FP16: return launch_pair_st<ST_FP16>(ae, ao, ps, walls, mp, th, sh, partial, stream);
    case ST_BF16: return launch_pair_st<ST_BF16>(ae, ao, ps, walls, mp, th, sh, partial, stream);
    }
    return hipErrorInvalidValue;
}

template <>
hipError_t launch_mask<BF_VS>(const uint8_t *ids, uint32_t *nmask, int X, int Y, int Z, int px, int py, int pz,
                              hipStream_t stream) {
    const int64_t n = (int64_t)X * Y * Z;
    const dim3 block(256), grid((unsigned)((n + 255) / 256));
    hipLaunchKernelGGL((k_mask<BF_VS>), grid, block, 0, stream, ids, nmask, X, Y, Z, px, py, pz);
    return hipGetLastError();
}

template <>
hipError_t launch_init<BF_VS>(const InitArgs &a, int storage, bool th, hipStream_t stream) {
    switch (storage) {
    case ST_FP32: return launch_init_st<ST_FP32>(a, th, stream);
    case ST_FP16: return launch_init_st<ST_FP16>(a, th, stream);
    case ST_BF16: return launch_init_st<ST_BF16>(a, th, stream);
    }
    return hipErrorInvalidValue;
}

}  // namespace bf
