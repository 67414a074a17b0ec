// bf_kernels.h — argument blocks and launchers of the HIP kernels.
//
// The kernels are templates over the velocity set (one translation unit per
// set, bf_step.hip compiled with -DBF_VS=<id>), the storage type, the tick
// parity and three feature switches; everything else is a runtime argument.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bf {

enum Storage : int { ST_FP16 = 0, ST_FP32 = 1, ST_BF16 = 2 };

// Per-object wall data, one entry per boundary id (LBM.cpp:613-629 holds the
// same quantities; only the translational velocity, temperature and effective
// density enter the update).
struct alignas(16) ObjectEntry {
    float ux, uy, uz, temperature;
};

constexpr int kMaxQ = 27, kMaxQT = 7;

// Lattice layout.  Slot i of cell c is element pop_index(c) of the array that
// starts at slot[i], with pop_index(c) = c + (c >> slab_log) * slab_skip:
//  - slabs: the cells are cut into slabs of 2^slab_log cells and the slots
//    into groups of W adjacent slots (opposite pairs never split); a group
//    stores, slab after slab, its W slots of that slab one after the other,
//    so slab_skip = (W - 1) 2^slab_log, and slot i + 1 sits 2^slab_log
//    elements after slot i within a group (pair_step);
//  - plain structure of arrays: slab_log = 31 (no cell reaches the second
//    slab), slab_skip = 0, pair_step = the slot stride.
// Every group stays below 4 GiB, so byte offsets from a slot base are 32-bit.
#ifndef BF_SLAB_LOG
#define BF_SLAB_LOG 13
#endif
constexpr int kSlabLog = BF_SLAB_LOG;  // 2^13 cells per slab

struct StepArgs {
    void *f;                 // lattice (storage type)
    float *g;                // thermal lattice, QT slots of `stride` floats (or null)
    int64_t stride;          // elements between thermal slots
    void *slot[kMaxQ];       // base of slot i
    uint32_t slab_log, slab_skip, pair_step;
    float *gslot[kMaxQT];    // base of thermal slot q (g + q * stride)
    int X, Y, Z;
    int z0;                  // first plane of the launch (grid.z planes from z0)
    int bdx, bdy;            // block shape of the launch (set by the launcher; scalar loads
                             // instead of the dispatch packet's 16-bit group size)
    int periodic_x, periodic_y, periodic_z;
    // Every neighbour of every cell that updates lies in the grid: the axes
    // are periodic, or every cell on the two boundary planes of a
    // non-periodic axis is solid (checked on the host from the ids).  The
    // launcher then runs the kernels without neighbour checks.
    int closed;
    const uint8_t *ids;      // boundary id per cell (or null)
    const uint32_t *nmask;   // solid-neighbour mask per cell (see launch_mask), with ids
    const ObjectEntry *objects;  // 256 entries
    int halfway;             // 1: halfway bounce-back redirect, 0: full-way (solid cells skipped only)
    int wall_ids;            // wall terms need the neighbours' ids (moving walls or thermal walls)
    int forcing;             // 0 none, 1 constant, 2 per-cell field
    float fx, fy, fz;        // constant force
    const float4 *force;     // per-cell force (xyz) when forcing == 2
    float tau;
    // 1 / tau and the Guo prefactor 1 - 1 / (2 tau), rounded once on the host
    // as the kernel would (IEEE division): a division in the kernel is a
    // dozen vector instructions per wave, although its operand is uniform
    float inv_tau, guo_pre;
    float ulim, uext;        // velocity clamp / equilibrium reset thresholds
    int source_on;           // Guo source active (forcing || multiphase || thermal)
    // Shan-Chen
    float G;
    const float *psi_prev;   // psi written on the previous tick
    float *psi_cur;          // psi written on this tick
    int write_psi;
    // thermal
    float tau_t, beta, t_ref, cs2_t;
    float inv_tau_t, k1_t, k2_t;  // 1 / tau_t, 1 / cs2_t, k1_t^2 / 2 (host-rounded, as inv_tau)
    // outputs
    int update;              // write populations
    float4 *vd;              // (u, rho) when non-null
    float4 *ft;              // (F, T) when non-null
    double *sums;            // 5 doubles accumulated when non-null
};

struct InitArgs {
    void *slot[kMaxQ];       // as in StepArgs
    uint32_t slab_log, slab_skip;
    float *g;
    int64_t stride;          // elements between thermal slots
    int64_t n;
    const float *rho;
    const float *u;          // 3 per cell
    const float *temperature;  // or null (T = 1)
    const uint8_t *ids;      // or null
    const ObjectEntry *objects;
    const float *rho_eff;    // 256 entries
    float *psi0, *psi1;      // or null
    float cs2_t;
};

// Cells updated per thread by the plain and wall kernels (the multiphase and
// thermal kernels update one).  Build-time tunable per storage width.
#ifndef BF_CPT32
#define BF_CPT32 1
#endif
#ifndef BF_CPT16
#define BF_CPT16 1
#endif
constexpr int cells_per_thread(int storage, bool multiphase, bool thermal) {
    return (multiphase || thermal) ? 1 : (storage == ST_FP32 ? BF_CPT32 : BF_CPT16);
}

struct LaunchShape {
    dim3 grid, block;
    size_t lds_bytes;
};

// Block shape for a grid of X x Y x Z cells.  Multiphase kernels stage the
// pseudopotential of the block tile plus a one-cell halo in LDS and use a
// 64 x 4 tile; the others cover whole x-rows (up to 512 cells per row and
// thread group) and then get `lds_bytes` of row buffer for the aligned even
// tick, or row segments of 256 cells (lds_bytes = 0) for longer rows.
LaunchShape step_shape(int X, int Y, int Z, bool multiphase, int cells_per_thread, int velocity_set);

// Directions with c_x != 0 of a velocity set (rows of the even-tick buffer).
int x_direction_count(int velocity_set);

// Launch one tick.  `partial` (5 doubles per block) receives FP64 block sums
// of rho, raw momentum and T when non-null.  Returns hipSuccess or the launch error.
template <int VS>
hipError_t launch_step(const StepArgs &a, int storage, bool odd, bool walls, bool multiphase,
                       bool thermal, const LaunchShape &shape, double *partial, hipStream_t stream);

// Synchronous Shan-Chen density pre-pass of the coming tick: the tick's reads
// of the main lattice, then psi of the floored density to a.psi_cur (no other
// writes).  Same shape and parity as the tick it precedes.
template <int VS>
hipError_t launch_density(const StepArgs &a, int storage, bool odd, bool walls, const LaunchShape &shape,
                          hipStream_t stream);

// Plane ranges of one launch of a chunked sweep, in dispatch order: an even
// tick, the odd tick after it, and for a sweep of two tick pairs the next even
// and odd ticks.  Range k covers n[k] planes from z0[k], modulo Z.
struct PairStages {
    int z0[4];
    int n[4];
};

// One launch of a chunked sweep (multiphase: the lagged Shan-Chen kernels,
// whose psi buffers `ae` and `ao` name in tick order): the ranges of `ps`
// in one grid, the even ranges with `ae`, the odd ranges with `ao`.  A range
// must not run beside another that touches its planes or their neighbours
// (the host's schedule keeps every tick of a plane in a later launch than
// the ticks it depends on; see k_pair).  `partial` receives the odd ticks'
// FP64 block sums when non-null.
template <int VS>
hipError_t launch_pair(const StepArgs &ae, const StepArgs &ao, const PairStages &ps, int storage, bool walls,
                       bool multiphase, bool thermal, const LaunchShape &shape, double *partial, hipStream_t stream);

// Neighbour mask of every cell: bit 0 = the cell is solid, bit i (1..Q-1) =
// the neighbour at x + c_i lies inside the domain (periodic axes wrap) and is
// solid.  Built once per boundary field; the even tick selects its bounce-back
// addresses from it with one 4-byte load instead of Q-1 neighbour id loads.
template <int VS>
hipError_t launch_mask(const uint8_t *ids, uint32_t *nmask, int X, int Y, int Z, int px, int py, int pz,
                       hipStream_t stream);

template <int VS>
hipError_t launch_init(const InitArgs &a, int storage, bool thermal, hipStream_t stream);

}  // namespace bf
