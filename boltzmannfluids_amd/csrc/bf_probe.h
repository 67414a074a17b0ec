/*
 * bf_probe.h — bandwidth probes of libbfprobe.so (measurement infrastructure
 * for the benchmark; not part of the solver's C ABI in include/bf/bf.h, and
 * not linked into libbfamd.so).  Return values are bf_status codes.
 */
#ifndef BF_PROBE_H
#define BF_PROBE_H

#include <stdint.h>

#include "../../include/bf/bf.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Device copy bandwidth: copies `bytes` between two device buffers `iters`
 * times and returns (read + write) GB/s: the achievable-HBM figure beside the
 * published 8 TB/s, and the known-traffic kernel for counter calibration. */
bf_status bfp_copy_bandwidth(int32_t device, int64_t bytes, int32_t iters, double *gbps);

/* In-place streaming: every cell reads and rewrites one FP32 element in each
 * of `slots` (5, 7, 9, 15, 19 or 27) slots of `cells` cells in the lattice's
 * slab layout.  *gbps = read + write bytes per second over `iters` launches:
 * the ceiling of the tick's access pattern without its arithmetic. */
bf_status bfp_inplace_bandwidth(int32_t device, int64_t cells, int32_t slots, int32_t iters, double *gbps);

/* The memory traffic of one D3Q19 tick of the in-place AA update on a periodic
 * X x Y x Z grid in the slab layout, without arithmetic.  even = 0: every cell
 * reads and rewrites its 19 slots at the cell itself (the odd tick); even = 1:
 * slot i is read in the row (y - c_iy, z - c_iz) and slot opp(i) written in
 * the row (y + c_iy, z + c_iz), at the cell's own x (the even tick, whose x
 * shift goes through LDS).  X <= 1024. */
bf_status bfp_aa_bandwidth(int32_t device, int32_t X, int32_t Y, int32_t Z, int32_t even, int32_t iters,
                           double *gbps);

/* Two-lattice pull pattern for comparison with the AA pattern: slot i of each
 * cell is read from one D3Q19 lattice in the row (y - c_iy, z - c_iz) and
 * written to a second lattice at the cell itself, at the cell's own x, on a
 * periodic X x Y x Z grid in the slab layout; launches alternate the two
 * lattices.  X <= 1024. */
bf_status bfp_pull_bandwidth(int32_t device, int32_t X, int32_t Y, int32_t Z, int32_t iters, double *gbps);

#ifdef __cplusplus
}
#endif

#endif /* BF_PROBE_H */
