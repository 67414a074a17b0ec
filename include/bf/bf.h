/*
 * bf.h — C ABI of the MI355X lattice Boltzmann solver (libbfamd.so).
 *
 * This is the drop-in boundary for the per-tick update of the reference
 * solver class `LBM` (BoltzmannFluids/Source/LBM/LBM.h:26-332).  Every entry
 * point takes plain pointers and sizes, returns a bf_status, never throws, and
 * copies host buffers synchronously.  A handle owns all device memory and one
 * HIP stream; it is thread-affine (one thread at a time per handle).
 *
 * Semantics of one tick (all velocity sets, all storage types):
 *   - the lattice is stored in place, one slot per direction; the device
 *     layout is internal (slabs of 8192 cells holding their slots one after
 *     the other, cell = z*X*Y + y*X + x), and population reads and checkpoint
 *     files use the logical order population * cells + cell.  The lattice
 *     is updated with the esoteric "AA" pattern of
 *     BoltzmannFluids/Source/GLSL/LBM/collide.comp:244-369: on even ticks a
 *     cell reads slot i at x - c_i and writes slot opp(i) at x + c_i, on odd
 *     ticks it reads slot opp(i) at x and writes slot i at x;
 *   - BGK collision with Guo forcing, velocity clamp and density floor as in
 *     collide.comp:676-793;
 *   - every storage type keeps the shifted value h_i = f_i - w_i and the
 *     tick computes on h in FP32 (rho = 1 + sum h_i), so low-Mach deviations
 *     keep the full mantissa; FP16/BF16 storage converts to FP32 on load.
 */
#ifndef BF_BF_H
#define BF_BF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BF_ABI_VERSION 1

typedef enum bf_status {
    BF_OK = 0,
    BF_ERR_INVALID = 1,      /* bad argument (reference: "[LBM Error]" + ASSERT) */
    BF_ERR_STATE = 2,        /* call out of order, e.g. bf_step before bf_init_fields */
    BF_ERR_UNSUPPORTED = 3,  /* feature combination not built */
    BF_ERR_DEVICE = 4,       /* a HIP call failed; bf_last_error has the text */
    BF_ERR_OOM = 5           /* device allocation failed */
} bf_status;

/* Velocity sets; values match enum VelocitySet in
 * BoltzmannFluids/Source/LBMConstants/VelocitySet.h:18-23.  Direction order and
 * weights follow VelocitySet.cpp:53-141 exactly. */
typedef enum bf_velocity_set {
    BF_D2Q9 = 2,
    BF_D3Q15 = 3,
    BF_D3Q19 = 4,
    BF_D3Q27 = 5
} bf_velocity_set;

/* Thermal (advection-diffusion) sets; VelocitySet.h:13-16, VelocitySet.cpp:180-203. */
typedef enum bf_thermal_set {
    BF_D2Q5 = 0,
    BF_D3Q7 = 1
} bf_thermal_set;

/* Population storage.  BF_FP32 and BF_FP16 keep the numeric values of enum
 * FloatingPointAccuracy (FloatingPointAccuracy.h:6-9); the reference stores
 * f unshifted, this library stores h = f - w (shifted) for every type. */
typedef enum bf_storage {
    BF_STORE_FP16 = 0,
    BF_STORE_FP32 = 1,
    BF_STORE_BF16 = 2
} bf_storage;

/* Wall treatment for solid (boundary id != 0) cells. */
typedef enum bf_boundary_mode {
    /* Halfway bounce-back with the moving-wall (Ladd) term of
     * stream.comp:487-496,576-580, applied inside the in-place AA update. */
    BF_BB_HALFWAY = 0,
    /* Reference esoteric behaviour (collide.comp:696-700): solid cells are
     * skipped and populations return to their sender two ticks later
     * (full-way bounce-back, no wall velocity). */
    BF_BB_FULLWAY_COMPAT = 1
} bf_boundary_mode;

typedef struct bf_config {
    int32_t resolution[3];     /* X, Y, Z (Z == 1 for D2Q9)           LBM.h:50 */
    float relaxation_time;     /* tau                                  LBM.h:51 */
    int32_t periodic[3];       /* periodic_x/y/z                       LBM.h:52-54 */
    int32_t velocity_set;      /* bf_velocity_set                      LBM.h:55 */
    int32_t storage;           /* bf_storage                           LBM.h:56 */
    int32_t multiphase;        /* Shan-Chen pseudopotential on/off     LBM.h:57 */
    int32_t boundary_mode;     /* bf_boundary_mode */
    int32_t device;            /* HIP device ordinal */
} bf_config;

/* Scalar / vector parameters for bf_set_param.  Defaults are the reference's. */
typedef enum bf_param {
    BF_PARAM_RELAXATION_TIME = 0,        /* 1 float; LBM::set_relaxation_time, LBM.cpp:285 */
    BF_PARAM_CONSTANT_FORCE = 1,         /* 3 floats; LBM::set_constant_force, LBM.cpp:366 */
    BF_PARAM_INTERACTION_STRENGTH = 2,   /* 1 float G (default -6); LBM.h:257, LBM.cpp:371 */
    BF_PARAM_THERMAL_EXPANSION = 3,      /* 1 float beta (default 0.5); LBM.h:243 */
    BF_PARAM_THERMAL_RELAXATION_TIME = 4,/* 1 float tau_T (default 2); LBM.cpp:461 */
    BF_PARAM_REFERENCE_TEMPERATURE = 5,  /* 1 float T_ref (default 1); collide.comp:713 */
    BF_PARAM_THERMAL_CS2 = 6,            /* 1 float c_s^2 of the thermal set (default 1/3); LBM.cpp:1611 */
    BF_PARAM_VELOCITY_LIMIT = 7,         /* 1 float (default 0.25); LBM.h:144 */
    BF_PARAM_VELOCITY_LIMIT_EXTREME = 8, /* 1 float (default 0.30); LBM.h:145 */
    /* 1 float, 0 or 1 (default 0).  0 = the reference's lagged neighbour
     * pseudopotential (psi of the previous tick, collide.comp:537,567 with the
     * ping-pong of LBM.cpp:1617-1621,1649-1650); 1 = synchronous: a density
     * pre-pass per tick gives every neighbour's psi of the current tick
     * (costs one extra read of the lattice per tick).  No counterpart in the
     * reference. */
    BF_PARAM_SHAN_CHEN_SYNCHRONOUS = 9,
    /* 1 float (default 0): how bf_step runs an even tick and the odd tick
     * after it.  -1 = each tick over the whole grid; k > 0 = the pair chunk
     * by chunk of k z-planes, chunk j's even tick in one launch with the odd
     * tick of the planes up to two before it, so that the odd tick reads
     * what the previous launch's even tick wrote from the last-level cache;
     * 0 = chunks of about 120 MB of lattice for every lattice that exceeds
     * 384 MB (where it was measured faster), else whole-grid ticks.  Results are
     * identical in every mode (each cell sees the same operations).  The
     * synchronous Shan-Chen mode always runs whole-grid ticks, and so does a
     * grid of fewer than four chunks; the lagged Shan-Chen mode pairs only
     * with k > 0 (0 runs it in whole-grid ticks, which measured faster).  No
     * counterpart in the reference. */
    BF_PARAM_TICK_PAIR_PLANES = 10,
    /* 1 float, 0, 2 or 4 (default 0): ticks per chunked sweep.  4 runs two
     * tick pairs in one sweep over the chunks: each launch also carries the
     * next pair's even tick and odd tick on the planes the tick before each
     * has finished, so every tick of a plane reads what the previous launch
     * wrote.  A bf_step call that saves fields or sums on its last tick ends
     * with a single pair.  0 = 4 with chunks of about 80 MB for 16-bit
     * lattices without boundary ids, a thermal lattice or Shan-Chen when
     * BF_PARAM_TICK_PAIR_PLANES is automatic (where it was measured faster),
     * else 2.  Results are identical in every mode.  No counterpart in the
     * reference. */
    BF_PARAM_TICK_SWEEP_TICKS = 11
} bf_param;

/* Fields for bf_read_field. */
typedef enum bf_field {
    /* float4 per cell (u.x, u.y, u.z, rho) saved on the last BF_STEP_SAVE_MACRO
     * tick; layout of velocity_density_texture (collide.comp:785, LBM.cpp:299). */
    BF_FIELD_VELOCITY_DENSITY = 0,
    /* uint8 per cell: boundary id (boundry_texture R8, LBM.cpp:300). */
    BF_FIELD_BOUNDARY = 1,
    /* float4 per cell (F.x, F.y, F.z, T); force_temperature_texture (collide.comp:790). */
    BF_FIELD_FORCE_TEMPERATURE = 2,
    /* Q*N floats: the in-place lattice's slots (the AA state, not the
     * streamed populations), population-major in cell order, converted to
     * unshifted FP32 (h + w in FP32). */
    BF_FIELD_POPULATIONS_RAW = 3,
    /* Qt*N floats: the in-place thermal lattice as stored (AA slot layout). */
    BF_FIELD_THERMAL_RAW = 4,
    /* N floats: pseudopotential psi written on the last tick (Shan-Chen). */
    BF_FIELD_PSI = 5,
    /* Q*N floats, population-major: the populations f_i(x) that the next
     * tick's collision at x uses (after streaming, before collision), derived
     * from the slots by that tick's AA read rules (collide.comp:244-305): on
     * an odd tick slot opp(i) at x; on an even tick slot i at x - c_i, 0.01
     * where x - c_i lies outside the domain, and with halfway bounce-back the
     * cell's own slot opp(i) plus the moving-wall term where x - c_i is
     * solid.  Solid cells report their slots as stored.  Computed on the
     * host from BF_FIELD_POPULATIONS_RAW. */
    BF_FIELD_POPULATIONS = 6
} bf_field;

/* Flags for bf_step. */
#define BF_STEP_SAVE_MACRO 1u     /* fill the macroscopic fields on the last tick (LBM.cpp:1626-1634) */
#define BF_STEP_CONSERVATION 2u   /* FP64 sums of rho and rho*u over fluid cells on the last tick */

typedef struct bf_solver bf_solver;

/* --- lifetime ----------------------------------------------------------- */

/* Library ABI version (BF_ABI_VERSION). */
int32_t bf_abi_version(void);

/* Number of populations Q of a velocity set, 0 if unknown
 * (get_VelocitySet_vector_count, VelocitySet.cpp:21-35). */
int32_t bf_velocity_count(int32_t velocity_set);

/* Copy the direction table of a velocity set: 4 floats per direction
 * (c.x, c.y, c.z, w) — the velocity_set_buffer UBO layout (LBM.cpp:869-875).
 * `thermal` selects bf_thermal_set values instead.  out must hold 4*Q floats. */
bf_status bf_velocity_table(int32_t set, int32_t thermal, float *out, int32_t capacity);

/* Create a solver: validates the config and allocates the lattice.
 * Replaces the allocation half of LBM::initialize_fields
 * (LBM.cpp:415-462 and _generate_lattice_buffer, LBM.cpp:808-876).
 * At most 2^30 - 2^14 cells (the lattice is split into slot groups that each
 * stay below 4 GiB of 32-bit offsets). */
bf_status bf_create(const bf_config *cfg, bf_solver **out);

/* Release everything the handle owns.  Safe on NULL. */
void bf_destroy(bf_solver *s);

/* Last error text of this handle (or of the last failed bf_create when s is NULL). */
const char *bf_last_error(const bf_solver *s);

/* --- setup -------------------------------------------------------------- */

/* Describe boundary object `id` (1..255): translational wall velocity,
 * angular velocity, centre of mass, wall temperature and effective density.
 * Replaces LBM::set_boundry_properties (LBM.h:60-80, LBM.cpp:1015-1065).
 * May be called before or after bf_init_fields. */
bf_status bf_set_object(bf_solver *s, uint32_t id, const float velocity[3],
                        const float angular_velocity[3], const float center_of_mass[3],
                        float temperature, float effective_density);

/* Forget all objects (LBM::clear_boundry_properties, LBM.cpp:1067-1070). */
bf_status bf_clear_objects(bf_solver *s);

/* Initialise all fields from host arrays in cell order (cell = z*X*Y + y*X + x).
 * Replaces the per-cell lambda of LBM::initialize_fields (LBM.cpp:415-506):
 *   rho[N]; u[3N] (xyz per cell); force[3N] or NULL; temperature[N] or NULL;
 *   boundary_id[N] or NULL (0 = fluid).
 * As in the reference, the force is detected constant or varying
 * (LBM.cpp:531-541,562), the flow is thermal when a temperature differs from
 * the temperature of cell 0 (LBM.cpp:542-545), solid cells start at rest
 * (LBM.cpp:534-535), populations start at equilibrium (set_equilibrium_populations.comp)
 * and the macroscopic fields are saved once without updating the lattice
 * (LBM.cpp:502).  Resets the tick counter to 0. */
bf_status bf_init_fields(bf_solver *s, const float *rho, const float *u,
                         const float *force, const float *temperature,
                         const uint8_t *boundary_id);

/* Set a runtime parameter (see bf_param). */
bf_status bf_set_param(bf_solver *s, int32_t key, const float *value, int32_t n);

/* Read a runtime parameter. */
bf_status bf_get_param(const bf_solver *s, int32_t key, float *value, int32_t n);

/* --- iterate ------------------------------------------------------------ */

/* Advance n ticks (n >= 0) with the fused in-place kernel.  Asynchronous on the
 * handle's stream.  Replaces LBM::iterate_time / _iterate_time / _collide
 * (LBM.cpp:220-240, 1484-1493, 1557-1653); wall-clock pacing stays with the
 * caller.  flags: BF_STEP_SAVE_MACRO, BF_STEP_CONSERVATION (both act on the
 * last of the n ticks). */
bf_status bf_step(bf_solver *s, int64_t n, uint32_t flags);

/* Same as bf_step but also times the n launches with HIP events recorded on the
 * handle's stream; blocks until done and returns the elapsed milliseconds. */
bf_status bf_step_timed(bf_solver *s, int64_t n, uint32_t flags, double *elapsed_ms);

/* Ticks advanced since bf_init_fields (LBM::get_total_ticks_elapsed, LBM.cpp:242). */
int64_t bf_ticks(const bf_solver *s);

/* Wait for all work on the handle's stream. */
bf_status bf_synchronize(bf_solver *s);

/* --- readback ----------------------------------------------------------- */

/* Copy a field to host memory (synchronous).  `bytes` must equal the field's
 * size.  Replaces get_velocity_density_texture / get_boundry_texture /
 * get_force_temperature_texture (LBM.h:97-99) for callers that read on the host. */
bf_status bf_read_field(bf_solver *s, int32_t field, void *dst, size_t bytes);

/* Size in bytes of a field (0 if the field does not exist for this solver). */
size_t bf_field_bytes(const bf_solver *s, int32_t field);

/* --- field textures (device to device) ---------------------------------- */

/* Copy a macroscopic field, in the encoding of the reference's GL texture for
 * it, into device memory `dst` of `bytes` bytes (exactly the field's size),
 * asynchronously on the handle's stream.  Replaces the texture side of the
 * macroscopic save (imageStore in collide.comp:784-792 into the textures of
 * LBM.cpp:957-1003, formats LBM.h:299-301) for consumers on the device:
 *   BF_FIELD_VELOCITY_DENSITY, BF_FIELD_FORCE_TEMPERATURE: RGBA32F texels, one
 *     float4 per cell in cell order (16 bytes per cell);
 *   BF_FIELD_BOUNDARY: R8 texels, what the reference's imageStore of
 *     vec4(boundary id) into an r8 image stores: 0 for fluid, 255 for any
 *     object (the store clamps to [0, 1]).
 * The MI355X has no image support and no graphics pipeline, so GL textures
 * cannot share its memory; a renderer on another device reads the fields with
 * bf_read_field.  BF_ERR_INVALID for another field or size. */
bf_status bf_copy_field_to_device(bf_solver *s, int32_t field, void *dst, size_t bytes);

/* FP64 conservation sums over fluid cells of the current state:
 * out[0] = sum rho, out[1..3] = sum of sum_i f_i c_i (raw momentum, no
 * half-force shift), out[4] = sum T (thermal flows, else 0).  Computed on the
 * GPU from the populations the next tick will read.  No reference counterpart. */
bf_status bf_conservation(bf_solver *s, double out[5]);

/* Sums recorded by the last bf_step(..., BF_STEP_CONSERVATION), same layout. */
bf_status bf_last_step_sums(bf_solver *s, double out[5]);

/* Logical description of the lattice for interop and tests. */
typedef struct bf_info {
    int32_t resolution[3];
    int32_t q;                 /* populations per cell */
    int32_t q_thermal;         /* thermal populations per cell, 0 if not thermal */
    int32_t storage;
    int32_t forcing;           /* 0 none, 1 constant, 2 per-cell field */
    int32_t thermal;
    int32_t multiphase;
    int32_t has_boundaries;
    int64_t cells;
    int64_t stride;            /* elements between slot i and slot i + 1 of one cell */
    int64_t device_bytes;      /* bytes of device memory the handle holds */
} bf_info;

bf_status bf_get_info(const bf_solver *s, bf_info *out);

/* Pack/unpack boundary ids with `bits` (1, 2, 4 or 8) bits per cell in uint32
 * words, the format of the reference's boundries buffer
 * (LBM.cpp:566-573,631-659; collide.comp:401-418).  words must hold
 * ceil(n*bits/32) entries.  Host-only helpers. */
int32_t bf_bits_per_boundary(int32_t object_count);
bf_status bf_pack_boundaries(const uint8_t *ids, int64_t n, int32_t bits, uint32_t *words);
bf_status bf_unpack_boundaries(const uint32_t *words, int64_t n, int32_t bits, uint8_t *ids);

/* --- differentiable channel (no counterpart in the reference) ------------
 *
 * A separate solver for gradient-based design: a channel with halfway
 * bounce-back walls at y = 0 and y = Y-1 (z periodic in 3-D), a Zou-He
 * velocity inlet at x = 0 with a Poiseuille profile of mean `inlet_velocity`,
 * a Zou-He pressure outlet at x = X-1, and a per-cell porosity gamma in [0, 1]
 * applied by partial bounce-back.  The objective J is the mean outlet u_x after
 * `steps` ticks from rest; bf_channel_gradient returns dJ/dgamma (per cell) and
 * dJ/d inlet_velocity from the hand-written discrete adjoint, recomputing the
 * forward states between checkpoints taken every `checkpoint_every` ticks. */

typedef struct bf_channel bf_channel;

typedef struct bf_channel_config {
    int32_t resolution[3];     /* X, Y, Z (Z == 1 for D2Q9) */
    int32_t velocity_set;      /* BF_D2Q9 or BF_D3Q19 */
    float relaxation_time;
    float inlet_velocity;      /* mean u_x of the inlet profile 6 s (1 - s) */
    float outlet_density;      /* rho imposed at the outlet */
    int32_t device;
} bf_channel_config;

typedef struct bf_adjoint_config {
    int64_t steps;                    /* ticks of the forward run (T) */
    int32_t checkpoint_every;         /* K: a checkpoint every K ticks */
    int64_t trajectory_budget_bytes;  /* device bytes for segments kept resident
                                         instead of recomputed (at least one
                                         segment buffer is always used, plus
                                         one more while segments are
                                         recomputed beside the adjoint sweep) */
} bf_adjoint_config;

bf_status bf_channel_create(const bf_channel_config *cfg, bf_channel **out);
void bf_channel_destroy(bf_channel *c);
const char *bf_channel_last_error(const bf_channel *c);
/* gamma: N floats in [0, 1], cell order; wall rows are ignored. */
bf_status bf_channel_set_porosity(bf_channel *c, const float *gamma);
bf_status bf_channel_set_inlet_velocity(bf_channel *c, float u);
/* Forward run of `steps` ticks from rest; *J = mean outlet u_x. */
bf_status bf_channel_forward(bf_channel *c, int64_t steps, double *J);
/* J, dJ/dgamma (N floats) and dJ/d inlet_velocity by the adjoint. */
bf_status bf_channel_gradient(bf_channel *c, const bf_adjoint_config *cfg, double *J, float *dJ_dgamma,
                              double *dJ_duin);
/* Pre-collision (u_x, u_y, u_z, rho) per cell after the last run (4N floats). */
bf_status bf_channel_read_moments(bf_channel *c, float *out, size_t bytes);
/* Device times of the last run in ms: [0] forward (with checkpoints),
 * [1] recomputation (summed over segments; it runs on a second stream beside
 * the adjoint sweep), [2] whole gradient (0 for a forward-only run). */
bf_status bf_channel_timing(const bf_channel *c, double out[3]);

/* --- checkpoint files ----------------------------------------------------
 * LBM::save_current_tick_macroscopic / save_current_tick_mesoscropic /
 * load_tick_macroscopic / load_tick_mesoscropic (LBM.h:123-126), declared but
 * never implemented by the reference.  One file per save in `directory`
 * (created if missing): tick_<ticks>.bfmeso or tick_<ticks>.bfmacro.
 * Mesoscopic files hold the raw state (populations as stored, thermal
 * populations, pseudopotential buffers, force field, ids, parameters and the
 * object table): loading one into a solver of the same configuration continues
 * the run bit for bit.  Macroscopic files hold the saved fields (u, rho),
 * (F, T), the ids and the external force (constant or per-cell field):
 * loading one re-initialises the populations at equilibrium (bf_init_fields
 * with the saved u, rho and T and the external force, not the saved F, which
 * includes buoyancy and the Shan-Chen force) and sets the tick count.  Loading checks that
 * the file's grid, velocity set, multiphase flag, boundary mode, periodicity
 * (and storage, for mesoscopic files) match the solver. */
bf_status bf_save_mesoscopic(bf_solver *solver, const char *directory);
bf_status bf_save_macroscopic(bf_solver *solver, const char *directory);
bf_status bf_load_mesoscopic(bf_solver *solver, const char *directory, int64_t tick);
bf_status bf_load_macroscopic(bf_solver *solver, const char *directory, int64_t tick);

/* --- iso-surface (marching cubes) ---------------------------------------
 * Replaces MarchingCubes::compute (MarchingCubes/MarchingCubes.cpp:3-48,
 * GLSL/MarchingCubes/marching_cubes.comp:41-98), which the reference never
 * dispatches.  Output is an indexed triangle mesh:
 *   vertices — 3 floats each, one per grid edge whose end values straddle the
 *              threshold, ordered by (grid point, axis x/y/z), linearly
 *              interpolated and placed at (p / resolution - 1/2) * mesh_scale
 *              (the reference's vertex placement, marching_cubes.comp:67-70);
 *   indices  — 3 uint32 per triangle, ordered by (cube, triangle of its case);
 *              points with value >= threshold are inside, and the winding gives
 *              normals toward lower values.
 * Passing NULL for both output arrays only returns the counts.  Too small a
 * capacity returns BF_ERR_INVALID with the needed counts set.  The text of the
 * last failure is in bf_marching_cubes_last_error(). */

/* Worst-case output sizes for a grid (get_vertex_buffer_needed_size_in_bytes /
 * get_index_buffer_needed_size_in_bytes, MarchingCubes.h:17-18). */
bf_status bf_marching_cubes_capacity(const int32_t resolution[3], int64_t *max_vertices, int64_t *max_indices);

/* Iso-surface of a host scalar field (X*Y*Z floats in cell order), on `device`. */
bf_status bf_marching_cubes(int32_t device, const float *volume, const int32_t resolution[3], float threshold,
                            const float mesh_scale[3], float *vertices, int64_t vertex_capacity, uint32_t *indices,
                            int64_t index_capacity, int64_t *vertex_count, int64_t *index_count);

/* Iso-surface of a solver's density as of its last save tick, computed on the
 * device from the resident field (no host copy of the volume). */
bf_status bf_isosurface(bf_solver *solver, float threshold, const float mesh_scale[3], float *vertices,
                        int64_t vertex_capacity, uint32_t *indices, int64_t index_capacity, int64_t *vertex_count,
                        int64_t *index_count);

/* Allocation-free variant of bf_marching_cubes, for callers that cannot
 * allocate device memory (tight serving or rendering loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  `volume` (X*Y*Z
 * floats), `vertices` and `indices` are device memory on `device`; `stream`
 * is a hipStream_t (NULL: the default stream), on which the passes run and
 * which is synchronised before return.  Same output, counts and size query
 * (NULL outputs) as bf_marching_cubes.  The work buffers (point and cube
 * classifications, prefix sums) come from a fixed device array of about
 * 40 MiB that the library holds on each GPU it runs on: at most
 * BF_MC_NOALLOC_MAX_POINTS grid points (BF_ERR_INVALID beyond), and calls
 * that overlap on one GPU race.  Calls no hipMalloc or hipFree. */
#define BF_MC_NOALLOC_MAX_POINTS 4194304
bf_status bf_marching_cubes_noalloc(int32_t device, const float *volume, const int32_t resolution[3], float threshold,
                                    const float mesh_scale[3], float *vertices, int64_t vertex_capacity,
                                    uint32_t *indices, int64_t index_capacity, int64_t *vertex_count,
                                    int64_t *index_count, void *stream);

/* The case table in use: triangle_count[256] and edges[256 * 15] (edge ids of
 * each triangle's corners, -1 padded).  Corner k of a cube sits at
 * (k & 1, k >> 1 & 1, k >> 2 & 1); edge 4a + j runs along axis a from the j-th
 * corner without bit a.  The case index has bit k set when corner k is inside. */
bf_status bf_marching_cubes_table(uint8_t *triangle_count, int8_t *edges);

const char *bf_marching_cubes_last_error(void);

/* Device time (ms, HIP events) of the last extraction's passes on this thread:
 * classification, prefix sums and the two emit passes (without transfers). */
double bf_marching_cubes_last_device_ms(void);

/* --- mesh voxeliser -------------------------------------------------------
 * Replaces MeshRasterizer3D::rasterize (MeshRasterizer3D/MeshRasterizer3D.cpp:
 * 7-100): marks the voxels inside a closed triangle mesh with `value` in
 * `volume` (X*Y*Z bytes, cell order z*X*Y + y*X + x, updated in place, other
 * voxels untouched), e.g. boundary ids for bf_init_fields.  `model` is a
 * column-major 4x4 matrix (glm::mat4 layout) applied to the vertices; the grid
 * spans the cube [-1, 1]^3 of model space as the reference's orthographic
 * camera does: voxel (i, j, k) covers x in [-1 + 2i/X, -1 + 2(i+1)/X], y
 * likewise and z in [1 - 2(k+1)/Z, 1 - 2k/Z] (slice 0 nearest +z).  A voxel is
 * inside when its z range overlaps, along the line through its x-y centre, an
 * interval between the 2m-th and (2m+1)-th crossing of the mesh.
 * *solid_count (optional) receives the number of voxels holding `value`
 * afterwards.  BF_ERR_UNSUPPORTED when a column crosses the mesh more than
 * 64 times. */
bf_status bf_voxelize_mesh(int32_t device, const float *vertices, int64_t vertex_count, const uint32_t *indices,
                           int64_t triangle_count, const float model[16], const int32_t resolution[3], uint8_t value,
                           uint8_t *volume, int64_t *solid_count);

const char *bf_voxelize_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* BF_BF_H */
