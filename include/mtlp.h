/*
 * mtlp.h -- C ABI of the MI355X (gfx950) PBF engine.
 *
 * Drop-in boundary for the reference's `class mtlphys::Engine`
 * (engine/include/mtlphys/Engine.hpp:27-100).  Every entry point below
 * names the reference member it replaces.  Conventions:
 *
 *   ownership  The engine owns its device buffers.  mtlp_get_buffer returns
 *              borrowed device pointers that stay valid (same address) until
 *              mtlp_reset or mtlp_destroy, like the reference's
 *              MTL::Buffer accessors (Engine.hpp:83-90).  Scan scratch is
 *              caller-owned, like runExclusiveScan's chunkSums
 *              (Engine.hpp:96-100).
 *   layout     Positions and velocities are float4 AoS: pos.w = 1,
 *              vel.w = 0 (Constraints.metal:351-352).  Buffers are padded
 *              to a multiple of 256 elements (Engine.cpp:383).
 *   streams    Work is enqueued on the caller's stream (a hipStream_t passed
 *              as mtlp_stream; NULL = the null stream), the analogue of the
 *              caller's MTL::CommandBuffer (Engine.hpp:52-56).  Calls return
 *              after enqueueing; the caller synchronises.  Only the
 *              *_read_* / *_write_* / *_upload_* helpers and
 *              mtlp_error_flags block.
 *   errors     Status codes, never abort.  Conditions on which the reference
 *              aborts (Engine.cpp:69-85, 407-413, 849-854) return
 *              MTLP_E_RANGE / MTLP_E_HIP; its silent no-ops (dt <= 0,
 *              Engine.cpp:686; n == 0, :848) return MTLP_OK.
 *   threads    One engine per host thread (Engine.hpp:4-6).
 */
#ifndef MTLP_H
#define MTLP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mtlp_engine mtlp_engine;
typedef void* mtlp_stream; /* hipStream_t */

typedef enum mtlp_status {
    MTLP_OK = 0,
    MTLP_E_INVALID = 1, /* null handle/pointer, bad enum, misaligned buffer  */
    MTLP_E_RANGE = 2,   /* size beyond a limit (e.g. grid, scan length)      */
    MTLP_E_NOMEM = 3,   /* device or host allocation failed                  */
    MTLP_E_HIP = 4,     /* a HIP runtime call failed (see mtlp_last_error)   */
    MTLP_E_DEVICE = 5   /* a kernel raised an error flag (mtlp_error_flags)  */
} mtlp_status;

/* Scene ids.  0-2 are the reference's (Engine.hpp:37-39, seeding at
 * Engine.cpp:454-549).  3-4 are the benchmark workloads: a 32^3 cube drop
 * at h = 0.1 and a 1M-particle dam break at h = 0.02, with seeded jitter
 * (their parameters come from mtlp_scene_params). */
enum {
    MTLP_SCENE_CUBE_DROP = 0,
    MTLP_SCENE_DAM_BREAK = 1,
    MTLP_SCENE_SPHERE_DROP = 2, /* the lake only; the sphere is not modelled */
    MTLP_SCENE_JITTER_CUBE = 3,
    MTLP_SCENE_JITTER_DAM = 4,
    MTLP_SCENE_EMPTY = 255      /* positions and velocities zeroed           */
};

/* Simulation parameters.  mtlp_default_params gives the reference values:
 * Engine.cpp:23-39 (bounds, cell size, h, rho0, epsilon, damping,
 * iterations), Engine.cpp:558 (gravity), Constraints.metal:158-160
 * (s_corr k, n, dq), :190 (0.3 h step cap), :325 (wall band), :340-345
 * (friction), :347 (speed cap h/dt). */
typedef struct mtlp_params {
    float gravity[3];
    float bounds_min[3];
    float bounds_max[3];
    float h;              /* smoothing length                               */
    float cell_size;      /* hash cell edge (the reference uses h)          */
    float rest_density;
    float epsilon;        /* constraint-force mixing                       */
    float scorr_k;
    float scorr_dq;       /* delta q / h                                    */
    float max_step;       /* |delta x| cap / h                              */
    float damping;        /* velocity factor at finalize                    */
    float wall_eps;       /* wall band width                                */
    float floor_friction; /* tangential factor on the floor                 */
    float wall_friction;  /* factor on side walls                           */
    float vmax;           /* speed cap / (h/dt)                             */
    uint32_t scorr_n;     /* s_corr exponent                                */
    uint32_t iterations;  /* solver iterations per step                     */
} mtlp_params;

/* Buffer ids = the reference accessors (Engine.hpp:83-90), plus the
 * sorted-order positions of the last step (gatherPositions' output,
 * Spatial.metal:190-198, holding the final solver iterate after a step). */
typedef enum mtlp_buffer {
    MTLP_BUF_POSITIONS = 0,       /* float4[n]                              */
    MTLP_BUF_VELOCITIES = 1,      /* float4[n]                              */
    MTLP_BUF_PREV_POSITIONS = 2,  /* float4[n], positions before the step   */
    MTLP_BUF_CELL_HASHES = 3,     /* uint32[n]                              */
    MTLP_BUF_CELL_START = 4,      /* uint32[total_cells + 1]                */
    MTLP_BUF_SORTED_INDEX = 5,    /* uint32[n]                              */
    MTLP_BUF_NEIGHBOR_COUNTS = 6, /* uint32[n]                              */
    MTLP_BUF_LAMBDAS = 7,         /* float[n], last iteration, orig. order  */
    MTLP_BUF_SORTED_POSITIONS = 8, /* float4[n]; .w of a step's final iterate
                                     holds that iterate's lambda input     */
    MTLP_BUF_FOAM_INTENSITY = 9   /* float[n], computeFoam of the last step
                                     (the reference's private _foamIntensity,
                                     Engine.hpp:177)                       */
} mtlp_buffer;

/* Largest scan length (Engine.hpp:93 caps the reference at 1024^2; the
 * single-pass scan here has no such cap below 2^31). */
#define MTLP_MAX_SCAN_LENGTH 0x7fffffffu
/* Largest particle count per engine. */
#define MTLP_MAX_PARTICLES 0x40000000u

const char* mtlp_status_string(mtlp_status s);
/* Message of the last failing call on this thread ("" if none). */
const char* mtlp_last_error(void);

mtlp_status mtlp_default_params(mtlp_params* out);
/* Parameters a scene id is defined with (defaults for scenes 0-2). */
mtlp_status mtlp_scene_params(uint32_t scene, uint32_t n, mtlp_params* out);
/* Host-only scene generator (Engine::seedParticles, Engine.cpp:454-549):
 * fills x4/v4 (4n floats each).  Needs no GPU. */
mtlp_status mtlp_generate_scene(uint32_t scene, uint32_t n, uint64_t seed,
                                const mtlp_params* p, float* x4, float* v4);
/* Grid of a parameter set (Engine::buildSpatialBuffers, Engine.cpp:398-413):
 * dims = ceil((bounds_max - bounds_min) / cell_size) per axis.  Host only. */
mtlp_status mtlp_grid_dims(const mtlp_params* p, uint32_t dims[3]);

/* Engine(MTL::Device*) (Engine.cpp:97-102).  Allocates nothing until
 * mtlp_reset. */
mtlp_status mtlp_create(int hip_device, const mtlp_params* p, mtlp_engine** out);
/* ~Engine (Engine.cpp:104-123) */
mtlp_status mtlp_destroy(mtlp_engine* e);
/* Engine::reset(n, scene) (Engine.cpp:129-134): (re)allocates every buffer
 * and seeds the scene; synchronous. */
mtlp_status mtlp_reset(mtlp_engine* e, uint32_t n, uint32_t scene, uint64_t seed);
/* Replaces the parameters (the reference writes PBFParams every step,
 * Engine.cpp:551-572).  A change of bounds or cell size that changes the
 * grid reallocates the cell buffers; synchronous. */
mtlp_status mtlp_set_params(mtlp_engine* e, const mtlp_params* p);
mtlp_status mtlp_get_params(const mtlp_engine* e, mtlp_params* out);
/* Replaces the reference's host writes into the Shared position and
 * velocity buffers (tests/test_pbf.cpp:197-200).  Either pointer may be
 * NULL.  Synchronous. */
mtlp_status mtlp_upload_state(mtlp_engine* e, const float* x4, const float* v4, uint32_t n);
/* The same from / to device memory (n float4 each, original order),
 * asynchronous on `stream`; either pointer may be NULL.  For callers whose
 * state lives in their own device tensors. */
mtlp_status mtlp_set_state_device(mtlp_engine* e, mtlp_stream stream, const float* d_x4, const float* d_v4);
mtlp_status mtlp_get_state_device(mtlp_engine* e, mtlp_stream stream, float* d_x4, float* d_v4);

/* Engine::step(cmd, dt) (Engine.cpp:685-811): predict, spatial hash,
 * iterations x (density/lambda, delta[, sphere collision]), scatter back,
 * finalize, foam.  dt <= 0 is a no-op returning MTLP_OK. */
mtlp_status mtlp_step(mtlp_engine* e, mtlp_stream stream, float dt);
/* Engine::integrate(cmd, dt) (Engine.hpp:48, Engine.cpp:578-595,
 * Integrate.metal:11-42): the legacy semi-implicit Euler step with
 * gravity -9.81, damping 0.999 and the box with restitution 0.4; on
 * positions and velocities in place (no spatial hash). */
mtlp_status mtlp_integrate(mtlp_engine* e, mtlp_stream stream, float dt);
/* Engine::applyMousePulse(cmd, rayOrigin, rayDir, force, radius)
 * (Engine.hpp:61-65, Engine.cpp:813-841, Constraints.metal:222-253): adds
 * force * exp(-d^2 / (0.4 radius^2)) to the velocity of every particle in
 * front of the ray origin within `radius` of the ray (ray_dir normalised). */
mtlp_status mtlp_apply_mouse_pulse(mtlp_engine* e, mtlp_stream stream, const float ray_origin[3],
                                   const float ray_dir[3], const float force[3], float radius);
/* The sphere-drop scene's kinematic sphere (Engine.hpp:180-184).  mtlp_reset
 * with MTLP_SCENE_SPHERE_DROP switches it on at rest at (0, 4.5, 0), radius
 * 0.6 (Engine.cpp:522-527); each mtlp_step then integrates it on the host
 * (Engine.cpp:693-712) and pushes solver iterates out of it after every
 * delta pass (sphereCollision, Constraints.metal:262-277).  Steps with the
 * sphere on are not replayed from a HIP graph, and the rollout adjoint
 * rejects them.  NULL arguments leave that part unchanged. */
mtlp_status mtlp_set_sphere(mtlp_engine* e, int enabled, const float center[3], const float velocity[3],
                            float radius);
mtlp_status mtlp_get_sphere(const mtlp_engine* e, int* enabled, float center[3], float velocity[3],
                            float* radius);
/* Engine::buildSpatialHash(cmd) (Engine.cpp:601-679) on the current
 * positions: hash, count, scan, stable scatter, gather, neighbour count. */
mtlp_status mtlp_build_spatial_hash(mtlp_engine* e, mtlp_stream stream);

/* Engine::runExclusiveScan(cmd, in, out, chunkSums, n) (Engine.cpp:843-882).
 * d_in and d_out are device pointers, 16-byte aligned, n elements; d_scratch
 * holds mtlp_scan_scratch_bytes(n) bytes.  n == 0 is a no-op; the engine may
 * be NULL. */
size_t mtlp_scan_scratch_bytes(uint32_t n);
mtlp_status mtlp_exclusive_scan(mtlp_engine* e, mtlp_stream stream, const uint32_t* d_in,
                                uint32_t* d_out, void* d_scratch, uint32_t n);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): mtlp_exclusive_scan without d_scratch and without an
 * engine.  Its working memory is a fixed array in the library's device
 * memory (128 KiB per GPU the library runs on, held until the process
 * ends), cleared by a kernel at the start of every call; the call allocates
 * nothing.  n <= MTLP_NOALLOC_SCAN_MAX, else MTLP_E_RANGE; d_in and d_out as
 * for mtlp_exclusive_scan.  Every call on one GPU shares that array: calls
 * that overlap on one GPU (on different streams, or from several threads)
 * race and give wrong results; calls in one stream's order do not. */
#define MTLP_NOALLOC_SCAN_MAX 0x10000000u
mtlp_status mtlp_exclusive_scan_noalloc(mtlp_stream stream, const uint32_t* d_in, uint32_t* d_out, uint32_t n);

/* Accessors (Engine.hpp:78-90). */
mtlp_status mtlp_get_buffer(mtlp_engine* e, mtlp_buffer which, void** d_ptr, size_t* bytes);
/* Synchronous device->host / host->device copies of a buffer (the tests'
 * readBuffer, tests/TestUtils.hpp:82-97).  bytes may be smaller than the
 * buffer. */
mtlp_status mtlp_read_buffer(mtlp_engine* e, mtlp_buffer which, void* host_dst, size_t bytes);
mtlp_status mtlp_write_buffer(mtlp_engine* e, mtlp_buffer which, const void* host_src, size_t bytes);
uint32_t mtlp_particle_count(const mtlp_engine* e);
uint32_t mtlp_total_cells(const mtlp_engine* e);
float mtlp_cell_size(const mtlp_engine* e);

/* Diagnostics (no reference counterpart; the reference's only timing is
 * GPUEndTime - GPUStartTime per command buffer, tests/bench_engine.cpp:43).
 * With profiling on, mtlp_step records HIP events on its stream around each
 * kernel; mtlp_stage_times synchronises and returns the last call's kernel
 * durations in ms with their names. */
mtlp_status mtlp_set_profiling(mtlp_engine* e, int enable);
mtlp_status mtlp_stage_times(mtlp_engine* e, float* ms, const char** names,
                             uint32_t capacity, uint32_t* count);
/* Device error flags raised by kernels since the last call (bit 0: scan
 * look-back gave up waiting; bit 1: cell offsets inconsistent; bit 2: the
 * rollout adjoint left the FP32 range); synchronises and clears them. */
mtlp_status mtlp_error_flags(mtlp_engine* e, uint32_t* flags);

/* ---- Batched ensembles (no reference counterpart: the reference steps one
 * scene per Engine; these entry points step many independent scenes, each
 * with its own particle count, bounds/grid, rest density, gravity and other
 * parameters, with one kernel launch per stage for all of them).  A batch is
 * an mtlp_engine: mtlp_step, mtlp_upload_state, mtlp_get_buffer and friends
 * act on all scenes laid end to end (scene s owns particles
 * [particle_offsets[s], particle_offsets[s+1]) in both original and sorted
 * order, and cells [cell_offsets[s], cell_offsets[s+1])).  Every scene must
 * use the same iteration count. */
mtlp_status mtlp_batch_create(int hip_device, uint32_t n_scenes, const uint32_t* counts,
                              const mtlp_params* per_scene, mtlp_engine** out);
/* Seeds scene s with mtlp_generate_scene(scene, counts[s], seed + s,
 * per_scene[s]); synchronous. */
mtlp_status mtlp_batch_seed(mtlp_engine* e, uint32_t scene, uint64_t seed);
/* Scene count and offsets (arrays of n_scenes + 1 entries, may be NULL). */
mtlp_status mtlp_batch_info(const mtlp_engine* e, uint32_t* n_scenes, uint32_t* particle_offsets,
                            uint32_t* cell_offsets);
mtlp_status mtlp_batch_set_params(mtlp_engine* e, uint32_t scene, const mtlp_params* p);

/* n_steps calls of mtlp_step with the same dt. */
mtlp_status mtlp_step_n(mtlp_engine* e, mtlp_stream stream, float dt, uint32_t n_steps);
/* With graphs on, mtlp_step captures its launches once per dt into a HIP
 * graph and replays it (the look-back words of the scan are cleared by a
 * kernel inside the graph).  Profiling is ignored while graphs are on. */
mtlp_status mtlp_set_graph(mtlp_engine* e, int enable);
/* Times eight steps at each launch configuration of the stencil kernels
 * (workgroup size 64, 128, 256 or 512 times tiles per pull 0, 1 or 4, see
 * mtlp_set_tile_pull, times packed positions on or off, see
 * mtlp_set_packed_positions), two passes over them, from the current state,
 * and keeps the configuration with the fastest pass, unless the engine's
 * current configuration is within 3% of it (then that one stays).
 * Afterwards every buffer a
 * caller can read (positions, velocities, previous positions, cell hashes,
 * cell offsets, sorted index and positions, neighbour counts, lambdas,
 * foam) holds what it held before the call.  With cache_path, reads a
 * previous choice for the same batch shape from that text file or appends
 * the new one.  *block receives the chosen workgroup size (may be NULL). */
mtlp_status mtlp_autotune(mtlp_engine* e, mtlp_stream stream, float dt, const char* cache_path,
                          uint32_t* block);
/* Workgroup size of the stencil kernels (64, 128, 256 or 512). */
mtlp_status mtlp_set_stencil_block(mtlp_engine* e, uint32_t block);
/* Launch mode of the stencil passes (no reference counterpart): 0 (the
 * default) launches one workgroup per tile of slots, taken in dispatch
 * order; k in 1..64 launches persistent workgroups (16 waves per compute
 * unit) that take k tiles at a time from one counter per XCD, each XCD
 * working through its own range of the tile order.  Results are the same
 * bits either way. */
mtlp_status mtlp_set_tile_pull(mtlp_engine* e, uint32_t tiles_per_pull);
/* Packed positions (no reference counterpart; on by default): with
 * neighbour lists on, each iterate also gets a packed 12-byte (x, y, z)
 * copy, written by the spatial hash's gather and by the delta pass, and the
 * density pass filters its stencil candidates from it (eight candidates in
 * six 16-byte loads instead of eight 12-byte ones).  Faster where the
 * stencil filter dominates (the 1M dam break); the autotuner tries both.
 * Results are the same bits either way.  MTLP_E_RANGE when enabling it on
 * an engine of more than 357,913,941 slots, which has no copies. */
mtlp_status mtlp_set_packed_positions(mtlp_engine* e, int enable);
/* The stencil kernels' current workgroup size, tiles per pull and packed
 * positions setting (any pointer may be NULL). */
mtlp_status mtlp_launch_config(const mtlp_engine* e, uint32_t* block, uint32_t* tiles_per_pull,
                               int* packed_positions);
/* Neighbour lists (on by default): each iteration's density pass records
 * up to 96 neighbours per particle, or the stencil chunks that hold them
 * (644 B of device memory per particle), and the delta pass walks them
 * instead of testing the 27-cell stencil's candidates again.  Results are
 * bit-identical with lists on or off. */
mtlp_status mtlp_set_neighbor_lists(mtlp_engine* e, int enable);

/* ---- state I/O (no reference counterpart) ------------------------------
 * A little-endian binary file: "MTLPSTAT", u32 version (1), u32 n_scenes,
 * u32 n (particles), u32 flags (bit 0: batch), u32 scene id, u32 sphere on,
 * f32 sphere centre[3], velocity[3], radius; per scene an mtlp_params (92
 * bytes) and a u32 particle count; then float4 positions[n] and
 * velocities[n] in original order.  mtlp_save_state synchronises and
 * writes the engine's state; mtlp_load_state reads it back: a single-scene
 * engine takes the file's parameters and particle count (reallocating if
 * they differ), a batch needs the same scene counts. */
#define MTLP_STATE_VERSION 1u
mtlp_status mtlp_save_state(mtlp_engine* e, const char* path);
mtlp_status mtlp_load_state(mtlp_engine* e, const char* path);

/* ---- differentiable rollout (single-scene engines) --------------------
 * The reference has no adjoint; these entries have no counterpart in it.
 * The adjoint differentiates mtlp_step as computed, with the neighbour
 * structure held fixed (binning and candidate sets are piecewise
 * constant), the box clamps as pass-through inside the box and 0 outside,
 * the 0.3 h step cap and the h/dt speed cap through their normalisation
 * Jacobians, and the wall responses as piecewise-constant multipliers.
 *
 * Device arrays of particles are float4 (x, y, z, w) in original order;
 * the w of every gradient output is 0 and the w of gradient inputs is
 * ignored. */

/* Steps the engine `steps` times (as mtlp_step) and keeps (x, v) before
 * steps 0, k, 2k, ... (k = ckpt_every) in engine memory for
 * mtlp_adjoint_rollout.  The engine ends at step `steps`.  Engines of more
 * than 2^28 particles get MTLP_E_RANGE. */
mtlp_status mtlp_rollout(mtlp_engine* e, mtlp_stream stream, float dt, uint32_t steps,
                         uint32_t ckpt_every);
/* Adjoint of the last mtlp_rollout (same dt, steps and ckpt_every): from
 * dL/dx_T and dL/dv_T (NULL = 0) to dL/dx_0, dL/dv_0 and dL/dgravity (3
 * floats), each written only if non-NULL.  Each segment of ckpt_every
 * steps is recomputed from its checkpoint with a tape, then walked
 * backwards.  The engine's state (x_T, v_T) is not changed. */
mtlp_status mtlp_adjoint_rollout(mtlp_engine* e, mtlp_stream stream, float dt, uint32_t steps,
                                 uint32_t ckpt_every, const float* d_xbar_T, const float* d_vbar_T,
                                 float* d_xbar_0, float* d_vbar_0, float* d_gbar);
/* Density rho_i = W(0) + sum_j W(|x_i - x_j|) (Constraints.metal:100-115)
 * at positions d_x4 (n float4), with the spatial hash rebuilt from d_x4 in
 * the engine's hash buffers; d_rho receives n floats. */
mtlp_status mtlp_density(mtlp_engine* e, mtlp_stream stream, const float* d_x4, float* d_rho);
/* dL/dx at d_x4 from dL/drho (n floats) for mtlp_density. */
mtlp_status mtlp_density_vjp(mtlp_engine* e, mtlp_stream stream, const float* d_x4, const float* d_rhobar,
                             float* d_xbar);

/* Tape of the taped forward, for inspection.  mtlp_tape_record runs
 * `steps` taped steps from the engine's current state without changing
 * it; mtlp_tape_buffer returns a borrowed device pointer into the tape of
 * step `step` (< steps): */
typedef enum mtlp_tape_kind {
    MTLP_TAPE_POSITIONS = 0,     /* float4, state at the step's start (original order) */
    MTLP_TAPE_VELOCITIES = 1,    /* float4 */
    MTLP_TAPE_SORTED_INDEX = 2,  /* u32 */
    MTLP_TAPE_CELL_START = 3,    /* u32, total_cells + 1 */
    MTLP_TAPE_ITERATE = 4,       /* float4 sorted, iteration 0..I; .w = lambda of that
                                    iteration for iteration < I */
    MTLP_TAPE_GRAD_SUMS = 5,     /* float4 sorted (sum_j grad W, D), iteration < I */
    MTLP_TAPE_DELTA = 6          /* float4 sorted, R before the 0.3 h cap, iteration < I */
} mtlp_tape_kind;
mtlp_status mtlp_tape_record(mtlp_engine* e, mtlp_stream stream, float dt, uint32_t steps);
mtlp_status mtlp_tape_buffer(mtlp_engine* e, mtlp_tape_kind kind, uint32_t step, uint32_t iteration,
                             void** d_ptr, size_t* bytes);
/* Synchronous copy of a tape array to host memory (bytes <= its size). */
mtlp_status mtlp_tape_read(mtlp_engine* e, mtlp_tape_kind kind, uint32_t step, uint32_t iteration,
                           void* host_dst, size_t bytes);

#ifdef __cplusplus
}
#endif

#endif /* MTLP_H */
