// mtlp_kernels.h -- launch interface between the C++ engine (engine.cpp,
// adjoint.cpp) and the gfx950 kernels (hash.hip, scan.hip, solver.hip,
// adjoint.hip).
//
// Every kernel works on a batch of independent scenes laid end to end: scene
// s owns particles [p_off[s], p_off[s+1]) in original order AND sorted slots
// [p_off[s], p_off[s+1]) in cell order (global cell keys are
// cell_off + local key, so a scene's cells, and therefore its slots, are
// contiguous), plus cells [cell_off, cell_off + C_s).  A single scene is the
// batch of one.  Constants live in a device table of SceneDev, one per scene;
// nothing is shared mutable state between steps (the reference rewrites a
// Shared PBFParams buffer every step, Engine.cpp:551-572).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mtlp {

constexpr uint32_t kBlock = 256;          // particle-kernel workgroup (4 waves)
constexpr uint32_t kScanBlock = 512;      // scan workgroup (8 waves)
constexpr uint32_t kScanItems = 32;       // elements per scan thread (8 x uint4) ...
constexpr uint32_t kScanItemsLarge = 64;  // ... and from kScanLargeFrom elements on (16 x uint4)
constexpr uint32_t kScanLargeFrom = 1u << 22;
constexpr uint32_t kScanTile = kScanBlock * kScanItems;  // 16384, the smaller tile (sizes the look-back words)
constexpr uint32_t kSmallCell = 256;      // cells up to this size are ranked in place
constexpr uint32_t kLdsSortMax = 16384;   // cells up to this size: LDS bitonic sort
constexpr uint32_t kBigBlock = 1024;      // workgroup of the big-cell kernel
constexpr uint32_t kBigGrid = 128;        // workgroups of the big-cell kernel
// Cells above kLdsSortMax ("huge") are ranked in index order: the particle
// indices are cut into kHugeBlocks blocks of ceil(n / kHugeBlocks); the
// scatter counts each huge cell's members per block into the cell's first
// kHugeBlocks slots of hcnt, the ranking pass turns those counts into the
// blocks' starting ranks (hoff), and one wave of the big-cell kernel per
// block places the block's members in index order.  Linear in n however
// many huge cells there are.
constexpr uint32_t kHugeBlocks = kBigGrid * (kBigBlock / 64);  // 2048, one per wave of the big-cell kernel
static_assert(kHugeBlocks <= kLdsSortMax, "a huge cell holds its block column");
__host__ __device__ inline uint32_t huge_block_len(uint32_t n) { return (n + kHugeBlocks - 1u) / kHugeBlocks; }
constexpr uint32_t kChunkSlack = 32;      // readable slots past the end of every sorted position buffer
// Largest packed (x, y, z) copy (slots incl. slack): its 12-byte slots are
// addressed with 32-bit byte offsets.  Larger engines run without copies.
constexpr uint32_t kMaxXyzSlots = 0xffffffffu / 12u;
constexpr uint32_t kNumXcd = 8;           // XCDs of an MI355X; workgroup b runs on XCD b % kNumXcd

// Error flag bits (device word, read by mtlp_error_flags).
constexpr uint32_t kErrScanLookback = 1u;  // a scan tile gave up waiting
constexpr uint32_t kErrHashRange = 2u;     // cell offsets inconsistent with n
constexpr uint32_t kErrAdjointRange = 4u;  // the adjoint left the FP32 range (inf / NaN)

// Grid of the spatial hash (Shared.h:40-49).
struct GridDev {
    float ox, oy, oz;     // origin = bounds_min
    float inv;            // 1 / cell_size (FP32)
    uint32_t dx, dy, dz;  // cells per axis
    uint32_t dxy;         // dx * dy
};

// Solver constants (writePBFParams, Engine.cpp:551-572; Constraints.metal).
// The density sum runs in FP64 (poly6_d, h2_d, w0_d, inv_rho0_d): rho / rho0
// - 1 cancels to a few per cent near rest, which FP32 resolves only to
// ~1e-5 relative in lambda.  The delta pass runs its per-neighbour terms in
// FP64 too: with a strong s_corr the terms cancel to a small sum that the
// 0.3 h cap then rescales, which turns FP32 term rounding into a direction
// error.
struct SolverDev {
    double poly6_d, h2_d, w0_d, inv_rho0_d;
    double h_d, spiky_d, inv_wq_d, scorr_k_d, max_step_d;
    float h, h2, spiky, inv_rho0, eps;
    uint32_t scorr_n;
    float bminx, bminy, bminz, bmaxx, bmaxy, bmaxz;
};

// dt-independent finalizeStep constants (Constraints.metal:305-353);
// 1/dt and the speed cap are formed in the kernel from the step's dt.
struct FinalDev {
    float bminx, bminy, bminz, bmaxx, bmaxy, bmaxz;
    float damping, weps, floor_f, wall_f;
    float vmax_h;         // vmax * h (the cap is vmax_h / dt, Constraints.metal:347)
    float foam_max_nbr;   // neighbour count of zero surface-ness (Engine.cpp:27)
};

struct SceneDev {
    SolverDev sd;
    GridDev g;
    FinalDev fd;
    float gx, gy, gz;     // gravity
    uint32_t cell_off;    // first global cell key of the scene
};

// Kinematic sphere of the sphere-drop scene (SphereParams, Shared.h:78-81):
// solver iterates inside it are pushed to its surface.
struct SphereDev {
    float cx, cy, cz, r;
    uint32_t on;
};

// Step-wide inputs shared by every scene.
struct StepArgs {
    const SceneDev* scenes;
    const uint32_t* p_off;  // n_scenes + 1 particle offsets
    uint32_t n_scenes;
    uint32_t n;             // total particles
    float dt;
    uint32_t predict;       // 0: hash positions as they are (buildSpatialHash)
    SphereDev sphere;       // single-scene engines only
};

// Buffers of one spatial-hash build.
struct HashBufs {
    const float4* x;      // positions (original order)
    const float4* v;      // velocities (original order)
    uint32_t* keys;       // global cell key per particle
    uint32_t* tickets;    // arrival rank within the cell from the count atomics
    uint32_t* counts;     // C + 1 counters, all zero on entry, zero on exit
    uint32_t* cell_start; // C + 1
    uint2* tmp;           // (index, key) grouped by cell, unordered within a cell
    float4* stage;        // binned positions in tmp's order (may be a free iterate buffer)
    uint32_t* sorted_index;
    float4* sorted;       // gathered (predicted) positions
    uint32_t* big_list;   // cells larger than kSmallCell
    uint32_t* big_count;
    uint32_t* err;
    uint32_t* slot_of;    // sorted slot of each particle (original order); may be NULL
    uint32_t* heavy_count; // 4 heavy-queue counters, cleared by the build; may be NULL
    float* xyz = nullptr;  // packed (x, y, z) of `sorted`, 12 bytes a slot; may be NULL
    uint32_t* hcnt = nullptr;      // n: huge cells' member counts per index block, zero between builds
    uint32_t* hoff = nullptr;      // n: their starting ranks per index block
    uint32_t* huge_flag = nullptr; // set when this build has a huge cell
    uint32_t* scan_words = nullptr; // look-back words of the scan that follows (cleared by the predict pass)
    uint32_t n_scan_words = 0;
};

// ---- launchers (return hipGetLastError after the launches) -------------
hipError_t launch_predict_hash_count(const HashBufs& b, const StepArgs& a, hipStream_t s);
hipError_t launch_scatter_gather(const HashBufs& b, const StepArgs& a, hipStream_t s);
hipError_t launch_count_neighbors(const float4* sorted, const uint32_t* cell_start,
                                  const uint32_t* sorted_index, const StepArgs& a, uint32_t* nbr,
                                  hipStream_t s);

// Exclusive scan of n u32 (in is zeroed after reading when zero_input).
// state/ticket: look-back words of scan_tiles(n) tiles plus one ticket
// counter; words are tagged with `epoch`, so a launch only needs them clear
// if it reuses an epoch (graph replays clear them with launch_fill_u32).
uint32_t scan_tiles(uint32_t n);
hipError_t launch_exclusive_scan(uint32_t* in, uint32_t* out, uint32_t n, bool zero_input,
                                 uint64_t* state, uint32_t* ticket, uint32_t epoch,
                                 uint32_t* err, hipStream_t s);
hipError_t launch_fill_u32(uint32_t* p, uint32_t value, size_t count, hipStream_t s);
// The allocation-free scan's working memory on the current device (a
// file-scope device array in scan.hip): 4 u32 words, then
// kNoAllocScanMax / kScanTile look-back words.
constexpr uint32_t kNoAllocScanMax = 1u << 28;
hipError_t noalloc_scan_scratch(uint64_t** out);

// Neighbour list of one solver iterate: the density pass records each
// slot's neighbours in its stencil-walk order, and the passes that follow
// on the same iterate (delta, and the adjoint's) walk the list instead of
// testing the 27-cell stencil's candidates again; the order is kept, so
// their sums are bit-identical either way.  Entries go in blocks of four
// (blk[b * stride + s] holds entries 4b..4b+3 of slot s, the unused words
// of the last block hold s), written by the lanes of a wave together;
// count[s] may exceed kMaxList, and then readers walk the stencil.
//
// The density pass walks the stencil in chunks of at most 32 consecutive
// candidate slots, (c0, mask) with bit k set when slot c0 + k is a
// neighbour, and keeps only the chunks with a neighbour (walk order): the
// first kLdsChunks of them in LDS (8 bytes each per thread) and kept chunks
// kLdsChunks .. kChunkScratch - 1 in spill[k * stride + s], before running
// the terms over them.  A slot with more than kMaxList neighbours gets count
// kCountChunks | K and all its K <= kChunkScratch kept chunks in spill,
// which the delta pass right after walks (the scratch is rewritten by the
// next density pass); more kept chunks than that (none in practice) gives
// kCountStencil, and readers walk the stencil.
constexpr uint32_t kMaxList = 96;
constexpr uint32_t kListBlocks = kMaxList / 4;
// 15 rather than 16: with 7,680 bytes of LDS per one-wave workgroup the LDS
// lets 21 workgroups share a CU (the kernel's VGPRs then allow 20), with
// 8,192 bytes 18 (measured).
constexpr uint32_t kLdsChunks = 15;
constexpr uint32_t kChunkScratch = 32;
//
// Such slots ("heavy": their walk is several times the typical one) also
// go to a queue, marked kCountQueued in the count, and the delta pass's
// ordinary workgroups skip them (one heavy lane in an ordinary wave would
// keep it running long after the rest of the pass).  Slots left to the
// stencil (more than kChunkScratch chunks) go to a queue of up to kHeavyCap
// whose slots get a whole wave each, 64 candidates at a time; slots whose
// chunks are in the scratch go to a queue of up to lanes_cap whose slots the
// delta pass's heavy-lane workgroups walk one per lane, so that those waves
// hold heavy slots only (a short queue goes to the heavy waves instead).  A slot that finds its queue full stays with its
// ordinary wave.  The queues are double-buffered by iteration parity: the
// delta pass of one iteration clears the counters the next density pass
// appends to.
constexpr uint32_t kCountChunks = 0x80000000u;   // | K: chunks in spill
constexpr uint32_t kCountQueued = 0x40000000u;   // in the heavy queue
constexpr uint32_t kCountStencil = kCountChunks | 0xffffu;
constexpr uint32_t kHeavyCap = 16384;
// Heavy-lane queue entries per iteration parity for n slots: n / 8, a
// multiple of 4,096 (so its workgroups are a multiple of 8 at every stencil
// workgroup size and keep the XCD mapping of the workgroups after them).
inline uint32_t lanes_cap_for(size_t n) { return (uint32_t)(((n >> 3) + 4095u) / 4096u * 4096u); }
struct NbrList {
    uint4* blk;       // kListBlocks * stride
    uint32_t* count;  // stride
    uint32_t stride;
    uint2* spill;     // kChunkScratch * stride: density-pass chunk scratch
    uint32_t* heavy;             // kHeavyCap queued slots of this iteration
    uint32_t* heavy_count;       // their number (may exceed kHeavyCap)
    uint32_t* heavy_count_next;  // the next iteration's counter
    uint32_t* lanes;             // lanes_cap queued heavy-lane slots of this iteration
    uint32_t* lanes_count;       // their number (may exceed lanes_cap)
    uint32_t* lanes_count_next;  // the next iteration's counter
    uint32_t lanes_cap;          // lanes_cap_for(stride)
    // Tile order: each density pass times its tiles (one stencil workgroup
    // of slots each) into cost buckets; the delta pass that follows turns
    // them into the order in which the next iteration's passes take the
    // tiles of each XCD's range, heaviest buckets first and slot order
    // inside a bucket, so that their last workgroups are short ones.  The
    // results do not depend on the order.
    uint32_t* tile_plan;        // per tile: its cost bucket
    const uint32_t* tile_order; // both passes: tile at each position (NULL: identity)
    uint32_t* tile_order_next;  // delta pass: writes the next iteration's order here (NULL: none)
    // Persistent launches (pull > 0): a fixed number of resident workgroups
    // take the positions of their XCD's range `pull` tiles at a time from
    // kTileQueues counters per XCD, instead of one workgroup per tile in
    // dispatch order.  tile_ctr holds kTileCtrs words for the density pass
    // (a counter every kTileCtrStride words), then kTileCtrs for the delta
    // pass; each persistent launch zeroes the other pass's, which no launch
    // reads until the next one of that pass.
    uint32_t* tile_ctr = nullptr;
    uint32_t pull = 0;
    uint32_t pull_waves = 0;  // waves of a persistent launch (host side)
    // Packed (x, y, z) copies of the iterates, 12 bytes a slot with
    // kChunkSlack slots of read slack: the density pass's stencil filter
    // reads its candidates from xyz (eight of them in six 16-byte loads
    // instead of eight 12-byte ones), and the delta pass writes the next
    // iterate's copy to xyz_next beside it.  NULL: the filter reads the
    // float4 iterate.
    const float* xyz = nullptr;
    float* xyz_next = nullptr;
};

constexpr uint32_t kCostBuckets = 32;
constexpr uint32_t kCostTicks = 500;  // bucket width: 5 us of the 100 MHz real-time clock
constexpr uint32_t kCostSpread = 8;   // a range whose heaviest tiles are at most this many buckets
                                      // above its median keeps the slot order

// Workgroup sizes the stencil kernels are built for (the autotuner picks).
constexpr uint32_t kStencilBlocks[] = {64, 128, 256, 512};
constexpr uint32_t kNumStencilBlocks = 4;
// Tiles per counter increment of a persistent launch (0: one workgroup per
// tile); the autotuner tries these.
constexpr uint32_t kTilePulls[] = {0, 1, 4};
constexpr uint32_t kNumTilePulls = 3;
constexpr uint32_t kMaxTilePull = 64;
constexpr uint32_t kTileQueues = 16;     // per XCD
constexpr uint32_t kTileCtrStride = 32;  // u32 words: one 128-byte line per counter
constexpr uint32_t kTileCtrs = kNumXcd * kTileQueues * kTileCtrStride;  // words per pass
// Waves a persistent stencil launch keeps resident per CU.
constexpr uint32_t kPersistentWavesPerCu = 16;
// Workgroups of a stencil launch of `tiles` tiles of kB threads: one per
// tile, or (persistent) nl.pull_waves waves' worth rounded up to a multiple
// of kNumXcd x kTileQueues, so every queue of every XCD has workgroups
// (a queue without one would leave its positions undone).
inline uint32_t stencil_grid(uint32_t tiles, uint32_t kB, const NbrList& nl) {
    if (!nl.pull) return tiles;
    constexpr uint32_t q = kNumXcd * kTileQueues;
    const uint32_t fill = nl.pull_waves / (kB / 64u);
    const uint32_t g = tiles < fill ? tiles : fill;
    return (g + q - 1u) / q * q;
}

// With nl.blk non-NULL the density pass writes the list and the delta pass
// reads it.
hipError_t launch_density_lambda(float4* cur, const uint32_t* cell_start,
                                 const uint32_t* sorted_index, const StepArgs& a,
                                 float* lambda_orig, uint32_t* nbr_orig, const NbrList& nl,
                                 uint32_t block, hipStream_t s);
hipError_t launch_apply_delta(const float4* cur, const uint32_t* cell_start, const StepArgs& a,
                              float4* next, const NbrList& nl, uint32_t block, hipStream_t s);
// Scatter back + finalize in original order: particle i takes its result
// from slot slot_of[i]; with foam non-NULL also computeFoam from the step's
// neighbour counts (nbr, original order); with lam_in_w (a step with
// solver iterations: the last delta pass leaves each slot's lambda in .w of
// the final iterate) it also writes lambda_orig.
hipError_t launch_finalize(const float4* final_sorted, const uint32_t* slot_of,
                           const StepArgs& a, float4* x, float4* v, float4* prev,
                           const uint32_t* nbr, float* foam, bool lam_in_w, float* lambda_orig,
                           hipStream_t s);
// integrate (Integrate.metal:11-42) on n particles; bounds from b.
hipError_t launch_integrate(float4* x, float4* v, uint32_t n, float dt, float gy, float damping,
                            const FinalDev& b, hipStream_t s);
// applyMousePulse (Constraints.metal:222-253) on n particles.
hipError_t launch_mouse_pulse(const float4* x, float4* v, uint32_t n, const float origin[3],
                              const float dir[3], const float force[3], float radius, hipStream_t s);

// ---- adjoint (adjoint.hip) ---------------------------------------------
// Lanes per particle of the backward stencil kernels, and the particle
// count their 32-bit thread numbers allow.
constexpr uint32_t kAdjSub = 4;
constexpr uint32_t kMaxAdjointParticles = 1u << 28;
// What the adjoint of one step reads from its taped forward: the hash, the
// iterates p^0 .. p^I (sorted float4; p^k.w = lambda^k for k < I), per
// iteration the density pass's gradient sums (sum_j grad W, D) and the
// delta pass's R (before the 0.3 h cap), and the step's input positions.
struct StepTape {
    const uint32_t* sorted_index;
    const uint32_t* cell_start;
    const float4* x_in;      // original order
    float4* iter;            // I + 1 arrays of `stride` float4
    float4* grad;            // I arrays
    float4* dlt;             // I arrays
    uint4* lblk;             // I neighbour lists (kListBlocks * stride each)
    uint32_t* lcnt;          // I count arrays
    uint2* spill;            // the density pass's chunk scratch (not part of the tape)
    uint32_t* heavy;         // 2 x kHeavyCap heavy-slot queues (iteration parity)
    uint32_t* heavy_count;   // 4 counters: heavy queue, heavy-lane queue, by parity
    uint32_t* lanes;         // 2 x lanes_cap_for(stride) heavy-lane queues
    size_t stride;
    float4* iterate(uint32_t k) const { return iter + k * stride; }
    float4* grad_sums(uint32_t k) const { return grad + k * stride; }
    float4* delta(uint32_t k) const { return dlt + k * stride; }
    NbrList list(uint32_t k) const {
        return NbrList{lblk + (size_t)k * kListBlocks * stride, lcnt + k * stride, (uint32_t)stride, spill,
                       heavy + (k & 1u) * kHeavyCap, heavy_count + (k & 1u), heavy_count + ((k + 1u) & 1u),
                       lanes + (size_t)(k & 1u) * lanes_cap_for(stride), heavy_count + 2u + (k & 1u),
                       heavy_count + 2u + ((k + 1u) & 1u), lanes_cap_for(stride)};
    }
};

// Remainder terms (a pair's share for a slot that cannot see the pair) go
// to the receiving slot's inbox: kInbox entries (lambda-bar, pbar.xyz) per
// slot, appended at positions handed out by an atomic counter and added up
// by the consuming stage in an order-independent way, so the adjoint's bits
// do not depend on the order of the atomics.  Entries past kInbox go to
// float atomics in prem / lrem instead (then the order matters again).
// Each such pair sends two terms per iteration (delta and density passes):
// the 64K differentiable workload needs at most 12 entries per pass, the
// 1M dam break with the benchmark constants up to 32 in its first 15 steps
// (FP32 oracle trajectories).
constexpr uint32_t kInbox = 32;

// Work buffers of the backward pass (sorted order unless noted).  The
// inbox counters, prem and lrem must be zero between steps (the pointwise
// stages that consume them clear them).
struct AdjointBufs {
    float4* pacc;   // adjoint of the current iterate, own part
    float4* prem;   // its remainder part past the inbox
    float4* rbar;   // adjoint of R
    float4* db;     // (2 qbar g, qbar)
    float* lbar;    // adjoint of lambda, own part
    float* lrem;    // its remainder part past the inbox
    double* gdev;   // 6 per slot: deviation terms of the gravity recursion
    double* gpart;  // 6 per block of kBlock slots
    double* gstate; // 9: sum xbar, sum vbar, the gravity gradient so far
    uint32_t* err;  // error flag word
    uint32_t* rcnt; // per slot: inbox entries appended (may exceed kInbox)
    float4* rval;   // kInbox entries per slot (slot-major)
};

// lambda_orig / nbr_orig as in launch_density_lambda (may be NULL).
hipError_t launch_density_lambda_tape(float4* cur, const uint32_t* cell_start, const uint32_t* sorted_index,
                                      const StepArgs& a, float4* aux, float* lambda_orig, uint32_t* nbr_orig,
                                      const NbrList& nl, uint32_t block, hipStream_t s);
hipError_t launch_apply_delta_tape(const float4* cur, const uint32_t* cell_start, const StepArgs& a,
                                   float4* next, float4* rr, const NbrList& nl, uint32_t block, hipStream_t s);
hipError_t launch_finalize_tape(const float4* fin, const uint32_t* sorted_index, const StepArgs& a,
                                const float4* x_in, float4* x_out, float4* v_out, hipStream_t s);
// Starts the gravity recursion from the adjoint inputs (xbar_T, vbar_T).
hipError_t launch_gravity_init(const AdjointBufs& b, const float4* xbar_T, const float4* vbar_T, uint32_t n,
                               float* gout, hipStream_t s);
// (xbar_out, vbar_out) of one step -> (xbar_in, vbar_in); the gravity
// gradient so far goes to b.gstate and, as float, to gout[3].
hipError_t launch_step_vjp(const StepTape& t, const AdjointBufs& b, const StepArgs& a, uint32_t iterations,
                           float damping, const float4* xbar_out, const float4* vbar_out, float4* xbar_in,
                           float4* vbar_in, float* gout, uint32_t block, hipStream_t s);
// Density at the positions of a fresh spatial hash (sorted, cell_start,
// sorted_index), original order, and its adjoint (sorted.w is overwritten).
hipError_t launch_density_rho(const float4* sorted, const uint32_t* cell_start, const uint32_t* sorted_index,
                              const StepArgs& a, float* rho, hipStream_t s);
hipError_t launch_density_rho_vjp(float4* sorted, const uint32_t* cell_start, const uint32_t* sorted_index,
                                  const StepArgs& a, const float* rhobar, float4* xbar, hipStream_t s);

}  // namespace mtlp
