// mtlp_kernels.h -- launch interface between the C++ engine (engine.cpp) and
// the gfx950 kernels (hash.hip, scan.hip, solver.hip).  Plain PODs passed to
// kernels by value: no shared mutable uniform buffer (the reference rewrites
// a Shared PBFParams buffer every step, Engine.cpp:551-572).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mtlp {

constexpr uint32_t kBlock = 256;          // particle-kernel workgroup (4 waves)
constexpr uint32_t kScanBlock = 256;      // scan workgroup
constexpr uint32_t kScanItems = 16;       // elements per scan thread (4 x uint4)
constexpr uint32_t kScanTile = kScanBlock * kScanItems;  // 4096
constexpr uint32_t kSmallCell = 256;      // cells up to this size are ranked in place
constexpr uint32_t kLdsSortMax = 16384;   // cells up to this size: LDS bitonic sort
constexpr uint32_t kBigBlock = 1024;      // workgroup of the big-cell kernel
constexpr uint32_t kBigGrid = 128;        // workgroups of the big-cell kernel

// Error flag bits (device word, read by mtlp_error_flags).
constexpr uint32_t kErrScanLookback = 1u;  // a scan tile gave up waiting
constexpr uint32_t kErrHashRange = 2u;     // cell offsets inconsistent with n

// Grid of the spatial hash (Shared.h:40-49).
struct GridDev {
    float ox, oy, oz;     // origin = bounds_min
    float inv;            // 1 / cell_size (FP32)
    uint32_t dx, dy, dz;  // cells per axis
    uint32_t dxy;         // dx * dy
};

// predictPositions inputs (Constraints.metal:54-66).  gdt = gravity * dt,
// rounded once in FP32 exactly as the kernel would.
struct PredictDev {
    float gdtx, gdty, gdtz, dt;
    uint32_t enable;      // 0: hash positions as they are (buildSpatialHash)
};

// Solver constants (writePBFParams, Engine.cpp:551-572; Constraints.metal).
// The density sum runs in FP64 (poly6_d, h2_d, w0_d, inv_rho0_d): rho / rho0
// - 1 cancels to a few per cent near rest, which FP32 resolves only to
// ~1e-5 relative in lambda.
struct SolverDev {
    double poly6_d, h2_d, w0_d, inv_rho0_d;
    // The delta pass runs its per-neighbour terms in FP64 too: with a
    // strong s_corr the terms cancel to a small sum that the 0.3 h cap then
    // rescales, which turns FP32 term rounding into a direction error.
    double h_d, spiky_d, inv_wq_d, scorr_k_d, max_step_d;
    float h, h2, poly6, spiky, inv_rho0, eps;
    float w0;             // poly6(0)
    float inv_wq;         // 1 / poly6((dq h)^2)
    float scorr_k;
    float max_step;       // max_step * h
    uint32_t scorr_n;
    float bminx, bminy, bminz, bmaxx, bmaxy, bmaxz;
};

// finalizeStep constants (Constraints.metal:305-353).
struct FinalDev {
    float bminx, bminy, bminz, bmaxx, bmaxy, bmaxz;
    float inv_dt, damping, weps, floor_f, wall_f, vmax;
};

// Buffers of one spatial-hash build.
struct HashBufs {
    const float4* x;      // positions (original order)
    const float4* v;      // velocities (original order)
    uint32_t* keys;       // cell key per particle
    uint32_t* tickets;    // arrival rank within the cell from the count atomics
    uint32_t* counts;     // C + 1 counters, all zero on entry, zero on exit
    uint32_t* cell_start; // C + 1
    uint32_t* tmp_index;  // unordered cell-grouped indices
    uint32_t* sorted_index;
    float4* sorted;       // gathered (predicted) positions
    uint32_t* big_list;   // cells larger than kSmallCell
    uint32_t* big_count;
    uint64_t* scan_state; // look-back words, one per scan tile
    uint32_t* scan_ticket;
    uint32_t* err;
};

// ---- launchers (return hipGetLastError after the launches) -------------
hipError_t launch_predict_hash_count(const HashBufs& b, uint32_t n, const PredictDev& pd,
                                     const GridDev& g, hipStream_t s);
hipError_t launch_scatter_gather(const HashBufs& b, uint32_t n, const PredictDev& pd,
                                 hipStream_t s);
hipError_t launch_count_neighbors(const float4* sorted, const uint32_t* cell_start,
                                  const uint32_t* sorted_index, uint32_t n, const GridDev& g,
                                  float r2max, uint32_t* nbr, hipStream_t s);

// Exclusive scan of n u32 (out may alias nothing; in is zeroed after
// reading when zero_input).  state/ticket: the look-back words of
// scan_tiles(n) tiles plus one ticket counter; with epoch-tagged words they
// need clearing only once (clear_scan_state) per allocation.
uint32_t scan_tiles(uint32_t n);
hipError_t launch_exclusive_scan(uint32_t* in, uint32_t* out, uint32_t n, bool zero_input,
                                 uint64_t* state, uint32_t* ticket, uint32_t epoch,
                                 uint32_t* err, hipStream_t s);
hipError_t launch_fill_u32(uint32_t* p, uint32_t value, size_t count, hipStream_t s);

hipError_t launch_density_lambda(float4* cur, const uint32_t* cell_start,
                                 const uint32_t* sorted_index, uint32_t n, const GridDev& g,
                                 const SolverDev& sd, float* lambda_orig, uint32_t* nbr_orig,
                                 hipStream_t s);
hipError_t launch_apply_delta(const float4* cur, const uint32_t* cell_start, uint32_t n,
                              const GridDev& g, const SolverDev& sd, float4* next,
                              hipStream_t s);
hipError_t launch_finalize(const float4* final_sorted, const uint32_t* sorted_index, uint32_t n,
                           const FinalDev& fd, float4* x, float4* v, float4* prev,
                           hipStream_t s);

}  // namespace mtlp
