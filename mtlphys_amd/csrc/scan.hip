// scan.hip -- single-pass exclusive scan of u32 for gfx950.
//
// Replaces the reference's three-kernel chunked scan (scanChunkExclusive /
// scanChunkSums / addChunkOffsets, Spatial.metal:82-161, sequenced at
// Engine.cpp:625-648 and :843-882), which is capped at 1024^2 elements
// (Engine.hpp:93).  Here one launch scans any length below 2^31:
//
//   * a tile is 512 threads x 8 coalesced uint4 loads (16,384 elements),
//     from 2^22 elements on 512 x 16 (32,768);
//   * inside a thread, 4 consecutive elements are summed serially;
//   * inside a wave64, thread sums are scanned with DPP row shifts and row
//     broadcasts (wave_incl_scan), replacing simd_prefix_exclusive_sum;
//   * the eight wave totals of each 2048-element column go through LDS;
//   * tiles chain by decoupled look-back (Merrill & Garland 2016): tile ids
//     come from an atomic ticket so a tile only ever waits on tiles that
//     already run; wave 0 inspects 64 predecessors per step.
//
// Look-back words are 64 bit: [63:34] epoch, [33:32] status (1 aggregate,
// 2 inclusive prefix), [31:0] value; a word of another epoch reads as "not
// ready".  Every caller clears the words before a launch and passes epoch 1:
// the scan entry points with a fill kernel, the engine's hash build in the
// predict pass before its scan (so a replayed graph needs no extra launch).
#include <mutex>

#include "mtlp_device.h"

namespace mtlp {

namespace {

constexpr uint32_t kStatusAgg = 1u;
constexpr uint32_t kStatusIncl = 2u;
// Upper bound on look-back polls before a tile gives up (flagging
// kErrScanLookback); far above any legitimate wait.
constexpr uint32_t kMaxPolls = 1u << 22;

__device__ __forceinline__ uint64_t pack_state(uint32_t epoch, uint32_t status, uint32_t value) {
    return ((uint64_t)epoch << 34) | ((uint64_t)status << 32) | (uint64_t)value;
}

__device__ __forceinline__ uint4 load4(const uint32_t* in, size_t e, size_t n) {
    if (e + 3 < n) return *reinterpret_cast<const uint4*>(in + e);
    uint4 r = make_uint4(0u, 0u, 0u, 0u);
    if (e < n) r.x = in[e];
    if (e + 1 < n) r.y = in[e + 1];
    if (e + 2 < n) r.z = in[e + 2];
    return r;
}

__device__ __forceinline__ void store4(uint32_t* out, size_t e, size_t n, uint4 v) {
    if (e + 3 < n) {
        *reinterpret_cast<uint4*>(out + e) = v;
        return;
    }
    if (e < n) out[e] = v.x;
    if (e + 1 < n) out[e + 1] = v.y;
    if (e + 2 < n) out[e + 2] = v.z;
}

// One tile of kB x kItems elements per workgroup: thread t holds uint4
// column k (k < kItems / 4) at tile offset 4 (k kB + t).
template <uint32_t kB, uint32_t kItems>
__global__ __launch_bounds__(kB) void k_exclusive_scan(
    uint32_t* in, uint32_t* out, uint32_t n, uint32_t zero_input,
    uint64_t* __restrict__ state, uint32_t* __restrict__ ticket, uint32_t epoch,
    uint32_t num_tiles, uint32_t* __restrict__ err) {
    static_assert(kB % 64 == 0 && kItems % 4 == 0 && kItems / 4 <= kB, "whole waves, uint4 columns, a thread per column");
    constexpr int kVec = (int)(kItems / 4);  // uint4 per thread
    constexpr uint32_t kWaves = kB / 64;
    constexpr size_t kTile = (size_t)kB * kItems;
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_wave[kWaves][kVec];  // wave totals per column, then their exclusive prefix
    __shared__ uint32_t s_col[kVec];           // column totals
    __shared__ uint32_t s_excl;

    const uint32_t t = threadIdx.x;
    const uint32_t lane = t & 63u;
    const uint32_t wave = t >> 6;

    if (t == 0) {
        const uint32_t id =
            __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        // The last ticket is issued after every other one: reset for the
        // next launch on the stream.
        if (id == num_tiles - 1)
            __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_tile = id;
    }
    __syncthreads();
    const uint32_t tile = s_tile;
    const size_t base = (size_t)tile * kTile;

    // The loads go out together before any is used, then the input is
    // zeroed: a zeroing store between them would hold each later load back
    // (the two may alias as far as the compiler knows), one memory latency
    // per load.  A full tile (the test is uniform over the workgroup) takes
    // plain 16-byte accesses, the last tile the bounds-checked ones.
    const bool full = base + kTile <= (size_t)n;
    uint4 v[kVec];
    if (full) {
#pragma unroll
        for (int k = 0; k < kVec; ++k)
            v[k] = *reinterpret_cast<const uint4*>(in + base + (size_t)k * (kB * 4) + 4u * t);
        if (zero_input) {
#pragma unroll
            for (int k = 0; k < kVec; ++k)
                *reinterpret_cast<uint4*>(in + base + (size_t)k * (kB * 4) + 4u * t) =
                    make_uint4(0u, 0u, 0u, 0u);
        }
    } else {
#pragma unroll
        for (int k = 0; k < kVec; ++k) v[k] = load4(in, base + (size_t)k * (kB * 4) + 4u * t, n);
        if (zero_input) {
#pragma unroll
            for (int k = 0; k < kVec; ++k)
                store4(in, base + (size_t)k * (kB * 4) + 4u * t, n, make_uint4(0u, 0u, 0u, 0u));
        }
    }
    // per column: the thread's exclusive prefix inside its wave, and the
    // wave's total into LDS (u32 sums wrap, so any grouping gives the same
    // result)
    uint32_t wex[kVec];
#pragma unroll
    for (int k = 0; k < kVec; ++k) {
        const uint32_t ts = v[k].x + v[k].y + v[k].z + v[k].w;
        const uint32_t wi = wave_incl_scan(ts);
        wex[k] = wi - ts;
        if (lane == 63) s_wave[wave][k] = wi;
    }
    __syncthreads();
    if (t < (uint32_t)kVec) {  // column t: exclusive prefix over the waves, and its total
        uint32_t run = 0;
#pragma unroll
        for (uint32_t w = 0; w < kWaves; ++w) {
            const uint32_t x = s_wave[w][t];
            s_wave[w][t] = run;
            run += x;
        }
        s_col[t] = run;
    }
    __syncthreads();

    if (wave == 0) {
        uint32_t agg = 0;
#pragma unroll
        for (int k = 0; k < kVec; ++k) agg += s_col[k];
        uint32_t excl = 0;
        if (tile == 0) {
            if (lane == 0)
                __hip_atomic_store(&state[0], pack_state(epoch, kStatusIncl, agg),
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            if (lane == 0)
                __hip_atomic_store(&state[tile], pack_state(epoch, kStatusAgg, agg),
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            int64_t j0 = (int64_t)tile - 1;
            uint32_t polls = 0;
            while (true) {
                const int64_t j = j0 - (int64_t)lane;
                const uint64_t st =
                    (j >= 0) ? __hip_atomic_load(&state[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                             : pack_state(epoch, kStatusIncl, 0u);
                const uint32_t status =
                    ((uint32_t)(st >> 34) == epoch) ? (uint32_t)(st >> 32) & 3u : 0u;
                const uint64_t incl = __ballot(status == kStatusIncl);
                const uint64_t invalid = __ballot(status == 0u);
                const uint32_t first = incl ? (uint32_t)__ffsll((unsigned long long)incl) - 1u : 64u;
                const uint64_t window = (first >= 63u) ? ~0ull : ((2ull << first) - 1ull);
                if (invalid & window) {
                    if (++polls > kMaxPolls) {
                        if (lane == 0) atomicOr(err, kErrScanLookback);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                excl += wave_sum(lane <= first ? (uint32_t)st : 0u);
                if (first < 64u) break;
                j0 -= 64;
            }
            if (lane == 0)
                __hip_atomic_store(&state[tile], pack_state(epoch, kStatusIncl, excl + agg),
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (lane == 0) s_excl = excl;
    }
    __syncthreads();

    uint32_t run = s_excl;
#pragma unroll
    for (int k = 0; k < kVec; ++k) {
        const size_t e = base + (size_t)k * (kB * 4) + 4u * t;
        uint4 o;
        o.x = run + s_wave[wave][k] + wex[k];
        o.y = o.x + v[k].x;
        o.z = o.y + v[k].y;
        o.w = o.z + v[k].z;
        if (full) *reinterpret_cast<uint4*>(out + e) = o;
        else store4(out, e, n, o);
        run += s_col[k];
    }
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t value, size_t count) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = value;
}

}  // namespace

uint32_t scan_tiles(uint32_t n) { return (uint32_t)(((uint64_t)n + kScanTile - 1) / kScanTile); }

// Working memory of the allocation-free scan (mtlp_exclusive_scan_noalloc):
// 4 u32 words (tile ticket, error flag, 2 unused), then one look-back word
// per tile of the longest scan it accepts, kNoAllocScanMax elements
// (16,384 tiles: 128 KiB + 16 bytes).  Each GPU the library runs on holds
// its own copy of this array until the process ends, and every call on a
// GPU shares it.
static __device__ __align__(256) uint64_t scratch[2 + kNoAllocScanMax / kScanTile];

hipError_t noalloc_scan_scratch(uint64_t** out) {
    // device address of `scratch` per device, looked up once per device with
    // that device current (the array's name is only a host-side stand-in)
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static uint64_t* addr[kMaxDevices] = {};
    *out = nullptr;
    int dev = 0;
    const hipError_t gd = hipGetDevice(&dev);
    if (gd != hipSuccess) return gd;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        addr[dev] = static_cast<uint64_t*>(p);
    }
    *out = addr[dev];
    return hipSuccess;
}

hipError_t launch_exclusive_scan(uint32_t* in, uint32_t* out, uint32_t n, bool zero_input,
                                 uint64_t* state, uint32_t* ticket, uint32_t epoch,
                                 uint32_t* err, hipStream_t s) {
    if (n == 0) return hipSuccess;
    // Tiles twice as long from kScanLargeFrom on: the inclusive prefixes
    // advance by one look-back window (64 tiles) per poll, so long scans
    // take fewer, longer tiles (2^24 elements: 32 us instead of 38 per
    // call); short ones keep more tiles in flight.  Either count is at most
    // scan_tiles(n), the look-back words allocated.
    if (n >= kScanLargeFrom) {
        constexpr uint32_t kTile = kScanBlock * kScanItemsLarge;
        const uint32_t tiles = (uint32_t)(((uint64_t)n + kTile - 1) / kTile);
        hipLaunchKernelGGL((k_exclusive_scan<kScanBlock, kScanItemsLarge>), dim3(tiles), dim3(kScanBlock), 0, s, in,
                           out, n, zero_input ? 1u : 0u, state, ticket, epoch, tiles, err);
    } else {
        const uint32_t tiles = scan_tiles(n);
        hipLaunchKernelGGL((k_exclusive_scan<kScanBlock, kScanItems>), dim3(tiles), dim3(kScanBlock), 0, s, in, out,
                           n, zero_input ? 1u : 0u, state, ticket, epoch, tiles, err);
    }
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t value, size_t count, hipStream_t s) {
    if (count == 0) return hipSuccess;
    const size_t blocks = (count + 255) / 256;
    const uint32_t grid = (uint32_t)(blocks < 4096 ? blocks : 4096);
    hipLaunchKernelGGL(k_fill_u32, dim3(grid), dim3(256), 0, s, p, value, count);
    return hipGetLastError();
}

}  // namespace mtlp
