// mtlp_device.h -- device helpers shared by the gfx950 kernels.
#pragma once

#include "mtlp_kernels.h"

namespace mtlp {

// cellCoordOf (Spatial.metal:35-39): q = (p - origin) * invCell as two
// separately rounded FP32 operations, floor, clamp to [0, dim-1].  q is
// limited to [-1, dim] first so the float->int conversion is defined for
// every input (NaN -> -1 -> cell 0); no result of the final clamp changes.
__device__ __forceinline__ int cell_axis(float p, float o, float inv, uint32_t dim) {
#pragma clang fp contract(off)
    float q = (p - o) * inv;
    q = fminf(fmaxf(q, -1.0f), (float)dim);
    const int c = (int)floorf(q);
    return min(max(c, 0), (int)dim - 1);
}

struct Cell3 {
    int x, y, z;
};

__device__ __forceinline__ Cell3 cell_of(float px, float py, float pz, const GridDev& g) {
    return Cell3{cell_axis(px, g.ox, g.inv, g.dx), cell_axis(py, g.oy, g.inv, g.dy),
                 cell_axis(pz, g.oz, g.inv, g.dz)};
}

// cellHashOf (Spatial.metal:41-43)
__device__ __forceinline__ uint32_t key_of(const Cell3& c, const GridDev& g) {
    return (uint32_t)c.x + (uint32_t)c.y * g.dx + (uint32_t)c.z * g.dxy;
}

// The neighbour test of every stencil walk and of countNeighbors: |r|^2 <
// h^2 (Constraints.metal:112-113, Spatial.metal:235-236: dot(r, r) < h^2),
// evaluated as ((rx rx - h^2) + ry ry) + rz rz, each step one fused
// multiply-add; the result's sign bit is set exactly for a neighbour.  Three
// FMAs instead of three multiplies, two adds and a subtract; the CPU oracle
// evaluates the same chain with fmaf.
__device__ __forceinline__ float nbr_key(float rx, float ry, float rz, float h2) {
    return __builtin_fmaf(rz, rz, __builtin_fmaf(ry, ry, __builtin_fmaf(rx, rx, -h2)));
}

// cell_start[key] and cell_start[key + 1] with one 8-byte load (4-byte
// aligned: one request per lane where two separate loads made two, which
// counts where the keys are in random order).
struct __attribute__((packed, aligned(4))) CellRange {
    uint32_t lo, hi;
};
__device__ __forceinline__ CellRange cell_range(const uint32_t* __restrict__ cell_start, uint32_t key) {
    return *reinterpret_cast<const CellRange*>(cell_start + key);
}

// Scene of particle (or sorted slot) i: the last s with p_off[s] <= i.
// With one scene the result is the constant 0, so scene fields are read
// with uniform (scalar) loads.
__device__ __forceinline__ uint32_t scene_of(uint32_t i, const StepArgs& a) {
    if (a.n_scenes == 1) return 0u;
    uint32_t lo = 0, hi = a.n_scenes;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a.p_off[mid] <= i) lo = mid;
        else hi = mid;
    }
    return lo;
}

// The scene table read through the constant address space: with a
// wave-uniform index these are scalar loads into SGPRs, like kernel
// arguments.  The table is written only by the host between launches.
typedef const __attribute__((address_space(4))) uint32_t* ConstWords;
static_assert(sizeof(SceneDev) % 4 == 0, "scene constants are copied as dwords");

__device__ __forceinline__ SceneDev load_scene_uniform(const SceneDev* table, uint32_t sid) {
    union {
        SceneDev s;
        uint32_t w[sizeof(SceneDev) / 4];
    } u;
    const ConstWords src = (ConstWords)(table + sid);
#pragma unroll
    for (uint32_t k = 0; k < sizeof(SceneDev) / 4; ++k) u.w[k] = src[k];
    return u.s;
}

// Calls body(constants of the scene of element i), the constants always
// coming through scalar loads: each trip takes the scene of the first active
// lane, runs the lanes of that scene, and leaves the others for the next trip
// (one trip for a wave inside one scene; a wave straddling k scenes takes k).
template <class Body>
__device__ __forceinline__ void with_scene(uint32_t i, const StepArgs& a, Body&& body) {
    const uint32_t sid = scene_of(i, a);
    // The trip condition is the wave-uniform mask of lanes still to run, so
    // the body stays inside the loop under a uniform scene index.  (Branching
    // on sid == sid0, or leaving the loop after the body, lets the compiler
    // put the per-lane sid back into the address.)
    uint64_t todo = __ballot(true);
    while (todo) {
        const uint32_t sid0 = __builtin_amdgcn_readlane(sid, (int)__builtin_ctzll(todo));
        const uint64_t same = __ballot(sid == sid0);
        if ((same >> __lane_id()) & 1ull) {
            const SceneDev sc = load_scene_uniform(a.scenes, sid0);
            body(sc);
        }
        todo &= ~same;
    }
}

// predictPositions (Constraints.metal:54-66): v' = v + g dt, x* = x + v' dt,
// every operation rounded separately (no FMA) so the hash of x* is
// reproducible on the CPU.
__device__ __forceinline__ float4 predict_pos(float4 x, float4 v, float gx, float gy, float gz,
                                              float dt) {
#pragma clang fp contract(off)
    const float gdtx = gx * dt, gdty = gy * dt, gdtz = gz * dt;
    const float vx = v.x + gdtx, vy = v.y + gdty, vz = v.z + gdtz;
    const float sx = vx * dt, sy = vy * dt, sz = vz * dt;
    return make_float4(x.x + sx, x.y + sy, x.z + sz, x.w);
}

// Position particle i is binned at: x* on the step path, x for a plain
// spatial-hash build.
__device__ __forceinline__ float4 binned_pos_of(const float4& xi, const float4& vi, uint32_t i, const StepArgs& a) {
    if (!a.predict) return xi;
    float4 out;
    with_scene(i, a, [&](const SceneDev& sc) { out = predict_pos(xi, vi, sc.gx, sc.gy, sc.gz, a.dt); });
    return out;
}
__device__ __forceinline__ float4 binned_pos(const float4* __restrict__ x, const float4* __restrict__ v,
                                             uint32_t i, const StepArgs& a) {
    const float4 xi = x[i];
    return binned_pos_of(xi, a.predict ? v[i] : xi, i, a);
}

// Slot s of a packed (x, y, z) position copy (12 bytes a slot; NULL: none).
// The byte offset is 32-bit from the uniform base (one VGPR for the
// address); copies are allocated for at most kMaxXyzSlots slots.
__device__ __forceinline__ void store_xyz(float* __restrict__ xyz, uint32_t s, const float4& p) {
    if (!xyz) return;
    float* q = reinterpret_cast<float*>(reinterpret_cast<char*>(xyz) + s * 12u);
    q[0] = p.x;
    q[1] = p.y;
    q[2] = p.z;
}

// (x, y, z, 0) of slot s of a packed copy: one 12-byte load (4-byte aligned)
// where the float4 iterate takes 16 bytes.
struct __attribute__((packed, aligned(4))) Words3 {
    float a, b, c;
};
__device__ __forceinline__ float4 load_xyz(const float* __restrict__ xyz, uint32_t s) {
    const Words3 w = *reinterpret_cast<const Words3*>(reinterpret_cast<const char*>(xyz) + s * 12u);
    return make_float4(w.a, w.b, w.c, 0.0f);
}

// Inclusive wave64 prefix sum with DPP: row_shr 1/2/4/8 inside each 16-lane
// row, then row_bcast:15 and row_bcast:31 carry row totals forward.  Lanes
// whose DPP source is invalid or whose row is masked off read `old` = 0.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false);  // row_shr:1
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false);  // row_shr:2
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false);  // row_shr:4
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false);  // row_shr:8
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);  // row_bcast:15
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);  // row_bcast:31
    return v;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_readlane((int)wave_incl_scan(v), 63);
}

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Bitonic sort of the wave's 64 (key, val) pairs by key, then val; lane l
// returns the l-th val.  Every lane of the wave must take part.
__device__ __forceinline__ uint32_t wave_sort_by_key(uint32_t key, uint32_t val) {
    const uint32_t lane = __lane_id();
#pragma unroll
    for (uint32_t k = 2; k <= 64; k <<= 1) {
#pragma unroll
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            const uint32_t ok = (uint32_t)__shfl_xor((int)key, (int)j);
            const uint32_t ov = (uint32_t)__shfl_xor((int)val, (int)j);
            const bool ascending = (lane & k) == 0u, lower = (lane & j) == 0u;
            const bool other_less = ok < key || (ok == key && ov < val);
            // the lower lane of the pair keeps the smaller pair when ascending
            if (lower ? other_less == ascending : other_less != ascending) {
                key = ok;
                val = ov;
            }
        }
    }
    return val;
}

// MI355X dispatches workgroup b to XCD b % 8 (kNumXcd), and each XCD has
// its own L2.  Renumber so that XCD x runs one contiguous run of logical
// blocks: slots that share stencil rows then share an L2.  A bijection on
// [0, nb).
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb) {
    const uint32_t x = b % kNumXcd, i = b / kNumXcd;
    const uint32_t q = nb / kNumXcd, r = nb % kNumXcd;
    return x * q + min(x, r) + i;
}

// First position of XCD x's range under xcd_block, and the XCD whose range
// holds position t.
__device__ __forceinline__ uint32_t xcd_range_start(uint32_t x, uint32_t nb) {
    const uint32_t q = nb / kNumXcd, r = nb % kNumXcd;
    return x * q + min(x, r);
}
__device__ __forceinline__ uint32_t xcd_of_position(uint32_t t, uint32_t nb) {
    const uint32_t q = nb / kNumXcd, r = nb % kNumXcd;
    const uint32_t big = r * (q + 1u);  // the first r ranges hold q + 1 positions
    return t < big ? t / (q + 1u) : r + (t - big) / q;  // q > 0 whenever t >= big
}

}  // namespace mtlp
