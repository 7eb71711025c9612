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

// predictPositions (Constraints.metal:54-66): v' = v + g dt, x* = x + v' dt,
// every operation rounded separately (no FMA) so the hash of x* is
// reproducible on the CPU.
__device__ __forceinline__ float4 predict_pos(float4 x, float4 v, const PredictDev& pd) {
#pragma clang fp contract(off)
    if (!pd.enable) return x;
    const float vx = v.x + pd.gdtx, vy = v.y + pd.gdty, vz = v.z + pd.gdtz;
    const float sx = vx * pd.dt, sy = vy * pd.dt, sz = vz * pd.dt;
    return make_float4(x.x + sx, x.y + sy, x.z + sz, x.w);
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

// MI355X dispatches workgroup b to XCD b % 8, and each XCD has its own L2.
// Renumber so that XCD x runs one contiguous run of logical blocks: slots
// that share stencil rows then share an L2.  A bijection on [0, nb).
constexpr uint32_t kNumXcd = 8;
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb) {
    const uint32_t x = b % kNumXcd, i = b / kNumXcd;
    const uint32_t q = nb / kNumXcd, r = nb % kNumXcd;
    return x * q + min(x, r) + i;
}

}  // namespace mtlp
