// solver.hip -- PBF constraint passes for gfx950, in sorted-slot order.
//
// densityLambdaSorted (Constraints.metal:82-131), applyDeltaSorted
// (:140-199) and scatterPositionsToOriginal + finalizeStep (:205-213,
// :305-353) fused into one pass.  The per-slot code is in mtlp_stencil.h.
//
// Every kernel covers all scenes of a batch; a slot's scene constants come
// from the scene table.  The stencil kernels are built for 64-, 128-, 256-
// and 512-thread workgroups (the autotuner picks per batch).
#include "mtlp_stencil.h"

namespace mtlp {

namespace {

template <uint32_t kB, bool kOrig, bool kCount, bool kList, bool kPersist, bool kXyz>
__global__ __launch_bounds__(kB) void k_density_lambda(
    float4* __restrict__ cur, const uint32_t* __restrict__ cell_start,
    const uint32_t* __restrict__ sorted_index, StepArgs a, float* __restrict__ lambda_orig,
    uint32_t* __restrict__ nbr_orig, NbrList nl) {
    __shared__ uint2 ent[kList ? kLdsChunks * kB : 1];  // each thread's recorded chunks
    const uint32_t n = a.n;
    const uint32_t nb = (n + kB - 1u) / kB;  // tiles
    if (kPersist && blockIdx.x == 0)  // the delta pass's counters
        for (uint32_t q = threadIdx.x; q < kNumXcd * kTileQueues; q += kB) nl.tile_ctr[kTileCtrs + q * kTileCtrStride] = 0u;
    auto tile_body = [&](uint32_t pos) {
        // the tile (kB slots) at this position: in the order the last delta
        // pass left (heaviest first within each XCD's range), else in place
        const uint32_t tile =
            kList && nl.tile_order ? min((uint32_t)__builtin_amdgcn_readfirstlane((int)nl.tile_order[pos]), nb - 1u)
                                   : pos;
        const TileTimer timer(kList ? nl : NbrList{}, tile);
        const uint32_t s0 = tile * kB + threadIdx.x;
        // The wave's slots are handed to its lanes in order of their current
        // cell, so lanes that walk the same stencil rows sit together: their
        // candidate loads then touch fewer cache lines per instruction.  Only
        // the lane assignment changes.  (Full waves inside one scene.)
        uint32_t s = s0;
        const uint32_t wbase = (uint32_t)__builtin_amdgcn_readfirstlane((int)(s0 - __lane_id()));
        if (wbase + 64u <= n) {
            const uint32_t sid = scene_of(wbase, a);
            if (sid == scene_of(wbase + 63u, a)) {
                const SceneDev sc = load_scene_uniform(a.scenes, sid);
                const float4 p = cur[s0];
                s = wave_sort_by_key(key_of(cell_of(p.x, p.y, p.z, sc.g), sc.g), s0);
            }
        }
        if (s >= n) return;
        with_scene(s, a, [&](const SceneDev& sc) {
            density_lambda_slot<kOrig, kCount, false, kList, kXyz>(s, sc, cur, cell_start, sorted_index, n, lambda_orig,
                                                             nbr_orig, nullptr, nl, ent + threadIdx.x, kB);
        });
    };
    for_each_tile_position<kPersist>(nl, nl.tile_ctr, blockIdx.x, gridDim.x, nb, tile_body);
}

// Five waves per SIMD: the list walk is latency-bound, and the packed-copy
// store of finish_delta would otherwise take the register count one past the
// 96 that allow five.
template <uint32_t kB, bool kList, bool kPersist>
__global__ __launch_bounds__(kB) __attribute__((amdgpu_waves_per_eu(5))) void k_apply_delta(const float4* __restrict__ cur,
                                                   const uint32_t* __restrict__ cell_start,
                                                   StepArgs a, float4* __restrict__ next, NbrList nl) {
    if (kList) {
        delta_launch_body<kB, false, kPersist>(cur, cell_start, a, next, nullptr, nl);
        return;
    }
    const uint32_t s = xcd_block(blockIdx.x, gridDim.x) * kB + threadIdx.x;
    if (s >= a.n) return;
    with_scene(s, a, [&](const SceneDev& sc) {
        apply_delta_slot<false, false>(s, sc, cur, cell_start, a.n, next, nullptr, nl, a.sphere);
    });
}

// Scatter back + finalize (+ computeFoam), in original order: particle i
// reads its final iterate at slot slot_of[i]; its old position (still in x)
// becomes prev; x and v get the step's result.  With lam_in_w, lambda of
// the last density pass (which the delta pass left in .w of the final
// iterate) goes to lambda_orig[i].
__global__ __launch_bounds__(kBlock) void k_finalize(
    const float4* __restrict__ fin, const uint32_t* __restrict__ slot_of, StepArgs a,
    float4* __restrict__ x, float4* __restrict__ v, float4* __restrict__ prev, const uint32_t* __restrict__ nbr,
    float* __restrict__ foam, bool lam_in_w, float* __restrict__ lambda_orig) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= a.n) return;
    with_scene(i, a, [&](const SceneDev& sc) {
        const uint32_t s = slot_of[i];
        const float4 old = x[i];
        prev[i] = old;
        const float4 pf = fin[s];
        const Finalized r = finalize_particle(sc.fd, a.dt, pf, old, foam != nullptr, foam ? nbr[i] : 0u);
        x[i] = r.x;
        v[i] = r.v;
        if (foam) foam[i] = r.foam;
        if (lam_in_w) lambda_orig[i] = pf.w;
    });
}

// Legacy semi-implicit Euler (Integrate.metal:11-42): v += g dt, v *= damping,
// x += v dt, then per axis clamp to the box and reflect with restitution 0.4.
__global__ __launch_bounds__(kBlock) void k_integrate(float4* __restrict__ x, float4* __restrict__ v, uint32_t n,
                                                     float dt, float gy, float damping, FinalDev b) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const float4 p0 = x[i], v0 = v[i];
    float p[3] = {p0.x, p0.y, p0.z}, w[3] = {v0.x, v0.y + gy * dt, v0.z};
    const float lo[3] = {b.bminx, b.bminy, b.bminz}, hi[3] = {b.bmaxx, b.bmaxy, b.bmaxz};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        w[k] *= damping;
        p[k] += w[k] * dt;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k)
        if (p[k] < lo[k]) {
            p[k] = lo[k];
            w[k] = -w[k] * 0.4f;
        }
#pragma unroll
    for (int k = 0; k < 3; ++k)
        if (p[k] > hi[k]) {
            p[k] = hi[k];
            w[k] = -w[k] * 0.4f;
        }
    x[i] = make_float4(p[0], p[1], p[2], 1.0f);
    v[i] = make_float4(w[0], w[1], w[2], 0.0f);
}

// applyMousePulse (Constraints.metal:222-253): velocity impulse with a
// Gaussian falloff around the cursor ray, for particles in front of its
// origin and within `radius` of it.
__global__ __launch_bounds__(kBlock) void k_mouse_pulse(const float4* __restrict__ x, float4* __restrict__ v,
                                                       uint32_t n, float ox, float oy, float oz, float dx,
                                                       float dy, float dz, float fx, float fy, float fz,
                                                       float radius) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const float4 p = x[i];
    const float tx = p.x - ox, ty = p.y - oy, tz = p.z - oz;
    const float t = tx * dx + ty * dy + tz * dz;
    if (t <= 0.0f) return;
    const float qx = p.x - (ox + dx * t), qy = p.y - (oy + dy * t), qz = p.z - (oz + dz * t);
    const float r2 = qx * qx + qy * qy + qz * qz;
    const float R2 = radius * radius;
    if (r2 < R2) {
        const float w = expf(-r2 / (R2 * 0.4f));
        const float4 vi = v[i];
        v[i] = make_float4(vi.x + fx * w, vi.y + fy * w, vi.z + fz * w, 0.0f);
    }
}

template <uint32_t kB, bool kList, bool kPersist, bool kXyz>
void density_lambda_b(float4* cur, const uint32_t* cs, const uint32_t* sidx, const StepArgs& a,
                      float* lo, uint32_t* nb, const NbrList& nl, hipStream_t s) {
    const dim3 grid(stencil_grid(blocks_for(a.n, kB), kB, nl)), block(kB);
    if (lo && nb)
        hipLaunchKernelGGL((k_density_lambda<kB, true, true, kList, kPersist, kXyz>), grid, block, 0, s, cur, cs,
                           sidx, a, lo, nb, nl);
    else if (lo)
        hipLaunchKernelGGL((k_density_lambda<kB, true, false, kList, kPersist, kXyz>), grid, block, 0, s, cur, cs,
                           sidx, a, lo, nb, nl);
    else if (nb)
        hipLaunchKernelGGL((k_density_lambda<kB, false, true, kList, kPersist, kXyz>), grid, block, 0, s, cur, cs,
                           sidx, a, lo, nb, nl);
    else
        hipLaunchKernelGGL((k_density_lambda<kB, false, false, kList, kPersist, kXyz>), grid, block, 0, s, cur, cs,
                           sidx, a, lo, nb, nl);
}

template <uint32_t kB>
void density_lambda_bl(float4* cur, const uint32_t* cs, const uint32_t* sidx, const StepArgs& a, float* lo,
                       uint32_t* nb, const NbrList& nl, hipStream_t s) {
    if (!nl.blk) density_lambda_b<kB, false, false, false>(cur, cs, sidx, a, lo, nb, nl, s);
    else if (nl.pull && nl.xyz) density_lambda_b<kB, true, true, true>(cur, cs, sidx, a, lo, nb, nl, s);
    else if (nl.pull) density_lambda_b<kB, true, true, false>(cur, cs, sidx, a, lo, nb, nl, s);
    else if (nl.xyz) density_lambda_b<kB, true, false, true>(cur, cs, sidx, a, lo, nb, nl, s);
    else density_lambda_b<kB, true, false, false>(cur, cs, sidx, a, lo, nb, nl, s);
}

template <uint32_t kB>
void apply_delta_b(const float4* cur, const uint32_t* cs, const StepArgs& a, float4* next, const NbrList& nl,
                   hipStream_t s) {
    // with lists, the heavy queue's workgroups in front
    const dim3 grid(stencil_grid(blocks_for(a.n, kB), kB, nl) +
                    (nl.blk ? heavy_blocks<kB>() + split_blocks<kB>() : 0u) + (nl.tile_order_next ? kNumXcd : 0u)),
        block(kB);
    if (nl.blk && nl.pull) hipLaunchKernelGGL((k_apply_delta<kB, true, true>), grid, block, 0, s, cur, cs, a, next, nl);
    else if (nl.blk) hipLaunchKernelGGL((k_apply_delta<kB, true, false>), grid, block, 0, s, cur, cs, a, next, nl);
    else hipLaunchKernelGGL((k_apply_delta<kB, false, false>), grid, block, 0, s, cur, cs, a, next, nl);
}

}  // namespace

hipError_t launch_density_lambda(float4* cur, const uint32_t* cell_start,
                                 const uint32_t* sorted_index, const StepArgs& a,
                                 float* lambda_orig, uint32_t* nbr_orig, const NbrList& nl, uint32_t block,
                                 hipStream_t s) {
    if (a.n == 0) return hipSuccess;
    switch (block) {
        case 64: density_lambda_bl<64>(cur, cell_start, sorted_index, a, lambda_orig, nbr_orig, nl, s); break;
        case 128: density_lambda_bl<128>(cur, cell_start, sorted_index, a, lambda_orig, nbr_orig, nl, s); break;
        case 512: density_lambda_bl<512>(cur, cell_start, sorted_index, a, lambda_orig, nbr_orig, nl, s); break;
        default: density_lambda_bl<256>(cur, cell_start, sorted_index, a, lambda_orig, nbr_orig, nl, s); break;
    }
    return hipGetLastError();
}

hipError_t launch_apply_delta(const float4* cur, const uint32_t* cell_start, const StepArgs& a,
                              float4* next, const NbrList& nl, uint32_t block, hipStream_t s) {
    if (a.n == 0) return hipSuccess;
    switch (block) {
        case 64: apply_delta_b<64>(cur, cell_start, a, next, nl, s); break;
        case 128: apply_delta_b<128>(cur, cell_start, a, next, nl, s); break;
        case 512: apply_delta_b<512>(cur, cell_start, a, next, nl, s); break;
        default: apply_delta_b<256>(cur, cell_start, a, next, nl, s); break;
    }
    return hipGetLastError();
}

hipError_t launch_finalize(const float4* final_sorted, const uint32_t* slot_of,
                           const StepArgs& a, float4* x, float4* v, float4* prev,
                           const uint32_t* nbr, float* foam, bool lam_in_w, float* lambda_orig,
                           hipStream_t s) {
    if (a.n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_finalize, dim3(blocks_for(a.n, kBlock)), dim3(kBlock), 0, s, final_sorted,
                       slot_of, a, x, v, prev, nbr, foam, lam_in_w, lambda_orig);
    return hipGetLastError();
}

hipError_t launch_integrate(float4* x, float4* v, uint32_t n, float dt, float gy, float damping,
                            const FinalDev& bounds, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_integrate, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, s, x, v, n, dt, gy, damping,
                       bounds);
    return hipGetLastError();
}

hipError_t launch_mouse_pulse(const float4* x, float4* v, uint32_t n, const float origin[3],
                              const float dir[3], const float force[3], float radius, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_mouse_pulse, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, s, x, v, n, origin[0],
                       origin[1], origin[2], dir[0], dir[1], dir[2], force[0], force[1], force[2], radius);
    return hipGetLastError();
}

}  // namespace mtlp
