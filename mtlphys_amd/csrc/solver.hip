// solver.hip -- PBF constraint passes for gfx950, in sorted-slot order.
//
// densityLambdaSorted (Constraints.metal:82-131), applyDeltaSorted
// (:140-199) and scatterPositionsToOriginal + finalizeStep (:205-213,
// :305-353) fused into one pass.
//
// Layout: the sorted positions are float4; the density pass stores lambda_i
// in .w of its own slot, so the delta pass reads (x_j, lambda_j) with one
// 16-byte load per candidate.  The 27-cell stencil is walked as 9 rows of
// three x-adjacent cells: their slots are contiguous in cell order, so each
// row is one slot range [cell_start[x0], cell_start[x1 + 1]), visited in the
// reference's z, y, x, slot order.
#include "mtlp_device.h"

namespace mtlp {

namespace {

struct Stencil {
    int x0, x1, y0, y1, z0, z1;
};

__device__ __forceinline__ Stencil stencil_of(float px, float py, float pz, const GridDev& g) {
    const Cell3 c = cell_of(px, py, pz, g);
    return Stencil{max(c.x - 1, 0), min(c.x + 1, (int)g.dx - 1), max(c.y - 1, 0),
                   min(c.y + 1, (int)g.dy - 1), max(c.z - 1, 0), min(c.z + 1, (int)g.dz - 1)};
}

// |r|^2 as ((rx rx + ry ry) + rz rz) without FMA: the neighbour test
// r2 < h^2 then decides exactly as the oracle's FP32 arithmetic does.
__device__ __forceinline__ float dist2(float rx, float ry, float rz) {
#pragma clang fp contract(off)
    return rx * rx + ry * ry + rz * rz;
}

__device__ __forceinline__ float ipow(float x, uint32_t n) {
    float result = 1.0f, b = x;
    while (n) {
        if (n & 1u) result *= b;
        n >>= 1;
        if (n) b *= b;
    }
    return result;
}

template <bool kOrig, bool kCount>
__global__ __launch_bounds__(kBlock) void k_density_lambda(
    float4* __restrict__ cur, const uint32_t* __restrict__ cell_start,
    const uint32_t* __restrict__ sorted_index, uint32_t n, GridDev g, SolverDev sd,
    float* __restrict__ lambda_orig, uint32_t* __restrict__ nbr_orig) {
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= n) return;
    const float4 pi = cur[s];
    const Stencil st = stencil_of(pi.x, pi.y, pi.z, g);
    // sum of (h^2 - r^2)^3 in FP64; r's components are differences of
    // nearby FP32 values (exact by Sterbenz away from the origin), so the
    // density is accurate to FP64 rounding.
    double wsum = 0.0;
    float gx = 0.f, gy = 0.f, gz = 0.f, gsq = 0.f;
    uint32_t cnt = 0;
    for (int z = st.z0; z <= st.z1; ++z)
        for (int y = st.y0; y <= st.y1; ++y) {
            const uint32_t row = (uint32_t)y * g.dx + (uint32_t)z * g.dxy;
            const uint32_t a = cell_start[row + (uint32_t)st.x0];
            const uint32_t b = min(cell_start[row + (uint32_t)st.x1 + 1], n);
            for (uint32_t j = a; j < b; ++j) {
                const float4 pj = cur[j];
                const float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
                const float r2 = dist2(rx, ry, rz);
                if (r2 < sd.h2 && j != s) {
                    ++cnt;
                    const double dxd = rx, dyd = ry, dzd = rz;
                    const double diff = sd.h2_d - (dxd * dxd + dyd * dyd + dzd * dzd);
                    wsum += diff * diff * diff;
                    const float rl = sqrtf(r2);
                    if (rl > 1e-6f && rl < sd.h) {
                        const float dd = sd.h - rl;
                        const float sc = -sd.spiky * dd * dd / rl;
                        const float ax = sc * rx, ay = sc * ry, az = sc * rz;
                        gx += ax;
                        gy += ay;
                        gz += az;
                        gsq += ax * ax + ay * ay + az * az;
                    }
                }
            }
        }
    const double constraint = (sd.w0_d + sd.poly6_d * wsum) * sd.inv_rho0_d - 1.0;
    const float denom = sd.inv_rho0 * sd.inv_rho0 * ((gx * gx + gy * gy + gz * gz) + gsq) + sd.eps;
    const float lambda = (float)(-constraint / (double)denom);
    reinterpret_cast<float*>(cur)[4 * (size_t)s + 3] = lambda;
    if (kOrig || kCount) {
        const uint32_t i = sorted_index[s];
        if (kOrig) lambda_orig[i] = lambda;
        if (kCount) nbr_orig[i] = cnt;
    }
}

__global__ __launch_bounds__(kBlock) void k_apply_delta(
    const float4* __restrict__ cur, const uint32_t* __restrict__ cell_start, uint32_t n, GridDev g,
    SolverDev sd, float4* __restrict__ next) {
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= n) return;
    const float4 pi = cur[s];
    const float lam_i = pi.w;
    const Stencil st = stencil_of(pi.x, pi.y, pi.z, g);
    float dx = 0.f, dy = 0.f, dz = 0.f;
    for (int z = st.z0; z <= st.z1; ++z)
        for (int y = st.y0; y <= st.y1; ++y) {
            const uint32_t row = (uint32_t)y * g.dx + (uint32_t)z * g.dxy;
            const uint32_t a = cell_start[row + (uint32_t)st.x0];
            const uint32_t b = min(cell_start[row + (uint32_t)st.x1 + 1], n);
            for (uint32_t j = a; j < b; ++j) {
                const float4 pj = cur[j];
                const float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
                const float r2 = dist2(rx, ry, rz);
                if (r2 < sd.h2 && j != s) {
                    const float rl = sqrtf(r2);
                    float sc = 0.f;
                    if (rl > 1e-6f && rl < sd.h) {
                        const float dd = sd.h - rl;
                        sc = -sd.spiky * dd * dd / rl;
                    }
                    const float diff = sd.h2 - r2;
                    const float wij = sd.poly6 * diff * diff * diff;
                    const float scorr = -sd.scorr_k * ipow(wij * sd.inv_wq, sd.scorr_n);
                    const float coef = (lam_i + pj.w + scorr) * sc;
                    dx += coef * rx;
                    dy += coef * ry;
                    dz += coef * rz;
                }
            }
        }
    dx *= sd.inv_rho0;
    dy *= sd.inv_rho0;
    dz *= sd.inv_rho0;
    const float dlen = sqrtf(dx * dx + dy * dy + dz * dz);
    if (dlen > sd.max_step) {
        const float f = sd.max_step / dlen;
        dx *= f;
        dy *= f;
        dz *= f;
    }
    next[s] = make_float4(fminf(fmaxf(pi.x + dx, sd.bminx), sd.bmaxx),
                          fminf(fmaxf(pi.y + dy, sd.bminy), sd.bmaxy),
                          fminf(fmaxf(pi.z + dz, sd.bminz), sd.bmaxz), 1.0f);
}

// Scatter back + finalize: slot s writes particle sorted_index[s].  The old
// position (still in x) becomes prev; x and v get the step's result.
__global__ __launch_bounds__(kBlock) void k_finalize(
    const float4* __restrict__ fin, const uint32_t* __restrict__ sorted_index, uint32_t n,
    FinalDev fd, float4* __restrict__ x, float4* __restrict__ v, float4* __restrict__ prev) {
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= n) return;
    const uint32_t i = sorted_index[s];
    const float4 pf = fin[s];
    const float4 old = x[i];
    prev[i] = old;
    const float px = fminf(fmaxf(pf.x, fd.bminx), fd.bmaxx);
    const float py = fminf(fmaxf(pf.y, fd.bminy), fd.bmaxy);
    const float pz = fminf(fmaxf(pf.z, fd.bminz), fd.bmaxz);
    float vx = (px - old.x) * fd.inv_dt * fd.damping;
    float vy = (py - old.y) * fd.inv_dt * fd.damping;
    float vz = (pz - old.z) * fd.inv_dt * fd.damping;
    const bool xmin = px <= fd.bminx + fd.weps, xmax = px >= fd.bmaxx - fd.weps;
    const bool ymin = py <= fd.bminy + fd.weps, ymax = py >= fd.bmaxy - fd.weps;
    const bool zmin = pz <= fd.bminz + fd.weps, zmax = pz >= fd.bmaxz - fd.weps;
    if (xmin) vx = fmaxf(vx, 0.f);
    if (xmax) vx = fminf(vx, 0.f);
    if (ymin) vy = fmaxf(vy, 0.f);
    if (ymax) vy = fminf(vy, 0.f);
    if (zmin) vz = fmaxf(vz, 0.f);
    if (zmax) vz = fminf(vz, 0.f);
    if (ymin) {
        vx *= fd.floor_f;
        vz *= fd.floor_f;
    } else if (xmin || xmax || zmin || zmax) {
        vx *= fd.wall_f;
        vy *= fd.wall_f;
        vz *= fd.wall_f;
    }
    const float vlen = sqrtf(vx * vx + vy * vy + vz * vz);
    if (vlen > fd.vmax) {
        const float f = fd.vmax / vlen;
        vx *= f;
        vy *= f;
        vz *= f;
    }
    x[i] = make_float4(px, py, pz, 1.0f);
    v[i] = make_float4(vx, vy, vz, 0.0f);
}

inline uint32_t blocks_for(uint32_t n) { return (n + kBlock - 1) / kBlock; }

}  // namespace

hipError_t launch_density_lambda(float4* cur, const uint32_t* cell_start,
                                 const uint32_t* sorted_index, uint32_t n, const GridDev& g,
                                 const SolverDev& sd, float* lambda_orig, uint32_t* nbr_orig,
                                 hipStream_t s) {
    if (n == 0) return hipSuccess;
    const dim3 grid(blocks_for(n)), block(kBlock);
    if (lambda_orig && nbr_orig)
        hipLaunchKernelGGL((k_density_lambda<true, true>), grid, block, 0, s, cur, cell_start,
                           sorted_index, n, g, sd, lambda_orig, nbr_orig);
    else if (lambda_orig)
        hipLaunchKernelGGL((k_density_lambda<true, false>), grid, block, 0, s, cur, cell_start,
                           sorted_index, n, g, sd, lambda_orig, nbr_orig);
    else if (nbr_orig)
        hipLaunchKernelGGL((k_density_lambda<false, true>), grid, block, 0, s, cur, cell_start,
                           sorted_index, n, g, sd, lambda_orig, nbr_orig);
    else
        hipLaunchKernelGGL((k_density_lambda<false, false>), grid, block, 0, s, cur, cell_start,
                           sorted_index, n, g, sd, lambda_orig, nbr_orig);
    return hipGetLastError();
}

hipError_t launch_apply_delta(const float4* cur, const uint32_t* cell_start, uint32_t n,
                              const GridDev& g, const SolverDev& sd, float4* next,
                              hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_apply_delta, dim3(blocks_for(n)), dim3(kBlock), 0, s, cur, cell_start, n,
                       g, sd, next);
    return hipGetLastError();
}

hipError_t launch_finalize(const float4* final_sorted, const uint32_t* sorted_index, uint32_t n,
                           const FinalDev& fd, float4* x, float4* v, float4* prev,
                           hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_finalize, dim3(blocks_for(n)), dim3(kBlock), 0, s, final_sorted,
                       sorted_index, n, fd, x, v, prev);
    return hipGetLastError();
}

}  // namespace mtlp
