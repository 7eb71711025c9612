// adjoint.hip -- taped forward passes and the adjoint (vector-Jacobian
// product) of the PBF step for gfx950.  The reference has no adjoint; this
// differentiates the step that solver.hip computes, with the neighbour
// structure held fixed (binning and candidate sets are piecewise constant),
// the box clamps as pass-through inside the box and 0 outside, the 0.3 h
// step cap and the h/dt speed cap through their normalisation Jacobians, and
// the wall responses as piecewise-constant multipliers.
//
// Pair terms, scatter form.  With the stale binning the neighbour relation
// is not symmetric (j in N_i does not imply i in N_j after iteration 0), so
// every pair term (i, j), j in N_i, adds to i and subtracts from j.  The
// kernel for slot s walks its own candidates c and handles both roles:
//   A (i = s, j = c): s keeps its part; c's part is delivered by c itself
//     when s is one of c's candidates (bin(s) in the stencil of c's cell),
//     otherwise as a remainder term into c's inbox (rare);
//   B (i = c, j = s): s takes c's term on s when s is in N_c, i.e. when
//     bin(s) lies in the stencil of c's current cell.
// The two roles share the pair's geometry (H is even in r, G odd), so each
// kernel evaluates them as one expression.  Terms and sums are FP32 (the
// terms' own rounding dominates; FP64 sums measured no closer to the FP64
// adjoint); a slot's own result is a plain store and its inbox is added up
// independently of arrival order, so the output is deterministic (unless a
// slot receives more than kInbox remainder terms in one pass).
//
// Math (r = p_i - p_j, f(r) = -k_s (h - r)^2 the spiky gradient magnitude,
// G = f r^ (Constraints.metal:35-39), its Hessian H v = (f' - f/r)(r^.v) r^
// + (f/r) v with f' = 2 k_s (h - r); W = poly6, grad W = gp r with
// gp = -6 k_p (h^2 - r^2)^2; s_corr = -k (W/Wq)^n, grad s = ds r):
//   delta:   R_i = (1/rho0) sum_j (l_i + l_j + s_ij) G_ij
//            Rbar -> lbar_i += G.Rbar_i/rho0, lbar_j += G.Rbar_i/rho0,
//                    pbar_i += (c/rho0) H Rbar_i + (G.Rbar_i/rho0) grad s
//   lambda:  l = -C/(D + eps), C = rho/rho0 - 1, D = (|g|^2 + q)/rho0^2,
//            g = sum_j G_ij, q = sum_j |G_ij|^2
//            rhobar = -lbar/((D+eps) rho0), qbar = -lbar l/((D+eps) rho0^2),
//            gbar = 2 qbar g; pbar_i += rhobar gp r + H (gbar + 2 qbar G)
#include "mtlp_stencil.h"

namespace mtlp {

namespace {

// ---- taped forward ---------------------------------------------------------

template <uint32_t kB, bool kOrig, bool kCount>
__global__ __launch_bounds__(kB) void k_density_lambda_tape(float4* __restrict__ cur,
                                                           const uint32_t* __restrict__ cell_start,
                                                           const uint32_t* __restrict__ sorted_index, StepArgs a,
                                                           float4* __restrict__ aux, float* __restrict__ lambda_orig,
                                                           uint32_t* __restrict__ nbr_orig, NbrList nl) {
    __shared__ uint2 ent[kLdsChunks * kB];  // each thread's recorded chunks
    const uint32_t s = xcd_block(blockIdx.x, gridDim.x) * kB + threadIdx.x;
    if (s >= a.n) return;
    with_scene(s, a, [&](const SceneDev& sc) {
        density_lambda_slot<kOrig, kCount, true, true>(s, sc, cur, cell_start, sorted_index, a.n, lambda_orig,
                                                       nbr_orig, aux, nl, ent + threadIdx.x, kB);
    });
}

template <uint32_t kB>
void density_lambda_tape_b(float4* cur, const uint32_t* cs, const uint32_t* sidx, const StepArgs& a, float4* aux,
                           float* lo, uint32_t* nb, const NbrList& nl, hipStream_t s) {
    const dim3 grid(blocks_for(a.n, kB)), block(kB);
    if (lo && nb)
        hipLaunchKernelGGL((k_density_lambda_tape<kB, true, true>), grid, block, 0, s, cur, cs, sidx, a, aux, lo, nb,
                           nl);
    else if (lo)
        hipLaunchKernelGGL((k_density_lambda_tape<kB, true, false>), grid, block, 0, s, cur, cs, sidx, a, aux, lo, nb,
                           nl);
    else if (nb)
        hipLaunchKernelGGL((k_density_lambda_tape<kB, false, true>), grid, block, 0, s, cur, cs, sidx, a, aux, lo, nb,
                           nl);
    else
        hipLaunchKernelGGL((k_density_lambda_tape<kB, false, false>), grid, block, 0, s, cur, cs, sidx, a, aux, lo,
                           nb, nl);
}

template <uint32_t kB>
__global__ __launch_bounds__(kB) void k_apply_delta_tape(const float4* __restrict__ cur,
                                                        const uint32_t* __restrict__ cell_start, StepArgs a,
                                                        float4* __restrict__ next, float4* __restrict__ rr,
                                                        NbrList nl) {
    delta_launch_body<kB, true>(cur, cell_start, a, next, rr, nl);
}

__global__ __launch_bounds__(kBlock) void k_finalize_tape(const float4* __restrict__ fin,
                                                         const uint32_t* __restrict__ sorted_index, StepArgs a,
                                                         const float4* __restrict__ x_in,
                                                         float4* __restrict__ x_out,
                                                         float4* __restrict__ v_out) {
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= a.n) return;
    with_scene(s, a, [&](const SceneDev& sc) {
        finalize_slot(s, sc.fd, a.dt, fin, sorted_index, x_in, x_out, v_out, nullptr);
    });
}

// ---- adjoint: pointwise stages ------------------------------------------

__device__ __forceinline__ float4 zero4() { return make_float4(0.f, 0.f, 0.f, 0.f); }

// out[0..5] = sums of g[0..5] over the kBlock threads of this block.
__device__ __forceinline__ void block_sum6(const double g[6], double* __restrict__ out) {
    __shared__ double red[6][kBlock / 64];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        double v = g[k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if (lane_id() == 0) red[k][threadIdx.x / 64] = v;
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        double t = 0.0;
#pragma unroll
        for (uint32_t w = 0; w < kBlock / 64; ++w) t += red[threadIdx.x][w];
        out[threadIdx.x] = t;
    }
}

// finalizeStep (Constraints.metal:305-353) backwards for slot s, particle i:
// (xbar_out, vbar_out)[i] -> pbar of the last iterate (sorted) and the
// prev-position part of xbar_in[i].
//
// Gravity gradient.  dL/dg = dt^2 sum_s pbar^0_s, but after a few chaotic
// steps |pbar| is ~1e9 times the sum, which FP32 values cannot resolve.  The
// sums follow a recursion instead: every pair term adds to one slot what it
// subtracts from another, and the per-particle maps are identities except
// at walls, caps and clamps, so with Sx = sum xbar_out, Sv = sum vbar_out,
// c = damping / dt:
//   P0 = sum pbar^0 = Sx + c Sv + sum_s (du_s + dp_s)
//   sum xbar_in = P0 - c Sv - sum_s du_s,  sum vbar_in = dt P0,
// where du_s = ubar_s - c vbar_s (nonzero only where friction, a wall gate
// or the speed cap acts) and dp_s collects -pbar of every clamped axis (at
// finalize and at each iteration).  gdev[6 s ..] = (du + dp, du) in FP64.
__global__ __launch_bounds__(kBlock) void k_finalize_vjp(
    const float4* __restrict__ fin, const uint32_t* __restrict__ sorted_index, StepArgs a,
    const float4* __restrict__ x_in, const float4* __restrict__ xbar_out,
    const float4* __restrict__ vbar_out, float4* __restrict__ pacc, float4* __restrict__ xbar_in,
    double* __restrict__ gdev) {
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= a.n) return;
    with_scene(s, a, [&](const SceneDev& sc) {
        const FinalDev& fd = sc.fd;
        const float inv_dt = 1.0f / a.dt;
        const float vmax = fd.vmax_h * inv_dt;
        const uint32_t i = sorted_index[s];
        const float4 pf = fin[s], old = x_in[i];
        const float p[3] = {fminf(fmaxf(pf.x, fd.bminx), fd.bmaxx), fminf(fmaxf(pf.y, fd.bminy), fd.bmaxy),
                            fminf(fmaxf(pf.z, fd.bminz), fd.bmaxz)};
        const float o[3] = {old.x, old.y, old.z};
        const float lo[3] = {fd.bminx, fd.bminy, fd.bminz}, hi[3] = {fd.bmaxx, fd.bmaxy, fd.bmaxz};
        const float raw[3] = {pf.x, pf.y, pf.z};
        bool mn[3], mx[3];
        float u[3], gate[3], w[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            u[k] = (p[k] - o[k]) * inv_dt * fd.damping;
            mn[k] = p[k] <= lo[k] + fd.weps;
            mx[k] = p[k] >= hi[k] - fd.weps;
            gate[k] = ((mn[k] && u[k] < 0.f) || (mx[k] && u[k] > 0.f)) ? 0.f : 1.f;
        }
        float fr[3] = {1.f, 1.f, 1.f};
        if (mn[1]) {
            fr[0] = fd.floor_f;
            fr[2] = fd.floor_f;
        } else if (mn[0] || mx[0] || mn[2] || mx[2]) {
            fr[0] = fr[1] = fr[2] = fd.wall_f;
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) w[k] = fr[k] * gate[k] * u[k];
        const float4 vb4 = vbar_out[i], xb4 = xbar_out[i];
        const float vb[3] = {vb4.x, vb4.y, vb4.z};
        float wb[3] = {vb4.x, vb4.y, vb4.z};
        const float xb[3] = {xb4.x, xb4.y, xb4.z};
        const float wl2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
        const float wl = sqrtf(wl2);
        if (wl > vmax) {
            const float dot = (w[0] * wb[0] + w[1] * wb[1] + w[2] * wb[2]) / wl2;
            const float f = vmax / wl;
#pragma unroll
            for (int k = 0; k < 3; ++k) wb[k] = f * (wb[k] - w[k] * dot);
        }
        float pb[3], ub[3];
        const double c = (1.0 / (double)a.dt) * (double)fd.damping;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            ub[k] = fr[k] * gate[k] * wb[k] * inv_dt * fd.damping;
            const bool inside = raw[k] >= lo[k] && raw[k] <= hi[k];
            pb[k] = inside ? xb[k] + ub[k] : 0.f;
            // deviations from the identity maps, for the gravity recursion
            const bool plain = fr[k] == 1.f && gate[k] == 1.f && wl <= vmax;
            const double du = plain ? 0.0 : ((double)(fr[k] * gate[k] * wb[k]) - (double)vb[k]) * c;
            const double dp = inside ? 0.0 : -((double)xb[k] + (double)vb[k] * c + du);
            gdev[6 * (size_t)s + k] = du + dp;
            gdev[6 * (size_t)s + 3 + k] = du;
        }
        pacc[s] = make_float4(pb[0], pb[1], pb[2], 0.f);
        xbar_in[i] = make_float4(-ub[0], -ub[1], -ub[2], 0.f);
    });
}

// A remainder term v = (lambda-bar, pbar.xyz) for slot c: into c's inbox,
// or past its kInbox entries by float atomics (the lambda part only when
// with_l).
__device__ __forceinline__ void remainder_put(uint32_t* __restrict__ rcnt, float4* __restrict__ rval,
                                              float* __restrict__ lrem, float4* __restrict__ prem, uint32_t c,
                                              const float4& v, bool with_l) {
    const uint32_t q = atomicAdd(rcnt + c, 1u);
    if (q < kInbox) {
        rval[(size_t)c * kInbox + q] = v;
        return;
    }
    if (with_l) atomicAdd(lrem + c, v.x);
    float* p = reinterpret_cast<float*>(prem + c);
    atomicAdd(p + 0, v.y);
    atomicAdd(p + 1, v.z);
    atomicAdd(p + 2, v.w);
}

// Order-independent sums of components kFirst .. kFirst + kC - 1 of the
// m <= kInbox entries e[0..m): every FP32 value of a component is scaled by
// 2^S to a 64-bit integer, S set by the largest exponent in the component
// (each term below 2^57 in magnitude, so |total| < 2^62 for 32 entries), the
// integers are added and the total is rounded back to FP32 once.  Any arrival order gives the same bits;
// parts below 2^-57 of the largest value are truncated toward zero.  A
// component holding an infinity or NaN is summed as floats (the result is
// not finite either way).
__device__ __forceinline__ uint32_t exp_bits(float v) { return (__float_as_uint(v) >> 23) & 0xffu; }

__device__ __forceinline__ double pow2_d(int e) {  // 2^e, -1022 <= e <= 1023
    return __longlong_as_double((long long)(e + 1023) << 52);
}

template <int kC, int kFirst>
__device__ __forceinline__ void inbox_sum(const float4* __restrict__ e, uint32_t m, float* out) {
    uint32_t emax[kC];
#pragma unroll
    for (int k = 0; k < kC; ++k) emax[k] = 1u;
#pragma unroll 1
    for (uint32_t q = 0; q < m; ++q) {
        const float4 v = e[q];
        const float c[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int k = 0; k < kC; ++k) emax[k] = max(emax[k], exp_bits(c[kFirst + k]));
    }
    long long acc[kC];
    float fsum[kC];
    double scale[kC];
#pragma unroll
    for (int k = 0; k < kC; ++k) {
        acc[k] = 0;
        fsum[k] = 0.f;
        scale[k] = pow2_d(183 - (int)min(emax[k], 254u));
    }
#pragma unroll 1
    for (uint32_t q = 0; q < m; ++q) {
        const float4 v = e[q];
        const float c[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int k = 0; k < kC; ++k) {
            const float x = c[kFirst + k];
            if (emax[k] < 255u) acc[k] += (long long)((double)x * scale[k]);
            fsum[k] += x;
        }
    }
#pragma unroll
    for (int k = 0; k < kC; ++k) out[k] = emax[k] < 255u ? (float)((double)acc[k] / scale[k]) : fsum[k];
}

// Box clamp and 0.3 h cap of applyDeltaSorted (Constraints.metal:185-197)
// backwards: pbar of iterate k+1 (own part in pacc plus the remainder) ->
// Rbar_s, and the pass-through part of pbar of iterate k (into pacc).  A
// clamped axis takes its pbar out of the gravity sum (gdev).
__global__ __launch_bounds__(kBlock) void k_delta_vjp_pre(const float4* __restrict__ cur,
                                                         const float4* __restrict__ rr, StepArgs a,
                                                         float4* __restrict__ pacc, float4* __restrict__ prem,
                                                         uint32_t* __restrict__ rcnt, const float4* __restrict__ rval,
                                                         float4* __restrict__ rbar, double* __restrict__ gdev) {
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= a.n) return;
    with_scene(s, a, [&](const SceneDev& sc) {
        const SolverDev& sd = sc.sd;
        const float4 pb4 = pacc[s], ro = prem[s], p4 = cur[s], r4 = rr[s];
        float in[3];
        inbox_sum<3, 1>(rval + (size_t)s * kInbox, min(rcnt[s], kInbox), in);
        const float4 rm = make_float4(in[0] + ro.x, in[1] + ro.y, in[2] + ro.z, 0.f);
        prem[s] = zero4();
        rcnt[s] = 0u;
        const double R[3] = {r4.x, r4.y, r4.z};
        const double rl2 = R[0] * R[0] + R[1] * R[1] + R[2] * R[2];
        const double rl = sqrt(rl2);
        const bool capped = rl > sd.max_step_d;
        const double f = capped ? sd.max_step_d / rl : 1.0;
        const double p[3] = {p4.x, p4.y, p4.z};
        const double lo[3] = {sd.bminx, sd.bminy, sd.bminz}, hi[3] = {sd.bmaxx, sd.bmaxy, sd.bmaxz};
        const float pb[3] = {pb4.x + rm.x, pb4.y + rm.y, pb4.z + rm.z};
        double tb[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double t = p[k] + R[k] * f;
            const bool inside = t >= lo[k] && t <= hi[k];
            tb[k] = inside ? (double)pb[k] : 0.0;
            if (!inside) gdev[6 * (size_t)s + k] -= (double)pb[k];
        }
        double rb[3] = {tb[0], tb[1], tb[2]};
        if (capped) {
            const double dot = (R[0] * tb[0] + R[1] * tb[1] + R[2] * tb[2]) / rl2;
#pragma unroll
            for (int k = 0; k < 3; ++k) rb[k] = f * (tb[k] - R[k] * dot);
        }
        rbar[s] = make_float4((float)rb[0], (float)rb[1], (float)rb[2], 0.f);
        pacc[s] = make_float4((float)tb[0], (float)tb[1], (float)tb[2], 0.f);
    });
}

// lambda = -C / (D + eps) backwards: lbar (own + remainder) -> rhobar (into
// cur[s].w, replacing lambda, which no later stage reads) and
// (gbar, qbar) into db[s].
__global__ __launch_bounds__(kBlock) void k_density_vjp_pre(float4* __restrict__ cur,
                                                           const float4* __restrict__ aux, StepArgs a,
                                                           const float* __restrict__ lbar,
                                                           float* __restrict__ lrem, const uint32_t* __restrict__ rcnt,
                                                           const float4* __restrict__ rval, float4* __restrict__ db) {
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= a.n) return;
    with_scene(s, a, [&](const SceneDev& sc) {
        const SolverDev& sd = sc.sd;
        float in;
        inbox_sum<1, 0>(rval + (size_t)s * kInbox, min(rcnt[s], kInbox), &in);
        const float lb = lbar[s] + (in + lrem[s]);
        lrem[s] = 0.f;
        const float4 ax = aux[s];
        const float lam = cur[s].w;
        const double dn = (double)(ax.w + sd.eps);  // the forward's FP32 denominator
        const double rho_bar = -(double)lb / dn * sd.inv_rho0_d;
        const double q_bar = -(double)lb * (double)lam / dn * sd.inv_rho0_d * sd.inv_rho0_d;
        const double g2 = 2.0 * q_bar;
        reinterpret_cast<float*>(cur)[4 * (size_t)s + 3] = (float)rho_bar;
        db[s] = make_float4((float)(g2 * ax.x), (float)(g2 * ax.y), (float)(g2 * ax.z), (float)q_bar);
    });
}

// predictPositions backwards (x* = x + (v + g dt) dt): pbar of iterate 0
// (own + remainder) -> xbar_in[i] (added to the prev part), vbar_in[i], and
// this block's sums of gdev (6 doubles) for the gravity recursion.  A
// non-finite xbar_in (the adjoint outgrew FP32) raises kErrAdjointRange.
__global__ __launch_bounds__(kBlock) void k_predict_vjp(const float4* __restrict__ pacc,
                                                       float4* __restrict__ prem, uint32_t* __restrict__ rcnt,
                                                       const float4* __restrict__ rval,
                                                       const uint32_t* __restrict__ sorted_index, StepArgs a,
                                                       float4* __restrict__ xbar_in, float4* __restrict__ vbar_in,
                                                       const double* __restrict__ gdev, double* __restrict__ gpart,
                                                       uint32_t* __restrict__ err) {
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    double g[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (s < a.n) {
        const float4 pb = pacc[s], ro = prem[s];
        float in[3];
        inbox_sum<3, 1>(rval + (size_t)s * kInbox, min(rcnt[s], kInbox), in);
        prem[s] = zero4();
        rcnt[s] = 0u;
        const float b[3] = {pb.x + (in[0] + ro.x), pb.y + (in[1] + ro.y), pb.z + (in[2] + ro.z)};
        const uint32_t i = sorted_index[s];
        const float4 xp = xbar_in[i];
        const float4 xn = make_float4(xp.x + b[0], xp.y + b[1], xp.z + b[2], 0.f);
        xbar_in[i] = xn;
        vbar_in[i] = make_float4(b[0] * a.dt, b[1] * a.dt, b[2] * a.dt, 0.f);
        if (!(isfinite(xn.x) && isfinite(xn.y) && isfinite(xn.z))) atomicOr(err, kErrAdjointRange);
#pragma unroll
        for (int k = 0; k < 6; ++k) g[k] = gdev[6 * (size_t)s + k];
    }
    block_sum6(g, gpart + 6 * (size_t)blockIdx.x);
}

// gpart[6 b ..] = this block's sums of (xbar.xyz, vbar.xyz) over particles.
__global__ __launch_bounds__(kBlock) void k_sum_state(const float4* __restrict__ xbar,
                                                     const float4* __restrict__ vbar, uint32_t n,
                                                     double* __restrict__ gpart) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    double g[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (i < n) {
        const float4 x = xbar[i], v = vbar[i];
        g[0] = x.x; g[1] = x.y; g[2] = x.z;
        g[3] = v.x; g[4] = v.y; g[5] = v.z;
    }
    block_sum6(g, gpart + 6 * (size_t)blockIdx.x);
}

// Sums gpart over blocks in a fixed order into d[6] (thread 0 holds them).
__device__ __forceinline__ void reduce_blocks6(const double* __restrict__ gpart, uint32_t nblocks, double d[6]) {
    __shared__ double red[6][kBlock];
    double g[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (uint32_t b = threadIdx.x; b < nblocks; b += kBlock)
#pragma unroll
        for (int k = 0; k < 6; ++k) g[k] += gpart[6 * (size_t)b + k];
#pragma unroll
    for (int k = 0; k < 6; ++k) red[k][threadIdx.x] = g[k];
    __syncthreads();
    for (uint32_t w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w)
#pragma unroll
            for (int k = 0; k < 6; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) d[k] = red[k][0];
}

// gstate = (Sx, Sv, gsum) <- (sums of the adjoint inputs, 0).
__global__ __launch_bounds__(kBlock) void k_init_gravity(const double* __restrict__ gpart, uint32_t nblocks,
                                                        double* __restrict__ gstate, float* __restrict__ gout) {
    double d[6];
    reduce_blocks6(gpart, nblocks, d);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < 6; ++k) gstate[k] = d[k];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            gstate[6 + k] = 0.0;
            gout[k] = 0.f;
        }
    }
}

// One step of the gravity recursion (see k_finalize_vjp); gout = (float)gsum.
__global__ __launch_bounds__(kBlock) void k_reduce_gravity(const double* __restrict__ gpart, uint32_t nblocks,
                                                          double dt, double c, double* __restrict__ gstate,
                                                          float* __restrict__ gout) {
    double d[6];
    reduce_blocks6(gpart, nblocks, d);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double sx = gstate[k], sv = gstate[3 + k];
            const double p0 = sx + c * sv + d[k];
            gstate[k] = p0 - c * sv - d[3 + k];
            gstate[3 + k] = dt * p0;
            gstate[6 + k] += dt * dt * p0;
            gout[k] = (float)gstate[6 + k];
        }
    }
}

// ---- adjoint: stencil stages --------------------------------------------

// Geometry of one pair from FP32 positions: r = p_s - p_c, 1/|r|, and the
// spiky gradient magnitude f and its derivative f' (0 outside 1e-6 < r < h).
// Two float4 loaded per neighbour (position and adjoint input).
struct Pair4 {
    float4 a, b;
};

struct PairGeo {
    float rx, ry, rz, r2, inv_r, ux, uy, uz, f, fp, f_r;
};

__device__ __forceinline__ PairGeo pair_geo(const float4& pi, const float4& pj, const SolverDev& sd) {
    PairGeo g;
    g.rx = pi.x - pj.x;
    g.ry = pi.y - pj.y;
    g.rz = pi.z - pj.z;
    g.r2 = dist2(g.rx, g.ry, g.rz);
    g.inv_r = __builtin_amdgcn_rsqf(g.r2);
    const float rl = g.r2 * g.inv_r;
    // coincident particles (r2 = 0: two clamped into the same corner) give
    // inv_r = inf and rl = NaN; like r <= 1e-6 they fall outside the range
    // and contribute no spiky term, with a zero unit vector
    if (rl > 1e-6f && rl < sd.h) {
        const float dd = sd.h - rl;
        g.ux = g.rx * g.inv_r;
        g.uy = g.ry * g.inv_r;
        g.uz = g.rz * g.inv_r;
        g.f = -sd.spiky * dd * dd;
        g.fp = 2.0f * sd.spiky * dd;
        g.f_r = g.f * g.inv_r;
    } else {
        g.ux = g.uy = g.uz = 0.f;
        g.f = g.fp = g.f_r = 0.f;
    }
    return g;
}

// bin(s) lies in the stencil of the cell of p: |bin - cell(p)| <= 1 per axis.
__device__ __forceinline__ bool in_stencil(const Cell3& bin, const float4& p, const GridDev& g) {
    const Cell3 c = cell_of(p.x, p.y, p.z, g);
    return abs(bin.x - c.x) <= 1 && abs(bin.y - c.y) <= 1 && abs(bin.z - c.z) <= 1;
}

__device__ __forceinline__ float ipow_f(float x, uint32_t n) {
    float result = 1.0f, b = x;
    while (n) {
        if (n & 1u) result *= b;
        n >>= 1;
        if (n) b *= b;
    }
    return result;
}

// The backward stencil kernels run kAdjSub lanes per slot: a slot's listed
// neighbours are split between them by list block (lane `sub` takes blocks
// sub, sub + kAdjSub, ...), and their partial sums are added across the
// lanes at the end.  With one lane per slot the 64K differentiable workload
// ran one wave per SIMD, each waiting on its own neighbour loads; four lanes
// per slot give each SIMD four waves with a quarter of the walk each.
// A slot whose neighbours are not listed is walked by its first lane alone.
static_assert(kAdjSub == 4, "group_sum adds over four lanes");

template <class Load, class Pair>
__device__ __forceinline__ void for_each_neighbor_split(uint32_t s, uint32_t sub, const float4& pi,
                                                        const SceneDev& sc, const float4* __restrict__ cur,
                                                        const uint32_t* __restrict__ cell_start, uint32_t n,
                                                        const NbrList& nl, Load&& load, Pair&& pair) {
    const uint32_t cnt = nl.count[s];
    if (cnt <= kMaxList) {
        const uint4* p = nl.blk + s;
#pragma unroll 1
        for (uint32_t b = sub; 4u * b < cnt; b += kAdjSub) {
            const uint4 e = p[(size_t)b * nl.stride];
            const auto v0 = load(e.x), v1 = load(e.y), v2 = load(e.z), v3 = load(e.w);
            pair(e.x, v0);
            if (4u * b + 1u < cnt) pair(e.y, v1);
            if (4u * b + 2u < cnt) pair(e.z, v2);
            if (4u * b + 3u < cnt) pair(e.w, v3);
        }
    } else if (sub == 0u) {
        for_each_neighbor_of<true, false>(s, pi, sc, cur, cell_start, n, nl, load, pair);
    }
}

// The sum of v over the kAdjSub lanes of this lane's slot (every lane of the
// group must be active).
__device__ __forceinline__ float group_sum(float v) {
    v += __shfl_xor(v, 1);
    v += __shfl_xor(v, 2);
    return v;
}

// applyDeltaSorted backwards for slot s.  cur = iterate k with lambda_k in
// .w, bins = iterate 0 (the binned positions), rbar from k_delta_vjp_pre.
//
// Both roles of a pair at once: with dR = Rbar_s - Rbar_c (role B present)
// or Rbar_s (role A only), H even and G odd in r, the pair adds
//   lbar_s += f (r^.dR) / rho0,
//   pbar_s += c H dR / rho0 + (f (r^.dR) / rho0) grad s,
// and without role B, c's share of role A (the same values, negated) goes
// to c's inbox.
template <uint32_t kB>
__global__ __launch_bounds__(kB) void k_delta_vjp(const float4* __restrict__ cur,
                                                 const float4* __restrict__ bins,
                                                 const uint32_t* __restrict__ cell_start, StepArgs a,
                                                 const float4* __restrict__ rbar, float* __restrict__ lbar,
                                                 float* __restrict__ lrem, float4* __restrict__ pacc,
                                                 float4* __restrict__ prem, uint32_t* __restrict__ rcnt,
                                                 float4* __restrict__ rval, NbrList nl) {
    const uint32_t t = xcd_block(blockIdx.x, gridDim.x) * kB + threadIdx.x;
    const uint32_t s = t / kAdjSub, sub = t % kAdjSub;
    if (s >= a.n) return;  // the slot's whole lane group
    float lb = 0.f, px = 0.f, py = 0.f, pz = 0.f;
    with_scene(s, a, [&](const SceneDev& sc) {
        const SolverDev& sd = sc.sd;
        const float4 pi = cur[s], rs = rbar[s], b0 = bins[s];
        const Cell3 bin = cell_of(b0.x, b0.y, b0.z, sc.g);
        const float inv_rho0 = sd.inv_rho0;
        const float poly6 = (float)sd.poly6_d, inv_wq = (float)sd.inv_wq_d, sk = (float)sd.scorr_k_d;
        const uint32_t sn = sd.scorr_n;
        // d s_corr / d r = dsk * ratio^(n-1) * gp, gp = -6 poly6 (h^2 - r^2)^2
        const float dsk = -sk * (float)sn * inv_wq * (-6.0f * poly6);
        auto pair = [&](uint32_t c, const float4& pc, const float4& rc) {
            const PairGeo g = pair_geo(pi, pc, sd);
            const float diff = sd.h2 - g.r2;
            const float ratio = poly6 * diff * diff * diff * inv_wq;
            const float rn1 = sn ? ipow_f(ratio, sn - 1) : 0.f;
            const float scorr = sn ? -sk * rn1 * ratio : -sk;
            const float ds = dsk * rn1 * diff * diff;
            const float cc = (pi.w + pc.w + scorr) * inv_rho0;
            const bool both = in_stencil(bin, pc, sc.g);
            const float dx = both ? rs.x - rc.x : rs.x;
            const float dy = both ? rs.y - rc.y : rs.y;
            const float dz = both ? rs.z - rc.z : rs.z;
            const float dot = g.ux * dx + g.uy * dy + g.uz * dz;
            const float gr = g.f * dot * inv_rho0;
            const float hd = (g.fp - g.f_r) * dot;
            const float tx = cc * (hd * g.ux + g.f_r * dx) + gr * ds * g.rx;
            const float ty = cc * (hd * g.uy + g.f_r * dy) + gr * ds * g.ry;
            const float tz = cc * (hd * g.uz + g.f_r * dz) + gr * ds * g.rz;
            lb += gr;
            px += tx;
            py += ty;
            pz += tz;
            if (!both) remainder_put(rcnt, rval, lrem, prem, c, make_float4(gr, -tx, -ty, -tz), true);
        };
        for_each_neighbor_split(
            s, sub, pi, sc, cur, cell_start, a.n, nl, [&](uint32_t j) { return Pair4{cur[j], rbar[j]}; },
            [&](uint32_t j, const Pair4& v) { pair(j, v.a, v.b); });
    });
    lb = group_sum(lb);
    px = group_sum(px);
    py = group_sum(py);
    pz = group_sum(pz);
    if (sub == 0u) {
        lbar[s] = lb;
        const float4 acc = pacc[s];
        pacc[s] = make_float4(acc.x + px, acc.y + py, acc.z + pz, 0.f);
    }
}

// densityLambdaSorted backwards for slot s: cur = iterate k with rhobar in
// .w, db = (gbar, qbar).  Both roles at once: with the sums and difference
// over the pair's two ends (role B present) or s's own values (role A
// only), the pair adds
//   (rhobar_s + rhobar_c) gp r + H (gbar_s - gbar_c) + 2 (qbar_s + qbar_c) f f' r^.
template <uint32_t kB>
__global__ __launch_bounds__(kB) void k_density_vjp(const float4* __restrict__ cur,
                                                   const float4* __restrict__ bins,
                                                   const uint32_t* __restrict__ cell_start, StepArgs a,
                                                   const float4* __restrict__ db, float4* __restrict__ pacc,
                                                   float4* __restrict__ prem, uint32_t* __restrict__ rcnt,
                                                   float4* __restrict__ rval, NbrList nl) {
    const uint32_t t = xcd_block(blockIdx.x, gridDim.x) * kB + threadIdx.x;
    const uint32_t s = t / kAdjSub, sub = t % kAdjSub;
    if (s >= a.n) return;  // the slot's whole lane group
    float px = 0.f, py = 0.f, pz = 0.f;
    with_scene(s, a, [&](const SceneDev& sc) {
        const SolverDev& sd = sc.sd;
        const float4 pi = cur[s], ds4 = db[s], b0 = bins[s];
        const Cell3 bin = cell_of(b0.x, b0.y, b0.z, sc.g);
        const float gp6 = -6.0f * (float)sd.poly6_d;
        auto pair = [&](uint32_t c, const float4& pc, const float4& dc) {
            const PairGeo g = pair_geo(pi, pc, sd);
            const float diff = sd.h2 - g.r2;
            const bool both = in_stencil(bin, pc, sc.g);
            const float rho_b = both ? pi.w + pc.w : pi.w;
            const float q_b = both ? ds4.w + dc.w : ds4.w;
            const float gx = both ? ds4.x - dc.x : ds4.x;
            const float gy = both ? ds4.y - dc.y : ds4.y;
            const float gz = both ? ds4.z - dc.z : ds4.z;
            const float k = rho_b * gp6 * diff * diff;
            const float u = (g.fp - g.f_r) * (g.ux * gx + g.uy * gy + g.uz * gz) + 2.0f * q_b * g.f * g.fp;
            const float tx = k * g.rx + u * g.ux + g.f_r * gx;
            const float ty = k * g.ry + u * g.uy + g.f_r * gy;
            const float tz = k * g.rz + u * g.uz + g.f_r * gz;
            px += tx;
            py += ty;
            pz += tz;
            if (!both) remainder_put(rcnt, rval, nullptr, prem, c, make_float4(0.f, -tx, -ty, -tz), false);
        };
        for_each_neighbor_split(
            s, sub, pi, sc, cur, cell_start, a.n, nl, [&](uint32_t j) { return Pair4{cur[j], db[j]}; },
            [&](uint32_t j, const Pair4& v) { pair(j, v.a, v.b); });
    });
    px = group_sum(px);
    py = group_sum(py);
    pz = group_sum(pz);
    if (sub == 0u) {
        const float4 acc = pacc[s];
        pacc[s] = make_float4(acc.x + px, acc.y + py, acc.z + pz, 0.f);
    }
}

// ---- density at arbitrary positions and its adjoint (fresh binning) ----

// rho_i = W(0) + sum_j W(r_ij) (Constraints.metal:100-115) in original order.
__global__ __launch_bounds__(kBlock) void k_density_rho(const float4* __restrict__ sorted,
                                                       const uint32_t* __restrict__ cell_start,
                                                       const uint32_t* __restrict__ sorted_index, StepArgs a,
                                                       float* __restrict__ rho) {
    const uint32_t s = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (s >= a.n) return;
    with_scene(s, a, [&](const SceneDev& sc) {
        const SolverDev& sd = sc.sd;
        const float4 pi = sorted[s];
        double wsum = 0.0;
        auto term = [&](const float4& pj) {
            const double rx = (double)(pi.x - pj.x), ry = (double)(pi.y - pj.y), rz = (double)(pi.z - pj.z);
            const double diff = sd.h2_d - (rx * rx + ry * ry + rz * rz);
            wsum += diff * diff * diff;
        };
        for_each_neighbor(s, pi, sc, sorted, cell_start, a.n, [&](uint32_t j1, uint32_t j2, bool two) {
            const float4 q1 = sorted[j1], q2 = sorted[j2];
            term(q1);
            if (two) term(q2);
        });
        rho[sorted_index[s]] = (float)(sd.w0_d + sd.poly6_d * wsum);
    });
}

// sorted[s].w = rhobar[sorted_index[s]]
__global__ __launch_bounds__(kBlock) void k_gather_w(float4* __restrict__ sorted,
                                                    const uint32_t* __restrict__ sorted_index, uint32_t n,
                                                    const float* __restrict__ w) {
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= n) return;
    reinterpret_cast<float*>(sorted)[4 * (size_t)s + 3] = w[sorted_index[s]];
}

// xbar_i = sum_j (rhobar_i + rhobar_j) grad W(r_ij): with fresh binning the
// neighbour relation is symmetric, so slot s gathers both roles itself.
__global__ __launch_bounds__(kBlock) void k_density_rho_vjp(const float4* __restrict__ sorted,
                                                           const uint32_t* __restrict__ cell_start,
                                                           const uint32_t* __restrict__ sorted_index, StepArgs a,
                                                           float4* __restrict__ xbar) {
    const uint32_t s = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (s >= a.n) return;
    with_scene(s, a, [&](const SceneDev& sc) {
        const SolverDev& sd = sc.sd;
        const float4 pi = sorted[s];
        double px = 0.0, py = 0.0, pz = 0.0;
        auto term = [&](const float4& pj) {
            const double rx = (double)(pi.x - pj.x), ry = (double)(pi.y - pj.y), rz = (double)(pi.z - pj.z);
            const double diff = sd.h2_d - (rx * rx + ry * ry + rz * rz);
            const double k = -6.0 * sd.poly6_d * diff * diff * ((double)pi.w + (double)pj.w);
            px += k * rx;
            py += k * ry;
            pz += k * rz;
        };
        for_each_neighbor(s, pi, sc, sorted, cell_start, a.n, [&](uint32_t j1, uint32_t j2, bool two) {
            const float4 q1 = sorted[j1], q2 = sorted[j2];
            term(q1);
            if (two) term(q2);
        });
        xbar[sorted_index[s]] = make_float4((float)px, (float)py, (float)pz, 0.f);
    });
}

}  // namespace

// ---- launchers ---------------------------------------------------------------

#define MTLP_STENCIL_LAUNCH(KERNEL, block, n, stream, ...)                                           \
    do {                                                                                             \
        switch (block) {                                                                             \
            case 64:                                                                                 \
                hipLaunchKernelGGL(KERNEL<64>, dim3(blocks_for(n, 64)), dim3(64), 0, stream,         \
                                   __VA_ARGS__);                                                     \
                break;                                                                               \
            case 128:                                                                                \
                hipLaunchKernelGGL(KERNEL<128>, dim3(blocks_for(n, 128)), dim3(128), 0, stream,      \
                                   __VA_ARGS__);                                                     \
                break;                                                                               \
            case 512:                                                                                \
                hipLaunchKernelGGL(KERNEL<512>, dim3(blocks_for(n, 512)), dim3(512), 0, stream,      \
                                   __VA_ARGS__);                                                     \
                break;                                                                               \
            default:                                                                                 \
                hipLaunchKernelGGL(KERNEL<256>, dim3(blocks_for(n, 256)), dim3(256), 0, stream,      \
                                   __VA_ARGS__);                                                     \
                break;                                                                               \
        }                                                                                            \
    } while (0)

hipError_t launch_density_lambda_tape(float4* cur, const uint32_t* cell_start, const uint32_t* sorted_index,
                                      const StepArgs& a, float4* aux, float* lambda_orig, uint32_t* nbr_orig,
                                      const NbrList& nl, uint32_t block, hipStream_t s) {
    if (a.n == 0) return hipSuccess;
    switch (block) {
        case 64: density_lambda_tape_b<64>(cur, cell_start, sorted_index, a, aux, lambda_orig, nbr_orig, nl, s); break;
        case 128: density_lambda_tape_b<128>(cur, cell_start, sorted_index, a, aux, lambda_orig, nbr_orig, nl, s); break;
        case 512: density_lambda_tape_b<512>(cur, cell_start, sorted_index, a, aux, lambda_orig, nbr_orig, nl, s); break;
        default: density_lambda_tape_b<256>(cur, cell_start, sorted_index, a, aux, lambda_orig, nbr_orig, nl, s); break;
    }
    return hipGetLastError();
}

hipError_t launch_apply_delta_tape(const float4* cur, const uint32_t* cell_start, const StepArgs& a,
                                   float4* next, float4* rr, const NbrList& nl, uint32_t block, hipStream_t s) {
    if (a.n == 0) return hipSuccess;
    switch (block) {  // the heavy queue's workgroups in front
        case 64:
            hipLaunchKernelGGL(k_apply_delta_tape<64>, dim3(blocks_for(a.n, 64) + heavy_blocks<64>() + split_blocks<64>()), dim3(64), 0,
                               s, cur, cell_start, a, next, rr, nl);
            break;
        case 128:
            hipLaunchKernelGGL(k_apply_delta_tape<128>, dim3(blocks_for(a.n, 128) + heavy_blocks<128>() + split_blocks<128>()), dim3(128),
                               0, s, cur, cell_start, a, next, rr, nl);
            break;
        case 512:
            hipLaunchKernelGGL(k_apply_delta_tape<512>, dim3(blocks_for(a.n, 512) + heavy_blocks<512>() + split_blocks<512>()), dim3(512),
                               0, s, cur, cell_start, a, next, rr, nl);
            break;
        default:
            hipLaunchKernelGGL(k_apply_delta_tape<256>, dim3(blocks_for(a.n, 256) + heavy_blocks<256>() + split_blocks<256>()), dim3(256),
                               0, s, cur, cell_start, a, next, rr, nl);
            break;
    }
    return hipGetLastError();
}

hipError_t launch_finalize_tape(const float4* fin, const uint32_t* sorted_index, const StepArgs& a,
                                const float4* x_in, float4* x_out, float4* v_out, hipStream_t s) {
    if (a.n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_finalize_tape, dim3(blocks_for(a.n, kBlock)), dim3(kBlock), 0, s, fin, sorted_index,
                       a, x_in, x_out, v_out);
    return hipGetLastError();
}

hipError_t launch_gravity_init(const AdjointBufs& b, const float4* xbar_T, const float4* vbar_T, uint32_t n,
                               float* gout, hipStream_t s) {
    const dim3 pg(blocks_for(n, kBlock)), pb(kBlock);
    if (n) hipLaunchKernelGGL(k_sum_state, pg, pb, 0, s, xbar_T, vbar_T, n, b.gpart);
    hipLaunchKernelGGL(k_init_gravity, dim3(1), pb, 0, s, b.gpart, n ? pg.x : 0u, b.gstate, gout);
    return hipGetLastError();
}

hipError_t launch_step_vjp(const StepTape& t, const AdjointBufs& b, const StepArgs& a, uint32_t iterations,
                           float damping, const float4* xbar_out, const float4* vbar_out, float4* xbar_in,
                           float4* vbar_in, float* gout, uint32_t block, hipStream_t s) {
    if (a.n == 0) return hipSuccess;
    const dim3 pg(blocks_for(a.n, kBlock)), pb(kBlock);
    hipLaunchKernelGGL(k_finalize_vjp, pg, pb, 0, s, t.iterate(iterations), t.sorted_index, a, t.x_in,
                       xbar_out, vbar_out, b.pacc, xbar_in, b.gdev);
    for (uint32_t k = iterations; k-- > 0;) {
        hipLaunchKernelGGL(k_delta_vjp_pre, pg, pb, 0, s, t.iterate(k), t.delta(k), a, b.pacc, b.prem, b.rcnt,
                           b.rval, b.rbar, b.gdev);
        MTLP_STENCIL_LAUNCH(k_delta_vjp, block, a.n * kAdjSub, s, t.iterate(k), t.iterate(0), t.cell_start, a,
                            b.rbar, b.lbar, b.lrem, b.pacc, b.prem, b.rcnt, b.rval, t.list(k));
        hipLaunchKernelGGL(k_density_vjp_pre, pg, pb, 0, s, t.iterate(k), t.grad_sums(k), a, b.lbar, b.lrem,
                           b.rcnt, b.rval, b.db);
        MTLP_STENCIL_LAUNCH(k_density_vjp, block, a.n * kAdjSub, s, t.iterate(k), t.iterate(0), t.cell_start, a,
                            b.db, b.pacc, b.prem, b.rcnt, b.rval, t.list(k));
    }
    hipLaunchKernelGGL(k_predict_vjp, pg, pb, 0, s, b.pacc, b.prem, b.rcnt, b.rval, t.sorted_index, a, xbar_in,
                       vbar_in, b.gdev, b.gpart, b.err);
    const double dt = (double)a.dt;
    hipLaunchKernelGGL(k_reduce_gravity, dim3(1), pb, 0, s, b.gpart, pg.x, dt, (1.0 / dt) * (double)damping,
                       b.gstate, gout);
    return hipGetLastError();
}

hipError_t launch_density_rho(const float4* sorted, const uint32_t* cell_start, const uint32_t* sorted_index,
                              const StepArgs& a, float* rho, hipStream_t s) {
    if (a.n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_density_rho, dim3(blocks_for(a.n, kBlock)), dim3(kBlock), 0, s, sorted, cell_start,
                       sorted_index, a, rho);
    return hipGetLastError();
}

hipError_t launch_density_rho_vjp(float4* sorted, const uint32_t* cell_start, const uint32_t* sorted_index,
                                  const StepArgs& a, const float* rhobar, float4* xbar, hipStream_t s) {
    if (a.n == 0) return hipSuccess;
    const dim3 pg(blocks_for(a.n, kBlock)), pb(kBlock);
    hipLaunchKernelGGL(k_gather_w, pg, pb, 0, s, sorted, sorted_index, a.n, rhobar);
    hipLaunchKernelGGL(k_density_rho_vjp, pg, pb, 0, s, sorted, cell_start, sorted_index, a, xbar);
    return hipGetLastError();
}

}  // namespace mtlp
