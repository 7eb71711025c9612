// mtlp_stencil.h -- per-slot device code of the PBF constraint passes,
// shared by the step kernels (solver.hip) and the taped forward and adjoint
// kernels (adjoint.hip).
//
// Layout: the sorted positions are float4; the density pass stores lambda_i
// in .w of its own slot, so the delta pass reads (x_j, lambda_j) with one
// 16-byte load per candidate.  The 27-cell stencil is walked as 9 rows of
// three x-adjacent cells: their slots are contiguous in cell order, so each
// row is one slot range [cell_start[x0], cell_start[x1 + 1]), visited in the
// reference's z, y, x, slot order.
#pragma once

#include "mtlp_device.h"

namespace mtlp {

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

// The chunk's mask from the sign bits shifted in by a filter's groups of
// eight: candidate k went in at bit 8 * groups - 1 - k; bits past the
// chunk's len candidates and the bit of s itself are cleared.
__device__ __forceinline__ uint32_t chunk_mask(uint32_t shifted, uint32_t c0, uint32_t len, uint32_t s) {
    uint32_t mask = __builtin_bitreverse32(shifted) >> (32u - 8u * ((min(len, 32u) + 7u) / 8u));
    mask &= len >= 32u ? 0xffffffffu : (1u << len) - 1u;
    const uint32_t self = s - c0;  // wraps to a large value when s < c0
    return self < 32u ? mask & ~(1u << self) : mask;
}

// Candidates are filtered before any SPH arithmetic: bit k of the result is
// set when slot c0 + k (k < c1 - c0 <= 32) is a neighbour of s, i.e.
// |p_s - p_j|^2 < h^2 and j != s.  The expensive terms then run only on set
// bits, so lanes of a wave stay busy on real neighbours.  Loads go out eight
// at a time from one base address before any is consumed; they may read up
// to kChunkSlack slots past the last particle (position buffers carry that
// slack), and the bits of those and of s itself are masked off at the end.
__device__ __forceinline__ uint32_t neighbor_mask(const float4* __restrict__ cur, uint32_t c0,
                                                  uint32_t c1, uint32_t s, float4 pi, float h2) {
    const float4* q = cur + c0;
    const uint32_t len = c1 - c0;  // >= 1
    uint32_t mask = 0;
#pragma unroll
    for (uint32_t k0 = 0; k0 < 32; k0 += 8) {
        if (k0 >= len) break;
        float4 v[8];
#pragma unroll
        for (uint32_t u = 0; u < 8; ++u) v[u] = q[k0 + u];
#pragma unroll
        for (uint32_t u = 0; u < 8; ++u) {
            // shift in the sign of the neighbour key: set for a neighbour
            const float d = nbr_key(pi.x - v[u].x, pi.y - v[u].y, pi.z - v[u].z, h2);
            mask = __builtin_amdgcn_alignbit(mask, __float_as_uint(d), 31);
        }
    }
    return chunk_mask(mask, c0, len, s);
}

// neighbor_mask over a packed (x, y, z) copy of the iterate: the eight
// candidates of a group are 96 contiguous bytes, fetched as six 16-byte
// loads (4-byte aligned) instead of eight 12-byte ones, which is the
// filter's share of the vector-memory instructions that bound the density
// pass.  Same values, same arithmetic, same bits as neighbor_mask.  Reads
// up to kChunkSlack slots past the last particle, like neighbor_mask.
struct __attribute__((packed, aligned(4))) Words4 {
    float a, b, c, d;
};
__device__ __forceinline__ uint32_t neighbor_mask_xyz(const float* __restrict__ xyz, uint32_t c0, uint32_t c1,
                                                      uint32_t s, float4 pi, float h2) {
    const float* q = xyz + 3 * (size_t)c0;
    const uint32_t len = c1 - c0;  // >= 1
    uint32_t mask = 0;
#pragma unroll
    for (uint32_t k0 = 0; k0 < 32; k0 += 8) {
        if (k0 >= len) break;
        float t[24];
#pragma unroll
        for (uint32_t u = 0; u < 6; ++u) {
            const Words4 w = reinterpret_cast<const Words4*>(q + 3 * k0)[u];
            t[4 * u] = w.a;
            t[4 * u + 1] = w.b;
            t[4 * u + 2] = w.c;
            t[4 * u + 3] = w.d;
        }
#pragma unroll
        for (uint32_t u = 0; u < 8; ++u) {
            const float d = nbr_key(pi.x - t[3 * u], pi.y - t[3 * u + 1], pi.z - t[3 * u + 2], h2);
            mask = __builtin_amdgcn_alignbit(mask, __float_as_uint(d), 31);
        }
    }
    return chunk_mask(mask, c0, len, s);
}

// Slot ranges of the 9 stencil rows (z-major, then y), fetched together.
// Rows outside the grid are empty.  Row (y, z) holds cells x0..x1, whose
// slots are contiguous in cell order.
struct Rows {
    uint32_t a[9], b[9];
};

__device__ __forceinline__ uint32_t ld_u32(const uint32_t* __restrict__ base, uint32_t i) {
    // 32-bit byte offset from a uniform base: one VGPR per address
    return *reinterpret_cast<const uint32_t*>(reinterpret_cast<const char*>(base) + i * 4u);
}

__device__ __forceinline__ Rows stencil_rows(const Stencil& st, const uint32_t* __restrict__ cell_start,
                                             const GridDev& g, uint32_t cell_off, uint32_t n) {
    Rows r;
    uint32_t ia[9], ib[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const int z = st.z0 + k / 3, y = st.y0 + k % 3;
        const bool ok = z <= st.z1 && y <= st.y1;
        const uint32_t row = cell_off + (ok ? (uint32_t)y * g.dx + (uint32_t)z * g.dxy : 0u);
        ia[k] = row + (uint32_t)st.x0;
        ib[k] = row + (uint32_t)st.x1 + 1u;
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        r.a[k] = ld_u32(cell_start, ia[k]);
        r.b[k] = ld_u32(cell_start, ib[k]);
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const int z = st.z0 + k / 3, y = st.y0 + k % 3;
        const bool ok = z <= st.z1 && y <= st.y1;
        r.a[k] = ok ? r.a[k] : 0u;
        r.b[k] = ok ? min(r.b[k], n) : 0u;
    }
    return r;
}

// Visits the stencil of slot s in walk order, one chunk of at most 32
// consecutive candidate slots at a time: body(c0, mask), bit k of mask set
// when slot c0 + k is a neighbour (neighbor_mask).
// With kUnrollRows the 9-row loop is unrolled, so the rows' slot ranges
// stay in registers: indexed by a loop counter they go to scratch memory
// when LDS holds other data (the density pass's chunk columns), 18 extra
// vector-memory reads per slot.  With kXyz the candidates come from xyz,
// the iterate's packed copy (a template choice, so each kernel carries one
// filter: both in one kernel measured slower on the ensemble's later steps).
template <bool kUnrollRows = false, bool kXyz = false, class Body>
__device__ __forceinline__ void for_each_chunk(uint32_t s, const float4& pi, const SceneDev& sc,
                                               const float4* __restrict__ cur,
                                               const uint32_t* __restrict__ cell_start, uint32_t n,
                                               Body&& body, const float* __restrict__ xyz = nullptr) {
    const Stencil st = stencil_of(pi.x, pi.y, pi.z, sc.g);
    const Rows rows = stencil_rows(st, cell_start, sc.g, sc.cell_off, n);
    if constexpr (kUnrollRows) {
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            const uint32_t b = rows.b[k];
            for (uint32_t c0 = rows.a[k]; c0 < b; c0 += 32) {
                const uint32_t c1 = min(c0 + 32u, b);
                if constexpr (kXyz) body(c0, neighbor_mask_xyz(xyz, c0, c1, s, pi, sc.sd.h2));
                else body(c0, neighbor_mask(cur, c0, c1, s, pi, sc.sd.h2));
            }
        }
    } else {
#pragma unroll 1
        for (int k = 0; k < 9; ++k) {
            const uint32_t b = rows.b[k];
            for (uint32_t c0 = rows.a[k]; c0 < b; c0 += 32)
                body(c0, neighbor_mask(cur, c0, min(c0 + 32u, b), s, pi, sc.sd.h2));
        }
    }
}

// Visits every neighbour j of slot s (|p_s - p_j|^2 < h^2 in FP32, j != s,
// j in the 27-cell stencil of p_s's cell), two per trip so both loads are in
// flight together: body(j1, j2, two) per trip, in slot order.
template <class Body>
__device__ __forceinline__ void for_each_neighbor(uint32_t s, const float4& pi, const SceneDev& sc,
                                                  const float4* __restrict__ cur,
                                                  const uint32_t* __restrict__ cell_start, uint32_t n,
                                                  Body&& body) {
    for_each_chunk(s, pi, sc, cur, cell_start, n, [&](uint32_t c0, uint32_t mask) {
        while (mask) {
            const uint32_t j1 = c0 + (uint32_t)__builtin_ctz(mask);
            mask &= mask - 1u;
            const bool two = mask != 0u;
            const uint32_t j2 = two ? c0 + (uint32_t)__builtin_ctz(mask) : j1;
            if (two) mask &= mask - 1u;
            body(j1, j2, two);
        }
    });
}

// Neighbours of slot s in walk order from chunk `first` of its stencil walk
// on (chunks as for_each_chunk cuts them), testing one candidate at a time:
// the case of more chunks than the density pass keeps (kChunkScratch), small
// in registers rather than fast.
template <class Fn>
__device__ __forceinline__ void neighbors_from_chunk(uint32_t s, const float4& pi, const SceneDev& sc,
                                                     const float4* __restrict__ cur,
                                                     const uint32_t* __restrict__ cell_start, uint32_t n,
                                                     uint32_t first, Fn&& fn) {
    const Stencil st = stencil_of(pi.x, pi.y, pi.z, sc.g);
    const GridDev& g = sc.g;
    uint32_t k0 = 0;  // chunks in the rows before row k
#pragma unroll 1
    for (int k = 0; k < 9; ++k) {
        const int z = st.z0 + k / 3, y = st.y0 + k % 3;
        if (z > st.z1 || y > st.y1) continue;
        const uint32_t row = sc.cell_off + (uint32_t)y * g.dx + (uint32_t)z * g.dxy;
        const uint32_t a = cell_start[row + (uint32_t)st.x0];
        const uint32_t b = min(cell_start[row + (uint32_t)st.x1 + 1], n);
        const uint32_t nch = b > a ? (b - a + 31u) / 32u : 0u;
#pragma unroll 1
        for (uint32_t c = first > k0 ? a + 32u * (first - k0) : a; c < b; ++c) {
            const float4 q = cur[c];
            if (c != s && nbr_key(pi.x - q.x, pi.y - q.y, pi.z - q.z, sc.sd.h2) < 0.0f) fn(c);
        }
        k0 += nch;
    }
}

// Walks slot s's stencil and keeps the chunks (c0, mask) that hold a
// neighbour: the first kLdsChunks of them in this lane's LDS column le
// (stride lst), kept chunks kLdsChunks .. kChunkScratch - 1 at their index
// in nl.spill.  Returns the number of kept chunks, the number of neighbours
// in nbrs and, when more than kChunkScratch chunks are kept, in over the
// walk index (counting every chunk for_each_chunk cuts) of the first one
// past them.  Keeping only non-empty chunks lets the terms walk fetch a
// lane's next chunk with one read instead of a loop (a corner row of the
// stencil often has no neighbour): 1M dam break 1.34 -> 1.25 ms per step,
// 1.81 -> 1.37 ms after impact.  With kXyz the candidates come from nl.xyz,
// the iterate's packed copy.
template <bool kXyz>
__device__ __forceinline__ uint32_t record_chunks(uint32_t s, const float4& pi, const SceneDev& sc,
                                                  const float4* __restrict__ cur,
                                                  const uint32_t* __restrict__ cell_start, uint32_t n,
                                                  const NbrList& nl, uint2* le, uint32_t lst, uint32_t& nbrs,
                                                  uint32_t& over) {
    uint32_t K = 0, c = 0, walked = 0;
    over = 0xffffffffu;
    for_each_chunk<true, kXyz>(s, pi, sc, cur, cell_start, n, [&](uint32_t c0, uint32_t mask) {
        if (mask) {
            if (K == kChunkScratch) over = walked;
            const uint2 e = make_uint2(c0, mask);
            // two stores under conditions of their own: as if / else if, the two
            // destinations merge into one generic pointer (flat stores)
            if (K < kLdsChunks) le[K * lst] = e;
            if (K - kLdsChunks < kChunkScratch - kLdsChunks) nl.spill[(size_t)K * nl.stride + s] = e;
            ++K;
            c += (uint32_t)__builtin_popcount(mask);
        }
        ++walked;
    }, nl.xyz);
    nbrs = c;
    return K;
}

// Visits the set bits of K <= kChunkScratch non-empty chunks (c0, mask) in order,
// chunk k from chunk(k): term(j, load(j)) per neighbour j.  Every lane moves
// on to its next chunk by itself, so a wave runs max(neighbours) / 4 trips
// rather than the sum over chunks of each chunk's largest count.  A trip
// takes four neighbours, with their loads in flight before any term runs,
// and with `listing` stores them as block t of slot s's list (padded with
// s), so the lanes of a wave store block t together.
// From chunk kAheadFrom on, the next chunk is fetched when this one starts,
// so a walk through the global scratch waits one memory latency per chunk
// less (the delta pass's heavy slots from the first chunk; the density pass
// once its walk leaves the chunks kept in LDS, which need no prefetch).
template <uint32_t kAheadFrom = 0xffffffffu, class Chunk, class Load, class Term>
__device__ __forceinline__ void visit_chunks(Chunk&& chunk, uint32_t K, uint32_t s, const NbrList& nl,
                                             bool listing, Load&& load, Term&& term) {
    uint32_t k = 0, b = 0, m = 0;
    uint2 ahead = kAheadFrom == 0u && K ? chunk(0u) : make_uint2(0u, 0u);
    auto next = [&](uint32_t& j) {
        if (m == 0u && k < K) {  // chunks are non-empty
            const uint2 e = k >= kAheadFrom ? ahead : chunk(k);
            if (k + 1u >= kAheadFrom && k + 1u < K) ahead = chunk(k + 1u);
            ++k;
            b = e.x;
            m = e.y;
        }
        if (m == 0u) return false;
        j = b + (uint32_t)__builtin_ctz(m);
        m &= m - 1u;
        return true;
    };
    uint4* out = nl.blk + s;
    const size_t st = nl.stride;
#pragma unroll 1
    for (uint32_t t = 0;; ++t) {
        uint32_t j0 = s, j1 = s, j2 = s, j3 = s;
        if (!next(j0)) break;
        const bool h1 = next(j1);
        const bool h2 = h1 && next(j2);
        const bool h3 = h2 && next(j3);
        const auto v0 = load(j0), v1 = load(j1), v2 = load(j2), v3 = load(j3);
        if (listing && t < kListBlocks) out[(size_t)t * st] = make_uint4(j0, j1, j2, j3);
        term(j0, v0);
        if (h1) term(j1, v1);
        if (h2) term(j2, v2);
        if (h3) term(j3, v3);
        if (!h3) break;
    }
}

// The cnt <= kMaxList listed neighbours of slot s in order, a block of four
// per trip: v = load(j) for all four entries (padding holds s), then
// term(j, v) for those below cnt.  The next block is fetched a trip ahead.
template <class Load, class Term>
__device__ __forceinline__ void for_each_listed(uint32_t s, uint32_t cnt, const NbrList& nl, Load&& load,
                                                Term&& term) {
    const uint4* p = nl.blk + s;
    const size_t st = nl.stride;
    uint4 e = p[0];  // read even when cnt = 0: in bounds, unused
#pragma unroll 1
    for (uint32_t k = 0; k < cnt; k += 4) {
        const auto v0 = load(e.x), v1 = load(e.y), v2 = load(e.z), v3 = load(e.w);
        const uint32_t b = min((k >> 2) + 1u, kListBlocks - 1u);
        const uint4 ce = e;
        e = p[(size_t)b * st];
        term(ce.x, v0);
        if (k + 1 < cnt) term(ce.y, v1);
        if (k + 2 < cnt) term(ce.z, v2);
        if (k + 3 < cnt) term(ce.w, v3);
    }
}

// The neighbours of slot s in walk order: from the list when it holds them
// all (kList and count <= kMaxList); with kScratch (the pass right after the
// density pass that wrote the list) from the chunks the density pass left in
// nl.spill when the count says so; else from the stencil.
template <bool kList, bool kScratch, class Load, class Term>
__device__ __forceinline__ void for_each_neighbor_of(uint32_t s, const float4& pi, const SceneDev& sc,
                                                     const float4* __restrict__ cur,
                                                     const uint32_t* __restrict__ cell_start, uint32_t n,
                                                     const NbrList& nl, Load&& load, Term&& term) {
    if (kList) {
        const uint32_t cnt = nl.count[s];
        if (cnt <= kMaxList) {
            for_each_listed(s, cnt, nl, load, term);
            return;
        }
        const uint32_t K = cnt & 0xffffu;
        if (kScratch && (cnt & kCountChunks) && K != 0xffffu) {
            visit_chunks<0u>([&](uint32_t k) { return nl.spill[(size_t)k * nl.stride + s]; }, K, s, nl, false,
                             load, term);
            return;
        }
    }
    for_each_neighbor(s, pi, sc, cur, cell_start, n, [&](uint32_t j1, uint32_t j2, bool two) {
        const auto v1 = load(j1), v2 = load(j2);
        term(j1, v1);
        if (two) term(j2, v2);
    });
}

// One neighbour's contribution to the density sum (FP64) and to the
// constraint gradient terms (Constraints.metal:110-119): its (h^2 - r^2)
// and squared in FP64 and grad W in FP32, added to the sums in walk order
// by add_density.  Every fused multiply-add is written out (contraction
// off), so the bits do not depend on the context the compiler inlines this
// into.
struct DensityTerm {
    double diff, diff2;
    float ax, ay, az;
};

__device__ __forceinline__ DensityTerm density_pair(const float4& pi, const float4& pj, const SolverDev& sd) {
#pragma clang fp contract(off)
    DensityTerm t;
    const float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
    const float r2 = dist2(rx, ry, rz);
    const double dxd = rx, dyd = ry, dzd = rz;
    t.diff = sd.h2_d - fma(dzd, dzd, fma(dyd, dyd, dxd * dxd));
    t.diff2 = t.diff * t.diff;
    // gradient terms only enter lambda's denominator: hardware 1/sqrt (~1 ulp)
    const float inv_r = __builtin_amdgcn_rsqf(r2);
    const float rl = r2 * inv_r;
    // outside (1e-6, h) the term is zero (a select: r2 = 0 makes inv_r inf)
    const float dd = sd.h - rl;
    const float sc = rl > 1e-6f && rl < sd.h ? -sd.spiky * dd * dd * inv_r : 0.0f;
    t.ax = sc * rx;
    t.ay = sc * ry;
    t.az = sc * rz;
    return t;
}

__device__ __forceinline__ void add_density(const DensityTerm& t, double& wsum, float& gx, float& gy, float& gz,
                                            float& gsq) {
#pragma clang fp contract(off)
    wsum = fma(t.diff, t.diff2, wsum);
    gx += t.ax;
    gy += t.ay;
    gz += t.az;
    gsq = fmaf(t.az, t.az, fmaf(t.ay, t.ay, fmaf(t.ax, t.ax, gsq)));
}

__device__ __forceinline__ void density_term(const float4& pi, const float4& pj, const SolverDev& sd,
                                             double& wsum, float& gx, float& gy, float& gz, float& gsq) {
    add_density(density_pair(pi, pj, sd), wsum, gx, gy, gz, gsq);
}

// lambda = -C / (D + eps) from the sums: C = rho / rho0 - 1 in FP64 with
// rho = W(0) + poly6 sum, D = (|sum grad W|^2 + sum |grad W|^2) / rho0^2.
// Returns lambda; dterm gets D.
__device__ __forceinline__ float lambda_of(double wsum, float gx, float gy, float gz, float gsq,
                                           const SolverDev& sd, float& dterm) {
#pragma clang fp contract(off)
    const double constraint = fma(sd.poly6_d, wsum, sd.w0_d) * sd.inv_rho0_d - 1.0;
    dterm = (sd.inv_rho0 * sd.inv_rho0) * (fmaf(gz, gz, fmaf(gy, gy, gx * gx)) + gsq);
    const float denom = dterm + sd.eps;
    return (float)(-constraint / (double)denom);
}

// The list count code of a slot whose walk had K chunks and cnt neighbours
// (see NbrList); a heavy slot also goes to its queue.
__device__ __forceinline__ uint32_t list_code(uint32_t s, uint32_t K, uint32_t cnt, const NbrList& nl) {
    uint32_t code = kCountStencil;
    if (K <= kChunkScratch) code = cnt <= kMaxList ? cnt : kCountChunks | K;
    if (code == kCountStencil) {  // a wave of the delta pass's heavy waves
        const uint32_t q = atomicAdd(nl.heavy_count, 1u);
        if (q < kHeavyCap) {
            nl.heavy[q] = s;
            code |= kCountQueued;
        }
    } else if (code > kMaxList) {  // a lane of its heavy-lane workgroups
        const uint32_t q = atomicAdd(nl.lanes_count, 1u);
        if (q < nl.lanes_cap) {
            nl.lanes[q] = s;
            code |= kCountQueued;
        }
    }
    return code;
}

__device__ __forceinline__ double ipow_d(double x, uint32_t n) {
    if (n == 4u) {  // the reference's s_corr exponent: the loop below gives (x^2)^2 too
        const double x2 = x * x;
        return x2 * x2;
    }
    double result = 1.0, b = x;
    while (n) {
        if (n & 1u) result *= b;
        n >>= 1;
        if (n) b *= b;
    }
    return result;
}

// One neighbour's term of the position correction (Constraints.metal:171-181),
// in FP64 from the FP32 positions and lambdas (all exact in FP64): the
// coefficient c and r = p_i - p_j, the term being c r.  Every fused
// multiply-add is written out (contraction off), so each path that
// computes the term (per lane, or a wave per slot) gets the same bits.
struct DeltaTerm {
    double c, rx, ry, rz;
};

__device__ __forceinline__ DeltaTerm delta_coef(const float4& pi, double lam_i, const float4& pj,
                                                const SolverDev& sd) {
#pragma clang fp contract(off)
    DeltaTerm t;
    t.rx = (double)pi.x - (double)pj.x;
    t.ry = (double)pi.y - (double)pj.y;
    t.rz = (double)pi.z - (double)pj.z;
    const double r2 = fma(t.rz, t.rz, fma(t.rx, t.rx, t.ry * t.ry));
    // 1/r from the hardware estimate (~2^-23) plus one Newton step (~1e-14);
    // r2 = 0 gives NaN here and fails the range test below, like r <= 1e-6.
    const double y0 = __builtin_amdgcn_rsq(r2);
    const double inv_r = y0 * fma(y0, (-0.5 * r2) * y0, 1.5);
    const double rl = r2 * inv_r;
    const double dd = fma(-r2, inv_r, sd.h_d);
    const double sc = rl > (double)1e-6f && rl < sd.h_d ? ((dd * -sd.spiky_d) * dd) * inv_r : 0.0;
    const double diff = sd.h2_d - r2;
    const double wij = ((sd.poly6_d * diff) * diff) * diff;
    const double lam = lam_i + (double)pj.w;
    t.c = sc * fma(-sd.scorr_k_d, ipow_d(wij * sd.inv_wq_d, sd.scorr_n), lam);
    return t;
}

__device__ __forceinline__ void add_term(const DeltaTerm& t, double& dx, double& dy, double& dz) {
    dx = fma(t.c, t.rx, dx);
    dy = fma(t.c, t.ry, dy);
    dz = fma(t.c, t.rz, dz);
}

// The delta pass's sum R = sum_j c_j r_j in walk order, in blocks of
// kSumBlock neighbours: each block is summed from zero with fused
// multiply-adds, and the sums of the non-empty blocks are added in block
// order.  Every path that forms R (a lane per slot, a wave per slot, the
// lanes of a heavy slot taking a block each) keeps this order, so R has one
// set of bits whichever path runs; splitting a slot into blocks cuts its
// dependency chain from its neighbour count to a block.
constexpr uint32_t kSumBlock = 32;
struct DeltaSum {
    double dx = 0.0, dy = 0.0, dz = 0.0;  // sum of the finished blocks
    double bx = 0.0, by = 0.0, bz = 0.0;  // the current block
    uint32_t n = 0;                       // terms in the current block
    __device__ __forceinline__ void add(const DeltaTerm& t) {
        add_term(t, bx, by, bz);
        if (++n == kSumBlock) fold();
    }
    __device__ __forceinline__ void fold() {
#pragma clang fp contract(off)
        dx += bx;
        dy += by;
        dz += bz;
        bx = by = bz = 0.0;
        n = 0;
    }
    __device__ __forceinline__ void finish() {
        if (n) fold();
    }
};

// densityLambdaSorted for slot s: lambda_s into cur[s].w.  With kTape the
// constraint-gradient sums the adjoint needs go to aux[s] = (sum_j grad W,
// D) with D = (|sum_j grad W|^2 + sum_j |grad W|^2) / rho0^2, so that
// lambda = -C / (D + eps).  With kList the neighbours go to nl; the walk's
// chunks are staged in this lane's LDS column le (stride lst, kLdsChunks
// entries); with kXyz the stencil filter reads nl.xyz.
template <bool kOrig, bool kCount, bool kTape, bool kList, bool kXyz = false>
__device__ __forceinline__ void density_lambda_slot(uint32_t s, const SceneDev& sc, float4* __restrict__ cur,
                                                    const uint32_t* __restrict__ cell_start,
                                                    const uint32_t* __restrict__ sorted_index, uint32_t n,
                                                    float* __restrict__ lambda_orig,
                                                    uint32_t* __restrict__ nbr_orig,
                                                    float4* __restrict__ aux, const NbrList& nl,
                                                    uint2* le = nullptr, uint32_t lst = 0) {
    const SolverDev& sd = sc.sd;
    const float4 pi = cur[s];
    // sum of (h^2 - r^2)^3 in FP64; r's components are differences of
    // nearby FP32 values (exact by Sterbenz away from the origin), so the
    // density is accurate to FP64 rounding.
    double wsum = 0.0;
    float gx = 0.f, gy = 0.f, gz = 0.f, gsq = 0.f;
    uint32_t cnt = 0;
    if (kList) {
        // filter first, keeping the chunks, then the terms over them; the
        // list for the passes that follow is written on the way
        uint32_t over;
        const uint32_t K = record_chunks<kXyz>(s, pi, sc, cur, cell_start, n, nl, le, lst, cnt, over);
        // the terms use x, y, z only: with kXyz each neighbour is a 12-byte
        // load from the packed copy (1M dam break, same bits: -0.6% step time
        // at steps 81-120, within noise at 6-25)
        auto load = [&](uint32_t j) { return kXyz ? load_xyz(nl.xyz, j) : cur[j]; };
        auto term = [&](uint32_t, const float4& q) { density_term(pi, q, sd, wsum, gx, gy, gz, gsq); };
        visit_chunks<kLdsChunks>(
            [&](uint32_t k) { return k < kLdsChunks ? le[k * lst] : nl.spill[(size_t)k * nl.stride + s]; },
            min(K, kChunkScratch), s, nl, K <= kChunkScratch, load, term);
        if (K > kChunkScratch)
            neighbors_from_chunk(s, pi, sc, cur, cell_start, n, over,
                                 [&](uint32_t j) { term(j, load(j)); });
        // The list holds the neighbours when there are at most kMaxList and
        // the walk fit the kept chunks (else visit_chunks did not list).
        // More neighbours than the list holds, in kept chunks: the pass that
        // follows takes them from the chunks, all of which go to the scratch.
        // More chunks than that: the stencil.
        if (K <= kChunkScratch && cnt > kMaxList)
            for (uint32_t k = 0; k < min(K, kLdsChunks); ++k) nl.spill[(size_t)k * nl.stride + s] = le[k * lst];
        nl.count[s] = list_code(s, K, cnt, nl);
    } else {
        for_each_neighbor(s, pi, sc, cur, cell_start, n, [&](uint32_t j1, uint32_t j2, bool two) {
            cnt += two ? 2u : 1u;
            const float4 q1 = cur[j1], q2 = cur[j2];
            density_term(pi, q1, sd, wsum, gx, gy, gz, gsq);
            if (two) density_term(pi, q2, sd, wsum, gx, gy, gz, gsq);
        });
    }
    float dterm;
    const float lambda = lambda_of(wsum, gx, gy, gz, gsq, sd, dterm);
    reinterpret_cast<float*>(cur)[4 * (size_t)s + 3] = lambda;
    if (kTape) aux[s] = make_float4(gx, gy, gz, dterm);
    if (kOrig || kCount) {
        const uint32_t i = sorted_index[s];
        if (kOrig) lambda_orig[i] = lambda;
        if (kCount) nbr_orig[i] = cnt;
    }
}

// Work item s of the delta pass: its slot unless the slot is queued as
// heavy (then a heavy wave or a heavy-lane workgroup of the launch takes
// it).  Item 0 clears the counters of the next iteration's queues.
__device__ __forceinline__ bool delta_takes_slot(uint32_t s, uint32_t n, const NbrList& nl) {
    if (s == 0) {
        *nl.heavy_count_next = 0u;
        *nl.lanes_count_next = 0u;
    }
    return s < n && (nl.count[s] & kCountQueued) == 0u;
}

// sphereCollision (Constraints.metal:262-277): a point inside the sphere
// moves to the surface along the ray from the centre.
__device__ __forceinline__ float4 sphere_push(float4 p, const SphereDev& sp) {
    const float dx = p.x - sp.cx, dy = p.y - sp.cy, dz = p.z - sp.cz;
    const float r2 = dx * dx + dy * dy + dz * dz;
    if (!(r2 < sp.r * sp.r)) return p;
    const float r = sqrtf(fmaxf(r2, 1e-8f));
    return make_float4(sp.cx + dx / r * sp.r, sp.cy + dy / r * sp.r, sp.cz + dz / r * sp.r, 1.0f);
}

// The end of applyDeltaSorted for slot s from R = the sum of its terms:
// next[s] = clamp(p_s + cap(R / rho0)), then the sphere-drop scene's
// sphereCollision (Engine.cpp:760-771); its packed copy goes to xyz_next
// when there is one.  With kTape R / rho0 (before the 0.3 h cap) goes to
// rr[s] for the adjoint.
template <bool kTape>
__device__ __forceinline__ void finish_delta(uint32_t s, const float4& pi, const SolverDev& sd, double dx,
                                             double dy, double dz, float4* __restrict__ next,
                                             float4* __restrict__ rr, const SphereDev& sphere,
                                             float* __restrict__ xyz_next) {
#pragma clang fp contract(off)
    dx *= sd.inv_rho0_d;
    dy *= sd.inv_rho0_d;
    dz *= sd.inv_rho0_d;
    if (kTape) rr[s] = make_float4((float)dx, (float)dy, (float)dz, 0.0f);
    const double dlen = sqrt(fma(dz, dz, fma(dy, dy, dx * dx)));
    if (dlen > sd.max_step_d) {
        const double f = sd.max_step_d / dlen;
        dx *= f;
        dy *= f;
        dz *= f;
    }
    // clamp to the box (Constraints.metal:197), one rounding to FP32
    float4 q = make_float4((float)fmin(fmax((double)pi.x + dx, (double)sd.bminx), (double)sd.bmaxx),
                           (float)fmin(fmax((double)pi.y + dy, (double)sd.bminy), (double)sd.bmaxy),
                           (float)fmin(fmax((double)pi.z + dz, (double)sd.bminz), (double)sd.bmaxz), 1.0f);
    if (sphere.on) q = sphere_push(q, sphere);
    q.w = pi.w;  // lambda_s, which this iterate was computed with (read by finalize)
    next[s] = q;
    store_xyz(xyz_next, s, q);
}

// applyDeltaSorted for slot s (Constraints.metal:140-199): R_s = sum_j
// (lambda_s + lambda_j + s_corr) grad W, summed in walk order.  With kList
// the neighbours come from the density pass's list.
template <bool kTape, bool kList>
__device__ __forceinline__ void apply_delta_slot(uint32_t s, const SceneDev& sc,
                                                 const float4* __restrict__ cur,
                                                 const uint32_t* __restrict__ cell_start, uint32_t n,
                                                 float4* __restrict__ next, float4* __restrict__ rr,
                                                 const NbrList& nl, const SphereDev& sphere) {
    const SolverDev& sd = sc.sd;
    const float4 pi = cur[s];
    const double lam_d = pi.w;
    DeltaSum r;
    for_each_neighbor_of<kList, true>(
        s, pi, sc, cur, cell_start, n, nl, [&](uint32_t j) { return cur[j]; },
        [&](uint32_t, const float4& q) { r.add(delta_coef(pi, lam_d, q, sd)); });
    r.finish();
    finish_delta<kTape>(s, pi, sd, r.dx, r.dy, r.dz, next, rr, sphere, nl.xyz_next);
}

__device__ __forceinline__ double readlane_d(double v, int lane) {
    const uint64_t u = (uint64_t)__double_as_longlong(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, lane);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), lane);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// applyDeltaSorted for a queued (heavy) slot s by a whole wave: the lanes
// compute the terms of up to 64 neighbours at a time (consecutive in walk
// order, in lane order), and the wave adds them up one by one in walk
// order (each term read across lanes into every lane), so the bits are
// those of apply_delta_slot.  The neighbours come from the chunks the
// density pass left in nl.spill, or from the stencil.  s is wave-uniform.
template <bool kTape>
__device__ __forceinline__ void apply_delta_heavy(uint32_t s, const SceneDev& sc,
                                                  const float4* __restrict__ cur,
                                                  const uint32_t* __restrict__ cell_start, uint32_t n,
                                                  float4* __restrict__ next, float4* __restrict__ rr,
                                                  const NbrList& nl, const SphereDev& sphere) {
    const SolverDev& sd = sc.sd;
    const uint32_t lane = __lane_id();
    const float4 pi = cur[s];
    const double lam_d = pi.w;
    DeltaSum r;  // the same in every lane
    // one group of candidates, lane l's being j (a neighbour when is_nbr)
    auto group = [&](bool is_nbr, uint32_t j) {
        DeltaTerm t{0.0, 0.0, 0.0, 0.0};
        if (is_nbr) t = delta_coef(pi, lam_d, cur[j], sd);
        for (uint64_t bal = __ballot(is_nbr); bal; bal &= bal - 1ull) {
            const int l = __builtin_ctzll(bal);
            r.add(DeltaTerm{readlane_d(t.c, l), readlane_d(t.rx, l), readlane_d(t.ry, l), readlane_d(t.rz, l)});
        }
    };
    const uint32_t K = nl.count[s] & 0xffffu;
    if (K != 0xffffu) {
        // chunks k and k + 1 per group: lanes 0-31 take chunk k's bits,
        // lanes 32-63 chunk k + 1's, which is walk order
        for (uint32_t k = 0; k < K; k += 2) {
            const uint32_t kk = k + (lane >> 5);
            const uint2 e = kk < K ? nl.spill[(size_t)kk * nl.stride + s] : make_uint2(0u, 0u);
            const uint32_t bit = lane & 31u;
            group(((e.y >> bit) & 1u) != 0u, e.x + bit);
        }
    } else {
        // the stencil rows in walk order, 64 candidates per group, with the
        // neighbour test of neighbor_mask
        const Stencil st = stencil_of(pi.x, pi.y, pi.z, sc.g);
        const Rows rows = stencil_rows(st, cell_start, sc.g, sc.cell_off, n);
        for (int k = 0; k < 9; ++k) {
            for (uint32_t c0 = rows.a[k]; c0 < rows.b[k]; c0 += 64) {
                const uint32_t j = c0 + lane;
                bool is_nbr = false;
                if (j < rows.b[k] && j != s) {
                    const float4 q = cur[j];
                    is_nbr = nbr_key(pi.x - q.x, pi.y - q.y, pi.z - q.z, sd.h2) < 0.0f;
                }
                group(is_nbr, j);
            }
        }
    }
    r.finish();
    if (lane == 0) finish_delta<kTape>(s, pi, sd, r.dx, r.dy, r.dz, next, rr, sphere, nl.xyz_next);
}

// The walk of block B (neighbours kSumBlock B .. kSumBlock (B + 1) - 1 in
// walk order) of a slot whose K <= kChunkScratch chunks are in nl.spill:
// its sum from zero into (bx, by, bz); returns false when the walk has no
// such neighbours.  The chunks' set bits are counted four chunk loads at a
// time to find the block's first neighbour.
__device__ __forceinline__ bool delta_block_sum(uint32_t s, uint32_t K, uint32_t B, const float4& pi, double lam_d,
                                                const SolverDev& sd, const float4* __restrict__ cur,
                                                const NbrList& nl, double& bx, double& by, double& bz) {
    const uint32_t target = kSumBlock * B;
    const uint2* sp = nl.spill + s;
    const size_t st = nl.stride;
    uint32_t k = 0, acc = 0, m = 0, base = 0;
    bool found = false;
#pragma unroll 1
    for (uint32_t k0 = 0; k0 < K && !found; k0 += 4) {
        uint2 e[4];
#pragma unroll
        for (uint32_t u = 0; u < 4; ++u) e[u] = k0 + u < K ? sp[(size_t)(k0 + u) * st] : make_uint2(0u, 0u);
#pragma unroll
        for (uint32_t u = 0; u < 4; ++u) {
            const uint32_t pc = (uint32_t)__builtin_popcount(e[u].y);
            if (!found && k0 + u < K) {
                if (acc + pc > target) {  // the chunk holding the block's first neighbour
                    found = true;
                    m = e[u].y;
                    base = e[u].x;
                    k = k0 + u;
                } else {
                    acc += pc;
                }
            }
        }
    }
    if (!found) return false;
    for (uint32_t d = target - acc; d; --d) m &= m - 1u;  // drop the neighbours of earlier blocks
    bx = by = bz = 0.0;
    uint32_t left = kSumBlock;
#pragma unroll 1
    while (left) {
        while (m == 0u && k + 1u < K) {
            ++k;
            const uint2 e = sp[(size_t)k * st];
            base = e.x;
            m = e.y;
        }
        if (m == 0u) break;
        // up to four neighbours per trip, their loads in flight together
        uint32_t j[4];
        uint32_t got = 0;
#pragma unroll
        for (uint32_t u = 0; u < 4; ++u) {
            j[u] = s;
            if (m != 0u && got < left) {
                j[u] = base + (uint32_t)__builtin_ctz(m);
                m &= m - 1u;
                ++got;
            }
        }
        const float4 q0 = cur[j[0]], q1 = cur[j[1]], q2 = cur[j[2]], q3 = cur[j[3]];
        add_term(delta_coef(pi, lam_d, q0, sd), bx, by, bz);
        if (got > 1) add_term(delta_coef(pi, lam_d, q1, sd), bx, by, bz);
        if (got > 2) add_term(delta_coef(pi, lam_d, q2, sd), bx, by, bz);
        if (got > 3) add_term(delta_coef(pi, lam_d, q3, sd), bx, by, bz);
        left -= got;
    }
    return true;
}

// applyDeltaSorted for the slots of the heavy-lane queue (more than kMaxList
// neighbours in at most kChunkScratch chunks): kSplitLanes lanes per slot,
// lane b of a slot summing blocks b, b + kSplitLanes, ... of its walk, one
// block per round; after each round the slot's first lane adds the round's
// block sums in block order, so the sum is DeltaSum's bit for bit.  q_ok is
// false for lanes past the queue: they take part in the wave's exchanges
// and compute nothing.  Every lane of the wave must be active.
constexpr uint32_t kSplitLanes = 8;

template <bool kTape>
__device__ __forceinline__ void apply_delta_split(uint32_t s, bool q_ok, const StepArgs& a,
                                                  const float4* __restrict__ cur, float4* __restrict__ next,
                                                  float4* __restrict__ rr, const NbrList& nl) {
    const uint32_t lane = __lane_id();
    const uint32_t b = lane % kSplitLanes;
    const uint32_t K = q_ok ? nl.count[s] & 0xffffu : 0u;
    const float4 pi = q_ok ? cur[s] : make_float4(0.f, 0.f, 0.f, 0.f);
    const double lam_d = pi.w;
    double dx = 0.0, dy = 0.0, dz = 0.0;  // the sum so far, in the slot's first lane
#pragma unroll 1
    for (uint32_t round = 0;; ++round) {
        double bx = 0.0, by = 0.0, bz = 0.0;
        bool has = false;
        if (q_ok) {
            with_scene(s, a, [&](const SceneDev& sc) {
                has = delta_block_sum(s, K, b + kSplitLanes * round, pi, lam_d, sc.sd, cur, nl, bx, by, bz);
            });
        }
        if (__ballot(has) == 0ull) break;
        // the slot's first lane adds its lanes' block sums in block order
        if (has && b == 0u) {
            dx += bx;
            dy += by;
            dz += bz;
        }
#pragma unroll
        for (uint32_t o = 1; o < kSplitLanes; ++o) {
            const double ox = __shfl(bx, (int)(lane + o)), oy = __shfl(by, (int)(lane + o)),
                         oz = __shfl(bz, (int)(lane + o));
            const bool oh = __shfl((int)has, (int)(lane + o)) != 0;
            if (b == 0u && oh) {
                dx += ox;
                dy += oy;
                dz += oz;
            }
        }
    }
    if (q_ok && b == 0u) {
        with_scene(s, a, [&](const SceneDev& sc) { finish_delta<kTape>(s, pi, sc.sd, dx, dy, dz, next, rr, a.sphere, nl.xyz_next); });
    }
}

// A delta launch over n slots with lists: workgroups [0, heavy_blocks<kB>)
// come first in dispatch order and run the heavy queue, a wave per queued
// slot (grid-stride over the queue, kHeavyWaves waves); the next
// split_blocks<kB> run the heavy-lane queue, kSplitLanes lanes per slot
// (grid-stride, kSplitWaves waves); then, with an order to write, one
// workgroup per XCD range for it; the others take the slots.
constexpr uint32_t kHeavyWaves = 2048;
constexpr uint32_t kSplitWaves = 4096;

template <uint32_t kB>
constexpr uint32_t heavy_blocks() {
    return kHeavyWaves / (kB / 64u);
}

template <uint32_t kB>
constexpr uint32_t split_blocks() {
    return kSplitWaves / (kB / 64u);
}

// Runs tile_fn(pos) for the tile positions (of nb) this workgroup takes:
// the one its index wg (of nwg, one per position) maps to, or, in a
// persistent launch (kPersist, nl.pull > 0), pulls from one of the kTileQueues
// queues of its XCD (wg % kNumXcd): queue j of XCD x hands out the
// positions lo_x + j, lo_x + j + kTileQueues, ... of the XCD's range in
// turn (so each queue keeps the range's heaviest-first order), nl.pull at
// a time, from a counter of its own cache line (a single counter per XCD
// serialised ~2,000 device-scope atomics on one address).  Every thread of
// the workgroup must call it.
template <bool kPersist, class TileFn>
__device__ __forceinline__ void for_each_tile_position(const NbrList& nl, uint32_t* ctr, uint32_t wg,
                                                       uint32_t nwg, uint32_t nb, TileFn&& tile_fn) {
    if constexpr (!kPersist) {
        tile_fn(xcd_block(wg, nwg));
        return;
    }
    __shared__ uint32_t sh_k;
    const uint32_t x = wg % kNumXcd, j = (wg / kNumXcd) % kTileQueues;
    const uint32_t lo = xcd_range_start(x, nb), hi = xcd_range_start(x + 1u, nb);
    uint32_t* c = ctr + (x * kTileQueues + j) * kTileCtrStride;
    for (;;) {
        if (threadIdx.x == 0) sh_k = nl.pull * atomicAdd(c, 1u);
        __syncthreads();
        const uint32_t k0 = sh_k;
        __syncthreads();
        if (lo + j + k0 * kTileQueues >= hi) break;
        for (uint32_t u = 0; u < nl.pull; ++u) {
            const uint32_t pos = lo + j + (k0 + u) * kTileQueues;
            if (pos >= hi) break;
            tile_fn(pos);
        }
    }
}

// Times a stencil workgroup's tile (its first wave, from the start to the
// exit of thread 0) into a cost bucket, for the order of the next
// iteration's passes (NbrList::tile_order).
struct TileTimer {
    uint64_t t0;
    uint32_t* plan;
    uint32_t tile;
    __device__ TileTimer(const NbrList& nl, uint32_t tile_)
        : t0(__builtin_amdgcn_s_memrealtime()), plan(nl.tile_plan), tile(tile_) {}
    __device__ ~TileTimer() {
        if (plan && threadIdx.x == 0) {
            const uint64_t dt = __builtin_amdgcn_s_memrealtime() - t0;
            plan[tile] = (uint32_t)min(dt / kCostTicks, (uint64_t)(kCostBuckets - 1u));
        }
    }
};

// A delta launch's order workgroup for XCD range x of nb tiles (wave 0
// works): the next iteration's order of the range from the cost buckets the
// density pass left, heavier buckets first and slot order inside a bucket;
// a range whose heaviest bucket is at most kCostSpread above its median
// keeps the slot order.  Within one wave LDS operations complete in order,
// so the histogram and the running bucket positions need no barrier.
__device__ __forceinline__ void order_range(const NbrList& nl, uint32_t x, uint32_t nb) {
    __shared__ uint32_t hist[kCostBuckets], next_pos[kCostBuckets];
    if (threadIdx.x >= 64u) return;
    const uint32_t lane = __lane_id();
    const uint32_t lo = xcd_range_start(x, nb), hi = xcd_range_start(x + 1u, nb);
    if (lane < kCostBuckets) hist[lane] = 0u;
    for (uint32_t t = lo + lane; t < hi; t += 64u) atomicAdd(&hist[min(nl.tile_plan[t], kCostBuckets - 1u)], 1u);
    const uint32_t cb = lane < kCostBuckets ? hist[lane] : 0u;
    const uint32_t incl = wave_incl_scan(cb);  // tiles of the range in buckets <= lane
    const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
    const uint64_t past_half = __ballot(lane < kCostBuckets && 2u * incl >= total);
    const uint64_t nonempty = __ballot(cb != 0u);
    const uint32_t med = past_half ? (uint32_t)__builtin_ctzll(past_half) : 0u;
    const uint32_t top = nonempty ? 63u - (uint32_t)__builtin_clzll(nonempty) : 0u;
    if (top <= med + kCostSpread) {
        for (uint32_t t = lo + lane; t < hi; t += 64u) nl.tile_order_next[t] = t;
        return;
    }
    if (lane < kCostBuckets) next_pos[lane] = lo + total - incl;  // first position of each bucket
    const uint64_t below_me = (1ull << lane) - 1ull;
    for (uint32_t c = lo; c < hi; c += 64u) {
        const uint32_t t = c + lane;
        const bool in = t < hi;
        const uint32_t b = in ? min(nl.tile_plan[t], kCostBuckets - 1u) : 0u;
        uint64_t same = __ballot(in);  // lanes of this chunk in my bucket
        for (uint32_t k = 0; k < 5; ++k) {
            const uint64_t m = __ballot(in && ((b >> k) & 1u));
            same &= ((b >> k) & 1u) ? m : ~m;
        }
        if (in) {
            const uint32_t p = next_pos[b] + (uint32_t)__popcll(same & below_me);
            if (p < hi) nl.tile_order_next[p] = t;
        }
        if (in && (same >> lane) == 1ull) next_pos[b] += (uint32_t)__popcll(same);  // the group's last lane
    }
}

template <uint32_t kB, bool kTape, bool kPersist = false>
__device__ __forceinline__ void delta_launch_body(const float4* __restrict__ cur,
                                                  const uint32_t* __restrict__ cell_start, const StepArgs& a,
                                                  float4* __restrict__ next, float4* __restrict__ rr,
                                                  const NbrList& nl) {
    constexpr uint32_t hb = heavy_blocks<kB>();
    constexpr uint32_t lb = split_blocks<kB>();
    const uint32_t wave = (uint32_t)__builtin_amdgcn_readfirstlane((int)(blockIdx.x * (kB / 64u) + threadIdx.x / 64u));
    if (blockIdx.x < hb) {
        const uint32_t h_end = min(*nl.heavy_count, kHeavyCap);
        for (uint32_t h = wave; h < h_end; h += kHeavyWaves) {
            const uint32_t s = (uint32_t)__builtin_amdgcn_readfirstlane((int)nl.heavy[h]);  // h is wave-uniform
            with_scene(s, a, [&](const SceneDev& sc) {
                apply_delta_heavy<kTape>(s, sc, cur, cell_start, a.n, next, rr, nl, a.sphere);
            });
        }
        return;
    }
    if (blockIdx.x < hb + lb) {
        // 64 / kSplitLanes queued slots per wave and trip
        const uint32_t l_end = min(*nl.lanes_count, nl.lanes_cap);
        constexpr uint32_t kPerWave = 64u / kSplitLanes;
        for (uint32_t q0 = (wave - hb * (kB / 64u)) * kPerWave; q0 < l_end; q0 += kSplitWaves * kPerWave) {
            const uint32_t q = q0 + __lane_id() / kSplitLanes;
            const bool q_ok = q < l_end;
            apply_delta_split<kTape>(q_ok ? nl.lanes[q] : 0u, q_ok, a, cur, next, rr, nl);
        }
        return;
    }
    const uint32_t ob = nl.tile_order_next ? kNumXcd : 0u;
    const uint32_t nbt = (a.n + kB - 1u) / kB;
    if (blockIdx.x < hb + lb + ob) {
        order_range(nl, blockIdx.x - hb - lb, nbt);
        return;
    }
    const uint32_t wg = blockIdx.x - hb - lb - ob;
    if (kPersist && wg == 0u)  // the density pass's counters
        for (uint32_t q = threadIdx.x; q < kNumXcd * kTileQueues; q += kB) nl.tile_ctr[q * kTileCtrStride] = 0u;
    for_each_tile_position<kPersist>(nl, nl.tile_ctr + kTileCtrs, wg, gridDim.x - hb - lb - ob, nbt, [&](uint32_t pos) {
        // the tiles in the density pass's order (heaviest first), else in place
        const uint32_t tile =
            nl.tile_order ? min((uint32_t)__builtin_amdgcn_readfirstlane((int)nl.tile_order[pos]), nbt - 1u) : pos;
        const uint32_t s = tile * kB + threadIdx.x;
        if (!delta_takes_slot(s, a.n, nl)) return;
        with_scene(s, a, [&](const SceneDev& sc) {
            apply_delta_slot<kTape, true>(s, sc, cur, cell_start, a.n, next, rr, nl, a.sphere);
        });
    });
}

// finalizeStep (Constraints.metal:305-353) for one particle: pf is its
// final iterate, old its position at the start of the step; returns the new
// position and velocity, and with want_foam the foam intensity from its
// neighbour count (computeFoam, Constraints.metal:287-299).
struct Finalized {
    float4 x, v;
    float foam;
};

__device__ __forceinline__ Finalized finalize_particle(const FinalDev& fd, float dt, const float4& pf,
                                                       const float4& old, bool want_foam, uint32_t nbr) {
    const float inv_dt = 1.0f / dt;       // Engine.cpp:561
    const float vmax = fd.vmax_h * inv_dt;  // Constraints.metal:347
    const float px = fminf(fmaxf(pf.x, fd.bminx), fd.bmaxx);
    const float py = fminf(fmaxf(pf.y, fd.bminy), fd.bmaxy);
    const float pz = fminf(fmaxf(pf.z, fd.bminz), fd.bmaxz);
    float vx = (px - old.x) * inv_dt * fd.damping;
    float vy = (py - old.y) * inv_dt * fd.damping;
    float vz = (pz - old.z) * inv_dt * fd.damping;
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
    if (vlen > vmax) {
        const float f = vmax / vlen;
        vx *= f;
        vy *= f;
        vz *= f;
    }
    Finalized r{make_float4(px, py, pz, 1.0f), make_float4(vx, vy, vz, 0.0f), 0.0f};
    if (want_foam) {
        // surface-ness from the neighbour count times smoothstep(1.2, 4.5, |v|)
        const float speed = sqrtf(vx * vx + vy * vy + vz * vz);
        const float surface = 1.0f - fminf(fmaxf((float)nbr / fd.foam_max_nbr, 0.0f), 1.0f);
        const float t = fminf(fmaxf((speed - 1.2f) / (4.5f - 1.2f), 0.0f), 1.0f);
        r.foam = surface * (t * t * (3.0f - 2.0f * t));
    }
    return r;
}

// Scatter back + finalize for slot s, particle i = sorted_index[s]: x_in[i]
// is the position at the start of the step; the step's result goes to
// x_out[i], v_out[i] (x_out may be x_in), and x_in[i] to prev[i] if given.
// With foam, also the foam intensity from nbr[i] and the new velocity.
__device__ __forceinline__ void finalize_slot(uint32_t s, const FinalDev& fd, float dt,
                                              const float4* __restrict__ fin,
                                              const uint32_t* __restrict__ sorted_index,
                                              const float4* x_in, float4* x_out, float4* __restrict__ v_out,
                                              float4* __restrict__ prev, const uint32_t* __restrict__ nbr = nullptr,
                                              float* __restrict__ foam = nullptr) {
    const uint32_t i = sorted_index[s];
    const float4 old = x_in[i];
    if (prev) prev[i] = old;
    const Finalized r = finalize_particle(fd, dt, fin[s], old, foam != nullptr, foam ? nbr[i] : 0u);
    x_out[i] = r.x;
    v_out[i] = r.v;
    if (foam) foam[i] = r.foam;
}

inline uint32_t blocks_for(uint32_t n, uint32_t b) { return (n + b - 1) / b; }

}  // namespace mtlp
