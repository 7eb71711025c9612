// hash.hip -- spatial hash for gfx950: fused predict + cell hash + atomic
// count, stable scatter into cell order, gather of sorted positions, and
// the neighbour count of buildSpatialHash.  One launch per stage covers every
// scene of a batch.
//
// Reference stages (Engine.cpp:601-679, 715-723): predictPositions,
// hashCells, countCells, [scan], scatterParticles, gatherPositions,
// countNeighbors.  Differences by design:
//   * predict, hash and count are one pass over x and v; the predicted
//     position is not stored (the scatter recomputes it bit for bit);
//   * the within-cell order is stable by original index.  The count atomic
//     returns each particle's arrival rank ("ticket"); after the scan the
//     particle lands at cell_start + ticket (unordered, no second atomic
//     pass) with its key and predicted position, read in index order; then
//     every slot ranks its index among its cell's indices and moves its
//     entry to cell_start + rank, so the random accesses of the pair are
//     the scatter's stores rather than gathers.  Cells larger than
//     kSmallCell are listed and sorted by a separate kernel (LDS bitonic
//     sort; past kLdsSortMax a rank per index block, see kHugeBlocks).
#include "mtlp_stencil.h"

namespace mtlp {

namespace {

__global__ __launch_bounds__(kBlock) void k_predict_hash_count(
    const float4* __restrict__ x, const float4* __restrict__ v, StepArgs a,
    uint32_t* __restrict__ keys, uint32_t* __restrict__ tickets, uint32_t* __restrict__ counts,
    uint32_t* __restrict__ big_count, uint32_t* __restrict__ heavy_count, uint32_t* __restrict__ huge_flag,
    uint32_t* __restrict__ scan_words, uint32_t n_scan_words) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    // the scan that follows starts from cleared look-back words: every build
    // scans with the same epoch, so a replayed graph needs no clearing
    // launch and no host-side epoch
    for (uint32_t j = i; j < n_scan_words; j += gridDim.x * kBlock) scan_words[j] = 0u;
    if (i == 0) {  // the big-cell list of this build starts empty, with no huge cell
        *big_count = 0u;
        *huge_flag = 0u;
    }
    if (i < 4 && heavy_count) heavy_count[i] = 0u;  // so do the heavy-slot queues of its iterations
    if (i >= a.n) return;
    const float4 xi = x[i];
    const float4 vi = a.predict ? v[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    uint32_t key;
    with_scene(i, a, [&](const SceneDev& sc) {
        const float4 p = a.predict ? predict_pos(xi, vi, sc.gx, sc.gy, sc.gz, a.dt) : xi;
        key = sc.cell_off + key_of(cell_of(p.x, p.y, p.z, sc.g), sc.g);
    });
    keys[i] = key;
    tickets[i] = __hip_atomic_fetch_add(&counts[key], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Offsets and tickets are consistent by construction (counts start at
// zero); the range checks only turn a broken invariant into an error flag
// instead of an out-of-bounds access.  Each particle lands at cell_start +
// ticket with its key and binned position, read here in index order, so
// the ranking pass that follows reads its slot's data in slot order.
__global__ __launch_bounds__(kBlock) void k_scatter_ticket(
    const uint32_t* __restrict__ keys, const uint32_t* __restrict__ tickets,
    const uint32_t* __restrict__ cell_start, const float4* __restrict__ x,
    const float4* __restrict__ v, StepArgs a, uint2* __restrict__ tmp, float4* __restrict__ stage,
    uint32_t* __restrict__ err, uint32_t* __restrict__ hcnt) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t n = a.n;
    if (i >= n) return;
    // every load that does not need the cell range goes out with the key,
    // so the pass waits two memory latencies (key, then range) instead of
    // one more for x and v after the range
    const uint32_t key = keys[i];
    const uint32_t ticket = tickets[i];
    const float4 xi = x[i];
    const float4 vi = a.predict ? v[i] : xi;
    const CellRange r = cell_range(cell_start, key);
    const uint32_t lo = r.lo;
    const uint32_t d = lo + ticket;
    if (d < n) {
        tmp[d] = make_uint2(i, key);
        stage[d] = binned_pos_of(xi, vi, i, a);
        // a huge cell's members, counted per index block in the cell's
        // first kHugeBlocks slots (i / block length < kHugeBlocks <= len)
        const uint32_t len = r.hi - lo;
        if (len > kLdsSortMax && len <= n - lo) atomicAdd(hcnt + lo + i / huge_block_len(n), 1u);
    } else {
        atomicOr(err, kErrHashRange);
    }
}

// The wave holding the first slot L of a huge cell turns the cell's member
// counts per index block (hcnt[L .. L + kHugeBlocks)) into exclusive
// prefix sums (hoff, the rank of each block's first member), clears the
// counts for the next build and raises huge_flag.  Every lane takes part.
__device__ __forceinline__ void huge_offsets(uint32_t* __restrict__ hcnt, uint32_t* __restrict__ hoff,
                                             uint32_t L, uint32_t* __restrict__ huge_flag) {
    constexpr uint32_t kPer = kHugeBlocks / 64u;
    const uint32_t lane = __lane_id();
    uint32_t* c = hcnt + L + kPer * lane;
    uint32_t* o = hoff + L + kPer * lane;
    uint32_t sum = 0;
#pragma unroll 1
    for (uint32_t k = 0; k < kPer; ++k) sum += c[k];
    uint32_t pre = sum;  // inclusive prefix over the lanes
#pragma unroll
    for (int w = 1; w < 64; w <<= 1) {
        const uint32_t u = (uint32_t)__shfl_up((int)pre, w);
        if ((int)lane >= w) pre += u;
    }
    pre -= sum;
#pragma unroll 1
    for (uint32_t k = 0; k < kPer; ++k) {
        const uint32_t ck = c[k];
        o[k] = pre;
        c[k] = 0u;
        pre += ck;
    }
    if (lane == 0) *huge_flag = 1u;
}

__global__ __launch_bounds__(kBlock) void k_rank_gather(
    const uint2* __restrict__ tmp, const float4* __restrict__ stage,
    const uint32_t* __restrict__ cell_start, uint32_t n, uint32_t* __restrict__ sorted_index,
    float4* __restrict__ sorted, uint32_t* __restrict__ big_list, uint32_t* __restrict__ big_count,
    uint32_t* __restrict__ err, uint32_t* __restrict__ slot_of, float* __restrict__ xyz,
    uint32_t* __restrict__ hcnt, uint32_t* __restrict__ hoff, uint32_t* __restrict__ huge_flag) {
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    uint32_t i = 0, key = 0, lo = 0, hi = 0;
    bool ok = s < n;
    if (ok) {
        const uint2 e = tmp[s];
        i = e.x;
        key = e.y;
        if (i >= n) {
            atomicOr(err, kErrHashRange);
            ok = false;
        }
    }
    if (ok) {
        const CellRange r = cell_range(cell_start, key);
        lo = r.lo;
        hi = r.hi;
        if (hi > n || lo > hi) {
            atomicOr(err, kErrHashRange);
            ok = false;
        }
    }
    const bool big = ok && hi - lo > kSmallCell;
    // a huge cell starting in this wave (at most one: it spans > kLdsSortMax
    // slots): the whole wave turns its per-block counts into ranks
    const uint64_t hs = __ballot(big && s == lo && hi - lo > kLdsSortMax);
    if (hs) huge_offsets(hcnt, hoff, (uint32_t)__builtin_amdgcn_readlane((int)lo, (int)__builtin_ctzll(hs)), huge_flag);
    if (!ok) return;
    if (big) {
        if (s == lo) big_list[atomicAdd(big_count, 1u)] = key;
        return;
    }
    // the slot's staged position is read before the ranking, and the cell's
    // indices four at a time, so their loads overlap
    const float4 p = stage[s];
    uint32_t rank = 0, m = lo;
    for (; m + 4u <= hi; m += 4u) {
        const uint32_t a0 = tmp[m].x, a1 = tmp[m + 1u].x, a2 = tmp[m + 2u].x, a3 = tmp[m + 3u].x;
        rank += (a0 < i ? 1u : 0u) + (a1 < i ? 1u : 0u) + (a2 < i ? 1u : 0u) + (a3 < i ? 1u : 0u);
    }
    for (; m < hi; ++m) rank += (tmp[m].x < i) ? 1u : 0u;
    const uint32_t dst = lo + rank;
    sorted_index[dst] = i;
    sorted[dst] = p;
    store_xyz(xyz, dst, p);
    if (slot_of) slot_of[i] = dst;
}

// One workgroup per listed cell (grid-stride over the list).
__global__ __launch_bounds__(kBigBlock) void k_big_cells(
    const uint2* __restrict__ tmp, const uint32_t* __restrict__ keys,
    const uint32_t* __restrict__ cell_start, const float4* __restrict__ x,
    const float4* __restrict__ v, StepArgs a, const uint32_t* __restrict__ big_list,
    const uint32_t* __restrict__ big_count, uint32_t* __restrict__ sorted_index,
    float4* __restrict__ sorted, uint32_t* __restrict__ slot_of, float* __restrict__ xyz,
    uint32_t* __restrict__ hoff, const uint32_t* __restrict__ huge_flag) {
    __shared__ uint32_t lds[kLdsSortMax];
    const uint32_t t = threadIdx.x;
    const uint32_t lane = t & 63u, wave = t >> 6;
    const uint32_t n = a.n;
    const uint32_t count = *big_count;
    for (uint32_t bi = blockIdx.x; bi < count; bi += gridDim.x) {
        const uint32_t key = big_list[bi];
        const uint32_t lo = cell_start[key], hi = cell_start[key + 1];
        if (hi > n || lo > hi) continue;  // flagged by k_rank_gather
        const uint32_t len = hi - lo;
        if (len <= kLdsSortMax) {
            uint32_t p2 = 1;
            while (p2 < len) p2 <<= 1;
            for (uint32_t k = t; k < p2; k += kBigBlock) {
                const uint32_t idx = (k < len) ? tmp[lo + k].x : 0xffffffffu;
                lds[k] = (k < len && idx >= n) ? 0u : idx;  // out-of-range ids were flagged
            }
            __syncthreads();
            // Bitonic sort of p2 keys, ascending.
            for (uint32_t kk = 2; kk <= p2; kk <<= 1) {
                for (uint32_t j = kk >> 1; j > 0; j >>= 1) {
                    for (uint32_t q = t; q < (p2 >> 1); q += kBigBlock) {
                        const uint32_t e0 = 2 * j * (q / j) + (q % j);
                        const uint32_t e1 = e0 + j;
                        const bool up = (e0 & kk) == 0;
                        const uint32_t va = lds[e0], vb = lds[e1];
                        if ((va > vb) == up) {
                            lds[e0] = vb;
                            lds[e1] = va;
                        }
                    }
                    __syncthreads();
                }
            }
            for (uint32_t k = t; k < len; k += kBigBlock) {
                const uint32_t i = lds[k];
                sorted_index[lo + k] = i;
                const float4 p = binned_pos(x, v, i, a);
                sorted[lo + k] = p;
                store_xyz(xyz, lo + k, p);
                if (slot_of) slot_of[i] = lo + k;
            }
            __syncthreads();
        }
    }
    if (!*huge_flag) return;
    // Huge cells: wave W places the members of index block W in index order.
    // A member's rank in its cell is the rank of the block's first member
    // (hoff) plus the members of the cell before it in the block, counted
    // 64 indices at a time: a ballot per huge cell present gives each lane
    // its group's first lane and its rank in the group, then one atomic
    // from each group's first lane (all groups at once) moves that cell's
    // running rank on.
    static_assert(kBigGrid * (kBigBlock / 64) == kHugeBlocks, "one wave per index block");
    const uint32_t W = blockIdx.x * (kBigBlock / 64) + wave;
    const uint32_t R = huge_block_len(n);
    const uint32_t b0 = W * R, b1 = min(b0 + R, n);
    for (uint32_t i0 = b0; i0 < b1; i0 += 64) {
        const uint32_t i = i0 + lane;
        uint32_t key = 0, lo = 0;
        bool huge = false;
        if (i < b1) {
            key = keys[i];
            const CellRange r = cell_range(cell_start, key);
            lo = r.lo;
            const uint32_t len = r.hi - lo;
            huge = lo <= n && len > kLdsSortMax && len <= n - lo;  // hi <= n, no wrap
        }
        uint64_t todo = __ballot(huge);
        if (!todo) continue;
        int first = 0;
        uint32_t rank = 0, group = 0;
        while (todo) {
            const int leader = (int)__builtin_ctzll(todo);
            const uint32_t hk = (uint32_t)__builtin_amdgcn_readlane((int)key, leader);
            const uint64_t peers = __ballot(huge && key == hk);
            if ((peers >> lane) & 1ull) {
                first = leader;
                rank = (uint32_t)__popcll(peers & ((1ull << lane) - 1ull));
            }
            if ((int)lane == leader) group = (uint32_t)__popcll(peers);
            todo &= ~peers;
        }
        uint32_t base = 0;
        if (group) base = atomicAdd(hoff + lo + W, group);
        base = (uint32_t)__shfl((int)base, first);
        if (huge) {
            const uint32_t dst = lo + base + rank;
            if (dst < n) {
                sorted_index[dst] = i;
                const float4 p = binned_pos(x, v, i, a);
                sorted[dst] = p;
                store_xyz(xyz, dst, p);
                if (slot_of) slot_of[i] = dst;
            }
        }
    }
}

// The neighbour count of slot s: the set bits of its stencil chunks, by the
// density pass's filter (eight candidate loads in flight per group, the
// same neighbour test), so the counts equal the iteration-0 density pass's.
// Reads up to kChunkSlack slots past the last particle, like that filter.
__device__ __forceinline__ void count_neighbors_slot(uint32_t s, const SceneDev& sc,
                                                     const float4* __restrict__ sorted,
                                                     const uint32_t* __restrict__ cell_start,
                                                     const uint32_t* __restrict__ sorted_index,
                                                     uint32_t n, uint32_t* __restrict__ nbr) {
    const float4 p = sorted[s];
    uint32_t cnt = 0;
    for_each_chunk<true>(s, p, sc, sorted, cell_start, n,
                         [&](uint32_t, uint32_t mask) { cnt += (uint32_t)__builtin_popcount(mask); });
    nbr[sorted_index[s]] = cnt;
}

// countNeighbors (Spatial.metal:207-243), per sorted slot: neighbours j != s
// in the 27-cell stencil of cell(p_s) with |p_j - p_s|^2 < h^2, by the
// neighbour test of the stencil passes (nbr_key), so the counts equal the
// iteration-0 density pass's.
__global__ __launch_bounds__(kBlock) void k_count_neighbors(
    const float4* __restrict__ sorted, const uint32_t* __restrict__ cell_start,
    const uint32_t* __restrict__ sorted_index, StepArgs a, uint32_t* __restrict__ nbr) {
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t n = a.n;
    if (s >= n) return;
    with_scene(s, a, [&](const SceneDev& sc) { count_neighbors_slot(s, sc, sorted, cell_start, sorted_index, n, nbr); });
}

}  // namespace

hipError_t launch_predict_hash_count(const HashBufs& b, const StepArgs& a, hipStream_t s) {
    const uint32_t grid = blocks_for(a.n > 0 ? a.n : 1, kBlock);
    hipLaunchKernelGGL(k_predict_hash_count, dim3(grid), dim3(kBlock), 0, s, b.x, b.v, a, b.keys,
                       b.tickets, b.counts, b.big_count, b.heavy_count, b.huge_flag, b.scan_words,
                       b.n_scan_words);
    return hipGetLastError();
}

hipError_t launch_scatter_gather(const HashBufs& b, const StepArgs& a, hipStream_t s) {
    if (a.n == 0) return hipSuccess;
    const uint32_t grid = blocks_for(a.n, kBlock);
    hipLaunchKernelGGL(k_scatter_ticket, dim3(grid), dim3(kBlock), 0, s, b.keys, b.tickets,
                       b.cell_start, b.x, b.v, a, b.tmp, b.stage, b.err, b.hcnt);
    hipLaunchKernelGGL(k_rank_gather, dim3(grid), dim3(kBlock), 0, s, b.tmp, b.stage, b.cell_start, a.n,
                       b.sorted_index, b.sorted, b.big_list, b.big_count, b.err, b.slot_of, b.xyz, b.hcnt,
                       b.hoff, b.huge_flag);
    // Only cells above kSmallCell particles exist past this point; with
    // n <= kSmallCell there can be none.
    if (a.n > kSmallCell)
        hipLaunchKernelGGL(k_big_cells, dim3(kBigGrid), dim3(kBigBlock), 0, s, b.tmp, b.keys,
                           b.cell_start, b.x, b.v, a, b.big_list, b.big_count, b.sorted_index,
                           b.sorted, b.slot_of, b.xyz, b.hoff, b.huge_flag);
    return hipGetLastError();
}

hipError_t launch_count_neighbors(const float4* sorted, const uint32_t* cell_start,
                                  const uint32_t* sorted_index, const StepArgs& a, uint32_t* nbr,
                                  hipStream_t s) {
    if (a.n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_count_neighbors, dim3(blocks_for(a.n, kBlock)), dim3(kBlock), 0, s, sorted,
                       cell_start, sorted_index, a, nbr);
    return hipGetLastError();
}

}  // namespace mtlp
