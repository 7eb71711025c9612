// engine_impl.h -- internal definitions shared by the host engine
// (engine.cpp) and the rollout / adjoint entry points (adjoint.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <string>
#include <vector>

#include "mtlp_host.h"
#include "mtlp_kernels.h"

namespace mtlp {

// Records msg as the thread's last error and returns s.
mtlp_status fail(mtlp_status s, const std::string& msg);


#define HIP_TRY(expr)                                                                     \
    do {                                                                                  \
        const hipError_t e_ = (expr);                                                     \
        if (e_ != hipSuccess)                                                             \
            return fail(MTLP_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));   \
    } while (0)

#define MTLP_TRY(expr)                          \
    do {                                        \
        const mtlp_status st_ = (expr);         \
        if (st_ != MTLP_OK) return st_;         \
    } while (0)

struct DeviceGuard {
    int prev = -1;
    bool ok = false;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        ok = hipSetDevice(dev) == hipSuccess;
    }
    ~DeviceGuard() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

constexpr uint32_t kPad = 256;  // particle buffers padded like Engine.cpp:383
constexpr uint32_t kMaxIterations = 256;
constexpr uint32_t kMaxScorrN = 64;
constexpr uint32_t kMaxScenes = 1u << 20;

inline uint32_t round_up(uint32_t n, uint32_t m) { return ((n + m - 1) / m) * m; }

template <class T>
inline hipError_t dev_alloc(T** p, size_t count) {
    *p = nullptr;
    return hipMalloc(reinterpret_cast<void**>(p), sizeof(T) * (count ? count : 1));
}

template <class T>
inline void dev_free(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

mtlp_status validate_params(const mtlp_params* p, uint32_t dims[3]);

}  // namespace mtlp

// Internal header of the two host translation units: the engine struct below
// uses the kernel interface types unqualified.
using namespace mtlp;

// Device storage of the differentiable rollout: checkpoints of (x, v) every
// k steps, the tape of one segment of k steps, and the backward work
// buffers.  Sized for the engine's particle and cell counts.
struct AdjointState {
    // the last mtlp_rollout
    bool recorded = false;
    float dt = 0.0f;
    uint32_t steps = 0, every = 0;
    uint32_t n_ckpt = 0;
    float4* ck_x = nullptr;  // n_ckpt * stride
    float4* ck_v = nullptr;
    // tape of up to tape_steps steps; tape_last_seg: it holds the last
    // segment of the recorded rollout (taped during mtlp_rollout)
    uint32_t tape_steps = 0, tape_iters = 0, tape_cells = 0;
    bool tape_last_seg = false;
    size_t stride = 0;       // padded particle count of every per-particle array
    float4* xs = nullptr;    // (tape_steps + 1) * stride: state at each step start
    float4* vs = nullptr;
    float4* iter = nullptr;  // tape_steps * (iters + 1) * stride
    float4* grad = nullptr;  // tape_steps * iters * stride
    float4* dlt = nullptr;   // tape_steps * iters * stride
    uint4* lblk = nullptr;     // tape_steps * iters * kListBlocks * stride: neighbour lists
    uint32_t* lcnt = nullptr;  // tape_steps * iters * stride
    uint32_t* sidx = nullptr;  // tape_steps * stride
    uint32_t* cs = nullptr;    // tape_steps * (cells + 1)
    // backward work
    size_t work_stride = 0;
    float4 *pacc = nullptr, *prem = nullptr, *rbar = nullptr, *db = nullptr;
    float *lbar = nullptr, *lrem = nullptr;
    uint32_t* rcnt = nullptr;  // remainder inbox counters (stride)
    float4* rval = nullptr;    // remainder inbox entries (kInbox * stride)
    double *gdev = nullptr, *gpart = nullptr, *gstate = nullptr;
    float* gout = nullptr;
    float4* xb[2] = {nullptr, nullptr};
    float4* vb[2] = {nullptr, nullptr};

    void release_tape() {
        dev_free(xs); dev_free(vs); dev_free(iter); dev_free(grad); dev_free(dlt); dev_free(lblk); dev_free(lcnt);
        dev_free(sidx); dev_free(cs);
        tape_steps = tape_iters = tape_cells = 0;
        tape_last_seg = false;
    }
    void release_ckpt() {
        dev_free(ck_x); dev_free(ck_v);
        n_ckpt = 0;
        recorded = false;
    }
    void release_work() {
        dev_free(pacc); dev_free(prem); dev_free(rbar); dev_free(db); dev_free(lbar); dev_free(lrem);
        dev_free(rcnt); dev_free(rval);
        dev_free(gdev); dev_free(gpart); dev_free(gstate); dev_free(gout);
        dev_free(xb[0]); dev_free(xb[1]); dev_free(vb[0]); dev_free(vb[1]);
        work_stride = 0;
    }
    void release() {
        release_tape();
        release_ckpt();
        release_work();
        stride = 0;
    }
};

struct mtlp_engine {
    int device = 0;
    // scenes (host)
    std::vector<mtlp_params> scene_params;
    std::vector<std::array<uint32_t, 3>> scene_dims;
    std::vector<uint32_t> scene_counts;
    std::vector<uint32_t> p_off;     // n_scenes + 1
    std::vector<uint32_t> cell_off;  // n_scenes + 1
    bool batch = false;
    uint32_t cells = 0;
    uint32_t n = 0;
    uint32_t n_alloc = 0;
    uint32_t scene = MTLP_SCENE_CUBE_DROP;
    bool allocated = false;
    uint32_t stencil_block = 64;  // one wave per workgroup: fastest on every workload measured

    // scene table (device)
    SceneDev* d_scenes = nullptr;
    uint32_t* d_p_off = nullptr;
    // particle buffers (original order)
    float4* x = nullptr;
    float4* v = nullptr;
    float4* prev = nullptr;
    uint32_t* keys = nullptr;
    uint32_t* tickets = nullptr;
    uint32_t* nbr = nullptr;
    float* lambdas = nullptr;
    float* foam = nullptr;
    // particle buffers (sorted order)
    uint2* tmp_index = nullptr;  // (index, key) of the hash's unstable scatter
    uint32_t* sorted_index = nullptr;
    uint32_t* slot_of = nullptr;  // sorted slot of each particle
    float4* sorted[2] = {nullptr, nullptr};
    float* xyz[2] = {nullptr, nullptr};  // packed (x, y, z) of sorted[0], sorted[1] (NbrList::xyz)
    bool packed_positions = true;        // filter from xyz (mtlp_set_packed_positions); off without copies
    uint32_t* big_list = nullptr;
    uint32_t* hcnt = nullptr;  // huge cells: member counts per index block (zero between builds)
    uint32_t* hoff = nullptr;  // huge cells: starting ranks per index block
    int final_buf = 0;
    // neighbour list of the current iterate (kListBlocks * n_alloc) and counts
    uint4* nlist = nullptr;
    uint32_t* ncount = nullptr;
    uint2* nspill = nullptr;  // kChunkScratch * n_alloc: density-pass chunk scratch
    uint32_t* heavy = nullptr;        // 2 x kHeavyCap heavy-slot queues
    uint32_t* heavy_count = nullptr;  // their 2 counters, then the heavy-lane queues' 2
    uint32_t* lanes = nullptr;        // 2 x lanes_cap_for(n_alloc) heavy-lane queues
    uint32_t* tile_plan = nullptr;    // ceil(n_alloc / 64): density-pass tile costs (NbrList::tile_plan)
    // two tile orders (ceil(n_alloc / 64) each): an iteration's density and
    // delta passes follow tile_order[tile_order_cur] while the delta pass
    // writes the other for the next iteration
    uint32_t* tile_order[2] = {nullptr, nullptr};
    uint32_t tile_order_tiles[2] = {0, 0};  // each holds a permutation of this many tiles (0: none)
    uint32_t tile_order_cur = 0;
    bool use_lists = true;
    // persistent stencil launches: tiles per counter increment (0: one
    // workgroup per tile) and the 2 x kNumXcd position counters
    uint32_t tile_pull = 0;
    uint32_t* tile_ctr = nullptr;
    uint32_t cus = 256;  // compute units of the device
    // cell buffers
    uint32_t* counts = nullptr;
    uint32_t* cell_start = nullptr;
    uint64_t* scan_state = nullptr;
    // misc device words: [0] big-cell count, [1] scan ticket, [2] error
    // flags, [3] huge-cell flag
    uint32_t* misc = nullptr;

    // profiling
    bool profiling = false;
    std::vector<hipEvent_t> events;
    std::vector<const char*> stage_names;
    uint32_t stages_recorded = 0;

    // graph replay
    bool use_graph = false;
    bool capturing = false;
    hipStream_t capture_stream = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    float graph_dt = 0.0f;

    // rollout checkpoints, tape and adjoint work buffers (adjoint.cpp)
    AdjointState adj;

    uint32_t iterations() const { return scene_params.empty() ? 0u : scene_params[0].iterations; }

    void drop_graph() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        graph_exec = nullptr;
    }
    void free_particles() {
        dev_free(x); dev_free(v); dev_free(prev); dev_free(keys); dev_free(tickets);
        dev_free(nbr); dev_free(lambdas); dev_free(foam); dev_free(tmp_index); dev_free(sorted_index);
        dev_free(sorted[0]); dev_free(sorted[1]); dev_free(xyz[0]); dev_free(xyz[1]); dev_free(slot_of); dev_free(big_list); dev_free(hcnt); dev_free(hoff); dev_free(nlist); dev_free(ncount); dev_free(nspill); dev_free(heavy); dev_free(heavy_count); dev_free(lanes); dev_free(tile_plan); dev_free(tile_order[0]); dev_free(tile_order[1]); dev_free(tile_ctr);
        tile_order_tiles[0] = tile_order_tiles[1] = 0;
        n_alloc = 0;
    }
    void free_cells() {
        dev_free(counts); dev_free(cell_start); dev_free(scan_state);
        cells = 0;
    }
    ~mtlp_engine() {
        DeviceGuard g(device);
        drop_graph();
        if (capture_stream) (void)hipStreamDestroy(capture_stream);
        free_particles();
        free_cells();
        adj.release();
        dev_free(misc);
        dev_free(d_scenes);
        dev_free(d_p_off);
        for (hipEvent_t ev : events) (void)hipEventDestroy(ev);
    }

    HashBufs hash_bufs() {
        HashBufs b{};
        b.x = x; b.v = v; b.keys = keys; b.tickets = tickets; b.counts = counts;
        b.cell_start = cell_start; b.tmp = tmp_index; b.sorted_index = sorted_index;
        b.stage = sorted[1];  // the second iterate buffer is free until the first delta pass
        b.sorted = sorted[0]; b.big_list = big_list; b.big_count = misc + 0; b.err = misc + 2;
        b.slot_of = slot_of;
        b.heavy_count = heavy_count;
        b.hcnt = hcnt; b.hoff = hoff; b.huge_flag = misc + 3;
        b.xyz = use_lists && packed_positions ? xyz[0] : nullptr;  // iteration 0's copy (may be NULL)
        return b;
    }

    uint32_t stencil_tiles() const { return (n + stencil_block - 1) / stencil_block; }

    // lists of solver iteration `it` (its heavy queue by parity); the
    // passes follow tile_order[tile_order_cur] when it is a permutation of
    // this launch's tiles
    // Iteration `it` reads iterate sorted[it & 1] and writes the other one;
    // with packed positions on, its density pass filters from that
    // iterate's packed copy and its delta pass writes the next one.
    NbrList nbr_list(uint32_t it) const {
        if (!use_lists) return NbrList{};
        const uint32_t c = tile_order_cur;
        const uint32_t lcap = lanes_cap_for(n_alloc);
        NbrList nl{nlist, ncount, n_alloc, nspill, heavy + (it & 1u) * kHeavyCap, heavy_count + (it & 1u),
                   heavy_count + ((it + 1u) & 1u), lanes + (size_t)(it & 1u) * lcap,
                   heavy_count + 2u + (it & 1u), heavy_count + 2u + ((it + 1u) & 1u), lcap, tile_plan,
                   tile_order_tiles[c] == stencil_tiles() ? tile_order[c] : nullptr, tile_order[c ^ 1u],
                   tile_ctr, tile_pull, cus * kPersistentWavesPerCu};
        if (packed_positions && xyz[0]) {
            nl.xyz = xyz[it & 1u];
            nl.xyz_next = xyz[(it + 1u) & 1u];
        }
        return nl;
    }
    // after enqueueing an iteration's passes with lists: its delta pass
    // wrote the other order
    void flip_tile_order() {
        tile_order_tiles[tile_order_cur ^ 1u] = stencil_tiles();
        tile_order_cur ^= 1u;
    }

    // Kinematic sphere of the sphere-drop scene (Engine.hpp:180-184),
    // integrated on the host each step like Engine.cpp:693-712.
    struct Sphere {
        bool on = false;
        float c[3] = {0.0f, 4.5f, 0.0f};
        float v[3] = {0.0f, 0.0f, 0.0f};
        float r = 0.6f;
    } sphere;

    StepArgs step_args(float dt, bool predict) const {
        StepArgs a{};
        a.scenes = d_scenes;
        a.p_off = d_p_off;
        a.n_scenes = (uint32_t)scene_params.size();
        a.n = n;
        a.dt = dt;
        a.predict = predict ? 1u : 0u;
        a.sphere = SphereDev{sphere.c[0], sphere.c[1], sphere.c[2], sphere.r, sphere.on ? 1u : 0u};
        return a;
    }
};

namespace mtlp {

// Records a HIP event after each kernel when profiling is on.
struct StageClock {
    mtlp_engine* e;
    hipStream_t s;
    bool on;
    uint32_t k = 0;
    StageClock(mtlp_engine* eng, hipStream_t stream)
        : e(eng), s(stream), on(eng->profiling && !eng->capturing) {
        if (!on) return;
        e->stage_names.clear();
        e->stages_recorded = 0;
        (void)hipEventRecord(e->events[0], s);
    }
    void mark(const char* name) {
        if (!on || k + 1 >= e->events.size()) return;
        ++k;
        (void)hipEventRecord(e->events[k], s);
        e->stage_names.push_back(name);
        e->stages_recorded = k;
    }
};

// Spatial hash of b.x (predicted when a.predict) into b's cell_start,
// sorted_index and sorted: hash + count, scan, stable scatter + gather.
mtlp_status enqueue_hash(mtlp_engine* e, hipStream_t s, const StepArgs& a, StageClock& clk, const HashBufs& b);
mtlp_status check_engine(const mtlp_engine* e);

}  // namespace mtlp
