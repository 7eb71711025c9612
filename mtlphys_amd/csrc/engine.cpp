// engine.cpp -- C++20 host engine on the HIP runtime and its C ABI
// (include/mtlp.h).
//
// Restates the orchestration of mtlphys::Engine (engine/src/Engine.cpp):
// buffer allocation (:381-452), parameter derivation (:551-572), the spatial
// hash (:601-679), the step (:685-811) and the standalone scan (:843-882).
// Every stage is enqueued on the caller's stream; nothing inside a step
// waits for the device.  An engine holds one or more independent scenes laid
// end to end (a batch); every stage is one launch for all of them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "engine_impl.h"

using namespace mtlp;

namespace mtlp {

namespace {
thread_local std::string g_last_error;
}  // namespace

mtlp_status fail(mtlp_status s, const std::string& msg) {
    g_last_error = msg;
    return s;
}

mtlp_status validate_params(const mtlp_params* p, uint32_t dims[3]) {
    if (!p) return fail(MTLP_E_INVALID, "params is NULL");
    if (!(p->h > 0.0f) || !std::isfinite(p->h)) return fail(MTLP_E_INVALID, "h must be > 0");
    if (!(p->cell_size > 0.0f) || !std::isfinite(p->cell_size))
        return fail(MTLP_E_INVALID, "cell_size must be > 0");
    if (!(p->rest_density > 0.0f)) return fail(MTLP_E_INVALID, "rest_density must be > 0");
    if (p->iterations > kMaxIterations) return fail(MTLP_E_RANGE, "iterations > 256");
    if (p->scorr_n > kMaxScorrN) return fail(MTLP_E_RANGE, "scorr_n > 64");
    uint64_t total = 1;
    for (int a = 0; a < 3; ++a) {
        if (!std::isfinite(p->bounds_min[a]) || !std::isfinite(p->bounds_max[a]) ||
            !(p->bounds_max[a] > p->bounds_min[a]))
            return fail(MTLP_E_INVALID, "bounds_max must exceed bounds_min on every axis");
        // Engine.cpp:399-402: FP32 extent / cell size, rounded up.
        const float extent = p->bounds_max[a] - p->bounds_min[a];
        const float cells = std::ceil(extent / p->cell_size);
        if (!(cells >= 1.0f) || cells > 16777216.0f)
            return fail(MTLP_E_RANGE, "grid axis outside [1, 2^24] cells");
        dims[a] = (uint32_t)cells;
        total *= dims[a];
    }
    // scan length is total + 1 (Engine.cpp:405)
    if (total + 1 > (uint64_t)MTLP_MAX_SCAN_LENGTH)
        return fail(MTLP_E_RANGE, "grid has more cells than the scan supports");
    return MTLP_OK;
}

GridDev make_grid(const mtlp_params& p, const uint32_t dims[3]) {
    GridDev g{};
    g.ox = p.bounds_min[0];
    g.oy = p.bounds_min[1];
    g.oz = p.bounds_min[2];
    g.inv = 1.0f / p.cell_size;  // Engine.cpp:446
    g.dx = dims[0];
    g.dy = dims[1];
    g.dz = dims[2];
    g.dxy = dims[0] * dims[1];
    return g;
}

// writePBFParams (Engine.cpp:551-572) plus the kernel-local constants of
// applyDeltaSorted (Constraints.metal:158-160, 190): FP32 like the
// reference, and FP64 for the terms the kernels evaluate in FP64.
SolverDev make_solver(const mtlp_params& p) {
    const float kPi = 3.14159265358979323846f;
    const float h = p.h, h2 = h * h, h3 = h2 * h, h6 = h3 * h3;
    SolverDev s{};
    s.h = h;
    s.h2 = h2;
    s.spiky = 45.0f / (kPi * h6);
    s.inv_rho0 = 1.0f / p.rest_density;
    s.eps = p.epsilon;
    s.scorr_n = p.scorr_n;
    const double pi = 3.14159265358979323846;
    const double hd = h, h2d = hd * hd, h3d = h2d * hd, h6d = h3d * h3d, h9d = h6d * h3d;
    s.h2_d = h2d;
    s.poly6_d = 315.0 / (64.0 * pi * h9d);
    s.w0_d = s.poly6_d * h2d * h2d * h2d;
    s.inv_rho0_d = 1.0 / (double)p.rest_density;
    s.h_d = hd;
    s.spiky_d = 45.0 / (pi * h6d);
    const double dqh = (double)p.scorr_dq * hd, dd = h2d - dqh * dqh;
    s.inv_wq_d = 1.0 / (s.poly6_d * dd * dd * dd);
    s.scorr_k_d = p.scorr_k;
    s.max_step_d = (double)p.max_step * hd;
    s.bminx = p.bounds_min[0];
    s.bminy = p.bounds_min[1];
    s.bminz = p.bounds_min[2];
    s.bmaxx = p.bounds_max[0];
    s.bmaxy = p.bounds_max[1];
    s.bmaxz = p.bounds_max[2];
    return s;
}

FinalDev make_final(const mtlp_params& p) {
    FinalDev f{};
    f.bminx = p.bounds_min[0];
    f.bminy = p.bounds_min[1];
    f.bminz = p.bounds_min[2];
    f.bmaxx = p.bounds_max[0];
    f.bmaxy = p.bounds_max[1];
    f.bmaxz = p.bounds_max[2];
    f.damping = p.damping;
    f.weps = p.wall_eps;
    f.floor_f = p.floor_friction;
    f.wall_f = p.wall_friction;
    f.vmax_h = p.vmax * p.h;  // Constraints.metal:347: 1.0f * h * invDt
    f.foam_max_nbr = 96.0f;   // kMaxNeighborsForViz, Engine.cpp:27
    return f;
}

}  // namespace mtlp

namespace {

// Scene layout from per-scene params and counts: grids, cell and particle
// offsets.  Validates that every scene shares the iteration count.
mtlp_status layout_scenes(mtlp_engine* e) {
    const size_t ns = e->scene_params.size();
    e->scene_dims.assign(ns, {0, 0, 0});
    e->p_off.assign(ns + 1, 0);
    e->cell_off.assign(ns + 1, 0);
    uint64_t cells = 0, parts = 0;
    for (size_t s = 0; s < ns; ++s) {
        uint32_t d[3];
        MTLP_TRY(validate_params(&e->scene_params[s], d));
        if (e->scene_params[s].iterations != e->scene_params[0].iterations)
            return fail(MTLP_E_INVALID, "every scene of a batch must use the same iteration count");
        e->scene_dims[s] = {d[0], d[1], d[2]};
        e->cell_off[s] = (uint32_t)cells;
        e->p_off[s] = (uint32_t)parts;
        cells += (uint64_t)d[0] * d[1] * d[2];
        parts += e->scene_counts[s];
        if (cells + 1 > (uint64_t)MTLP_MAX_SCAN_LENGTH)
            return fail(MTLP_E_RANGE, "total cell count above what the scan supports");
        if (parts > (uint64_t)MTLP_MAX_PARTICLES)
            return fail(MTLP_E_RANGE, "particle count above MTLP_MAX_PARTICLES");
    }
    e->cell_off[ns] = (uint32_t)cells;
    e->p_off[ns] = (uint32_t)parts;
    return MTLP_OK;
}

mtlp_status upload_scene_table(mtlp_engine* e) {
    const size_t ns = e->scene_params.size();
    std::vector<SceneDev> table(ns);
    for (size_t s = 0; s < ns; ++s) {
        const mtlp_params& p = e->scene_params[s];
        SceneDev& d = table[s];
        d.sd = make_solver(p);
        d.g = make_grid(p, e->scene_dims[s].data());
        d.fd = make_final(p);
        d.gx = p.gravity[0];
        d.gy = p.gravity[1];
        d.gz = p.gravity[2];
        d.cell_off = e->cell_off[s];
    }
    dev_free(e->d_scenes);
    dev_free(e->d_p_off);
    HIP_TRY(dev_alloc(&e->d_scenes, ns));
    HIP_TRY(dev_alloc(&e->d_p_off, ns + 1));
    HIP_TRY(hipMemcpy(e->d_scenes, table.data(), sizeof(SceneDev) * ns, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(e->d_p_off, e->p_off.data(), sizeof(uint32_t) * (ns + 1), hipMemcpyHostToDevice));
    e->drop_graph();
    return MTLP_OK;
}

mtlp_status alloc_cells(mtlp_engine* e, uint32_t C) {
    e->free_cells();
    const size_t tiles = scan_tiles(C + 1);
    HIP_TRY(dev_alloc(&e->counts, (size_t)C + 1));
    HIP_TRY(dev_alloc(&e->cell_start, (size_t)C + 1));
    HIP_TRY(dev_alloc(&e->scan_state, tiles));
    HIP_TRY(hipMemset(e->counts, 0, sizeof(uint32_t) * ((size_t)C + 1)));
    HIP_TRY(hipMemset(e->cell_start, 0, sizeof(uint32_t) * ((size_t)C + 1)));
    HIP_TRY(hipMemset(e->scan_state, 0, sizeof(uint64_t) * tiles));
    e->cells = C;
    return MTLP_OK;
}

mtlp_status alloc_particles(mtlp_engine* e, uint32_t n) {
    e->free_particles();
    const uint32_t na = round_up(n > 0 ? n : 1, kPad);
    HIP_TRY(dev_alloc(&e->x, na));
    HIP_TRY(dev_alloc(&e->v, na));
    HIP_TRY(dev_alloc(&e->prev, na));
    HIP_TRY(dev_alloc(&e->keys, na));
    HIP_TRY(dev_alloc(&e->tickets, na));
    HIP_TRY(dev_alloc(&e->nbr, na));
    HIP_TRY(dev_alloc(&e->lambdas, na));
    HIP_TRY(dev_alloc(&e->foam, na));
    HIP_TRY(dev_alloc(&e->tmp_index, na));
    HIP_TRY(dev_alloc(&e->sorted_index, na));
    HIP_TRY(dev_alloc(&e->slot_of, na));
    HIP_TRY(dev_alloc(&e->sorted[0], na + kChunkSlack));  // the stencil filter reads past the end
    HIP_TRY(dev_alloc(&e->sorted[1], na + kChunkSlack));
    if ((size_t)na + kChunkSlack <= kMaxXyzSlots) {
        HIP_TRY(dev_alloc(&e->xyz[0], 3 * ((size_t)na + kChunkSlack)));
        HIP_TRY(dev_alloc(&e->xyz[1], 3 * ((size_t)na + kChunkSlack)));
        HIP_TRY(hipMemset(e->xyz[0], 0, 3 * sizeof(float) * (na + kChunkSlack)));
        HIP_TRY(hipMemset(e->xyz[1], 0, 3 * sizeof(float) * (na + kChunkSlack)));
    }
    HIP_TRY(dev_alloc(&e->big_list, na / kSmallCell + 1));
    HIP_TRY(dev_alloc(&e->hcnt, na));
    HIP_TRY(dev_alloc(&e->hoff, na));
    HIP_TRY(hipMemset(e->hcnt, 0, sizeof(uint32_t) * na));
    HIP_TRY(dev_alloc(&e->nlist, (size_t)kListBlocks * na));
    HIP_TRY(dev_alloc(&e->ncount, na));
    HIP_TRY(dev_alloc(&e->nspill, (size_t)kChunkScratch * na));
    HIP_TRY(dev_alloc(&e->heavy, 2 * (size_t)kHeavyCap));
    HIP_TRY(dev_alloc(&e->heavy_count, 4));
    HIP_TRY(hipMemset(e->heavy_count, 0, 4 * sizeof(uint32_t)));
    HIP_TRY(dev_alloc(&e->lanes, 2 * (size_t)lanes_cap_for(na)));
    HIP_TRY(dev_alloc(&e->tile_plan, (na + 63) / 64));
    HIP_TRY(dev_alloc(&e->tile_ctr, 2 * (size_t)kTileCtrs));
    HIP_TRY(hipMemset(e->tile_ctr, 0, 2 * kTileCtrs * sizeof(uint32_t)));
    HIP_TRY(dev_alloc(&e->tile_order[0], (na + 63) / 64));
    HIP_TRY(dev_alloc(&e->tile_order[1], (na + 63) / 64));
    e->tile_order_tiles[0] = e->tile_order_tiles[1] = 0;
    {
        // both orders start as the identity, so a buffer marked valid always
        // holds a permutation of the tiles
        std::vector<uint32_t> iota((na + 63) / 64);
        for (size_t t = 0; t < iota.size(); ++t) iota[t] = (uint32_t)t;
        HIP_TRY(hipMemcpy(e->tile_order[0], iota.data(), sizeof(uint32_t) * iota.size(), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(e->tile_order[1], iota.data(), sizeof(uint32_t) * iota.size(), hipMemcpyHostToDevice));
    }
    HIP_TRY(hipMemset(e->x, 0, sizeof(float4) * na));
    HIP_TRY(hipMemset(e->v, 0, sizeof(float4) * na));
    HIP_TRY(hipMemset(e->prev, 0, sizeof(float4) * na));
    HIP_TRY(hipMemset(e->keys, 0, sizeof(uint32_t) * na));
    HIP_TRY(hipMemset(e->nbr, 0, sizeof(uint32_t) * na));
    HIP_TRY(hipMemset(e->lambdas, 0, sizeof(float) * na));
    HIP_TRY(hipMemset(e->foam, 0, sizeof(float) * na));
    HIP_TRY(hipMemset(e->sorted_index, 0, sizeof(uint32_t) * na));
    HIP_TRY(hipMemset(e->sorted[0], 0, sizeof(float4) * (na + kChunkSlack)));
    HIP_TRY(hipMemset(e->sorted[1], 0, sizeof(float4) * (na + kChunkSlack)));
    e->n_alloc = na;
    e->n = n;
    return MTLP_OK;
}

// (Re)allocates every buffer for the current scene list.
mtlp_status allocate_all(mtlp_engine* e) {
    DeviceGuard g(e->device);
    HIP_TRY(hipDeviceSynchronize());
    e->allocated = false;
    e->adj.release();
    MTLP_TRY(layout_scenes(e));
    MTLP_TRY(alloc_particles(e, e->p_off.back()));
    MTLP_TRY(alloc_cells(e, e->cell_off.back()));
    MTLP_TRY(upload_scene_table(e));
    e->final_buf = 0;
    e->allocated = true;
    return MTLP_OK;
}

// Seeds every scene with the generator of `scene`; synchronous.
mtlp_status seed_scenes(mtlp_engine* e, uint32_t scene, uint64_t seed) {
    const size_t ns = e->scene_params.size();
    std::vector<float> hx(4 * (size_t)e->n), hv(4 * (size_t)e->n);
    for (size_t s = 0; s < ns; ++s) {
        const size_t off = 4 * (size_t)e->p_off[s];
        const mtlp_status st = generate_scene(scene, e->scene_counts[s], seed + (ns > 1 ? s : 0),
                                              e->scene_params[s], hx.data() + off, hv.data() + off);
        if (st != MTLP_OK) return fail(st, "unknown scene id");
    }
    DeviceGuard g(e->device);
    if (e->n) {
        HIP_TRY(hipMemcpy(e->x, hx.data(), sizeof(float) * 4 * (size_t)e->n, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(e->v, hv.data(), sizeof(float) * 4 * (size_t)e->n, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(e->prev, hx.data(), sizeof(float) * 4 * (size_t)e->n, hipMemcpyHostToDevice));
    }
    e->scene = scene;
    return MTLP_OK;
}

}  // namespace

namespace mtlp {

// Engine.cpp:601-667 and :715-726.
mtlp_status enqueue_hash(mtlp_engine* e, hipStream_t s, const StepArgs& a, StageClock& clk, const HashBufs& b) {
    // the predict pass clears the scan's look-back words, which the scan
    // then uses with epoch 1 (the same in eager launches and graph replays)
    HashBufs hb = b;
    hb.scan_words = reinterpret_cast<uint32_t*>(e->scan_state);
    hb.n_scan_words = 2 * scan_tiles(e->cells + 1);
    HIP_TRY(launch_predict_hash_count(hb, a, s));
    clk.mark("predict_hash_count");
    HIP_TRY(launch_exclusive_scan(b.counts, b.cell_start, e->cells + 1, true, e->scan_state,
                                  e->misc + 1, 1u, e->misc + 2, s));
    clk.mark("exclusive_scan");
    HIP_TRY(launch_scatter_gather(b, a, s));
    clk.mark("stable_scatter_gather");
    return MTLP_OK;
}

mtlp_status check_engine(const mtlp_engine* e) {
    if (!e) return fail(MTLP_E_INVALID, "engine is NULL");
    if (!e->allocated) return fail(MTLP_E_INVALID, "engine has no buffers: call mtlp_reset first");
    return MTLP_OK;
}

}  // namespace mtlp

namespace {

// Engine::step (Engine.cpp:685-811) for every scene.
mtlp_status enqueue_step(mtlp_engine* e, hipStream_t s, float dt) {
    StageClock clk(e, s);
    const StepArgs a = e->step_args(dt, true);
    MTLP_TRY(enqueue_hash(e, s, a, clk, e->hash_bufs()));
    const uint32_t iters = e->iterations();
    int cur = 0;
    for (uint32_t it = 0; it < iters; ++it) {  // Engine.cpp:732-774
        const NbrList nl = e->nbr_list(it);
        HIP_TRY(launch_density_lambda(e->sorted[cur], e->cell_start, e->sorted_index, a, nullptr,
                                      it == 0 ? e->nbr : nullptr, nl, e->stencil_block, s));
        clk.mark("density_lambda");
        HIP_TRY(launch_apply_delta(e->sorted[cur], e->cell_start, a, e->sorted[cur ^ 1], nl, e->stencil_block, s));
        clk.mark("apply_delta");
        if (nl.blk) e->flip_tile_order();
        cur ^= 1;
    }
    if (iters == 0) {
        HIP_TRY(launch_count_neighbors(e->sorted[0], e->cell_start, e->sorted_index, a, e->nbr, s));
        clk.mark("count_neighbors");
    }
    // the last delta pass left each slot's lambda in .w of the final iterate
    HIP_TRY(launch_finalize(e->sorted[cur], e->slot_of, a, e->x, e->v, e->prev, e->nbr, e->foam, iters > 0,
                            e->lambdas, s));
    clk.mark("scatter_finalize");
    e->final_buf = cur;
    return MTLP_OK;
}

// Sphere kinematics of the sphere-drop scene (Engine.cpp:693-712): gravity,
// 0.998 drag per step, walls with restitution 0.4 (the scene's bounds).
void advance_sphere(mtlp_engine::Sphere& sp, const mtlp_params& p, float dt) {
    sp.v[1] += -9.81f * dt;
    for (int a = 0; a < 3; ++a) sp.v[a] *= 0.998f;
    for (int a = 0; a < 3; ++a) sp.c[a] += sp.v[a] * dt;
    for (int a = 0; a < 3; ++a) {
        if (sp.c[a] - sp.r < p.bounds_min[a]) {
            sp.c[a] = p.bounds_min[a] + sp.r;
            sp.v[a] = -sp.v[a] * 0.4f;
        }
        if (sp.c[a] + sp.r > p.bounds_max[a]) {
            sp.c[a] = p.bounds_max[a] - sp.r;
            sp.v[a] = -sp.v[a] * 0.4f;
        }
    }
}

mtlp_status capture_graph(mtlp_engine* e, float dt) {
    e->drop_graph();
    if (!e->capture_stream) HIP_TRY(hipStreamCreateWithFlags(&e->capture_stream, hipStreamNonBlocking));
    HIP_TRY(hipStreamBeginCapture(e->capture_stream, hipStreamCaptureModeThreadLocal));
    // Capturing records the kernels without running them, so the tile
    // orders it marks as written are not: the host's tile-order state is
    // put back afterwards, whether the capture succeeds or fails (the graph
    // keeps the buffers it was captured with; every buffer it writes holds
    // a permutation of the tiles).
    const uint32_t order_cur = e->tile_order_cur;
    const uint32_t order_tiles[2] = {e->tile_order_tiles[0], e->tile_order_tiles[1]};
    e->capturing = true;
    const mtlp_status st = enqueue_step(e, e->capture_stream, dt);
    e->capturing = false;
    e->tile_order_cur = order_cur;
    e->tile_order_tiles[0] = order_tiles[0];
    e->tile_order_tiles[1] = order_tiles[1];
    hipGraph_t graph = nullptr;
    const hipError_t end = hipStreamEndCapture(e->capture_stream, &graph);
    if (st != MTLP_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return st;
    }
    HIP_TRY(end);
    const hipError_t inst = hipGraphInstantiate(&e->graph_exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIP_TRY(inst);
    e->graph_dt = dt;
    return MTLP_OK;
}

mtlp_status new_engine(int hip_device, mtlp_engine** out) {
    int count = 0;
    HIP_TRY(hipGetDeviceCount(&count));
    if (hip_device < 0 || hip_device >= count) return fail(MTLP_E_INVALID, "no such HIP device");
    mtlp_engine* e = new (std::nothrow) mtlp_engine();
    if (!e) return fail(MTLP_E_NOMEM, "host allocation failed");
    e->device = hip_device;
    DeviceGuard g(hip_device);
    int cus = 0;
    if (g.ok && hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, hip_device) == hipSuccess &&
        cus > 0)
        e->cus = (uint32_t)cus;
    if (!g.ok || hipMalloc(reinterpret_cast<void**>(&e->misc), 4 * sizeof(uint32_t)) != hipSuccess ||
        hipMemset(e->misc, 0, 4 * sizeof(uint32_t)) != hipSuccess) {
        delete e;
        return fail(MTLP_E_NOMEM, "device allocation failed");
    }
    *out = e;
    return MTLP_OK;
}

}  // namespace

extern "C" {

const char* mtlp_status_string(mtlp_status s) {
    switch (s) {
        case MTLP_OK: return "MTLP_OK";
        case MTLP_E_INVALID: return "MTLP_E_INVALID";
        case MTLP_E_RANGE: return "MTLP_E_RANGE";
        case MTLP_E_NOMEM: return "MTLP_E_NOMEM";
        case MTLP_E_HIP: return "MTLP_E_HIP";
        case MTLP_E_DEVICE: return "MTLP_E_DEVICE";
    }
    return "unknown status";
}

const char* mtlp_last_error(void) { return g_last_error.c_str(); }

mtlp_status mtlp_default_params(mtlp_params* out) {
    if (!out) return fail(MTLP_E_INVALID, "out is NULL");
    default_params(out);
    return MTLP_OK;
}

mtlp_status mtlp_scene_params(uint32_t scene, uint32_t n, mtlp_params* out) {
    if (!out) return fail(MTLP_E_INVALID, "out is NULL");
    const mtlp_status st = scene_params(scene, n, out);
    if (st != MTLP_OK) return fail(st, "unknown scene id");
    return MTLP_OK;
}

mtlp_status mtlp_generate_scene(uint32_t scene, uint32_t n, uint64_t seed, const mtlp_params* p,
                                float* x4, float* v4) {
    if (!p || (n && (!x4 || !v4))) return fail(MTLP_E_INVALID, "NULL argument");
    const mtlp_status st = generate_scene(scene, n, seed, *p, x4, v4);
    if (st != MTLP_OK) return fail(st, "unknown scene id");
    return MTLP_OK;
}

mtlp_status mtlp_grid_dims(const mtlp_params* p, uint32_t dims[3]) {
    if (!dims) return fail(MTLP_E_INVALID, "dims is NULL");
    return validate_params(p, dims);
}

mtlp_status mtlp_create(int hip_device, const mtlp_params* p, mtlp_engine** out) {
    if (!out) return fail(MTLP_E_INVALID, "out is NULL");
    *out = nullptr;
    mtlp_params params;
    if (p) params = *p; else default_params(&params);
    uint32_t dims[3];
    MTLP_TRY(validate_params(&params, dims));
    mtlp_engine* e = nullptr;
    MTLP_TRY(new_engine(hip_device, &e));
    e->scene_params = {params};
    e->scene_counts = {0};
    *out = e;
    return MTLP_OK;
}

mtlp_status mtlp_batch_create(int hip_device, uint32_t n_scenes, const uint32_t* counts,
                              const mtlp_params* per_scene, mtlp_engine** out) {
    if (!out) return fail(MTLP_E_INVALID, "out is NULL");
    *out = nullptr;
    if (n_scenes == 0 || !counts || !per_scene) return fail(MTLP_E_INVALID, "empty batch");
    if (n_scenes > kMaxScenes) return fail(MTLP_E_RANGE, "too many scenes");
    mtlp_engine* e = nullptr;
    MTLP_TRY(new_engine(hip_device, &e));
    e->batch = true;
    e->scene_params.assign(per_scene, per_scene + n_scenes);
    e->scene_counts.assign(counts, counts + n_scenes);
    const mtlp_status st = allocate_all(e);
    if (st != MTLP_OK) {
        const std::string msg = g_last_error;
        delete e;
        return fail(st, msg);
    }
    e->scene = MTLP_SCENE_EMPTY;
    *out = e;
    return MTLP_OK;
}

mtlp_status mtlp_destroy(mtlp_engine* e) {
    if (!e) return fail(MTLP_E_INVALID, "engine is NULL");
    {
        DeviceGuard g(e->device);
        (void)hipDeviceSynchronize();
    }
    delete e;
    return MTLP_OK;
}

mtlp_status mtlp_reset(mtlp_engine* e, uint32_t n, uint32_t scene, uint64_t seed) {
    if (!e) return fail(MTLP_E_INVALID, "engine is NULL");
    if (n > MTLP_MAX_PARTICLES) return fail(MTLP_E_RANGE, "particle count above MTLP_MAX_PARTICLES");
    mtlp_params probe;
    if (scene_params(scene, 0, &probe) != MTLP_OK) return fail(MTLP_E_INVALID, "unknown scene id");
    if (e->batch) {
        if (n != e->n) return fail(MTLP_E_RANGE, "a batch keeps its particle count; use mtlp_batch_seed");
    } else {
        e->scene_counts = {n};
        MTLP_TRY(allocate_all(e));
    }
    // the sphere-drop scene's sphere starts at rest above the lake (Engine.cpp:522-527)
    e->sphere = mtlp_engine::Sphere{};
    e->sphere.on = !e->batch && scene == MTLP_SCENE_SPHERE_DROP;
    return seed_scenes(e, scene, seed);
}

mtlp_status mtlp_batch_seed(mtlp_engine* e, uint32_t scene, uint64_t seed) {
    MTLP_TRY(check_engine(e));
    return seed_scenes(e, scene, seed);
}

mtlp_status mtlp_batch_info(const mtlp_engine* e, uint32_t* n_scenes, uint32_t* particle_offsets,
                            uint32_t* cell_offsets) {
    if (!e) return fail(MTLP_E_INVALID, "engine is NULL");
    if (n_scenes) *n_scenes = (uint32_t)e->scene_params.size();
    if (particle_offsets && !e->p_off.empty())
        std::memcpy(particle_offsets, e->p_off.data(), sizeof(uint32_t) * e->p_off.size());
    if (cell_offsets && !e->cell_off.empty())
        std::memcpy(cell_offsets, e->cell_off.data(), sizeof(uint32_t) * e->cell_off.size());
    return MTLP_OK;
}

mtlp_status mtlp_batch_set_params(mtlp_engine* e, uint32_t scene, const mtlp_params* p) {
    if (!e) return fail(MTLP_E_INVALID, "engine is NULL");
    if (scene >= e->scene_params.size()) return fail(MTLP_E_RANGE, "no such scene");
    uint32_t dims[3];
    MTLP_TRY(validate_params(p, dims));
    const mtlp_params old = e->scene_params[scene];
    e->scene_params[scene] = *p;
    e->adj.recorded = false;  // checkpoints were taken under the old parameters
    if (!e->allocated) return MTLP_OK;
    const auto& d = e->scene_dims[scene];
    if (dims[0] != d[0] || dims[1] != d[1] || dims[2] != d[2]) {
        // new grid: re-lay the cells (particle buffers keep their contents)
        DeviceGuard g(e->device);
        HIP_TRY(hipDeviceSynchronize());
        mtlp_status st = layout_scenes(e);
        if (st == MTLP_OK) st = alloc_cells(e, e->cell_off.back());
        if (st != MTLP_OK) {
            e->scene_params[scene] = old;
            e->allocated = false;
            return st;
        }
    } else {
        mtlp_status st = layout_scenes(e);  // checks the shared iteration count
        if (st != MTLP_OK) {
            e->scene_params[scene] = old;
            (void)layout_scenes(e);
            return st;
        }
    }
    DeviceGuard g(e->device);
    HIP_TRY(hipDeviceSynchronize());
    return upload_scene_table(e);
}

mtlp_status mtlp_set_params(mtlp_engine* e, const mtlp_params* p) {
    if (!e) return fail(MTLP_E_INVALID, "engine is NULL");
    if (e->batch) return fail(MTLP_E_INVALID, "use mtlp_batch_set_params on a batch");
    return mtlp_batch_set_params(e, 0, p);
}

mtlp_status mtlp_get_params(const mtlp_engine* e, mtlp_params* out) {
    if (!e || !out) return fail(MTLP_E_INVALID, "NULL argument");
    *out = e->scene_params[0];
    return MTLP_OK;
}

mtlp_status mtlp_upload_state(mtlp_engine* e, const float* x4, const float* v4, uint32_t n) {
    MTLP_TRY(check_engine(e));
    if (n != e->n) return fail(MTLP_E_RANGE, "n differs from the engine's particle count");
    if (n == 0) return MTLP_OK;
    DeviceGuard g(e->device);
    HIP_TRY(hipDeviceSynchronize());
    if (x4) HIP_TRY(hipMemcpy(e->x, x4, sizeof(float) * 4 * (size_t)n, hipMemcpyHostToDevice));
    if (v4) HIP_TRY(hipMemcpy(e->v, v4, sizeof(float) * 4 * (size_t)n, hipMemcpyHostToDevice));
    return MTLP_OK;
}

mtlp_status mtlp_set_state_device(mtlp_engine* e, mtlp_stream stream, const float* d_x4, const float* d_v4) {
    MTLP_TRY(check_engine(e));
    if (e->n == 0) return MTLP_OK;
    DeviceGuard g(e->device);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const size_t bytes = sizeof(float4) * e->n;
    if (d_x4) HIP_TRY(hipMemcpyAsync(e->x, d_x4, bytes, hipMemcpyDeviceToDevice, s));
    if (d_v4) HIP_TRY(hipMemcpyAsync(e->v, d_v4, bytes, hipMemcpyDeviceToDevice, s));
    return MTLP_OK;
}

mtlp_status mtlp_get_state_device(mtlp_engine* e, mtlp_stream stream, float* d_x4, float* d_v4) {
    MTLP_TRY(check_engine(e));
    if (e->n == 0) return MTLP_OK;
    DeviceGuard g(e->device);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const size_t bytes = sizeof(float4) * e->n;
    if (d_x4) HIP_TRY(hipMemcpyAsync(d_x4, e->x, bytes, hipMemcpyDeviceToDevice, s));
    if (d_v4) HIP_TRY(hipMemcpyAsync(d_v4, e->v, bytes, hipMemcpyDeviceToDevice, s));
    return MTLP_OK;
}

mtlp_status mtlp_step(mtlp_engine* e, mtlp_stream stream, float dt) {
    MTLP_TRY(check_engine(e));
    if (std::isnan(dt) || std::isinf(dt)) return fail(MTLP_E_INVALID, "dt is not finite");
    if (dt <= 0.0f) return MTLP_OK;  // Engine.cpp:686
    if (e->n == 0) return MTLP_OK;
    DeviceGuard g(e->device);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (e->sphere.on) {
        advance_sphere(e->sphere, e->scene_params[0], dt);
        return enqueue_step(e, s, dt);  // the sphere moves every step: no graph replay
    }
    if (e->use_graph) {
        if (!e->graph_exec || e->graph_dt != dt) MTLP_TRY(capture_graph(e, dt));
        HIP_TRY(hipGraphLaunch(e->graph_exec, s));
        e->final_buf = (int)(e->iterations() & 1u);
        return MTLP_OK;
    }
    return enqueue_step(e, s, dt);
}

mtlp_status mtlp_integrate(mtlp_engine* e, mtlp_stream stream, float dt) {
    MTLP_TRY(check_engine(e));
    if (e->batch) return fail(MTLP_E_INVALID, "integrate runs on single-scene engines");
    if (e->n == 0) return MTLP_OK;
    DeviceGuard g(e->device);
    // Engine.cpp:578-595: fixed gravity and damping, the scene's box
    HIP_TRY(launch_integrate(e->x, e->v, e->n, dt, -9.81f, 0.999f, make_final(e->scene_params[0]),
                             reinterpret_cast<hipStream_t>(stream)));
    return MTLP_OK;
}

mtlp_status mtlp_apply_mouse_pulse(mtlp_engine* e, mtlp_stream stream, const float ray_origin[3],
                                   const float ray_dir[3], const float force[3], float radius) {
    MTLP_TRY(check_engine(e));
    if (!ray_origin || !ray_dir || !force) return fail(MTLP_E_INVALID, "NULL argument");
    if (e->n == 0) return MTLP_OK;
    DeviceGuard g(e->device);
    HIP_TRY(launch_mouse_pulse(e->x, e->v, e->n, ray_origin, ray_dir, force, radius,
                               reinterpret_cast<hipStream_t>(stream)));
    return MTLP_OK;
}

mtlp_status mtlp_set_sphere(mtlp_engine* e, int enabled, const float center[3], const float velocity[3],
                            float radius) {
    if (!e) return fail(MTLP_E_INVALID, "engine is NULL");
    if (enabled && e->batch) return fail(MTLP_E_INVALID, "the kinematic sphere is for single-scene engines");
    if (enabled && !(radius > 0.0f)) return fail(MTLP_E_INVALID, "sphere radius must be > 0");
    e->sphere.on = enabled != 0;
    for (int a = 0; a < 3; ++a) {
        if (center) e->sphere.c[a] = center[a];
        if (velocity) e->sphere.v[a] = velocity[a];
    }
    if (radius > 0.0f) e->sphere.r = radius;
    return MTLP_OK;
}

mtlp_status mtlp_get_sphere(const mtlp_engine* e, int* enabled, float center[3], float velocity[3],
                            float* radius) {
    if (!e) return fail(MTLP_E_INVALID, "engine is NULL");
    if (enabled) *enabled = e->sphere.on ? 1 : 0;
    for (int a = 0; a < 3; ++a) {
        if (center) center[a] = e->sphere.c[a];
        if (velocity) velocity[a] = e->sphere.v[a];
    }
    if (radius) *radius = e->sphere.r;
    return MTLP_OK;
}

mtlp_status mtlp_step_n(mtlp_engine* e, mtlp_stream stream, float dt, uint32_t n_steps) {
    for (uint32_t k = 0; k < n_steps; ++k) MTLP_TRY(mtlp_step(e, stream, dt));
    return MTLP_OK;
}

mtlp_status mtlp_set_graph(mtlp_engine* e, int enable) {
    if (!e) return fail(MTLP_E_INVALID, "engine is NULL");
    DeviceGuard g(e->device);
    e->use_graph = enable != 0;
    if (!e->use_graph) e->drop_graph();
    return MTLP_OK;
}

mtlp_status mtlp_set_neighbor_lists(mtlp_engine* e, int enable) {
    if (!e) return fail(MTLP_E_INVALID, "engine is NULL");
    if (e->use_lists != (enable != 0)) e->drop_graph();
    e->use_lists = enable != 0;
    return MTLP_OK;
}

mtlp_status mtlp_set_stencil_block(mtlp_engine* e, uint32_t block) {
    if (!e) return fail(MTLP_E_INVALID, "engine is NULL");
    bool ok = false;
    for (uint32_t b : kStencilBlocks) ok |= b == block;
    if (!ok) return fail(MTLP_E_INVALID, "stencil block must be 64, 128, 256 or 512");
    if (e->stencil_block != block) e->drop_graph();
    e->stencil_block = block;
    return MTLP_OK;
}

mtlp_status mtlp_set_tile_pull(mtlp_engine* e, uint32_t tiles_per_pull) {
    if (!e) return fail(MTLP_E_INVALID, "engine is NULL");
    if (tiles_per_pull > kMaxTilePull) return fail(MTLP_E_RANGE, "tiles per pull above 64");
    if (e->tile_pull == tiles_per_pull) return MTLP_OK;
    // the counters are left as the last persistent launch of each pass left
    // them: zero them for the next
    if (e->tile_ctr) {
        DeviceGuard g(e->device);
        HIP_TRY(hipDeviceSynchronize());
        HIP_TRY(hipMemset(e->tile_ctr, 0, 2 * kTileCtrs * sizeof(uint32_t)));
    }
    e->drop_graph();
    e->tile_pull = tiles_per_pull;
    return MTLP_OK;
}

mtlp_status mtlp_set_packed_positions(mtlp_engine* e, int enable) {
    if (!e) return fail(MTLP_E_INVALID, "engine is NULL");
    if (enable && e->allocated && !e->xyz[0])
        return fail(MTLP_E_RANGE, "engine too large for packed position copies");
    // every step's hash writes iteration 0's copy and each delta pass the
    // next one before a density pass reads it, so switching needs no sync
    if (e->packed_positions != (enable != 0)) e->drop_graph();
    e->packed_positions = enable != 0;
    return MTLP_OK;
}

mtlp_status mtlp_launch_config(const mtlp_engine* e, uint32_t* block, uint32_t* tiles_per_pull,
                               int* packed_positions) {
    if (!e) return fail(MTLP_E_INVALID, "engine is NULL");
    if (block) *block = e->stencil_block;
    if (tiles_per_pull) *tiles_per_pull = e->tile_pull;
    if (packed_positions) *packed_positions = e->packed_positions && (!e->allocated || e->xyz[0]) ? 1 : 0;
    return MTLP_OK;
}

mtlp_status mtlp_autotune(mtlp_engine* e, mtlp_stream stream, float dt, const char* cache_path,
                          uint32_t* block) {
    MTLP_TRY(check_engine(e));
    if (!(dt > 0.0f) || !std::isfinite(dt)) return fail(MTLP_E_INVALID, "dt must be > 0");
    std::ostringstream key;
    key << "launch,n=" << e->n << ",scenes=" << e->scene_params.size() << ",cells=" << e->cells
        << ",iters=" << e->iterations();
    if (cache_path) {
        std::ifstream in(cache_path);
        std::string line;
        while (std::getline(in, line)) {
            std::istringstream ls(line);
            std::string k;
            uint32_t b = 0, pull = 0, packed = 0;
            if (ls >> k >> b >> pull >> packed && k == key.str() && mtlp_set_stencil_block(e, b) == MTLP_OK &&
                mtlp_set_tile_pull(e, pull) == MTLP_OK && mtlp_set_packed_positions(e, (int)packed) == MTLP_OK) {
                if (block) *block = b;
                return MTLP_OK;
            }
        }
    }
    if (e->n == 0) {
        if (block) *block = e->stencil_block;
        return MTLP_OK;
    }
    DeviceGuard g(e->device);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    // The tuning steps run from a copy of the state and every buffer a
    // caller can read is put back afterwards: positions, velocities and
    // previous positions, and the last step's outputs (cell hashes and
    // offsets, sorted index and positions, neighbour counts, lambdas, foam).
    const size_t n = e->n, c1 = (size_t)e->cells + 1;
    struct Part {
        void* buf;
        size_t bytes;
    };
    const Part parts[] = {{e->x, 16 * n}, {e->v, 16 * n}, {e->prev, 16 * n}, {e->keys, 4 * n},
                          {e->cell_start, 4 * c1}, {e->sorted_index, 4 * n}, {e->nbr, 4 * n},
                          {e->lambdas, 4 * n}, {e->foam, 4 * n}, {e->sorted[e->final_buf], 16 * n}};
    size_t total = 0;
    for (const Part& q : parts) total += (q.bytes + 255) / 256 * 256;
    struct Scratch {  // freed on every return path
        char* p = nullptr;
        hipEvent_t t0 = nullptr, t1 = nullptr;
        ~Scratch() {
            if (p) (void)hipFree(p);
            if (t0) (void)hipEventDestroy(t0);
            if (t1) (void)hipEventDestroy(t1);
        }
    } sc;
    HIP_TRY(hipMalloc(reinterpret_cast<void**>(&sc.p), total));
    HIP_TRY(hipEventCreate(&sc.t0));
    HIP_TRY(hipEventCreate(&sc.t1));
    auto copy_parts = [&](bool save) {
        size_t off = 0;
        for (const Part& q : parts) {
            void* dst = save ? sc.p + off : q.buf;
            const void* src = save ? q.buf : sc.p + off;
            const hipError_t r = hipMemcpyAsync(dst, src, q.bytes, hipMemcpyDeviceToDevice, s);
            if (r != hipSuccess) return r;
            off += (q.bytes + 255) / 256 * 256;
        }
        return hipSuccess;
    };
    const int final_buf = e->final_buf;
    HIP_TRY(copy_parts(true));
    const bool graph = e->use_graph, prof = e->profiling;
    const uint32_t old_block = e->stencil_block, old_pull = e->tile_pull;
    const bool old_packed = e->packed_positions;
    const uint32_t packs = e->xyz[0] ? 2u : 1u;  // packed positions on, off (only off without copies)
    e->use_graph = false;
    e->profiling = false;
    e->drop_graph();
    constexpr uint32_t kLaunches = kNumStencilBlocks * kNumTilePulls;
    const uint32_t configs = kLaunches * packs;
    float cfg_ms[2 * kLaunches];
    for (float& m : cfg_ms) m = 1e30f;
    mtlp_status st = MTLP_OK;
    // two passes over the configurations (workgroup size x tiles per pull x
    // packed positions), each keeping its faster pass, so that neither their
    // order nor one noisy timing decides
    for (uint32_t pass = 0; pass < 2 * configs && st == MTLP_OK; ++pass) {
        const uint32_t ci = pass % configs, li = ci % kLaunches;
        e->stencil_block = kStencilBlocks[li / kNumTilePulls];
        e->packed_positions = packs == 2u && ci < kLaunches;
        st = mtlp_set_tile_pull(e, kTilePulls[li % kNumTilePulls]);
        float ms = 0.0f;
        // rep 0 warms up with one step; rep 1 times kTuneSteps steps, both
        // from the saved state (a single step from a fresh state can
        // favour a configuration that later steps do not)
        constexpr int kTuneSteps = 8;
        for (int rep = 0; rep < 2 && st == MTLP_OK; ++rep) {
            if (hipMemcpyAsync(e->x, sc.p, 16 * n, hipMemcpyDeviceToDevice, s) != hipSuccess ||
                hipMemcpyAsync(e->v, sc.p + (16 * n + 255) / 256 * 256, 16 * n, hipMemcpyDeviceToDevice, s) !=
                    hipSuccess ||
                hipEventRecord(sc.t0, s) != hipSuccess) {
                st = fail(MTLP_E_HIP, "autotune copy/event failed");
                break;
            }
            for (int k = 0; k < (rep ? kTuneSteps : 1) && st == MTLP_OK; ++k) st = enqueue_step(e, s, dt);
            if (st != MTLP_OK || hipEventRecord(sc.t1, s) != hipSuccess || hipEventSynchronize(sc.t1) != hipSuccess ||
                hipEventElapsedTime(&ms, sc.t0, sc.t1) != hipSuccess) {
                if (st == MTLP_OK) st = fail(MTLP_E_HIP, "autotune timing failed");
                break;
            }
        }
        if (st == MTLP_OK) cfg_ms[ci] = std::min(cfg_ms[ci], ms);
    }
    // The configuration the engine ran with stays unless another is more
    // than kTuneMargin faster: neighbouring workgroup sizes are within a few
    // per cent of each other on the dam break, and short timings flip them.
    constexpr float kTuneMargin = 0.03f;
    uint32_t best = 0;
    for (uint32_t ci = 1; ci < configs; ++ci)
        if (cfg_ms[ci] < cfg_ms[best]) best = ci;
    uint32_t incumbent = configs;
    for (uint32_t ci = 0; ci < configs; ++ci) {
        const uint32_t li = ci % kLaunches;
        if (kStencilBlocks[li / kNumTilePulls] == old_block && kTilePulls[li % kNumTilePulls] == old_pull &&
            (packs == 2u && ci < kLaunches) == old_packed)
            incumbent = ci;
    }
    if (incumbent < configs && cfg_ms[best] > cfg_ms[incumbent] * (1.0f - kTuneMargin)) best = incumbent;
    const hipError_t restore = copy_parts(false);
    const hipError_t sync = hipStreamSynchronize(s);
    e->final_buf = final_buf;
    e->use_graph = graph;
    e->profiling = prof;
    if (st != MTLP_OK) {
        e->stencil_block = old_block;
        e->packed_positions = old_packed;
        (void)mtlp_set_tile_pull(e, old_pull);
        return st;
    }
    HIP_TRY(restore);
    HIP_TRY(sync);
    e->stencil_block = kStencilBlocks[(best % kLaunches) / kNumTilePulls];
    e->packed_positions = packs == 2u && best < kLaunches;
    MTLP_TRY(mtlp_set_tile_pull(e, kTilePulls[(best % kLaunches) % kNumTilePulls]));
    e->drop_graph();
    if (cache_path) {
        std::ofstream out(cache_path, std::ios::app);
        out << key.str() << " " << e->stencil_block << " " << e->tile_pull << " " << (e->packed_positions ? 1 : 0)
            << "\n";
    }
    if (block) *block = e->stencil_block;
    return MTLP_OK;
}

mtlp_status mtlp_build_spatial_hash(mtlp_engine* e, mtlp_stream stream) {
    MTLP_TRY(check_engine(e));
    if (e->n == 0) return MTLP_OK;
    DeviceGuard g(e->device);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    StageClock clk(e, s);
    const StepArgs a = e->step_args(0.0f, false);
    MTLP_TRY(enqueue_hash(e, s, a, clk, e->hash_bufs()));
    HIP_TRY(launch_count_neighbors(e->sorted[0], e->cell_start, e->sorted_index, a, e->nbr, s));
    clk.mark("count_neighbors");
    e->final_buf = 0;
    return MTLP_OK;
}

size_t mtlp_scan_scratch_bytes(uint32_t n) {
    return (size_t)scan_tiles(n > 0 ? n : 1) * sizeof(uint64_t) + 4 * sizeof(uint32_t);
}

mtlp_status mtlp_exclusive_scan(mtlp_engine* e, mtlp_stream stream, const uint32_t* d_in,
                                uint32_t* d_out, void* d_scratch, uint32_t n) {
    if (n == 0) return MTLP_OK;  // Engine.cpp:848
    if (n > MTLP_MAX_SCAN_LENGTH) return fail(MTLP_E_RANGE, "scan length above MTLP_MAX_SCAN_LENGTH");
    if (!d_in || !d_out || !d_scratch) return fail(MTLP_E_INVALID, "NULL buffer");
    if ((reinterpret_cast<uintptr_t>(d_in) | reinterpret_cast<uintptr_t>(d_out)) & 15u)
        return fail(MTLP_E_INVALID, "scan buffers must be 16-byte aligned");
    if (reinterpret_cast<uintptr_t>(d_scratch) & 7u)
        return fail(MTLP_E_INVALID, "scan scratch must be 8-byte aligned");
    int dev = 0;
    if (e) dev = e->device; else HIP_TRY(hipGetDevice(&dev));
    DeviceGuard g(dev);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const uint32_t tiles = scan_tiles(n);
    uint64_t* state = static_cast<uint64_t*>(d_scratch);
    uint32_t* words = reinterpret_cast<uint32_t*>(state + tiles);  // [0] ticket, [1] error
    HIP_TRY(launch_fill_u32(static_cast<uint32_t*>(d_scratch), 0u, 2 * (size_t)tiles + 4, s));
    uint32_t* err = e ? e->misc + 2 : words + 1;
    HIP_TRY(launch_exclusive_scan(const_cast<uint32_t*>(d_in), d_out, n, false, state, words, 1u,
                                  err, s));
    return MTLP_OK;
}

mtlp_status mtlp_exclusive_scan_noalloc(mtlp_stream stream, const uint32_t* d_in, uint32_t* d_out, uint32_t n) {
    if (n == 0) return MTLP_OK;
    if (n > MTLP_NOALLOC_SCAN_MAX)
        return fail(MTLP_E_RANGE, "scan length above MTLP_NOALLOC_SCAN_MAX (2^28) for the allocation-free scan");
    if (!d_in || !d_out) return fail(MTLP_E_INVALID, "NULL buffer");
    if ((reinterpret_cast<uintptr_t>(d_in) | reinterpret_cast<uintptr_t>(d_out)) & 15u)
        return fail(MTLP_E_INVALID, "scan buffers must be 16-byte aligned");
    static_assert(MTLP_NOALLOC_SCAN_MAX == kNoAllocScanMax, "header and kernel limits agree");
    // the current device's copy of the scan's working memory (no allocation)
    uint64_t* scratch = nullptr;
    HIP_TRY(noalloc_scan_scratch(&scratch));
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const uint32_t tiles = scan_tiles(n);
    uint32_t* words = reinterpret_cast<uint32_t*>(scratch);  // [0] ticket, [1] error
    uint64_t* state = scratch + 2;
    // the last call's look-back words and ticket are cleared by a kernel (not
    // a memset), so the call may also be captured into a graph
    HIP_TRY(launch_fill_u32(words, 0u, 4 + 2 * (size_t)tiles, s));
    HIP_TRY(launch_exclusive_scan(const_cast<uint32_t*>(d_in), d_out, n, false, state, words, 1u, words + 1, s));
    return MTLP_OK;
}

mtlp_status mtlp_get_buffer(mtlp_engine* e, mtlp_buffer which, void** d_ptr, size_t* bytes) {
    MTLP_TRY(check_engine(e));
    if (!d_ptr) return fail(MTLP_E_INVALID, "d_ptr is NULL");
    void* p = nullptr;
    size_t b = 0;
    const size_t n = e->n;
    switch (which) {
        case MTLP_BUF_POSITIONS: p = e->x; b = n * sizeof(float4); break;
        case MTLP_BUF_VELOCITIES: p = e->v; b = n * sizeof(float4); break;
        case MTLP_BUF_PREV_POSITIONS: p = e->prev; b = n * sizeof(float4); break;
        case MTLP_BUF_CELL_HASHES: p = e->keys; b = n * sizeof(uint32_t); break;
        case MTLP_BUF_CELL_START: p = e->cell_start; b = ((size_t)e->cells + 1) * sizeof(uint32_t); break;
        case MTLP_BUF_SORTED_INDEX: p = e->sorted_index; b = n * sizeof(uint32_t); break;
        case MTLP_BUF_NEIGHBOR_COUNTS: p = e->nbr; b = n * sizeof(uint32_t); break;
        case MTLP_BUF_LAMBDAS: p = e->lambdas; b = n * sizeof(float); break;
        case MTLP_BUF_SORTED_POSITIONS: p = e->sorted[e->final_buf]; b = n * sizeof(float4); break;
        case MTLP_BUF_FOAM_INTENSITY: p = e->foam; b = n * sizeof(float); break;
        default: return fail(MTLP_E_INVALID, "unknown buffer id");
    }
    *d_ptr = p;
    if (bytes) *bytes = b;
    return MTLP_OK;
}

mtlp_status mtlp_read_buffer(mtlp_engine* e, mtlp_buffer which, void* host_dst, size_t bytes) {
    void* p = nullptr;
    size_t b = 0;
    MTLP_TRY(mtlp_get_buffer(e, which, &p, &b));
    if (bytes > b) return fail(MTLP_E_RANGE, "read larger than the buffer");
    if (bytes == 0) return MTLP_OK;
    if (!host_dst) return fail(MTLP_E_INVALID, "host_dst is NULL");
    DeviceGuard g(e->device);
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(host_dst, p, bytes, hipMemcpyDeviceToHost));
    return MTLP_OK;
}

mtlp_status mtlp_write_buffer(mtlp_engine* e, mtlp_buffer which, const void* host_src, size_t bytes) {
    void* p = nullptr;
    size_t b = 0;
    MTLP_TRY(mtlp_get_buffer(e, which, &p, &b));
    if (bytes > b) return fail(MTLP_E_RANGE, "write larger than the buffer");
    if (bytes == 0) return MTLP_OK;
    if (!host_src) return fail(MTLP_E_INVALID, "host_src is NULL");
    DeviceGuard g(e->device);
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(p, host_src, bytes, hipMemcpyHostToDevice));
    return MTLP_OK;
}

uint32_t mtlp_particle_count(const mtlp_engine* e) { return e ? e->n : 0u; }

uint32_t mtlp_total_cells(const mtlp_engine* e) {
    if (!e) return 0u;
    if (e->allocated) return e->cells;
    uint32_t d[3];
    return validate_params(&e->scene_params[0], d) == MTLP_OK ? d[0] * d[1] * d[2] : 0u;
}

float mtlp_cell_size(const mtlp_engine* e) { return e ? e->scene_params[0].cell_size : 0.0f; }

mtlp_status mtlp_set_profiling(mtlp_engine* e, int enable) {
    if (!e) return fail(MTLP_E_INVALID, "engine is NULL");
    DeviceGuard g(e->device);
    if (enable && e->events.empty()) {
        const size_t need = 8 + 2 * (size_t)kMaxIterations;
        e->events.resize(need, nullptr);
        for (auto& ev : e->events) HIP_TRY(hipEventCreate(&ev));
    }
    e->profiling = enable != 0;
    return MTLP_OK;
}

mtlp_status mtlp_stage_times(mtlp_engine* e, float* ms, const char** names, uint32_t capacity,
                             uint32_t* count) {
    if (!e || !count) return fail(MTLP_E_INVALID, "NULL argument");
    *count = 0;
    if (!e->profiling || e->stages_recorded == 0) return MTLP_OK;
    DeviceGuard g(e->device);
    HIP_TRY(hipEventSynchronize(e->events[e->stages_recorded]));
    const uint32_t k = e->stages_recorded < capacity ? e->stages_recorded : capacity;
    for (uint32_t i = 0; i < k; ++i) {
        float t = 0.0f;
        HIP_TRY(hipEventElapsedTime(&t, e->events[i], e->events[i + 1]));
        if (ms) ms[i] = t;
        if (names) names[i] = e->stage_names[i];
    }
    *count = k;
    return MTLP_OK;
}

mtlp_status mtlp_error_flags(mtlp_engine* e, uint32_t* flags) {
    if (!e || !flags) return fail(MTLP_E_INVALID, "NULL argument");
    DeviceGuard g(e->device);
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(flags, e->misc + 2, sizeof(uint32_t), hipMemcpyDeviceToHost));
    const uint32_t zero = 0;
    HIP_TRY(hipMemcpy(e->misc + 2, &zero, sizeof(uint32_t), hipMemcpyHostToDevice));
    return MTLP_OK;
}

}  // extern "C"
