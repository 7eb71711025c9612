// engine.cpp -- C++20 host engine on the HIP runtime and its C ABI
// (include/mtlp.h).
//
// Restates the orchestration of mtlphys::Engine (engine/src/Engine.cpp):
// buffer allocation (:381-452), parameter derivation (:551-572), the spatial
// hash (:601-679), the step (:685-811) and the standalone scan (:843-882).
// Every stage is enqueued on the caller's stream with its constants passed
// by value; nothing inside a step waits for the device.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "mtlp_host.h"
#include "mtlp_kernels.h"

using namespace mtlp;

namespace {

thread_local std::string g_last_error;

mtlp_status fail(mtlp_status s, const std::string& msg) {
    g_last_error = msg;
    return s;
}

#define HIP_TRY(expr)                                                                     \
    do {                                                                                  \
        const hipError_t e_ = (expr);                                                     \
        if (e_ != hipSuccess)                                                             \
            return fail(MTLP_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));   \
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
constexpr uint32_t kEpochLimit = (1u << 30) - 1u;

uint32_t round_up(uint32_t n, uint32_t m) { return ((n + m - 1) / m) * m; }

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
// applyDeltaSorted (Constraints.metal:158-160, 190), all in FP32.
SolverDev make_solver(const mtlp_params& p) {
    const float kPi = 3.14159265358979323846f;
    const float h = p.h, h2 = h * h, h3 = h2 * h, h6 = h3 * h3, h9 = h6 * h3;
    SolverDev s{};
    s.h = h;
    s.h2 = h2;
    s.poly6 = 315.0f / (64.0f * kPi * h9);
    s.spiky = 45.0f / (kPi * h6);
    s.inv_rho0 = 1.0f / p.rest_density;
    s.eps = p.epsilon;
    s.w0 = s.poly6 * h2 * h2 * h2;
    const float dqh = p.scorr_dq * h;
    const float dq2 = dqh * dqh;
    const float dd = h2 - dq2;
    s.inv_wq = 1.0f / (s.poly6 * dd * dd * dd);
    s.scorr_k = p.scorr_k;
    s.scorr_n = p.scorr_n;
    s.max_step = p.max_step * h;
    const double hd = h, h2d = hd * hd, h3d = h2d * hd, h6d = h3d * h3d, h9d = h6d * h3d;
    s.h2_d = h2d;
    s.poly6_d = 315.0 / (64.0 * 3.14159265358979323846 * h9d);
    s.w0_d = s.poly6_d * h2d * h2d * h2d;
    s.inv_rho0_d = 1.0 / (double)p.rest_density;
    s.h_d = hd;
    s.spiky_d = 45.0 / (3.14159265358979323846 * h6d);
    const double dqhd = (double)p.scorr_dq * hd, ddd = h2d - dqhd * dqhd;
    s.inv_wq_d = 1.0 / (s.poly6_d * ddd * ddd * ddd);
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

FinalDev make_final(const mtlp_params& p, float dt) {
    FinalDev f{};
    f.bminx = p.bounds_min[0];
    f.bminy = p.bounds_min[1];
    f.bminz = p.bounds_min[2];
    f.bmaxx = p.bounds_max[0];
    f.bmaxy = p.bounds_max[1];
    f.bmaxz = p.bounds_max[2];
    f.inv_dt = 1.0f / dt;  // Engine.cpp:561
    f.damping = p.damping;
    f.weps = p.wall_eps;
    f.floor_f = p.floor_friction;
    f.wall_f = p.wall_friction;
    f.vmax = p.vmax * p.h * f.inv_dt;  // Constraints.metal:347
    return f;
}

PredictDev make_predict(const mtlp_params& p, float dt, bool enable) {
    PredictDev d{};
    d.gdtx = p.gravity[0] * dt;
    d.gdty = p.gravity[1] * dt;
    d.gdtz = p.gravity[2] * dt;
    d.dt = dt;
    d.enable = enable ? 1u : 0u;
    return d;
}

template <class T>
hipError_t dev_alloc(T** p, size_t count) {
    *p = nullptr;
    return hipMalloc(reinterpret_cast<void**>(p), sizeof(T) * (count ? count : 1));
}

template <class T>
void dev_free(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

}  // namespace

struct mtlp_engine {
    int device = 0;
    mtlp_params params{};
    uint32_t dims[3] = {0, 0, 0};
    uint32_t cells = 0;
    uint32_t n = 0;
    uint32_t n_alloc = 0;
    uint32_t scene = MTLP_SCENE_CUBE_DROP;
    bool allocated = false;

    // particle buffers (original order)
    float4* x = nullptr;
    float4* v = nullptr;
    float4* prev = nullptr;
    uint32_t* keys = nullptr;
    uint32_t* tickets = nullptr;
    uint32_t* nbr = nullptr;
    float* lambdas = nullptr;
    // particle buffers (sorted order)
    uint32_t* tmp_index = nullptr;
    uint32_t* sorted_index = nullptr;
    float4* sorted[2] = {nullptr, nullptr};
    uint32_t* big_list = nullptr;
    int final_buf = 0;
    // cell buffers
    uint32_t* counts = nullptr;
    uint32_t* cell_start = nullptr;
    uint64_t* scan_state = nullptr;
    uint32_t scan_epoch = 0;
    // misc device words: [0] big-cell count, [1] scan ticket, [2] error flags
    uint32_t* misc = nullptr;

    // profiling
    bool profiling = false;
    std::vector<hipEvent_t> events;
    std::vector<const char*> stage_names;
    uint32_t stages_recorded = 0;

    void free_particles() {
        dev_free(x); dev_free(v); dev_free(prev); dev_free(keys); dev_free(tickets);
        dev_free(nbr); dev_free(lambdas); dev_free(tmp_index); dev_free(sorted_index);
        dev_free(sorted[0]); dev_free(sorted[1]); dev_free(big_list);
        n_alloc = 0;
    }
    void free_cells() {
        dev_free(counts); dev_free(cell_start); dev_free(scan_state);
        cells = 0;
    }
    ~mtlp_engine() {
        DeviceGuard g(device);
        free_particles();
        free_cells();
        dev_free(misc);
        for (hipEvent_t ev : events) (void)hipEventDestroy(ev);
    }

    HashBufs hash_bufs() {
        HashBufs b{};
        b.x = x; b.v = v; b.keys = keys; b.tickets = tickets; b.counts = counts;
        b.cell_start = cell_start; b.tmp_index = tmp_index; b.sorted_index = sorted_index;
        b.sorted = sorted[0]; b.big_list = big_list; b.big_count = misc + 0;
        b.scan_state = scan_state; b.scan_ticket = misc + 1; b.err = misc + 2;
        return b;
    }

    uint32_t next_epoch(hipStream_t s, mtlp_status* st) {
        *st = MTLP_OK;
        if (scan_epoch >= kEpochLimit) {
            const hipError_t e = launch_fill_u32(reinterpret_cast<uint32_t*>(scan_state), 0u,
                                                 2 * (size_t)scan_tiles(cells + 1), s);
            if (e != hipSuccess) *st = fail(MTLP_E_HIP, hipGetErrorString(e));
            scan_epoch = 0;
        }
        return ++scan_epoch;
    }
};

namespace {

// Records a HIP event after each kernel when profiling is on.
struct StageClock {
    mtlp_engine* e;
    hipStream_t s;
    uint32_t k = 0;
    StageClock(mtlp_engine* eng, hipStream_t stream) : e(eng), s(stream) {
        if (!e->profiling) return;
        e->stage_names.clear();
        e->stages_recorded = 0;
        (void)hipEventRecord(e->events[0], s);
    }
    void mark(const char* name) {
        if (!e->profiling) return;
        if (k + 1 >= e->events.size()) return;
        ++k;
        (void)hipEventRecord(e->events[k], s);
        e->stage_names.push_back(name);
        e->stages_recorded = k;
    }
};

mtlp_status alloc_cells(mtlp_engine* e, const uint32_t dims[3]) {
    e->free_cells();
    const uint32_t C = dims[0] * dims[1] * dims[2];
    const size_t tiles = scan_tiles(C + 1);
    HIP_TRY(dev_alloc(&e->counts, (size_t)C + 1));
    HIP_TRY(dev_alloc(&e->cell_start, (size_t)C + 1));
    HIP_TRY(dev_alloc(&e->scan_state, tiles));
    HIP_TRY(hipMemset(e->counts, 0, sizeof(uint32_t) * ((size_t)C + 1)));
    HIP_TRY(hipMemset(e->cell_start, 0, sizeof(uint32_t) * ((size_t)C + 1)));
    HIP_TRY(hipMemset(e->scan_state, 0, sizeof(uint64_t) * tiles));
    e->scan_epoch = 0;
    std::memcpy(e->dims, dims, sizeof(e->dims));
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
    HIP_TRY(dev_alloc(&e->tmp_index, na));
    HIP_TRY(dev_alloc(&e->sorted_index, na));
    HIP_TRY(dev_alloc(&e->sorted[0], na));
    HIP_TRY(dev_alloc(&e->sorted[1], na));
    HIP_TRY(dev_alloc(&e->big_list, na / kSmallCell + 1));
    HIP_TRY(hipMemset(e->x, 0, sizeof(float4) * na));
    HIP_TRY(hipMemset(e->v, 0, sizeof(float4) * na));
    HIP_TRY(hipMemset(e->prev, 0, sizeof(float4) * na));
    HIP_TRY(hipMemset(e->keys, 0, sizeof(uint32_t) * na));
    HIP_TRY(hipMemset(e->nbr, 0, sizeof(uint32_t) * na));
    HIP_TRY(hipMemset(e->lambdas, 0, sizeof(float) * na));
    HIP_TRY(hipMemset(e->sorted_index, 0, sizeof(uint32_t) * na));
    HIP_TRY(hipMemset(e->sorted[0], 0, sizeof(float4) * na));
    HIP_TRY(hipMemset(e->sorted[1], 0, sizeof(float4) * na));
    e->n_alloc = na;
    e->n = n;
    return MTLP_OK;
}

// Spatial hash into sorted[0]: predict (optional) + hash + count, scan,
// stable scatter + gather.  Engine.cpp:601-667 and :715-726.
mtlp_status enqueue_hash(mtlp_engine* e, hipStream_t s, const PredictDev& pd, StageClock& clk) {
    const GridDev g = make_grid(e->params, e->dims);
    HashBufs b = e->hash_bufs();
    HIP_TRY(launch_predict_hash_count(b, e->n, pd, g, s));
    clk.mark("predict_hash_count");
    mtlp_status st;
    const uint32_t epoch = e->next_epoch(s, &st);
    if (st != MTLP_OK) return st;
    HIP_TRY(launch_exclusive_scan(e->counts, e->cell_start, e->cells + 1, true, e->scan_state,
                                  e->misc + 1, epoch, e->misc + 2, s));
    clk.mark("exclusive_scan");
    HIP_TRY(launch_scatter_gather(b, e->n, pd, s));
    clk.mark("stable_scatter_gather");
    return MTLP_OK;
}

mtlp_status check_engine(const mtlp_engine* e) {
    if (!e) return fail(MTLP_E_INVALID, "engine is NULL");
    if (!e->allocated) return fail(MTLP_E_INVALID, "engine has no buffers: call mtlp_reset first");
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
    mtlp_status st = validate_params(&params, dims);
    if (st != MTLP_OK) return st;
    int count = 0;
    HIP_TRY(hipGetDeviceCount(&count));
    if (hip_device < 0 || hip_device >= count) return fail(MTLP_E_INVALID, "no such HIP device");
    mtlp_engine* e = new (std::nothrow) mtlp_engine();
    if (!e) return fail(MTLP_E_NOMEM, "host allocation failed");
    e->device = hip_device;
    e->params = params;
    std::memcpy(e->dims, dims, sizeof(dims));
    DeviceGuard g(hip_device);
    if (!g.ok) { delete e; return fail(MTLP_E_HIP, "hipSetDevice failed"); }
    if (hipMalloc(reinterpret_cast<void**>(&e->misc), 4 * sizeof(uint32_t)) != hipSuccess ||
        hipMemset(e->misc, 0, 4 * sizeof(uint32_t)) != hipSuccess) {
        delete e;
        return fail(MTLP_E_NOMEM, "device allocation failed");
    }
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
    std::vector<float> hx(4 * (size_t)n), hv(4 * (size_t)n);
    mtlp_status st = generate_scene(scene, n, seed, e->params, hx.data(), hv.data());
    if (st != MTLP_OK) return fail(st, "unknown scene id");
    DeviceGuard g(e->device);
    HIP_TRY(hipDeviceSynchronize());
    e->allocated = false;
    st = alloc_particles(e, n);
    if (st != MTLP_OK) return st;
    const uint32_t dims[3] = {e->dims[0], e->dims[1], e->dims[2]};
    st = alloc_cells(e, dims);
    if (st != MTLP_OK) return st;
    if (n) {
        HIP_TRY(hipMemcpy(e->x, hx.data(), sizeof(float) * 4 * (size_t)n, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(e->v, hv.data(), sizeof(float) * 4 * (size_t)n, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(e->prev, hx.data(), sizeof(float) * 4 * (size_t)n, hipMemcpyHostToDevice));
    }
    e->scene = scene;
    e->final_buf = 0;
    e->allocated = true;
    return MTLP_OK;
}

mtlp_status mtlp_set_params(mtlp_engine* e, const mtlp_params* p) {
    if (!e) return fail(MTLP_E_INVALID, "engine is NULL");
    uint32_t dims[3];
    mtlp_status st = validate_params(p, dims);
    if (st != MTLP_OK) return st;
    const bool regrid = dims[0] != e->dims[0] || dims[1] != e->dims[1] || dims[2] != e->dims[2];
    e->params = *p;
    if (regrid) {
        std::memcpy(e->dims, dims, sizeof(dims));
        if (e->allocated) {
            DeviceGuard g(e->device);
            HIP_TRY(hipDeviceSynchronize());
            st = alloc_cells(e, dims);
            if (st != MTLP_OK) { e->allocated = false; return st; }
        }
    }
    return MTLP_OK;
}

mtlp_status mtlp_get_params(const mtlp_engine* e, mtlp_params* out) {
    if (!e || !out) return fail(MTLP_E_INVALID, "NULL argument");
    *out = e->params;
    return MTLP_OK;
}

mtlp_status mtlp_upload_state(mtlp_engine* e, const float* x4, const float* v4, uint32_t n) {
    mtlp_status st = check_engine(e);
    if (st != MTLP_OK) return st;
    if (n != e->n) return fail(MTLP_E_RANGE, "n differs from the engine's particle count");
    if (n == 0) return MTLP_OK;
    DeviceGuard g(e->device);
    HIP_TRY(hipDeviceSynchronize());
    if (x4) HIP_TRY(hipMemcpy(e->x, x4, sizeof(float) * 4 * (size_t)n, hipMemcpyHostToDevice));
    if (v4) HIP_TRY(hipMemcpy(e->v, v4, sizeof(float) * 4 * (size_t)n, hipMemcpyHostToDevice));
    return MTLP_OK;
}

mtlp_status mtlp_step(mtlp_engine* e, mtlp_stream stream, float dt) {
    mtlp_status st = check_engine(e);
    if (st != MTLP_OK) return st;
    if (std::isnan(dt) || std::isinf(dt)) return fail(MTLP_E_INVALID, "dt is not finite");
    if (dt <= 0.0f) return MTLP_OK;  // Engine.cpp:686
    if (e->n == 0) return MTLP_OK;
    DeviceGuard g(e->device);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    StageClock clk(e, s);
    const mtlp_params& p = e->params;
    const GridDev grid = make_grid(p, e->dims);
    const SolverDev sd = make_solver(p);

    st = enqueue_hash(e, s, make_predict(p, dt, true), clk);
    if (st != MTLP_OK) return st;

    int cur = 0;
    for (uint32_t it = 0; it < p.iterations; ++it) {  // Engine.cpp:732-774
        const bool last = it + 1 == p.iterations;
        HIP_TRY(launch_density_lambda(e->sorted[cur], e->cell_start, e->sorted_index, e->n, grid, sd,
                                      last ? e->lambdas : nullptr, it == 0 ? e->nbr : nullptr, s));
        clk.mark("density_lambda");
        HIP_TRY(launch_apply_delta(e->sorted[cur], e->cell_start, e->n, grid, sd, e->sorted[cur ^ 1], s));
        clk.mark("apply_delta");
        cur ^= 1;
    }
    if (p.iterations == 0) {
        HIP_TRY(launch_count_neighbors(e->sorted[0], e->cell_start, e->sorted_index, e->n, grid,
                                       p.h * p.h, e->nbr, s));
        clk.mark("count_neighbors");
    }
    HIP_TRY(launch_finalize(e->sorted[cur], e->sorted_index, e->n, make_final(p, dt), e->x, e->v,
                            e->prev, s));
    clk.mark("scatter_finalize");
    e->final_buf = cur;
    return MTLP_OK;
}

mtlp_status mtlp_build_spatial_hash(mtlp_engine* e, mtlp_stream stream) {
    mtlp_status st = check_engine(e);
    if (st != MTLP_OK) return st;
    if (e->n == 0) return MTLP_OK;
    DeviceGuard g(e->device);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    StageClock clk(e, s);
    st = enqueue_hash(e, s, make_predict(e->params, 0.0f, false), clk);
    if (st != MTLP_OK) return st;
    // neighborRadiusSq = kSmoothingH^2 (Engine.cpp:447)
    HIP_TRY(launch_count_neighbors(e->sorted[0], e->cell_start, e->sorted_index, e->n,
                                   make_grid(e->params, e->dims), e->params.h * e->params.h, e->nbr, s));
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

mtlp_status mtlp_get_buffer(mtlp_engine* e, mtlp_buffer which, void** d_ptr, size_t* bytes) {
    mtlp_status st = check_engine(e);
    if (st != MTLP_OK) return st;
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
        default: return fail(MTLP_E_INVALID, "unknown buffer id");
    }
    *d_ptr = p;
    if (bytes) *bytes = b;
    return MTLP_OK;
}

mtlp_status mtlp_read_buffer(mtlp_engine* e, mtlp_buffer which, void* host_dst, size_t bytes) {
    void* p = nullptr;
    size_t b = 0;
    mtlp_status st = mtlp_get_buffer(e, which, &p, &b);
    if (st != MTLP_OK) return st;
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
    mtlp_status st = mtlp_get_buffer(e, which, &p, &b);
    if (st != MTLP_OK) return st;
    if (bytes > b) return fail(MTLP_E_RANGE, "write larger than the buffer");
    if (bytes == 0) return MTLP_OK;
    if (!host_src) return fail(MTLP_E_INVALID, "host_src is NULL");
    DeviceGuard g(e->device);
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(p, host_src, bytes, hipMemcpyHostToDevice));
    return MTLP_OK;
}

uint32_t mtlp_particle_count(const mtlp_engine* e) { return e ? e->n : 0u; }
uint32_t mtlp_total_cells(const mtlp_engine* e) { return e ? e->dims[0] * e->dims[1] * e->dims[2] : 0u; }
float mtlp_cell_size(const mtlp_engine* e) { return e ? e->params.cell_size : 0.0f; }

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
