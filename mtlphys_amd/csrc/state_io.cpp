// state_io.cpp -- engine state to and from a binary file (include/mtlp.h
// describes the format).  The reference keeps its state in GPU buffers only
// and has no counterpart.
#include <cstdio>
#include <cstring>
#include <memory>
#include <vector>

#include "engine_impl.h"

namespace {

constexpr char kMagic[8] = {'M', 'T', 'L', 'P', 'S', 'T', 'A', 'T'};

struct Header {
    char magic[8];
    uint32_t version, n_scenes, n, flags, scene, sphere_on;
    float sphere_c[3], sphere_v[3], sphere_r;
};
static_assert(sizeof(Header) == 8 + 6 * 4 + 7 * 4, "state header layout");
static_assert(sizeof(mtlp_params) == 92, "mtlp_params layout");

struct FileCloser {
    void operator()(std::FILE* f) const {
        if (f) std::fclose(f);
    }
};
using File = std::unique_ptr<std::FILE, FileCloser>;

}  // namespace

extern "C" {

mtlp_status mtlp_save_state(mtlp_engine* e, const char* path) {
    MTLP_TRY(check_engine(e));
    if (!path) return fail(MTLP_E_INVALID, "path is NULL");
    Header h{};
    std::memcpy(h.magic, kMagic, 8);
    h.version = MTLP_STATE_VERSION;
    h.n_scenes = (uint32_t)e->scene_params.size();
    h.n = e->n;
    h.flags = e->batch ? 1u : 0u;
    h.scene = e->scene;
    h.sphere_on = e->sphere.on ? 1u : 0u;
    for (int a = 0; a < 3; ++a) {
        h.sphere_c[a] = e->sphere.c[a];
        h.sphere_v[a] = e->sphere.v[a];
    }
    h.sphere_r = e->sphere.r;
    std::vector<float> x(4 * (size_t)e->n), v(4 * (size_t)e->n);
    if (e->n) {
        DeviceGuard g(e->device);
        HIP_TRY(hipDeviceSynchronize());
        HIP_TRY(hipMemcpy(x.data(), e->x, sizeof(float4) * e->n, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(v.data(), e->v, sizeof(float4) * e->n, hipMemcpyDeviceToHost));
    }
    File f(std::fopen(path, "wb"));
    if (!f) return fail(MTLP_E_INVALID, std::string("cannot open ") + path + " for writing");
    bool ok = std::fwrite(&h, sizeof(h), 1, f.get()) == 1;
    for (size_t s = 0; ok && s < e->scene_params.size(); ++s) {
        ok = std::fwrite(&e->scene_params[s], sizeof(mtlp_params), 1, f.get()) == 1 &&
             std::fwrite(&e->scene_counts[s], sizeof(uint32_t), 1, f.get()) == 1;
    }
    if (ok && e->n) {
        ok = std::fwrite(x.data(), sizeof(float), x.size(), f.get()) == x.size() &&
             std::fwrite(v.data(), sizeof(float), v.size(), f.get()) == v.size();
    }
    if (!ok) return fail(MTLP_E_INVALID, std::string("write to ") + path + " failed");
    return MTLP_OK;
}

mtlp_status mtlp_load_state(mtlp_engine* e, const char* path) {
    if (!e) return fail(MTLP_E_INVALID, "engine is NULL");
    if (!path) return fail(MTLP_E_INVALID, "path is NULL");
    File f(std::fopen(path, "rb"));
    if (!f) return fail(MTLP_E_INVALID, std::string("cannot open ") + path);
    Header h{};
    if (std::fread(&h, sizeof(h), 1, f.get()) != 1 || std::memcmp(h.magic, kMagic, 8) != 0)
        return fail(MTLP_E_INVALID, std::string(path) + " is not an mtlp state file");
    if (h.version != MTLP_STATE_VERSION) return fail(MTLP_E_INVALID, "unsupported state file version");
    if (h.n_scenes == 0 || h.n_scenes > (1u << 20) || h.n > MTLP_MAX_PARTICLES)
        return fail(MTLP_E_RANGE, "state file sizes out of range");
    std::vector<mtlp_params> params(h.n_scenes);
    std::vector<uint32_t> counts(h.n_scenes);
    uint64_t total = 0;
    for (uint32_t s = 0; s < h.n_scenes; ++s) {
        if (std::fread(&params[s], sizeof(mtlp_params), 1, f.get()) != 1 ||
            std::fread(&counts[s], sizeof(uint32_t), 1, f.get()) != 1)
            return fail(MTLP_E_INVALID, "truncated state file");
        uint32_t dims[3];
        MTLP_TRY(validate_params(&params[s], dims));
        total += counts[s];
    }
    if (total != h.n) return fail(MTLP_E_INVALID, "state file counts do not add up");
    std::vector<float> x(4 * (size_t)h.n), v(4 * (size_t)h.n);
    if (h.n && (std::fread(x.data(), sizeof(float), x.size(), f.get()) != x.size() ||
                std::fread(v.data(), sizeof(float), v.size(), f.get()) != v.size()))
        return fail(MTLP_E_INVALID, "truncated state file");
    if (e->batch) {
        if (h.n_scenes != e->scene_params.size() || counts != e->scene_counts)
            return fail(MTLP_E_RANGE, "a batch loads only a file with its own scene counts");
        for (uint32_t s = 0; s < h.n_scenes; ++s) MTLP_TRY(mtlp_batch_set_params(e, s, &params[s]));
    } else {
        if (h.n_scenes != 1) return fail(MTLP_E_RANGE, "a single-scene engine loads a one-scene file");
        if (!e->allocated || e->n != h.n) {
            e->scene_params = {params[0]};
            MTLP_TRY(mtlp_reset(e, h.n, MTLP_SCENE_EMPTY, 0));
        } else {
            MTLP_TRY(mtlp_set_params(e, &params[0]));
        }
    }
    if (h.n) MTLP_TRY(mtlp_upload_state(e, x.data(), v.data(), h.n));
    e->scene = h.scene;
    e->sphere.on = h.sphere_on != 0 && !e->batch;
    for (int a = 0; a < 3; ++a) {
        e->sphere.c[a] = h.sphere_c[a];
        e->sphere.v[a] = h.sphere_v[a];
    }
    e->sphere.r = h.sphere_r;
    return MTLP_OK;
}

}  // extern "C"
