// scenes.cpp -- host-side scene seeding and scene parameter presets.
//
// Scenes 0-2 restate Engine::seedParticles (engine/src/Engine.cpp:454-549):
// lattice fills at spacing h/2 in y, z, x order that stop at N, with the
// reference's FP32 arithmetic.  Scenes 3-4 are the benchmark workloads
// (a jittered 32^3 cube drop at h = 0.1 and a jittered dam break at
// h = 0.02); their box is scaled so the lattice holds N particles and their
// rest density is rho0 (0.1 / h)^3 (particle mass (h / 0.1)^3), which leaves
// every formula of the solver unchanged.  Jitter comes from PCG32
// (O'Neill 2014, pcg32_srandom_r / pcg32_random_r), whose stream is the
// same on every platform.
#include "mtlp_host.h"

#include <cmath>
#include <cstring>

namespace mtlp {

namespace {

struct Pcg32 {
    uint64_t state = 0, inc = 0;
    Pcg32(uint64_t initstate, uint64_t initseq) {
        inc = (initseq << 1u) | 1u;
        next();
        state += initstate;
        next();
    }
    uint32_t next() {
        const uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        const uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        const uint32_t rot = (uint32_t)(old >> 59u);
        return (xorshifted >> rot) | (xorshifted << ((-rot) & 31u));
    }
    // Uniform in [0, 1) with 24 random bits.
    float uniform() { return (float)(next() >> 8) * (1.0f / 16777216.0f); }
};

// Stream id of the jitter generator (the "initseq" of pcg32_srandom_r).
constexpr uint64_t kJitterStream = 54u;

void put(float* x4, uint32_t i, float x, float y, float z, float w) {
    x4[4 * (size_t)i + 0] = x;
    x4[4 * (size_t)i + 1] = y;
    x4[4 * (size_t)i + 2] = z;
    x4[4 * (size_t)i + 3] = w;
}

// Engine.cpp:530-548
void seed_cube_drop(uint32_t n, const mtlp_params& p, float* x4) {
    const float spacing = p.h * 0.5f;
    const uint32_t side = (uint32_t)std::ceil(std::cbrt((double)n));
    const float half_blk = 0.5f * (float)(side - 1) * spacing;
    const float top_y = 2.0f;
    uint32_t i = 0;
    for (uint32_t y = 0; y < side && i < n; ++y)
        for (uint32_t z = 0; z < side && i < n; ++z)
            for (uint32_t x = 0; x < side && i < n; ++x, ++i)
                put(x4, i, -half_blk + (float)x * spacing, top_y - (float)y * spacing,
                    -half_blk + (float)z * spacing, 1.0f);
}

// Engine.cpp:463-491
void seed_dam_break(uint32_t n, const mtlp_params& p, float* x4) {
    const float spacing = p.h * 0.5f;
    const float max_height = 5.5f;
    const uint32_t layers = (uint32_t)(max_height / spacing);
    const uint32_t per_layer = (n + layers - 1) / layers;
    const uint32_t side = (uint32_t)std::ceil(std::sqrt((double)per_layer));
    const float base_extent = (float)(side - 1) * spacing;
    const float x_min = p.bounds_min[0] + 0.05f;
    const float y_min = p.bounds_min[1] + 0.05f;
    const float z_start = -base_extent * 0.5f;
    uint32_t i = 0;
    for (uint32_t y = 0; y < layers && i < n; ++y)
        for (uint32_t z = 0; z < side && i < n; ++z)
            for (uint32_t x = 0; x < side && i < n; ++x, ++i)
                put(x4, i, x_min + (float)x * spacing, y_min + (float)y * spacing,
                    z_start + (float)z * spacing, 1.0f);
}

// Engine.cpp:494-528 (the lake; the kinematic sphere is not modelled)
void seed_lake(uint32_t n, const mtlp_params& p, float* x4) {
    const float spacing = p.h * 0.5f;
    const float x_min = p.bounds_min[0] + 0.05f, x_max = p.bounds_max[0] - 0.05f;
    const float z_min = p.bounds_min[2] + 0.05f, z_max = p.bounds_max[2] - 0.05f;
    const float y_min = p.bounds_min[1] + 0.05f;
    const uint32_t nx = (uint32_t)((x_max - x_min) / spacing);
    const uint32_t nz = (uint32_t)((z_max - z_min) / spacing);
    uint32_t i = 0;
    for (uint32_t y = 0;; ++y) {
        const float y_pos = y_min + (float)y * spacing;
        if (y_pos > p.bounds_max[1] - 0.5f || i >= n) break;
        for (uint32_t z = 0; z < nz && i < n; ++z)
            for (uint32_t x = 0; x < nx && i < n; ++x, ++i)
                put(x4, i, x_min + (float)x * spacing, y_pos, z_min + (float)z * spacing, 1.0f);
    }
}

// Jittered lattice block: nx x ny x nz points at origin + (k + off) * spacing,
// filled y, z, x until n; each coordinate gets U(-j, j) from PCG32.
void seed_jittered_block(uint32_t n, uint64_t seed, float spacing, float off, const float origin[3],
                         uint32_t nx, uint32_t nz, float jitter, float* x4) {
    Pcg32 rng(seed, kJitterStream);
    uint32_t i = 0;
    for (uint32_t y = 0; i < n; ++y)
        for (uint32_t z = 0; z < nz && i < n; ++z)
            for (uint32_t x = 0; x < nx && i < n; ++x, ++i) {
                const float jx = (2.0f * rng.uniform() - 1.0f) * jitter;
                const float jy = (2.0f * rng.uniform() - 1.0f) * jitter;
                const float jz = (2.0f * rng.uniform() - 1.0f) * jitter;
                put(x4, i, origin[0] + ((float)x + off) * spacing + jx,
                    origin[1] + ((float)y + off) * spacing + jy,
                    origin[2] + ((float)z + off) * spacing + jz, 1.0f);
            }
}

uint32_t cube_side(uint32_t n) {
    uint32_t side = (uint32_t)std::ceil(std::cbrt((double)n));
    while (side > 1 && (uint64_t)(side - 1) * (side - 1) * (side - 1) >= n) --side;
    while ((uint64_t)side * side * side < n) ++side;
    return side < 2 ? 2 : side;
}

// Dam-break column: a 1 x 1 x 0.5 box whose column spans x in [0, 0.3],
// z over the whole depth, scaled by 2.027 at 1M particles and by
// 2.027 cbrt(N / 1M) otherwise.
float dam_scale(uint32_t n) { return 2.027f * (float)std::cbrt((double)n / 1.0e6); }

}  // namespace

mtlp_status scene_params(uint32_t scene, uint32_t n, mtlp_params* out) {
    default_params(out);
    switch (scene) {
        case MTLP_SCENE_CUBE_DROP:
        case MTLP_SCENE_DAM_BREAK:
        case MTLP_SCENE_SPHERE_DROP:
        case MTLP_SCENE_EMPTY:
            return MTLP_OK;
        case MTLP_SCENE_JITTER_CUBE: {
            // 32^3 cube at spacing 0.05 occupying [0.2 L, 0.6 L]^3 of [0, L]^3,
            // L = (side - 1) * spacing / 0.4 (3.875 at N = 32768).
            const float h = 0.1f;
            const uint32_t side = cube_side(n);
            const float L = (float)(side - 1) * (0.5f * h) / 0.4f;
            for (int a = 0; a < 3; ++a) {
                out->bounds_min[a] = 0.0f;
                out->bounds_max[a] = L;
            }
            out->gravity[1] = -9.8f;
            out->h = h;
            out->cell_size = h;
            out->rest_density = 6378.0f;
            out->epsilon = 600.0f;
            out->scorr_k = 0.1f;
            out->scorr_n = 4;
            out->scorr_dq = 0.2f;
            out->iterations = 3;
            return MTLP_OK;
        }
        case MTLP_SCENE_JITTER_DAM: {
            const float h = 0.02f;
            const float S = dam_scale(n);
            out->bounds_min[0] = out->bounds_min[1] = out->bounds_min[2] = 0.0f;
            out->bounds_max[0] = S;
            out->bounds_max[1] = S;
            out->bounds_max[2] = 0.5f * S;
            out->gravity[1] = -9.8f;
            out->h = h;
            out->cell_size = h;
            const float m = 0.1f / h;
            out->rest_density = 6378.0f * m * m * m;
            out->epsilon = 600.0f;
            out->scorr_k = 0.1f;
            out->scorr_n = 4;
            out->scorr_dq = 0.2f;
            out->iterations = 4;
            return MTLP_OK;
        }
        default:
            return MTLP_E_INVALID;
    }
}

mtlp_status generate_scene(uint32_t scene, uint32_t n, uint64_t seed, const mtlp_params& p,
                           float* x4, float* v4) {
    std::memset(x4, 0, sizeof(float) * 4 * (size_t)n);
    std::memset(v4, 0, sizeof(float) * 4 * (size_t)n);
    if (n == 0) return MTLP_OK;
    switch (scene) {
        case MTLP_SCENE_CUBE_DROP:
            seed_cube_drop(n, p, x4);
            return MTLP_OK;
        case MTLP_SCENE_DAM_BREAK:
            seed_dam_break(n, p, x4);
            return MTLP_OK;
        case MTLP_SCENE_SPHERE_DROP:
            seed_lake(n, p, x4);
            return MTLP_OK;
        case MTLP_SCENE_EMPTY:
            return MTLP_OK;
        case MTLP_SCENE_JITTER_CUBE: {
            const float spacing = 0.5f * p.h;
            const uint32_t side = cube_side(n);
            const float L = p.bounds_max[0] - p.bounds_min[0];
            const float c0 = p.bounds_min[0] + 0.2f * L;
            const float origin[3] = {c0, p.bounds_min[1] + 0.2f * L, p.bounds_min[2] + 0.2f * L};
            seed_jittered_block(n, seed, spacing, 0.0f, origin, side, side, 0.01f * p.h, x4);
            return MTLP_OK;
        }
        case MTLP_SCENE_JITTER_DAM: {
            const float spacing = 0.5f * p.h;
            const float wx = p.bounds_max[0] - p.bounds_min[0];
            const float wz = p.bounds_max[2] - p.bounds_min[2];
            uint32_t nx = (uint32_t)(0.3f * wx / spacing);
            uint32_t nz = (uint32_t)(wz / spacing);
            if (nx == 0) nx = 1;
            if (nz == 0) nz = 1;
            seed_jittered_block(n, seed, spacing, 0.5f, p.bounds_min, nx, nz, 0.01f * p.h, x4);
            return MTLP_OK;
        }
        default:
            return MTLP_E_INVALID;
    }
}

void default_params(mtlp_params* p) {
    // Engine.cpp:23-39, :558; Constraints.metal:158-160, 190, 325, 340-347
    std::memset(p, 0, sizeof(*p));
    p->gravity[1] = -9.81f;
    p->bounds_min[0] = -3.0f; p->bounds_min[1] = -3.0f; p->bounds_min[2] = -3.0f;
    p->bounds_max[0] = 3.0f;  p->bounds_max[1] = 6.0f;  p->bounds_max[2] = 3.0f;
    p->h = 0.10f;
    p->cell_size = 0.10f;
    p->rest_density = 8000.0f;
    p->epsilon = 800.0f;
    p->scorr_k = 1e-4f;
    p->scorr_dq = 0.2f;
    p->max_step = 0.3f;
    p->damping = 0.996f;
    p->wall_eps = 1e-3f;
    p->floor_friction = 0.55f;
    p->wall_friction = 0.97f;
    p->vmax = 1.0f;
    p->scorr_n = 4;
    p->iterations = 5;
}

}  // namespace mtlp
