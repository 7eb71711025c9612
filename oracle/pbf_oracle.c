/*
 * pbf_oracle.c -- CPU oracle for the PBF step (test infrastructure only;
 * see pbf_oracle.h for what it restates and what pins it).
 *
 * Build: oracle/build.py (gcc -O2 -fopenmp -ffp-contract=off).  The
 * -ffp-contract=off is part of the contract: predict and the cell hash are
 * defined as separately rounded FP32 operations.
 */
#include "pbf_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

void orc_default_params(orc_params* p)
{
    /* Engine.cpp:23-39, writePBFParams :558, Constraints.metal:158-160,190,
     * 325, 340-347 */
    p->gravity[0] = 0.0f; p->gravity[1] = -9.81f; p->gravity[2] = 0.0f;
    p->bounds_min[0] = -3.0f; p->bounds_min[1] = -3.0f; p->bounds_min[2] = -3.0f;
    p->bounds_max[0] =  3.0f; p->bounds_max[1] =  6.0f; p->bounds_max[2] =  3.0f;
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

void orc_grid_make(const orc_params* p, orc_grid* g)
{
    /* Engine.cpp:398-403, 443-450: FP32 extent / cell, ceil per axis. */
    for (int a = 0; a < 3; ++a) {
        const float extent = p->bounds_max[a] - p->bounds_min[a];
        g->origin[a] = p->bounds_min[a];
        g->dim[a] = (uint32_t)ceilf(extent / p->cell_size);
    }
    g->cell_size = p->cell_size;
    g->inv_cell = 1.0f / p->cell_size;
    g->total_cells = g->dim[0] * g->dim[1] * g->dim[2];
}

/* cellCoordOf (Spatial.metal:35-39): q = (p - origin) * invCell in FP32,
 * floor, clamp to [0, dim-1].  q is first limited to [-1, dim] so the
 * float->int conversion is defined for every input (NaN -> -1, +-inf ->
 * the limits); that changes no result of the final clamp. */
static void cell_coord_f32(const float p[3], const orc_grid* g, int c[3])
{
    for (int a = 0; a < 3; ++a) {
        float q = (p[a] - g->origin[a]) * g->inv_cell;
        if (!(q >= -1.0f)) q = -1.0f;
        if (q > (float)g->dim[a]) q = (float)g->dim[a];
        int ci = (int)floorf(q);
        if (ci < 0) ci = 0;
        if (ci > (int)g->dim[a] - 1) ci = (int)g->dim[a] - 1;
        c[a] = ci;
    }
}

/* cellHashOf (Spatial.metal:41-43) */
/* The neighbour test of the FP32 stencil walks and countNeighbors: the
 * reference's dot(r, r) < h^2 (Spatial.metal:235-236, Constraints.metal:
 * 112-113; Metal's fast math lets dot contract into fused multiply-adds) as
 * ((r0 r0 - h^2) + r1 r1) + r2 r2, one correctly rounded fmaf per step; a
 * negative result is a neighbour.  (Built with -mfma, fmaf is the hardware
 * instruction.) */
static float nbr_key_f32(float r0, float r1, float r2, float h2)
{
    return fmaf(r2, r2, fmaf(r1, r1, fmaf(r0, r0, -h2)));
}

static uint32_t key_of(const int c[3], const orc_grid* g)
{
    return (uint32_t)c[0] + (uint32_t)c[1] * g->dim[0] + (uint32_t)c[2] * g->dim[0] * g->dim[1];
}

uint32_t orc_cell_key(const float p[3], const orc_grid* g)
{
    int c[3];
    cell_coord_f32(p, g, c);
    return key_of(c, g);
}

/* predictPositions (Constraints.metal:54-66): v' = v + g dt; x* = x + v' dt,
 * w kept.  Separately rounded FP32 operations. */
void orc_predict_f32(const float* x4, const float* v4, uint32_t n,
                     const float gravity[3], float dt, float* xs4)
{
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        const size_t i = (size_t)ii;
        for (int a = 0; a < 3; ++a) {
            const float gdt = gravity[a] * dt;
            const float vn = v4[4 * i + a] + gdt;
            const float step = vn * dt;
            xs4[4 * i + a] = x4[4 * i + a] + step;
        }
        xs4[4 * i + 3] = x4[4 * i + 3];
    }
}

/* hashCells (Spatial.metal:49-57) */
void orc_hash(const float* p4, uint32_t n, const orc_grid* g, uint32_t* keys)
{
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii)
        keys[ii] = orc_cell_key(p4 + 4 * (size_t)ii, g);
}

/* countCells (Spatial.metal:63-70) after the zero fill (Engine.cpp:606).
 * counts[total_cells] stays 0, so the exclusive scan ends with N. */
void orc_count(const uint32_t* keys, uint32_t n, uint32_t total_cells, uint32_t* counts)
{
    memset(counts, 0, sizeof(uint32_t) * ((size_t)total_cells + 1));
    for (uint32_t i = 0; i < n; ++i) counts[keys[i]]++;
}

/* Exclusive scan (Spatial.metal:82-161; cpuExclusiveScan, TestUtils.hpp:52-61).
 * u32 wrap-around arithmetic, like the GPU. */
void orc_exclusive_scan(const uint32_t* in, uint32_t* out, uint32_t n)
{
    uint32_t run = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t v = in[i];
        out[i] = run;
        run += v;
    }
}

/* scatterParticles (Spatial.metal:169-180) with the within-cell order
 * fixed to increasing original index: a serial counting sort. */
void orc_stable_scatter(const uint32_t* keys, uint32_t n, const uint32_t* cell_start,
                        uint32_t* sorted_index)
{
    uint32_t maxk = 0;
    for (uint32_t i = 0; i < n; ++i) if (keys[i] > maxk) maxk = keys[i];
    uint32_t* cursor = (uint32_t*)calloc((size_t)maxk + 1, sizeof(uint32_t));
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t c = keys[i];
        sorted_index[cell_start[c] + cursor[c]++] = i;
    }
    free(cursor);
}

/* buildSpatialHash (Engine.cpp:601-679): hash, count, scan, scatter,
 * gatherPositions (Spatial.metal:190-198), countNeighbors
 * (Spatial.metal:207-243), all on p4 as given. */
void orc_build_spatial_hash(const float* p4, uint32_t n, const orc_grid* g,
                            uint32_t* keys, uint32_t* cell_start,
                            uint32_t* sorted_index, float* sorted4,
                            float r2max, uint32_t* neighbor_counts)
{
    const uint32_t C = g->total_cells;
    uint32_t* cnt = (uint32_t*)malloc(sizeof(uint32_t) * ((size_t)C + 1));
    orc_hash(p4, n, g, keys);
    orc_count(keys, n, C, cnt);
    orc_exclusive_scan(cnt, cell_start, C + 1);
    orc_stable_scatter(keys, n, cell_start, sorted_index);
    for (uint32_t s = 0; s < n; ++s)
        memcpy(sorted4 + 4 * (size_t)s, p4 + 4 * (size_t)sorted_index[s], 4 * sizeof(float));
    free(cnt);
    if (!neighbor_counts) return;
    /* r2max = neighborRadiusSq = kSmoothingH^2 (Engine.cpp:447) */
    const uint32_t dx = g->dim[0], dxy = g->dim[0] * g->dim[1];
#pragma omp parallel for schedule(static)
    for (int64_t gi = 0; gi < (int64_t)n; ++gi) {
        const uint32_t gid = (uint32_t)gi;
        const float* p = p4 + 4 * (size_t)gid;
        int c[3], lo[3], hi[3];
        cell_coord_f32(p, g, c);
        for (int a = 0; a < 3; ++a) {
            lo[a] = c[a] - 1 < 0 ? 0 : c[a] - 1;
            hi[a] = c[a] + 1 > (int)g->dim[a] - 1 ? (int)g->dim[a] - 1 : c[a] + 1;
        }
        uint32_t count = 0;
        for (int z = lo[2]; z <= hi[2]; ++z)
            for (int y = lo[1]; y <= hi[1]; ++y)
                for (int x = lo[0]; x <= hi[0]; ++x) {
                    const uint32_t cc = (uint32_t)x + (uint32_t)y * dx + (uint32_t)z * dxy;
                    for (uint32_t i = cell_start[cc]; i < cell_start[cc + 1]; ++i) {
                        if (sorted_index[i] == gid) continue;
                        const float* q = sorted4 + 4 * (size_t)i;
                        const float d0 = q[0] - p[0], d1 = q[1] - p[1], d2 = q[2] - p[2];
                        if (nbr_key_f32(d0, d1, d2, r2max) < 0.0f) ++count;
                    }
                }
        neighbor_counts[gid] = count;
    }
}

void orc_sphere_advance(float state[6], const orc_params* p, float dt)
{
    /* Engine.cpp:693-712 (bounds: the scene's, kBoundsMin/Max for scene 2) */
    float* c = state;
    float* v = state + 3;
    const float r = p->sphere[3];
    v[1] += -9.81f * dt;
    for (int a = 0; a < 3; ++a) v[a] *= 0.998f;
    for (int a = 0; a < 3; ++a) c[a] += v[a] * dt;
    for (int a = 0; a < 3; ++a) {
        if (c[a] - r < p->bounds_min[a]) { c[a] = p->bounds_min[a] + r; v[a] = -v[a] * 0.4f; }
        if (c[a] + r > p->bounds_max[a]) { c[a] = p->bounds_max[a] - r; v[a] = -v[a] * 0.4f; }
    }
}

void orc_foam_f64(const double* v4, const uint32_t* nbr, uint32_t n, double max_nbr, double* foam)
{
    /* Constraints.metal:287-299: (1 - clamp(nbr / max, 0, 1)) * smoothstep(1.2, 4.5, |v|),
     * smoothstep(e0, e1, x) = t^2 (3 - 2t), t = clamp((x - e0) / (e1 - e0), 0, 1) */
    for (uint32_t i = 0; i < n; ++i) {
        const double* v = v4 + 4 * (size_t)i;
        const double speed = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        double surf = (double)nbr[i] / max_nbr;
        surf = 1.0 - (surf < 0 ? 0 : (surf > 1 ? 1 : surf));
        double t = (speed - 1.2) / (4.5 - 1.2);
        t = t < 0 ? 0 : (t > 1 ? 1 : t);
        foam[i] = surf * t * t * (3.0 - 2.0 * t);
    }
}

void orc_mouse_pulse_f64(const double* x4, double* v4, uint32_t n, const double origin[3],
                         const double dir[3], const double force[3], double radius)
{
    /* Constraints.metal:231-253 */
    const double R2 = radius * radius;
    for (uint32_t i = 0; i < n; ++i) {
        const double* x = x4 + 4 * (size_t)i;
        const double to[3] = { x[0] - origin[0], x[1] - origin[1], x[2] - origin[2] };
        const double t = to[0] * dir[0] + to[1] * dir[1] + to[2] * dir[2];
        if (t <= 0.0) continue;
        double r2 = 0;
        for (int a = 0; a < 3; ++a) {
            const double d = x[a] - (origin[a] + dir[a] * t);
            r2 += d * d;
        }
        if (r2 < R2) {
            const double w = exp(-r2 / (R2 * 0.4));
            for (int a = 0; a < 3; ++a) v4[4 * (size_t)i + a] += force[a] * w;
            v4[4 * (size_t)i + 3] = 0.0;
        }
    }
}

void orc_integrate_f64(double* x4, double* v4, uint32_t n, const orc_params* p, double dt)
{
    /* Integrate.metal:11-42 with Engine.cpp:578-595's gravity and damping */
    for (uint32_t i = 0; i < n; ++i) {
        double* x = x4 + 4 * (size_t)i;
        double* v = v4 + 4 * (size_t)i;
        v[1] += -9.81 * dt;
        for (int a = 0; a < 3; ++a) { v[a] *= 0.999; x[a] += v[a] * dt; }
        for (int a = 0; a < 3; ++a)
            if (x[a] < p->bounds_min[a]) { x[a] = p->bounds_min[a]; v[a] = -v[a] * 0.4; }
        for (int a = 0; a < 3; ++a)
            if (x[a] > p->bounds_max[a]) { x[a] = p->bounds_max[a]; v[a] = -v[a] * 0.4; }
        x[3] = 1.0;
        v[3] = 0.0;
    }
}

void orc_set_threads(int threads)
{
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
}

int orc_get_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ---- precision-generic solver ------------------------------------------ */
#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(CAT(name, _), SFX)

/* FP64: the plain test on the exact square (the contract); FP32: the fused
 * chain of the GPU kernels (nbr_key_f32). */
#define REAL double
#define SFX f64
#define SQRT_OF(x) sqrt(x)
#define NOT_NEIGHBOR(r, r2, h2) ((r2) >= (h2))
#include "pbf_solver_impl.h"
#undef REAL
#undef SFX
#undef SQRT_OF
#undef NOT_NEIGHBOR

#define REAL float
#define SFX f32
#define SQRT_OF(x) sqrtf(x)
#define NOT_NEIGHBOR(r, r2, h2) (!(nbr_key_f32((r)[0], (r)[1], (r)[2], (h2)) < 0.0f))
#include "pbf_solver_impl.h"
#undef REAL
#undef SFX
#undef SQRT_OF
#undef NOT_NEIGHBOR

/* ---- FP64 rollout step and adjoint (uses the f64 helpers above) -------- */
#include "pbf_adjoint_impl.h"
