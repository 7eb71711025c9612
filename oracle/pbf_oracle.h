/*
 * pbf_oracle.h -- CPU oracle for the PBF step of rakdcolon/mtlphys.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: tests/, the
 * smoke() entry and bench.py's cpu_baseline leg load it.  The product
 * library (mtlphys_amd/libmtlp.so) never links, loads or calls it.
 *
 * It restates, in plain C, the reference's Metal kernels and the engine
 * orchestration that sequences them:
 *   engine/shaders/Spatial.metal      :35-43, 49-198, 207-243
 *   engine/shaders/Constraints.metal  :29-66, 82-213, 305-353
 *   engine/src/Engine.cpp             :398-452 (grid), :551-572 (derived
 *                                      constants), :601-679 and :685-811
 *                                      (stage order, ping-pong), :843-882
 *                                      (scan contract)
 *   tests/TestUtils.cpp:30-42, tests/TestUtils.hpp:52-61 (CPU hash, scan)
 *
 * Precision contract
 *   - predict, cell hash and the stencil centre are computed in FP32 with
 *     no FMA contraction (built with -ffp-contract=off) in BOTH
 *     instantiations: cell keys, counts, scan offsets and the permutation
 *     are integer results of FP32 inputs and must be bit-exact.
 *   - the solver (density/lambda, delta, finalize) is instantiated twice:
 *     *_f64 (the contract) and *_f32 (for error estimates and the CPU
 *     baseline).  The FP64 solver consumes the FP32 predicted positions
 *     (upcast).  The stencil centre of an FP64 iterate is the FP32 cell of
 *     that iterate rounded to float.
 *   - the within-cell order of the permutation is stable by original
 *     index (the reference's atomic order, Spatial.metal:166-180, is not
 *     defined; the counting sort here fixes it).
 *
 * Parity pins (each function is checked by tests/test_oracle_*.py):
 *   hash              reference known answers (test_spatial.cpp:52-70), face
 *                     and out-of-grid clamp cases
 *   count/scan        numpy cumsum / bincount, reference scan cases
 *                     (test_scan.cpp:59-110)
 *   permutation       numpy argsort(kind="stable")
 *   neighbour sets    O(N^2) brute force at iteration 0; reference cases
 *                     (test_spatial.cpp:102-145)
 *   density/lambda    closed forms: isolated particle, 26-neighbour lattice
 *                     (5.15625 W(0)), pair formula; sign case
 *                     (test_pbf.cpp:67-107)
 *   delta             pair closed form incl. s_corr at r = dq; antisymmetry
 *                     sum_i dx_i = 0; step clamp |dx| = max_step
 *   predict/finalize  free fall (test_pbf.cpp:47-65), per-wall hand cases,
 *                     CFL cap
 *   sphere, foam,     hand cases: a point inside the sphere lands on its
 *   mouse pulse       surface along the centre ray; foam 0 below 1.2 m/s and
 *                     for >= 96 neighbours, smoothstep values between; pulse
 *                     weight 1 on the ray, exp(-2.5) at the rim, 0 behind
 */
#ifndef PBF_ORACLE_H
#define PBF_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Regular grid covering [origin, origin + dim*cell) (Shared.h:37-49). */
typedef struct orc_grid {
    float    origin[3];
    float    cell_size;
    float    inv_cell;      /* 1.0f / cell_size, FP32 (Engine.cpp:446)      */
    uint32_t dim[3];
    uint32_t total_cells;
} orc_grid;

/* Physical parameters.  Defaults are the reference's (Engine.cpp:23-39,
 * Constraints.metal:158-160,190,325,340-347). */
typedef struct orc_params {
    float    gravity[3];
    float    bounds_min[3];
    float    bounds_max[3];
    float    h;
    float    cell_size;
    float    rest_density;
    float    epsilon;
    float    scorr_k;
    float    scorr_dq;       /* delta q as a fraction of h                   */
    float    max_step;       /* |dx| cap as a fraction of h                  */
    float    damping;
    float    wall_eps;
    float    floor_friction;
    float    wall_friction;
    float    vmax;           /* speed cap as a fraction of h/dt              */
    uint32_t scorr_n;
    uint32_t iterations;
    /* sphere-drop scene (SphereParams, Shared.h:78-81): centre xyz, radius;
     * when sphere_on, every delta pass ends with sphereCollision */
    float    sphere[4];
    uint32_t sphere_on;
} orc_params;

void orc_default_params(orc_params* p);
/* Grid rule of Engine::buildSpatialBuffers (Engine.cpp:398-452). */
void orc_grid_make(const orc_params* p, orc_grid* g);

/* ---- integer stages (exact) ------------------------------------------- */
void orc_predict_f32(const float* x4, const float* v4, uint32_t n,
                     const float gravity[3], float dt, float* xs4);
uint32_t orc_cell_key(const float p[3], const orc_grid* g);
void orc_hash(const float* p4, uint32_t n, const orc_grid* g, uint32_t* keys);
void orc_count(const uint32_t* keys, uint32_t n, uint32_t total_cells,
               uint32_t* counts /* total_cells + 1 */);
void orc_exclusive_scan(const uint32_t* in, uint32_t* out, uint32_t n);
void orc_stable_scatter(const uint32_t* keys, uint32_t n,
                        const uint32_t* cell_start, uint32_t* sorted_index);
/* buildSpatialHash incl. countNeighbors (Engine.cpp:601-679) on p4 as given. */
void orc_build_spatial_hash(const float* p4, uint32_t n, const orc_grid* g,
                            uint32_t* keys, uint32_t* cell_start,
                            uint32_t* sorted_index, float* sorted4,
                            float r2max, uint32_t* neighbor_counts);

/* ---- solver, two precisions ------------------------------------------- */
typedef struct orc_trace_f64 {
    uint32_t* keys;          /* n                                            */
    uint32_t* cell_start;    /* total_cells + 1                              */
    uint32_t* sorted_index;  /* n                                            */
    uint32_t* nbr_count;     /* n, original order, iteration-0 stencil       */
    float*    xstar4;        /* 4n, predicted positions, original order      */
    double*   lambda_iter;   /* iterations * n, sorted order                 */
    double*   sx_iter;       /* (iterations + 1) * 4n, sorted positions      */
    double*   lambda_orig;   /* n, last iteration, original order            */
} orc_trace_f64;

typedef struct orc_trace_f32 {
    uint32_t* keys;
    uint32_t* cell_start;
    uint32_t* sorted_index;
    uint32_t* nbr_count;
    float*    xstar4;
    float*    lambda_iter;
    float*    sx_iter;
    float*    lambda_orig;
} orc_trace_f32;

void orc_density_lambda_f64(const double* sx4, const uint32_t* cell_start,
                            const uint32_t* sorted_index, uint32_t n,
                            const orc_grid* g, const orc_params* p,
                            double* lambda_sorted, double* lambda_orig,
                            uint32_t* nbr_orig);
void orc_density_lambda_f32(const float* sx4, const uint32_t* cell_start,
                            const uint32_t* sorted_index, uint32_t n,
                            const orc_grid* g, const orc_params* p,
                            float* lambda_sorted, float* lambda_orig,
                            uint32_t* nbr_orig);
void orc_apply_delta_f64(const double* sx4, const double* lambda_sorted,
                         const uint32_t* cell_start, uint32_t n,
                         const orc_grid* g, const orc_params* p,
                         double* sx4_next);
void orc_apply_delta_f32(const float* sx4, const float* lambda_sorted,
                         const uint32_t* cell_start, uint32_t n,
                         const orc_grid* g, const orc_params* p,
                         float* sx4_next);
/* finalize on original-order positions x4 (after the scatter back) */
void orc_finalize_f64(double* x4, double* v4, const double* prev4, uint32_t n,
                      const orc_params* p, float dt);
void orc_finalize_f32(float* x4, float* v4, const float* prev4, uint32_t n,
                      const orc_params* p, float dt);

/* Engine::step (Engine.cpp:685-811) without the foam / sphere stages.
 * Inputs are the engine state (FP32 float4 AoS).  dt <= 0 is a no-op that
 * copies the inputs through (Engine.cpp:686).  Returns 0. */
int orc_step_f64(const float* x4, const float* v4, uint32_t n,
                 const orc_params* p, float dt,
                 double* x4_out, double* v4_out, orc_trace_f64* trace);
int orc_step_f32(const float* x4, const float* v4, uint32_t n,
                 const orc_params* p, float dt,
                 float* x4_out, float* v4_out, orc_trace_f32* trace);

/* ---- FP64 rollout step and its adjoint (pbf_adjoint_impl.h) ----------- */
/* What the adjoint of one step needs from its forward pass. */
typedef struct orc_tape {
    uint32_t* sorted_index;  /* n                                            */
    uint32_t* cell_start;    /* total_cells + 1                              */
    double*   p;             /* (iterations + 1) * 4n, sorted iterates       */
    double*   x_in;          /* 4n, positions at the start of the step       */
} orc_tape;

/* The step in FP64 throughout (predict included), binned at the FP32
 * rounding of the FP64 prediction.  dt must be > 0. */
int orc_step_d(const double* x4, const double* v4, uint32_t n, const orc_params* p, float dt,
               double* x4_out, double* v4_out, orc_tape* tape);
/* Vector-Jacobian product of orc_step_d at its tape: given the adjoints of
 * (x_out, v_out), the adjoints of (x_in, v_in) and of gravity.  sums: NULL
 * (gravity from the direct sum) or the 6 running sums of the adjoints, see
 * pbf_adjoint_impl.h. */
int orc_step_vjp_d(const orc_tape* tape, uint32_t n, const orc_params* p, float dt,
                   const double* xbar_out, const double* vbar_out,
                   double* xbar_in, double* vbar_in, double gbar[3], double* sums);
/* Branch signature (bins, stencil cells, clamps, caps, wall flags, gates)
 * of a taped step per particle: (2 + iterations) words each. */
void orc_step_kinks_d(const orc_tape* tape, uint32_t n, const orc_params* p, float dt, uint64_t* out);
/* rho at x (fresh binning of x), original order, and its VJP. */
void orc_density_at_d(const double* x4, uint32_t n, const orc_params* p, double* rho);
void orc_density_vjp_d(const double* x4, uint32_t n, const orc_params* p,
                       const double* rhobar, double* xbar);

/* ---- scene extras ------------------------------------------------------ */
/* Sphere kinematics of one step (Engine.cpp:693-712), FP32 like the
 * reference: state = centre xyz, velocity xyz; radius and bounds from p. */
void orc_sphere_advance(float state[6], const orc_params* p, float dt);
/* computeFoam (Constraints.metal:287-299) from velocities and neighbour
 * counts, original order; max_nbr = kMaxNeighborsForViz (Engine.cpp:27). */
void orc_foam_f64(const double* v4, const uint32_t* nbr, uint32_t n, double max_nbr, double* foam);
/* applyMousePulse (Constraints.metal:222-253) on v4, in place. */
void orc_mouse_pulse_f64(const double* x4, double* v4, uint32_t n, const double origin[3],
                         const double dir[3], const double force[3], double radius);

/* integrate (Integrate.metal:11-42, Engine.cpp:578-595) in FP64, in place. */
void orc_integrate_f64(double* x4, double* v4, uint32_t n, const orc_params* p, double dt);

/* OpenMP thread count for the per-particle stages (0 = library default). */
void orc_set_threads(int threads);
int  orc_get_threads(void);

#ifdef __cplusplus
}
#endif

#endif /* PBF_ORACLE_H */
