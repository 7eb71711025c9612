/*
 * pbf_solver_impl.h -- precision-generic part of the oracle (test
 * infrastructure only; see pbf_oracle.h).  Included twice by pbf_oracle.c
 * with REAL = double / float and SFX = f64 / f32.
 *
 * Needs from the includer: REAL, SFX, SQRT_OF(x) = sqrt in REAL,
 * NOT_NEIGHBOR(r, r2, h2) = the neighbour test failing for the pair
 * difference r (|r|^2 = r2), cell_coord_f32(), key_of(), and the CAT/FN
 * macros.
 */

/* Derived constants of writePBFParams (Engine.cpp:551-572), in REAL. */
typedef struct FN(consts) {
    REAL h, h2, poly6_norm, spiky_norm, inv_rho0, eps;
    REAL w0;            /* poly6(0) = poly6Norm * h2^3                      */
    REAL inv_wq;        /* 1 / poly6((dq h)^2), Constraints.metal:159-160  */
    REAL scorr_k;
    uint32_t scorr_n;
    REAL max_step;      /* max_step * h, Constraints.metal:190             */
    REAL bmin[3], bmax[3];
} FN(consts);

static REAL FN(poly6)(REAL r2, REAL h2, REAL norm)
{   /* Constraints.metal:29-33 */
    if (r2 >= h2) return (REAL)0;
    const REAL diff = h2 - r2;
    return norm * diff * diff * diff;
}

/* Constraints.metal:35-39: -norm (h - r)^2 r_vec / r for 1e-6 < r < h. */
static void FN(spiky_grad)(const REAL r[3], REAL rlen, REAL h, REAL norm,
                           REAL out[3])
{
    if (rlen <= (REAL)1e-6f || rlen >= h) { out[0] = out[1] = out[2] = 0; return; }
    const REAL diff = h - rlen;
    const REAL s = -norm * diff * diff;
    out[0] = s * (r[0] / rlen);
    out[1] = s * (r[1] / rlen);
    out[2] = s * (r[2] / rlen);
}

/* x^n by squaring; for n = 4 this is (x*x)*(x*x), the reference's
 * ratio2 * ratio2 (Constraints.metal:178-180). */
static REAL FN(ipow)(REAL x, uint32_t n)
{
    REAL result = 1, b = x;
    while (n) {
        if (n & 1u) result *= b;
        n >>= 1;
        if (n) b *= b;
    }
    return result;
}

static void FN(make_consts)(const orc_params* p, FN(consts)* c)
{
    const REAL h  = (REAL)p->h;
    const REAL h2 = h * h;
    const REAL h3 = h2 * h;
    const REAL h6 = h3 * h3;
    const REAL h9 = h6 * h3;
    const REAL pi = (REAL)3.14159265358979323846;
    c->h = h; c->h2 = h2;
    c->poly6_norm = (REAL)315 / ((REAL)64 * pi * h9);
    c->spiky_norm = (REAL)45 / (pi * h6);
    c->inv_rho0   = (REAL)1 / (REAL)p->rest_density;
    c->eps        = (REAL)p->epsilon;
    c->w0         = FN(poly6)((REAL)0, h2, c->poly6_norm);
    const REAL dqh = (REAL)p->scorr_dq * h;
    c->inv_wq     = (REAL)1 / FN(poly6)(dqh * dqh, h2, c->poly6_norm);
    c->scorr_k    = (REAL)p->scorr_k;
    c->scorr_n    = p->scorr_n;
    c->max_step   = (REAL)p->max_step * h;
    for (int a = 0; a < 3; ++a) {
        c->bmin[a] = (REAL)p->bounds_min[a];
        c->bmax[a] = (REAL)p->bounds_max[a];
    }
}

/* Stencil bounds around the FP32 cell of p (Constraints.metal:94-98). */
static void FN(stencil)(const REAL* p, const orc_grid* g, int lo[3], int hi[3])
{
    const float pf[3] = { (float)p[0], (float)p[1], (float)p[2] };
    int c[3];
    cell_coord_f32(pf, g, c);
    for (int a = 0; a < 3; ++a) {
        lo[a] = c[a] - 1 < 0 ? 0 : c[a] - 1;
        hi[a] = c[a] + 1 > (int)g->dim[a] - 1 ? (int)g->dim[a] - 1 : c[a] + 1;
    }
}

static REAL FN(clampr)(REAL v, REAL lo, REAL hi)
{
    return v < lo ? lo : (v > hi ? hi : v);
}

/* densityLambdaSorted, Constraints.metal:82-131.  Optionally also the
 * neighbour count of countNeighbors (Spatial.metal:207-243), which on the
 * step path sees the same positions, stencil and radius as iteration 0. */
void CAT(orc_density_lambda_, SFX)(const REAL* sx4, const uint32_t* cell_start,
                                  const uint32_t* sorted_index, uint32_t n,
                                  const orc_grid* g, const orc_params* p,
                                  REAL* lambda_sorted, REAL* lambda_orig,
                                  uint32_t* nbr_orig)
{
    FN(consts) k;
    FN(make_consts)(p, &k);
    const uint32_t dx = g->dim[0], dxy = g->dim[0] * g->dim[1];
#pragma omp parallel for schedule(static)
    for (int64_t si = 0; si < (int64_t)n; ++si) {
        const uint32_t s = (uint32_t)si;
        const REAL* pi = sx4 + 4 * (size_t)s;
        int lo[3], hi[3];
        FN(stencil)(pi, g, lo, hi);
        REAL density = k.w0;
        REAL gs[3] = { 0, 0, 0 };
        REAL gsq = 0;
        uint32_t cnt = 0;
        for (int z = lo[2]; z <= hi[2]; ++z)
            for (int y = lo[1]; y <= hi[1]; ++y)
                for (int x = lo[0]; x <= hi[0]; ++x) {
                    const uint32_t c = (uint32_t)x + (uint32_t)y * dx + (uint32_t)z * dxy;
                    for (uint32_t j = cell_start[c]; j < cell_start[c + 1]; ++j) {
                        if (j == s) continue;
                        const REAL* pj = sx4 + 4 * (size_t)j;
                        const REAL r[3] = { pi[0] - pj[0], pi[1] - pj[1], pi[2] - pj[2] };
                        const REAL r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
                        if (NOT_NEIGHBOR(r, r2, k.h2)) continue;
                        ++cnt;
                        density += FN(poly6)(r2, k.h2, k.poly6_norm);
                        REAL gr[3];
                        FN(spiky_grad)(r, SQRT_OF(r2), k.h, k.spiky_norm, gr);
                        gs[0] += gr[0]; gs[1] += gr[1]; gs[2] += gr[2];
                        gsq += gr[0] * gr[0] + gr[1] * gr[1] + gr[2] * gr[2];
                    }
                }
        const REAL constraint = density * k.inv_rho0 - (REAL)1;
        const REAL denom = k.inv_rho0 * k.inv_rho0 *
                           ((gs[0] * gs[0] + gs[1] * gs[1] + gs[2] * gs[2]) + gsq) + k.eps;
        const REAL lambda = -constraint / denom;
        if (lambda_sorted) lambda_sorted[s] = lambda;
        if (lambda_orig) lambda_orig[sorted_index[s]] = lambda;
        if (nbr_orig) nbr_orig[sorted_index[s]] = cnt;
    }
}

/* applyDeltaSorted, Constraints.metal:140-199, then sphereCollision when
 * the sphere is on (Engine.cpp:760-771). */
void CAT(orc_apply_delta_, SFX)(const REAL* sx4, const REAL* lambda_sorted,
                               const uint32_t* cell_start, uint32_t n,
                               const orc_grid* g, const orc_params* p,
                               REAL* sx4_next)
{
    FN(consts) k;
    FN(make_consts)(p, &k);
    const uint32_t dx = g->dim[0], dxy = g->dim[0] * g->dim[1];
#pragma omp parallel for schedule(static)
    for (int64_t si = 0; si < (int64_t)n; ++si) {
        const uint32_t s = (uint32_t)si;
        const REAL* pi = sx4 + 4 * (size_t)s;
        const REAL lam_i = lambda_sorted[s];
        int lo[3], hi[3];
        FN(stencil)(pi, g, lo, hi);
        REAL d[3] = { 0, 0, 0 };
        for (int z = lo[2]; z <= hi[2]; ++z)
            for (int y = lo[1]; y <= hi[1]; ++y)
                for (int x = lo[0]; x <= hi[0]; ++x) {
                    const uint32_t c = (uint32_t)x + (uint32_t)y * dx + (uint32_t)z * dxy;
                    for (uint32_t j = cell_start[c]; j < cell_start[c + 1]; ++j) {
                        if (j == s) continue;
                        const REAL* pj = sx4 + 4 * (size_t)j;
                        const REAL r[3] = { pi[0] - pj[0], pi[1] - pj[1], pi[2] - pj[2] };
                        const REAL r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
                        if (NOT_NEIGHBOR(r, r2, k.h2)) continue;
                        const REAL lam_j = lambda_sorted[j];
                        REAL gr[3];
                        FN(spiky_grad)(r, SQRT_OF(r2), k.h, k.spiky_norm, gr);
                        const REAL wij = FN(poly6)(r2, k.h2, k.poly6_norm);
                        const REAL ratio = wij * k.inv_wq;
                        const REAL scorr = -k.scorr_k * FN(ipow)(ratio, k.scorr_n);
                        const REAL coef = lam_i + lam_j + scorr;
                        d[0] += coef * gr[0]; d[1] += coef * gr[1]; d[2] += coef * gr[2];
                    }
                }
        d[0] *= k.inv_rho0; d[1] *= k.inv_rho0; d[2] *= k.inv_rho0;
        const REAL dlen = SQRT_OF(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        if (dlen > k.max_step) {
            const REAL sc = k.max_step / dlen;
            d[0] *= sc; d[1] *= sc; d[2] *= sc;
        }
        REAL* out = sx4_next + 4 * (size_t)s;
        for (int a = 0; a < 3; ++a) out[a] = FN(clampr)(pi[a] + d[a], k.bmin[a], k.bmax[a]);
        out[3] = (REAL)1;
        if (p->sphere_on) {
            /* sphereCollision (Constraints.metal:262-277) on the new iterate */
            REAL q[3], r2 = 0;
            for (int a = 0; a < 3; ++a) { q[a] = out[a] - (REAL)p->sphere[a]; r2 += q[a] * q[a]; }
            const REAL rr = (REAL)p->sphere[3] * (REAL)p->sphere[3];
            if (r2 < rr) {
                const REAL r = SQRT_OF(r2 > (REAL)1e-8f ? r2 : (REAL)1e-8f);
                for (int a = 0; a < 3; ++a) out[a] = (REAL)p->sphere[a] + q[a] / r * (REAL)p->sphere[3];
            }
        }
    }
}

/* finalizeStep, Constraints.metal:305-353, on original-order positions. */
void CAT(orc_finalize_, SFX)(REAL* x4, REAL* v4, const REAL* prev4, uint32_t n,
                            const orc_params* p, float dt)
{
    FN(consts) k;
    FN(make_consts)(p, &k);
    const REAL inv_dt = (REAL)1 / (REAL)dt;
    const REAL damping = (REAL)p->damping;
    const REAL weps = (REAL)p->wall_eps;
    const REAL ff = (REAL)p->floor_friction, wf = (REAL)p->wall_friction;
    const REAL vmax = (REAL)p->vmax * k.h * inv_dt;
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        const size_t i = (size_t)ii;
        REAL pp[3], v[3];
        for (int a = 0; a < 3; ++a) {
            pp[a] = FN(clampr)(x4[4 * i + a], k.bmin[a], k.bmax[a]);
            v[a] = (pp[a] - prev4[4 * i + a]) * inv_dt;
            v[a] *= damping;
        }
        const int xmin = pp[0] <= k.bmin[0] + weps, xmax = pp[0] >= k.bmax[0] - weps;
        const int ymin = pp[1] <= k.bmin[1] + weps, ymax = pp[1] >= k.bmax[1] - weps;
        const int zmin = pp[2] <= k.bmin[2] + weps, zmax = pp[2] >= k.bmax[2] - weps;
        if (xmin && v[0] < 0) v[0] = 0;
        if (xmax && v[0] > 0) v[0] = 0;
        if (ymin && v[1] < 0) v[1] = 0;
        if (ymax && v[1] > 0) v[1] = 0;
        if (zmin && v[2] < 0) v[2] = 0;
        if (zmax && v[2] > 0) v[2] = 0;
        if (ymin) {
            v[0] *= ff; v[2] *= ff;
        } else if (xmin || xmax || zmin || zmax) {
            v[0] *= wf; v[1] *= wf; v[2] *= wf;
        }
        const REAL vlen = SQRT_OF(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        if (vlen > vmax) {
            const REAL sc = vmax / vlen;
            v[0] *= sc; v[1] *= sc; v[2] *= sc;
        }
        x4[4 * i + 0] = pp[0]; x4[4 * i + 1] = pp[1]; x4[4 * i + 2] = pp[2]; x4[4 * i + 3] = 1;
        v4[4 * i + 0] = v[0];  v4[4 * i + 1] = v[1];  v4[4 * i + 2] = v[2];  v4[4 * i + 3] = 0;
    }
}

/* Engine::step (Engine.cpp:685-811): predict -> hash -> I x (lambda, delta,
 * ping-pong) -> scatter back -> finalize.  Foam and the sphere are out. */
int CAT(orc_step_, SFX)(const float* x4, const float* v4, uint32_t n,
                        const orc_params* p, float dt,
                        REAL* x4_out, REAL* v4_out, CAT(orc_trace_, SFX)* tr)
{
    if (!(dt > 0.0f)) {
        for (size_t i = 0; i < 4 * (size_t)n; ++i) { x4_out[i] = x4[i]; v4_out[i] = v4[i]; }
        return 0;
    }
    orc_grid g;
    orc_grid_make(p, &g);
    const uint32_t C = g.total_cells;
    float*    xs    = (float*)malloc(sizeof(float) * 4 * (size_t)(n ? n : 1));
    uint32_t* keys  = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n ? n : 1));
    uint32_t* cnt   = (uint32_t*)malloc(sizeof(uint32_t) * ((size_t)C + 1));
    uint32_t* cs    = (uint32_t*)malloc(sizeof(uint32_t) * ((size_t)C + 1));
    uint32_t* sidx  = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n ? n : 1));
    REAL*     cur   = (REAL*)malloc(sizeof(REAL) * 4 * (size_t)(n ? n : 1));
    REAL*     nxt   = (REAL*)malloc(sizeof(REAL) * 4 * (size_t)(n ? n : 1));
    REAL*     lam   = (REAL*)malloc(sizeof(REAL) * (size_t)(n ? n : 1));

    orc_predict_f32(x4, v4, n, p->gravity, dt, xs);                 /* :715-723 */
    orc_hash(xs, n, &g, keys);                                       /* :609-616 */
    orc_count(keys, n, C, cnt);                                      /* :617-624 */
    orc_exclusive_scan(cnt, cs, C + 1);                              /* :625-648 */
    orc_stable_scatter(keys, n, cs, sidx);                           /* :649-658 */
    for (uint32_t s = 0; s < n; ++s)                                 /* :659-667 */
        for (int a = 0; a < 4; ++a) cur[4 * (size_t)s + a] = (REAL)xs[4 * (size_t)sidx[s] + a];

    if (tr) {
        if (tr->keys) memcpy(tr->keys, keys, sizeof(uint32_t) * n);
        if (tr->cell_start) memcpy(tr->cell_start, cs, sizeof(uint32_t) * ((size_t)C + 1));
        if (tr->sorted_index) memcpy(tr->sorted_index, sidx, sizeof(uint32_t) * n);
        if (tr->xstar4) memcpy(tr->xstar4, xs, sizeof(float) * 4 * (size_t)n);
        if (tr->sx_iter) memcpy(tr->sx_iter, cur, sizeof(REAL) * 4 * (size_t)n);
    }

    for (uint32_t it = 0; it < p->iterations; ++it) {                /* :732-774 */
        const int last = (it + 1 == p->iterations);
        CAT(orc_density_lambda_, SFX)(cur, cs, sidx, n, &g, p, lam,
                                      (last && tr) ? tr->lambda_orig : NULL,
                                      (it == 0 && tr) ? tr->nbr_count : NULL);
        if (tr && tr->lambda_iter) memcpy(tr->lambda_iter + (size_t)it * n, lam, sizeof(REAL) * n);
        CAT(orc_apply_delta_, SFX)(cur, lam, cs, n, &g, p, nxt);
        REAL* t = cur; cur = nxt; nxt = t;
        if (tr && tr->sx_iter)
            memcpy(tr->sx_iter + (size_t)(it + 1) * 4 * n, cur, sizeof(REAL) * 4 * (size_t)n);
    }
    if (p->iterations == 0 && tr && tr->nbr_count)
        CAT(orc_density_lambda_, SFX)(cur, cs, sidx, n, &g, p, lam, NULL, tr->nbr_count);

    for (uint32_t s = 0; s < n; ++s)                                 /* :777-785 */
        for (int a = 0; a < 4; ++a) x4_out[4 * (size_t)sidx[s] + a] = cur[4 * (size_t)s + a];
    REAL* prev = (REAL*)malloc(sizeof(REAL) * 4 * (size_t)(n ? n : 1));
    for (size_t i = 0; i < 4 * (size_t)n; ++i) { v4_out[i] = (REAL)v4[i]; prev[i] = (REAL)x4[i]; }
    CAT(orc_finalize_, SFX)(x4_out, v4_out, prev, n, p, dt);        /* :788-796 */
    free(prev);

    free(xs); free(keys); free(cnt); free(cs); free(sidx); free(cur); free(nxt); free(lam);
    return 0;
}
