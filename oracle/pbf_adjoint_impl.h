/*
 * pbf_adjoint_impl.h -- FP64 rollout step with a tape, and its adjoint
 * (vector-Jacobian product).  Test infrastructure only (see pbf_oracle.h).
 * Included by pbf_oracle.c after the FP64 solver instantiation, whose static
 * helpers (consts_f64, poly6_f64, stencil_f64, ipow_f64, clampr_f64) it uses.
 *
 * The reference has no adjoint; this one differentiates the oracle's FP64
 * step (orc_step_d) with the neighbour structure held fixed (binning and the
 * candidate sets are piecewise constant), the clamps differentiated as
 * pass-through inside [bmin, bmax] and 0 outside, the 0.3h step cap and the
 * h/dt speed cap through their normalisation Jacobians, and the wall
 * responses as piecewise-constant multipliers.  Every pair term is scattered
 * from particle i's own sum (i -> i and i -> j), which is exact even where
 * the stale-binned neighbour relation is not symmetric.  It is pinned by
 * central finite differences of orc_step_d (tests/test_oracle_adjoint.py).
 */

/* Pair quantities for |r|^2 < h^2: Poly6 gradient, Spiky gradient and its
 * Hessian, s_corr and its gradient (Constraints.metal:29-39, 158-180). */
typedef struct pairq {
    double G[3], H[9], Gp[3], s, ds[3];
} pairq;

static void pair_q(const double r[3], double r2, const consts_f64* k, pairq* q)
{
    const double diff = k->h2 - r2;
    const double W = k->poly6_norm * diff * diff * diff;
    const double gpc = -6.0 * k->poly6_norm * diff * diff;
    for (int a = 0; a < 3; ++a) q->Gp[a] = gpc * r[a];
    const double rl = sqrt(r2);
    if (rl > (double)1e-6f && rl < k->h) {
        const double f = -k->spiky_norm * (k->h - rl) * (k->h - rl);
        const double fp = 2.0 * k->spiky_norm * (k->h - rl);
        const double rh[3] = { r[0] / rl, r[1] / rl, r[2] / rl };
        for (int a = 0; a < 3; ++a) {
            q->G[a] = f * rh[a];
            for (int b = 0; b < 3; ++b)
                q->H[3 * a + b] = fp * rh[a] * rh[b] + (f / rl) * ((a == b ? 1.0 : 0.0) - rh[a] * rh[b]);
        }
    } else {
        for (int a = 0; a < 3; ++a) q->G[a] = 0.0;
        for (int a = 0; a < 9; ++a) q->H[a] = 0.0;
    }
    const double ratio = W * k->inv_wq;
    q->s = -k->scorr_k * ipow_f64(ratio, k->scorr_n);
    const double dsc = k->scorr_n ? -k->scorr_k * (double)k->scorr_n * ipow_f64(ratio, k->scorr_n - 1) * k->inv_wq
                                  : 0.0;
    for (int a = 0; a < 3; ++a) q->ds[a] = dsc * q->Gp[a];
}

static void matvec3(const double H[9], const double v[3], double out[3])
{
    for (int a = 0; a < 3; ++a) out[a] = H[3 * a] * v[0] + H[3 * a + 1] * v[1] + H[3 * a + 2] * v[2];
}

/* Visits every neighbour j of sorted slot s (stencil of the FP32 cell of
 * p_s, |r|^2 < h^2 in FP64, j != s), the candidate set of the forward pass. */
#define FOR_NEIGHBOURS(sx4, s, cell_start, g, k, ...)                                   \
    do {                                                                              \
        const double* pi_ = (sx4) + 4 * (size_t)(s);                                  \
        int lo_[3], hi_[3];                                                           \
        stencil_f64(pi_, (g), lo_, hi_);                                              \
        const uint32_t dx_ = (g)->dim[0], dxy_ = (g)->dim[0] * (g)->dim[1];           \
        for (int z_ = lo_[2]; z_ <= hi_[2]; ++z_)                                     \
            for (int y_ = lo_[1]; y_ <= hi_[1]; ++y_)                                 \
                for (int x_ = lo_[0]; x_ <= hi_[0]; ++x_) {                           \
                    const uint32_t c_ = (uint32_t)x_ + (uint32_t)y_ * dx_ + (uint32_t)z_ * dxy_; \
                    for (uint32_t j = (cell_start)[c_]; j < (cell_start)[c_ + 1]; ++j) { \
                        if (j == (s)) continue;                                       \
                        const double* pj_ = (sx4) + 4 * (size_t)j;                    \
                        const double r[3] = { pi_[0] - pj_[0], pi_[1] - pj_[1], pi_[2] - pj_[2] }; \
                        const double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];    \
                        if (r2 >= (k)->h2) continue;                                  \
                        __VA_ARGS__                                                   \
                    }                                                                 \
                }                                                                     \
    } while (0)

/* FP64 predict, binning at the FP32 rounding of x*, FP64 iterations and
 * finalize.  Fills the tape the adjoint needs. */
int orc_step_d(const double* x4, const double* v4, uint32_t n, const orc_params* p, float dt,
               double* x4_out, double* v4_out, orc_tape* tape)
{
    orc_grid g;
    orc_grid_make(p, &g);
    const uint32_t C = g.total_cells, I = p->iterations;
    const size_t nn = n ? n : 1;
    float* xsf = (float*)malloc(sizeof(float) * 4 * nn);
    double* xs = (double*)malloc(sizeof(double) * 4 * nn);
    uint32_t* keys = (uint32_t*)malloc(sizeof(uint32_t) * nn);
    uint32_t* cnt = (uint32_t*)malloc(sizeof(uint32_t) * ((size_t)C + 1));
    double* cur = (double*)malloc(sizeof(double) * 4 * nn);
    double* nxt = (double*)malloc(sizeof(double) * 4 * nn);
    double* lam = (double*)malloc(sizeof(double) * nn);
    const double ddt = (double)dt;
    for (size_t i = 0; i < n; ++i) {
        for (int a = 0; a < 3; ++a) {
            xs[4 * i + a] = x4[4 * i + a] + (v4[4 * i + a] + (double)p->gravity[a] * ddt) * ddt;
            xsf[4 * i + a] = (float)xs[4 * i + a];
        }
        xs[4 * i + 3] = 1.0;
        xsf[4 * i + 3] = 1.0f;
    }
    orc_hash(xsf, n, &g, keys);
    orc_count(keys, n, C, cnt);
    orc_exclusive_scan(cnt, tape->cell_start, C + 1);
    orc_stable_scatter(keys, n, tape->cell_start, tape->sorted_index);
    for (uint32_t s = 0; s < n; ++s)
        for (int a = 0; a < 4; ++a) cur[4 * (size_t)s + a] = xs[4 * (size_t)tape->sorted_index[s] + a];
    memcpy(tape->p, cur, sizeof(double) * 4 * (size_t)n);
    for (uint32_t it = 0; it < I; ++it) {
        orc_density_lambda_f64(cur, tape->cell_start, tape->sorted_index, n, &g, p, lam, NULL, NULL);
        orc_apply_delta_f64(cur, lam, tape->cell_start, n, &g, p, nxt);
        double* t = cur; cur = nxt; nxt = t;
        memcpy(tape->p + (size_t)(it + 1) * 4 * n, cur, sizeof(double) * 4 * (size_t)n);
    }
    for (uint32_t s = 0; s < n; ++s)
        for (int a = 0; a < 4; ++a) x4_out[4 * (size_t)tape->sorted_index[s] + a] = cur[4 * (size_t)s + a];
    memcpy(tape->x_in, x4, sizeof(double) * 4 * (size_t)n);
    for (size_t i = 0; i < 4 * (size_t)n; ++i) v4_out[i] = v4[i];
    orc_finalize_f64(x4_out, v4_out, x4, n, p, dt);
    free(xsf); free(xs); free(keys); free(cnt); free(cur); free(nxt); free(lam);
    return 0;
}

static int inside_box(double v, double lo, double hi) { return v >= lo && v <= hi; }

/* Adjoint of orc_step_d: (xbar_out, vbar_out) -> (xbar_in, vbar_in, gbar).
 *
 * gbar = dt^2 sum_s pbar^0_s.  With sums == NULL that sum is taken
 * directly.  With sums = (sum xbar_out, sum vbar_out) (6, updated to the
 * sums of xbar_in, vbar_in) it follows the recursion that holds because
 * every pair term adds to one particle what it subtracts from another:
 * P0 = Sx + c Sv + sum(du + dp), c = damping / dt, where du = ubar - c vbar
 * (nonzero only where friction, a wall gate or the speed cap acts) and dp
 * collects -pbar of every clamped axis; then sum xbar_in = P0 - c Sv -
 * sum du and sum vbar_in = dt P0.  After chaotic steps, where |pbar| is
 * many orders above the sum, the recursion keeps the sum's accuracy. */
int orc_step_vjp_d(const orc_tape* tape, uint32_t n, const orc_params* p, float dt,
                   const double* xbar_out, const double* vbar_out,
                   double* xbar_in, double* vbar_in, double gbar[3], double* sums)
{
    orc_grid g;
    orc_grid_make(p, &g);
    consts_f64 k;
    make_consts_f64(p, &k);
    const uint32_t I = p->iterations;
    const size_t nn = n ? n : 1;
    const double ddt = (double)dt, inv_dt = 1.0 / ddt, damp = (double)p->damping;
    const double weps = (double)p->wall_eps, vmax = (double)p->vmax * k.h * inv_dt;
    const uint32_t* sidx = tape->sorted_index;
    const uint32_t* cs = tape->cell_start;

    double* pbar = (double*)calloc(4 * nn, sizeof(double));   /* adjoint of the current iterate (sorted) */
    double* pnew = (double*)calloc(4 * nn, sizeof(double));
    double* Rbar = (double*)calloc(4 * nn, sizeof(double));
    double* lam = (double*)calloc(nn, sizeof(double));
    double* lbar = (double*)calloc(nn, sizeof(double));
    double* aux = (double*)calloc(8 * nn, sizeof(double));    /* C, D, g[3], rho_bar, gbar-free slots */
    double* R = (double*)calloc(4 * nn, sizeof(double));

    for (int a = 0; a < 3; ++a) gbar[a] = 0.0;
    double dev_u[3] = { 0, 0, 0 }, dev_p[3] = { 0, 0, 0 };

    /* ---- finalize (Constraints.metal:305-353) ---- */
    const double* pI = tape->p + (size_t)I * 4 * n;
    for (uint32_t s = 0; s < n; ++s) {
        const uint32_t i = sidx[s];
        double xp[3], u[3], w[3], f[3], pass[3];
        for (int a = 0; a < 3; ++a) {
            xp[a] = clampr_f64(pI[4 * (size_t)s + a], k.bmin[a], k.bmax[a]);
            u[a] = (xp[a] - tape->x_in[4 * (size_t)i + a]) * inv_dt * damp;
        }
        const int mn[3] = { xp[0] <= k.bmin[0] + weps, xp[1] <= k.bmin[1] + weps, xp[2] <= k.bmin[2] + weps };
        const int mx[3] = { xp[0] >= k.bmax[0] - weps, xp[1] >= k.bmax[1] - weps, xp[2] >= k.bmax[2] - weps };
        for (int a = 0; a < 3; ++a) {
            pass[a] = 1.0;
            if (mn[a] && u[a] < 0) pass[a] = 0.0;
            if (mx[a] && u[a] > 0) pass[a] = 0.0;
            f[a] = 1.0;
        }
        if (mn[1]) { f[0] = (double)p->floor_friction; f[2] = (double)p->floor_friction; }
        else if (mn[0] || mx[0] || mn[2] || mx[2]) { f[0] = f[1] = f[2] = (double)p->wall_friction; }
        for (int a = 0; a < 3; ++a) w[a] = f[a] * pass[a] * u[a];
        double wb[3] = { vbar_out[4 * (size_t)i], vbar_out[4 * (size_t)i + 1], vbar_out[4 * (size_t)i + 2] };
        const double wl = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
        if (wl > vmax) {
            const double dotwv = (w[0] * wb[0] + w[1] * wb[1] + w[2] * wb[2]) / (wl * wl);
            for (int a = 0; a < 3; ++a) wb[a] = (vmax / wl) * (wb[a] - w[a] * dotwv);
        }
        for (int a = 0; a < 3; ++a) {
            const double ub = f[a] * pass[a] * wb[a] * inv_dt * damp;
            const double xb = xbar_out[4 * (size_t)i + a] + ub;
            xbar_in[4 * (size_t)i + a] = -ub;
            const int in = inside_box(pI[4 * (size_t)s + a], k.bmin[a], k.bmax[a]);
            pbar[4 * (size_t)s + a] = in ? xb : 0.0;
            const int plain = f[a] == 1.0 && pass[a] == 1.0 && !(wl > vmax);
            const double du = plain ? 0.0 : ub - vbar_out[4 * (size_t)i + a] * inv_dt * damp;
            dev_u[a] += du;
            dev_p[a] += du + (in ? 0.0 : -xb);
        }
        xbar_in[4 * (size_t)i + 3] = 0.0;
    }

    /* ---- iterations, last to first ---- */
    for (int it = (int)I - 1; it >= 0; --it) {
        const double* pk = tape->p + (size_t)it * 4 * n;
        /* forward quantities at p^k: lambda, C, D, g, R */
        for (uint32_t s = 0; s < n; ++s) {
            double rho = k.w0, gs[3] = { 0, 0, 0 }, gsq = 0;
            FOR_NEIGHBOURS(pk, s, cs, &g, &k, {
                pairq q;
                pair_q(r, r2, &k, &q);
                const double diff = k.h2 - r2;
                rho += k.poly6_norm * diff * diff * diff;
                for (int a = 0; a < 3; ++a) gs[a] += q.G[a];
                gsq += q.G[0] * q.G[0] + q.G[1] * q.G[1] + q.G[2] * q.G[2];
            });
            const double Cc = rho * k.inv_rho0 - 1.0;
            const double D = k.inv_rho0 * k.inv_rho0 * ((gs[0] * gs[0] + gs[1] * gs[1] + gs[2] * gs[2]) + gsq);
            lam[s] = -Cc / (D + k.eps);
            double* ax = aux + 8 * (size_t)s;
            ax[0] = Cc; ax[1] = D; ax[2] = gs[0]; ax[3] = gs[1]; ax[4] = gs[2];
        }
        for (uint32_t s = 0; s < n; ++s) {
            double Rs[3] = { 0, 0, 0 };
            FOR_NEIGHBOURS(pk, s, cs, &g, &k, {
                pairq q;
                pair_q(r, r2, &k, &q);
                const double c = lam[s] + lam[j] + q.s;
                for (int a = 0; a < 3; ++a) Rs[a] += c * q.G[a];
            });
            for (int a = 0; a < 3; ++a) R[4 * (size_t)s + a] = Rs[a] * k.inv_rho0;
        }
        /* P1: box clamp and step cap */
        for (uint32_t s = 0; s < n; ++s) {
            const double* Rs = R + 4 * (size_t)s;
            const double Rl = sqrt(Rs[0] * Rs[0] + Rs[1] * Rs[1] + Rs[2] * Rs[2]);
            double d[3] = { Rs[0], Rs[1], Rs[2] };
            const int capped = Rl > k.max_step;
            if (capped) for (int a = 0; a < 3; ++a) d[a] *= k.max_step / Rl;
            double tb[3];
            for (int a = 0; a < 3; ++a) {
                const double t = pk[4 * (size_t)s + a] + d[a];
                const int in = inside_box(t, k.bmin[a], k.bmax[a]);
                tb[a] = in ? pbar[4 * (size_t)s + a] : 0.0;
                if (!in) dev_p[a] -= pbar[4 * (size_t)s + a];
                pnew[4 * (size_t)s + a] = tb[a];
            }
            if (capped) {
                const double dot = (Rs[0] * tb[0] + Rs[1] * tb[1] + Rs[2] * tb[2]) / (Rl * Rl);
                for (int a = 0; a < 3; ++a) Rbar[4 * (size_t)s + a] = (k.max_step / Rl) * (tb[a] - Rs[a] * dot);
            } else {
                for (int a = 0; a < 3; ++a) Rbar[4 * (size_t)s + a] = tb[a];
            }
        }
        /* P2: delta pair terms, scattered from i's own sum */
        memset(lbar, 0, sizeof(double) * nn);
        for (uint32_t s = 0; s < n; ++s) {
            const double* Rb = Rbar + 4 * (size_t)s;
            FOR_NEIGHBOURS(pk, s, cs, &g, &k, {
                pairq q;
                pair_q(r, r2, &k, &q);
                const double gr = (q.G[0] * Rb[0] + q.G[1] * Rb[1] + q.G[2] * Rb[2]) * k.inv_rho0;
                lbar[s] += gr;
                lbar[j] += gr;
                const double c = lam[s] + lam[j] + q.s;
                double HR[3];
                matvec3(q.H, Rb, HR);
                for (int a = 0; a < 3; ++a) {
                    const double rb = c * HR[a] * k.inv_rho0 + gr * q.ds[a];
                    pnew[4 * (size_t)s + a] += rb;
                    pnew[4 * (size_t)j + a] -= rb;
                }
            });
        }
        /* P3 + P4: lambda = -C / (D + eps), C = rho/rho0 - 1, D = (|g|^2 + q)/rho0^2 */
        for (uint32_t s = 0; s < n; ++s) {
            double* ax = aux + 8 * (size_t)s;
            const double Cc = ax[0], D = ax[1], de = D + k.eps;
            const double Cb = -lbar[s] / de, Db = lbar[s] * Cc / (de * de);
            ax[5] = Cb * k.inv_rho0;                      /* rho bar */
            ax[6] = Db * k.inv_rho0 * k.inv_rho0;         /* q bar */
            ax[7] = 2.0 * ax[6];                          /* g bar = 2 Dbar g / rho0^2 -> factor */
        }
        for (uint32_t s = 0; s < n; ++s) {
            const double* ax = aux + 8 * (size_t)s;
            const double gb[3] = { ax[7] * ax[2], ax[7] * ax[3], ax[7] * ax[4] };
            FOR_NEIGHBOURS(pk, s, cs, &g, &k, {
                pairq q;
                pair_q(r, r2, &k, &q);
                double Hg[3], HG[3];
                matvec3(q.H, gb, Hg);
                matvec3(q.H, q.G, HG);
                for (int a = 0; a < 3; ++a) {
                    const double rb = ax[5] * q.Gp[a] + Hg[a] + 2.0 * ax[6] * HG[a];
                    pnew[4 * (size_t)s + a] += rb;
                    pnew[4 * (size_t)j + a] -= rb;
                }
            });
        }
        double* t = pbar; pbar = pnew; pnew = t;
        memset(pnew, 0, sizeof(double) * 4 * nn);
    }

    /* ---- predict: x* = x + (v + g dt) dt ---- */
    for (uint32_t s = 0; s < n; ++s) {
        const uint32_t i = sidx[s];
        for (int a = 0; a < 3; ++a) {
            const double xb = pbar[4 * (size_t)s + a];
            xbar_in[4 * (size_t)i + a] += xb;
            vbar_in[4 * (size_t)i + a] = xb * ddt;
            if (!sums) gbar[a] += xb * ddt * ddt;
        }
        vbar_in[4 * (size_t)i + 3] = 0.0;
    }
    if (sums) {
        const double c = inv_dt * damp;
        for (int a = 0; a < 3; ++a) {
            const double p0 = sums[a] + c * sums[3 + a] + dev_p[a];
            sums[a] = p0 - c * sums[3 + a] - dev_u[a];
            sums[3 + a] = ddt * p0;
            gbar[a] = ddt * ddt * p0;
        }
    }
    free(pbar); free(pnew); free(Rbar); free(lam); free(lbar); free(aux); free(R);
    return 0;
}

/* Density at x (fresh hash at the FP32 rounding of x), original order:
 * rho_i = W(0) + sum_j W(r_ij^2) (Constraints.metal:100-115). */
void orc_density_at_d(const double* x4, uint32_t n, const orc_params* p, double* rho)
{
    orc_grid g;
    orc_grid_make(p, &g);
    consts_f64 k;
    make_consts_f64(p, &k);
    const size_t nn = n ? n : 1;
    float* xf = (float*)malloc(sizeof(float) * 4 * nn);
    uint32_t* keys = (uint32_t*)malloc(sizeof(uint32_t) * nn);
    uint32_t* cs = (uint32_t*)malloc(sizeof(uint32_t) * ((size_t)g.total_cells + 1));
    uint32_t* sidx = (uint32_t*)malloc(sizeof(uint32_t) * nn);
    float* s4 = (float*)malloc(sizeof(float) * 4 * nn);
    double* sx = (double*)malloc(sizeof(double) * 4 * nn);
    for (size_t i = 0; i < 4 * (size_t)n; ++i) xf[i] = (float)x4[i];
    orc_build_spatial_hash(xf, n, &g, keys, cs, sidx, s4, 0.0f, NULL);
    for (uint32_t s = 0; s < n; ++s)
        for (int a = 0; a < 4; ++a) sx[4 * (size_t)s + a] = x4[4 * (size_t)sidx[s] + a];
    for (uint32_t s = 0; s < n; ++s) {
        double r_ = k.w0;
        FOR_NEIGHBOURS(sx, s, cs, &g, &k, {
            (void)r;
            const double diff = k.h2 - r2;
            r_ += k.poly6_norm * diff * diff * diff;
        });
        rho[sidx[s]] = r_;
    }
    free(xf); free(keys); free(cs); free(sidx); free(s4); free(sx);
}

/* Adjoint of orc_density_at_d: xbar (original order) from rhobar. */
void orc_density_vjp_d(const double* x4, uint32_t n, const orc_params* p, const double* rhobar, double* xbar)
{
    orc_grid g;
    orc_grid_make(p, &g);
    consts_f64 k;
    make_consts_f64(p, &k);
    const size_t nn = n ? n : 1;
    float* xf = (float*)malloc(sizeof(float) * 4 * nn);
    uint32_t* keys = (uint32_t*)malloc(sizeof(uint32_t) * nn);
    uint32_t* cs = (uint32_t*)malloc(sizeof(uint32_t) * ((size_t)g.total_cells + 1));
    uint32_t* sidx = (uint32_t*)malloc(sizeof(uint32_t) * nn);
    float* s4 = (float*)malloc(sizeof(float) * 4 * nn);
    double* sx = (double*)malloc(sizeof(double) * 4 * nn);
    double* sb = (double*)calloc(4 * nn, sizeof(double));
    for (size_t i = 0; i < 4 * (size_t)n; ++i) xf[i] = (float)x4[i];
    orc_build_spatial_hash(xf, n, &g, keys, cs, sidx, s4, 0.0f, NULL);
    for (uint32_t s = 0; s < n; ++s)
        for (int a = 0; a < 4; ++a) sx[4 * (size_t)s + a] = x4[4 * (size_t)sidx[s] + a];
    for (uint32_t s = 0; s < n; ++s) {
        const double rb = rhobar[sidx[s]];
        FOR_NEIGHBOURS(sx, s, cs, &g, &k, {
            const double diff = k.h2 - r2;
            const double gpc = -6.0 * k.poly6_norm * diff * diff * rb;
            for (int a = 0; a < 3; ++a) {
                sb[4 * (size_t)s + a] += gpc * r[a];
                sb[4 * (size_t)j + a] -= gpc * r[a];
            }
        });
    }
    for (uint32_t s = 0; s < n; ++s)
        for (int a = 0; a < 4; ++a) xbar[4 * (size_t)sidx[s] + a] = a < 3 ? sb[4 * (size_t)s + a] : 0.0;
    free(xf); free(keys); free(cs); free(sidx); free(s4); free(sx); free(sb);
}

/* Branch signature of one taped step, per particle (original order), for
 * counting kink crossings between finite-difference runs: out[(2 + I) i]
 * = FP32 bin key of x*; [+1] = finalize bits (wall flags min/max per axis,
 * bits 0-5; the velocity gates, 6-8; speed cap, 9); [+2+k] = iteration k:
 * (FP32 cell key of p^k << 8) | box-clamp flags per axis (bits 0-2) | step
 * cap (bit 3).  Equal signatures mean the same branch of every
 * piecewise-defined stage. */
void orc_step_kinks_d(const orc_tape* tape, uint32_t n, const orc_params* p, float dt, uint64_t* out)
{
    orc_grid g;
    orc_grid_make(p, &g);
    consts_f64 k;
    make_consts_f64(p, &k);
    const uint32_t I = p->iterations, W = 2 + I;
    const size_t nn = n ? n : 1;
    const uint32_t* sidx = tape->sorted_index;
    const uint32_t* cs = tape->cell_start;
    double* lam = (double*)malloc(sizeof(double) * nn);
    const double ddt = (double)dt, inv_dt = 1.0 / ddt, damp = (double)p->damping;
    const double weps = (double)p->wall_eps, vmax = (double)p->vmax * k.h * inv_dt;
    for (uint32_t s = 0; s < n; ++s) {
        const double* b = tape->p + 4 * (size_t)s;
        const float bf[3] = { (float)b[0], (float)b[1], (float)b[2] };
        out[(size_t)W * sidx[s]] = orc_cell_key(bf, &g);
    }
    for (uint32_t it = 0; it < I; ++it) {
        const double* pk = tape->p + (size_t)it * 4 * n;
        orc_density_lambda_f64(pk, cs, sidx, n, &g, p, lam, NULL, NULL);
        for (uint32_t s = 0; s < n; ++s) {
            double R[3] = { 0, 0, 0 };
            FOR_NEIGHBOURS(pk, s, cs, &g, &k, {
                pairq q;
                pair_q(r, r2, &k, &q);
                const double c = lam[s] + lam[j] + q.s;
                for (int a = 0; a < 3; ++a) R[a] += c * q.G[a];
            });
            for (int a = 0; a < 3; ++a) R[a] *= k.inv_rho0;
            const double Rl = sqrt(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
            const int capped = Rl > k.max_step;
            uint64_t bits = capped ? 8u : 0u;
            for (int a = 0; a < 3; ++a) {
                const double t = pk[4 * (size_t)s + a] + (capped ? R[a] * (k.max_step / Rl) : R[a]);
                if (t < k.bmin[a] || t > k.bmax[a]) bits |= 1u << a;
            }
            const float pf[3] = { (float)pk[4 * (size_t)s], (float)pk[4 * (size_t)s + 1], (float)pk[4 * (size_t)s + 2] };
            out[(size_t)W * sidx[s] + 2 + it] = ((uint64_t)orc_cell_key(pf, &g) << 8) | bits;
        }
    }
    const double* pI = tape->p + (size_t)I * 4 * n;
    for (uint32_t s = 0; s < n; ++s) {
        const uint32_t i = sidx[s];
        double xp[3], u[3], w[3];
        uint64_t bits = 0;
        int mn[3], mx[3];
        for (int a = 0; a < 3; ++a) {
            xp[a] = clampr_f64(pI[4 * (size_t)s + a], k.bmin[a], k.bmax[a]);
            u[a] = (xp[a] - tape->x_in[4 * (size_t)i + a]) * inv_dt * damp;
            mn[a] = xp[a] <= k.bmin[a] + weps;
            mx[a] = xp[a] >= k.bmax[a] - weps;
            if (mn[a]) bits |= 1u << a;
            if (mx[a]) bits |= 1u << (3 + a);
            const int gate = (mn[a] && u[a] < 0) || (mx[a] && u[a] > 0);
            if (gate) bits |= 1u << (6 + a);
            w[a] = gate ? 0.0 : u[a];
        }
        double f[3] = { 1, 1, 1 };
        if (mn[1]) { f[0] = (double)p->floor_friction; f[2] = (double)p->floor_friction; }
        else if (mn[0] || mx[0] || mn[2] || mx[2]) { f[0] = f[1] = f[2] = (double)p->wall_friction; }
        for (int a = 0; a < 3; ++a) w[a] *= f[a];
        if (sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]) > vmax) bits |= 1u << 9;
        out[(size_t)W * i + 1] = bits;
    }
    free(lam);
}
