/*
 * ORACLE -- test infrastructure, not product code.
 *
 * Flat entry points over the oracle for the Python test suite (ctypes): whole arrays in, whole
 * arrays out, layouts matching the product's ([k][n] planar).  `nm` selects the Nelder-Mead
 * routine: NULL for the oracle's restatement, or the reference's own nelmin loaded from
 * oracle/_ref (same calling convention).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "photic_oracle.h"

typedef struct {
    double *x;        /* [n_params][n] */
    float *depth, *rrs_error, *model_error, *k_min, *iod, *albedo, *bottom_type;
    float *b_pct;     /* [3][n] */
    int *n_iter, *converged, *n_evals, *n_starts, *status;
    double *misfit;   /* objective at the best point, in double */
} orc_py_out;

/* Invert n single-date pixels exactly as the product's single-date path is specified. */
long orc_py_invert(int mode, int n_bottoms, const int *bottom_idx, const double *wl, int nb,
                   long n, const float *rrs /*[nb][n]*/, const float *sec_sun /*[n] or NULL*/,
                   double sec_sun0, double reqmin, int konvge, int kcount, int n_hstarts,
                   const double *hstart, double early_exit, float nodata, orc_nelmin_fn nm,
                   const float *seed /*[np][n] or NULL*/, const double *seed_rel,
                   const double *seed_min, orc_py_out *o)
{
    orc_bands bands;
    orc_band_tables(wl, nb, &bands);
    orc_nm_cfg cfg;
    orc_default_cfg(&cfg, 1, reqmin, konvge, kcount);
    cfg.n_hstarts = n_hstarts;
    for (int k = 0; k < n_hstarts && k < 16; k++)
        cfg.hstart[k] = hstart[k];
    if (early_exit > 0.0)
        cfg.early_exit = early_exit;
    long done = 0;
#pragma omp parallel for schedule(dynamic, 8) reduction(+ : done)
    for (long p = 0; p < n; p++) {
        double r[ORC_MAXB], ss = sec_sun ? (double)sec_sun[p] : sec_sun0, sv = 1.0;
        int bad = 0;
        for (int b = 0; b < nb; b++) {
            float v = rrs[(long)b * n + p];
            if (orc_approx_equal(v, nodata, 1.0e-6f) || v < 0.0f)
                bad = 1;
            r[b] = (double)v;
        }
        orc_pixel px;
        orc_pixel_init(&px, mode, n_bottoms, bottom_idx, 1, &bands, &ss, &sv, r);
        int np_ = orc_n_params(&px);
        if (bad) {
            if (o->status) o->status[p] = -1;
            if (o->depth) o->depth[p] = 0.0f;
            if (o->x)
                for (int i = 0; i < np_; i++) o->x[(long)i * n + p] = 0.0;
            continue;
        }
        orc_result res;
        if (seed) {
            double st[ORC_MAXN], sp[ORC_MAXN];
            int off = mode == ORC_MODE_MIX ? 2 : 1 + 2 * n_bottoms;
            for (int i = 0; i < np_; i++) {
                int cls = i == 0 ? 0 : (i < off ? 1 : 2 + (i - off) % 3);
                st[i] = (double)seed[(long)i * n + p];
                double s = fabs(st[i]) * seed_rel[cls];
                if (s < seed_min[cls]) s = seed_min[cls];
                sp[i] = s;
            }
            orc_optimise_seeded(&px, &cfg, nm, st, sp, &res);
        } else {
            orc_optimise(&px, &cfg, nm, &res);
        }
        if (o->x)
            for (int i = 0; i < np_; i++) o->x[(long)i * n + p] = res.x[i];
        if (o->depth) o->depth[p] = (float)res.depth;
        if (o->rrs_error) o->rrs_error[p] = (float)res.Rrs_error;
        if (o->model_error) o->model_error[p] = (float)res.model_error;
        if (o->k_min) o->k_min[p] = (float)res.K_min;
        if (o->iod) o->iod[p] = (float)res.iod;
        if (o->albedo) o->albedo[p] = (float)res.bottom_albedo;
        if (o->bottom_type) o->bottom_type[p] = (float)res.bottom_type;
        if (o->b_pct)
            for (int k = 0; k < 3; k++) o->b_pct[(long)k * n + p] = (float)res.B_percent[k];
        if (o->n_iter) o->n_iter[p] = res.n_iterations;
        if (o->converged) o->converged[p] = res.converged;
        if (o->n_evals) o->n_evals[p] = (int)res.n_evals;
        if (o->n_starts) o->n_starts[p] = res.n_starts;
        if (o->status) o->status[p] = res.converged ? 0 : 2;
        if (o->misfit) o->misfit[p] = res.model_error;
        done++;
    }
    return done;
}

/* Forward model: params [n][5] = (H, P, G, X, B mix), sec_sun [n], Y [n] -> Rrs [n][nb]. */
void orc_py_forward(const double *wl, int nb, long n, const double *params,
                    const double *sec_sun, const double *Y, double *out)
{
    orc_bands bands;
    orc_band_tables(wl, nb, &bands);
    for (long i = 0; i < n; i++) {
        const double *p = params + i * 5;
        for (int b = 0; b < nb; b++) {
            double rho = p[4] * bands.bottom[0][b] + (1.0 - p[4]) * bands.bottom[1][b];
            out[i * nb + b] = orc_rrs_core(p[1], p[2], p[3], rho, p[0], Y[i], &bands, b,
                                           sec_sun[i], 1.0, NULL);
        }
    }
}

/* Objective of MIX-layout vectors x [n][5] for pixels rrs [n][nb], sec_sun [n]. */
void orc_py_objective(const double *wl, int nb, long n, const double *x, const float *rrs,
                      const float *sec_sun, double *out, double *rrs_error)
{
    orc_bands bands;
    orc_band_tables(wl, nb, &bands);
    for (long i = 0; i < n; i++) {
        double r[ORC_MAXB], ss = sec_sun[i], sv = 1.0, xv[5];
        for (int b = 0; b < nb; b++)
            r[b] = (double)rrs[i * nb + b];
        orc_pixel px;
        orc_pixel_init(&px, ORC_MODE_MIX, 2, NULL, 1, &bands, &ss, &sv, r);
        for (int k = 0; k < 5; k++)
            xv[k] = x[i * 5 + k];
        out[i] = orc_objective(xv, &px);
        if (rrs_error)
            rrs_error[i] = px.Rrs_error;
    }
}

/* Synthetic single-date pixels pix0 .. pix0+n-1: rrs [nb][n], sec_sun [n], truth [6][n]
 * (H, P, G, X, B, theta), nodata flag [n], Y [n]. */
void orc_py_gen(uint64_t seed, long pix0, long n, const double *wl, int nb, double sigma,
                double deep_frac, float *rrs, float *sec_sun, double *truth, int *nodata,
                double *Y)
{
    orc_bands bands;
    orc_band_tables(wl, nb, &bands);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; i++) {
        orc_truth t;
        float r[ORC_MAXB];
        double y;
        int nd = orc_gen_pixel(seed, (uint64_t)(pix0 + i), &bands, sigma, deep_frac, &t, r, &y);
        for (int b = 0; b < nb; b++)
            rrs[(long)b * n + i] = r[b];
        sec_sun[i] = (float)orc_sec_deg(t.theta_sun);
        if (truth) {
            truth[0 * n + i] = t.H;
            truth[1 * n + i] = t.P;
            truth[2 * n + i] = t.G;
            truth[3 * n + i] = t.X;
            truth[4 * n + i] = t.B;
            truth[5 * n + i] = t.theta_sun;
        }
        if (nodata) nodata[i] = nd;
        if (Y) Y[i] = y;
    }
}

/* Band tables at the given wavelengths: out [13][nb] = a0, a1, aw, bbw, 8 bottoms, and
 * aw640 in out2. */
void orc_py_band_tables(const double *wl, int nb, double *out, double *aw640)
{
    orc_bands t;
    orc_band_tables(wl, nb, &t);
    for (int b = 0; b < nb; b++) {
        out[0 * nb + b] = t.a0[b];
        out[1 * nb + b] = t.a1[b];
        out[2 * nb + b] = t.aw[b];
        out[3 * nb + b] = t.bbw[b];
        for (int k = 0; k < ORC_NBOTTOM_TYPES; k++)
            out[(4 + k) * nb + b] = t.bottom[k][b];
    }
    *aw640 = t.aw640;
}

/* Test functions for the minimiser: kind 0 = quadratic sum_i w_i (x_i - c_i)^2, kind 1 =
 * Rosenbrock chain, kind 2 = a quadratic floored at truncation steps (exercises the
 * (long)(10 f) local-minimum test). */
typedef struct {
    int kind, n;
    double w[32], c[32];
    long evals;
} orc_testfn;

static double orc_testfn_eval(double x[], void *ctx)
{
    orc_testfn *t = (orc_testfn *)ctx;
    double f = 0.0;
    t->evals++;
    if (t->kind == 1) {
        for (int i = 0; i + 1 < t->n; i++) {
            double a = x[i + 1] - x[i] * x[i], b = 1.0 - x[i];
            f += 100.0 * a * a + b * b;
        }
    } else {
        for (int i = 0; i < t->n; i++) {
            double d = x[i] - t->c[i];
            f += t->w[i] * d * d;
        }
        if (t->kind == 2)
            f = floor(f * 10.0) / 10.0;
    }
    return f;
}

void orc_py_nelmin_test(orc_nelmin_fn nm, int kind, int n, const double *w, const double *c,
                        double *start, const double *step, double reqmin, int konvge,
                        int kcount, double *xmin, double *ynewlo, int *icount, int *numres,
                        int *ifault, long *evals)
{
    orc_testfn t;
    memset(&t, 0, sizeof(t));
    t.kind = kind;
    t.n = n;
    for (int i = 0; i < n; i++) {
        t.w[i] = w ? w[i] : 1.0;
        t.c[i] = c ? c[i] : 0.0;
    }
    double st[32];
    memcpy(st, step, sizeof(double) * (size_t)n);
    if (!nm)
        nm = orc_nelmin;
    nm(orc_testfn_eval, &t, n, start, xmin, ynewlo, reqmin, st, konvge, kcount, icount,
       numres, ifault);
    *evals = t.evals;
}

/* The oracle's restatement and the reference's helpers side by side. */
double orc_py_interp(const double *X, const double *Y, int n, double x)
{
    return orc_interp_1d(X, Y, n, x);
}

int orc_py_approx_equal(float a, float b, float eps) { return orc_approx_equal(a, b, eps); }

void orc_py_philox(const uint32_t *ctr, const uint32_t *key, uint32_t *out)
{
    orc_philox4x32_10(ctr, key, out);
}

void orc_py_mlp_forward(const double *w, int n, const double *in, double *out)
{
    orc_mlp_forward(w, n, in, out);
}

double orc_py_mlp_loss_grad(const double *w, int n, const double *in, const double *target,
                            double *grad)
{
    return orc_mlp_loss_grad(w, n, in, target, grad);
}

int orc_py_mlp_nparam(void) { return ORC_MLP_NPARAM; }

#ifdef _OPENMP
#include <omp.h>
#endif

/* Threads of the oracle's OpenMP loops (<= 0: leave the runtime's default). */
int orc_py_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

/* Y(chi) of model/samodel.c:2898-2901. */
double orc_py_Y(double chi) { return orc_Y_of_chi(chi); }

/* Start point and steps of one depth start (model/samodel.c:2246-2354) for a single-date
 * pixel: mode, n_bottoms (native), wavelengths wl [nb], Rrs rrs [nb]; start, step [n]. */
int orc_py_start_step(int mode, int n_bottoms, const double *wl, int nb, const double *rrs,
                      double h0, double *start, double *step)
{
    orc_bands bands;
    orc_band_tables(wl, nb, &bands);
    orc_pixel px;
    double ss = 1.0, sv = 1.0;
    orc_pixel_init(&px, mode, n_bottoms, NULL, 1, &bands, &ss, &sv, rrs);
    orc_start_step(&px, h0, start, step);
    return orc_n_params(&px);
}

/* Stack scene pixels pix0 .. pix0+n-1: rrs [n_dates][nb][n], sec_sun [n_dates][n] (sec of
 * the drawn zenith), valid [n_dates][n], truth [2 + 4 n_dates][n] (may be NULL). */
void orc_py_gen_stack(uint64_t seed, long pix0, long n, const double *wl, int nb, double sigma,
                      int n_dates, double p_cloud, int k_min, float *rrs, float *sec_sun,
                      uint8_t *valid, double *truth)
{
    orc_bands bands;
    orc_band_tables(wl, nb, &bands);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; i++) {
        double t[2 + 4 * ORC_MAXD];
        float r[ORC_MAXD * ORC_MAXB];
        uint8_t v[ORC_MAXD];
        orc_gen_stack_pixel(seed, (uint64_t)(pix0 + i), &bands, sigma, n_dates, p_cloud, k_min,
                            t, r, v);
        for (int d = 0; d < n_dates; d++) {
            for (int b = 0; b < nb; b++)
                rrs[((long)d * nb + b) * n + i] = r[d * nb + b];
            sec_sun[(long)d * n + i] = (float)orc_sec_deg(t[5 + 4 * d]);
            valid[(long)d * n + i] = v[d];
        }
        if (truth)
            for (int k = 0; k < 2 + 4 * n_dates; k++)
                truth[(long)k * n + i] = t[k];
    }
}

/* Invert n multi-date stacks (MIX layout: [H, B, (P, G, X) per valid date]).  rrs [n_dates]
 * [nb][n], sec_sun [n_dates][n], valid [n_dates][n]; only the valid dates of a pixel enter its
 * problem, in date order (k = their count, 2 + 3k parameters, early exit at 2.5 k).  seed
 * [2 + 3 n_dates][n] or NULL: H, B, then (P, G, X) at each date's slot (invalid slots
 * ignored).  x [2 + 3 n_dates][n]: H, B and each valid date's P, G, X at its date slot (0 at
 * invalid slots).  Other outputs as orc_py_invert. */
/* Multi-date stacks: each pixel's valid dates, in order, are one N_s = k problem
 * (model/samodel.c:1808-1815): mode 1 the mix layout [H, B, (P, G, X) per date], mode 0 the
 * native layout [H, B_k, q_k, (P, G, X) per date] with n_bottoms bottoms; per-date sec(view)
 * [n_dates][n] or NULL (nadir); hstart_px [n]: one depth start per pixel, or NULL (the list). */
long orc_py_invert_stack_ex(const double *wl, int nb, int n_dates, long n, const float *rrs,
                            const float *sec_sun, const float *sec_view, const uint8_t *valid,
                            int mode, int n_bottoms, const int *bottom_idx, double reqmin,
                            int konvge, int kcount, int n_hstarts, const double *hstart,
                            const float *hstart_px, float nodata, const float *seed,
                            const double *seed_rel, const double *seed_min, orc_py_out *o)
{
    orc_bands bands;
    orc_band_tables(wl, nb, &bands);
    const int off = mode == ORC_MODE_MIX ? 2 : 1 + 2 * n_bottoms;
    const int np_max = off + 3 * n_dates;
    long done = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : done)
    for (long p = 0; p < n; p++) {
        double r[ORC_MAXD * ORC_MAXB], ss[ORC_MAXD], sv[ORC_MAXD];
        int slot[ORC_MAXD], k = 0, bad = 0;
        for (int d = 0; d < n_dates; d++) {
            if (!valid[(long)d * n + p])
                continue;
            for (int b = 0; b < nb; b++) {
                float v = rrs[((long)d * nb + b) * n + p];
                if (orc_approx_equal(v, nodata, 1.0e-6f) || v < 0.0f)
                    bad = 1;
                r[k * nb + b] = (double)v;
            }
            ss[k] = (double)sec_sun[(long)d * n + p];
            sv[k] = sec_view ? (double)sec_view[(long)d * n + p] : 1.0;
            slot[k++] = d;
        }
        if (o->x)
            for (int i = 0; i < np_max; i++) o->x[(long)i * n + p] = 0.0;
        if (bad || k == 0) {
            if (o->status) o->status[p] = -1;
            if (o->depth) o->depth[p] = 0.0f;
            continue;
        }
        orc_pixel px;
        orc_pixel_init(&px, mode, n_bottoms, bottom_idx, k, &bands, ss, sv, r);
        orc_nm_cfg cfg;
        orc_default_cfg(&cfg, k, reqmin, konvge, kcount);
        cfg.n_hstarts = n_hstarts;
        for (int h = 0; h < n_hstarts && h < 16; h++)
            cfg.hstart[h] = hstart[h];
        if (hstart_px) {
            cfg.n_hstarts = 1;
            cfg.hstart[0] = (double)hstart_px[p];
        }
        orc_result res;
        const int np_ = orc_n_params(&px);
        if (seed) {
            double st[ORC_MAXN], sp[ORC_MAXN];
            for (int i = 0; i < np_; i++) {
                int src = i < off ? i : off + 3 * slot[(i - off) / 3] + (i - off) % 3;
                int cls = i == 0 ? 0 : (i < off ? 1 : 2 + (i - off) % 3);
                st[i] = (double)seed[(long)src * n + p];
                double s = fabs(st[i]) * seed_rel[cls];
                if (s < seed_min[cls]) s = seed_min[cls];
                sp[i] = s;
            }
            orc_optimise_seeded(&px, &cfg, NULL, st, sp, &res);
        } else {
            orc_optimise(&px, &cfg, NULL, &res);
        }
        if (o->x)
            for (int i = 0; i < np_; i++) {
                int dst = i < off ? i : off + 3 * slot[(i - off) / 3] + (i - off) % 3;
                o->x[(long)dst * n + p] = res.x[i];
            }
        if (o->depth) o->depth[p] = (float)res.depth;
        if (o->rrs_error) o->rrs_error[p] = (float)res.Rrs_error;
        if (o->model_error) o->model_error[p] = (float)res.model_error;
        if (o->k_min) o->k_min[p] = (float)res.K_min;
        if (o->iod) o->iod[p] = (float)res.iod;
        if (o->albedo) o->albedo[p] = (float)res.bottom_albedo;
        if (o->bottom_type) o->bottom_type[p] = (float)res.bottom_type;
        if (o->b_pct)
            for (int q = 0; q < 3; q++) o->b_pct[(long)q * n + p] = (float)res.B_percent[q];
        if (o->n_iter) o->n_iter[p] = res.n_iterations;
        if (o->converged) o->converged[p] = res.converged;
        if (o->n_evals) o->n_evals[p] = (int)res.n_evals;
        if (o->n_starts) o->n_starts[p] = res.n_starts;
        if (o->status) o->status[p] = res.converged ? 0 : 2;
        if (o->misfit) o->misfit[p] = res.model_error;
        done++;
    }
    return done;
}

long orc_py_invert_stack(const double *wl, int nb, int n_dates, long n, const float *rrs,
                         const float *sec_sun, const uint8_t *valid, double reqmin,
                         int konvge, int kcount, int n_hstarts, const double *hstart,
                         float nodata, const float *seed, const double *seed_rel,
                         const double *seed_min, orc_py_out *o)
{
    return orc_py_invert_stack_ex(wl, nb, n_dates, n, rrs, sec_sun, NULL, valid, ORC_MODE_MIX,
                                  2, NULL, reqmin, konvge, kcount, n_hstarts, hstart, NULL,
                                  nodata, seed, seed_rel, seed_min, o);
}

/* Objective of one mix-layout vector over k dates: x [2 + 3k], rrs [k][nb], sec_sun [k];
 * returns the objective and E_Rrs in *rrs_error. */
/* The objective of one k-date vector x (mode 1 mix, mode 0 native with n_bottoms bottoms). */
double orc_py_objective_stack_ex(const double *wl, int nb, int k, int mode, int n_bottoms,
                                 const double *x, const float *rrs, const float *sec_sun,
                                 double *rrs_error)
{
    orc_bands bands;
    orc_band_tables(wl, nb, &bands);
    double r[ORC_MAXD * ORC_MAXB], ss[ORC_MAXD], sv[ORC_MAXD], xv[ORC_MAXN];
    for (int s = 0; s < k; s++) {
        for (int b = 0; b < nb; b++)
            r[s * nb + b] = (double)rrs[s * nb + b];
        ss[s] = (double)sec_sun[s];
        sv[s] = 1.0;
    }
    orc_pixel px;
    orc_pixel_init(&px, mode, n_bottoms, NULL, k, &bands, ss, sv, r);
    for (int i = 0; i < orc_n_params(&px); i++)
        xv[i] = x[i];
    double f = orc_objective(xv, &px);
    if (rrs_error)
        *rrs_error = px.Rrs_error;
    return f;
}

double orc_py_objective_stack(const double *wl, int nb, int k, const double *x,
                              const float *rrs, const float *sec_sun, double *rrs_error)
{
    return orc_py_objective_stack_ex(wl, nb, k, ORC_MODE_MIX, 2, x, rrs, sec_sun, rrs_error);
}

/* Start point and steps of one depth start for a k-date pixel (rrs [k][nb]). */
int orc_py_start_step_stack(int mode, int n_bottoms, const double *wl, int nb, int k,
                            const double *rrs, double h0, double *start, double *step)
{
    orc_bands bands;
    orc_band_tables(wl, nb, &bands);
    orc_pixel px;
    double ss[ORC_MAXD], sv[ORC_MAXD];
    for (int s = 0; s < k; s++)
        ss[s] = sv[s] = 1.0;
    orc_pixel_init(&px, mode, n_bottoms, NULL, k, &bands, ss, sv, rrs);
    orc_start_step(&px, h0, start, step);
    return orc_n_params(&px);
}

/* The no-data disc mask at the pixel indices idx[m] (row * n_cols + col). */
void orc_py_disc_mask(uint64_t seed, long n_rows, long n_cols, int n_discs, double r_min,
                      double r_max, const long *idx, long m, uint8_t *mask)
{
#pragma omp parallel for schedule(static)
    for (long k = 0; k < m; k++)
        mask[k] = (uint8_t)orc_in_nodata_disc(seed, idx[k] / n_cols, idx[k] % n_cols, n_rows,
                                              n_cols, n_discs, r_min, r_max);
}

/* ---- spatial regions (orc_region.c) ---- */

/* Invert the pixels idx[0..n-1] (row * n_cols + col) of an n_rows x n_cols scene with their
 * (2 n_spatial - 1)^2 neighbourhoods: rrs [nd][nb][n_rows * n_cols] (already smoothed),
 * nodata [nd * nb] (one value per band grid), sec_sun / sec_view [nd] per scene, hstart_px [n]
 * one depth start per pixel or NULL (the list).  screen_rrs (same layout as rrs, e.g. the
 * grids before smoothing) or NULL: first apply the reference's per-pixel rule to the pixel's
 * own values there (any band of any scene no-data or negative: not modelled,
 * model/samodel.c:933-947).  Pixels not modelled or with no region left get status -1.
 * x [stride][n] (stride >= the largest parameter count) receives |best| in the reference's
 * layout; n_regions [n] and origin [n] the gather's result.  Other outputs as orc_py_invert. */
/* The same with a band set per scene: nbs[s] bands in scene s at wavelengths wl (the scenes'
 * lists one after another); rrs, screen_rrs [sum nbs][n_rows * n_cols] and nodata [sum nbs]
 * hold the scenes' band grids one after another. */
long orc_py_invert_regions_bands(const double *wl, const int *nbs, int nd, int n_rows,
                                 int n_cols,
                           const float *rrs, const float *nodata, const double *sec_sun,
                           const double *sec_view, int n_spatial, int n_bottoms,
                           const int *bottom_idx, double reqmin, int konvge, int kcount,
                           int n_hstarts, const double *hstart, const float *hstart_px,
                           long n, const long *idx, const float *screen_rrs, int x_stride,
                           int *n_regions,
                           int *origin, orc_py_out *o)
{
    orc_bands *bands = (orc_bands *)calloc((size_t)nd, sizeof(orc_bands));
    const orc_bands *sb[ORC_RMAXD];
    int nrow = 0;
    for (int s = 0; s < nd; s++) {
        orc_band_tables(wl + nrow, nbs[s], &bands[s]);
        sb[s] = &bands[s];
        nrow += nbs[s];
    }
    long done = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : done)
    for (long p = 0; p < n; p++) {
        orc_region_pixel *px = (orc_region_pixel *)calloc(1, sizeof(orc_region_pixel));
        orc_region_result *res = (orc_region_result *)calloc(1, sizeof(orc_region_result));
        const int i = (int)(idx[p] / n_cols), j = (int)(idx[p] % n_cols);
        int bad = 0;
        for (int t = 0; screen_rrs && t < nrow; t++) {
            float v = screen_rrs[(long)t * n_rows * n_cols + idx[p]];
            if (orc_approx_equal(v, nodata[t], 1.0e-6f) || v < 0.0f)
                bad = 1;
        }
        int R = bad ? 0 : orc_region_gather_scenes(px, rrs, nodata, nd, nbs, n_rows, n_cols, i,
                                                   j, n_spatial, 0.0f, NULL);
        if (n_regions) n_regions[p] = R;
        if (origin) origin[p] = px->origin;
        if (o->x)
            for (int k = 0; k < x_stride; k++) o->x[(long)k * n + p] = 0.0;
        if (R == 0) {
            if (o->status) o->status[p] = -1;
            if (o->depth) o->depth[p] = 0.0f;
            free(px);
            free(res);
            continue;
        }
        orc_region_init_scenes(px, n_bottoms, bottom_idx, nd, sb, sec_sun, sec_view);
        orc_nm_cfg cfg;
        orc_default_cfg(&cfg, nd, reqmin, konvge, kcount);
        cfg.n_hstarts = n_hstarts;
        for (int h = 0; h < n_hstarts && h < 16; h++)
            cfg.hstart[h] = hstart[h];
        if (hstart_px) {
            cfg.n_hstarts = 1;
            cfg.hstart[0] = (double)hstart_px[p];
        }
        orc_region_optimise(px, &cfg, NULL, res);
        if (o->x)
            for (int k = 0; k < res->n_params && k < x_stride; k++)
                o->x[(long)k * n + p] = res->x[k];
        if (o->depth) o->depth[p] = (float)res->depth;
        if (o->rrs_error) o->rrs_error[p] = (float)res->Rrs_error;
        if (o->model_error) o->model_error[p] = (float)res->model_error;
        if (o->k_min) o->k_min[p] = (float)res->K_min;
        if (o->iod) o->iod[p] = (float)res->iod;
        if (o->albedo) o->albedo[p] = (float)res->bottom_albedo;
        if (o->bottom_type) o->bottom_type[p] = (float)res->bottom_type;
        if (o->b_pct)
            for (int q = 0; q < 3; q++) o->b_pct[(long)q * n + p] = (float)res->B_percent[q];
        if (o->n_iter) o->n_iter[p] = res->n_iterations;
        if (o->converged) o->converged[p] = res->converged;
        if (o->n_evals) o->n_evals[p] = (int)res->n_evals;
        if (o->n_starts) o->n_starts[p] = res->n_starts;
        if (o->status) o->status[p] = res->converged ? 0 : 2;
        if (o->misfit) o->misfit[p] = res->model_error;
        free(px);
        free(res);
        done++;
    }
    free(bands);
    return done;
}

long orc_py_invert_regions(const double *wl, int nb, int nd, int n_rows, int n_cols,
                           const float *rrs, const float *nodata, const double *sec_sun,
                           const double *sec_view, int n_spatial, int n_bottoms,
                           const int *bottom_idx, double reqmin, int konvge, int kcount,
                           int n_hstarts, const double *hstart, const float *hstart_px,
                           long n, const long *idx, const float *screen_rrs, int x_stride,
                           int *n_regions,
                           int *origin, orc_py_out *o)
{
    int nbs[ORC_RMAXD];
    double *w = (double *)malloc(sizeof(double) * (size_t)(nd * nb));
    for (int s = 0; s < nd; s++) {
        nbs[s] = nb;
        for (int b = 0; b < nb; b++) w[s * nb + b] = wl[b];
    }
    long r = orc_py_invert_regions_bands(w, nbs, nd, n_rows, n_cols, rrs, nodata, sec_sun,
                                         sec_view, n_spatial, n_bottoms, bottom_idx, reqmin,
                                         konvge, kcount, n_hstarts, hstart, hstart_px, n, idx,
                                         screen_rrs, x_stride, n_regions, origin, o);
    free(w);
    return r;
}

static orc_region_pixel *orc_py_region_px(const double *wl, int nb, int nd, int R, int origin,
                                          const double *rrs /*[R][nd][nb]*/,
                                          const double *sec_sun, const double *sec_view,
                                          int n_bottoms, const int *bottom_idx,
                                          orc_bands *bands)
{
    orc_band_tables(wl, nb, bands);
    orc_region_pixel *px = (orc_region_pixel *)calloc(1, sizeof(orc_region_pixel));
    px->n_regions = R;
    px->origin = origin;
    for (int r = 0; r < R; r++)
        for (int s = 0; s < nd; s++)
            for (int b = 0; b < nb; b++)
                px->rrs[r][s][b] = rrs[((long)r * nd + s) * nb + b];
    orc_region_init(px, n_bottoms, bottom_idx, nd, bands, sec_sun, sec_view);
    return px;
}

/* The region objective of one vector x for one pixel's region data rrs [R][nd][nb]; terms
 * [6] = E_Rrs, depth penalty, bottom penalty, K penalty, albedo (the last region's), model
 * error. */
double orc_py_region_objective(const double *wl, int nb, int nd, int R, int origin,
                               const double *rrs, const double *sec_sun,
                               const double *sec_view, int n_bottoms, const int *bottom_idx,
                               const double *x, double *terms)
{
    orc_bands bands;
    orc_region_pixel *px = orc_py_region_px(wl, nb, nd, R, origin, rrs, sec_sun, sec_view,
                                            n_bottoms, bottom_idx, &bands);
    double xv[ORC_RMAXN];
    for (int i = 0; i < orc_region_n_params(px); i++)
        xv[i] = x[i];
    double f = orc_region_objective(xv, px);
    if (terms) {
        terms[0] = px->Rrs_error;
        terms[1] = px->depth_error;
        terms[2] = px->bottom_error;
        terms[3] = px->K_error;
        terms[4] = px->bottom_albedo;
        terms[5] = px->model_error;
    }
    free(px);
    return f;
}

/* orc_py_region_objective with a band set per scene: nbs[s] bands in scene s at wl (the
 * scenes' lists one after another); rrs [R][sum nbs] each region's values, scene by scene. */
double orc_py_region_objective_bands(const double *wl, const int *nbs, int nd, int R,
                                     int origin, const double *rrs, const double *sec_sun,
                                     const double *sec_view, int n_bottoms,
                                     const int *bottom_idx, const double *x, double *terms)
{
    orc_bands *bands = (orc_bands *)calloc((size_t)nd, sizeof(orc_bands));
    const orc_bands *sb[ORC_RMAXD];
    int rows = 0;
    for (int s = 0; s < nd; s++) {
        orc_band_tables(wl + rows, nbs[s], &bands[s]);
        sb[s] = &bands[s];
        rows += nbs[s];
    }
    orc_region_pixel *px = (orc_region_pixel *)calloc(1, sizeof(orc_region_pixel));
    px->n_regions = R;
    px->origin = origin;
    for (int r = 0; r < R; r++) {
        int at = 0;
        for (int s = 0; s < nd; s++)
            for (int b = 0; b < nbs[s]; b++)
                px->rrs[r][s][b] = rrs[(long)r * rows + at++];
    }
    orc_region_init_scenes(px, n_bottoms, bottom_idx, nd, sb, sec_sun, sec_view);
    double xv[ORC_RMAXN];
    for (int i = 0; i < orc_region_n_params(px); i++)
        xv[i] = x[i];
    double f = orc_region_objective(xv, px);
    if (terms) {
        terms[0] = px->Rrs_error;
        terms[1] = px->depth_error;
        terms[2] = px->bottom_error;
        terms[3] = px->K_error;
        terms[4] = px->bottom_albedo;
        terms[5] = px->model_error;
    }
    free(px);
    free(bands);
    return f;
}

/* Start point and steps of one depth start for one pixel's region data. */
int orc_py_region_start_step(const double *wl, int nb, int nd, int R, const double *rrs,
                             int n_bottoms, double h0, double *start, double *step)
{
    orc_bands bands;
    double ss[ORC_RMAXD], sv[ORC_RMAXD];
    for (int s = 0; s < nd; s++)
        ss[s] = sv[s] = 1.0;
    orc_region_pixel *px = orc_py_region_px(wl, nb, nd, R, 0, rrs, ss, sv, n_bottoms, NULL,
                                            &bands);
    orc_region_start_step(px, h0, start, step);
    int n = orc_region_n_params(px);
    free(px);
    return n;
}

/* One pixel's multi-start inversion from its region data; out [8] = depth, E_Rrs, model
 * error, K_min, albedo, IOD, bottom type, evaluations; x receives |best|.  Returns n_params. */
int orc_py_region_optimise(const double *wl, int nb, int nd, int R, int origin,
                           const double *rrs, const double *sec_sun, const double *sec_view,
                           int n_bottoms, const int *bottom_idx, double reqmin, int konvge,
                           int kcount, int n_hstarts, const double *hstart, double *x,
                           double *out)
{
    orc_bands bands;
    orc_region_pixel *px = orc_py_region_px(wl, nb, nd, R, origin, rrs, sec_sun, sec_view,
                                            n_bottoms, bottom_idx, &bands);
    orc_nm_cfg cfg;
    orc_default_cfg(&cfg, nd, reqmin, konvge, kcount);
    cfg.n_hstarts = n_hstarts;
    for (int h = 0; h < n_hstarts && h < 16; h++)
        cfg.hstart[h] = hstart[h];
    orc_region_result *res = (orc_region_result *)calloc(1, sizeof(orc_region_result));
    orc_region_optimise(px, &cfg, NULL, res);
    for (int i = 0; i < res->n_params; i++)
        x[i] = res->x[i];
    out[0] = res->depth;
    out[1] = res->Rrs_error;
    out[2] = res->model_error;
    out[3] = res->K_min;
    out[4] = res->bottom_albedo;
    out[5] = res->iod;
    out[6] = res->bottom_type;
    out[7] = (double)res->n_evals;
    int n = res->n_params;
    free(res);
    free(px);
    return n;
}

/* The gather of pixel (i, j): returns the region count, *origin, and rrs_out [25][nd][nb]. */
int orc_py_region_gather(const float *rrs, const float *nodata, int nd, int nb, int n_rows,
                         int n_cols, int i, int j, int n_spatial, int *origin, double *rrs_out)
{
    orc_region_pixel *px = (orc_region_pixel *)calloc(1, sizeof(orc_region_pixel));
    int R = orc_region_gather(px, rrs, nodata, nd, nb, n_rows, n_cols, i, j, n_spatial, 0.0f,
                              NULL);
    *origin = px->origin;
    for (int r = 0; r < R; r++)
        for (int s = 0; s < nd; s++)
            for (int b = 0; b < nb; b++)
                rrs_out[((long)r * nd + s) * nb + b] = px->rrs[r][s][b];
    free(px);
    return R;
}

/* Every pixel of a grid [n_rows * n_cols] smoothed with radius r. */
void orc_py_smooth_grid(const float *grid, int n_rows, int n_cols, int r, float spval,
                        float *out)
{
    for (int i = 0; i < n_rows; i++)
        for (int j = 0; j < n_cols; j++)
            out[(long)i * n_cols + j] = orc_smoothed_index_2d(grid, i, j, n_rows, n_cols, r, spval);
}

/* The region workload's coherent scene: rrs [nb][n], truth [6][n] (H, P, G, X, B, theta) or
 * NULL, nodata [n] (1 where a band went negative), n = n_rows * n_cols. */
void orc_py_gen_field(uint64_t seed, long n_rows, long n_cols, double theta, const double *wl,
                      int nb, double sigma, float *rrs, double *truth, uint8_t *nodata)
{
    orc_bands bands;
    orc_band_tables(wl, nb, &bands);
    const long n = n_rows * n_cols;
#pragma omp parallel for schedule(static)
    for (long q = 0; q < n; q++) {
        orc_truth t;
        float r[ORC_MAXB];
        int bad = orc_gen_field_pixel(seed, q / n_cols, q % n_cols, n_cols, theta, &bands, sigma,
                                      &t, r);
        for (int b = 0; b < nb; b++) rrs[(long)b * n + q] = r[b];
        if (nodata) nodata[q] = (uint8_t)bad;
        if (truth) {
            truth[0 * n + q] = t.H;
            truth[1 * n + q] = t.P;
            truth[2 * n + q] = t.G;
            truth[3 * n + q] = t.X;
            truth[4 * n + q] = t.B;
            truth[5 * n + q] = t.theta_sun;
        }
    }
}

/* The depth-error pass (orc_depth_sigma) with the reference's Nelder-Mead settings given:
 * depth [npix] positive, grid [nd][nb][npix] smoothed, nodata / rsig [nd * nb], sec_sun /
 * sec_view [nd], emp [npix] or NULL.  table [120], sigma [npix]; returns the interval count. */
int orc_py_depth_sigma(const double *wl, int nb, int nd, int n_rows, int n_cols,
                       const float *depth, const float *grid, const float *nodata,
                       const double *rsig, const double *sec_sun, const double *sec_view,
                       int n_spatial, int n_bottoms, const float *emp, float emp_nodata,
                       double h_prev, double reqmin, int konvge, int kcount, double *table,
                       float *sigma)
{
    orc_bands bands;
    orc_band_tables(wl, nb, &bands);
    orc_nm_cfg cfg;
    orc_default_cfg(&cfg, nd, reqmin, konvge, kcount);
    return orc_depth_sigma(depth, n_rows, n_cols, grid, nodata, rsig, nd, &bands, sec_sun,
                           sec_view, n_spatial, n_bottoms, emp, emp_nodata, h_prev, &cfg, table,
                           sigma);
}

/* One perturbed region inversion (the depth-error pass's sample): pixel (i, j) of grid
 * [nd][nb][n_rows * n_cols] with n_sigma R_sigma added; returns the depth in double (0 when
 * no region is left), x [n_params] |best|, *n_evals. */
double orc_py_region_sample(const double *wl, int nb, int nd, int n_rows, int n_cols,
                            const float *grid, const float *nodata, const double *rsig,
                            const double *sec_sun, const double *sec_view, int n_spatial,
                            int n_bottoms, int i, int j, float n_sigma, double reqmin,
                            int konvge, int kcount, double *x, long *n_evals)
{
    orc_bands bands;
    orc_band_tables(wl, nb, &bands);
    orc_nm_cfg cfg;
    orc_default_cfg(&cfg, nd, reqmin, konvge, kcount);
    orc_region_pixel *px = (orc_region_pixel *)calloc(1, sizeof(orc_region_pixel));
    orc_region_result *res = (orc_region_result *)calloc(1, sizeof(orc_region_result));
    const int bidx[3] = {0, 1, 2};
    double depth = 0.0;
    int R = orc_region_gather(px, grid, nodata, nd, nb, n_rows, n_cols, i, j, n_spatial, n_sigma,
                              rsig);
    if (R > 0) {
        orc_region_init(px, n_bottoms, bidx, nd, &bands, sec_sun, sec_view);
        orc_region_optimise(px, &cfg, NULL, res);
        depth = res->depth;
        for (int k = 0; x && k < res->n_params; k++) x[k] = res->x[k];
        if (n_evals) *n_evals = res->n_evals;
    }
    free(px);
    free(res);
    return depth;
}

/* The objective of one k-date vector x (rrs [k][nb], sec_sun, sec_view [k]) and the outputs
 * derived from it: out [16] = objective, E_Rrs, K penalty, depth, K_min, IOD, albedo, bottom
 * type, B_percent[8]; K [k][nb] the attenuation of every (date, band). */
void orc_py_eval(int mode, int n_bottoms, const int *bottom_idx, const double *wl, int nb, int k,
                 const double *x, const float *rrs, const double *sec_sun,
                 const double *sec_view, double *out, double *K)
{
    orc_bands bands;
    orc_band_tables(wl, nb, &bands);
    double r[ORC_MAXD * ORC_MAXB], ss[ORC_MAXD], sv[ORC_MAXD];
    for (int s = 0; s < k; s++) {
        for (int b = 0; b < nb; b++)
            r[s * nb + b] = (double)rrs[s * nb + b];
        ss[s] = sec_sun[s];
        sv[s] = sec_view[s];
    }
    orc_pixel px;
    orc_pixel_init(&px, mode, n_bottoms, bottom_idx, k, &bands, ss, sv, r);
    orc_result res;
    orc_eval_outputs(&px, x, &res);
    out[0] = px.model_error;
    out[1] = px.Rrs_error;
    out[2] = px.K_error;
    out[3] = res.depth;
    out[4] = res.K_min;
    out[5] = res.iod;
    out[6] = res.bottom_albedo;
    out[7] = res.bottom_type;
    for (int t = 0; t < ORC_NBOTTOM_TYPES; t++)
        out[8 + t] = res.B_percent[t];
    for (int s = 0; s < k; s++)
        for (int b = 0; b < nb; b++)
            K[s * nb + b] = px.K[s][b];
}

/* The region objective of one vector x over region data rrs [R][nd][nb] and the region
 * outputs: out [4] = objective, depth, K_min, IOD. */
void orc_py_region_eval(const double *wl, int nb, int nd, int R, int origin, const double *rrs,
                        const double *sec_sun, const double *sec_view, int n_bottoms,
                        const int *bottom_idx, const double *x, double *out)
{
    orc_bands bands;
    orc_region_pixel *px = orc_py_region_px(wl, nb, nd, R, origin, rrs, sec_sun, sec_view,
                                            n_bottoms, bottom_idx, &bands);
    orc_region_result res;
    orc_region_eval_outputs(px, x, &res);
    out[0] = px->model_error;
    out[1] = res.depth;
    out[2] = res.K_min;
    out[3] = res.iod;
    free(px);
}

/* The ordered inversion (orc_invert_ordered) with the given settings; noise [4] or NULL. */
long orc_py_invert_ordered(int mode, int n_bottoms, const int *bottom_idx, const double *wl,
                           int nb, long n, const float *rrs, const float *sec_sun,
                           double sec_sun0, double reqmin, int konvge, int kcount, float nodata,
                           int run, int window, int use_lut, double match_max,
                           double entry_max, double restart, const double *noise, double *x,
                           float *depth, float *rrs_error, float *model_error, int *status,
                           int *n_evals, int *n_starts, int *role)
{
    orc_bands bands;
    orc_band_tables(wl, nb, &bands);
    orc_nm_cfg cfg;
    orc_default_cfg(&cfg, 1, reqmin, konvge, kcount);
    return orc_invert_ordered(mode, n_bottoms, bottom_idx, &bands, n, rrs, sec_sun, sec_sun0,
                              &cfg, nodata, run, window, use_lut, match_max, entry_max, restart,
                              noise, x, depth, rrs_error, model_error, status, n_evals,
                              n_starts, role);
}

float orc_py_spectral_diff_mean(const double *meas, const double *model)
{
    return orc_spectral_diff_mean(meas, model);
}

float orc_py_order_key(const float *r, const double *noise) { return orc_order_key(r, noise); }
