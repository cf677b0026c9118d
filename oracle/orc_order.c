/*
 * ORACLE -- test infrastructure, not product code.
 *
 * The ordered inversion (the library's photic_invert_ordered): the reference's ordering by a
 * spectral index (model/samodel.c:701-770), its look-up table matched by
 * model_spectral_diff_mean (:980-1004, :2795-2842) and its hot start with restart
 * (:1013-1031, :2233-2241), recast as runs of the sorted sequence whose first pixel (the
 * leader) is inverted with the full depth list and whose other pixels (the followers) either
 * adopt the best-matching earlier leader's solution or start once from their own leader's
 * depth.  Restated here sequentially, pixel by pixel, from the rules alone, to check the GPU
 * batches against.
 */
#include "photic_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* the spectral index of one pixel's Rrs r[4] (bands 1, 2, 3 are blue, green, red,
 * model/samodel.c:361-363); noise: the per-band R_sigma when all are >= 1e-6, else NULL */
float orc_order_key(const float *r, const double *noise)
{
    const float bs = r[1], gs = r[2], rs = r[3];
    float k = (float)sqrt(pow(bs, 2) + pow(gs, 2) + pow(rs, 2));
    if (noise) {
        double ang = bs * noise[1] + gs * noise[2] + rs * noise[3];
        ang /= k * sqrt(pow(noise[1], 2) + pow(noise[2], 2) + pow(noise[3], 2));
        ang = acos(ang);
        k *= ang;
    }
    return k;
}

/* model_spectral_diff_mean for one region and one scene (model/samodel.c:2795-2842). */
float orc_spectral_diff_mean(const double *meas, const double *model)
{
    double error = 0.0;
    int n_total_bands = 0;
    for (int b = 0; b < 4; b++) {
        if (!(b == 1 || b == 2 || b == 3)) continue;
        double current = 0.0, previous = 0.0, nobs = 0.0;
        if (!orc_approx_equal((float)meas[b], 0.0f, 1.0e-6f) &&
            !orc_approx_equal((float)model[b], 0.0f, 1.0e-6f)) {
            current += meas[b];
            previous += model[b];
            nobs += 1.0;
        }
        if (nobs > 0.5) {
            current /= nobs;
            previous /= nobs;
        }
        error += pow((current - previous) / current, 2.0);
        n_total_bands++;
    }
    return (float)(100.0 * sqrt(error / (double)n_total_bands));
}

typedef struct {
    float key;
    long pix;
} orc_keyed;

static int orc_keyed_cmp(const void *a, const void *b)
{
    const orc_keyed *x = (const orc_keyed *)a, *y = (const orc_keyed *)b;
    if (x->key < y->key) return -1;
    if (x->key > y->key) return 1;
    return x->pix < y->pix ? -1 : (x->pix > y->pix ? 1 : 0); /* ties in pixel order */
}

/* What the library writes for one pixel. */
typedef struct {
    double x[ORC_MAXN];
    float depth, rrs_error, model_error;
    int status, n_evals, n_starts, role;
    double Rrs_mod[4];
} orc_order_out;

static void orc_order_take(orc_pixel *px, const orc_result *res, orc_order_out *o)
{
    const int n = orc_n_params(px);
    memcpy(o->x, res->x, sizeof(double) * (size_t)n);
    o->depth = (float)res->depth;
    o->rrs_error = (float)res->Rrs_error;
    o->model_error = (float)res->model_error;
    o->status = res->converged ? 0 : 2;
    o->n_evals = (int)res->n_evals;
    o->n_starts = res->n_starts;
    for (int b = 0; b < 4; b++) o->Rrs_mod[b] = px->Rrs_mod[0][b];
}

/* The ordered inversion of n single-date pixels rrs [4][n] (sec_sun [n] or NULL: sec_sun0);
 * cfg the depth-start list and Nelder-Mead settings.  Writes per pixel (in pixel order) x
 * [np][n], depth, rrs_error, model_error, status (-1 not modelled), n_evals, n_starts and
 * role (-1 .. 3 as the library's).  Returns the number of modelled pixels. */
long orc_invert_ordered(int mode, int n_bottoms, const int *bottom_idx, const orc_bands *bands,
                        long n, const float *rrs, const float *sec_sun, double sec_sun0,
                        const orc_nm_cfg *cfg, float nodata, int run, int window, int use_lut,
                        double match_max, double entry_max, double restart,
                        const double *noise, double *x, float *depth, float *rrs_error,
                        float *model_error, int *status, int *n_evals, int *n_starts, int *role)
{
    orc_keyed *seq = (orc_keyed *)malloc(sizeof(orc_keyed) * (size_t)(n ? n : 1));
    orc_order_out *res = (orc_order_out *)calloc((size_t)(n ? n : 1), sizeof(orc_order_out));
    long nv = 0;
    int np_ = mode == ORC_MODE_MIX ? 5 : 1 + 2 * n_bottoms + 3;
    for (long p = 0; p < n; p++) {
        float r[4];
        int bad = 0;
        for (int b = 0; b < 4; b++) {
            r[b] = rrs[(long)b * n + p];
            if (orc_approx_equal(r[b], nodata, 1.0e-6f) || r[b] < 0.0f) bad = 1;
        }
        res[p].role = -1;
        res[p].status = -1;
        if (bad) continue;
        seq[nv].key = orc_order_key(r, noise);
        seq[nv].pix = p;
        nv++;
    }
    qsort(seq, (size_t)nv, sizeof(orc_keyed), orc_keyed_cmp);

    const long n_lead = (nv + run - 1) / run;
#pragma omp parallel for schedule(dynamic, 1)
    for (long k = 0; k < n_lead; k++) {
        const long p = seq[k * run].pix;
        double r[4], ss = sec_sun ? (double)sec_sun[p] : sec_sun0, sv = 1.0;
        for (int b = 0; b < 4; b++) r[b] = (double)rrs[(long)b * n + p];
        orc_pixel px;
        orc_pixel_init(&px, mode, n_bottoms, bottom_idx, 1, bands, &ss, &sv, r);
        orc_result o;
        orc_optimise(&px, cfg, NULL, &o);
        orc_order_take(&px, &o, &res[p]);
        res[p].role = 0;
    }
#pragma omp parallel for schedule(dynamic, 8)
    for (long t = 0; t < nv; t++) {
        if (t % run == 0) continue;
        const long p = seq[t].pix, k = t / run;
        double r[4], ss = sec_sun ? (double)sec_sun[p] : sec_sun0, sv = 1.0;
        for (int b = 0; b < 4; b++) r[b] = (double)rrs[(long)b * n + p];
        double best = 1.0e4;
        long bk = -1;
        if (use_lut)
            for (long kk = (k - window < 0 ? 0 : k - window); kk <= k; kk++) {
                const orc_order_out *l = &res[seq[kk * run].pix];
                if (!(l->status >= 0 && (double)l->model_error < entry_max && l->depth < 32.0f))
                    continue;
                const double d = orc_spectral_diff_mean(r, l->Rrs_mod);
                if (d < best) {
                    best = d;
                    bk = kk;
                }
            }
        orc_pixel px;
        orc_pixel_init(&px, mode, n_bottoms, bottom_idx, 1, bands, &ss, &sv, r);
        orc_result o;
        if (bk >= 0 && !(best > match_max)) {
            /* a match: the leader's vector, evaluated on this pixel's own data */
            const orc_order_out *l = &res[seq[bk * run].pix];
            orc_eval_outputs(&px, l->x, &o);
            orc_order_take(&px, &o, &res[p]);
            res[p].status = l->status;
            res[p].n_evals = 1;
            res[p].n_starts = 0;
            res[p].role = 1;
            continue;
        }
        /* the hot start: one start from the run leader's depth */
        const orc_order_out *l = &res[seq[k * run].pix];
        orc_nm_cfg hc = *cfg;
        hc.n_hstarts = 1;
        hc.hstart[0] = (double)(l->depth > 0.0f ? l->depth : 1.0f);
        orc_optimise(&px, &hc, NULL, &o);
        orc_order_take(&px, &o, &res[p]);
        res[p].role = 2;
        if (restart > 0.0 && (double)res[p].model_error > restart) {
            const int ev = res[p].n_evals, st = res[p].n_starts;
            orc_pixel px2;
            orc_pixel_init(&px2, mode, n_bottoms, bottom_idx, 1, bands, &ss, &sv, r);
            orc_optimise(&px2, cfg, NULL, &o);
            orc_order_take(&px2, &o, &res[p]);
            res[p].n_evals += ev;
            res[p].n_starts += st;
            res[p].role = 3;
        }
    }
    for (long p = 0; p < n; p++) {
        const orc_order_out *o = &res[p];
        for (int i = 0; i < np_; i++) x[(long)i * n + p] = o->role >= 0 ? o->x[i] : 0.0;
        depth[p] = o->role >= 0 ? o->depth : 0.0f;
        rrs_error[p] = o->role >= 0 ? o->rrs_error : 0.0f;
        model_error[p] = o->role >= 0 ? o->model_error : 0.0f;
        status[p] = o->status;
        n_evals[p] = o->role >= 0 ? o->n_evals : 0;
        n_starts[p] = o->role >= 0 ? o->n_starts : 0;
        role[p] = o->role;
    }
    free(seq);
    free(res);
    return nv;
}
