/*
 * ORACLE -- test infrastructure, not product code.  See photic_oracle.h for what it is and
 * who may load it.
 *
 * Spatial regions: the reference's n_spatial > 1 run, in which a pixel is inverted jointly
 * with its (2 n_spatial - 1)^2 neighbours, one depth and one bottom per region and one set of
 * water constituents per scene shared by all regions (model/samodel.c:1808-1815 layout with
 * N_r > 1).  Restated here:
 *   orc_smoothed_index_2d  smoothed_index_2d, model/common.c:232-276 (float arithmetic)
 *   orc_region_gather      extract_Rrs_data, model/samodel.c:2957-3027 (n_sigma = 0)
 *   orc_region_objective   samodel_error, model/samodel.c:2432-2759, every term
 *   orc_region_start_step  model/samodel.c:2246-2354 with region means
 *   orc_region_optimise    model/samodel.c:2119-2427 multi-start and the outputs of
 *                          model/samodel.c:1992-2079 (origin-weighted depth, the origin's
 *                          bottom percentages, IOD over every region)
 * Pinning (tests/test_oracle_region.py): with one region every function reduces bit for bit
 * to the pinned single-region oracle (orc_objective, orc_start_step, orc_optimise); the
 * continuity penalties against hand-computed values; the gather on hand-built grids (edges,
 * clamped duplicates, no-data neighbours, the origin's index); the smoothing bit for bit
 * against the reference's own compiled common.c (oracle/_ref).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "photic_oracle.h"

/* The (2r - 1)^2 box mean around (i, j) with clamp-to-edge indexing, no-data values skipped,
 * all no-data giving the no-data value, in float as model/common.c:232-276 computes it;
 * grid [n_rows * n_cols] row-major. */
float orc_smoothed_index_2d(const float *grid, int i, int j, int n_rows, int n_cols, int r,
                            float spval)
{
    float sum = 0.0f, nobs = 0.0f;
    for (int ii = 1 - r; ii < r; ii++) {
        int iii = i + ii < 0 ? 0 : (i + ii > n_rows - 1 ? n_rows - 1 : i + ii);
        for (int jj = 1 - r; jj < r; jj++) {
            int jjj = j + jj < 0 ? 0 : (j + jj > n_cols - 1 ? n_cols - 1 : j + jj);
            float v = grid[(long)iii * n_cols + jjj];
            if (!orc_approx_equal(v, spval, 1.0e-6f)) {
                sum += v;
                nobs += 1.0f;
            }
        }
    }
    return nobs < 0.5f ? spval : sum / nobs;
}

int orc_region_n_params(const orc_region_pixel *px)
{
    return px->n_regions * (1 + 2 * px->n_bottoms) + 3 * px->n_dates;
}

/* extract_Rrs_data (model/samodel.c:2957-3027): candidates (ii, jj) over [1 - ns, ns - 1]^2
 * in row-major order, clamped to the grid; a candidate with any no-data band (approx_equal
 * only -- negative values are kept) is dropped; the origin is the index the centre candidate
 * had when it was visited (the last such candidate at a clamped edge).  rsig [nd * nb] or NULL:
 * a perturbed gather adds n_sigma R_sigma to each kept value (model/samodel.c:3003-3007). */
int orc_region_gather_scenes(orc_region_pixel *px, const float *rrs, const float *nodata,
                             int nd, const int *nbs, int n_rows, int n_cols, int i, int j,
                             int n_spatial, float n_sigma, const double *rsig)
{
    const long npix = (long)n_rows * n_cols;
    int ns = n_spatial == 0 ? 1 : n_spatial;
    int kr = 0;
    px->origin = 0;
    for (int ii = 1 - ns; ii < ns; ii++) {
        int iii = i + ii < 0 ? 0 : (i + ii > n_rows - 1 ? n_rows - 1 : i + ii);
        for (int jj = 1 - ns; jj < ns; jj++) {
            int jjj = j + jj < 0 ? 0 : (j + jj > n_cols - 1 ? n_cols - 1 : j + jj);
            int bad = 0, row = 0;
            for (int s = 0; s < nd && !bad; s++) {
                for (int b = 0; b < nbs[s]; b++) {
                    float v = rrs[(long)(row + b) * npix + (long)iii * n_cols + jjj];
                    if (!orc_approx_equal(v, nodata[row + b], 1.0e-6f)) {
                        px->rrs[kr][s][b] = (double)v;
                        if (rsig && !orc_approx_equal(n_sigma, 0.0f, 1.0e-6f))
                            px->rrs[kr][s][b] += n_sigma * rsig[row + b];
                    } else {
                        bad = 1;
                        break;
                    }
                }
                row += nbs[s];
            }
            if (i == iii && j == jjj)
                px->origin = kr;
            if (!bad)
                kr++;
        }
    }
    px->n_regions = kr;
    return kr;
}

int orc_region_gather(orc_region_pixel *px, const float *rrs /*[nd][nb][n_rows*n_cols]*/,
                      const float *nodata /*[nd*nb]*/, int nd, int nb, int n_rows, int n_cols,
                      int i, int j, int n_spatial, float n_sigma, const double *rsig)
{
    int nbs[ORC_RMAXD];
    for (int s = 0; s < nd; s++)
        nbs[s] = nb;
    return orc_region_gather_scenes(px, rrs, nodata, nd, nbs, n_rows, n_cols, i, j, n_spatial,
                                    n_sigma, rsig);
}

/* Band tables, geometry and the per-region Rrs at 440/490/550/640/750 nm
 * (model/samodel.c:1785-1799); px->rrs, n_regions and origin already set. */
void orc_region_init_scenes(orc_region_pixel *px, int n_bottoms, const int *bottom_idx,
                            int n_dates, const orc_bands *const *sb, const double *sec_sun,
                            const double *sec_view)
{
    px->n_bottoms = n_bottoms;
    for (int k = 0; k < n_bottoms; k++)
        px->bottom_idx[k] = bottom_idx ? bottom_idx[k] : k;
    px->n_dates = n_dates;
    px->bands = sb[0];
    for (int s = 0; s < n_dates; s++) {
        px->sb[s] = sb[s];
        px->sec_sun[s] = sec_sun[s];
        px->sec_view[s] = sec_view[s];
    }
    for (int r = 0; r < px->n_regions; r++)
        for (int s = 0; s < n_dates; s++) {
            const orc_bands *t = sb[s];
            const int nb = t->n_bands;
            px->r440[r][s] = orc_interp_1d(t->wl, px->rrs[r][s], nb, 440.0);
            px->r490[r][s] = orc_interp_1d(t->wl, px->rrs[r][s], nb, 490.0);
            px->r550[r][s] = orc_interp_1d(t->wl, px->rrs[r][s], nb, 550.0);
            px->r640[r][s] = orc_interp_1d(t->wl, px->rrs[r][s], nb, 640.0);
            px->r750[r][s] = orc_interp_1d(t->wl, px->rrs[r][s], nb, 750.0);
            if (px->r440[r][s] < 0.0)
                px->r440[r][s] = 0.0001;
        }
    px->n_evals = 0;
}

void orc_region_init(orc_region_pixel *px, int n_bottoms, const int *bottom_idx, int n_dates,
                     const orc_bands *bands, const double *sec_sun, const double *sec_view)
{
    const orc_bands *sb[ORC_RMAXD];
    for (int s = 0; s < n_dates; s++)
        sb[s] = bands;
    orc_region_init_scenes(px, n_bottoms, bottom_idx, n_dates, sb, sec_sun, sec_view);
}

/* samodel_error (model/samodel.c:2432-2759) over every region: the misfit over (region,
 * scene, band) in that order, the depth and bottom continuity penalties, and the K penalty
 * from the origin's depth and the K of the last region evaluated (the reference keeps one
 * K[scene][band] array, overwritten region by region). */
double orc_region_objective(double *x, void *ctx)
{
    orc_region_pixel *px = (orc_region_pixel *)ctx;
    const int R = px->n_regions, nd = px->n_dates, nbot = px->n_bottoms;
    const int off = R + 2 * nbot * R;
    px->n_evals++;

    double err = 0.0, tot = 0.0;
    int n_total = 0;
    for (int r = 0; r < R; r++) {
        for (int s = 0; s < nd; s++) {
            const orc_bands *t = px->sb[s];
            const int nbd = t->n_bands;
            double H = fabs(x[r]);
            double P = 0.01 * fabs(x[off + 3 * s]);
            double G = 0.01 * fabs(x[off + 1 + 3 * s]);
            double X = 0.01 * fabs(x[off + 2 + 3 * s]);
            double B[ORC_RMAXBOT], q[ORC_RMAXBOT], qsum = 0.0;
            for (int k = 0; k < nbot; k++)
                B[k] = 0.01 * fabs(x[R + r * nbot + k]);
            for (int k = 0; k < nbot; k++) {
                q[k] = fabs(x[R + R * nbot + r * nbot + k]);
                qsum += q[k];
            }
            for (int k = 0; k < nbot; k++)
                q[k] /= qsum;
            double Y = orc_Y_of_chi(px->r440[r][s] / px->r490[r][s]);
            for (int b = 0; b < nbd; b++) {
                double rho = 0.0, alb = 0.0, side[4];
                for (int k = 0; k < nbot; k++) {
                    alb += q[k] * B[k];
                    rho += q[k] * B[k] * t->bottom[px->bottom_idx[k]][b];
                }
                px->bottom_albedo = alb;
                double m = orc_rrs_core(P, G, X, rho, H, Y, t, b, px->sec_sun[s],
                                        px->sec_view[s], side);
                px->K[s][b] = side[0];
                px->rrs_bot[r][s][b] = side[2];
                px->rrs_mod[r][s][b] = side[3];
                double d = m - px->rrs[r][s][b];
                err += d * d;
                tot += px->rrs[r][s][b];
                n_total++;
            }
        }
    }
    double mean = tot / (double)n_total;
    double e_rrs = 100.0 * sqrt(err / (double)n_total) / mean;
    px->Rrs_error = e_rrs;

    /* depth continuity (model/samodel.c:2599-2630) */
    double ed = 0.0, dm = 0.0, nout = 0.0, thr;
    for (int r = 0; r < R; r++)
        dm += fabs(x[r]);
    dm /= (double)R;
    if (dm < 4.0) thr = 0.4;
    else if (dm < 8.0) thr = 0.2;
    else if (dm < 12.0) thr = 0.1;
    else thr = 0.05;
    for (int r = 0; r < R; r++) {
        double h = fabs(x[r]);
        if (h < (1.0 - thr) * dm || h > (1.0 + thr) * dm) {
            ed += (h - dm) * (h - dm);
            nout += 1.0;
        }
    }
    if (nout > 0.5)
        ed = 100.0 * sqrt(ed / nout) / dm;
    px->depth_error = ed;

    /* bottom continuity (model/samodel.c:2632-2692) over q_k B_k / sum q */
    double bthr;
    if (dm < 5.0) bthr = 0.25;
    else if (dm < 10.0) bthr = 0.1;
    else if (dm < 15.0) bthr = 0.05;
    else bthr = 0.01;
    double eb = 0.0, btot = 0.0;
    nout = 0.0;
    for (int k = 0; k < nbot; k++) {
        double bm = 0.0;
        for (int r = 0; r < R; r++) {
            double qs = 0.0;
            for (int kb = 0; kb < nbot; kb++)
                qs += fabs(x[R + R * nbot + r * nbot + kb]);
            bm += fabs(x[R + r * nbot + k]) * fabs(x[R + R * nbot + r * nbot + k]) / qs;
        }
        bm /= (double)R;
        btot += bm;
        for (int r = 0; r < R; r++) {
            double qs = 0.0;
            for (int kb = 0; kb < nbot; kb++)
                qs += fabs(x[R + R * nbot + r * nbot + kb]);
            double bo = fabs(x[R + r * nbot + k]) * fabs(x[R + R * nbot + r * nbot + k]) / qs;
            if (bo < (1.0 - bthr) * bm || bo > (1.0 + bthr) * bm) {
                eb += (bo - bm) * (bo - bm);
                nout += 1.0;
            }
        }
    }
    if (nout > 0.5) {
        double bm = btot / (double)nbot;
        eb = 100.0 * sqrt(eb / nout) / bm;
    }
    px->bottom_error = eb;

    /* K penalty (model/samodel.c:2696-2732): the origin's depth; per scene the smallest
     * non-zero K of the last region; the last scene's minimum drives the K > 0.7 term */
    const double kmm = 0.275, kmn = 0.185;
    double H = fabs(x[px->origin]);
    double Kerr = 0.0, Kmin = 1.0e4;
    for (int s = 0; s < nd; s++) {
        Kmin = 1.0e4;
        for (int b = 0; b < px->sb[s]->n_bands; b++) {
            double k = px->K[s][b];
            if (!orc_approx_equal((float)k, 0.0f, 1.0e-6f) && k < Kmin)
                Kmin = k;
        }
        double u;
        if (H < 1.0 && Kmin < kmm) {
            u = 1.0 / (0.01 + Kmin) - 1.0 / (0.01 + kmn);
            Kerr += 100.0 * (u * u);
        } else if (H < 2.0 && Kmin < 0.5 * (1.5 * kmn + 0.5 * kmm)) {
            u = 1.0 / (0.01 + Kmin) - 1.0 / (0.01 + 0.5 * (1.5 * kmn + 0.5 * kmm));
            Kerr += 100.0 * (u * u);
        } else if (H < 3.0 && Kmin < 0.5 * (1.25 * kmn + 0.75 * kmm)) {
            u = 1.0 / (0.01 + Kmin) - 1.0 / (0.01 + 0.5 * (1.25 * kmn + 0.75 * kmm));
            Kerr += 100.0 * (u * u);
        } else if (H < 4.0 && Kmin < 0.5 * (1.5 * kmn + 0.5 * kmm)) {
            u = 1.0 / (0.01 + Kmin) - 1.0 / (0.01 + 0.5 * (1.5 * kmn + 0.5 * kmm));
            Kerr += 100.0 * (u * u);
        } else if (H < 5.0 && Kmin < 0.5 * (1.75 * kmn + 0.25 * kmm)) {
            u = 1.0 / (0.01 + Kmin) - 1.0 / (0.01 + 0.5 * (1.75 * kmn + 0.25 * kmm));
            Kerr += 100.0 * (u * u);
        }
    }
    if (Kmin > 0.7) {
        double u = 4.0 * (Kmin - 0.7);
        Kerr += 100.0 * (u * u);
    }
    px->K_error = Kerr;

    double e = 80.0 * (e_rrs * 1.0) + 15.0 * ed + 10.0 * eb + 15.0 * Kerr;
    e /= 80.0 + 15.0 + 10.0 + 15.0;
    px->model_error = e;
    return e;
}

/* Start point and steps for one depth start (model/samodel.c:2246-2354): every region's
 * depth at h0 (step 1.25 h0); every amplitude at 100 * 4 * the mean Rrs(490) over regions and
 * scenes (step 1.5x); every weight 1 (step 0.5); per scene P, G, X from the region means of
 * Rrs(490), Rrs(550), Rrs(640) (steps 2x). */
void orc_region_start_step(const orc_region_pixel *px, double h0, double *start, double *step)
{
    const int R = px->n_regions, nd = px->n_dates, nbot = px->n_bottoms;
    for (int r = 0; r < R; r++) {
        start[r] = h0;
        step[r] = 1.25 * start[r];
    }
    double m490 = 0.0;
    for (int r = 0; r < R; r++)
        for (int s = 0; s < nd; s++)
            m490 += px->r490[r][s];
    m490 /= (double)R * nd;
    for (int k = 0; k < nbot; k++)
        for (int r = 0; r < R; r++) {
            int i = R + r * nbot + k;
            start[i] = 100.0 * 4.0 * m490;
            step[i] = 1.5 * start[i];
        }
    for (int k = 0; k < nbot; k++)
        for (int r = 0; r < R; r++) {
            int i = R + nbot * R + r * nbot + k;
            start[i] = 1.0;
            step[i] = 0.5 * start[i];
        }
    const int off = R + 2 * nbot * R;
    for (int s = 0; s < nd; s++) {
        double a = 0.0, b = 0.0, c = 0.0;
        for (int r = 0; r < R; r++) {
            a += px->r490[r][s];
            b += px->r550[r][s];
            c += px->r640[r][s];
        }
        a /= (double)R;
        b /= (double)R;
        c /= (double)R;
        start[off + 3 * s] = 100.0 * 0.072 * pow(a / b, -1.7);
        start[off + 1 + 3 * s] = 1.5 * start[off + 3 * s];
        start[off + 2 + 3 * s] = 100.0 * 30.0 * px->bands->aw640 * c;
        step[off + 3 * s] = 2.0 * start[off + 3 * s];
        step[off + 1 + 3 * s] = 2.0 * start[off + 1 + 3 * s];
        step[off + 2 + 3 * s] = 2.0 * start[off + 2 + 3 * s];
    }
}

/* The outputs of model/samodel.c:1992-2079 after the final re-evaluation at best. */
static void orc_region_outputs(orc_region_pixel *px, const double *best, orc_region_result *res)
{
    const int R = px->n_regions, nd = px->n_dates, nbot = px->n_bottoms;
    const int n = orc_region_n_params(px);
    res->n_params = n;
    res->n_regions = R;
    for (int i = 0; i < n; i++)
        res->x[i] = fabs(best[i]);
    res->model_error = px->model_error;
    res->Rrs_error = px->Rrs_error;
    res->bottom_albedo = px->bottom_albedo;

    double kmin = 1.0e10;
    for (int s = 0; s < nd; s++)
        for (int b = 0; b < px->sb[s]->n_bands; b++) {
            double k = px->K[s][b];
            if (!orc_approx_equal((float)k, 0.0f, 1.0e-4f) && k < kmin)
                kmin = k;
        }
    res->K_min = (kmin == 1.0e10) ? 0.0 : kmin;

    /* depth: regions weighted 1, the origin sqrt(N_r) (model/samodel.c:2000-2017) */
    double ow = sqrt((double)R), d = 0.0;
    for (int r = 0; r < R; r++) {
        double H = fabs(best[r]);
        if (r == px->origin)
            d += ow * H;
        else
            d += H;
    }
    res->depth = d / (ow + ((double)R) - 1.0);

    for (int k = 0; k < ORC_NBOTTOM_TYPES; k++)
        res->B_percent[k] = 0.0;
    double qsum = 0.0;
    for (int k = 0; k < nbot; k++)
        qsum += fabs(best[R + R * nbot + px->origin * nbot + k]);
    for (int k = 0; k < nbot; k++)
        res->B_percent[px->bottom_idx[k]] =
            100.0 * fabs(best[R + R * nbot + px->origin * nbot + k]) / qsum;
    double largest = 0.0;
    res->bottom_type = 0.0;
    for (int k = 0; k < nbot; k++)
        if (res->B_percent[px->bottom_idx[k]] > largest) {
            largest = res->B_percent[px->bottom_idx[k]];
            res->bottom_type = 1.0 + px->bottom_idx[k];
        }

    /* IOD in (scene, region, band) order (model/samodel.c:2051-2064) */
    double iod = 0.0, nobs = 0.0;
    for (int s = 0; s < nd; s++)
        for (int r = 0; r < R; r++)
            for (int b = 0; b < px->sb[s]->n_bands; b++) {
                iod += px->rrs_bot[r][s][b] / px->rrs_mod[r][s][b];
                nobs += 1.0;
            }
    res->iod = 100.0 * iod / nobs;
}

void orc_region_eval_outputs(orc_region_pixel *px, const double *x, orc_region_result *res)
{
    double xv[ORC_RMAXN];
    memset(res, 0, sizeof(*res));
    memcpy(xv, x, sizeof(double) * (size_t)orc_region_n_params(px));
    orc_region_objective(xv, px);
    orc_region_outputs(px, xv, res);
}

void orc_region_optimise(orc_region_pixel *px, const orc_nm_cfg *cfg, orc_nelmin_fn nm,
                         orc_region_result *res)
{
    const int n = orc_region_n_params(px);
    double *buf = (double *)malloc(sizeof(double) * 4 * (size_t)n);
    double *start = buf, *step = buf + n, *xmin = buf + 2 * n, *best = buf + 3 * n;
    double lowest = 1.0e4;
    long evals0 = px->n_evals;
    if (!nm)
        nm = orc_nelmin;
    memset(res, 0, sizeof(*res));
    memset(best, 0, sizeof(double) * (size_t)n);
    for (int k = 0; k < cfg->n_hstarts; k++) {
        double err;
        int icount = 0, numres = 0, ifault = 0;
        orc_region_start_step(px, cfg->hstart[k], start, step);
        err = orc_region_objective(start, px); /* the reference evaluates the start first */
        nm(orc_region_objective, px, n, start, xmin, &err, cfg->reqmin, step, cfg->konvge,
           cfg->kcount, &icount, &numres, &ifault);
        res->n_starts++;
        if (err < lowest) {
            lowest = err;
            memcpy(best, xmin, sizeof(double) * (size_t)n);
            res->n_iterations = icount;
            res->converged = (ifault == 0);
            if (lowest < cfg->early_exit)
                break;
        }
    }
    orc_region_objective(best, px);
    orc_region_outputs(px, best, res);
    res->n_evals = px->n_evals - evals0;
    free(buf);
}

/* The depth-error pass of model/samodel.c:1376-1482: per 0.25 m interval up to the deepest
 * depth (at most 30 m), 128 samples, each a pixel drawn with up to sqrt(n_rows n_cols) tries
 * (Philox4x32-10, counter (interval * 128 + sample, try, 300, 0), where the reference calls
 * rand()) whose depth lies strictly inside the interval, perturbed by n_sigma R_sigma with
 * n_sigma = +-u (sign from u_0 < 0.5, counter (.., 0, 301, 0)), and inverted again; the
 * interval's sigma is the standard deviation of the sampled depths (mean rounded to float,
 * zeros skipped, as vec_mean2_double / vec_stddev_double).
 *
 * The starts follow the reference's start_at_previous rule: the first sample inverted (the
 * first in interval and sample order) runs the main pass's start rules; after it the
 * reference sets start_at_previous (:1454), so every later sample is a single start
 * (:2233-2241) from depth_prev -- which the pass itself never updates (it is set only in the
 * main loop, :1069, :1115), so it stays the depth of the main pass's last modelled pixel:
 * h_prev here.  With an empirical depth grid every sample starts from its own empirical
 * depth (the rule then keeps Hstart_quick).  h_prev < 0 instead inverts every sample with the
 * main pass's start rules (the independent-sample variant, for comparison).  depth [npix] the
 * positive retrieved depths (0:
 * unmodelled); emp [npix] the empirical depths or NULL.  Writes table [n_int] and sigma
 * [npix]; returns n_int. */
#define ORC_SIGMA_KEY 0x0000414D47495344ull

static void orc_sigma_uniforms(uint32_t c0, uint32_t c1, uint32_t c2, double *u0, double *u1)
{
    const uint32_t ctr[4] = {c0, c1, c2, 0};
    const uint32_t key[2] = {(uint32_t)ORC_SIGMA_KEY, (uint32_t)(ORC_SIGMA_KEY >> 32)};
    uint32_t o[4];
    orc_philox4x32_10(ctr, key, o);
    *u0 = (double)((((uint64_t)o[0] << 32) | o[1]) >> 11) * 0x1p-53;
    *u1 = (double)((((uint64_t)o[2] << 32) | o[3]) >> 11) * 0x1p-53;
}

int orc_depth_sigma(const float *depth, int n_rows, int n_cols, const float *grid,
                    const float *nodata, const double *rsig, int nd, const orc_bands *bands,
                    const double *sec_sun, const double *sec_view, int n_spatial, int n_bottoms,
                    const float *emp, float emp_nodata, double h_prev, const orc_nm_cfg *nm,
                    double *table, float *sigma)
{
    const long npix = (long)n_rows * n_cols;
    const int nb = bands->n_bands;
    float maxf = -1.0e30f;
    for (long p = 0; p < npix; p++) {
        if (orc_approx_equal(depth[p], 0.0f, 1.0e-4f)) continue;
        if (depth[p] > maxf) maxf = depth[p];
    }
    for (long p = 0; p < npix; p++) sigma[p] = 0.0f;
    if (!(maxf > 0.0f)) return 0;
    double maxd = maxf > 1.0e9f ? 1.0e9 : maxf;
    maxd = 0.25 * ((int)maxd / 0.25);
    if (maxd > 30.0) maxd = 30.0;
    const int n_int = (int)((int)maxd / 0.25);
    if (n_int <= 0) return 0;
    const int n_trials = (int)sqrt((double)n_rows * n_cols);
    const long n_s = (long)n_int * 128;
    double *trial = (double *)calloc((size_t)n_s, sizeof(double));
    long *pick = (long *)malloc(sizeof(long) * (size_t)n_s);

    /* every sample's pixel (or -1: none found in the interval) */
#pragma omp parallel for schedule(dynamic, 16)
    for (long q = 0; q < n_s; q++) {
        const double d = 0.25 * (double)(q / 128);
        pick[q] = -1;
        for (int t = 0; t < n_trials; t++) {
            double u0, u1;
            orc_sigma_uniforms((uint32_t)q, (uint32_t)t, 300, &u0, &u1);
            const int i = (int)(u0 * n_rows), j = (int)(u1 * n_cols);
            const float v = depth[(long)i * n_cols + j];
            if (v > d && v < d + 0.25) {
                pick[q] = (long)i * n_cols + j;
                break;
            }
        }
    }
    /* the first sample the pass inverts (a found pixel with, given empirical depths, one) */
    long first = -1;
    for (long q = 0; q < n_s && first < 0; q++)
        if (pick[q] >= 0 && (!emp || !orc_approx_equal(emp[pick[q]], emp_nodata, 1.0e-6f)))
            first = q;

#pragma omp parallel for schedule(dynamic, 1)
    for (long q = 0; q < n_s; q++) {
        const long pix = pick[q];
        if (pix < 0) continue;
        double us, um;
        orc_sigma_uniforms((uint32_t)q, 0, 301, &us, &um);
        const float n_sigma = us < 0.5 ? (float)um : -(float)um;
        orc_nm_cfg cfg = *nm;
        int nbot = n_bottoms, bidx[3] = {0, 1, 2};
        if (emp) {
            const float e = emp[pix];
            if (orc_approx_equal(e, emp_nodata, 1.0e-6f)) continue;
            const float h = e > -1.0f ? 1.0f : fabsf(e);
            cfg.n_hstarts = 1;
            cfg.hstart[0] = (double)h;
            if (h > 8.0) nbot = 1;
        } else if (q != first && h_prev >= 0.0) {
            cfg.n_hstarts = 1; /* start_at_previous: one start from the main pass's last depth */
            cfg.hstart[0] = h_prev;
        }
        orc_region_pixel *px = (orc_region_pixel *)calloc(1, sizeof(orc_region_pixel));
        orc_region_result *res = (orc_region_result *)calloc(1, sizeof(orc_region_result));
        const int R = orc_region_gather(px, grid, nodata, nd, nb, n_rows, n_cols,
                                        (int)(pix / n_cols), (int)(pix % n_cols), n_spatial,
                                        n_sigma, rsig);
        if (R > 0) {
            orc_region_init(px, nbot, bidx, nd, bands, sec_sun, sec_view);
            orc_region_optimise(px, &cfg, NULL, res);
            trial[q] = res->depth;
        }
        free(px);
        free(res);
    }
    for (int k = 0; k < n_int; k++) {
        const double *v = trial + (long)k * 128;
        double total = 0.0, sumdev = 0.0;
        int count = 0;
        for (int s = 0; s < 128; s++)
            if (!orc_approx_equal((float)v[s], 0.0f, 1.0e-4f)) {
                total += v[s];
                count++;
            }
        const float mean = count == 0 ? 0.0f : (float)(total / (double)count);
        count = 0;
        for (int s = 0; s < 128; s++)
            if (!orc_approx_equal((float)v[s], 0.0f, 1.0e-4f)) {
                sumdev += (v[s] - (double)mean) * (v[s] - (double)mean);
                count++;
            }
        table[k] = count == 0 ? 0.0 : sqrt(sumdev / (double)count);
    }
    free(pick);
    for (long p = 0; p < npix; p++) {
        const float v = depth[p];
        if (!(v > 0.0f)) continue;
        int k = 0;
        for (double d = 0.0; d < maxd; d += 0.25) {
            if (v > d && v <= d + 0.25) {
                sigma[p] = (float)table[k];
                break;
            }
            k++;
        }
    }
    free(trial);
    return n_int;
}
