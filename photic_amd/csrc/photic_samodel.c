/* photic_amd -- C host code: the drop-in samodel() of include/photic_samodel.h.
 *
 * Grid in, grid out, like model/samodel.c:289-1763: the caller's row-pointer band grids are
 * screened for no-data on every scene (model/samodel.c:933-947), smoothed as
 * model/common.c:232-276 does when n_smoothing_radius > 1, packed into planar pixel arrays per
 * depth-start group, inverted on the GPU -- one scene by photic_invert_h, up to eight (the
 * reference's N_s = nscenes run, model/samodel.c:1808-1815) by photic_invert_stack_h with
 * each scene's sun and view angles -- and the outputs scattered back into the caller's grids
 * with the reference's defaults and sign conventions.  With n_spatial > 1 each pixel is
 * inverted jointly with its (2 n_spatial - 1)^2 neighbours (model/samodel.c:2957-3027) by
 * photic_invert_regions, on the smoothed band grids held on the device; so is every pixel of
 * a run of more than eight scenes, with one region.  When the caller
 * passes a depth_sigma grid, the depth-error pass of model/samodel.c:1376-1482 fills it
 * (sm_sigma_pass). */
#define _POSIX_C_SOURCE 199309L /* clock_gettime */
#include "photic_samodel.h"

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "photic.h"
#include "ph_launch.h"

#define SM_MAX_SCENES PH_MAXS /* the reference's MAX_SCENES (model/common.h:46) */

/* PHOTIC_TIMING set in the environment: the drop-in prints the wall time of its phases to
 * stderr (host gather, device grids, inversion with the scatter, depth-error pass). */
static double sm_ms(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return 1.0e3 * (double)t.tv_sec + 1.0e-6 * (double)t.tv_nsec;
}

static int sm_feq(float a, float b, float eps)
{
    float d = a - b;
    double fa = fabs((double)a), fb = fabs((double)b);
    return fabs((double)d) <= ((fa < fb ? fb : fa) * (double)eps);
}

/* smoothed_index_2d (model/common.c:232-276): (2r-1)^2 box mean with clamp-to-edge indexing,
 * skipping no-data; all no-data gives the no-data value. */
static float sm_smoothed(float **a, int i, int j, int nrows, int ncols, int r, float spval)
{
    float sum = 0.0f, nobs = 0.0f;
    for (int ii = 1 - r; ii < r; ii++) {
        int iii = i + ii < 0 ? 0 : (i + ii > nrows - 1 ? nrows - 1 : i + ii);
        for (int jj = 1 - r; jj < r; jj++) {
            int jjj = j + jj < 0 ? 0 : (j + jj > ncols - 1 ? ncols - 1 : j + jj);
            float v = a[iii][jjj];
            if (!sm_feq(v, spval, 1.0e-6f)) {
                sum += v;
                nobs += 1.0f;
            }
        }
    }
    return nobs < 0.5f ? spval : sum / nobs;
}

typedef struct {
    long n;
    long *idx;      /* pixel index i * ncols + j */
    float *rrs;     /* [rows][n]: the scenes' band values one after another */
    float *hstart;  /* [n] or NULL */
} sm_group;

/* The scenes' band sets: scene d has nb[d] bands at wl[off[d] .. off[d] + nb[d]), the band
 * grids of all scenes one after another (rows in all). */
typedef struct {
    int rows;
    int nb[SM_MAX_SCENES], off[SM_MAX_SCENES];
    double wl[SM_MAX_SCENES * PH_MAXB];
} sm_bands;

/* nd > 1: the scenes' sec(sun) and sec(view), one value per scene */
static int sm_run_group(const photic_cfg *cfg, sm_group *g, int nd, const double *sec_sun,
                        const double *sec_view, int ncols, float **depth,
                        float **model_error, float **bottom_albedo, float **bottom_sand,
                        float **bottom_seagrass, float **bottom_coral, float **K_min,
                        float **bottom_type, float **index_optical_depth)
{
    if (g->n == 0) return PHOTIC_OK;
    const long n = g->n;
    float *d_rrs = NULL, *d_h = NULL, *d_out = NULL, *d_geo = NULL, *geo = NULL;
    unsigned char *d_valid = NULL, *valid = NULL;
    int rc = photic_device_malloc((void **)&d_rrs, sizeof(float) * 4 * nd * n);
    if (!rc && g->hstart) rc = photic_device_malloc((void **)&d_h, sizeof(float) * n);
    /* outputs: depth, rrs_error, albedo, k_min, iod, bottom_type, b_pct[3] */
    if (!rc) rc = photic_device_malloc((void **)&d_out, sizeof(float) * 9 * n);
    if (!rc) rc = photic_memcpy_htod(d_rrs, g->rrs, sizeof(float) * 4 * nd * n);
    if (!rc && d_h) rc = photic_memcpy_htod(d_h, g->hstart, sizeof(float) * n);
    if (!rc && nd > 1) {
        /* every scene valid (a pixel with no data on any scene was not gathered); the
         * scenes' geometry as [nd][n] sec(sun) then [nd][n] sec(view) */
        geo = (float *)malloc(sizeof(float) * 2 * nd * n);
        valid = (unsigned char *)malloc((size_t)nd * n);
        if (!geo || !valid) rc = PHOTIC_ENOMEM;
        for (long q = 0; !rc && q < (long)nd * n; q++) {
            geo[q] = (float)sec_sun[q / n];
            geo[(long)nd * n + q] = (float)sec_view[q / n];
            valid[q] = 1;
        }
        if (!rc) rc = photic_device_malloc((void **)&d_geo, sizeof(float) * 2 * nd * n);
        if (!rc) rc = photic_device_malloc((void **)&d_valid, (size_t)nd * n);
        if (!rc) rc = photic_memcpy_htod(d_geo, geo, sizeof(float) * 2 * nd * n);
        if (!rc) rc = photic_memcpy_htod(d_valid, valid, (size_t)nd * n);
    }
    if (!rc) {
        photic_outputs o;
        memset(&o, 0, sizeof(o));
        o.depth = d_out;
        o.rrs_error = d_out + n;
        o.albedo = d_out + 2 * n;
        o.k_min = d_out + 3 * n;
        o.iod = d_out + 4 * n;
        o.bottom_type = d_out + 5 * n;
        o.b_pct = d_out + 6 * n;
        if (nd == 1)
            rc = photic_invert_h(cfg, n, d_rrs, NULL, NULL, NULL, d_h, &o, NULL);
        else
            rc = photic_invert_stack_h(cfg, n, nd, d_rrs, d_geo, d_geo + (long)nd * n, d_valid,
                                       d_h, &o, NULL);
        if (!rc) rc = photic_synchronize();
    }
    float *h = NULL;
    if (!rc) {
        h = (float *)malloc(sizeof(float) * 9 * n);
        if (!h) rc = PHOTIC_ENOMEM;
    }
    if (!rc) rc = photic_memcpy_dtoh(h, d_out, sizeof(float) * 9 * n);
    if (!rc) {
#pragma omp parallel for schedule(static)
        for (long p = 0; p < n; p++) {
            long k = g->idx[p];
            int i = (int)(k / ncols), j = (int)(k % ncols);
            depth[i][j] = -h[p]; /* negative below the surface, model/samodel.c:1486-1490 */
            model_error[i][j] = h[n + p];
            bottom_albedo[i][j] = h[2 * n + p];
            K_min[i][j] = h[3 * n + p];
            index_optical_depth[i][j] = h[4 * n + p];
            bottom_type[i][j] = h[5 * n + p];
            bottom_sand[i][j] = h[6 * n + p];
            bottom_seagrass[i][j] = h[7 * n + p];
            bottom_coral[i][j] = h[8 * n + p];
        }
    }
    free(h);
    free(geo);
    free(valid);
    photic_device_free(d_rrs);
    if (d_h) photic_device_free(d_h);
    if (d_out) photic_device_free(d_out);
    if (d_geo) photic_device_free(d_geo);
    if (d_valid) photic_device_free(d_valid);
    return rc;
}

/* Inverts the pixels idx[0..n-1] with their neighbourhoods on the smoothed grids held on the
 * device (ph_invert_regions_pert; pert [n] host n_sigma per pixel, or NULL); h receives the
 * host outputs [9][n] (depth, rrs_error, albedo, k_min, iod, bottom_type, b_pct[3]) and st
 * [n] the status. */
static int sm_invert_regions(const photic_cfg *cfg, const sm_bands *bd, const float *d_grid,
                             const float *nodata,
                             int nrows, int ncols, int nd, const double *sec_sun,
                             const double *sec_view, int n_spatial, long n, const long *idx,
                             const float *hstart, const float *pert, const double *rsig,
                             float *h, int *st, double *depth64)
{
    int64_t *d_idx = NULL;
    float *d_h = NULL, *d_out = NULL, *d_pert = NULL;
    double *d_d64 = NULL;
    int *d_st = NULL;
    int64_t *ix = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    int rc = ix ? PHOTIC_OK : PHOTIC_ENOMEM;
    for (long q = 0; !rc && q < n; q++) ix[q] = (int64_t)idx[q];
    if (!rc) rc = photic_device_malloc((void **)&d_idx, sizeof(int64_t) * (size_t)n);
    if (!rc && hstart) rc = photic_device_malloc((void **)&d_h, sizeof(float) * (size_t)n);
    if (!rc && pert) rc = photic_device_malloc((void **)&d_pert, sizeof(float) * (size_t)n);
    if (!rc) rc = photic_device_malloc((void **)&d_out, sizeof(float) * 9 * (size_t)n);
    if (!rc) rc = photic_device_malloc((void **)&d_st, sizeof(int) * (size_t)n);
    if (!rc && depth64) rc = photic_device_malloc((void **)&d_d64, sizeof(double) * (size_t)n);
    if (!rc) rc = photic_memcpy_htod(d_idx, ix, sizeof(int64_t) * (size_t)n);
    if (!rc && d_h) rc = photic_memcpy_htod(d_h, hstart, sizeof(float) * (size_t)n);
    if (!rc && d_pert) rc = photic_memcpy_htod(d_pert, pert, sizeof(float) * (size_t)n);
    /* every pixel starts at the reference's defaults with status -1, so a pixel the kernel
     * never reaches reads as not modelled */
    for (long q = 0; !rc && q < n; q++) {
        for (int k = 0; k < 5; k++) h[(long)k * n + q] = 0.0f;
        for (int k = 5; k < 9; k++) h[(long)k * n + q] = -9999.0f;
        st[q] = -1;
    }
    if (!rc) rc = photic_memcpy_htod(d_out, h, sizeof(float) * 9 * (size_t)n);
    if (!rc) rc = photic_memcpy_htod(d_st, st, sizeof(int) * (size_t)n);
    if (!rc) {
        photic_outputs o;
        memset(&o, 0, sizeof(o));
        o.depth = d_out;
        o.rrs_error = d_out + n;
        o.albedo = d_out + 2 * n;
        o.k_min = d_out + 3 * n;
        o.iod = d_out + 4 * n;
        o.bottom_type = d_out + 5 * n;
        o.b_pct = d_out + 6 * n;
        o.status = d_st;
        rc = ph_invert_regions_bands(cfg, nrows, ncols, nd, bd->nb, bd->wl, d_grid, nodata,
                                     sec_sun, sec_view, n_spatial, n, d_idx, d_h, d_pert, rsig,
                                     &o, NULL, d_d64, NULL);
        if (!rc) rc = photic_synchronize();
    }
    if (!rc) rc = photic_memcpy_dtoh(h, d_out, sizeof(float) * 9 * (size_t)n);
    if (!rc) rc = photic_memcpy_dtoh(st, d_st, sizeof(int) * (size_t)n);
    if (!rc && depth64) rc = photic_memcpy_dtoh(depth64, d_d64, sizeof(double) * (size_t)n);
    free(ix);
    if (d_d64) photic_device_free(d_d64);
    if (d_idx) photic_device_free(d_idx);
    if (d_h) photic_device_free(d_h);
    if (d_pert) photic_device_free(d_pert);
    if (d_out) photic_device_free(d_out);
    if (d_st) photic_device_free(d_st);
    return rc;
}

/* n_spatial > 1: each depth-start group's pixel list inverted with its neighbourhoods, the
 * outputs scattered into the caller's grids. */
static int sm_run_regions(const photic_cfg *cfg, const sm_bands *bd, const float *d_grid,
                          const float *nodata,
                          int nrows, int ncols, int nd, const double *sec_sun,
                          const double *sec_view, int n_spatial, sm_group *g, float **depth,
                          float **model_error, float **bottom_albedo, float **bottom_sand,
                          float **bottom_seagrass, float **bottom_coral, float **K_min,
                          float **bottom_type, float **index_optical_depth)
{
    if (g->n == 0) return PHOTIC_OK;
    const long n = g->n;
    float *h = (float *)malloc(sizeof(float) * 9 * (size_t)n);
    int *st = (int *)malloc(sizeof(int) * (size_t)n);
    int rc = (h && st) ? PHOTIC_OK : PHOTIC_ENOMEM;
    if (!rc)
        rc = sm_invert_regions(cfg, bd, d_grid, nodata, nrows, ncols, nd, sec_sun, sec_view,
                               n_spatial, n, g->idx, g->hstart, NULL, NULL, h, st, NULL);
    for (long p = 0; !rc && p < n; p++) {
        if (st[p] < 0) continue; /* no region left: the defaults stay */
        long k = g->idx[p];
        int i = (int)(k / ncols), j = (int)(k % ncols);
        depth[i][j] = -h[p]; /* negative below the surface, model/samodel.c:1486-1490 */
        model_error[i][j] = h[n + p];
        bottom_albedo[i][j] = h[2 * n + p];
        K_min[i][j] = h[3 * n + p];
        index_optical_depth[i][j] = h[4 * n + p];
        bottom_type[i][j] = h[5 * n + p];
        bottom_sand[i][j] = h[6 * n + p];
        bottom_seagrass[i][j] = h[7 * n + p];
        bottom_coral[i][j] = h[8 * n + p];
    }
    free(h);
    free(st);
    return rc;
}

/* ---- the depth-error estimate (model/samodel.c:1376-1482) ----
 * Per 0.25 m depth interval up to the deepest retrieved depth (at most 30 m): 128 samples, each
 * a pixel drawn at random (up to sqrt(n_rows n_cols) tries) whose depth lies in the interval,
 * its reflectances perturbed by n_sigma R_sigma with n_sigma uniform in (-1, 1), and inverted
 * again; the interval's sigma is the standard deviation of the sampled depths, and every
 * modelled pixel gets its interval's sigma.  The starts are the reference's: the first sample
 * inverted runs the main pass's start rules, every later one a single start (start_at_previous,
 * :1454, :2233-2241) from depth_prev, which the pass never updates -- the depth of the main
 * pass's last modelled pixel (the reference's loop order; row-major here) -- unless an
 * empirical depth grid gives each sample its own start.  So the samples are independent of
 * one another and are inverted in one batched call.  The reference draws with rand() seeded
 * from the clock; here the draws come from Philox4x32-10 with a fixed key. */
#define SM_SIGMA_SAMPLES 128
#define SM_SIGMA_INTERVAL 0.25
#define SM_SIGMA_KEY 0x0000414D47495344ull /* the key of the sample draws */

static void sm_philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                      uint32_t k1, uint32_t out[4])
{
    for (int rr = 0; rr < 10; rr++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

/* Two uniforms in [0, 1) from counter (c0, c1, c2, 0) (53 bits each). */
static void sm_uniforms(uint32_t c0, uint32_t c1, uint32_t c2, double *u0, double *u1)
{
    uint32_t o[4];
    sm_philox(c0, c1, c2, 0, (uint32_t)SM_SIGMA_KEY, (uint32_t)(SM_SIGMA_KEY >> 32), o);
    *u0 = (double)((((uint64_t)o[0] << 32) | o[1]) >> 11) * 0x1p-53;
    *u1 = (double)((((uint64_t)o[2] << 32) | o[3]) >> 11) * 0x1p-53;
}

static int sm_sigma_pass(const photic_cfg *cfg, const sm_bands *bd,
                         const struct photic_scene *scene_data,
                         const int *scene_indexes, int nd, int nrows, int ncols, int n_spatial,
                         int empirical_depth_present, struct photic_geogrid empirical_depths,
                         const double *sec_sun, const double *sec_view, const float *d_grid,
                         const float *nodata, float **depth, float **depth_sigma)
{
    /* the deepest retrieved depth (the grid holds -H; unmodelled pixels 0), array_max2 */
    float maxf = -1.0e30f;
    for (int i = 0; i < nrows; i++)
        for (int j = 0; j < ncols; j++) {
            float v = -depth[i][j];
            if (sm_feq(v, 0.0f, 1.0e-4f)) continue;
            if (v > maxf) maxf = v;
        }
    if (!(maxf > 0.0f)) return PHOTIC_OK;
    const double t_start = sm_ms();
    double maxd = maxf > 1.0e9f ? 1.0e9 : maxf; /* only an int's range matters below */
    maxd = SM_SIGMA_INTERVAL * ((int)maxd / SM_SIGMA_INTERVAL);
    if (maxd > 30.0) maxd = 30.0;
    const int n_int = (int)((int)maxd / SM_SIGMA_INTERVAL);
    if (n_int <= 0) return PHOTIC_OK;
    const int n_trials = (int)sqrt((double)nrows * ncols);
    const long n_s = (long)n_int * SM_SIGMA_SAMPLES;

    long *pix = (long *)malloc(sizeof(long) * (size_t)n_s);
    float *ns = (float *)malloc(sizeof(float) * (size_t)n_s);
    int *grp = (int *)malloc(sizeof(int) * (size_t)n_s);
    float *hs = (float *)malloc(sizeof(float) * (size_t)n_s);
    double *trial = (double *)calloc((size_t)n_s, sizeof(double));
    double *table = (double *)calloc((size_t)n_int, sizeof(double));
    long *gidx = (long *)malloc(sizeof(long) * (size_t)n_s);
    long *gpos = (long *)malloc(sizeof(long) * (size_t)n_s);
    float *gns = (float *)malloc(sizeof(float) * (size_t)n_s);
    float *ghs = (float *)malloc(sizeof(float) * (size_t)n_s);
    float *h = (float *)malloc(sizeof(float) * 9 * (size_t)n_s);
    int *st = (int *)malloc(sizeof(int) * (size_t)n_s);
    double *d64 = (double *)malloc(sizeof(double) * (size_t)n_s);
    int rc = (pix && ns && grp && hs && trial && table && gidx && gpos && gns && ghs && h && st &&
              d64) ? PHOTIC_OK : PHOTIC_ENOMEM;
    double rsig[SM_MAX_SCENES * PH_MAXB];
    for (int d = 0; d < nd; d++)
        for (int b = 0; b < bd->nb[d]; b++)
            rsig[bd->off[d] + b] = scene_data[scene_indexes[d]].R_sigma[b];

    /* the samples: a pixel in the interval, its n_sigma, its depth-start group (each sample
     * from its own Philox counters, so the samples are independent of one another) */
    if (!rc) {
#pragma omp parallel for schedule(dynamic, 8)
        for (long q = 0; q < n_s; q++) {
            const int k = (int)(q / SM_SIGMA_SAMPLES);
            const double d = SM_SIGMA_INTERVAL * k;
            const uint32_t ctr = (uint32_t)q;
            pix[q] = -1;
            for (int t = 0; t < n_trials; t++) {
                double u0, u1;
                sm_uniforms(ctr, (uint32_t)t, 300, &u0, &u1);
                const int i = (int)(u0 * nrows), j = (int)(u1 * ncols);
                const float v = -depth[i][j];
                if (v > d && v < d + SM_SIGMA_INTERVAL) {
                    pix[q] = (long)i * ncols + j;
                    break;
                }
            }
            if (pix[q] < 0) continue;
            double us, um;
            sm_uniforms(ctr, 0, 301, &us, &um);
            ns[q] = us < 0.5 ? (float)um : -(float)um;
            grp[q] = 0;
            if (empirical_depth_present) {
                const long p = pix[q];
                const float e = empirical_depths.array[p / ncols][p % ncols];
                if (sm_feq(e, empirical_depths.nodata_value, 1.0e-6f)) {
                    pix[q] = -1; /* the reference skips such a sample */
                    continue;
                }
                hs[q] = e > -1.0f ? 1.0f : fabsf(e);
                grp[q] = hs[q] > 8.0 ? 2 : 1;
            }
        }
    }
    /* start_at_previous: after the first sample inverted, each sample without an empirical
     * depth starts once, from the main pass's last modelled depth (row-major) */
    if (!rc && !empirical_depth_present) {
        float h_prev = 0.0f;
        for (long p = (long)nrows * ncols - 1; p >= 0; p--) {
            const float v = -depth[p / ncols][p % ncols];
            if (v > 0.0f) {
                h_prev = v;
                break;
            }
        }
        long first = -1;
        for (long q = 0; q < n_s; q++) {
            if (pix[q] < 0) continue;
            if (first < 0) {
                first = q;
                continue;
            }
            hs[q] = h_prev;
            grp[q] = 1;
        }
    }
    const double t_sampled = sm_ms();
    /* one batched inversion per depth-start group */
    for (int k = 0; !rc && k < 3; k++) {
        long m = 0;
        for (long q = 0; q < n_s; q++)
            if (pix[q] >= 0 && grp[q] == k) {
                gidx[m] = pix[q];
                gpos[m] = q;
                gns[m] = ns[q];
                ghs[m] = hs[q];
                m++;
            }
        if (m == 0) continue;
        photic_cfg c = *cfg;
        if (k == 2) {
            c.n_bottoms = 1;
            c.bottom_idx[0] = 0; /* sand */
        }
        rc = sm_invert_regions(&c, bd, d_grid, nodata, nrows, ncols, nd, sec_sun, sec_view,
                               n_spatial < 1 ? 1 : n_spatial, m, gidx, k ? ghs : NULL, gns, rsig,
                               h, st, d64);
        for (long p = 0; !rc && p < m; p++) trial[gpos[p]] = st[p] >= 0 ? d64[p] : 0.0;
    }
    if (getenv("PHOTIC_TIMING"))
        fprintf(stderr, "photic_samodel: depth-error pass: %d intervals, %ld samples, sampling "
                        "%.1f ms, inversion %.1f ms\n",
                n_int, n_s, t_sampled - t_start, sm_ms() - t_sampled);
    /* per interval: the mean (vec_mean2_double, returned as a float) and the standard
     * deviation (vec_stddev_double) of the sampled depths, zeros skipped */
    for (int k = 0; !rc && k < n_int; k++) {
        const double *v = trial + (long)k * SM_SIGMA_SAMPLES;
        double total = 0.0, sumdev = 0.0;
        int count = 0;
        for (int s = 0; s < SM_SIGMA_SAMPLES; s++)
            if (!sm_feq((float)v[s], 0.0f, 1.0e-4f)) {
                total += v[s];
                count++;
            }
        const float mean = count == 0 ? 0.0f : (float)(total / (double)count);
        count = 0;
        for (int s = 0; s < SM_SIGMA_SAMPLES; s++)
            if (!sm_feq((float)v[s], 0.0f, 1.0e-4f)) {
                const double e = v[s] - (double)mean;
                sumdev += e * e;
                count++;
            }
        table[k] = count == 0 ? 0.0 : sqrt(sumdev / (double)count);
    }
    /* every modelled pixel gets its interval's sigma: the k with 0.25 k < H <= 0.25 (k + 1),
     * k < n_int, i.e. k = ceil(4 H) - 1 (4 H and the interval ends are exact in double) */
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nrows; i++)
        for (int j = 0; !rc && j < ncols; j++) {
            const float v = -depth[i][j];
            depth_sigma[i][j] = 0.0f;
            if (!(v > 0.0f)) continue;
            const double k = ceil(4.0 * (double)v) - 1.0;
            if (k < (double)n_int) depth_sigma[i][j] = (float)table[(int)k];
        }
    free(pix);
    free(ns);
    free(grp);
    free(hs);
    free(trial);
    free(table);
    free(gidx);
    free(gpos);
    free(gns);
    free(ghs);
    free(h);
    free(st);
    free(d64);
    return rc;
}

int photic_samodel(struct photic_scene scene_data[], struct photic_geogrid gridded_data[],
                   int *scene_indexes, int nscenes, int empirical_depth_present,
                   struct photic_geogrid empirical_depths, int n_smoothing_radius,
                   int n_spatial, int n_bottoms, float **depth, float **depth_sigma,
                   float **model_error, float **bottom_albedo, float **bottom_sand,
                   float **bottom_seagrass, float **bottom_coral, float **K_min,
                   float **bottom_type, float **index_optical_depth)
{
    if (!scene_data || !gridded_data || !scene_indexes || nscenes < 1) return PHOTIC_EINVAL;
    const struct photic_scene *sc = &scene_data[scene_indexes[0]];
    const int nrows = sc->nrows, ncols = sc->ncols, nb = sc->n_bands;

    /* the reference's defaults for every pixel (model/samodel.c:819-841) */
    for (int i = 0; i < nrows; i++)
        for (int j = 0; j < ncols; j++) {
            depth[i][j] = 0.0f;
            if (depth_sigma) depth_sigma[i][j] = 0.0f;
            model_error[i][j] = 0.0f;
            bottom_albedo[i][j] = 0.0f;
            bottom_sand[i][j] = -9999.0f;
            bottom_seagrass[i][j] = -9999.0f;
            bottom_coral[i][j] = -9999.0f;
            K_min[i][j] = 0.0f;
            bottom_type[i][j] = -9999.0f;
            index_optical_depth[i][j] = 0.0f;
        }
    if (nscenes > SM_MAX_SCENES) {
        fprintf(stderr, "photic samodel: at most %d scenes (got %d); outputs left at their "
                        "defaults\n", SM_MAX_SCENES, nscenes);
        return PHOTIC_EUNSUPPORTED;
    }
    /* each scene's band set (model/samodel.c:384-527): 2..8 bands at ascending wavelengths,
     * every scene on the first one's grid */
    sm_bands bd;
    memset(&bd, 0, sizeof(bd));
    int bands_ok = 1, same_bands = 1;
    for (int k = 0; k < nscenes; k++) {
        const struct photic_scene *t = &scene_data[scene_indexes[k]];
        if (t->n_bands < 2 || t->n_bands > PH_MAXB || t->nrows != nrows || t->ncols != ncols) {
            bands_ok = 0;
            break;
        }
        bd.nb[k] = t->n_bands;
        bd.off[k] = bd.rows;
        for (int b = 0; b < t->n_bands; b++) {
            bd.wl[bd.rows + b] = (double)t->wavelengths[b];
            if (b > 0 && !(t->wavelengths[b] > t->wavelengths[b - 1])) bands_ok = 0;
        }
        bd.rows += t->n_bands;
        if (t->n_bands != 4 || t->n_bands != nb) same_bands = 0;
        for (int b = 0; same_bands && b < 4; b++)
            if (t->wavelengths[b] != sc->wavelengths[b]) same_bands = 0;
    }
    const int nsr = n_spatial < 1 ? 1 : n_spatial;
    const int n_reg = (2 * nsr - 1) * (2 * nsr - 1);
    const int n_par = n_reg * (1 + 2 * n_bottoms) + 3 * nscenes;
    if (!bands_ok || n_spatial < 0 || n_bottoms < 1 || n_bottoms > 3 || n_reg > PH_REG_MAXR ||
        n_par > PH_REG_MAXN) {
        fprintf(stderr, "photic samodel: supported are 1..%d scenes of one grid with 2..%d "
                        "ascending bands each, 1..3 bottoms, n_spatial <= 4 and at most %d "
                        "parameters (got %d scenes, n_spatial %d, %d bottoms); outputs left "
                        "at their defaults\n",
                SM_MAX_SCENES, PH_MAXB, PH_REG_MAXN, nscenes, n_spatial, n_bottoms);
        return PHOTIC_EUNSUPPORTED;
    }
    /* one scene: the single-date kernel; up to PH_MAXD scenes: the stack kernel -- both for
     * scenes sharing the 4-band set they are compiled for; more scenes, other band sets or
     * any neighbourhood: the region kernel (one region when n_spatial <= 1) */
    const int use_regions = n_spatial > 1 || nscenes > PH_MAXD || !same_bands;
    const int nd = nscenes;
    const int rows = bd.rows;
    const int r = n_smoothing_radius < 1 ? 1 : n_smoothing_radius;

    photic_cfg cfg;
    photic_cfg_default(&cfg, bd.nb[0], bd.wl); /* the reference's Nelder-Mead settings */
    cfg.mode = PHOTIC_MODE_NATIVE;
    cfg.n_bottoms = n_bottoms;
    /* the pixels are screened below on their raw values (model/samodel.c:933-947); the
     * kernels invert the (smoothed) values they are given without a second screen */
    cfg.prescreened = 1;
    /* a kernel wavefront that finishes no pixel within the watchdog's limit (600 s, or
     * PHOTIC_WATCHDOG_S seconds from the environment) gives up; the call then fails and no
     * output of the failed inversion reaches the caller's grids */
    cfg.check_watchdog = 1;
    const char *wd = getenv("PHOTIC_WATCHDOG_S");
    if (wd && atof(wd) > 0.0) cfg.watchdog_s = atof(wd);
    /* each scene's subsurface sun and view angles, as secants (model/samodel.c:534-548) */
    double sec_sun[SM_MAX_SCENES], sec_view[SM_MAX_SCENES];
    for (int k = 0; k < nd; k++) {
        const struct photic_scene *t = &scene_data[scene_indexes[k]];
        double a = t->theta_w;
        a *= 3.141592653589793 / 180.0;
        sec_sun[k] = 1.0 / cos(a);
        a = t->theta_v;
        a *= 3.141592653589793 / 180.0;
        sec_view[k] = 1.0 / cos(a);
    }
    cfg.sec_sun = sec_sun[0];
    cfg.sec_view = sec_view[0];

    /* groups: 0 = no empirical depth (all depth starts), 1 = empirical depth <= 8 m, 2 =
     * empirical depth > 8 m (sand only), model/samodel.c:1826, :2208-2231, :960-976 */
    long npix = (long)nrows * ncols;
    sm_group gr[3];
    memset(gr, 0, sizeof(gr));
    const double t0 = sm_ms();
    int rc = PHOTIC_OK;
    for (int k = 0; k < 3 && !rc; k++) {
        gr[k].idx = (long *)malloc(sizeof(long) * (size_t)(npix ? npix : 1));
        gr[k].rrs = (float *)malloc(sizeof(float) * rows * (size_t)(npix ? npix : 1));
        gr[k].hstart = k ? (float *)malloc(sizeof(float) * (size_t)(npix ? npix : 1)) : NULL;
        if (!gr[k].idx || !gr[k].rrs || (k && !gr[k].hstart)) rc = PHOTIC_ENOMEM;
    }
    /* pixel screening over every scene and gather, per group, in row-major order: each row
     * screens its pixels into a row buffer (OpenMP over rows), the per-row counts give every
     * row its place in each group, and the rows are then copied there (rows [scene][band] of
     * each group at stride npix, compacted to n below) */
    signed char *code = NULL;
    float *hrow = NULL, *vrow = NULL;
    long *cnt = NULL;
    if (!rc) {
        code = (signed char *)malloc((size_t)(npix ? npix : 1));
        hrow = (float *)malloc(sizeof(float) * (size_t)(npix ? npix : 1));
        vrow = (float *)malloc(sizeof(float) * rows * (size_t)(npix ? npix : 1));
        cnt = (long *)calloc(3 * (size_t)nrows + 3, sizeof(long));
        if (!code || !hrow || !vrow || !cnt) rc = PHOTIC_ENOMEM;
    }
    if (!rc) {
#pragma omp parallel for schedule(dynamic, 4)
        for (int i = 0; i < nrows; i++)
            for (int j = 0; j < ncols; j++) {
                const long p = (long)i * ncols + j;
                int bad = 0;
                for (int d = 0; d < nd; d++) {
                    const struct photic_scene *t = &scene_data[scene_indexes[d]];
                    for (int b = 0; b < bd.nb[d]; b++) {
                        const struct photic_geogrid *g = &gridded_data[t->band_indexes[b]];
                        float refl = g->array[i][j];
                        if (sm_feq(refl, g->nodata_value, 1.0e-6f) || refl < 0.0f) bad = 1;
                        float sv = r > 1 ? sm_smoothed(g->array, i, j, nrows, ncols, r,
                                                       g->nodata_value)
                                         : refl;
                        /* one region: a smoothed no-data value drops it (with several, the
                         * region gather on the device decides) */
                        if (n_spatial <= 1 && sm_feq(sv, g->nodata_value, 1.0e-6f)) bad = 1;
                        vrow[(long)(bd.off[d] + b) * npix + p] = sv;
                    }
                }
                code[p] = -1;
                if (bad) continue;
                int k = 0;
                float h = 0.0f;
                if (empirical_depth_present &&
                    !sm_feq(empirical_depths.array[i][j], empirical_depths.nodata_value,
                            1.0e-6f)) {
                    float e = empirical_depths.array[i][j];
                    h = e > -1.0f ? 1.0f : fabsf(e);
                    k = h > 8.0 ? 2 : 1;
                }
                code[p] = (signed char)k;
                hrow[p] = h;
                cnt[3 * (long)(i + 1) + k]++;
            }
        /* cnt[3 (i + 1) + k]: row i's pixels in group k; prefix sums give row i's start */
        for (int i = 0; i < nrows; i++)
            for (int k = 0; k < 3; k++) cnt[3 * (long)(i + 1) + k] += cnt[3 * (long)i + k];
        for (int k = 0; k < 3; k++) gr[k].n = cnt[3 * (long)nrows + k];
#pragma omp parallel for schedule(static)
        for (int i = 0; i < nrows; i++) {
            long at[3] = {cnt[3 * (long)i], cnt[3 * (long)i + 1], cnt[3 * (long)i + 2]};
            for (int j = 0; j < ncols; j++) {
                const long p = (long)i * ncols + j;
                const int k = code[p];
                if (k < 0) continue;
                sm_group *g = &gr[k];
                const long m = at[k]++;
                g->idx[m] = p;
                for (int q = 0; q < rows; q++)
                    g->rrs[(long)q * npix + m] = vrow[(long)q * npix + p];
                if (k) g->hstart[m] = hrow[p];
            }
        }
    }
    free(code);
    free(hrow);
    free(vrow);
    free(cnt);
    /* compact the planar (scene, band) rows to n per group */
    for (int k = 0; k < 3 && !rc; k++)
        for (int q = 1; q < rows; q++)
            memmove(gr[k].rrs + (long)q * gr[k].n, gr[k].rrs + (long)q * npix,
                    sizeof(float) * (size_t)gr[k].n);

    /* every (scene, band) grid, smoothed, and its no-data value, on the device for the
     * neighbourhood gather (n_spatial > 1) and the depth-error pass */
    float *grid = NULL, *d_grid = NULL, nod[SM_MAX_SCENES * PH_MAXB];
    const double t1 = sm_ms();
    if (!rc && (use_regions || depth_sigma)) {
        grid = (float *)malloc(sizeof(float) * rows * (size_t)(npix ? npix : 1));
        if (!grid) rc = PHOTIC_ENOMEM;
        for (int d = 0; !rc && d < nd; d++)
            for (int b = 0; b < bd.nb[d]; b++) {
                const struct photic_geogrid *g =
                    &gridded_data[scene_data[scene_indexes[d]].band_indexes[b]];
                float *dst = grid + (long)(bd.off[d] + b) * npix;
                nod[bd.off[d] + b] = g->nodata_value;
                for (int i = 0; i < nrows; i++)
                    for (int j = 0; j < ncols; j++)
                        dst[(long)i * ncols + j] =
                            r > 1 ? sm_smoothed(g->array, i, j, nrows, ncols, r, g->nodata_value)
                                  : g->array[i][j];
            }
        if (!rc) rc = photic_device_malloc((void **)&d_grid, sizeof(float) * rows * npix);
        if (!rc) rc = photic_memcpy_htod(d_grid, grid, sizeof(float) * rows * npix);
    }
    const double t2 = sm_ms();
    for (int k = 0; k < 3 && !rc; k++) {
        photic_cfg c = cfg;
        if (k == 2) {
            c.n_bottoms = 1;
            c.bottom_idx[0] = 0; /* sand */
        }
        if (use_regions)
            rc = sm_run_regions(&c, &bd, d_grid, nod, nrows, ncols, nd, sec_sun, sec_view,
                                n_spatial,
                                &gr[k], depth, model_error, bottom_albedo, bottom_sand,
                                bottom_seagrass, bottom_coral, K_min, bottom_type,
                                index_optical_depth);
        else
            rc = sm_run_group(&c, &gr[k], nd, sec_sun, sec_view, ncols, depth, model_error,
                              bottom_albedo, bottom_sand, bottom_seagrass, bottom_coral, K_min,
                              bottom_type, index_optical_depth);
    }
    const double t3 = sm_ms();
    if (!rc && depth_sigma)
        rc = sm_sigma_pass(&cfg, &bd, scene_data, scene_indexes, nd, nrows, ncols, n_spatial,
                           empirical_depth_present, empirical_depths, sec_sun, sec_view, d_grid,
                           nod, depth, depth_sigma);
    if (getenv("PHOTIC_TIMING"))
        fprintf(stderr, "photic_samodel: gather %.1f ms, device grids %.1f ms, inversion and "
                        "scatter %.1f ms, depth-error pass %.1f ms\n",
                t1 - t0, t2 - t1, t3 - t2, sm_ms() - t3);
    free(grid);
    if (d_grid) photic_device_free(d_grid);
    for (int k = 0; k < 3; k++) {
        free(gr[k].idx);
        free(gr[k].rrs);
        free(gr[k].hstart);
    }
    if (rc)
        fprintf(stderr, "photic samodel: inversion failed (%d): %s\n", rc, photic_last_error());
    return rc;
}

void samodel(struct photic_scene scene_data[], struct photic_geogrid gridded_data[],
             int *scene_indexes, int nscenes, int empirical_depth_present,
             struct photic_geogrid empirical_depths, int n_smoothing_radius, int n_spatial,
             int n_bottoms, float **depth, float **depth_sigma, float **model_error,
             float **bottom_albedo, float **bottom_sand, float **bottom_seagrass,
             float **bottom_coral, float **K_min, float **bottom_type,
             float **index_optical_depth, float pagesize, int background, int linewidth)
{
    (void)pagesize;
    (void)background;
    (void)linewidth;
    photic_samodel(scene_data, gridded_data, scene_indexes, nscenes, empirical_depth_present,
                   empirical_depths, n_smoothing_radius, n_spatial, n_bottoms, depth, depth_sigma,
                   model_error, bottom_albedo, bottom_sand, bottom_seagrass, bottom_coral, K_min,
                   bottom_type, index_optical_depth);
}
