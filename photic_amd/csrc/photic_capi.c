/* photic_amd -- C host code: the public C ABI of include/photic.h.
 *
 * Turns a photic_cfg into the uniform constants of a launch (band tables interpolated at the
 * band wavelengths as model/samodel.c:505-618 does, Nelder-Mead settings, the interpolation
 * segments of model/samodel.c:1785-1799), checks arguments, launches the gfx950 kernels of
 * ph_kernels.hip on the caller's stream, and wraps the HIP runtime for callers that have no
 * allocator of their own.  No computation happens here on the CPU apart from those
 * per-call constants; without a GPU every computing entry point returns PHOTIC_EDEVICE. */
#include "photic.h"

#include <hip/hip_runtime_api.h>
#include <math.h>
#include <pthread.h>
#include <stdarg.h>
#include <stdlib.h>
#include <stdio.h>
#include <string.h>

#include "ph_launch.h"
#include "ph_tables.h"

/* ------------------------------------------------------------------------------------ */
/* errors                                                                               */
/* ------------------------------------------------------------------------------------ */

static _Thread_local char ph_err[512];

const char *photic_last_error(void) { return ph_err; }

int ph_error(int code, const char *msg)
{
    snprintf(ph_err, sizeof(ph_err), "%s", msg);
    return code;
}

static int ph_fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(ph_err, sizeof(ph_err), fmt, ap);
    va_end(ap);
    return code;
}

static int ph_hip(hipError_t e, const char *what)
{
    if (e == hipSuccess) return PHOTIC_OK;
    if (e == hipErrorOutOfMemory)
        return ph_fail(PHOTIC_ENOMEM, "%s: %s", what, hipGetErrorString(e));
    return ph_fail(PHOTIC_EDEVICE, "%s: %s", what, hipGetErrorString(e));
}

static int ph_launch_status(int rc, const char *what)
{
    if (rc == PH_LAUNCH_UNSUPPORTED)
        return ph_fail(PHOTIC_EUNSUPPORTED, "%s: configuration not compiled", what);
    return ph_hip((hipError_t)rc, what);
}

static int ph_have_device(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1)
        return ph_fail(PHOTIC_EDEVICE, "no GPU device available");
    return PHOTIC_OK;
}

/* ------------------------------------------------------------------------------------ */
/* the reference's helpers, model/common.c:298-335 and :392-406                          */
/* ------------------------------------------------------------------------------------ */

static int ph_feq(float a, float b, float eps)
{
    float d = a - b;
    double fa = fabs((double)a), fb = fabs((double)b);
    return fabs((double)d) <= ((fa < fb ? fb : fa) * (double)eps);
}

/* Segment and weights of the reference's interp_1d for query x: value = Y[i0] w0 + Y[i1] w1
 * (node hits return the node value: i1 = i0, w0 = 1, w1 = 0). */
static void ph_interp_weights(const double *X, int n, double x, int *i0, int *i1, double *w0,
                              double *w1)
{
    int a = 0, b = 1;
    float xf = (float)x;
    if (ph_feq(xf, (float)X[0], 1.0e-5f)) {
        *i0 = *i1 = 0; *w0 = 1.0; *w1 = 0.0;
        return;
    }
    if (ph_feq(xf, (float)X[n - 1], 1.0e-5f)) {
        *i0 = *i1 = n - 1; *w0 = 1.0; *w1 = 0.0;
        return;
    }
    if (X[0] < X[n - 1] && x < X[0]) {
        a = 0; b = 1;
    } else if (X[n - 1] > X[0] && x > X[n - 1]) {
        a = n - 2; b = n - 1;
    } else {
        for (int i = 0; i < n - 1; i++) {
            float xi = (float)X[i], xj = (float)X[i + 1];
            int le_i = xi < xf || ph_feq(xi, xf, 1.0e-5f);
            int ge_j = xj > xf || ph_feq(xj, xf, 1.0e-5f);
            int ge_i = xi > xf || ph_feq(xi, xf, 1.0e-5f);
            int le_j = xj < xf || ph_feq(xj, xf, 1.0e-5f);
            if ((le_i && ge_j) || (ge_i && le_j)) {
                a = i; b = i + 1;
                break;
            }
        }
    }
    double alpha = (x - X[a]) / (X[b] - X[a]);
    *i0 = a; *i1 = b; *w0 = 1.0 - alpha; *w1 = alpha;
}

static double ph_interp(const double *X, const double *Y, int n, double x)
{
    int i0, i1;
    double w0, w1;
    ph_interp_weights(X, n, x, &i0, &i1, &w0, &w1);
    return Y[i0] * w0 + Y[i1] * w1;
}

/* ------------------------------------------------------------------------------------ */
/* configuration                                                                          */
/* ------------------------------------------------------------------------------------ */

void photic_cfg_default(photic_cfg *cfg, int n_bands, const double *wavelengths)
{
    static const double hs[8] = {40.0, 30.0, 20.0, 15.0, 10.0, 7.5, 2.5, 1.0};
    memset(cfg, 0, sizeof(*cfg));
    cfg->n_bands = n_bands;
    for (int b = 0; b < n_bands && b < PHOTIC_MAX_BANDS; b++)
        cfg->wavelengths[b] = wavelengths[b];
    cfg->mode = PHOTIC_MODE_MIX;
    cfg->n_bottoms = 2;
    for (int k = 0; k < PHOTIC_N_BOTTOM_TYPES; k++)
        cfg->bottom_idx[k] = k;
    cfg->reqmin = 1.0e-2;
    cfg->konvge = 100;
    cfg->kcount = 5000;
    cfg->n_hstarts = 8;
    for (int i = 0; i < 8; i++)
        cfg->hstart[i] = hs[i];
    cfg->early_exit = 0.0;
    cfg->nodata = -9999.0f;
    cfg->sec_sun = 1.0;
    cfg->sec_view = 1.0;
    cfg->precision = PHOTIC_FP64;
    /* seeded starts: H +-30 %, B +-0.25, P/G/X a factor 2 either side */
    const double rel[5] = {0.3, 0.0, 1.0, 1.0, 1.0};
    const double mn[5] = {0.5, 0.25, 0.1, 0.1, 0.02};
    for (int k = 0; k < 5; k++) {
        cfg->seed_rel_step[k] = rel[k];
        cfg->seed_min_step[k] = mn[k];
    }
}

int photic_n_params(const photic_cfg *cfg)
{
    if (cfg->mode == PHOTIC_MODE_MIX) return 5;
    return 1 + 2 * cfg->n_bottoms + 3;
}

static int ph_make_const(const photic_cfg *cfg, int n_dates, int has_seed, PhConst *c)
{
    if (!cfg) return ph_fail(PHOTIC_EINVAL, "cfg is NULL");
    if (cfg->n_bands < 2 || cfg->n_bands > PH_MAXB)
        return ph_fail(PHOTIC_EINVAL, "n_bands %d outside 2..%d", cfg->n_bands, PH_MAXB);
    for (int b = 1; b < cfg->n_bands; b++)
        if (!(cfg->wavelengths[b] > cfg->wavelengths[b - 1]))
            return ph_fail(PHOTIC_EINVAL, "wavelengths must ascend");
    if (cfg->mode != PHOTIC_MODE_MIX && cfg->mode != PHOTIC_MODE_NATIVE)
        return ph_fail(PHOTIC_EINVAL, "mode %d unknown", cfg->mode);
    if (cfg->mode == PHOTIC_MODE_NATIVE) {
        if (cfg->n_bottoms < 1 || cfg->n_bottoms > 3)
            return ph_fail(PHOTIC_EUNSUPPORTED, "native mode supports 1..3 bottoms, not %d",
                           cfg->n_bottoms);
        for (int k = 0; k < cfg->n_bottoms; k++)
            if (cfg->bottom_idx[k] < 0 || cfg->bottom_idx[k] >= PH_NBOT)
                return ph_fail(PHOTIC_EINVAL, "bottom_idx[%d] = %d", k, cfg->bottom_idx[k]);
    }
    if (!(cfg->reqmin > 0.0) || cfg->konvge < 1 || cfg->kcount < 1)
        return ph_fail(PHOTIC_EINVAL, "reqmin must be > 0, konvge and kcount >= 1");
    if (cfg->n_hstarts < 0 || cfg->n_hstarts > PH_MAXH || (cfg->n_hstarts == 0 && !has_seed))
        return ph_fail(PHOTIC_EINVAL, "n_hstarts %d outside 1..%d", cfg->n_hstarts, PH_MAXH);
    if (cfg->precision != PHOTIC_FP64 && cfg->precision != PHOTIC_FP32)
        return ph_fail(PHOTIC_EINVAL, "precision %d unknown", cfg->precision);

    memset(c, 0, sizeof(*c));
    const int nb = cfg->n_bands;
    c->n_bands = nb;
    c->n_dates = n_dates;
    c->mode = cfg->mode == PHOTIC_MODE_MIX ? 1 : 0;
    c->n_bottoms = cfg->mode == PHOTIC_MODE_MIX ? 2 : cfg->n_bottoms;
    for (int k = 0; k < PH_NBOT; k++)
        c->bottom_idx[k] = cfg->mode == PHOTIC_MODE_MIX ? k : cfg->bottom_idx[k];
    c->konvge = cfg->konvge;
    c->kcount = cfg->kcount;
    c->n_hstarts = cfg->n_hstarts;
    c->has_seed = has_seed;
    c->reqmin = cfg->reqmin;
    int n = cfg->mode == PHOTIC_MODE_MIX ? 2 + 3 * n_dates : 1 + 2 * cfg->n_bottoms + 3 * n_dates;
    c->rq = cfg->reqmin * (double)n; /* asa047.c:170 */
    /* the early exit of a problem over n_dates dates or scenes: 2.5 (or the configured level)
     * per date, model/samodel.c:2404 */
    c->early_exit = (cfg->early_exit > 0.0 ? cfg->early_exit : 2.5) * (double)(float)n_dates;
    c->nodata = (double)cfg->nodata;
    c->prescreened = cfg->prescreened ? 1 : 0;
    c->seed_konvge = cfg->seed_konvge > 0 ? cfg->seed_konvge : cfg->konvge;
    c->sec_sun0 = cfg->sec_sun;
    c->sec_view0 = cfg->sec_view;
    for (int i = 0; i < PH_MAXH; i++)
        c->hstart[i] = i < cfg->n_hstarts ? cfg->hstart[i] : 0.0;
    for (int b = 0; b < nb; b++) {
        double wl = cfg->wavelengths[b];
        c->wl[b] = wl;
        c->a0[b] = ph_interp(ph_ref_wl, ph_tab_a0, PH_NREF, wl);
        c->a1[b] = ph_interp(ph_ref_wl, ph_tab_a1, PH_NREF, wl);
        c->bbw[b] = ph_interp(ph_ref_wl, ph_tab_bbw, PH_NREF, wl);
        c->aw[b] = ph_interp(ph_ref_wl, ph_tab_aw, PH_NREF, wl);
        c->eg[b] = exp(-0.015 * (wl - 440.0));
        for (int k = 0; k < PH_NBOT; k++)
            c->bot[k][b] = ph_interp(ph_ref_wl, ph_tab_bottom[k], PH_NREF, wl);
    }
    c->aw640 = ph_interp(ph_ref_wl, ph_tab_aw, PH_NREF, 640.0);
    const double q[5] = {440.0, 490.0, 550.0, 640.0, 750.0};
    for (int k = 0; k < 5; k++)
        ph_interp_weights(cfg->wavelengths, nb, q[k], &c->ip_i0[k], &c->ip_i1[k], &c->ip_w0[k],
                          &c->ip_w1[k]);
    for (int k = 0; k < 5; k++) {
        c->seed_rel_step[k] = cfg->seed_rel_step[k];
        c->seed_min_step[k] = cfg->seed_min_step[k];
    }
    /* never below the early exit: a seed start given up must not end the pixel as accepted */
    c->seed_abort = cfg->seed_abort_misfit > 0.0 ? cfg->seed_abort_misfit : 0.0;
    if (c->seed_abort > 0.0 && c->seed_abort < c->early_exit) c->seed_abort = c->early_exit;
    return PHOTIC_OK;
}

int ph_build_const(const struct photic_cfg *cfg, int n_dates, int has_seed, PhConst *out)
{
    return ph_make_const(cfg, n_dates, has_seed, out);
}

int ph_const_size(void) { return (int)sizeof(PhConst); }

int photic_band_coefficients(const photic_cfg *cfg, double *out, double *aw640)
{
    PhConst c;
    int rc = ph_make_const(cfg, 1, 0, &c);
    if (rc) return rc;
    if (!out) return ph_fail(PHOTIC_EINVAL, "out is NULL");
    const int nb = c.n_bands;
    for (int b = 0; b < nb; b++) {
        out[0 * nb + b] = c.a0[b];
        out[1 * nb + b] = c.a1[b];
        out[2 * nb + b] = c.aw[b];
        out[3 * nb + b] = c.bbw[b];
        for (int k = 0; k < PH_NBOT; k++) out[(4 + k) * nb + b] = c.bot[k][b];
    }
    if (aw640) *aw640 = c.aw640;
    return PHOTIC_OK;
}

/* ------------------------------------------------------------------------------------ */
/* inversion                                                                              */
/* ------------------------------------------------------------------------------------ */

static int ph_n_cu(int *n_cu)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(n_cu, hipDeviceAttributeMultiprocessorCount, dev);
    return ph_hip(e, "hipDeviceGetAttribute");
}

int photic_kernel_info(const photic_cfg *cfg, int *vgprs, int *scratch_bytes,
                       int *blocks_per_cu, int *n_cu)
{
    PhConst c;
    int rc = ph_have_device();
    if (rc) return rc;
    rc = ph_make_const(cfg, 1, 0, &c);
    if (rc) return rc;
    int fp32 = cfg->precision == PHOTIC_FP32, nbot = c.mode ? 0 : c.n_bottoms;
    int sg = 0, mt = 0, block = cfg->block > 0 ? cfg->block : 256;
    rc = ph_kernel_attrs(fp32, c.n_bands, nbot, vgprs, &sg, scratch_bytes, &mt);
    if (rc) return ph_launch_status(rc, "kernel attributes");
    rc = ph_kernel_occupancy(fp32, c.n_bands, nbot, 0, block, blocks_per_cu);
    if (rc) return ph_launch_status(rc, "occupancy");
    return ph_n_cu(n_cu);
}

int photic_invert(const photic_cfg *cfg, long n_pix, const float *rrs, const float *sec_sun,
                  const float *sec_view, const float *seed, const photic_outputs *out,
                  void *stream)
{
    return photic_invert_h(cfg, n_pix, rrs, sec_sun, sec_view, seed, NULL, out, stream);
}

/* Launch shapes chosen by photic_autotune / photic_autotune_stack, per kernel instance
 * (zero: not tuned, the defaults apply).  Results never depend on the shape. */
typedef struct {
    int block, bpc, chunk;
} ph_shape;
static ph_shape ph_tuned_single[2][4][2]; /* [fp32][n_bottoms (0 = mix)][seeded] */
static ph_shape ph_tuned_stack[2];        /* [seeded] */

/* The single-date inversion.  scratch: the launch's work counter and watchdog flag (two
 * words of device memory) from the caller, or NULL to allocate them on the stream. */
static int ph_invert_single(const photic_cfg *cfg, long n_pix, const float *rrs,
                            const float *sec_sun, const float *sec_view, const float *seed,
                            const float *hstart, const photic_outputs *out, void *stream,
                            unsigned long long *scratch)
{
    PhInvertArgs a;
    int rc = ph_have_device();
    if (rc) return rc;
    if (n_pix < 0) return ph_fail(PHOTIC_EINVAL, "n_pix < 0");
    if (!out) return ph_fail(PHOTIC_EINVAL, "out is NULL");
    if (n_pix > 0 && !rrs) return ph_fail(PHOTIC_EINVAL, "rrs is NULL");
    memset(&a, 0, sizeof(a));
    if (hstart) {
        /* one depth start per pixel: the reference's empirical-depth mode (samodel.c:2223) */
        photic_cfg one = *cfg;
        one.n_hstarts = 1;
        rc = ph_make_const(&one, 1, seed != NULL, &a.c);
    } else {
        rc = ph_make_const(cfg, 1, seed != NULL, &a.c);
    }
    if (rc) return rc;
    if (n_pix == 0) return PHOTIC_OK;

    const int fp32 = cfg->precision == PHOTIC_FP32;
    const int nbot = a.c.mode ? 0 : a.c.n_bottoms;
    if (fp32 && nbot != 0)
        return ph_fail(PHOTIC_EUNSUPPORTED, "FP32 forward model is compiled for MIX mode only");
    /* block size: the lanes' simplexes live in LDS, so the largest multiple of 64 lanes
     * (up to 256) whose LDS fits one workgroup */
    int dev = 0, max_lds = 0;
    rc = ph_hip(hipGetDevice(&dev), "hipGetDevice");
    if (rc) return rc;
    rc = ph_hip(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev),
                "hipDeviceGetAttribute");
    if (rc) return rc;
    const int lds_lane = ph_lds_bytes(nbot, 1);
    const ph_shape *tu = &ph_tuned_single[fp32][nbot][seed != NULL];
    int block = cfg->block > 0 ? cfg->block : (max_lds / lds_lane) / 64 * 64;
    if (cfg->block <= 0 && block > 256) block = 256;
    if (cfg->block <= 0 && tu->block > 0) block = tu->block;
    if (block < 64 || block % 64 != 0 || block > 256)
        return ph_fail(PHOTIC_EINVAL, "block %d must be a multiple of 64 in 64..256", block);
    if ((long)lds_lane * block > max_lds)
        return ph_fail(PHOTIC_EUNSUPPORTED, "block %d needs %d bytes of LDS, the device has %d",
                       block, lds_lane * block, max_lds);
    int n_cu = 0, bpc = 0;
    rc = ph_n_cu(&n_cu);
    if (rc) return rc;
    rc = ph_kernel_occupancy(fp32, a.c.n_bands, nbot, seed != NULL, block, &bpc);
    if (rc) return ph_launch_status(rc, "occupancy");
    if (bpc < 1) return ph_fail(PHOTIC_EUNSUPPORTED, "kernel does not fit a CU at block %d", block);
    if (cfg->grid <= 0 && cfg->block <= 0 && tu->bpc > 0 && tu->bpc < bpc) bpc = tu->bpc;
    long grid = cfg->grid > 0 ? cfg->grid : (long)n_cu * bpc;
    long waves_needed = (n_pix + 63) / 64;
    long blocks_needed = (waves_needed * 64 + block - 1) / block;
    if (grid > blocks_needed) grid = blocks_needed;
    if (grid < 1) grid = 1;

    a.n_pix = n_pix;
    a.rrs = rrs;
    a.sec_sun = sec_sun;
    a.sec_view = sec_view;
    a.seed = seed;
    a.hstart = hstart;
    a.chunk = cfg->chunk > 0 ? cfg->chunk : (tu->chunk > 0 ? tu->chunk : 64);
    /* a wave sets up new pixels once two lanes are idle: the set-up (five correctly rounded
     * pows) then runs for two pixels at a time (+1% at 4096^2, seeded and plain alike) */
    a.refill_min = 2;
    a.depth = out->depth;
    a.rrs_error = out->rrs_error;
    a.model_error = out->model_error;
    a.albedo = out->albedo;
    a.k_min = out->k_min;
    a.iod = out->iod;
    a.b_pct = out->b_pct;
    a.bottom_type = out->bottom_type;
    a.x_out = out->x;
    a.status = out->status;
    a.n_evals = out->n_evals;
    a.n_iter = out->n_iter;
    a.n_starts = out->n_starts;
    a.rrs_model = out->rrs_model;

    /* watchdog: a wavefront that finishes no pixel for this long gives up (it exits, leaving
     * its pixels unwritten); with cfg->check_watchdog the call waits for the kernel and
     * reports PHOTIC_EDEVICE when that happened */
    int wrate = 0;
    if (hipDeviceGetAttribute(&wrate, hipDeviceAttributeWallClockRate, dev) != hipSuccess ||
        wrate <= 0)
        wrate = 100000; /* kHz */
    double limit_s = cfg->watchdog_s > 0.0 ? cfg->watchdog_s : 600.0;
    a.max_ticks = (unsigned long long)(limit_s * 1000.0 * (double)wrate);

    hipStream_t st = (hipStream_t)stream;
    unsigned long long *queue = scratch;
    if (!queue) {
        rc = ph_hip(hipMallocAsync((void **)&queue, 2 * sizeof(unsigned long long), st),
                    "hipMallocAsync");
        if (rc) return rc;
    }
    /* cleared by a fill kernel, which a graph capture records as a kernel node */
    rc = ph_launch_status(ph_launch_zero_u64(queue, 2, st), "zero work counter");
    if (rc == PHOTIC_OK) {
        a.queue = queue;
        rc = ph_launch_status(ph_launch_invert(&a, fp32, a.c.n_bands, nbot, (int)grid, block, st),
                              "inversion kernel");
    }
    if (rc == PHOTIC_OK && cfg->check_watchdog) {
        unsigned long long flag = 0;
        rc = ph_hip(hipMemcpyAsync(&flag, queue + 1, sizeof(flag), hipMemcpyDeviceToHost, st),
                    "watchdog read");
        if (rc == PHOTIC_OK) rc = ph_hip(hipStreamSynchronize(st), "hipStreamSynchronize");
        if (rc == PHOTIC_OK && flag)
            rc = ph_fail(PHOTIC_EDEVICE, "inversion kernel stopped by its %.0f s watchdog",
                         limit_s);
    }
    if (scratch) return rc;
    int rc2 = ph_hip(hipFreeAsync(queue, st), "hipFreeAsync");
    return rc ? rc : rc2;
}

int photic_invert_h(const photic_cfg *cfg, long n_pix, const float *rrs, const float *sec_sun,
                    const float *sec_view, const float *seed, const float *hstart,
                    const photic_outputs *out, void *stream)
{
    return ph_invert_single(cfg, n_pix, rrs, sec_sun, sec_view, seed, hstart, out, stream, NULL);
}

int photic_invert_noalloc(const photic_cfg *cfg, long n_pix, const float *rrs,
                          const float *sec_sun, const float *sec_view, const float *seed,
                          const float *hstart, const photic_outputs *out, void *stream)
{
    int rc = ph_have_device();
    if (rc) return rc;
    /* the launch needs two words of the device's scratch array: its work counter and its
     * watchdog flag, whatever n_pix is */
    const int need = 2;
    if (need > PH_SCRATCH_WORDS)
        return ph_fail(PHOTIC_EUNSUPPORTED, "the inversion needs %d scratch words, the device "
                                            "array holds %d", need, PH_SCRATCH_WORDS);
    unsigned long long *scratch = NULL;
    rc = ph_launch_status(ph_scratch_words(&scratch), "scratch symbol address");
    if (rc) return rc;
    return ph_invert_single(cfg, n_pix, rrs, sec_sun, sec_view, seed, hstart, out, stream,
                            scratch);
}

/* ------------------------------------------------------------------------------------ */
/* multi-date stacks                                                                      */
/* ------------------------------------------------------------------------------------ */

/* Two streams per device on which photic_invert_stack spreads its per-date-count launches,
 * created on first use and kept (destroying a stream may wait for its work, which would make
 * the call synchronous).  Calls on one device share them, so their launches queue there in
 * call order; each call still waits only on events recorded after its own launches. */
#define PH_SIDE 2
#define PH_MAX_DEVICES 64
static hipStream_t ph_side_tab[PH_MAX_DEVICES][PH_SIDE];
static pthread_mutex_t ph_side_lock = PTHREAD_MUTEX_INITIALIZER;

static int ph_side_streams(int dev, hipStream_t *out)
{
    if (dev < 0 || dev >= PH_MAX_DEVICES) return ph_fail(PHOTIC_EDEVICE, "device %d", dev);
    int rc = PHOTIC_OK;
    pthread_mutex_lock(&ph_side_lock);
    for (int s = 0; s < PH_SIDE && !rc; s++) {
        if (!ph_side_tab[dev][s])
            rc = ph_hip(hipStreamCreateWithFlags(&ph_side_tab[dev][s], hipStreamNonBlocking),
                        "hipStreamCreate");
        out[s] = ph_side_tab[dev][s];
    }
    pthread_mutex_unlock(&ph_side_lock);
    return rc;
}

static int ph_invert_stack(const photic_cfg *cfg, long n_pix, int n_dates, const float *rrs,
                           const float *sec_sun, const float *sec_view,
                           const unsigned char *valid, const float *seed, const float *hstart,
                           const photic_outputs *out, void *stream)
{
    int rc = ph_have_device();
    if (rc) return rc;
    if (!cfg || !out) return ph_fail(PHOTIC_EINVAL, "cfg and out are required");
    if (n_pix < 0) return ph_fail(PHOTIC_EINVAL, "n_pix < 0");
    if (n_dates < 1 || n_dates > PH_MAXD)
        return ph_fail(PHOTIC_EINVAL, "n_dates %d outside 1..%d", n_dates, PH_MAXD);
    const int native = cfg->mode == PHOTIC_MODE_NATIVE;
    if (cfg->n_bands != 4 || cfg->precision != PHOTIC_FP64 ||
        (native && (cfg->n_bottoms < 1 || cfg->n_bottoms > 3)))
        return ph_fail(PHOTIC_EUNSUPPORTED,
                       "stacks are compiled for 4 bands, FP64, the mix layout or 1..3 native "
                       "bottoms");
    if (native && seed)
        return ph_fail(PHOTIC_EUNSUPPORTED, "seeded stacks are compiled for the mix layout");
    if (n_pix == 0) return PHOTIC_OK;
    if (!rrs || !sec_sun || !valid)
        return ph_fail(PHOTIC_EINVAL, "rrs, sec_sun and valid are required");
    if (n_pix > 2147483647L) return ph_fail(PHOTIC_EINVAL, "n_pix above 2^31 - 1");
    const int nbot = native ? cfg->n_bottoms : 0;

    PhStackArgs a;
    memset(&a, 0, sizeof(a));
    a.n_pix = n_pix;
    a.n_dates = n_dates;
    a.rrs = rrs;
    a.sec_sun = sec_sun;
    a.sec_view = sec_view;
    a.hstart = hstart;
    a.x_off = native ? 1 + 2 * nbot : 2;
    a.valid = valid;
    a.seed = seed;
    a.chunk = cfg->chunk > 0 ? cfg->chunk : (ph_tuned_stack[seed != NULL].chunk > 0
                                                 ? ph_tuned_stack[seed != NULL].chunk : 64);
    a.depth = out->depth;
    a.rrs_error = out->rrs_error;
    a.model_error = out->model_error;
    a.albedo = out->albedo;
    a.k_min = out->k_min;
    a.iod = out->iod;
    a.b_pct = out->b_pct;
    a.bottom_type = out->bottom_type;
    a.x_out = out->x;
    a.status = out->status;
    a.n_evals = out->n_evals;
    a.n_iter = out->n_iter;
    a.n_starts = out->n_starts;

    int dev = 0, wrate = 0, n_cu = 0, bpc = 0;
    rc = ph_hip(hipGetDevice(&dev), "hipGetDevice");
    if (rc) return rc;
    if (hipDeviceGetAttribute(&wrate, hipDeviceAttributeWallClockRate, dev) != hipSuccess ||
        wrate <= 0)
        wrate = 100000; /* kHz */
    const double limit_s = cfg->watchdog_s > 0.0 ? cfg->watchdog_s : 600.0;
    a.max_ticks = (unsigned long long)(limit_s * 1000.0 * (double)wrate);
    rc = ph_n_cu(&n_cu);
    if (rc) return rc;
    const ph_shape *tu = &ph_tuned_stack[seed != NULL];
    const int block = cfg->block > 0 ? cfg->block : (tu->block > 0 ? tu->block : 256);
    if (block < 64 || block > 256 || block % 64)
        return ph_fail(PHOTIC_EINVAL, "block %d must be a multiple of 64 in 64..256", block);

    hipStream_t st = (hipStream_t)stream;
    unsigned int *counts = NULL;
    int *lists = NULL;
    unsigned long long *queues = NULL;
    rc = ph_hip(hipMallocAsync((void **)&counts, sizeof(unsigned int) * 16, st), "hipMallocAsync");
    if (!rc)
        rc = ph_hip(hipMallocAsync((void **)&lists, sizeof(int) * (size_t)(n_dates + 1) *
                                                        (size_t)n_pix, st),
                    "hipMallocAsync");
    if (!rc)
        rc = ph_hip(hipMallocAsync((void **)&queues, sizeof(unsigned long long) * 2 *
                                                         (size_t)(n_dates + 1), st),
                    "hipMallocAsync");
    if (!rc) rc = ph_launch_status(ph_launch_zero_u64((unsigned long long *)counts, 8, st), "zero");
    if (!rc)
        rc = ph_launch_status(ph_launch_zero_u64(queues, 2 * (n_dates + 1), st), "zero");
    if (!rc) rc = ph_launch_status(ph_launch_stack_bucket(&a, counts, lists, st), "bucketing");
    /* One persistent launch per date count, each pulling its bucket's stacks; the group size
     * and so the LDS per block, hence the resident blocks, depend on the date count.  The
     * launches alternate between two side streams forked from the caller's, so one launch's
     * tail (its last stacks finishing while most lanes idle) overlaps the next launch. */
    hipStream_t side[PH_SIDE] = {NULL, NULL};
    hipEvent_t ev_fork = NULL, ev_join[PH_SIDE] = {NULL, NULL};
    if (!rc) rc = ph_side_streams(dev, side);
    if (!rc) rc = ph_hip(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming), "event");
    if (!rc) rc = ph_hip(hipEventRecord(ev_fork, st), "hipEventRecord");
    for (int s = 0; s < PH_SIDE && !rc; s++) {
        rc = ph_hip(hipEventCreateWithFlags(&ev_join[s], hipEventDisableTiming), "event");
        if (!rc) rc = ph_hip(hipStreamWaitEvent(side[s], ev_fork, 0), "hipStreamWaitEvent");
    }
    for (int k = n_dates; k >= 1 && !rc; k--) {
        photic_cfg one = *cfg;
        rc = ph_make_const(&one, k, seed != NULL, &a.c);
        if (rc) break;
        a.idx = lists + (long)k * n_pix;
        a.count = counts + k;
        a.queue = queues + 2 * k;
        if (hstart) a.c.n_hstarts = 1; /* the stack's own depth start replaces the list */
        rc = ph_launch_status(ph_stack_occupancy(k, nbot, seed != NULL, block, &bpc),
                              "occupancy");
        if (rc) break;
        if (bpc < 1) {
            rc = ph_fail(PHOTIC_EDEVICE, "stack kernel (%d dates) cannot be resident", k);
            break;
        }
        const long grid = cfg->grid > 0 ? cfg->grid : (long)n_cu * bpc;
        rc = ph_launch_status(ph_launch_stack(&a, k, nbot, seed != NULL, (int)grid, block,
                                              side[(n_dates - k) % PH_SIDE]),
                              "stack kernel");
    }
    /* join: the caller's stream waits for both side streams (also on error, so the frees
     * below stay ordered after every launch that was made) */
    for (int s = 0; s < PH_SIDE; s++) {
        if (!side[s]) continue;
        if (ev_join[s] && hipEventRecord(ev_join[s], side[s]) == hipSuccess)
            hipStreamWaitEvent(st, ev_join[s], 0);
        else
            hipStreamSynchronize(side[s]);
    }
    if (rc == PHOTIC_OK && cfg->check_watchdog) {
        unsigned long long flags[2 * (PH_MAXD + 1)];
        rc = ph_hip(hipMemcpyAsync(flags, queues, sizeof(unsigned long long) * 2 *
                                                      (size_t)(n_dates + 1),
                                   hipMemcpyDeviceToHost, st),
                    "watchdog read");
        if (rc == PHOTIC_OK) rc = ph_hip(hipStreamSynchronize(st), "hipStreamSynchronize");
        for (int k = 1; rc == PHOTIC_OK && k <= n_dates; k++)
            if (flags[2 * k + 1])
                rc = ph_fail(PHOTIC_EDEVICE, "stack kernel (%d dates) stopped by its %.0f s "
                                             "watchdog", k, limit_s);
    }
    if (queues) hipFreeAsync(queues, st);
    if (lists) hipFreeAsync(lists, st);
    if (counts) hipFreeAsync(counts, st);
    for (int s = 0; s < PH_SIDE; s++)
        if (ev_join[s]) hipEventDestroy(ev_join[s]); /* returns at once, released later */
    if (ev_fork) hipEventDestroy(ev_fork);
    return rc;
}

int photic_invert_stack(const photic_cfg *cfg, long n_pix, int n_dates, const float *rrs,
                        const float *sec_sun, const unsigned char *valid, const float *seed,
                        const photic_outputs *out, void *stream)
{
    return ph_invert_stack(cfg, n_pix, n_dates, rrs, sec_sun, NULL, valid, seed, NULL, out,
                           stream);
}

int photic_invert_stack_h(const photic_cfg *cfg, long n_pix, int n_dates, const float *rrs,
                          const float *sec_sun, const float *sec_view,
                          const unsigned char *valid, const float *hstart,
                          const photic_outputs *out, void *stream)
{
    return ph_invert_stack(cfg, n_pix, n_dates, rrs, sec_sun, sec_view, valid, NULL, hstart,
                           out, stream);
}

/* ------------------------------------------------------------------------------------ */
/* spatial regions                                                                        */
/* ------------------------------------------------------------------------------------ */

int photic_invert_regions(const photic_cfg *cfg, int n_rows, int n_cols, int n_scenes,
                          const float *rrs, const float *nodata, const double *sec_sun,
                          const double *sec_view, int n_spatial, long n, const int64_t *idx,
                          const float *hstart, const photic_outputs *out, int *n_regions,
                          void *stream)
{
    return ph_invert_regions_pert(cfg, n_rows, n_cols, n_scenes, rrs, nodata, sec_sun, sec_view,
                                  n_spatial, n, idx, hstart, NULL, NULL, out, n_regions, NULL,
                                  stream);
}

int photic_invert_regions_bands(const photic_cfg *cfg, int n_rows, int n_cols, int n_scenes,
                                const int *n_bands, const double *wavelengths, const float *rrs,
                                const float *nodata, const double *sec_sun,
                                const double *sec_view, int n_spatial, long n,
                                const int64_t *idx, const float *hstart,
                                const photic_outputs *out, int *n_regions, void *stream)
{
    return ph_invert_regions_bands(cfg, n_rows, n_cols, n_scenes, n_bands, wavelengths, rrs,
                                   nodata, sec_sun, sec_view, n_spatial, n, idx, hstart, NULL,
                                   NULL, out, n_regions, NULL, stream);
}

int ph_invert_regions_pert(const photic_cfg *cfg, int n_rows, int n_cols, int n_scenes,
                           const float *rrs, const float *nodata, const double *sec_sun,
                           const double *sec_view, int n_spatial, long n, const int64_t *idx,
                           const float *hstart, const float *pert, const double *rsig,
                           const photic_outputs *out, int *n_regions, double *depth64,
                           void *stream)
{
    if (!cfg) return ph_fail(PHOTIC_EINVAL, "cfg is NULL");
    if (n_scenes < 1 || n_scenes > PH_MAXS)
        return ph_fail(PHOTIC_EINVAL, "n_scenes %d outside 1..%d", n_scenes, PH_MAXS);
    if (cfg->n_bands < 2 || cfg->n_bands > PH_MAXB)
        return ph_fail(PHOTIC_EINVAL, "n_bands %d outside 2..%d", cfg->n_bands, PH_MAXB);
    /* every scene with cfg's bands */
    int nbs[PH_MAXS];
    double wl[PH_MAXS * PH_MAXB];
    for (int s = 0; s < n_scenes; s++) {
        nbs[s] = cfg->n_bands;
        for (int b = 0; b < cfg->n_bands; b++) wl[s * cfg->n_bands + b] = cfg->wavelengths[b];
    }
    return ph_invert_regions_bands(cfg, n_rows, n_cols, n_scenes, nbs, wl, rrs, nodata, sec_sun,
                                   sec_view, n_spatial, n, idx, hstart, pert, rsig, out,
                                   n_regions, depth64, stream);
}

int ph_invert_regions_bands(const photic_cfg *cfg, int n_rows, int n_cols, int n_scenes,
                            const int *n_bands, const double *wavelengths, const float *rrs,
                            const float *nodata, const double *sec_sun, const double *sec_view,
                            int n_spatial, long n, const int64_t *idx, const float *hstart,
                            const float *pert, const double *rsig, const photic_outputs *out,
                            int *n_regions, double *depth64, void *stream)
{
    int rc = ph_have_device();
    if (rc) return rc;
    if (!cfg || !out || !nodata || !sec_sun || !sec_view || !n_bands || !wavelengths)
        return ph_fail(PHOTIC_EINVAL, "cfg, out, nodata, sec_sun, sec_view, n_bands and "
                                      "wavelengths are required");
    if (n_scenes < 1 || n_scenes > PH_MAXS)
        return ph_fail(PHOTIC_EINVAL, "n_scenes %d outside 1..%d", n_scenes, PH_MAXS);
    int n_slots = 0;
    for (int s = 0; s < n_scenes; s++) {
        if (n_bands[s] < 2 || n_bands[s] > PH_MAXB)
            return ph_fail(PHOTIC_EINVAL, "scene %d: %d bands, outside 2..%d", s, n_bands[s],
                           PH_MAXB);
        n_slots += n_bands[s];
    }
    if (n_rows < 1 || n_cols < 1 || n < 0 || n_spatial < 0)
        return ph_fail(PHOTIC_EINVAL, "n_rows, n_cols >= 1, n >= 0 and n_spatial >= 0 required");
    if ((long)n_rows * n_cols > 2147483647L)
        return ph_fail(PHOTIC_EINVAL, "scene above 2^31 - 1 pixels");
    if (!idx && n != (long)n_rows * n_cols)
        return ph_fail(PHOTIC_EINVAL, "without idx, n must be n_rows * n_cols");
    if (cfg->mode != PHOTIC_MODE_NATIVE || cfg->precision != PHOTIC_FP64 ||
        cfg->n_bottoms < 1 || cfg->n_bottoms > 3)
        return ph_fail(PHOTIC_EUNSUPPORTED, "regions are compiled for the native layout with "
                                            "1..3 bottoms, FP64");
    const int ns = n_spatial == 0 ? 1 : n_spatial;
    const int nbot = cfg->n_bottoms;
    const int r_max = (2 * ns - 1) * (2 * ns - 1);
    const int n_max = r_max * (1 + 2 * nbot) + 3 * n_scenes;
    if (r_max > PH_REG_MAXR || n_max > PH_REG_MAXN)
        return ph_fail(PHOTIC_EUNSUPPORTED, "%d regions of %d bottoms and %d scenes need %d "
                                            "parameters; the kernels hold at most %d regions "
                                            "and %d parameters",
                       r_max, nbot, n_scenes, n_max, PH_REG_MAXR, PH_REG_MAXN);
    if (n == 0) return PHOTIC_OK;
    if (!rrs) return ph_fail(PHOTIC_EINVAL, "rrs is required");

    PhRegionArgs a;
    memset(&a, 0, sizeof(a));
    rc = ph_make_const(cfg, n_scenes, 0, &a.c);
    if (rc) return rc;
    if (hstart) a.c.n_hstarts = 1; /* the pixel's own depth start replaces the list */
    a.n_rows = n_rows;
    a.n_cols = n_cols;
    a.n_spatial = n_spatial;
    a.n_scenes = n_scenes;
    a.n_slots = n_slots;
    a.r_max = r_max;
    a.n_max = n_max;
    a.np_max = r_max * n_slots;
    /* the wave's LDS block: the simplex, then each array at an even offset */
    int off = (n_max + 1) * n_max;
#define PH_TAKE(field, count)          \
    do {                               \
        a.field = off;                 \
        off += ((count) + 1) & ~1;     \
    } while (0)
    PH_TAKE(o_x, n_max);
    PH_TAKE(o_d, a.np_max);
    PH_TAKE(o_q, a.np_max);
    PH_TAKE(o_rm, a.np_max);
    PH_TAKE(o_pw, a.np_max);
    PH_TAKE(o_k, n_slots);
    PH_TAKE(o_m, 3 * r_max * n_scenes);
    PH_TAKE(o_pix, r_max);
    PH_TAKE(o_bc, PH_REG_SLOT * n_slots);
    PH_TAKE(o_geo, 2 * n_scenes);
    PH_TAKE(o_sc, 2 * n_scenes);
    PH_TAKE(o_hd, r_max);
    PH_TAKE(o_bqv, r_max * nbot);
    PH_TAKE(o_bd, r_max * nbot);
#undef PH_TAKE
    a.lds_doubles = off;
    /* the tables the kernel reads from device memory (PhRegionArgs::scn): geometry, then per
     * band slot R_sigma, the no-data value and the band's coefficients (model/samodel.c:
     * 518-603: each scene's own wavelengths), then per scene its slots and interpolation
     * weights (model/common.c:298-335) */
    if (pert && !rsig) return ph_fail(PHOTIC_EINVAL, "a perturbation needs R_sigma");
    const size_t scn_n = (size_t)(2 * n_scenes + (2 + PH_REG_SLOT) * n_slots +
                                  PH_REG_SCENE * n_scenes);
    double scn[2 * PH_MAXS + (2 + PH_REG_SLOT) * PH_MAXS * PH_MAXB + PH_REG_SCENE * PH_MAXS];
    memset(scn, 0, sizeof(double) * scn_n);
    {
        double *ss = scn, *sr = scn + 2 * n_scenes, *sn = sr + n_slots, *sl = sn + n_slots;
        double *sc = sl + PH_REG_SLOT * n_slots;
        int slot = 0;
        for (int s = 0; s < n_scenes && !rc; s++) {
            ss[s] = sec_sun[s];
            ss[n_scenes + s] = sec_view[s];
            photic_cfg one = *cfg;
            one.n_bands = n_bands[s];
            for (int b = 0; b < n_bands[s]; b++) one.wavelengths[b] = wavelengths[slot + b];
            PhConst cs;
            rc = ph_make_const(&one, 1, 0, &cs);
            if (rc) break;
            double *st = sc + PH_REG_SCENE * s;
            st[0] = slot;
            st[1] = n_bands[s];
            for (int k = 0; k < 5; k++) {
                st[2 + 4 * k] = cs.ip_i0[k];
                st[3 + 4 * k] = cs.ip_i1[k];
                st[4 + 4 * k] = cs.ip_w0[k];
                st[5 + 4 * k] = cs.ip_w1[k];
            }
            for (int b = 0; b < n_bands[s]; b++, slot++) {
                sr[slot] = rsig ? rsig[slot] : 0.0;
                sn[slot] = (double)nodata[slot];
                double *t = sl + PH_REG_SLOT * slot;
                t[0] = cs.a0[b];
                t[1] = cs.a1[b];
                t[2] = cs.aw[b];
                t[3] = cs.bbw[b];
                t[4] = cs.eg[b];
                for (int k = 0; k < nbot; k++) t[5 + k] = cs.bot[cfg->bottom_idx[k]][b];
                t[8] = 440.0 / cs.wl[b];
                t[9] = 16 * s + b;
            }
        }
    }
    if (rc) return rc;
    a.n = n;
    a.rrs = rrs;
    a.idx = idx;
    a.hstart = hstart;
    a.pert = pert;
    a.depth = out->depth;
    a.rrs_error = out->rrs_error;
    a.model_error = out->model_error;
    a.albedo = out->albedo;
    a.k_min = out->k_min;
    a.iod = out->iod;
    a.b_pct = out->b_pct;
    a.bottom_type = out->bottom_type;
    a.x_out = out->x;
    a.status = out->status;
    a.n_evals = out->n_evals;
    a.n_iter = out->n_iter;
    a.n_starts = out->n_starts;
    a.n_regions = n_regions;
    a.depth64 = depth64;

    int dev = 0, wrate = 0, n_cu = 0, bpc = 0, max_lds = 0;
    rc = ph_hip(hipGetDevice(&dev), "hipGetDevice");
    if (rc) return rc;
    if (hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) !=
            hipSuccess || max_lds <= 0)
        max_lds = 65536;
    /* a wave's block lives in LDS when it fits there and the problem fits one coordinate slot
     * per lane; else its simplex is in device memory, one per wave of the grid, and the other
     * arrays stay in LDS when they fit there (their offsets then counted from the LDS base) or
     * follow the simplex.  Problems of two slots (65-128 parameters) fill the LDS with a
     * single wave per SIMD in the first form and run faster in the second, at three waves
     * per SIMD.  One-slot problems stay in LDS when three blocks per SIMD fit there (their
     * kernel is compiled for three waves); larger one-slot blocks, which the LDS would hold
     * to fewer than two waves per SIMD, run faster in the device-memory form.  The results
     * are the same in every form.  PHOTIC_REGION_DEVICE_BLOCKS (test hooks): 0 keeps blocks
     * of up to two slots in LDS where they fit, 1 forces the device-memory simplex, 2 the
     * whole block into device memory. */
    const char *force_dm = getenv("PHOTIC_REGION_DEVICE_BLOCKS");
    const int force = force_dm && *force_dm ? atoi(force_dm) : -1;
    const int fits_lds = (long)a.lds_doubles * 8 <= max_lds && n_max <= 128;
    const int fits_lds3 = (long)a.lds_doubles * 8 * 12 <= max_lds; /* 3 waves x 4 SIMDs */
    const int in_device = force == 1 || force == 2 || !fits_lds ||
                          (force != 0 && (n_max > 64 || !fits_lds3));
    a.work_doubles = 0;
    a.tables_in_lds = 0;
    if (in_device) {
        const int simplex = (n_max + 1) * n_max; /* even: the arrays stay 16-byte aligned */
        const long tables = (long)a.lds_doubles - simplex;
        if (tables * 8 <= max_lds && force != 2) {
            a.tables_in_lds = 1;
            a.work_doubles = simplex;
            a.lds_doubles = (int)tables;
            int *o[] = {&a.o_x, &a.o_d, &a.o_q, &a.o_rm, &a.o_pw, &a.o_k, &a.o_m,
                        &a.o_pix, &a.o_bc, &a.o_geo, &a.o_sc, &a.o_hd, &a.o_bqv, &a.o_bd};
            for (size_t i = 0; i < sizeof(o) / sizeof(o[0]); i++) *o[i] -= simplex;
        } else {
            a.work_doubles = a.lds_doubles;
        }
    }
    if (hipDeviceGetAttribute(&wrate, hipDeviceAttributeWallClockRate, dev) != hipSuccess ||
        wrate <= 0)
        wrate = 100000; /* kHz */
    const double limit_s = cfg->watchdog_s > 0.0 ? cfg->watchdog_s : 600.0;
    a.max_ticks = (unsigned long long)(limit_s * 1000.0 * (double)wrate);
    rc = ph_n_cu(&n_cu);
    if (rc) return rc;
    rc = ph_launch_status(ph_region_occupancy(&a, in_device, &bpc), "occupancy");
    if (rc) return rc;
    if (bpc < 1) return ph_fail(PHOTIC_EUNSUPPORTED, "region kernel does not fit a CU");
    long grid = cfg->grid > 0 ? cfg->grid : (long)n_cu * bpc;
    if (grid > n) grid = n;

    hipStream_t st = (hipStream_t)stream;
    /* one allocation: the work counter and watchdog flag, the per-scene values, and with
     * device-memory blocks one block per wave of the grid */
    const size_t scn_off = 2 * sizeof(unsigned long long);
    const size_t work_off = scn_off + sizeof(double) * scn_n;
    const size_t bytes = work_off + (in_device ? sizeof(double) * (size_t)a.work_doubles *
                                                     (size_t)grid : 0);
    char *buf = NULL;
    rc = ph_hip(hipMallocAsync((void **)&buf, bytes, st), "hipMallocAsync");
    if (rc) return rc;
    unsigned long long *queue = (unsigned long long *)buf;
    /* scn is pageable (this frame's array): HIP copies from pageable memory synchronously,
     * so the array may go out of scope as soon as the call returns */
    rc = ph_hip(hipMemcpyAsync(buf + scn_off, scn, sizeof(double) * scn_n,
                               hipMemcpyHostToDevice, st),
                "per-call tables");
    if (rc == PHOTIC_OK)
        rc = ph_launch_status(ph_launch_zero_u64(queue, 2, st), "zero work counter");
    if (rc == PHOTIC_OK) {
        a.queue = queue;
        a.scn = (const double *)(buf + scn_off);
        a.work = in_device ? (double *)(buf + work_off) : NULL;
        rc = ph_launch_status(ph_launch_region(&a, (int)grid, st), "region kernel");
    }
    if (rc == PHOTIC_OK && cfg->check_watchdog) {
        unsigned long long flag = 0;
        rc = ph_hip(hipMemcpyAsync(&flag, queue + 1, sizeof(flag), hipMemcpyDeviceToHost, st),
                    "watchdog read");
        if (rc == PHOTIC_OK) rc = ph_hip(hipStreamSynchronize(st), "hipStreamSynchronize");
        if (rc == PHOTIC_OK && flag)
            rc = ph_fail(PHOTIC_EDEVICE, "region kernel stopped by its %.0f s watchdog", limit_s);
    }
    int rc2 = ph_hip(hipFreeAsync(buf, st), "hipFreeAsync");
    return rc ? rc : rc2;
}

int photic_gen_stack(const photic_cfg *cfg, uint64_t seed, long pix0, long n, int n_dates,
                     double sigma, double p_cloud, int k_min, float *rrs, float *sec_sun,
                     unsigned char *valid, float *truth, void *stream)
{
    PhConst c;
    int rc = ph_have_device();
    if (rc) return rc;
    rc = ph_make_const(cfg, 1, 0, &c);
    if (rc) return rc;
    if (n < 0 || pix0 < 0) return ph_fail(PHOTIC_EINVAL, "negative pixel range");
    if (n_dates < 1 || n_dates > PH_MAXD)
        return ph_fail(PHOTIC_EINVAL, "n_dates %d outside 1..%d", n_dates, PH_MAXD);
    if (n > 0 && (!rrs || !sec_sun || !valid))
        return ph_fail(PHOTIC_EINVAL, "rrs, sec_sun and valid are required");
    return ph_launch_status(ph_launch_gen_stack(&c, seed, pix0, n, n_dates, sigma, p_cloud,
                                                k_min, rrs, sec_sun, valid, truth, stream),
                            "stack generator");
}

/* ------------------------------------------------------------------------------------ */
/* the launch-shape autotuner                                                             */
/* ------------------------------------------------------------------------------------ */

/* Time one inversion call of the candidate shape on `st` (HIP events), in ms. */
static int ph_time_call(int stack, const photic_cfg *c, long n, int n_dates, const float *rrs,
                        const float *ss, const unsigned char *valid, const float *seed,
                        const photic_outputs *o, hipStream_t st, float *ms)
{
    hipEvent_t e0, e1;
    int rc = ph_hip(hipEventCreate(&e0), "hipEventCreate");
    if (rc) return rc;
    rc = ph_hip(hipEventCreate(&e1), "hipEventCreate");
    if (rc) {
        hipEventDestroy(e0);
        return rc;
    }
    rc = ph_hip(hipEventRecord(e0, st), "hipEventRecord");
    if (!rc)
        rc = stack ? photic_invert_stack(c, n, n_dates, rrs, ss, valid, seed, o, st)
                   : photic_invert_h(c, n, rrs, ss, NULL, seed, NULL, o, st);
    if (!rc) rc = ph_hip(hipEventRecord(e1, st), "hipEventRecord");
    if (!rc) rc = ph_hip(hipEventSynchronize(e1), "hipEventSynchronize");
    if (!rc) rc = ph_hip(hipEventElapsedTime(ms, e0, e1), "hipEventElapsedTime");
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return rc;
}

static int ph_autotune(int stack, const photic_cfg *cfg, long n_pix, int n_dates,
                       const float *rrs, const float *ss, const unsigned char *valid,
                       const float *seed, photic_tuning *best, void *stream)
{
    int rc = ph_have_device();
    if (rc) return rc;
    if (!cfg || n_pix < 64 || !rrs) return ph_fail(PHOTIC_EINVAL, "a sample of >= 64 pixels");
    const int fp32 = cfg->precision == PHOTIC_FP32;
    const int nbot = cfg->mode == PHOTIC_MODE_MIX ? 0 : cfg->n_bottoms;
    if (nbot < 0 || nbot > 3) return ph_fail(PHOTIC_EINVAL, "n_bottoms %d", nbot);
    ph_shape *slot = stack ? &ph_tuned_stack[seed != NULL] : &ph_tuned_single[fp32][nbot][seed != NULL];
    hipStream_t st = (hipStream_t)stream;
    int dev = 0, max_lds = 0, n_cu = 0;
    rc = ph_hip(hipGetDevice(&dev), "hipGetDevice");
    if (!rc) rc = ph_hip(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev), "attr");
    if (!rc) rc = ph_n_cu(&n_cu);
    if (rc) return rc;
    int *d_status = NULL;
    rc = ph_hip(hipMalloc((void **)&d_status, sizeof(int) * (size_t)n_pix), "hipMalloc");
    if (rc) return rc;
    photic_outputs o;
    memset(&o, 0, sizeof(o));
    o.status = d_status;
    photic_cfg c = *cfg;
    c.check_watchdog = 0;
    const int blocks[4] = {64, 128, 192, 256};
    const int chunks[5] = {16, 64, 32, 128, 256};
    ph_shape win = {0, 0, 0};
    float win_ms = 1e30f, ms = 0.0f;
    int tried = 0;
    *slot = win; /* candidates run with explicit shapes */
    /* stacks: every date count's launch runs at its own full occupancy (the group size, and
     * with it the LDS per block, depends on the date count), so only the block size and the
     * chunk are swept */
    const int n_blocks = 4, n_chunks = stack ? 2 : 5, n_halves = stack ? 1 : 2;
    for (int bi = 0; bi < n_blocks && !rc; bi++) {
        const int block = blocks[bi];
        if (stack && block == 192) continue;
        if (!stack && (long)ph_lds_bytes(nbot, 1) * block > max_lds) continue;
        int bpc = 0;
        rc = stack ? ph_launch_status(ph_stack_occupancy(n_dates, nbot, seed != NULL, block, &bpc), "occupancy")
                   : ph_launch_status(ph_kernel_occupancy(fp32, cfg->n_bands, nbot, seed != NULL,
                                                          block, &bpc), "occupancy");
        if (rc || bpc < 1) continue;
        for (int half = 0; half < n_halves && !rc; half++) {
            const int b = half ? bpc / 2 : bpc;
            if (b < 1 || (half && b == bpc)) continue;
            for (int ci = 0; ci < n_chunks && !rc; ci++) {
                c.block = block;
                c.grid = stack ? 0 : n_cu * b;
                c.chunk = chunks[ci];
                if (tried == 0) /* first launch of the instance: load the code object */
                    rc = ph_time_call(stack, &c, n_pix, n_dates, rrs, ss, valid, seed, &o, st, &ms);
                if (!rc)
                    rc = ph_time_call(stack, &c, n_pix, n_dates, rrs, ss, valid, seed, &o, st, &ms);
                tried++;
                if (!rc && ms < win_ms) {
                    win_ms = ms;
                    win.block = block;
                    win.bpc = stack ? 0 : b;
                    win.chunk = chunks[ci];
                }
            }
        }
    }
    hipFree(d_status);
    if (rc) return rc;
    *slot = win;
    if (best) {
        best->block = win.block;
        best->blocks_per_cu = win.bpc;
        best->chunk = win.chunk;
        best->ms = win_ms;
        best->n_tried = tried;
    }
    return PHOTIC_OK;
}

int photic_autotune(const photic_cfg *cfg, long n_pix, const float *rrs, const float *sec_sun,
                    const float *seed, photic_tuning *best, void *stream)
{
    return ph_autotune(0, cfg, n_pix, 1, rrs, sec_sun, NULL, seed, best, stream);
}

int photic_autotune_stack(const photic_cfg *cfg, long n_pix, int n_dates, const float *rrs,
                          const float *sec_sun, const unsigned char *valid, const float *seed,
                          photic_tuning *best, void *stream)
{
    return ph_autotune(1, cfg, n_pix, n_dates, rrs, sec_sun, valid, seed, best, stream);
}

int photic_gen_scene(const photic_cfg *cfg, uint64_t seed, long pix0, long n, double sigma,
                     double deep_frac, float *rrs, float *sec_sun, float *truth, void *stream)
{
    PhConst c;
    int rc = ph_have_device();
    if (rc) return rc;
    rc = ph_make_const(cfg, 1, 0, &c);
    if (rc) return rc;
    if (n < 0 || pix0 < 0) return ph_fail(PHOTIC_EINVAL, "negative pixel range");
    if (n > 0 && (!rrs || !sec_sun)) return ph_fail(PHOTIC_EINVAL, "rrs and sec_sun are required");
    return ph_launch_status(
        ph_launch_gen(&c, seed, pix0, n, sigma, deep_frac, rrs, sec_sun, truth, stream),
        "scene generator");
}

int photic_gen_field(const photic_cfg *cfg, uint64_t seed, long n_rows, long n_cols,
                     double theta_sun_deg, double sigma, float *rrs, float *truth, void *stream)
{
    PhConst c;
    int rc = ph_have_device();
    if (rc) return rc;
    rc = ph_make_const(cfg, 1, 0, &c);
    if (rc) return rc;
    if (n_rows < 0 || n_cols < 0) return ph_fail(PHOTIC_EINVAL, "negative scene size");
    if (n_rows * n_cols > 0 && !rrs) return ph_fail(PHOTIC_EINVAL, "rrs is required");
    return ph_launch_status(ph_launch_gen_field(&c, seed, n_rows, n_cols, theta_sun_deg, sigma,
                                                rrs, truth, stream),
                            "field generator");
}

int photic_gen_nodata_discs(uint64_t seed, long n_rows, long n_cols, int n_discs, double r_min,
                            double r_max, int n_bands, float nodata, float *rrs, void *stream)
{
    int rc = ph_have_device();
    if (rc) return rc;
    if (n_rows < 0 || n_cols < 0 || n_discs < 0 || n_bands < 1 || n_bands > PH_MAXB ||
        !(r_min >= 0.0) || !(r_max >= r_min))
        return ph_fail(PHOTIC_EINVAL, "bad disc mask arguments");
    if (n_rows * n_cols > 0 && !rrs) return ph_fail(PHOTIC_EINVAL, "rrs is required");
    return ph_launch_status(ph_launch_nodata_discs(seed, n_rows, n_cols, n_discs, r_min, r_max,
                                                   n_bands, nodata, rrs, stream),
                            "no-data discs");
}

int photic_forward(const photic_cfg *cfg, long n, const double *params, const double *sec_sun,
                   const double *Y, double *out, int precision, void *stream)
{
    PhConst c;
    int rc = ph_have_device();
    if (rc) return rc;
    rc = ph_make_const(cfg, 1, 0, &c);
    if (rc) return rc;
    if (n < 0) return ph_fail(PHOTIC_EINVAL, "n < 0");
    return ph_launch_status(ph_launch_fwd(&c, n, params, sec_sun, Y, out,
                                          precision == PHOTIC_FP32, stream),
                            "forward kernel");
}

int photic_objective(const photic_cfg *cfg, long n, const double *x, const float *rrs,
                     const float *sec_sun, double *out, double *rrs_error, int precision,
                     void *stream)
{
    PhConst c;
    int rc = ph_have_device();
    if (rc) return rc;
    rc = ph_make_const(cfg, 1, 0, &c);
    if (rc) return rc;
    if (n < 0) return ph_fail(PHOTIC_EINVAL, "n < 0");
    return ph_launch_status(ph_launch_objective(&c, n, x, rrs, sec_sun, out, rrs_error,
                                                precision == PHOTIC_FP32, c.n_bands, stream),
                            "objective kernel");
}

/* ------------------------------------------------------------------------------------ */
/* the learned warm start                                                                 */
/* ------------------------------------------------------------------------------------ */

struct photic_mlp {
    float *w, *m, *v, *g; /* FP32 masters, Adam moments, gradient */
    float *part;          /* the gradient's partial sums, one copy per batch slice */
    unsigned short *pk;   /* packed BF16 weights */
    float *pb;            /* packed FP32 biases */
    float *loss;
    int step;
    float skip_h; /* seeds with a predicted depth below this get no seed start (0: none) */
};

int photic_mlp_set_seed_skip(photic_mlp *m, double h_max)
{
    if (!m || !(h_max >= 0.0)) return ph_fail(PHOTIC_EINVAL, "mlp and h_max >= 0 required");
    m->skip_h = (float)h_max;
    return PHOTIC_OK;
}

int photic_mlp_n_params(void) { return ph_mlp_nparam(); }

static int ph_mlp_pack(photic_mlp *m, void *stream)
{
    return ph_launch_status(ph_launch_mlp_adam(m->w, m->m, m->v, NULL, 0.f, 0.9f, 0.999f, 1e-8f, 1,
                                               m->pk, m->pb, stream), "pack weights");
}

void photic_mlp_destroy(photic_mlp *m)
{
    if (!m) return;
    hipFree(m->w);
    hipFree(m->m);
    hipFree(m->v);
    hipFree(m->g);
    hipFree(m->part);
    hipFree(m->pk);
    hipFree(m->pb);
    hipFree(m->loss);
    free(m);
}

int photic_mlp_create(photic_mlp **out, uint64_t init_seed)
{
    int rc = ph_have_device();
    if (rc) return rc;
    if (!out) return ph_fail(PHOTIC_EINVAL, "out is NULL");
    photic_mlp *m = (photic_mlp *)calloc(1, sizeof(*m));
    if (!m) return ph_fail(PHOTIC_ENOMEM, "host allocation");
    int np = ph_mlp_nparam(), nbf, nf;
    ph_mlp_packed_sizes(&nbf, &nf);
    rc = ph_hip(hipMalloc((void **)&m->w, sizeof(float) * np), "hipMalloc");
    if (!rc) rc = ph_hip(hipMalloc((void **)&m->m, sizeof(float) * np), "hipMalloc");
    if (!rc) rc = ph_hip(hipMalloc((void **)&m->v, sizeof(float) * np), "hipMalloc");
    if (!rc) rc = ph_hip(hipMalloc((void **)&m->g, sizeof(float) * np), "hipMalloc");
    if (!rc)
        rc = ph_hip(hipMalloc((void **)&m->part, sizeof(float) * np * PH_MLP_KSPLIT), "hipMalloc");
    if (!rc) rc = ph_hip(hipMalloc((void **)&m->pk, sizeof(unsigned short) * nbf), "hipMalloc");
    if (!rc) rc = ph_hip(hipMalloc((void **)&m->pb, sizeof(float) * nf), "hipMalloc");
    if (!rc) rc = ph_hip(hipMalloc((void **)&m->loss, sizeof(float)), "hipMalloc");
    if (!rc) rc = ph_hip(hipMemset(m->m, 0, sizeof(float) * np), "hipMemset");
    if (!rc) rc = ph_hip(hipMemset(m->v, 0, sizeof(float) * np), "hipMemset");
    if (!rc) rc = ph_hip(hipMemset(m->pk, 0, sizeof(unsigned short) * nbf), "hipMemset");
    if (!rc) rc = ph_hip(hipMemset(m->pb, 0, sizeof(float) * nf), "hipMemset");
    if (!rc) rc = ph_launch_status(ph_launch_mlp_init(init_seed, m->w, NULL), "init weights");
    if (!rc) rc = ph_mlp_pack(m, NULL);
    if (!rc) rc = ph_hip(hipDeviceSynchronize(), "hipDeviceSynchronize");
    if (rc) {
        photic_mlp_destroy(m);
        return rc;
    }
    *out = m;
    return PHOTIC_OK;
}

int photic_mlp_set_weights(photic_mlp *m, const float *host_w)
{
    if (!m || !host_w) return ph_fail(PHOTIC_EINVAL, "NULL argument");
    int rc = ph_hip(hipMemcpy(m->w, host_w, sizeof(float) * ph_mlp_nparam(), hipMemcpyHostToDevice),
                    "hipMemcpy");
    if (!rc) rc = ph_mlp_pack(m, NULL);
    if (!rc) rc = ph_hip(hipDeviceSynchronize(), "hipDeviceSynchronize");
    return rc;
}

int photic_mlp_get_weights(photic_mlp *m, float *host_w)
{
    if (!m || !host_w) return ph_fail(PHOTIC_EINVAL, "NULL argument");
    return ph_hip(hipMemcpy(host_w, m->w, sizeof(float) * ph_mlp_nparam(), hipMemcpyDeviceToHost),
                  "hipMemcpy");
}

int photic_mlp_trainset(const photic_cfg *cfg, uint64_t seed, long pix0, long n, double sigma,
                        float *x, float *t, void *stream)
{
    PhConst c;
    int rc = ph_have_device();
    if (rc) return rc;
    rc = ph_make_const(cfg, 1, 0, &c);
    if (rc) return rc;
    if (cfg->n_bands != 4) return ph_fail(PHOTIC_EUNSUPPORTED, "the network takes 4 bands");
    return ph_launch_status(ph_launch_mlp_trainset(&c, seed, pix0, n, sigma, x, t, stream),
                            "training set");
}

/* one step on samples [x, t) of n: gradient into m->g, loss into m->loss (accumulated) */
static int ph_mlp_grad_step(photic_mlp *m, const float *x, const float *t, long n,
                            unsigned short *gbuf, void *stream)
{
    float gscale = 2.0f / ((float)n * 5.0f);
    int rc = ph_launch_status(ph_launch_mlp_fwdbwd(m->pk, m->pb, x, t, NULL, n, gscale, gbuf,
                                                   m->loss, 0, stream), "network forward/backward");
    if (!rc)
        rc = ph_launch_status(ph_launch_mlp_wgrad(gbuf, n, m->g, m->part, stream),
                              "weight gradients");
    return rc;
}

int photic_mlp_gradient(photic_mlp *m, const float *x, const float *t, long n, float *host_grad,
                        float *loss)
{
    if (!m || n <= 0 || n % 16) return ph_fail(PHOTIC_EINVAL, "n must be a positive multiple of 16");
    unsigned short *gbuf = NULL;
    int rc = ph_hip(hipMalloc((void **)&gbuf, sizeof(unsigned short) * ph_mlp_gbuf_elems(n)), "hipMalloc");
    if (rc) return rc;
    rc = ph_hip(hipMemset(m->loss, 0, sizeof(float)), "hipMemset");
    if (!rc) rc = ph_hip(hipMemset(m->g, 0, sizeof(float) * ph_mlp_nparam()), "hipMemset");
    if (!rc) rc = ph_mlp_grad_step(m, x, t, n, gbuf, NULL);
    if (!rc) rc = ph_hip(hipDeviceSynchronize(), "hipDeviceSynchronize");
    if (!rc && host_grad)
        rc = ph_hip(hipMemcpy(host_grad, m->g, sizeof(float) * ph_mlp_nparam(), hipMemcpyDeviceToHost),
                    "hipMemcpy");
    if (!rc && loss) {
        float l = 0.f;
        rc = ph_hip(hipMemcpy(&l, m->loss, sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy");
        *loss = l / ((float)n * 5.0f);
    }
    hipFree(gbuf);
    return rc;
}

int photic_mlp_train(photic_mlp *m, const float *x, const float *t, long n, long batch,
                     int epochs, float lr, float lr_final, float *final_loss, void *stream)
{
    if (!m || batch <= 0 || batch % 16 || n < batch || epochs < 1)
        return ph_fail(PHOTIC_EINVAL, "need 16 | batch <= n and epochs >= 1");
    unsigned short *gbuf = NULL;
    int rc = ph_hip(hipMalloc((void **)&gbuf, sizeof(unsigned short) * ph_mlp_gbuf_elems(batch)),
                    "hipMalloc");
    if (rc) return rc;
    const long steps_per_epoch = n / batch;
    const long total = steps_per_epoch * epochs;
    const double decay = (lr_final > 0.f && lr > 0.f) ? pow((double)lr_final / lr, 1.0 / (double)total) : 1.0;
    double cur = lr;
    float last = 0.f;
    for (int e = 0; e < epochs && !rc; e++) {
        rc = ph_hip(hipMemsetAsync(m->loss, 0, sizeof(float), (hipStream_t)stream), "hipMemsetAsync");
        for (long k = 0; k < steps_per_epoch && !rc; k++) {
            const float *xb = x + k * batch * 5, *tb = t + k * batch * 5;
            rc = ph_mlp_grad_step(m, xb, tb, batch, gbuf, stream);
            if (!rc) {
                m->step++;
                rc = ph_launch_status(ph_launch_mlp_adam(m->w, m->m, m->v, m->g, (float)cur, 0.9f,
                                                         0.999f, 1e-8f, m->step, m->pk, m->pb,
                                                         stream), "adam");
            }
            cur *= decay;
        }
        if (!rc) {
            float l = 0.f;
            rc = ph_hip(hipMemcpyAsync(&l, m->loss, sizeof(float), hipMemcpyDeviceToHost,
                                       (hipStream_t)stream), "hipMemcpyAsync");
            if (!rc) rc = ph_hip(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
            last = l / ((float)steps_per_epoch * (float)batch * 5.0f);
        }
    }
    if (final_loss) *final_loss = last;
    hipFree(gbuf);
    return rc;
}

int photic_mlp_forward(photic_mlp *m, const float *x, long n, float *y, void *stream)
{
    if (!m || n < 0) return ph_fail(PHOTIC_EINVAL, "bad arguments");
    return ph_launch_status(ph_launch_mlp_fwdbwd(m->pk, m->pb, x, NULL, y, n, 0.f, NULL, NULL, 0,
                                                 stream), "network forward");
}

int photic_mlp_seed(photic_mlp *m, const photic_cfg *cfg, long n, const float *rrs,
                    const float *sec_sun, float *seed, void *stream)
{
    if (!m || !cfg || n < 0) return ph_fail(PHOTIC_EINVAL, "bad arguments");
    if (cfg->n_bands != 4 || cfg->mode != PHOTIC_MODE_MIX)
        return ph_fail(PHOTIC_EUNSUPPORTED, "seeds are made for 4 bands in MIX mode");
    if (n == 0) return PHOTIC_OK;
    hipStream_t st = (hipStream_t)stream;
    float *xb = NULL, *yb = NULL;
    int rc = ph_hip(hipMallocAsync((void **)&xb, sizeof(float) * 5 * n, st), "hipMallocAsync");
    if (rc) return rc;
    rc = ph_hip(hipMallocAsync((void **)&yb, sizeof(float) * 5 * n, st), "hipMallocAsync");
    if (!rc) rc = ph_launch_status(ph_launch_mlp_prep(rrs, sec_sun, (float)cfg->sec_sun, n, xb, st),
                                   "network inputs");
    if (!rc) rc = photic_mlp_forward(m, xb, n, yb, st);
    if (!rc) rc = ph_launch_status(ph_launch_mlp_seed(yb, n, m->skip_h, seed, st), "seeds");
    hipFreeAsync(xb, st);
    if (yb) hipFreeAsync(yb, st);
    return rc;
}

int photic_mlp_seed_stack(photic_mlp *m, const photic_cfg *cfg, long n, int n_dates,
                          const float *rrs, const float *sec_sun, const unsigned char *valid,
                          float *seed, void *stream)
{
    if (!m || !cfg || n < 0) return ph_fail(PHOTIC_EINVAL, "bad arguments");
    if (cfg->n_bands != 4 || cfg->mode != PHOTIC_MODE_MIX)
        return ph_fail(PHOTIC_EUNSUPPORTED, "seeds are made for 4 bands in MIX mode");
    if (n_dates < 1 || n_dates > PH_MAXD)
        return ph_fail(PHOTIC_EINVAL, "n_dates %d outside 1..%d", n_dates, PH_MAXD);
    if (n == 0) return PHOTIC_OK;
    if (!rrs || !sec_sun || !valid || !seed)
        return ph_fail(PHOTIC_EINVAL, "rrs, sec_sun, valid and seed are required");
    hipStream_t st = (hipStream_t)stream;
    float *xb = NULL, *yb = NULL;
    int rc = ph_hip(hipMallocAsync((void **)&xb, sizeof(float) * 5 * n, st), "hipMallocAsync");
    if (rc) return rc;
    rc = ph_hip(hipMallocAsync((void **)&yb, sizeof(float) * 5 * n, st), "hipMallocAsync");
    for (int d = 0; d < n_dates && !rc; d++) {
        rc = ph_launch_status(ph_launch_mlp_prep(rrs + (long)d * 4 * n, sec_sun + (long)d * n,
                                                 (float)cfg->sec_sun, n, xb, st),
                              "network inputs");
        if (!rc) rc = photic_mlp_forward(m, xb, n, yb, st);
        if (!rc)
            rc = ph_launch_status(ph_launch_mlp_stack_seed(yb, n, d, valid, seed, st), "seeds");
    }
    if (!rc)
        rc = ph_launch_status(ph_launch_mlp_stack_mean(n, n_dates, valid, seed, st), "seeds");
    hipFreeAsync(xb, st);
    if (yb) hipFreeAsync(yb, st);
    return rc;
}

/* ------------------------------------------------------------------------------------ */
/* streamed stack inversion of a host-resident scene                                      */
/* ------------------------------------------------------------------------------------ */

/* One output field of the stream: host array, element size, rows of n. */
typedef struct {
    void *host;
    size_t esz;
    int rows;
    size_t off; /* byte offset of the field in a tile's output block (rows x tile pixels) */
} ph_sfield;

typedef struct {
    unsigned char *pin_in, *dev_in, *pin_out, *dev_out;
    float *dev_seed;
    hipEvent_t ev_in, ev_comp, ev_done;
    long p0, nt; /* the tile in flight: first pixel and pixel count (nt 0: none) */
} ph_sslot;

int photic_invert_stack_stream(const photic_cfg *cfg, photic_mlp *mlp, long n_rows, long n_cols,
                               int n_dates, const float *rrs, const float *sec_sun,
                               const unsigned char *valid, long tile_rows,
                               const photic_outputs *out, photic_progress_fn progress,
                               void *user)
{
    int rc = ph_have_device();
    if (rc) return rc;
    if (!cfg || !out) return ph_fail(PHOTIC_EINVAL, "cfg and out are required");
    if (n_rows < 0 || n_cols < 0 || tile_rows < 1)
        return ph_fail(PHOTIC_EINVAL, "n_rows, n_cols >= 0 and tile_rows >= 1");
    if (n_dates < 1 || n_dates > PH_MAXD)
        return ph_fail(PHOTIC_EINVAL, "n_dates %d outside 1..%d", n_dates, PH_MAXD);
    const long n = n_rows * n_cols;
    if (n == 0) return PHOTIC_OK;
    if (!rrs || !sec_sun || !valid)
        return ph_fail(PHOTIC_EINVAL, "rrs, sec_sun and valid are required");
    const long D = n_dates;
    const long tile_px = (tile_rows < n_rows ? tile_rows : n_rows) * n_cols;
    const long n_tiles = (n_rows + tile_rows - 1) / tile_rows;

    ph_sfield f[13] = {
        {out->depth, 4, 1, 0},       {out->rrs_error, 4, 1, 0}, {out->model_error, 4, 1, 0},
        {out->albedo, 4, 1, 0},      {out->k_min, 4, 1, 0},     {out->iod, 4, 1, 0},
        {out->b_pct, 4, 3, 0},       {out->bottom_type, 4, 1, 0},
        {out->x, 8, (int)(2 + 3 * D), 0},
        {out->status, 4, 1, 0},      {out->n_evals, 4, 1, 0},   {out->n_iter, 4, 1, 0},
        {out->n_starts, 4, 1, 0}};
    size_t out_per_px = 0;
    for (int k = 0; k < 13; k++) {
        f[k].off = out_per_px; /* times the tile's pixel count */
        if (f[k].host) out_per_px += f[k].esz * (size_t)f[k].rows;
    }
    const size_t in_per_px = (size_t)D * (16 + 4 + 1);

    hipStream_t cs = NULL, xs = NULL; /* compute, copy */
    ph_sslot sl[2];
    memset(sl, 0, sizeof(sl));
    rc = ph_hip(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "hipStreamCreate");
    if (!rc) rc = ph_hip(hipStreamCreateWithFlags(&xs, hipStreamNonBlocking), "hipStreamCreate");
    for (int s = 0; s < 2 && !rc; s++) {
        rc = ph_hip(hipHostMalloc((void **)&sl[s].pin_in, in_per_px * tile_px, 0), "hipHostMalloc");
        if (!rc) rc = ph_hip(hipMalloc((void **)&sl[s].dev_in, in_per_px * tile_px), "hipMalloc");
        if (!rc && out_per_px)
            rc = ph_hip(hipHostMalloc((void **)&sl[s].pin_out, out_per_px * tile_px, 0),
                        "hipHostMalloc");
        if (!rc && out_per_px)
            rc = ph_hip(hipMalloc((void **)&sl[s].dev_out, out_per_px * tile_px), "hipMalloc");
        if (!rc && mlp)
            rc = ph_hip(hipMalloc((void **)&sl[s].dev_seed, sizeof(float) * (2 + 3 * D) * tile_px),
                        "hipMalloc");
        if (!rc) rc = ph_hip(hipEventCreateWithFlags(&sl[s].ev_in, hipEventDisableTiming), "event");
        if (!rc) rc = ph_hip(hipEventCreateWithFlags(&sl[s].ev_comp, hipEventDisableTiming), "event");
        if (!rc) rc = ph_hip(hipEventCreateWithFlags(&sl[s].ev_done, hipEventDisableTiming), "event");
    }
    /* cfg->check_watchdog makes every tile's inversion wait for its kernels (to read their
     * watchdog flags), so the next tile's staging then no longer overlaps the compute */
    const photic_cfg c1 = *cfg;

    /* Tile t is staged, copied in, inverted and copied out in slot t % 2; slot t % 2's
     * previous tile (t - 2) is retired first: its outputs leave the pinned buffer, which
     * the CPU refills with tile t while the GPU works on tile t - 1. */
    long done = 0;
    for (long t = 0; t < n_tiles + 2 && !rc; t++) {
        ph_sslot *s = &sl[t % 2];
        if (s->nt > 0) {
            rc = ph_hip(hipEventSynchronize(s->ev_done), "tile");
            if (rc) break;
            for (int k = 0; k < 13; k++) {
                if (!f[k].host) continue;
                for (int r = 0; r < f[k].rows; r++)
                    memcpy((char *)f[k].host + ((size_t)r * n + s->p0) * f[k].esz,
                           s->pin_out + f[k].off * s->nt + (size_t)r * s->nt * f[k].esz,
                           (size_t)s->nt * f[k].esz);
            }
            s->nt = 0;
            done++;
            if (progress) progress(done, n_tiles, user);
        }
        if (t >= n_tiles) continue;
        const long p0 = t * tile_px;
        const long nt = (n - p0 < tile_px) ? n - p0 : tile_px;
        /* staging: rrs [D][4][nt] | sec_sun [D][nt] | valid [D][nt] */
        float *pr = (float *)s->pin_in, *ps = pr + 4 * D * nt;
        unsigned char *pv = (unsigned char *)(ps + D * nt);
        for (long q = 0; q < 4 * D; q++)
            memcpy(pr + q * nt, rrs + q * n + p0, sizeof(float) * (size_t)nt);
        for (long d = 0; d < D; d++) {
            memcpy(ps + d * nt, sec_sun + d * n + p0, sizeof(float) * (size_t)nt);
            memcpy(pv + d * nt, valid + d * n + p0, (size_t)nt);
        }
        rc = ph_hip(hipMemcpyAsync(s->dev_in, s->pin_in, in_per_px * (size_t)nt,
                                   hipMemcpyHostToDevice, xs), "tile in");
        if (!rc) rc = ph_hip(hipEventRecord(s->ev_in, xs), "hipEventRecord");
        if (!rc) rc = ph_hip(hipStreamWaitEvent(cs, s->ev_in, 0), "hipStreamWaitEvent");
        const float *dr = (const float *)s->dev_in, *dss = dr + 4 * D * nt;
        const unsigned char *dv = (const unsigned char *)(dss + D * nt);
        if (!rc && mlp)
            rc = photic_mlp_seed_stack(mlp, cfg, nt, n_dates, dr, dss, dv, s->dev_seed, cs);
        photic_outputs o;
        memset(&o, 0, sizeof(o));
        void **dst[13] = {(void **)&o.depth, (void **)&o.rrs_error, (void **)&o.model_error,
                          (void **)&o.albedo, (void **)&o.k_min, (void **)&o.iod,
                          (void **)&o.b_pct, (void **)&o.bottom_type, (void **)&o.x,
                          (void **)&o.status, (void **)&o.n_evals, (void **)&o.n_iter,
                          (void **)&o.n_starts};
        for (int k = 0; k < 13; k++)
            if (f[k].host) *dst[k] = s->dev_out + f[k].off * nt;
        if (!rc)
            rc = photic_invert_stack(&c1, nt, n_dates, dr, dss, dv, mlp ? s->dev_seed : NULL,
                                     &o, cs);
        if (!rc) rc = ph_hip(hipEventRecord(s->ev_comp, cs), "hipEventRecord");
        if (!rc) rc = ph_hip(hipStreamWaitEvent(xs, s->ev_comp, 0), "hipStreamWaitEvent");
        if (!rc && out_per_px)
            rc = ph_hip(hipMemcpyAsync(s->pin_out, s->dev_out, out_per_px * (size_t)nt,
                                       hipMemcpyDeviceToHost, xs), "tile out");
        if (!rc) rc = ph_hip(hipEventRecord(s->ev_done, xs), "hipEventRecord");
        if (!rc) {
            s->p0 = p0;
            s->nt = nt;
        }
    }
    if (cs) hipStreamSynchronize(cs);
    if (xs) hipStreamSynchronize(xs);
    for (int s = 0; s < 2; s++) {
        if (sl[s].pin_in) hipHostFree(sl[s].pin_in);
        if (sl[s].pin_out) hipHostFree(sl[s].pin_out);
        if (sl[s].dev_in) hipFree(sl[s].dev_in);
        if (sl[s].dev_out) hipFree(sl[s].dev_out);
        if (sl[s].dev_seed) hipFree(sl[s].dev_seed);
        if (sl[s].ev_in) hipEventDestroy(sl[s].ev_in);
        if (sl[s].ev_comp) hipEventDestroy(sl[s].ev_comp);
        if (sl[s].ev_done) hipEventDestroy(sl[s].ev_done);
    }
    if (cs) hipStreamDestroy(cs);
    if (xs) hipStreamDestroy(xs);
    return rc;
}

/* ------------------------------------------------------------------------------------ */
/* device plumbing                                                                        */
/* ------------------------------------------------------------------------------------ */

int photic_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int photic_device_malloc(void **ptr, size_t bytes)
{
    int rc = ph_have_device();
    if (rc) return rc;
    return ph_hip(hipMalloc(ptr, bytes ? bytes : 1), "hipMalloc");
}

int photic_device_free(void *ptr) { return ph_hip(hipFree(ptr), "hipFree"); }

int photic_host_malloc(void **ptr, size_t bytes)
{
    int rc = ph_have_device();
    if (rc) return rc;
    return ph_hip(hipHostMalloc(ptr, bytes ? bytes : 1, 0), "hipHostMalloc");
}

int photic_host_free(void *ptr) { return ph_hip(hipHostFree(ptr), "hipHostFree"); }

int photic_memcpy_htod(void *dst, const void *src, size_t bytes)
{
    return ph_hip(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "hipMemcpy H2D");
}

int photic_memcpy_dtoh(void *dst, const void *src, size_t bytes)
{
    return ph_hip(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "hipMemcpy D2H");
}

int photic_memset_zero(void *dst, size_t bytes)
{
    return ph_hip(hipMemset(dst, 0, bytes), "hipMemset");
}

int photic_device_copy(void *dst, const void *src, size_t bytes, void *stream)
{
    if (bytes % 16 != 0 || ((uintptr_t)dst | (uintptr_t)src) % 16 != 0)
        return ph_fail(PHOTIC_EINVAL, "device copy needs 16-byte sizes and alignment");
    return ph_launch_status(ph_launch_copy16(src, dst, (long)(bytes / 16), stream), "copy kernel");
}

int photic_synchronize(void) { return ph_hip(hipDeviceSynchronize(), "hipDeviceSynchronize"); }

int photic_stream_create(void **stream)
{
    int rc = ph_have_device();
    if (rc) return rc;
    return ph_hip(hipStreamCreateWithFlags((hipStream_t *)stream, hipStreamNonBlocking),
                  "hipStreamCreate");
}

int photic_stream_destroy(void *stream)
{
    return ph_hip(hipStreamDestroy((hipStream_t)stream), "hipStreamDestroy");
}

int photic_event_create(void **event)
{
    int rc = ph_have_device();
    if (rc) return rc;
    return ph_hip(hipEventCreate((hipEvent_t *)event), "hipEventCreate");
}

int photic_event_destroy(void *event)
{
    return ph_hip(hipEventDestroy((hipEvent_t)event), "hipEventDestroy");
}

int photic_event_record(void *event, void *stream)
{
    return ph_hip(hipEventRecord((hipEvent_t)event, (hipStream_t)stream), "hipEventRecord");
}

int photic_event_elapsed_ms(void *start, void *stop, float *ms)
{
    hipError_t e = hipEventSynchronize((hipEvent_t)stop);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop);
    return ph_hip(e, "hipEventElapsedTime");
}
