/*
 * photic.h -- MI355X (gfx950) inversion of photic's per-pixel semi-analytical shallow-water
 * model: the HOPE-style forward model of model/samodel.c:2846-2949 minimised per pixel by
 * the modified AS 47 Nelder-Mead of model/asa047.c:10-502, driven by the multi-start of
 * model/samodel.c:2119-2427.
 *
 * Plain C ABI: plain pointers and sizes, int status codes, no exit().  Every array argument
 * of photic_invert and of the test entry points is a DEVICE pointer (allocate with
 * photic_device_malloc or any HIP allocator); arrays are planar, [k][n_pix].  The drop-in
 * for the reference's own entry point, samodel(), is declared in photic_samodel.h.
 *
 * There is no CPU fallback: when no GPU is present every entry point that computes returns
 * PHOTIC_EDEVICE.
 */
#ifndef PHOTIC_H
#define PHOTIC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PHOTIC_MAX_BANDS 8
#define PHOTIC_MAX_HSTARTS 16
#define PHOTIC_N_BOTTOM_TYPES 8

enum photic_status {
    PHOTIC_OK = 0,
    PHOTIC_EINVAL = 1,       /* bad argument (message in photic_last_error) */
    PHOTIC_EUNSUPPORTED = 2, /* a configuration the compiled kernels do not cover */
    PHOTIC_EDEVICE = 3,      /* no GPU, or a HIP runtime error */
    PHOTIC_ENOMEM = 4
};

/* Parameter vector of one pixel.  NATIVE is the reference's own layout with one region
 * (model/samodel.c:1808-1815): [H, B_0..B_{nb-1}, q_0..q_{nb-1}, (P, G, X) per date], P, G,
 * X and B scaled by 100.  MIX is the five-parameter form [H, B, P, G, X]: the reference's
 * two-bottom objective with both amplitudes fixed at 1 (B_k = 100) and q = (|B|, |1 - B|),
 * i.e. a sand/seagrass mix. */
enum photic_mode { PHOTIC_MODE_NATIVE = 0, PHOTIC_MODE_MIX = 1 };

/* Arithmetic of the forward model.  FP64 follows the reference's operations in order (the
 * default, and what parity is claimed for).  FP32 takes exp and log from the GPU's
 * single-precision hardware units (with an exact FP64 range reduction) and keeps all other
 * arithmetic, the misfit sum and every simplex decision in FP64: the forward model stays
 * within 1e-6 relative of FP64, but Nelder-Mead endpoints then differ from the reference's
 * along flat directions of the misfit. */
enum photic_precision { PHOTIC_FP64 = 0, PHOTIC_FP32 = 1 };

typedef struct photic_cfg {
    int n_bands;                             /* bands per date (compiled: 4) */
    double wavelengths[PHOTIC_MAX_BANDS];    /* nm, ascending */
    int mode;                                /* enum photic_mode */
    int n_bottoms;                           /* NATIVE: 1..3 */
    int bottom_idx[PHOTIC_N_BOTTOM_TYPES];   /* NATIVE: rows of the bottom table used */
    double reqmin;                           /* simplex variance limit */
    int konvge;                              /* iterations between convergence checks */
    int kcount;                              /* evaluations per Nelder-Mead run */
    int n_hstarts;                           /* depth starts, tried in order */
    double hstart[PHOTIC_MAX_HSTARTS];
    double early_exit;                       /* accept when best misfit < this; <= 0: 2.5 */
    float nodata;                            /* no-data value of the Rrs input */
    double sec_sun, sec_view;                /* used where the per-pixel arrays are NULL */
    int precision;                           /* enum photic_precision */
    /* Seeded starts: simplex step = max(rel * |seed|, min) per parameter class
     * (H, bottom, P, G, X). */
    double seed_rel_step[5], seed_min_step[5];
    int block, grid, chunk;                  /* launch shape; 0 = tuned defaults */
    double watchdog_s;                       /* a kernel running longer stops (<= 0: 600 s) */
    int check_watchdog;                      /* 1: photic_invert waits for the kernel and
                                                returns PHOTIC_EDEVICE if the watchdog fired */
} photic_cfg;

/* The reference's settings (model/samodel.c:2142-2147, 2216-2218): reqmin 1e-2, konvge 100,
 * kcount 5000, depth starts {40, 30, 20, 15, 10, 7.5, 2.5, 1} m, early exit at misfit 2.5;
 * MIX mode, nodata -9999, nadir geometry. */
void photic_cfg_default(photic_cfg *cfg, int n_bands, const double *wavelengths);
int photic_n_params(const photic_cfg *cfg);

/* Per-pixel outputs (device pointers, each may be NULL), as samodel_optimise leaves them
 * (model/samodel.c:1992-2079).  Unmodelled (no-data) pixels get the reference's defaults
 * (model/samodel.c:819-841) and status -1. */
typedef struct photic_outputs {
    float *depth;       /* H (m), positive */
    float *rrs_error;   /* E_Rrs (%) at the optimum: the reference's model_error grid */
    float *model_error; /* weighted objective at the optimum */
    float *albedo;      /* sum q_i B_i */
    float *k_min;       /* smallest K over bands */
    float *iod;         /* index of optical depth, % */
    float *b_pct;       /* [3][n]: sand, seagrass, coral percent */
    float *bottom_type; /* 1-based index of the dominant bottom */
    double *x;          /* [n_params][n]: |best| parameter vector */
    int *status;        /* 0 converged, 2 not converged (ifault 2), -1 no data */
    int *n_evals;       /* objective evaluations made */
    int *n_iter;        /* icount of the accepted start */
    int *n_starts;      /* Nelder-Mead runs made */
} photic_outputs;

/* Invert n_pix single-date pixels.  rrs: [n_bands][n_pix]; sec_sun, sec_view: [n_pix] or
 * NULL; seed: [n_params][n_pix] start points tried before the depth starts, or NULL.
 * Asynchronous on `stream` (NULL = default stream). */
int photic_invert(const photic_cfg *cfg, long n_pix, const float *rrs, const float *sec_sun,
                  const float *sec_view, const float *seed, const photic_outputs *out,
                  void *stream);

/* Synthetic single-date scene (Philox4x32-10 keyed by seed, counter = pixel index): draws
 * H ~ U(0.5, 25) (deep_frac of pixels U(30, 60)), log-normal P, G, X, B ~ U(0, 1), sun zenith
 * ~ U(20, 50) deg, FP64 forward model with self-consistent Y, N(0, sigma) noise.  rrs:
 * [n_bands][n]; sec_sun: [n]; truth: [6][n] (H, P, G, X, B, theta) or NULL.  Pixel indices
 * pix0 .. pix0 + n - 1. */
int photic_gen_scene(const photic_cfg *cfg, uint64_t seed, long pix0, long n, double sigma,
                     double deep_frac, float *rrs, float *sec_sun, float *truth, void *stream);

/* Test entry points: the forward model alone (params [n][5] = H, P, G, X, B; sec_sun [n];
 * Y [n]; out [n][n_bands]) and the misfit alone (x [n][5] MIX layout; rrs [n][n_bands];
 * sec_sun [n]; out, rrs_error [n]). */
int photic_forward(const photic_cfg *cfg, long n, const double *params, const double *sec_sun,
                   const double *Y, double *out, int precision, void *stream);
int photic_objective(const photic_cfg *cfg, long n, const double *x, const float *rrs,
                     const float *sec_sun, double *out, double *rrs_error, int precision,
                     void *stream);

/* Launch-shape facts of the compiled inversion kernel for cfg. */
int photic_kernel_info(const photic_cfg *cfg, int *vgprs, int *scratch_bytes,
                       int *blocks_per_cu, int *n_cu);

/* Device plumbing for callers without a HIP allocator of their own. */
int photic_device_count(void);
int photic_device_malloc(void **ptr, size_t bytes);
int photic_device_free(void *ptr);
int photic_memcpy_htod(void *dst, const void *src, size_t bytes);
int photic_memcpy_dtoh(void *dst, const void *src, size_t bytes);
int photic_memset_zero(void *dst, size_t bytes);
/* Device-to-device copy by a plain kernel (16-byte aligned sizes): also the known-traffic
 * launch that calibrates the profiler's HBM byte counters. */
int photic_device_copy(void *dst, const void *src, size_t bytes, void *stream);
int photic_synchronize(void);
int photic_stream_create(void **stream);
int photic_stream_destroy(void *stream);
int photic_event_create(void **event);
int photic_event_destroy(void *event);
int photic_event_record(void *event, void *stream);
int photic_event_elapsed_ms(void *start, void *stop, float *ms);
const char *photic_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
