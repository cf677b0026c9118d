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
    double watchdog_s;                       /* a kernel wavefront that finishes no pixel for
                                                this long stops (<= 0: 600 s), leaving its
                                                pixels unwritten */
    int check_watchdog;                      /* 1: the call waits for its kernels and returns
                                                PHOTIC_EDEVICE if a watchdog fired */
    /* Seeded single-date inversions only: a seed start whose best misfit is still above this
     * at its first convergence check is given up (as not converged) and the depth starts
     * follow.  0 (the default) runs every seed start to its end; a level below the early exit
     * acts as the early exit.  The seed start is this library's addition; the reference's
     * depth starts are never cut short. */
    double seed_abort_misfit;
    /* 1: the caller has screened its pixels already (as samodel() does, on the raw values
     * before smoothing, model/samodel.c:933-947): the single-date and stack kernels then invert
     * every pixel and valid date with the values as given, negative ones included, instead of
     * treating a value near cfg->nodata or below zero as no data.  0 (the default): they
     * screen. */
    int prescreened;
    /* Seeded single-date inversions only: iterations between the seed start's convergence
     * checks (0: konvge, as the depth starts). */
    int seed_konvge;
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
    double *rrs_model;  /* [n_bands][n]: the modelled Rrs at the optimum, the reference's
                           Rrs_modelled (single-date inversions; 0 where not modelled) */
} photic_outputs;

/* Invert n_pix single-date pixels.  rrs: [n_bands][n_pix]; sec_sun, sec_view: [n_pix] or
 * NULL; seed: [n_params][n_pix] start points tried before the depth starts, or NULL (a pixel
 * whose seed depth is NaN has no seed start).  Asynchronous on `stream` (NULL = default
 * stream). */
int photic_invert(const photic_cfg *cfg, long n_pix, const float *rrs, const float *sec_sun,
                  const float *sec_view, const float *seed, const photic_outputs *out,
                  void *stream);

/* As photic_invert, with one depth start per pixel (hstart [n_pix], metres) in place of the
 * configured list: the reference's mode when an empirical depth grid is given
 * (model/samodel.c:2208-2231). */
int photic_invert_h(const photic_cfg *cfg, long n_pix, const float *rrs, const float *sec_sun,
                    const float *sec_view, const float *seed, const float *hstart,
                    const photic_outputs *out, void *stream);

/* As photic_invert_h, allocating nothing: no counterpart in the reference; for callers that
 * cannot allocate device memory (code under stream or graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The launch's work counter and watchdog flag live in a
 * fixed 256-byte scratch array that each device holds for the life of the process (2 of its
 * 32 words are used, whatever n_pix is); every call on a device shares that array, so calls
 * that overlap on one device race -- serialise them (one stream, or events).  The array is
 * cleared by a fill kernel on `stream` before each launch.  Results are those of
 * photic_invert_h. */
int photic_invert_noalloc(const photic_cfg *cfg, long n_pix, const float *rrs,
                          const float *sec_sun, const float *sec_view, const float *seed,
                          const float *hstart, const photic_outputs *out, void *stream);

/* Multi-date stacks: each pixel seen on n_dates (<= 8) dates, some of them invalid (clouds).
 * A pixel's problem is the reference's N_s = k run over its k valid dates, in date order
 * (model/samodel.c:1808-1815 layout, one region): H and the bottom shared, P, G, X per date,
 * the misfit over every (date, band), early exit at 2.5 k (cfg->early_exit > 0: that value
 * times k).  The bottom is PHOTIC_MODE_MIX's B (2 + 3k parameters) or PHOTIC_MODE_NATIVE's
 * amplitudes and weights of cfg->n_bottoms (1..3) bottoms (1 + 2 n_bottoms + 3k).  Device
 * arrays: rrs [n_dates][4][n_pix], sec_sun [n_dates][n_pix], valid [n_dates][n_pix] (0 / 1),
 * seed [2 + 3 n_dates][n_pix] or NULL (mix layout only: H, B, then P, G, X at each date's
 * slot; invalid slots ignored).  Outputs as photic_invert; out->x is [off + 3 n_dates][n_pix]
 * (off = the bottom's coordinates plus one) with each valid date's P, G, X at its slot and 0 at
 * invalid slots; pixels with no valid date get the no-data defaults.  4 bands, FP64. */
int photic_invert_stack(const photic_cfg *cfg, long n_pix, int n_dates, const float *rrs,
                        const float *sec_sun, const unsigned char *valid, const float *seed,
                        const photic_outputs *out, void *stream);

/* As photic_invert_stack (unseeded), with per-date view geometry sec_view [n_dates][n_pix]
 * (NULL: cfg->sec_view) and one depth start per pixel, hstart [n_pix] in metres (NULL: the
 * configured list): the reference's multi-scene run with per-scene view angles and an
 * empirical depth grid (model/samodel.c:534-548, 2208-2231). */
int photic_invert_stack_h(const photic_cfg *cfg, long n_pix, int n_dates, const float *rrs,
                          const float *sec_sun, const float *sec_view,
                          const unsigned char *valid, const float *hstart,
                          const photic_outputs *out, void *stream);

/* Spatial regions: the reference's n_spatial > 1 run, replacing the per-pixel loop of
 * model/samodel.c:901-1128 with extract_Rrs_data (:2957-3027), samodel_error over every region
 * (:2432-2759, with the depth and bottom continuity penalties) and the region outputs of
 * samodel_optimise (:1992-2079).  Pixel idx[q] (row * n_cols + col) of an n_rows x n_cols scene
 * is inverted jointly with its (2 n_spatial - 1)^2 neighbours, clamped at the edges, a
 * neighbour with any no-data band dropped: one depth, n_bottoms amplitudes and n_bottoms
 * weights per region, one P, G, X per scene; early exit at 2.5 n_scenes (cfg->early_exit > 0:
 * that value times n_scenes).  rrs [n_scenes][4][n_rows * n_cols] (device; already smoothed
 * when the caller smooths), nodata [4 n_scenes] (host) the no-data value of each band grid,
 * sec_sun and sec_view [n_scenes] (host) each scene's subsurface secants.  idx [n] (device) or
 * NULL for every pixel (n = n_rows * n_cols), each then screened by the reference's rule (any
 * band no-data or negative: not modelled, model/samodel.c:933-947); a listed index outside the
 * scene is not modelled (status -1).  hstart [n] (device): one
 * depth start per listed pixel (an empirical depth), or NULL for the configured list.
 * Outputs per listed pixel as photic_invert: depth is the regions' depths weighted 1 and the
 * centre region sqrt(N_r), bottom percentages and type are the centre region's, IOD is over
 * every region; out->x is [R_max (1 + 2 n_bottoms) + 3 n_scenes][n] with R_max = (2 n_spatial
 * - 1)^2, in the reference's layout [H_r], [B_{r,k}], [q_{r,k}], [(P, G, X)_s] and zero past
 * the pixel's own parameter count.  n_regions [n] (device) or NULL: each pixel's region count
 * (0: not modelled, status -1).  PHOTIC_MODE_NATIVE with 1..3 bottoms, 4 bands, FP64, 1..32
 * scenes (the reference's MAX_SCENES), n_spatial <= 4 (R_max <= 64) and R_max (1 + 2
 * n_bottoms) + 3 n_scenes <= 384 parameters.  A pixel's simplex and working arrays live in
 * LDS for problems of at most 64 parameters whose blocks fit three per SIMD; other problems
 * keep the simplex in device memory allocated for the call (one per wavefront of the launch)
 * and the working arrays in LDS where they fit; the results are the same. */
int photic_invert_regions(const photic_cfg *cfg, int n_rows, int n_cols, int n_scenes,
                          const float *rrs, const float *nodata, const double *sec_sun,
                          const double *sec_view, int n_spatial, long n, const int64_t *idx,
                          const float *hstart, const photic_outputs *out, int *n_regions,
                          void *stream);

/* photic_invert_regions with a band set per scene, as the reference reads each scene's own
 * band count and wavelengths and derives its coefficients from them (model/samodel.c:384-527):
 * scene s has n_bands[s] (2..8) bands at wavelengths[...] (host; the scenes' lists one after
 * another, ascending within a scene); rrs [sum n_bands][n_rows * n_cols] (device) and nodata
 * [sum n_bands] (host) hold the scenes' band grids one after another.  The misfit runs over
 * every (region, scene, band) in that order, each scene's K penalty over its own bands, IOD
 * over every (scene, region, band).  cfg's own band list is not used; everything else is as
 * photic_invert_regions. */
int photic_invert_regions_bands(const photic_cfg *cfg, int n_rows, int n_cols, int n_scenes,
                                const int *n_bands, const double *wavelengths, const float *rrs,
                                const float *nodata, const double *sec_sun,
                                const double *sec_view, int n_spatial, long n,
                                const int64_t *idx, const float *hstart,
                                const photic_outputs *out, int *n_regions, void *stream);

/* Synthetic stack scene (Philox as photic_gen_scene, counter word 3 = date): H ~ U(0.5, 25),
 * B ~ U(0, 1) per pixel; per date sun zenith ~ U(20, 50) deg and log-normal P, G, X; a
 * Bernoulli(p_cloud) cloud mask with at least k_min clear dates; a date whose noisy Rrs goes
 * negative is invalid.  rrs [n_dates][4][n], sec_sun [n_dates][n], valid [n_dates][n], truth
 * [2 + 4 n_dates][n] (H, B, then P, G, X, theta per date) or NULL. */
int photic_gen_stack(const photic_cfg *cfg, uint64_t seed, long pix0, long n, int n_dates,
                     double sigma, double p_cloud, int k_min, float *rrs, float *sec_sun,
                     unsigned char *valid, float *truth, void *stream);

/* Synthetic single-date scene (Philox4x32-10 keyed by seed, counter = pixel index): draws
 * H ~ U(0.5, 25) (deep_frac of pixels U(30, 60)), log-normal P, G, X, B ~ U(0, 1), sun zenith
 * ~ U(20, 50) deg, FP64 forward model with self-consistent Y, N(0, sigma) noise.  rrs:
 * [n_bands][n]; sec_sun: [n]; truth: [6][n] (H, P, G, X, B, theta) or NULL.  Pixel indices
 * pix0 .. pix0 + n - 1. */
int photic_gen_scene(const photic_cfg *cfg, uint64_t seed, long pix0, long n, double sigma,
                     double deep_frac, float *rrs, float *sec_sun, float *truth, void *stream);

/* Spatially coherent synthetic scene of n_rows x n_cols pixels (pixel index row * n_cols + col)
 * for the region workload: each of H, B, ln P, ln G, ln X is a smooth field over the grid (0.5
 * plus two plane waves of amplitude 0.25, wavelengths 64-256 pixels, random directions and
 * phases from seed) mapped into its range (H 0.5-25 m, B 0-1, P, G, X between the single-date
 * scene's bounds), one sun zenith theta_sun_deg for the scene, view at nadir, the FP64 forward
 * model with self-consistent Y, and N(0, sigma) noise per pixel and band.  rrs [n_bands][n];
 * truth [6][n] (H, P, G, X, B, theta) or NULL.  This build's test data; no reference
 * counterpart. */
int photic_gen_field(const photic_cfg *cfg, uint64_t seed, long n_rows, long n_cols,
                     double theta_sun_deg, double sigma, float *rrs, float *truth, void *stream);

/* No-data blobs for a synthetic scene of n_rows x n_cols pixels (pixel index row * n_cols +
 * col): n_discs discs keyed by seed (Philox, counter = disc index), disc i centred at
 * (u0 n_cols, u1 n_rows) with radius r_min + (r_max - r_min) u2; every band of a pixel whose
 * centre lies inside a disc is set to nodata in rrs [n_bands][n] (device).  The masked
 * fraction is about 1 - exp(-n_discs pi E[r^2] / (n_rows n_cols)). */
int photic_gen_nodata_discs(uint64_t seed, long n_rows, long n_cols, int n_discs, double r_min,
                            double r_max, int n_bands, float nodata, float *rrs, void *stream);

/* Test entry points: the forward model alone (params [n][5] = H, P, G, X, B; sec_sun [n];
 * Y [n]; out [n][n_bands]) and the misfit alone (x [n][5] MIX layout; rrs [n][n_bands];
 * sec_sun [n]; out, rrs_error [n]). */
int photic_forward(const photic_cfg *cfg, long n, const double *params, const double *sec_sun,
                   const double *Y, double *out, int precision, void *stream);
int photic_objective(const photic_cfg *cfg, long n, const double *x, const float *rrs,
                     const float *sec_sun, double *out, double *rrs_error, int precision,
                     void *stream);

/* ---- the learned warm start --------------------------------------------------------
 * A 5-128-128-128-5 GELU network (BF16 operands, FP32 accumulation on the matrix cores)
 * mapping ln Rrs at the four bands and the sun zenith to a starting point (H, B, P, G, X)
 * for the simplex.  Inputs x = ((ln max(Rrs, 1e-4) + 4) / 1.5 per band, (theta_deg - 35) / 15);
 * targets t = ((ln H - 1.25), (ln P + 3.45) / 1.85, (ln G + 3.91) / 2.30, (ln X + 4.95) / 1.96,
 * 2 B - 1).  Parameters are FP32 masters in the layout W1[128][5] b1 W2[128][128] b2
 * W3[128][128] b3 W4[5][128] b4 (34,437 values) with a packed BF16 copy for the kernels. */
typedef struct photic_mlp photic_mlp;

int photic_mlp_create(photic_mlp **mlp, uint64_t init_seed);
void photic_mlp_destroy(photic_mlp *mlp);
int photic_mlp_n_params(void);
int photic_mlp_set_weights(photic_mlp *mlp, const float *host_w);
int photic_mlp_get_weights(photic_mlp *mlp, float *host_w);
/* Training pairs (device x, t: [n][5]) drawn from the scene generator's distributions
 * (single date, configured bands), pixels pix0 .. pix0 + n - 1 of seed. */
int photic_mlp_trainset(const photic_cfg *cfg, uint64_t seed, long pix0, long n, double sigma,
                        float *x, float *t, void *stream);
/* Mean-squared-error loss and its gradient (host_grad: photic_mlp_n_params() floats) over
 * n samples (a multiple of 16). */
int photic_mlp_gradient(photic_mlp *mlp, const float *x, const float *t, long n,
                        float *host_grad, float *loss);
/* Adam (lr decaying geometrically to lr_final) over `epochs` passes of n samples in
 * mini-batches of `batch` (a multiple of 16); *final_loss = the last epoch's mean loss. */
int photic_mlp_train(photic_mlp *mlp, const float *x, const float *t, long n, long batch,
                     int epochs, float lr, float lr_final, float *final_loss, void *stream);
/* Raw network outputs y [n][5] for inputs x [n][5]. */
int photic_mlp_forward(photic_mlp *mlp, const float *x, long n, float *y, void *stream);
/* Simplex seeds [5][n] in the MIX layout (H, B, 100 P, 100 G, 100 X) for a single-date scene
 * (rrs [4][n], sec_sun [n] or NULL for cfg->sec_sun).  A pixel whose predicted depth is below
 * the network's seed-skip depth (photic_mlp_set_seed_skip) gets a NaN depth: no seed start. */
int photic_mlp_seed(photic_mlp *mlp, const photic_cfg *cfg, long n, const float *rrs,
                    const float *sec_sun, float *seed, void *stream);
/* The depth below which photic_mlp_seed gives no seed (0, the default: every pixel is
 * seeded).  Shallow clear water cannot meet the early exit under the reference's K penalty
 * (model/samodel.c:2696-2732), so its pixels run every depth start whatever their seed; a
 * seed start there is work the inversion throws away.  The level is this build's, learned
 * from labelled forward-model samples (bench.py fits it on a scene other than the one it
 * times). */
int photic_mlp_set_seed_skip(photic_mlp *mlp, double h_max);
/* Stack seeds [2 + 3 D][n] (H, B, then 100 P, 100 G, 100 X at each date's slot) for a D-date
 * scene (rrs [D][4][n], sec_sun [D][n], valid [D][n]): the network runs on every date; H and B
 * are the means of its per-date values over the pixel's valid dates, P, G, X the date's own.
 * The seed input of photic_invert_stack.  No reference counterpart (the learned warm start is
 * this build's addition). */
int photic_mlp_seed_stack(photic_mlp *mlp, const photic_cfg *cfg, long n, int n_dates,
                          const float *rrs, const float *sec_sun, const unsigned char *valid,
                          float *seed, void *stream);

/* Streamed stack inversion of a host-resident scene of n_rows x n_cols pixels (pixel index
 * row * n_cols + col): rrs [n_dates][4][n], sec_sun [n_dates][n], valid [n_dates][n] in host
 * memory (pageable or pinned).  The scene goes through HBM in tiles of tile_rows rows: a
 * tile's inputs are staged into one of two pinned buffers and copied to the device on a copy
 * stream while the previous tile is inverted on a compute stream (MLP stack seeds from mlp
 * when it is not NULL, then photic_invert_stack), and its outputs come back the same way
 * into out's HOST arrays (photic_outputs layout over the whole scene, each may be NULL).
 * progress, when not NULL, is called from the calling thread after each tile's outputs are
 * in place.  Returns when the whole scene is done.  With cfg->check_watchdog each tile's
 * inversion is waited for, so staging overlaps only the copies.  The tiling is this build's
 * (the reference inverts a scene held in memory, model/samodel.c:289-1763); each pixel's
 * result is the one photic_invert_stack gives for it. */
typedef void (*photic_progress_fn)(long tiles_done, long n_tiles, void *user);
int photic_invert_stack_stream(const photic_cfg *cfg, photic_mlp *mlp, long n_rows, long n_cols,
                               int n_dates, const float *rrs, const float *sec_sun,
                               const unsigned char *valid, long tile_rows,
                               const photic_outputs *out, photic_progress_fn progress,
                               void *user);

/* ---- the ordered inversion: ordering, look-up table and hot start -------------------
 * The reference inverts a scene's pixels in the order of a spectral index -- the norm of the
 * summed Rrs at bands 1, 2, 3, times the spectral angle to the noise spectrum when every
 * R_sigma is at least 1e-6 (model/samodel.c:701-770) -- copies the solution of a recent pixel
 * from a look-up table when its modelled Rrs matches within 1.5 % (model_spectral_diff_mean,
 * :980-1004, :2795-2842), and otherwise starts its one Nelder-Mead run from the previous
 * pixel's depth, inverting again with the full depth list when the misfit stays above 5
 * (start_at_previous and RESTART, :1013-1031, :2233-2241).  photic_invert_ordered recasts that
 * for the GPU: the modelled pixels are sorted by the same index (a stable sort; the reference's
 * qsort leaves ties in no set order) and cut into runs of `run`; each run's first pixel (its
 * leader) is inverted with the full depth list; each other pixel (a follower) compares its
 * Rrs with the modelled Rrs of its own leader and the `window` leaders before it that are LUT
 * entries (objective below entry_max, depth below 32 m, :1034) and, at a difference of at most
 * match_max, takes the best one's parameter vector as its solution (its outputs are that
 * vector's on its own data: its own misfit, where the reference copies the entry's); every
 * other follower is inverted with one start from its own leader's depth, and, with a restart
 * level set, again with the full depth list if its objective stays above it.  Leaders, hot starts and
 * restarts are each one batch.  Single-date pixels (rrs [4][n], sec_sun [n] or NULL, device
 * arrays), FP64, either parameter layout, unseeded.  role [n] (device) or NULL: -1 not
 * modelled, 0 leader, 1 LUT match, 2 hot start, 3 hot start inverted again. */
typedef struct photic_order_cfg {
    int run;               /* pixels per run: a leader and run - 1 followers (64) */
    int window;            /* earlier leaders a follower may match, besides its own (16) */
    int use_lut;           /* 0: no look-up table, every follower hot-starts (1) */
    double match_max;      /* largest spectral difference of a match, % (1.5) */
    double entry_max;      /* a leader is a LUT entry below this objective (1.5) */
    double restart_misfit; /* > 0: a hot start above this objective is inverted again with
                              the full depth list (the reference's RESTART, 5, :895); 0 (the
                              default): no restart, as the reference compiles it (:98) */
    double r_sigma[4];     /* noise per band; all >= 1e-6: the spectral angle enters (0) */
} photic_order_cfg;

typedef struct photic_order_stats {
    long n_valid, n_leaders, n_lut, n_hot, n_restarted;
} photic_order_stats;

/* The reference's values (model/samodel.c:895-897; no restart), 64-pixel runs, a 16-leader
 * window. */
void photic_order_cfg_default(photic_order_cfg *ocfg);
int photic_invert_ordered(const photic_cfg *cfg, const photic_order_cfg *ocfg, long n_pix,
                          const float *rrs, const float *sec_sun, const photic_outputs *out,
                          int *role, photic_order_stats *stats, void *stream);

/* The band coefficients the kernels use for cfg's wavelengths, interpolated from the
 * reference's 41-node tables (model/samodel.c:103-286) as model/samodel.c:560-618 does:
 * out [12][n_bands] = a0, a1, a_w, b_bw, then the 8 bottom reflectances in the order of
 * model/samodel.c:466-478; *aw640 = a_w(640 nm) (may be NULL).  Host only, no GPU needed. */
int photic_band_coefficients(const photic_cfg *cfg, double *out, double *aw640);

/* ---- the launch-shape autotuner ----------------------------------------------------
 * Times the inversion on a device-resident sample (n_pix >= 64 pixels or stacks) for every
 * candidate shape -- lanes per block 64 / 128 / 192 / 256 (as the simplexes' LDS allows),
 * the occupancy's blocks per CU or half of them, queue chunks of 16 .. 256 pixels; for
 * stacks blocks of 64 / 128 / 256 at full occupancy and chunks of 16 and 64 -- and
 * keeps the fastest for the kernel instance cfg selects (precision, layout, seeded or not).
 * Later calls whose cfg->block, grid and chunk are 0 use it.  Per-pixel results do not
 * depend on the shape.  Process-wide; not to be run concurrently with inversions. */
typedef struct photic_tuning {
    int block, blocks_per_cu, chunk;
    float ms;     /* the winner's time on the sample */
    int n_tried;  /* shapes timed */
} photic_tuning;

int photic_autotune(const photic_cfg *cfg, long n_pix, const float *rrs, const float *sec_sun,
                    const float *seed, photic_tuning *best, void *stream);
int photic_autotune_stack(const photic_cfg *cfg, long n_pix, int n_dates, const float *rrs,
                          const float *sec_sun, const unsigned char *valid, const float *seed,
                          photic_tuning *best, void *stream);

/* Launch-shape facts of the compiled inversion kernel for cfg. */
int photic_kernel_info(const photic_cfg *cfg, int *vgprs, int *scratch_bytes,
                       int *blocks_per_cu, int *n_cu);

/* Device plumbing for callers without a HIP allocator of their own. */
int photic_device_count(void);
int photic_device_malloc(void **ptr, size_t bytes);
int photic_device_free(void *ptr);
/* Page-locked host memory: device-to-host copies into it run at the link's rate instead
 * of going through the runtime's staging buffer. */
int photic_host_malloc(void **ptr, size_t bytes);
int photic_host_free(void *ptr);
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
