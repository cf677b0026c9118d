/*
 * ORACLE -- test infrastructure, not product code.
 *
 * A plain-C, FP64 restatement of photic's per-pixel semi-analytical inversion, written from
 * the reference's algorithm (model/samodel.c, model/asa047.c, model/common.c) to check the
 * HIP path against.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and only as the checker or the timed CPU baseline; the product (photic_amd/)
 * never links it and has no CPU fallback.
 *
 * Pinning (what makes this oracle trustworthy, each check lives in tests/):
 *   - orc_nelmin             bit-identical to the reference's own nelmin (asa047.c, compiled
 *                            into oracle/_ref) on the same objective: xmin, ynewlo, icount,
 *                            numres, ifault (tests/test_oracle_ref.py).
 *   - orc_interp_1d,         bit-identical to the reference's compiled common.c.
 *     orc_approx_equal
 *   - band tables            equal to the numbers in the reference's text (golden fixture
 *                            tests/golden/reference_tables.json, tests/test_oracle_forward.py).
 *   - orc_rrs_core           closed forms: H = 0 -> r = rho/pi; deep pure water -> r = r_dp
 *                            with u = b_bw/(a_w+b_bw); each constituent alone (a_phi with a0
 *                            and a1, a_g's slope, b_bp's (440/lambda)^Y); affine in rho; the
 *                            attenuation coefficients and the K clamp from dr/dH at H = 0.
 *   - orc_Y_of_chi           clamps and a hand value.
 *   - orc_objective          E_Rrs's definition and the 80/120 weight; every K-penalty tier
 *                            and the K_min > 0.7 term against hand-computed values.
 *   - orc_start_step         hand-computed start points and steps, both parameter layouts.
 *   - orc_optimise           equal to a brute-force restatement from single-start runs (order,
 *                            early exit at 2.5, first minimum kept); with the reference's
 *                            nelmin plugged in it gives the same results as with orc_nelmin.
 *   - generator              Philox4x32-10 known-answer vectors (Salmon et al. 2011).
 *   - MLP                    finite differences on the FP64 forward.
 *   - orc_region_*           with one region, bit-identical to orc_objective, orc_start_step
 *                            and orc_optimise; continuity penalties against hand-computed
 *                            values; the gather on hand-built grids (tests/test_oracle_region.py).
 * samodel.c itself cannot be built in this image (it needs netcdf.h and cpgplot.h), so the
 * forward model and objective are pinned by the closed forms above, not by linking.
 */
#ifndef PHOTIC_ORACLE_H
#define PHOTIC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NREF 41
#define ORC_NBOTTOM_TYPES 8
#define ORC_MAXB 16      /* bands per date */
#define ORC_MAXD 32      /* dates (scenes) per pixel */
#define ORC_MAXBOT 8     /* bottom types in one combination */
#define ORC_MAXN (1 + 2 * ORC_MAXBOT + 3 * ORC_MAXD)

#define ORC_PI 3.141592653589793 /* model/common.h:19 */

extern const double orc_ref_wl[ORC_NREF];
extern const double orc_aw[ORC_NREF];
extern const double orc_bbw[ORC_NREF];
extern const double orc_a0[ORC_NREF];
extern const double orc_a1[ORC_NREF];
extern const double orc_bottom[ORC_NBOTTOM_TYPES][ORC_NREF];

/* Parameter-vector layouts. */
enum {
    ORC_MODE_NATIVE = 0, /* [H, B_0..B_{nb-1}, q_0..q_{nb-1}, (P,G,X)_0 .. (P,G,X)_{nd-1}] */
    ORC_MODE_MIX = 1     /* [H, B, (P,G,X)_0 ..]: sand/seagrass mix, amplitude fixed at 1 */
};

/* Band tables at the band wavelengths (model/samodel.c:505-618). */
typedef struct {
    int n_bands;
    double wl[ORC_MAXB];
    double a0[ORC_MAXB], a1[ORC_MAXB], aw[ORC_MAXB], bbw[ORC_MAXB];
    double bottom[ORC_NBOTTOM_TYPES][ORC_MAXB];
    double aw640;
} orc_bands;

/* Everything one pixel's objective needs, plus the side effects of the last evaluation
 * (the reference keeps the same quantities in model_data, model/common.h:116-188). */
typedef struct {
    int mode;
    int n_bottoms;                  /* native mode: bottoms in the combination */
    int bottom_idx[ORC_MAXBOT];     /* native mode: indexes into the bottom table */
    int n_dates;
    const orc_bands *bands;
    double sec_sun[ORC_MAXD], sec_view[ORC_MAXD];
    double rrs[ORC_MAXD][ORC_MAXB]; /* measured */
    double r440[ORC_MAXD], r490[ORC_MAXD], r550[ORC_MAXD], r640[ORC_MAXD], r750[ORC_MAXD];
    /* side effects of the last objective evaluation */
    double K[ORC_MAXD][ORC_MAXB];
    double Rrs_mod[ORC_MAXD][ORC_MAXB];
    double rrs_mod[ORC_MAXD][ORC_MAXB];
    double rrs_bot[ORC_MAXD][ORC_MAXB];
    double rrs_dp[ORC_MAXD][ORC_MAXB];
    double bottom_albedo, Rrs_error, K_error, model_error;
    long n_evals;
} orc_pixel;

/* Nelder-Mead settings (model/samodel.c:2142-2147 for the reference's own values). */
typedef struct {
    double reqmin;
    int konvge;
    int kcount;
    int n_hstarts;            /* length of hstart[] */
    double hstart[16];        /* depth starts, tried in order */
    double early_exit;        /* accept when the best misfit is below this (2.5 * n_dates) */
} orc_nm_cfg;

/* Per-pixel results, as samodel_optimise leaves them (model/samodel.c:1992-2079). */
typedef struct {
    double x[ORC_MAXN];       /* |best| */
    int n_params;
    double model_error;       /* objective at best */
    double Rrs_error;         /* E_Rrs (%) at best -- the reference's model_error grid */
    double depth;
    double K_min;
    double bottom_albedo;
    double B_percent[ORC_NBOTTOM_TYPES];
    double bottom_type;
    double iod;
    double P[ORC_MAXD], G[ORC_MAXD], X[ORC_MAXD];
    int n_iterations;         /* icount of the accepted start */
    int converged;            /* ifault == 0 of the accepted start */
    int n_starts;             /* NM runs made */
    long n_evals;             /* objective evaluations made (all starts, incl. re-evals) */
} orc_result;

/* A Nelder-Mead routine with the reference's calling convention (model/asa047.h:4-7); the
 * context pointer stands where the reference passes model_data *. */
typedef double (*orc_objective_fn)(double x[], void *ctx);
typedef void (*orc_nelmin_fn)(orc_objective_fn fn, void *ctx, int n, double start[],
                              double xmin[], double *ynewlo, double reqmin, double step[],
                              int konvge, int kcount, int *icount, int *numres, int *ifault);

/* --- helpers (model/common.c) --- */
int orc_approx_equal(float a, float b, float eps);
double orc_interp_1d(const double *X, const double *Y, int n, double x);

/* --- tables / pixel set-up --- */
void orc_band_tables(const double *wl, int n_bands, orc_bands *out);
double orc_sec_deg(double theta_deg);
void orc_pixel_init(orc_pixel *px, int mode, int n_bottoms, const int *bottom_idx,
                    int n_dates, const orc_bands *bands, const double *sec_sun,
                    const double *sec_view, const double *rrs /* [n_dates][n_bands] */);

/* --- forward model and objective --- */
double orc_Y_of_chi(double chi);
double orc_rrs_core(double P, double G, double X, double rho, double H, double Y,
                    const orc_bands *t, int band, double sec_sun, double sec_view,
                    double *side /* K, rrs_dp, rrs_B, rrs; may be NULL */);
double orc_rrs_band(double P, double G, double X, const double *B, const double *q,
                    double H, orc_pixel *px, int date, int band);
double orc_objective(double *x, void *px);
int orc_n_params(const orc_pixel *px);

/* --- minimiser --- */
void orc_nelmin(orc_objective_fn fn, void *ctx, int n, double start[], double xmin[],
                double *ynewlo, double reqmin, double step[], int konvge, int kcount,
                int *icount, int *numres, int *ifault);
void orc_default_cfg(orc_nm_cfg *cfg, int n_dates, double reqmin, int konvge, int kcount);
void orc_start_step(const orc_pixel *px, double hstart, double *start, double *step);
void orc_optimise(orc_pixel *px, const orc_nm_cfg *cfg, orc_nelmin_fn nm, orc_result *res);
/* The objective at x and the outputs samodel_optimise derives from it (a test hook: the
 * outputs of a given vector, with no minimisation). */
void orc_eval_outputs(orc_pixel *px, const double *x, orc_result *res);
void orc_optimise_seeded(orc_pixel *px, const orc_nm_cfg *cfg, orc_nelmin_fn nm,
                         const double *seed_start, const double *seed_step,
                         orc_result *res);

/* --- scene-level driver (the OpenMP pixel loop, model/samodel.c:901-1128) --- */
typedef struct {
    int mode, n_bottoms, bottom_idx[ORC_MAXBOT];
    int n_dates, n_bands;
    double wl[ORC_MAXB];
    orc_nm_cfg nm;
    float nodata;
} orc_scene_cfg;

/* rrs: [n_dates][n_bands][n_pix] float; sec_sun: [n_dates][n_pix] (or NULL -> 1/cos 0);
 * valid: [n_dates][n_pix] bytes (NULL = every date valid).  Outputs [n_pix] each; x_out
 * [n_pix][n_params_max] may be NULL.  threads <= 0 uses the OpenMP default.
 * Returns the number of pixels inverted. */
long orc_invert_scene(const orc_scene_cfg *cfg, long n_pix, const float *rrs,
                      const float *sec_sun, const uint8_t *valid, int threads,
                      float *depth, float *rrs_error, float *model_error, int *status,
                      int *n_evals, double *x_out, int x_stride);

/* --- spatial regions (n_spatial > 1), orc_region.c --- */
#define ORC_MAXR 49    /* regions: (2 n_spatial - 1)^2 for n_spatial <= 4 */
#define ORC_RMAXD 32   /* scenes (the reference's MAX_SCENES) */
#define ORC_RMAXBOT 3  /* bottoms */
#define ORC_RMAXN (ORC_MAXR * (1 + 2 * ORC_RMAXBOT) + 3 * ORC_RMAXD)

/* One pixel's region problem (model_data with N_r regions, model/common.h:116-188). */
typedef struct {
    int n_regions, origin;
    int n_bottoms, bottom_idx[ORC_RMAXBOT];
    int n_dates;
    const orc_bands *bands;            /* scene 0's bands */
    const orc_bands *sb[ORC_RMAXD];    /* each scene's bands (model/samodel.c:384-527) */
    double sec_sun[ORC_RMAXD], sec_view[ORC_RMAXD];
    double rrs[ORC_MAXR][ORC_RMAXD][ORC_MAXB]; /* measured, [region][scene][band] */
    double r440[ORC_MAXR][ORC_RMAXD], r490[ORC_MAXR][ORC_RMAXD], r550[ORC_MAXR][ORC_RMAXD];
    double r640[ORC_MAXR][ORC_RMAXD], r750[ORC_MAXR][ORC_RMAXD];
    /* side effects of the last evaluation */
    double K[ORC_RMAXD][ORC_MAXB];             /* the last region's, as the reference keeps it */
    double rrs_mod[ORC_MAXR][ORC_RMAXD][ORC_MAXB], rrs_bot[ORC_MAXR][ORC_RMAXD][ORC_MAXB];
    double bottom_albedo, Rrs_error, depth_error, bottom_error, K_error, model_error;
    long n_evals;
} orc_region_pixel;

typedef struct {
    double x[ORC_RMAXN]; /* |best|: [H_r], [B_{r,k}], [q_{r,k}], [(P, G, X)_s] */
    int n_params, n_regions;
    double model_error, Rrs_error, depth, K_min, bottom_albedo;
    double B_percent[ORC_NBOTTOM_TYPES];
    double bottom_type, iod;
    int n_iterations, converged, n_starts;
    long n_evals;
} orc_region_result;

float orc_smoothed_index_2d(const float *grid, int i, int j, int n_rows, int n_cols, int r,
                            float spval);
int orc_region_n_params(const orc_region_pixel *px);
int orc_region_gather(orc_region_pixel *px, const float *rrs, const float *nodata, int nd,
                      int nb, int n_rows, int n_cols, int i, int j, int n_spatial,
                      float n_sigma, const double *rsig);
/* The same with nbs[s] bands in scene s: rrs holds the scenes' band grids one after another
 * ([sum nbs][n_rows * n_cols]), nodata and rsig one value per band grid. */
int orc_region_gather_scenes(orc_region_pixel *px, const float *rrs, const float *nodata,
                             int nd, const int *nbs, int n_rows, int n_cols, int i, int j,
                             int n_spatial, float n_sigma, const double *rsig);
void orc_region_init(orc_region_pixel *px, int n_bottoms, const int *bottom_idx, int n_dates,
                     const orc_bands *bands, const double *sec_sun, const double *sec_view);
/* The same with each scene's own band set sb[s]. */
void orc_region_init_scenes(orc_region_pixel *px, int n_bottoms, const int *bottom_idx,
                            int n_dates, const orc_bands *const *sb, const double *sec_sun,
                            const double *sec_view);
double orc_region_objective(double *x, void *px);
void orc_region_start_step(const orc_region_pixel *px, double h0, double *start, double *step);
/* The region objective at x and the region outputs derived from it (a test hook). */
void orc_region_eval_outputs(orc_region_pixel *px, const double *x, orc_region_result *res);
void orc_region_optimise(orc_region_pixel *px, const orc_nm_cfg *cfg, orc_nelmin_fn nm,
                         orc_region_result *res);
/* The ordered inversion (orc_order.c): the spectral index, model_spectral_diff_mean for one
 * region and scene, and the whole run-and-leader inversion. */
float orc_order_key(const float *r, const double *noise);
float orc_spectral_diff_mean(const double *meas, const double *model);
long orc_invert_ordered(int mode, int n_bottoms, const int *bottom_idx, const orc_bands *bands,
                        long n, const float *rrs, const float *sec_sun, double sec_sun0,
                        const orc_nm_cfg *cfg, float nodata, int run, int window, int use_lut,
                        double match_max, double entry_max, double restart,
                        const double *noise, double *x, float *depth, float *rrs_error,
                        float *model_error, int *status, int *n_evals, int *n_starts, int *role);

/* The depth-error pass (model/samodel.c:1376-1482); h_prev: the depth of the main pass's
 * last modelled pixel, the start of every sample after the first when there is no empirical
 * depth grid. */
int orc_depth_sigma(const float *depth, int n_rows, int n_cols, const float *grid,
                    const float *nodata, const double *rsig, int nd, const orc_bands *bands,
                    const double *sec_sun, const double *sec_view, int n_spatial, int n_bottoms,
                    const float *emp, float emp_nodata, double h_prev, const orc_nm_cfg *nm,
                    double *table, float *sigma);

/* --- synthetic scenes (Philox4x32-10, Salmon et al. 2011) --- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

typedef struct {
    double H, P, G, X, B, theta_sun;
} orc_truth;

/* Draws the generating parameters of one pixel of the single-date scene (seeded, keyed by
 * pixel index) and its float32 Rrs at the given bands; returns 1 if the pixel is nodata
 * (a band went negative after noise).  deep_frac > 0 redraws H ~ U(30, 60) for that
 * fraction of pixels. */
int orc_gen_pixel(uint64_t seed, uint64_t pix, const orc_bands *bands, double noise_sigma,
                  double deep_frac, orc_truth *truth, float *rrs_out, double *Y_out);
void orc_fwd_truth(const orc_truth *t, const orc_bands *bands, int n_fixed_point,
                   double *rrs_out, double *Y_out);
/* Pixel (i, j) of the spatially coherent scene of the region workload: H, B, ln P, ln G, ln X
 * smooth fields over the grid (blocks 200-219 of pixel 0), one sun zenith theta, the noise of
 * pixel i * n_cols + j; returns 1 if the pixel is nodata. */
int orc_gen_field_pixel(uint64_t seed, long i, long j, long n_cols, double theta,
                        const orc_bands *bands, double noise_sigma, orc_truth *truth,
                        float *rrs_out);
/* 1 when pixel (row, col) of an n_rows x n_cols scene lies inside one of its n_discs no-data
 * discs (Philox keyed by seed, counter = disc index). */
int orc_in_nodata_disc(uint64_t seed, long row, long col, long n_rows, long n_cols, int n_discs,
                       double r_min, double r_max);
/* One pixel of a multi-date stack scene: shared H and B, per-date P, G, X and sun zenith,
 * a Bernoulli(p_cloud) cloud mask with at least k_min clear dates.  truth [2 + 4 n_dates]:
 * H, B, then (P, G, X, theta_sun) per date; returns the number of valid dates. */
int orc_gen_stack_pixel(uint64_t seed, uint64_t pix, const orc_bands *bands,
                        double noise_sigma, int n_dates, double p_cloud, int k_min,
                        double *truth, float *rrs_out, uint8_t *valid);

/* --- MLP warm start, FP64 reference (5-128-128-128-5, GELU tanh form) --- */
#define ORC_MLP_IN 5
#define ORC_MLP_HID 128
#define ORC_MLP_OUT 5
#define ORC_MLP_NPARAM (ORC_MLP_IN * ORC_MLP_HID + ORC_MLP_HID + 2 * (ORC_MLP_HID * ORC_MLP_HID + ORC_MLP_HID) + ORC_MLP_HID * ORC_MLP_OUT + ORC_MLP_OUT)
/* Weight packing (row-major, out x in): W1[128][5] b1[128] W2[128][128] b2[128]
 * W3[128][128] b3[128] W4[5][128] b4[5]. */
void orc_mlp_forward(const double *w, int n, const double *in /*[n][5]*/, double *out /*[n][5]*/);
/* Mean-squared-error loss over n samples and its gradient w.r.t. every weight. */
double orc_mlp_loss_grad(const double *w, int n, const double *in, const double *target,
                         double *grad /* ORC_MLP_NPARAM, may be NULL */);

#ifdef __cplusplus
}
#endif
#endif
