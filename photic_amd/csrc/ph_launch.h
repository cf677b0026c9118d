/* photic_amd -- the internal boundary between the C host code and the HIP kernels.
 *
 * Plain C structs: the host fills them, the launchers pass them to the kernels by value, so
 * every uniform constant of a call lives in the kernel-argument segment and is read through
 * the scalar cache.  Internal to libphotic; the public API is include/photic.h. */
#ifndef PH_LAUNCH_H
#define PH_LAUNCH_H

#include <stddef.h>
#include <stdint.h>

#define PH_MAXB 8  /* bands per date */
#define PH_NBOT 8  /* bottom types in the table */
#define PH_MAXH 16 /* depth starts */
#define PH_MAXD 8  /* dates per pixel (stacks) */
#define PH_MAXS 32 /* scenes of a region problem (the reference's MAX_SCENES) */
#define PH_REG_MAXR 64  /* regions of a pixel: (2 n_spatial - 1)^2, n_spatial <= 4 */
#define PH_REG_MAXN 384 /* parameters of a region problem: 64 per lane slot, 6 slots */
#define PH_REG_SLOT 10  /* doubles per band slot in PhRegionArgs::scn */
#define PH_REG_SCENE 22 /* doubles per scene in PhRegionArgs::scn */

#define PH_LAUNCH_UNSUPPORTED 100001

/* Uniform constants of one inversion call. */
typedef struct PhConst {
    int n_bands;
    int n_dates;
    int mode;                /* 0 = native bottoms, 1 = sand/seagrass mix */
    int n_bottoms;           /* native mode */
    int bottom_idx[PH_NBOT]; /* native mode */
    int konvge, kcount, n_hstarts, has_seed;
    int prescreened;         /* 1: no no-data / negativity screen of the input values */
    int seed_konvge;         /* iterations between convergence checks in a seed start */
    double reqmin, rq, early_exit, nodata;
    double sec_sun0, sec_view0; /* geometry where the per-pixel arrays are NULL */
    double hstart[PH_MAXH];
    double wl[PH_MAXB], a0[PH_MAXB], a1[PH_MAXB], aw[PH_MAXB], bbw[PH_MAXB];
    double eg[PH_MAXB]; /* exp(-0.015 (lambda - 440)), evaluated on the host with libm */
    double bot[PH_NBOT][PH_MAXB];
    double aw640;
    /* Rrs at 440/490/550/640/750 nm from the band values: the reference's interp_1d segment
     * and weights depend only on the wavelengths, so the host resolves them once:
     * value = R[i0] * w0 + R[i1] * w1, with w0 = 1 - alpha, w1 = alpha (model/common.c:298). */
    int ip_i0[5], ip_i1[5];
    double ip_w0[5], ip_w1[5];
    double seed_rel_step[8]; /* seeded start: simplex step as a fraction of the seed value, */
    double seed_min_step[8]; /* and its floor, per class H, bottom, P, G, X              */
    double seed_abort;       /* give a seed start up above this misfit at its first check */
} PhConst;

/* One single-date inversion launch.  Arrays are device pointers, planar per band / per
 * parameter ([k][n_pix]); any output may be NULL. */
typedef struct PhInvertArgs {
    PhConst c;
    long n_pix;
    const float *rrs;      /* [n_bands][n_pix] */
    const float *sec_sun;  /* [n_pix] or NULL (nadir sun) */
    const float *sec_view; /* [n_pix] or NULL (nadir view) */
    const float *seed;     /* [n_params][n_pix] when c.has_seed */
    const float *hstart;   /* [n_pix]: one depth start per pixel (c.n_hstarts == 1), or NULL */
    unsigned long long *queue; /* [0] work counter, [1] watchdog flag; zero at launch */
    int chunk;
    int refill_min; /* idle lanes a wave gathers before it sets up new pixels (<= 1: any) */
    unsigned long long max_ticks; /* watchdog: wall-clock ticks a wave may go without progress */
    float *depth, *rrs_error, *model_error, *albedo, *k_min, *iod;
    float *b_pct; /* [3][n_pix]: sand, seagrass, coral percent */
    float *bottom_type;
    double *x_out; /* [n_params][n_pix], |best| */
    int *status, *n_evals, *n_iter, *n_starts;
    double *rrs_model; /* [n_bands][n_pix]: modelled Rrs at the optimum */
} PhInvertArgs;

/* One launch of the stack inversion for the pixels with exactly k valid dates (k = the
 * instance's date count, c.n_dates == k).  The full scene arrays are read through the
 * bucket's pixel list; outputs are written at the pixels' own indices. */
typedef struct PhStackArgs {
    PhConst c;
    long n_pix;                 /* pixels in the scene arrays */
    int n_dates;                /* dates in the scene arrays (D) */
    const float *rrs;           /* [D][n_bands][n_pix] */
    const float *sec_sun;       /* [D][n_pix] */
    const float *sec_view;      /* [D][n_pix], or NULL: c.sec_view0 */
    const float *hstart;        /* [n_pix]: one depth start per stack, or NULL: c.hstart */
    int x_off;                  /* first P row of x_out: 2 (mix) or 1 + 2 n_bottoms */
    const unsigned char *valid; /* [D][n_pix] */
    const float *seed;          /* [2 + 3 D][n_pix]: H, B, (P, G, X) per date slot, or NULL */
    const int *idx;             /* the bucket's pixel indices */
    const unsigned int *count;  /* the bucket's size (device) */
    unsigned long long *queue;  /* [0] work counter, [1] watchdog flag; zero at launch */
    int chunk;
    unsigned long long max_ticks;
    float *depth, *rrs_error, *model_error, *albedo, *k_min, *iod;
    float *b_pct, *bottom_type;
    double *x_out; /* [2 + 3 D][n_pix]: H, B, then each valid date's P, G, X at its slot */
    int *status, *n_evals, *n_iter, *n_starts;
} PhStackArgs;

/* One launch of the spatial-region inversion (ph_region.hip): the pixels idx[q] (or q) of an
 * n_rows x n_cols scene, each inverted with its (2 n_spatial - 1)^2 neighbourhood, one
 * wavefront per pixel.  A wave's block of lds_doubles doubles in LDS holds the simplex
 * [(n_max + 1) n_max] from offset 0, then the arrays at the o_* offsets (in doubles).  For
 * problems whose block does not fit the LDS (the device-memory instances), the simplex is
 * at work + wave * work_doubles in device memory and the arrays in LDS (lds_doubles, the
 * o_* offsets counted from the LDS base) when they fit there, else the whole block is in
 * device memory (tables_in_lds 0, work_doubles the block).  The arrays:
 * o_x the trial point [n_max], o_d squared residual per pair,
 * o_q rB / r per pair, o_rm measured Rrs per pair, o_pw (440 / lambda)^Y per pair
 * [np_max each; pair j = (region j / n_slots, band slot j % n_slots)], o_k the last region's
 * K per band slot [n_slots], o_m Rrs at 490 / 550 / 640 nm per (region, scene) [3 r_max
 * n_scenes], o_pix the regions' pixel indices [r_max], o_bc the band slots' coefficients
 * [n_slots][PH_REG_SLOT] (as in scn), o_geo each scene's sec(sun) then sec(view) [2 n_scenes],
 * o_sc each scene's first band slot and band count [2 n_scenes], o_hd each region's
 * depth-penalty term [r_max], o_bqv and o_bd each (bottom, region)'s fraction and penalty term
 * [n_bottoms r_max each]. */
typedef struct PhRegionArgs {
    PhConst c;                     /* c.n_dates = n_scenes, c.n_bottoms native bottoms */
    int n_rows, n_cols, n_spatial, n_scenes;
    int n_slots;                   /* band slots: the scenes' band counts summed */
    int r_max, n_max, np_max;      /* regions, parameters, (region, scene, band) pairs */
    int lds_doubles;
    int work_doubles;              /* device-memory instances: doubles per wave in work */
    int tables_in_lds;             /* device-memory instances: the arrays are in LDS */
    int o_x, o_d, o_q, o_rm, o_pw, o_k, o_m, o_pix, o_bc, o_geo, o_sc, o_hd, o_bqv, o_bd;
    const double *scn;             /* device tables: sec(sun) [n_scenes], sec(view)
                                      [n_scenes]; per band slot (scene s's band b, the
                                      scenes' bands one after another) R_sigma [n_slots] and
                                      the grid's no-data value [n_slots]; per slot a0, a1,
                                      a_w, b_bw, exp(-0.015 (lambda - 440)), the bottoms'
                                      reflectances (3), 440 / lambda, 16 s + b
                                      [n_slots][PH_REG_SLOT]; per scene its first slot, its
                                      band count and the interpolation (i0, i1, w0, w1) at
                                      440, 490, 550, 640, 750 nm [n_scenes][PH_REG_SCENE] */
    double *work;                  /* device blocks [grid][work_doubles] when they live there */
    long n;                        /* pixels to invert */
    const float *rrs;              /* [n_scenes][4][n_rows * n_cols] */
    const int64_t *idx;            /* [n] pixel indices, or NULL: pixel q is q, screened */
    const float *hstart;           /* [n]: one depth start per pixel, or NULL: c.hstart */
    const float *pert;             /* [n]: n_sigma per pixel (R_sigma perturbation), or NULL */
    unsigned long long *queue;     /* [0] work counter, [1] watchdog flag; zero at launch */
    unsigned long long max_ticks;
    float *depth, *rrs_error, *model_error, *albedo, *k_min, *iod, *b_pct, *bottom_type;
    double *x_out;                 /* [n_max][n] */
    double *depth64;               /* [n]: the depth before its rounding to float, or NULL */
    int *status, *n_evals, *n_iter, *n_starts, *n_regions;
} PhRegionArgs;

struct photic_cfg;
struct photic_outputs;

#ifdef __cplusplus
extern "C" {
#endif

int ph_launch_invert(const PhInvertArgs *a, int fp32, int n_bands, int n_bottoms, int grid,
                     int block, void *stream);
/* The no-allocation entry point's scratch: a file-scope device array of PH_SCRATCH_WORDS
 * 64-bit words (ph_kernels.hip), one copy per device.  *out: its address on the current
 * device, looked up once per device with that device current. */
#define PH_SCRATCH_WORDS 32
int ph_scratch_words(unsigned long long **out);
int ph_launch_zero_u64(unsigned long long *p, int n, void *stream);
int ph_launch_copy16(const void *src, void *dst, long n16, void *stream);
int ph_launch_fwd(const PhConst *c, long n, const double *params, const double *sec_sun,
                  const double *Y, double *out, int fp32, void *stream);
int ph_launch_objective(const PhConst *c, long n, const double *x, const float *rrs,
                        const float *sec_sun, double *out, double *rrs_error, int fp32,
                        int n_bands, void *stream);
int ph_launch_nodata_discs(uint64_t seed, long n_rows, long n_cols, int n_discs, double r_min,
                           double r_max, int n_bands, float nodata, float *rrs, void *stream);
int ph_launch_gen(const PhConst *c, uint64_t seed, long pix0, long n, double sigma,
                  double deep_frac, float *rrs, float *sec_sun, float *truth, void *stream);
int ph_launch_gen_field(const PhConst *c, uint64_t seed, long n_rows, long n_cols, double theta,
                        double sigma, float *rrs, float *truth, void *stream);
int ph_kernel_attrs(int fp32, int n_bands, int n_bottoms, int *vgprs, int *sgprs,
                    int *scratch_bytes, int *max_threads);
int ph_kernel_occupancy(int fp32, int n_bands, int n_bottoms, int seeded, int block,
                        int *blocks_per_cu);
int ph_lds_bytes(int n_bottoms, int block);

/* multi-date stacks (ph_stack.hip) */
int ph_launch_stack_bucket(const PhStackArgs *a, unsigned int *counts, int *lists,
                           void *stream);
int ph_launch_stack(const PhStackArgs *a, int k, int nbot, int seeded, int grid, int block,
                    void *stream);
int ph_stack_occupancy(int k, int nbot, int seeded, int block, int *blocks_per_cu);
int ph_launch_gen_stack(const PhConst *c, uint64_t seed, long pix0, long n, int n_dates,
                        double sigma, double p_cloud, int k_min, float *rrs, float *sec_sun,
                        unsigned char *valid, float *truth, void *stream);

/* spatial regions (ph_region.hip) */
int ph_launch_region(const PhRegionArgs *a, int grid, void *stream);
/* Blocks per CU of the instance a launch would use: device_blocks selects the form whose blocks
 * live in device memory (ph_launch_region takes it when a->work is set). */
int ph_region_occupancy(const PhRegionArgs *a, int device_blocks, int *blocks_per_cu);

/* the ordered inversion (ph_order.hip) */
int ph_launch_order_key(const PhInvertArgs *a, int noise, const double *nsig, float *key, int *pix,
                        unsigned int *n_valid, int *role_px, void *stream);
int ph_order_sort(const float *key_in, float *key_out, const int *val_in, int *val_out, long n,
                  void *tmp, size_t *tmp_bytes, void *stream);
int ph_launch_order_decide(long n_valid, int run, int window, int use_lut, double match_max,
                           double entry_max, const float *rrs, long n_pix, const int *order,
                           const float *l_depth, const float *l_objective, const int *l_status,
                           const double *l_rrs, long n_lead, int *role, int *src, float *hstart,
                           void *stream);
int ph_launch_order_list(long n, const int *role, int want, const int *flag, const int *map,
                         int *list, unsigned int *count, void *stream);
int ph_launch_order_roles(long n, const int *list, const int *role_s, const int *order,
                          int *role_px, void *stream);
int ph_launch_order_gather(long m, const int *list, long stride, const int *order,
                           const float *rrs, const float *sec_sun, long n_pix, const float *h_in,
                           float *rrs_out, float *ss_out, float *h_out, void *stream);
int ph_launch_order_adopt(const PhInvertArgs *a, int n_bottoms, long m, const int *list,
                          const int *order, const int *src, const double *l_x,
                          const int *l_status, long n_lead, void *stream);
int ph_launch_order_scatter(const struct photic_outputs *dst, const struct photic_outputs *from,
                            long m, const int *list, long stride, const int *order, long n_pix,
                            int n_params, int accumulate, void *stream);
int ph_launch_order_restart(long m, const float *objective, double level, int *flag,
                            void *stream);

/* the warm-start network (ph_mlp.hip) */
int ph_mlp_nparam(void);
int ph_mlp_packed_sizes(int *n_bf16, int *n_f32);
long ph_mlp_gbuf_elems(long batch);
int ph_launch_mlp_init(uint64_t seed, float *w, void *stream);
int ph_launch_mlp_adam(float *w, float *m, float *v, const float *g, float lr, float b1,
                       float b2, float eps, int step, unsigned short *pk, float *pb,
                       void *stream);
int ph_launch_mlp_fwdbwd(const unsigned short *pk, const float *pb, const float *x,
                         const float *t, float *y, long n, float gscale, unsigned short *gbuf,
                         float *loss, int grid, void *stream);
#define PH_MLP_KSPLIT 16 /* batch slices of the weight-gradient kernels */
int ph_launch_mlp_wgrad(const unsigned short *gbuf, long n, float *grad, float *part,
                        void *stream);
int ph_launch_mlp_trainset(const PhConst *c, uint64_t seed, long pix0, long n, double sigma,
                           float *x, float *t, void *stream);
int ph_launch_mlp_prep(const float *rrs, const float *ss, float ss0, long n, float *x,
                       void *stream);
int ph_launch_mlp_seed(const float *y, long n, float skip_h, float *seed, void *stream);
int ph_launch_mlp_stack_seed(const float *y, long n, int d, const unsigned char *valid,
                             float *seed, void *stream);
int ph_launch_mlp_stack_mean(long n, int n_dates, const unsigned char *valid, float *seed,
                             void *stream);

/* photic_invert_regions with a perturbation: pert [n] (device) n_sigma per listed pixel, rsig
 * [4 n_scenes] (host) R_sigma per (scene, band); every region's values become
 * Rrs + n_sigma R_sigma (model/samodel.c:3003-3007); depth64 [n] (device) or NULL receives each
 * depth in double, as the reference keeps it.  For the drop-in's depth-error pass.  Every
 * scene has cfg's bands. */
int ph_invert_regions_pert(const struct photic_cfg *cfg, int n_rows, int n_cols, int n_scenes,
                           const float *rrs, const float *nodata, const double *sec_sun,
                           const double *sec_view, int n_spatial, long n, const int64_t *idx,
                           const float *hstart, const float *pert, const double *rsig,
                           const struct photic_outputs *out, int *n_regions, double *depth64,
                           void *stream);
/* The same with a band set per scene (model/samodel.c:384-527): scene s has n_bands[s] (2..8)
 * bands at wavelengths[...] (the scenes' lists one after another, ascending within a scene);
 * rrs [sum n_bands][n_rows * n_cols] (device), nodata and rsig [sum n_bands] (host) hold the
 * scenes' band grids one after another.  cfg's own band list is not used. */
int ph_invert_regions_bands(const struct photic_cfg *cfg, int n_rows, int n_cols, int n_scenes,
                            const int *n_bands, const double *wavelengths, const float *rrs,
                            const float *nodata, const double *sec_sun, const double *sec_view,
                            int n_spatial, long n, const int64_t *idx, const float *hstart,
                            const float *pert, const double *rsig,
                            const struct photic_outputs *out, int *n_regions, double *depth64,
                            void *stream);

/* Host side (photic_capi.c): the launch constants for a photic_cfg; exported for the test
 * suite's host build of the lane engine. */
/* Sets photic_last_error()'s message and returns code (for the host modules beside
 * photic_capi.c). */
int ph_error(int code, const char *msg);
int ph_build_const(const struct photic_cfg *cfg, int n_dates, int has_seed, PhConst *out);
int ph_const_size(void);

#ifdef __cplusplus
}
#endif
#endif
