/*
 * photic_samodel.h -- drop-in for the reference's per-pixel inversion entry point.
 *
 * void samodel(...) below has the signature of model/samodel.h:8-19 (model/samodel.c:
 * 289-300) and takes the reference's own structures: photic_geogrid and photic_scene have the
 * layouts of `geogrid` (model/common.h:69-84) and `scene` (model/common.h:194-218), and the
 * reference's `bool` is an int (model/common.h:58), so a caller compiled against the
 * reference's headers links against libphotic.so unchanged.
 *
 * Semantics: the reference's inversion of every modelled pixel (nodata test of
 * model/samodel.c:933-947, Rrs gather of :2957-3027 with the box smoothing of
 * model/common.c:232-276, the native parameter vector with n_bottoms bottoms -- or sand only
 * where an empirical depth deeper than 8 m is given, :1824-1858 -- the multi-start of
 * :2119-2427 with the reference's settings, the outputs of :1992-2079, the grids of
 * :1120-1128 with depth negated as at :1486-1490), each pixel inverted independently on the
 * GPU -- with n_spatial > 1 jointly with its (2 n_spatial - 1)^2 neighbours and their
 * continuity penalties (:2432-2759), with several scenes as one problem (:1808-1815).  When
 * depth_sigma is not NULL the depth-error pass of :1376-1482 fills it, with Philox draws in
 * place of the reference's clock-seeded rand(); as in the reference, the first sample inverted
 * runs the main pass's start rules and every later one a single start from the depth of the
 * main pass's last modelled pixel (start_at_previous, :1454, :2233-2241).  Not carried over: the look-up table and hot start
 * (order-dependent and sequential in the reference; photic_invert_ordered is their batched
 * form) and the netCDF files the reference writes as a side effect (photic_pack_int16 is
 * their packing).  Supported: 1-32 scenes of one grid with the same 4 bands, 1-3 bottoms,
 * n_spatial <= 4 and at most 384 parameters per pixel ((2 n_spatial - 1)^2 (1 + 2 n_bottoms)
 * + 3 nscenes); anything else prints a message and leaves the outputs at the reference's
 * defaults.  Use photic_samodel() for the status code.
 */
#ifndef PHOTIC_SAMODEL_H
#define PHOTIC_SAMODEL_H

#ifdef __cplusplus
extern "C" {
#endif

#define PHOTIC_MAX_GRIDS 265
#define PHOTIC_MAX_STRING_LEN 2048

struct photic_geogrid { /* = geogrid, model/common.h:69-84 */
    int nrows;
    int ncols;
    float cellsize;
    float wlon;
    float slat;
    float elon;
    float nlat;
    float nodata_value;
    float lambda;
    float theta_v;
    float theta_w;
    float **array;
};

struct photic_scene { /* = scene, model/common.h:194-218 */
    char scene_name[PHOTIC_MAX_STRING_LEN];
    int n_bands;
    int nrows;
    int ncols;
    int band_indexes[PHOTIC_MAX_GRIDS];
    int wavelengths[PHOTIC_MAX_GRIDS];
    double theta_w; /* subsurface solar zenith, degrees */
    double theta_v; /* subsurface view angle from nadir, degrees */
    double H_tide;
    double R_inf[PHOTIC_MAX_GRIDS];
    double R_sigma[PHOTIC_MAX_GRIDS];
    double K[PHOTIC_MAX_GRIDS];
    double K_sigma[PHOTIC_MAX_GRIDS];
    double ratio_min;
    double ratio_max;
    double slope_min;
    double slope_max;
    int pgx_present;
    int ps_p;
    int ps_g;
    int ps_x;
};

void samodel(struct photic_scene scene_data[], struct photic_geogrid gridded_data[],
             int *scene_indexes, int nscenes, int empirical_depth_present,
             struct photic_geogrid empirical_depths, int n_smoothing_radius, int n_spatial,
             int n_bottoms, float **depth, float **depth_sigma, float **model_error,
             float **bottom_albedo, float **bottom_sand, float **bottom_seagrass,
             float **bottom_coral, float **K_min, float **bottom_type,
             float **index_optical_depth, float pagesize, int background, int linewidth);

/* The same, returning a photic_status code (samodel() itself is void like the reference). */
int photic_samodel(struct photic_scene scene_data[], struct photic_geogrid gridded_data[],
                   int *scene_indexes, int nscenes, int empirical_depth_present,
                   struct photic_geogrid empirical_depths, int n_smoothing_radius,
                   int n_spatial, int n_bottoms, float **depth, float **depth_sigma,
                   float **model_error, float **bottom_albedo, float **bottom_sand,
                   float **bottom_seagrass, float **bottom_coral, float **K_min,
                   float **bottom_type, float **index_optical_depth);

#ifdef __cplusplus
}
#endif
#endif
