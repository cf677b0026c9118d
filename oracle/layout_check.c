/*
 * Compile-time check that the drop-in's structures (include/photic_samodel.h) have the
 * reference's memory layout (geogrid and scene, model/common.h:69-84, 194-218): the same size
 * and every field at the same offset, so the reference's caller (bam.c:3236-3241) can pass its
 * own arrays of them.  Test infrastructure: built with -fsyntax-only by `make -C oracle layout`
 * where the reference tree is mounted (tests/test_abi.py); nothing is linked or run.
 */
#include <stddef.h>

#include "common.h"          /* the reference's model/common.h */
#include "photic_samodel.h"  /* this library's mirror */

#define SAME_SIZE(a, b) _Static_assert(sizeof(a) == sizeof(b), "size of " #b)
#define SAME_FIELD(a, b, f)                                                            \
    _Static_assert(offsetof(a, f) == offsetof(b, f) &&                                 \
                       sizeof(((a *)0)->f) == sizeof(((b *)0)->f),                     \
                   #b "." #f)

SAME_SIZE(geogrid, struct photic_geogrid);
SAME_FIELD(geogrid, struct photic_geogrid, nrows);
SAME_FIELD(geogrid, struct photic_geogrid, ncols);
SAME_FIELD(geogrid, struct photic_geogrid, cellsize);
SAME_FIELD(geogrid, struct photic_geogrid, wlon);
SAME_FIELD(geogrid, struct photic_geogrid, slat);
SAME_FIELD(geogrid, struct photic_geogrid, elon);
SAME_FIELD(geogrid, struct photic_geogrid, nlat);
SAME_FIELD(geogrid, struct photic_geogrid, nodata_value);
SAME_FIELD(geogrid, struct photic_geogrid, lambda);
SAME_FIELD(geogrid, struct photic_geogrid, theta_v);
SAME_FIELD(geogrid, struct photic_geogrid, theta_w);
SAME_FIELD(geogrid, struct photic_geogrid, array);

SAME_SIZE(scene, struct photic_scene);
SAME_FIELD(scene, struct photic_scene, scene_name);
SAME_FIELD(scene, struct photic_scene, n_bands);
SAME_FIELD(scene, struct photic_scene, nrows);
SAME_FIELD(scene, struct photic_scene, ncols);
SAME_FIELD(scene, struct photic_scene, band_indexes);
SAME_FIELD(scene, struct photic_scene, wavelengths);
SAME_FIELD(scene, struct photic_scene, theta_w);
SAME_FIELD(scene, struct photic_scene, theta_v);
SAME_FIELD(scene, struct photic_scene, H_tide);
SAME_FIELD(scene, struct photic_scene, R_inf);
SAME_FIELD(scene, struct photic_scene, R_sigma);
SAME_FIELD(scene, struct photic_scene, K);
SAME_FIELD(scene, struct photic_scene, K_sigma);
SAME_FIELD(scene, struct photic_scene, ratio_min);
SAME_FIELD(scene, struct photic_scene, ratio_max);
SAME_FIELD(scene, struct photic_scene, slope_min);
SAME_FIELD(scene, struct photic_scene, slope_max);
SAME_FIELD(scene, struct photic_scene, pgx_present);
SAME_FIELD(scene, struct photic_scene, ps_p);
SAME_FIELD(scene, struct photic_scene, ps_g);
SAME_FIELD(scene, struct photic_scene, ps_x);
