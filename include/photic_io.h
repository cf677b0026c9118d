/*
 * photic_io.h -- ESRI ASCII grids in and out, the reference's real-scene format for band,
 * depth and output grids (model/io.c:143-269).  Host code only; no GPU needed.
 *
 * The grid is held as the reference holds it: separately allocated rows behind a float**,
 * row 0 the SOUTHERN edge (the file lists rows north to south, model/io.c:224-240).  Errors are
 * returned as PHOTIC_* codes with a message in photic_last_error(); nothing calls exit().
 */
#ifndef PHOTIC_IO_H
#define PHOTIC_IO_H

#ifdef __cplusplus
extern "C" {
#endif

/* Reads an ESRI ASCII grid: the header "ncols", "nrows", "xllcorner", "yllcorner",
 * "cellsize", "NODATA_value" in that order, then nrows x ncols values, north row first.
 * *grid receives nrows separately allocated rows (free with photic_free_grid), row 0 south.
 * Replaces read_asc (model/io.c:143-232), which exits on a malformed header; here that is
 * PHOTIC_EINVAL, as is a file with fewer than nrows x ncols values (the reference leaves the
 * rest unset). */
int photic_read_asc(const char *path, float ***grid, int *ncols, int *nrows, double *xllcorner,
                    double *yllcorner, double *cellsize, double *nodata_value);

/* Writes grid (row 0 south) as an ESRI ASCII grid, north row first, values as "%.2f " -- the
 * layout of write_asc (model/io.c:235-269). */
int photic_write_asc(const char *path, float **grid, int ncols, int nrows, double xllcorner,
                     double yllcorner, double cellsize, double nodata_value);

/* Frees a grid from photic_read_asc. */
void photic_free_grid(float **grid, int nrows);

#ifdef __cplusplus
}
#endif
#endif
