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

/* The int16 packing of the reference's netCDF output grids (compress_2d, model/nc.c:265-300;
 * samodel() writes each output grid packed this way, model/samodel.c:1494-1689): n values
 * (row-major) are stored as short integers c with value = c * scale_factor + add_offset, and
 * values equal to (float)spval as *missing_value = SHRT_MIN.  The arithmetic is the
 * reference's, in float: add_offset is the smallest value, scale_factor = (largest - smallest)
 * / 32767, c = rint((v - add_offset) / scale_factor) -- including its scan, which starts the
 * largest value at FLT_MIN (so a grid of negative values keeps FLT_MIN as its largest) and
 * tests a value for the largest only when it is not a new smallest (so a first value that is
 * the grid's maximum is not seen as the largest), and its conversion to short through a
 * 32-bit integer keeping the low 16 bits, as the reference's x86-64 build does for a quotient
 * out of range.  Replaces compress_2d, which is not visible outside nc.c; no netCDF library is
 * needed. */
int photic_pack_int16(const float *values, long n, double spval, short *packed,
                      float *add_offset, float *scale_factor, short *missing_value);

/* The inverse (decompress_2d, model/nc.c:244-262): values[k] = (float)spval where packed[k]
 * equals missing_value, else (float)packed[k] * scale_factor + add_offset in float. */
int photic_unpack_int16(const short *packed, long n, float add_offset, float scale_factor,
                        short missing_value, double spval, float *values);

#ifdef __cplusplus
}
#endif
#endif
