/* photic_amd -- C host code: ESRI ASCII grids (include/photic_io.h), the reference's
 * real-scene grid format (model/io.c:143-269), and the int16 packing of its netCDF output
 * grids (model/nc.c:244-300). */
#include "photic_io.h"

#include <float.h>
#include <limits.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "photic.h"
#include "ph_launch.h"

void photic_free_grid(float **grid, int nrows)
{
    if (!grid) return;
    for (int i = 0; i < nrows; i++) free(grid[i]);
    free(grid);
}

/* One "keyword value" header line; the keyword must match. */
static int io_header(FILE *fp, const char *key, double *v)
{
    char word[64];
    if (fscanf(fp, "%63s %lf", word, v) != 2 || strcmp(word, key) != 0) return 0;
    return 1;
}

int photic_read_asc(const char *path, float ***grid, int *ncols, int *nrows, double *xllcorner,
                    double *yllcorner, double *cellsize, double *nodata_value)
{
    if (!path || !grid || !ncols || !nrows || !xllcorner || !yllcorner || !cellsize ||
        !nodata_value)
        return ph_error(PHOTIC_EINVAL, "photic_read_asc: a NULL argument");
    *grid = NULL;
    FILE *fp = fopen(path, "r");
    if (!fp) return ph_error(PHOTIC_EINVAL, "photic_read_asc: cannot open the file");
    double nc = 0.0, nr = 0.0;
    int ok = io_header(fp, "ncols", &nc) && io_header(fp, "nrows", &nr) &&
             io_header(fp, "xllcorner", xllcorner) && io_header(fp, "yllcorner", yllcorner) &&
             io_header(fp, "cellsize", cellsize) && io_header(fp, "NODATA_value", nodata_value);
    if (!ok || nc < 1.0 || nr < 1.0 || nc > 1.0e9 || nr > 1.0e9 || nc != (double)(int)nc ||
        nr != (double)(int)nr) {
        fclose(fp);
        return ph_error(PHOTIC_EINVAL, "photic_read_asc: malformed ascii grid header");
    }
    const int c = (int)nc, r = (int)nr;
    float **g = (float **)calloc((size_t)r, sizeof(float *));
    int rc = g ? PHOTIC_OK : PHOTIC_ENOMEM;
    for (int i = 0; !rc && i < r; i++) {
        g[i] = (float *)malloc(sizeof(float) * (size_t)c);
        if (!g[i]) rc = PHOTIC_ENOMEM;
    }
    /* values north row first: file row k is grid row r - 1 - k (model/io.c:224-232) */
    long have = 0;
    const long want = (long)r * c;
    float v;
    while (!rc && have < want && fscanf(fp, "%f", &v) == 1) {
        g[r - 1 - (int)(have / c)][have % c] = v;
        have++;
    }
    fclose(fp);
    if (!rc && have < want) rc = ph_error(PHOTIC_EINVAL, "photic_read_asc: too few values");
    if (rc == PHOTIC_ENOMEM) ph_error(rc, "photic_read_asc: out of memory");
    if (rc) {
        if (g) photic_free_grid(g, r);
        return rc;
    }
    *grid = g;
    *ncols = c;
    *nrows = r;
    return PHOTIC_OK;
}

int photic_write_asc(const char *path, float **grid, int ncols, int nrows, double xllcorner,
                     double yllcorner, double cellsize, double nodata_value)
{
    if (!path || !grid || ncols < 1 || nrows < 1)
        return ph_error(PHOTIC_EINVAL, "photic_write_asc: bad arguments");
    FILE *fp = fopen(path, "w");
    if (!fp) return ph_error(PHOTIC_EINVAL, "photic_write_asc: cannot create the file");
    fprintf(fp, "ncols        %d\n", ncols);
    fprintf(fp, "nrows        %d\n", nrows);
    fprintf(fp, "xllcorner    %lf\n", xllcorner);
    fprintf(fp, "yllcorner    %lf\n", yllcorner);
    fprintf(fp, "cellsize     %lf\n", cellsize);
    fprintf(fp, "NODATA_value %lf\n", nodata_value);
    for (int i = nrows - 1; i >= 0; i--) {
        for (int j = 0; j < ncols; j++) fprintf(fp, "%.2f ", grid[i][j]);
        fprintf(fp, "\n");
    }
    if (fclose(fp) != 0) return ph_error(PHOTIC_EINVAL, "photic_write_asc: write failed");
    return PHOTIC_OK;
}

/* rint(q) converted to short as the reference's x86-64 build converts it: through a 32-bit
 * integer (cvttsd2si; a NaN or a value outside int's range gives INT_MIN), keeping its low 16
 * bits. */
static short io_to_short(double r)
{
    int32_t t = (r > -2147483649.0 && r < 2147483648.0) ? (int32_t)r : INT32_MIN;
    return (short)(int16_t)(uint16_t)(uint32_t)t;
}

int photic_pack_int16(const float *values, long n, double spval, short *packed,
                      float *add_offset, float *scale_factor, short *missing_value)
{
    if (n < 0 || (n > 0 && (!values || !packed)) || !add_offset || !scale_factor ||
        !missing_value)
        return ph_error(PHOTIC_EINVAL, "photic_pack_int16: a NULL argument or n < 0");
    const float fspv = (float)spval;
    *missing_value = SHRT_MIN;
    float lo = FLT_MAX, hi = FLT_MIN;
    for (long k = 0; k < n; k++) {
        const float v = values[k];
        if (v == fspv) continue;
        if (v < lo) lo = v;
        else if (v > hi) hi = v;
    }
    const float scale = (hi - lo) / (float)SHRT_MAX;
    *add_offset = lo;
    *scale_factor = scale;
    for (long k = 0; k < n; k++) {
        const float v = values[k];
        packed[k] = v == fspv ? *missing_value : io_to_short(rint((double)((v - lo) / scale)));
    }
    return PHOTIC_OK;
}

int photic_unpack_int16(const short *packed, long n, float add_offset, float scale_factor,
                        short missing_value, double spval, float *values)
{
    if (n < 0 || (n > 0 && (!packed || !values)))
        return ph_error(PHOTIC_EINVAL, "photic_unpack_int16: a NULL argument or n < 0");
    for (long k = 0; k < n; k++) {
        if (packed[k] == missing_value) {
            values[k] = (float)spval;
        } else {
            const float c = (float)packed[k];
            const float t = c * scale_factor;
            values[k] = t + add_offset;
        }
    }
    return PHOTIC_OK;
}
