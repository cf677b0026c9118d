/*
 * ORACLE -- test infrastructure, not product code.
 *
 * Synthetic scenes: counter-based Philox4x32-10 (Salmon, Moraes, Dror & Shaw, "Parallel
 * random numbers: as easy as 1, 2, 3", SC 2011) keyed by (seed) with counter (pixel, block,
 * date), so every pixel's draws are independent of tiling and of the device that makes them.
 * The product's GPU generator must reproduce these draws; tests compare the two.
 *
 * Draw map per pixel (counter word 2 = block, word 3 = date):
 *   block 0      : u_H, u_B              block 1 : u_theta, u_deep
 *   block 2      : u_Hdeep, (unused)
 *   blocks 16+a  : P attempt a (a < 32)  blocks 48+a : G attempt a   blocks 80+a : X attempt a
 *   blocks 112+b : noise of band b
 *   stacks       : cloud mask of date d from block 3 (redraw r from block 3 + r, r = 1..16)
 * Each block yields two 53-bit uniforms; a normal uses one block (Box-Muller, cosine branch).
 */
#include "photic_oracle.h"

#include <math.h>
#include <string.h>

void orc_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4])
{
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)M0 * c0, p1 = (uint64_t)M1 * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += W0;
        k1 += W1;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

/* Two uniforms in [0, 1) with 53 random bits each from one Philox block. */
static void orc_block_uniforms(uint64_t seed, uint64_t pix, uint32_t block, uint32_t date,
                               double *u0, double *u1)
{
    uint32_t ctr[4] = {(uint32_t)pix, (uint32_t)(pix >> 32), block, date};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t o[4];
    orc_philox4x32_10(ctr, key, o);
    uint64_t a = ((uint64_t)o[0] << 32 | o[1]) >> 11;
    uint64_t b = ((uint64_t)o[2] << 32 | o[3]) >> 11;
    *u0 = (double)a * 0x1p-53;
    *u1 = (double)b * 0x1p-53;
}

static double orc_block_normal(uint64_t seed, uint64_t pix, uint32_t block, uint32_t date)
{
    double u0, u1;
    orc_block_uniforms(seed, pix, block, date, &u0, &u1);
    double r = sqrt(-2.0 * log(1.0 - u0)); /* 1 - u0 is in (0, 1] */
    return r * cos(6.283185307179586 * u1);
}

/* Log-normal truncated to [lo, hi]: mu = mean of the log bounds, sigma = a quarter of
 * their spread; out-of-range draws are redrawn (32 attempts, then clamped). */
static double orc_trunc_lognormal(uint64_t seed, uint64_t pix, uint32_t base, uint32_t date,
                                  double lo, double hi)
{
    double llo = log(lo), lhi = log(hi);
    double mu = 0.5 * (llo + lhi), sg = 0.25 * (lhi - llo);
    double v = 0.0;
    for (uint32_t a = 0; a < 32; a++) {
        double z = orc_block_normal(seed, pix, base + a, date);
        v = exp(mu + sg * z);
        if (v >= lo && v <= hi)
            return v;
    }
    return v < lo ? lo : hi;
}

/* Forward model of the generating parameters, with the backscatter exponent Y made
 * self-consistent with the modelled Rrs by n_fixed_point iterations from Y = 1. */
void orc_fwd_truth(const orc_truth *t, const orc_bands *bands, int n_fixed_point,
                   double *rrs_out, double *Y_out)
{
    int nb = bands->n_bands;
    double rho[ORC_MAXB], ss = orc_sec_deg(t->theta_sun), sv = 1.0;
    for (int b = 0; b < nb; b++)
        rho[b] = t->B * bands->bottom[0][b] + (1.0 - t->B) * bands->bottom[1][b];
    double Y = 1.0;
    for (int it = 0; it <= n_fixed_point; it++) {
        for (int b = 0; b < nb; b++)
            rrs_out[b] = orc_rrs_core(t->P, t->G, t->X, rho[b], t->H, Y, bands, b, ss, sv, NULL);
        if (it == n_fixed_point)
            break;
        double r440 = orc_interp_1d(bands->wl, rrs_out, nb, 440.0);
        double r490 = orc_interp_1d(bands->wl, rrs_out, nb, 490.0);
        if (r440 < 0.0)
            r440 = 0.0001;
        Y = orc_Y_of_chi(r440 / r490);
    }
    if (Y_out)
        *Y_out = Y;
}

int orc_gen_pixel(uint64_t seed, uint64_t pix, const orc_bands *bands, double noise_sigma,
                  double deep_frac, orc_truth *truth, float *rrs_out, double *Y_out)
{
    double uH, uB, uT, uD, uH2, unused;
    orc_block_uniforms(seed, pix, 0, 0, &uH, &uB);
    orc_block_uniforms(seed, pix, 1, 0, &uT, &uD);
    orc_block_uniforms(seed, pix, 2, 0, &uH2, &unused);
    orc_truth t;
    t.H = 0.5 + 24.5 * uH;
    if (deep_frac > 0.0 && uD < deep_frac)
        t.H = 30.0 + 30.0 * uH2;
    t.B = uB;
    t.theta_sun = 20.0 + 30.0 * uT;
    t.P = orc_trunc_lognormal(seed, pix, 16, 0, 0.005, 0.2);
    t.G = orc_trunc_lognormal(seed, pix, 48, 0, 0.002, 0.2);
    t.X = orc_trunc_lognormal(seed, pix, 80, 0, 0.001, 0.05);
    if (truth)
        *truth = t;
    double rrs[ORC_MAXB];
    orc_fwd_truth(&t, bands, 20, rrs, Y_out);
    int nodata = 0;
    for (int b = 0; b < bands->n_bands; b++) {
        double v = rrs[b] + noise_sigma * orc_block_normal(seed, pix, 112 + (uint32_t)b, 0);
        float f = (float)v;
        if (f < 0.0f)
            nodata = 1;
        rrs_out[b] = f;
    }
    return nodata;
}

/* A smooth field in [0, 1] for the region workload: 0.5 plus two plane waves of amplitude
 * 0.25; wave m of field q takes direction 2 pi u0, wavelength 64 + 192 u1 pixels and phase
 * 2 pi u2 from blocks 200 + 4q + 2m and 201 + 4q + 2m of pixel 0. */
static double orc_field(uint64_t seed, int q, long i, long j)
{
    double s = 0.5;
    for (int m = 0; m < 2; m++) {
        double u0, u1, u2, u3;
        orc_block_uniforms(seed, 0, 200 + 4 * (uint32_t)q + 2 * (uint32_t)m, 0, &u0, &u1);
        orc_block_uniforms(seed, 0, 201 + 4 * (uint32_t)q + 2 * (uint32_t)m, 0, &u2, &u3);
        double a = 6.283185307179586 * u0, L = 64.0 + 192.0 * u1, ph = 6.283185307179586 * u2;
        s += 0.25 * sin(6.283185307179586 * ((double)i * cos(a) + (double)j * sin(a)) / L + ph);
    }
    return s;
}

int orc_gen_field_pixel(uint64_t seed, long i, long j, long n_cols, double theta,
                        const orc_bands *bands, double noise_sigma, orc_truth *truth,
                        float *rrs_out)
{
    const uint64_t pix = (uint64_t)(i * n_cols + j);
    orc_truth t;
    t.H = 0.5 + 24.5 * orc_field(seed, 0, i, j);
    t.B = orc_field(seed, 1, i, j);
    t.P = exp(log(0.005) + orc_field(seed, 2, i, j) * (log(0.2) - log(0.005)));
    t.G = exp(log(0.002) + orc_field(seed, 3, i, j) * (log(0.2) - log(0.002)));
    t.X = exp(log(0.001) + orc_field(seed, 4, i, j) * (log(0.05) - log(0.001)));
    t.theta_sun = theta;
    if (truth)
        *truth = t;
    double rrs[ORC_MAXB];
    orc_fwd_truth(&t, bands, 20, rrs, NULL);
    int nodata = 0;
    for (int b = 0; b < bands->n_bands; b++) {
        float f = (float)(rrs[b] + noise_sigma * orc_block_normal(seed, pix, 112 + (uint32_t)b, 0));
        if (f < 0.0f)
            nodata = 1;
        rrs_out[b] = f;
    }
    return nodata;
}

/* Multi-date stacks.  Per pixel: H, B from block 0 of date 0 (as the single-date scene); per
 * date d: sun zenith (block 1), P, G, X (blocks 16+a, 48+a, 80+a) and noise (112+b) with
 * counter word 3 = d; cloud mask from block 3 of date d (clear when u >= p_cloud).  A mask
 * with fewer than k_min clear dates is redrawn from blocks 4 + r (r < 16), then the first
 * k_min dates are forced clear.  A date whose noisy Rrs goes negative is marked invalid.
 * truth: H, B, then (P, G, X, theta_sun) per date. */
int orc_gen_stack_pixel(uint64_t seed, uint64_t pix, const orc_bands *bands,
                        double noise_sigma, int n_dates, double p_cloud, int k_min,
                        double *truth, float *rrs_out /*[n_dates][nb]*/, uint8_t *valid)
{
    double uH, uB, u0, u1;
    orc_block_uniforms(seed, pix, 0, 0, &uH, &uB);
    const double H = 0.5 + 24.5 * uH, B = uB;
    truth[0] = H;
    truth[1] = B;
    int nclear = 0;
    for (int r = 0; r < 17; r++) {
        nclear = 0;
        for (int d = 0; d < n_dates; d++) {
            orc_block_uniforms(seed, pix, 3u + (uint32_t)r, (uint32_t)d, &u0, &u1);
            valid[d] = u0 >= p_cloud;
            nclear += valid[d];
        }
        if (nclear >= k_min)
            break;
    }
    if (nclear < k_min)
        for (int d = 0; d < n_dates; d++)
            valid[d] = d < k_min;
    int nb = bands->n_bands, k = 0;
    for (int d = 0; d < n_dates; d++) {
        orc_truth t;
        orc_block_uniforms(seed, pix, 1, (uint32_t)d, &u0, &u1);
        t.H = H;
        t.B = B;
        t.theta_sun = 20.0 + 30.0 * u0;
        t.P = orc_trunc_lognormal(seed, pix, 16, (uint32_t)d, 0.005, 0.2);
        t.G = orc_trunc_lognormal(seed, pix, 48, (uint32_t)d, 0.002, 0.2);
        t.X = orc_trunc_lognormal(seed, pix, 80, (uint32_t)d, 0.001, 0.05);
        truth[2 + 4 * d] = t.P;
        truth[3 + 4 * d] = t.G;
        truth[4 + 4 * d] = t.X;
        truth[5 + 4 * d] = t.theta_sun;
        double rrs[ORC_MAXB];
        orc_fwd_truth(&t, bands, 20, rrs, NULL);
        for (int b = 0; b < nb; b++) {
            double v = rrs[b] + noise_sigma * orc_block_normal(seed, pix, 112 + (uint32_t)b,
                                                               (uint32_t)d);
            float f = (float)v;
            if (f < 0.0f)
                valid[d] = 0;
            rrs_out[d * nb + b] = f;
        }
        k += valid[d];
    }
    return k;
}

/* No-data discs (the build's scene-scale masks): disc i's centre (u0 n_cols, u1 n_rows) from
 * block 250 of counter i, radius r_min + (r_max - r_min) u from block 251; 1 when the pixel's
 * centre lies strictly inside a disc.  Not a reference function: the masks are this build's
 * synthetic input, drawn identically on both sides. */
int orc_in_nodata_disc(uint64_t seed, long row, long col, long n_rows, long n_cols, int n_discs,
                       double r_min, double r_max)
{
    const double y = (double)row + 0.5, x = (double)col + 0.5;
    for (int i = 0; i < n_discs; i++) {
        double u0, u1, u2, u3;
        orc_block_uniforms(seed, (uint64_t)i, 250u, 0u, &u0, &u1);
        orc_block_uniforms(seed, (uint64_t)i, 251u, 0u, &u2, &u3);
        const double cx = u0 * (double)n_cols, cy = u1 * (double)n_rows;
        const double r = r_min + (r_max - r_min) * u2;
        const double dx = x - cx, dy = y - cy;
        if (dx * dx + dy * dy < r * r)
            return 1;
    }
    return 0;
}
