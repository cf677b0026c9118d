// photic_amd -- synthetic scenes on the GPU.
//
// Counter-based Philox4x32-10 (Salmon, Moraes, Dror & Shaw, SC 2011) keyed by the scene seed
// with counter (pixel, block, date), so every pixel's draws are independent of tiling and of
// which device makes them.  Per pixel: H ~ U(0.5, 25) m (or U(30, 60) for the optically deep
// fraction), P, G, X log-normal truncated to [0.005, 0.2], [0.002, 0.2], [0.001, 0.05] 1/m
// (mu = mean of the log bounds, sigma = a quarter of their spread, redrawn when outside),
// B ~ U(0, 1) sand/seagrass mix, sun zenith ~ U(20, 50) deg, view at nadir; Rrs from the FP64
// forward model with the backscatter exponent Y made self-consistent with the modelled Rrs
// (20 fixed-point passes from Y = 1), plus N(0, sigma) noise per band, rounded to float.
//
// Draw map (counter word 2 = block, word 3 = date):
//   block 0: u_H, u_B          block 1: u_theta, u_deep     block 2: u_Hdeep
//   blocks 16+a / 48+a / 80+a: P / G / X attempt a (a < 32)  blocks 112+b: noise of band b
//   stacks: cloud mask of date d from block 3 (redraw r from block 3 + r, r = 1..16)
//   no-data discs: disc i (counter word 0 = i) centre from block 250, radius from block 251
//   smooth fields: field q, wave m from blocks 200 + 4q + 2m and 201 + 4q + 2m (pixel 0)
#pragma once

#include "ph_engine.h"

PH_INLINE void ph_philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                          uint32_t k0, uint32_t k1, uint32_t out[4])
{
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

PH_INLINE void ph_uniforms(uint64_t seed, uint64_t pix, uint32_t block,
                                            uint32_t date, double* u0, double* u1)
{
    uint32_t o[4];
    ph_philox((uint32_t)pix, (uint32_t)(pix >> 32), block, date, (uint32_t)seed,
              (uint32_t)(seed >> 32), o);
    uint64_t a = (((uint64_t)o[0] << 32) | o[1]) >> 11;
    uint64_t b = (((uint64_t)o[2] << 32) | o[3]) >> 11;
    *u0 = (double)a * 0x1p-53;
    *u1 = (double)b * 0x1p-53;
}

PH_INLINE double ph_normal(uint64_t seed, uint64_t pix, uint32_t block,
                                            uint32_t date)
{
    double u0, u1;
    ph_uniforms(seed, pix, block, date, &u0, &u1);
    double r = sqrt(-2.0 * log(1.0 - u0));
    return r * cos(6.283185307179586 * u1);
}

PH_INLINE double ph_trunc_lognormal(uint64_t seed, uint64_t pix, uint32_t base,
                                                     uint32_t date, double lo, double hi)
{
    double llo = log(lo), lhi = log(hi);
    double mu = 0.5 * (llo + lhi), sg = 0.25 * (lhi - llo);
    double v = 0.0;
    for (uint32_t a = 0; a < 32; a++) {
        v = exp(mu + sg * ph_normal(seed, pix, base + a, date));
        if (v >= lo && v <= hi) return v;
    }
    return v < lo ? lo : hi;
}

struct PhTruth {
    double H, P, G, X, B, theta;
};

// Rrs of the generating parameters with self-consistent Y (bands from c), FP64.
PH_INLINE void ph_truth_rrs(const PhTruth& t, const PhConst& c, int nb,
                                             double* rrs)
{
    double ss, th = t.theta;
    th *= 3.141592653589793 / 180.0;
    ss = 1.0 / cos(th);
    double Y = 1.0;
    for (int it = 0; it <= 20; it++) {
        for (int b = 0; b < nb; b++) {
            double rho = t.B * c.bot[0][b] + (1.0 - t.B) * c.bot[1][b];
            double pw = pow(440.0 / c.wl[b], Y);
            double K, rB, r;
            rrs[b] = ph_band64(t.P, t.G, t.X, rho, t.H, c.a0[b], c.a1[b], c.aw[b], c.bbw[b],
                               c.eg[b], pw, ss, 1.0, &K, &rB, &r);
        }
        if (it == 20) break;
        double r440 = ph_rrs_at_ptr(c, rrs, 0);
        double r490 = ph_rrs_at_ptr(c, rrs, 1);
        if (r440 < 0.0) r440 = 0.0001;
        double chi = r440 / r490;
        Y = 3.44 * (1.0 - 3.17 * exp(-2.01 * chi));
        if (Y < 0.0) Y = 0.0;
        if (Y > 2.5) Y = 2.5;
    }
}

// One single-date pixel; returns true when a band went negative after noise (written as the
// nodata value).
PH_INLINE bool ph_gen_pixel(uint64_t seed, uint64_t pix, const PhConst& c,
                                             int nb, double sigma, double deep_frac,
                                             PhTruth* truth, float* rrs_out)
{
    double uH, uB, uT, uD, uH2, unused;
    ph_uniforms(seed, pix, 0, 0, &uH, &uB);
    ph_uniforms(seed, pix, 1, 0, &uT, &uD);
    ph_uniforms(seed, pix, 2, 0, &uH2, &unused);
    PhTruth t;
    t.H = 0.5 + 24.5 * uH;
    if (deep_frac > 0.0 && uD < deep_frac) t.H = 30.0 + 30.0 * uH2;
    t.B = uB;
    t.theta = 20.0 + 30.0 * uT;
    t.P = ph_trunc_lognormal(seed, pix, 16, 0, 0.005, 0.2);
    t.G = ph_trunc_lognormal(seed, pix, 48, 0, 0.002, 0.2);
    t.X = ph_trunc_lognormal(seed, pix, 80, 0, 0.001, 0.05);
    *truth = t;
    double rrs[PH_MAXB];
    ph_truth_rrs(t, c, nb, rrs);
    bool nodata = false;
    for (int b = 0; b < nb; b++) {
        float f = (float)(rrs[b] + sigma * ph_normal(seed, pix, 112 + (uint32_t)b, 0));
        if (f < 0.0f) {
            nodata = true;
            f = (float)c.nodata;
        }
        rrs_out[b] = f;
    }
    return nodata;
}

// A smooth field in [0, 1] over the grid for the region workload: 0.5 plus two plane waves of
// amplitude 0.25, wave m of field q with direction 2 pi u0, wavelength 64 + 192 u1 pixels and
// phase 2 pi u2 drawn from the scene seed.
PH_INLINE double ph_field(uint64_t seed, int q, long i, long j)
{
    double s = 0.5;
    for (int m = 0; m < 2; m++) {
        double u0, u1, u2, u3;
        ph_uniforms(seed, 0, 200 + 4 * (uint32_t)q + 2 * (uint32_t)m, 0, &u0, &u1);
        ph_uniforms(seed, 0, 201 + 4 * (uint32_t)q + 2 * (uint32_t)m, 0, &u2, &u3);
        const double a = 6.283185307179586 * u0, L = 64.0 + 192.0 * u1;
        const double ph = 6.283185307179586 * u2;
        s += 0.25 * sin(6.283185307179586 * ((double)i * cos(a) + (double)j * sin(a)) / L + ph);
    }
    return s;
}

// Pixel (i, j) of a spatially coherent single-date scene: H = 0.5 + 24.5 s_H m, B = s_B,
// P, G, X log-uniform over their bounds by s_P, s_G, s_X, one sun zenith for the scene, the
// noise of pixel i * n_cols + j as in the single-date scene.  Returns true when a band went
// negative after noise (written as the nodata value).
PH_INLINE bool ph_gen_field_pixel(uint64_t seed, long i, long j, long n_cols, double theta,
                                  const PhConst& c, int nb, double sigma, PhTruth* truth,
                                  float* rrs_out)
{
    const uint64_t pix = (uint64_t)(i * n_cols + j);
    PhTruth t;
    t.H = 0.5 + 24.5 * ph_field(seed, 0, i, j);
    t.B = ph_field(seed, 1, i, j);
    t.P = exp(log(0.005) + ph_field(seed, 2, i, j) * (log(0.2) - log(0.005)));
    t.G = exp(log(0.002) + ph_field(seed, 3, i, j) * (log(0.2) - log(0.002)));
    t.X = exp(log(0.001) + ph_field(seed, 4, i, j) * (log(0.05) - log(0.001)));
    t.theta = theta;
    *truth = t;
    double rrs[PH_MAXB];
    ph_truth_rrs(t, c, nb, rrs);
    bool nodata = false;
    for (int b = 0; b < nb; b++) {
        float f = (float)(rrs[b] + sigma * ph_normal(seed, pix, 112 + (uint32_t)b, 0));
        if (f < 0.0f) {
            nodata = true;
            f = (float)c.nodata;
        }
        rrs_out[b] = f;
    }
    return nodata;
}

// One pixel of a multi-date stack: shared H and B (block 0 of date 0), per date d the sun
// zenith (block 1), P, G, X and the band noise with counter word 3 = d; a cloud mask
// (clear when u >= p_cloud) with at least k_min clear dates, redrawn up to 16 times and then
// forced clear on the first k_min dates; a date whose noisy Rrs goes negative is invalid.
// truth: H, B, then (P, G, X, theta) per date.  Returns the number of valid dates.
PH_INLINE int ph_gen_stack_pixel(uint64_t seed, uint64_t pix, const PhConst& c, int nb,
                                 double sigma, int n_dates, double p_cloud, int k_min,
                                 double* truth, float* rrs_out, unsigned char* valid)
{
    double uH, uB, u0, u1;
    ph_uniforms(seed, pix, 0, 0, &uH, &uB);
    const double H = 0.5 + 24.5 * uH, B = uB;
    truth[0] = H;
    truth[1] = B;
    int nclear = 0;
    for (int r = 0; r < 17; r++) {
        nclear = 0;
        for (int d = 0; d < n_dates; d++) {
            ph_uniforms(seed, pix, 3u + (uint32_t)r, (uint32_t)d, &u0, &u1);
            valid[d] = u0 >= p_cloud;
            nclear += valid[d];
        }
        if (nclear >= k_min) break;
    }
    if (nclear < k_min)
        for (int d = 0; d < n_dates; d++) valid[d] = d < k_min;
    int k = 0;
    for (int d = 0; d < n_dates; d++) {
        PhTruth t;
        ph_uniforms(seed, pix, 1, (uint32_t)d, &u0, &u1);
        t.H = H;
        t.B = B;
        t.theta = 20.0 + 30.0 * u0;
        t.P = ph_trunc_lognormal(seed, pix, 16, (uint32_t)d, 0.005, 0.2);
        t.G = ph_trunc_lognormal(seed, pix, 48, (uint32_t)d, 0.002, 0.2);
        t.X = ph_trunc_lognormal(seed, pix, 80, (uint32_t)d, 0.001, 0.05);
        truth[2 + 4 * d] = t.P;
        truth[3 + 4 * d] = t.G;
        truth[4 + 4 * d] = t.X;
        truth[5 + 4 * d] = t.theta;
        double rrs[PH_MAXB];
        ph_truth_rrs(t, c, nb, rrs);
        for (int b = 0; b < nb; b++) {
            float f = (float)(rrs[b] + sigma * ph_normal(seed, pix, 112 + (uint32_t)b,
                                                         (uint32_t)d));
            if (f < 0.0f) valid[d] = 0;
            rrs_out[d * nb + b] = f;
        }
        k += valid[d];
    }
    return k;
}

// No-data blobs of a scene of n_rows x n_cols pixels (pixel index row * n_cols + col): n_discs
// discs, disc i centred at (u0 n_cols, u1 n_rows) (block 250 of counter i) with radius
// r_min + (r_max - r_min) u2 (block 251); a pixel is masked when its centre lies strictly
// inside a disc.  Expected masked fraction about 1 - exp(-n_discs pi E[r^2] / area).
PH_INLINE bool ph_in_nodata_disc(uint64_t seed, long row, long col, long n_rows, long n_cols,
                                 int n_discs, double r_min, double r_max)
{
    const double y = (double)row + 0.5, x = (double)col + 0.5;
    for (int i = 0; i < n_discs; i++) {
        double u0, u1, u2, u3;
        ph_uniforms(seed, (uint64_t)i, 250u, 0u, &u0, &u1);
        ph_uniforms(seed, (uint64_t)i, 251u, 0u, &u2, &u3);
        const double cx = u0 * (double)n_cols, cy = u1 * (double)n_rows;
        const double r = r_min + (r_max - r_min) * u2;
        const double dx = x - cx, dy = y - cy;
        if (dx * dx + dy * dy < r * r) return true;
    }
    return false;
}
