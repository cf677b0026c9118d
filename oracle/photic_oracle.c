/*
 * ORACLE -- test infrastructure, not product code.  See photic_oracle.h for what it is,
 * who may load it and how each function is pinned.
 *
 * Arithmetic follows the reference's order of operations so that results can be compared
 * bit for bit where the reference is compiled (nelmin, interp_1d) and to rounding elsewhere.
 * Build with -ffp-contract=off: the reference's x86-64 build has no fused multiply-add.
 */
#include "photic_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------------------ */
/* helpers: model/common.c:392-406 (approx_equal and friends), :298-335 (interp_1d)      */
/* ------------------------------------------------------------------------------------ */

/* Knuth's relative test with both operands rounded to float first, as the reference's
 * approx_equal(float, float, float) does (model/common.c:392-394). */
int orc_approx_equal(float a, float b, float eps)
{
    float d = a - b;
    double fa = fabs((double)a), fb = fabs((double)b);
    return fabs((double)d) <= ((fa < fb ? fb : fa) * (double)eps);
}

static int orc_approx_le(float a, float b, float eps)
{
    return a < b || orc_approx_equal(a, b, eps);
}

static int orc_approx_ge(float a, float b, float eps)
{
    return a > b || orc_approx_equal(a, b, eps);
}

/* Piecewise-linear interpolation with linear extrapolation at both ends, node tests made
 * in float with relative tolerance 1e-5 (model/common.c:298-335). */
double orc_interp_1d(const double *X, const double *Y, int n, double x)
{
    double x0 = 0.0, x1 = 1.0, y0 = 0.0, y1 = 0.0, alpha;
    if (orc_approx_equal((float)x, (float)X[0], 1.0e-5f))
        return Y[0];
    if (orc_approx_equal((float)x, (float)X[n - 1], 1.0e-5f))
        return Y[n - 1];
    if (X[0] < X[n - 1] && x < X[0]) {
        x0 = X[0]; x1 = X[1]; y0 = Y[0]; y1 = Y[1];
    } else if (X[n - 1] > X[0] && x > X[n - 1]) {
        x0 = X[n - 2]; x1 = X[n - 1]; y0 = Y[n - 2]; y1 = Y[n - 1];
    } else {
        for (int i = 0; i < n - 1; i++) {
            float xi = (float)X[i], xj = (float)X[i + 1], xf = (float)x;
            if ((orc_approx_le(xi, xf, 1.0e-5f) && orc_approx_ge(xj, xf, 1.0e-5f)) ||
                (orc_approx_ge(xi, xf, 1.0e-5f) && orc_approx_le(xj, xf, 1.0e-5f))) {
                x0 = X[i]; x1 = X[i + 1]; y0 = Y[i]; y1 = Y[i + 1];
                break;
            }
        }
    }
    alpha = (x - x0) / (x1 - x0);
    return y0 * (1.0 - alpha) + y1 * alpha;
}

/* ------------------------------------------------------------------------------------ */
/* set-up: band tables (model/samodel.c:505-618), geometry (:534-548), Rrs at 440/490/   */
/* 550/640/750 nm (:1785-1799)                                                           */
/* ------------------------------------------------------------------------------------ */

void orc_band_tables(const double *wl, int n_bands, orc_bands *out)
{
    memset(out, 0, sizeof(*out));
    out->n_bands = n_bands;
    for (int b = 0; b < n_bands; b++) {
        out->wl[b] = wl[b];
        out->a0[b] = orc_interp_1d(orc_ref_wl, orc_a0, ORC_NREF, wl[b]);
        out->a1[b] = orc_interp_1d(orc_ref_wl, orc_a1, ORC_NREF, wl[b]);
        out->bbw[b] = orc_interp_1d(orc_ref_wl, orc_bbw, ORC_NREF, wl[b]);
        out->aw[b] = orc_interp_1d(orc_ref_wl, orc_aw, ORC_NREF, wl[b]);
        for (int k = 0; k < ORC_NBOTTOM_TYPES; k++)
            out->bottom[k][b] = orc_interp_1d(orc_ref_wl, orc_bottom[k], ORC_NREF, wl[b]);
    }
    out->aw640 = orc_interp_1d(orc_ref_wl, orc_aw, ORC_NREF, 640.0);
}

/* secant of a subsurface angle given in degrees: theta *= PI/180; 1/cos(theta). */
double orc_sec_deg(double theta_deg)
{
    double t = theta_deg;
    t *= ORC_PI / 180.0;
    return 1.0 / cos(t);
}

void orc_pixel_init(orc_pixel *px, int mode, int n_bottoms, const int *bottom_idx,
                    int n_dates, const orc_bands *bands, const double *sec_sun,
                    const double *sec_view, const double *rrs)
{
    memset(px, 0, sizeof(*px));
    px->mode = mode;
    if (mode == ORC_MODE_MIX) {
        px->n_bottoms = 2;
        px->bottom_idx[0] = 0; /* sand */
        px->bottom_idx[1] = 1; /* seagrass */
    } else {
        px->n_bottoms = n_bottoms;
        for (int k = 0; k < n_bottoms; k++)
            px->bottom_idx[k] = bottom_idx ? bottom_idx[k] : k;
    }
    px->n_dates = n_dates;
    px->bands = bands;
    int nb = bands->n_bands;
    for (int s = 0; s < n_dates; s++) {
        px->sec_sun[s] = sec_sun[s];
        px->sec_view[s] = sec_view[s];
        for (int b = 0; b < nb; b++)
            px->rrs[s][b] = rrs[s * nb + b];
        px->r440[s] = orc_interp_1d(bands->wl, px->rrs[s], nb, 440.0);
        px->r490[s] = orc_interp_1d(bands->wl, px->rrs[s], nb, 490.0);
        px->r550[s] = orc_interp_1d(bands->wl, px->rrs[s], nb, 550.0);
        px->r640[s] = orc_interp_1d(bands->wl, px->rrs[s], nb, 640.0);
        px->r750[s] = orc_interp_1d(bands->wl, px->rrs[s], nb, 750.0);
        if (px->r440[s] < 0.0)
            px->r440[s] = 0.0001;
    }
}

int orc_n_params(const orc_pixel *px)
{
    if (px->mode == ORC_MODE_MIX)
        return 2 + 3 * px->n_dates;
    return 1 + 2 * px->n_bottoms + 3 * px->n_dates;
}

/* ------------------------------------------------------------------------------------ */
/* forward model: model/samodel.c:2846-2949 (one date, one band, one region)             */
/* ------------------------------------------------------------------------------------ */

/* Y(chi) of model/samodel.c:2898-2901. */
double orc_Y_of_chi(double chi)
{
    double Y = 3.44 * (1.0 - 3.17 * exp(-2.01 * chi));
    if (Y < 0.0) Y = 0.0;
    if (Y > 2.5) Y = 2.5;
    return Y;
}

/* The band computation of samodel_Rrs for given bottom reflectance rho and exponent Y;
 * side[0..3] receive K, rrs_dp, rrs_B, rrs. */
double orc_rrs_core(double P, double G, double X, double rho, double H, double Y,
                    const orc_bands *t, int b, double sec_sun, double sec_view, double *side)
{
    const double S = 0.015;
    double lambda = t->wl[b];

    double a_phi = (t->a0[b] + t->a1[b] * log(fabs(P))) * fabs(P);
    double a_g = fabs(G) * exp(-S * (lambda - 440.0));
    double a = t->aw[b] + a_phi + a_g;

    double b_p = X * pow(440.0 / lambda, Y);
    double bb = t->bbw[b] + b_p;

    double u = bb / (a + bb);
    double K = a + bb;
    if (K < 0.0) K = 0.0;
    if (K > 2.5) K = 2.5;

    double rrs_dp = (0.084 + 0.170 * u) * u;
    double DuC = 1.03 * sqrt(1.0 + 2.4 * u);
    double DuB = 1.04 * sqrt(1.0 + 5.4 * u);

    double M = sec_sun + DuC * sec_view;
    double rrs_C = rrs_dp * (1.0 - exp(-M * K * H));
    M = sec_sun + DuB * sec_view;
    double rrs_B = rho / ORC_PI * exp(-M * K * H);

    double rrs = rrs_C + rrs_B;
    if (side) {
        side[0] = K;
        side[1] = rrs_dp;
        side[2] = rrs_B;
        side[3] = rrs;
    }
    return 0.5 * rrs / (1.0 - 1.5 * rrs) + 0.0;
}

double orc_rrs_band(double P, double G, double X, const double *B, const double *q,
                    double H, orc_pixel *px, int s, int b)
{
    const orc_bands *t = px->bands;
    double rho = 0.0, alb = 0.0, side[4];

    for (int i = 0; i < px->n_bottoms; i++) {
        alb += q[i] * B[i];
        rho += q[i] * B[i] * t->bottom[px->bottom_idx[i]][b];
    }
    px->bottom_albedo = alb;

    double Y = orc_Y_of_chi(px->r440[s] / px->r490[s]);
    double Rrs = orc_rrs_core(P, G, X, rho, H, Y, t, b, px->sec_sun[s], px->sec_view[s], side);
    px->K[s][b] = side[0];
    px->rrs_dp[s][b] = side[1];
    px->rrs_bot[s][b] = side[2];
    px->rrs_mod[s][b] = side[3];
    px->Rrs_mod[s][b] = Rrs;
    return Rrs;
}

/* ------------------------------------------------------------------------------------ */
/* objective: model/samodel.c:2432-2759 with one region (depth and bottom continuity     */
/* penalties are identically zero for one region, SAM is compiled out, Delta = 0)        */
/* ------------------------------------------------------------------------------------ */

double orc_objective(double *x, void *ctx)
{
    orc_pixel *px = (orc_pixel *)ctx;
    const int nd = px->n_dates, nbd = px->bands->n_bands;
    int nbot, off;
    double B[ORC_MAXBOT], q[ORC_MAXBOT], qsum = 0.0;

    px->n_evals++;
    if (px->mode == ORC_MODE_MIX) {
        /* the reference's two-bottom vector [H, 100, 100, |B|, |1-B|, P, G, X, ...] */
        nbot = 2;
        off = 2;
        B[0] = 0.01 * fabs(100.0);
        B[1] = 0.01 * fabs(100.0);
        q[0] = fabs(fabs(x[1]));
        q[1] = fabs(fabs(1.0 - x[1]));
    } else {
        nbot = px->n_bottoms;
        off = 1 + 2 * nbot;
        for (int k = 0; k < nbot; k++) {
            B[k] = 0.01 * fabs(x[1 + k]);
            q[k] = fabs(x[1 + nbot + k]);
        }
    }
    for (int k = 0; k < nbot; k++)
        qsum += q[k];
    for (int k = 0; k < nbot; k++)
        q[k] /= qsum;

    double err = 0.0, tot = 0.0;
    int n_total = 0;
    double H = fabs(x[0]);
    for (int s = 0; s < nd; s++) {
        double P = 0.01 * fabs(x[off + 3 * s]);
        double G = 0.01 * fabs(x[off + 3 * s + 1]);
        double X = 0.01 * fabs(x[off + 3 * s + 2]);
        for (int b = 0; b < nbd; b++) {
            double m = orc_rrs_band(P, G, X, B, q, H, px, s, b);
            double d = m - px->rrs[s][b];
            err += d * d;
            tot += px->rrs[s][b];
            n_total++;
        }
    }
    double mean = tot / (double)n_total;
    double e_rrs = 100.0 * sqrt(err / (double)n_total) / mean;
    px->Rrs_error = e_rrs;

    /* K penalty (model/samodel.c:2696-2732): per date, the smallest non-zero K; the last
     * date's minimum also drives the K > 0.7 term. */
    const double kmm = 0.275, kmn = 0.185;
    double Kerr = 0.0, Kmin = 1.0e4;
    for (int s = 0; s < nd; s++) {
        Kmin = 1.0e4;
        for (int b = 0; b < nbd; b++) {
            double k = px->K[s][b];
            if (!orc_approx_equal((float)k, 0.0f, 1.0e-6f) && k < Kmin)
                Kmin = k;
        }
        double t;
        if (H < 1.0 && Kmin < kmm) {
            t = 1.0 / (0.01 + Kmin) - 1.0 / (0.01 + kmn);
            Kerr += 100.0 * (t * t);
        } else if (H < 2.0 && Kmin < 0.5 * (1.5 * kmn + 0.5 * kmm)) {
            t = 1.0 / (0.01 + Kmin) - 1.0 / (0.01 + 0.5 * (1.5 * kmn + 0.5 * kmm));
            Kerr += 100.0 * (t * t);
        } else if (H < 3.0 && Kmin < 0.5 * (1.25 * kmn + 0.75 * kmm)) {
            t = 1.0 / (0.01 + Kmin) - 1.0 / (0.01 + 0.5 * (1.25 * kmn + 0.75 * kmm));
            Kerr += 100.0 * (t * t);
        } else if (H < 4.0 && Kmin < 0.5 * (1.5 * kmn + 0.5 * kmm)) {
            t = 1.0 / (0.01 + Kmin) - 1.0 / (0.01 + 0.5 * (1.5 * kmn + 0.5 * kmm));
            Kerr += 100.0 * (t * t);
        } else if (H < 5.0 && Kmin < 0.5 * (1.75 * kmn + 0.25 * kmm)) {
            t = 1.0 / (0.01 + Kmin) - 1.0 / (0.01 + 0.5 * (1.75 * kmn + 0.25 * kmm));
            Kerr += 100.0 * (t * t);
        }
    }
    if (Kmin > 0.7) {
        double t = 4.0 * (Kmin - 0.7);
        Kerr += 100.0 * (t * t);
    }
    px->K_error = Kerr;

    double e = 80.0 * (e_rrs * 1.0) + 15.0 * 0.0 + 10.0 * 0.0 + 15.0 * Kerr;
    e /= 80.0 + 15.0 + 10.0 + 15.0;
    px->model_error = e;
    return e;
}

/* ------------------------------------------------------------------------------------ */
/* Nelder-Mead simplex minimisation, as published by J. A. Nelder and R. Mead (Computer  */
/* Journal 7, 1965) in the form of R. O'Neill's Algorithm AS 47 (Applied Statistics 20,  */
/* 1971), with the changes the reference makes to it (model/asa047.c:10-502): a local-   */
/* minimum probe whose comparisons are on (long)(10 f), and, when a probe point is lower, */
/* a restart from that point with a simplex of relative size 1e-6.                       */
/*                                                                                        */
/* The arithmetic is the published algorithm's, in its order (the result must agree bit  */
/* for bit with the reference's compiled routine, tests/test_oracle_ref.py):              */
/*   centroid  c = (sum of every vertex - the worst) / n, coordinate by coordinate;       */
/*   trial     t = c + coef (s - c): the reflection of the worst vertex w is             */
/*             c + 1 (c - w); expansion and contractions take s = the reflected point or  */
/*             w with coef 2 or 1/2;                                                      */
/*   shrink    every vertex v <- (v + best) / 2, in index order;                          */
/*   converged when sum (f - mean f)^2 <= reqmin n, tested every konvge iterations.       */
/* ------------------------------------------------------------------------------------ */

/* (long)(10 v) as the reference's x86-64 build computes it (cvttsd2si: a NaN or a value
 * outside the range of long gives LONG_MIN). */
static long orc_tenths(double v)
{
    const double t = 10.0 * v;
    if (!(fabs(t) < 9.2233720368547758e18)) return (long)(-9223372036854775807L - 1L);
    return (long)t;
}

typedef struct {
    int n;                 /* coordinates */
    double *vx;            /* vertices, [n + 1][n] */
    double *fx;            /* their values, [n + 1] */
    double *mid;           /* centroid of every vertex but the worst */
    double *refl;          /* reflected point */
    double *trial;         /* expanded or contracted point */
    orc_objective_fn fn;
    void *ctx;
    int evals;
} orc_simplex;

static double *orc_vtx(orc_simplex *s, int k) { return s->vx + (size_t)k * (size_t)s->n; }

static double orc_eval(orc_simplex *s, double *pt)
{
    s->evals++;
    return s->fn(pt, s->ctx);
}

/* first index of the smallest (worst = 0) or the largest (worst = 1) value */
static int orc_extreme(const orc_simplex *s, int worst)
{
    int k = 0;
    for (int j = 1; j <= s->n; j++)
        if (worst ? (s->fx[k] < s->fx[j]) : (s->fx[j] < s->fx[k])) k = j;
    return k;
}

static void orc_put(orc_simplex *s, int k, const double *pt, double f)
{
    memcpy(orc_vtx(s, k), pt, sizeof(double) * (size_t)s->n);
    s->fx[k] = f;
}

/* dst = mid + coef (src - mid) */
static void orc_toward(const orc_simplex *s, double *dst, const double *src, double coef)
{
    for (int i = 0; i < s->n; i++) dst[i] = s->mid[i] + coef * (src[i] - s->mid[i]);
}

/* The simplex around `base`: the base itself is the last vertex (evaluated first), vertex j
 * moves coordinate j by scale * step[j]; `base` is returned unchanged. */
static void orc_build(orc_simplex *s, double *base, const double *step, double scale)
{
    const int n = s->n;
    memcpy(orc_vtx(s, n), base, sizeof(double) * (size_t)n);
    s->fx[n] = orc_eval(s, base);
    for (int j = 0; j < n; j++) {
        const double keep = base[j];
        base[j] = base[j] + step[j] * scale;
        memcpy(orc_vtx(s, j), base, sizeof(double) * (size_t)n);
        s->fx[j] = orc_eval(s, base);
        base[j] = keep;
    }
}

/* Sum of squared deviations of the vertex values from their mean. */
static double orc_spread(const orc_simplex *s)
{
    const int m = s->n + 1;
    double sum = 0.0;
    for (int j = 0; j < m; j++) sum = sum + s->fx[j];
    const double mean = sum / (double)m;
    double dev = 0.0;
    for (int j = 0; j < m; j++) dev = dev + (s->fx[j] - mean) * (s->fx[j] - mean);
    return dev;
}

/* One iteration after the worst vertex `hi` is known; *lo and *flo are the best vertex and
 * its value as last recorded (the value is kept apart from the vertex, as the published
 * algorithm keeps it: when every value is equal the worst vertex is also the best, and
 * replacing it leaves the recorded value standing).  Returns 1 when the iteration ended in
 * a shrink (which does not count towards the convergence interval). */
static int orc_iterate(orc_simplex *s, int hi, int *lo, double *flo)
{
    const int n = s->n;
    double *w = orc_vtx(s, hi);
    for (int i = 0; i < n; i++) {
        double z = 0.0;
        for (int j = 0; j <= n; j++) z = z + orc_vtx(s, j)[i];
        z = z - w[i];
        s->mid[i] = z / (double)n;
    }
    for (int i = 0; i < n; i++) s->refl[i] = s->mid[i] + 1.0 * (s->mid[i] - w[i]);
    const double fr = orc_eval(s, s->refl);

    if (fr < *flo) {
        /* better than the best: try twice as far */
        orc_toward(s, s->trial, s->refl, 2.0);
        const double fe = orc_eval(s, s->trial);
        if (fr < fe) orc_put(s, hi, s->refl, fr);
        else orc_put(s, hi, s->trial, fe);
    } else {
        int beaten = 0; /* vertices worse than the reflected point */
        for (int j = 0; j <= n; j++)
            if (fr < s->fx[j]) beaten++;
        if (beaten > 1) {
            orc_put(s, hi, s->refl, fr);
        } else if (beaten == 1) {
            /* contract on the reflection's side */
            orc_toward(s, s->trial, s->refl, 0.5);
            const double fc = orc_eval(s, s->trial);
            if (fc <= fr) orc_put(s, hi, s->trial, fc);
            else orc_put(s, hi, s->refl, fr);
        } else {
            /* contract on the worst vertex's side; if even that is worse, shrink */
            orc_toward(s, s->trial, w, 0.5);
            const double fc = orc_eval(s, s->trial);
            if (s->fx[hi] < fc) {
                for (int j = 0; j <= n; j++) {
                    double *v = orc_vtx(s, j);
                    const double *b = orc_vtx(s, *lo);
                    for (int i = 0; i < n; i++) v[i] = (v[i] + b[i]) * 0.5;
                    s->fx[j] = orc_eval(s, v);
                }
                *lo = orc_extreme(s, 0);
                *flo = s->fx[*lo];
                return 1;
            }
            orc_put(s, hi, s->trial, fc);
        }
    }
    if (s->fx[hi] < *flo) {
        *flo = s->fx[hi];
        *lo = hi;
    }
    return 0;
}

void orc_nelmin(orc_objective_fn fn, void *ctx, int n, double start[], double xmin[],
                double *ynewlo, double reqmin, double step[], int konvge, int kcount,
                int *icount, int *numres, int *ifault)
{
    if (reqmin <= 0.0 || n < 1 || konvge < 1) {
        *ifault = 1;
        return;
    }
    orc_simplex s;
    s.n = n;
    s.fn = fn;
    s.ctx = ctx;
    s.evals = 0;
    s.vx = (double *)malloc(sizeof(double) * (size_t)(n + 1) * (size_t)n);
    s.fx = (double *)malloc(sizeof(double) * (size_t)(n + 1));
    s.mid = (double *)malloc(sizeof(double) * (size_t)n);
    s.refl = (double *)malloc(sizeof(double) * (size_t)n);
    s.trial = (double *)malloc(sizeof(double) * (size_t)n);
    const double limit = reqmin * (double)n, probe = 1.0e-6;
    double scale = 1.0;
    int restarts = 0, status = 0;

    for (;;) {
        orc_build(&s, start, step, scale);
        int lo = orc_extreme(&s, 0);
        double flo = s.fx[lo];
        int until_check = konvge;
        /* iterate until converged or the evaluations are spent */
        while (s.evals < kcount) {
            const int hi = orc_extreme(&s, 1);
            *ynewlo = s.fx[hi];
            if (orc_iterate(&s, hi, &lo, &flo)) continue;
            if (--until_check > 0) continue;
            if (s.evals > kcount) continue;
            until_check = konvge;
            if (orc_spread(&s) <= limit) break;
        }
        /* the best vertex, and the probe one step * 1e-6 either side of it per coordinate */
        memcpy(xmin, orc_vtx(&s, lo), sizeof(double) * (size_t)n);
        *ynewlo = s.fx[lo];
        if (s.evals > kcount) {
            status = 2;
            break;
        }
        const long best10 = orc_tenths(s.fx[lo]);
        status = 0;
        for (int i = 0; i < n && status == 0; i++) {
            const double d = step[i] * probe;
            xmin[i] = xmin[i] + d;
            if (orc_tenths(orc_eval(&s, xmin)) < best10) {
                status = 2;
                break;
            }
            xmin[i] = xmin[i] - d - d;
            if (orc_tenths(orc_eval(&s, xmin)) < best10) {
                status = 2;
                break;
            }
            xmin[i] = xmin[i] + d;
        }
        if (status == 0) break;
        /* a lower point next to the best vertex: start again from it, tiny simplex */
        memcpy(start, xmin, sizeof(double) * (size_t)n);
        scale = probe;
        restarts++;
    }
    *icount = s.evals;
    *numres = restarts;
    *ifault = status;
    free(s.vx);
    free(s.fx);
    free(s.mid);
    free(s.refl);
    free(s.trial);
}

/* ------------------------------------------------------------------------------------ */
/* multi-start driver: model/samodel.c:2119-2427 (no empirical depth, no hot start, one  */
/* restart per depth start), and the derived outputs of :1992-2079                       */
/* ------------------------------------------------------------------------------------ */

void orc_default_cfg(orc_nm_cfg *cfg, int n_dates, double reqmin, int konvge, int kcount)
{
    static const double hs[8] = {40.0, 30.0, 20.0, 15.0, 10.0, 7.5, 2.5, 1.0};
    memset(cfg, 0, sizeof(*cfg));
    cfg->reqmin = reqmin;
    cfg->konvge = konvge;
    cfg->kcount = kcount;
    cfg->n_hstarts = 8;
    for (int i = 0; i < 8; i++)
        cfg->hstart[i] = hs[i];
    cfg->early_exit = 2.5 * (double)(float)n_dates;
}

/* Initial point and simplex steps for one depth start (model/samodel.c:2246-2354).  In mix
 * mode the mix parameter starts at 0.5 with step 0.5, as the reference starts each q at 1
 * with step 0.5. */
void orc_start_step(const orc_pixel *px, double h0, double *start, double *step)
{
    int nd = px->n_dates, off;
    start[0] = h0;
    step[0] = 1.25 * start[0];
    if (px->mode == ORC_MODE_MIX) {
        start[1] = 0.5;
        step[1] = 0.5;
        off = 2;
    } else {
        int nb = px->n_bottoms;
        double m490 = 0.0;
        for (int s = 0; s < nd; s++)
            m490 += px->r490[s];
        m490 /= (double)1 * nd;
        for (int k = 0; k < nb; k++) {
            start[1 + k] = 100.0 * 4.0 * m490;
            step[1 + k] = 1.5 * start[1 + k];
        }
        for (int k = 0; k < nb; k++) {
            start[1 + nb + k] = 1.0;
            step[1 + nb + k] = 0.5 * start[1 + nb + k];
        }
        off = 1 + 2 * nb;
    }
    for (int s = 0; s < nd; s++) {
        double m490 = px->r490[s] / (double)1;
        double m550 = px->r550[s] / (double)1;
        double m640 = px->r640[s] / (double)1;
        start[off + 3 * s] = 100.0 * 0.072 * pow(m490 / m550, -1.7);
        start[off + 3 * s + 1] = 1.5 * start[off + 3 * s];
        start[off + 3 * s + 2] = 100.0 * 30.0 * px->bands->aw640 * m640;
        step[off + 3 * s] = 2.0 * start[off + 3 * s];
        step[off + 3 * s + 1] = 2.0 * start[off + 3 * s + 1];
        step[off + 3 * s + 2] = 2.0 * start[off + 3 * s + 2];
    }
}

/* Outputs derived from the best vector after the final re-evaluation. */
static void orc_outputs(orc_pixel *px, const double *best, orc_result *res)
{
    int n = orc_n_params(px), nd = px->n_dates, nbd = px->bands->n_bands;
    res->n_params = n;
    for (int i = 0; i < n; i++)
        res->x[i] = fabs(best[i]);
    res->model_error = px->model_error;
    res->Rrs_error = px->Rrs_error;
    res->bottom_albedo = px->bottom_albedo;

    /* array_min_double2(K, ..., 0.0): smallest K not within 1e-4 of zero (common.c:979) */
    double kmin = 1.0e10;
    for (int s = 0; s < nd; s++)
        for (int b = 0; b < nbd; b++) {
            double k = px->K[s][b];
            if (!orc_approx_equal((float)k, 0.0f, 1.0e-4f) && k < kmin)
                kmin = k;
        }
    res->K_min = (kmin == 1.0e10) ? 0.0 : kmin;

    /* one region: origin weight sqrt(1), depth = H */
    double ow = sqrt(1.0);
    res->depth = (0.0 + ow * res->x[0]) / (ow + 1.0 - 1.0);

    for (int k = 0; k < ORC_NBOTTOM_TYPES; k++)
        res->B_percent[k] = 0.0;
    double qv[ORC_MAXBOT], qsum = 0.0;
    int nbot;
    if (px->mode == ORC_MODE_MIX) {
        nbot = 2;
        qv[0] = fabs(best[1]);
        qv[1] = fabs(1.0 - best[1]);
    } else {
        nbot = px->n_bottoms;
        for (int k = 0; k < nbot; k++)
            qv[k] = res->x[1 + nbot + k];
    }
    for (int k = 0; k < nbot; k++)
        qsum += qv[k];
    for (int k = 0; k < nbot; k++)
        res->B_percent[px->bottom_idx[k]] = 100.0 * qv[k] / qsum;
    double largest = 0.0;
    res->bottom_type = 0.0;
    for (int k = 0; k < nbot; k++)
        if (res->B_percent[px->bottom_idx[k]] > largest) {
            largest = res->B_percent[px->bottom_idx[k]];
            res->bottom_type = 1.0 + px->bottom_idx[k];
        }

    double iod = 0.0, nobs = 0.0;
    for (int s = 0; s < nd; s++)
        for (int b = 0; b < nbd; b++) {
            iod += px->rrs_bot[s][b] / px->rrs_mod[s][b];
            nobs += 1.0;
        }
    res->iod = 100.0 * iod / nobs;

    int off = (px->mode == ORC_MODE_MIX) ? 2 : 1 + 2 * px->n_bottoms;
    for (int s = 0; s < nd; s++) {
        res->P[s] = 0.01 * res->x[off + 3 * s];
        res->G[s] = 0.01 * res->x[off + 3 * s + 1];
        res->X[s] = 0.01 * res->x[off + 3 * s + 2];
    }
}

static void orc_run_starts(orc_pixel *px, const orc_nm_cfg *cfg, orc_nelmin_fn nm,
                           const double *seed_start, const double *seed_step,
                           orc_result *res)
{
    int n = orc_n_params(px);
    double start[ORC_MAXN], step[ORC_MAXN], xmin[ORC_MAXN], best[ORC_MAXN];
    double lowest = 1.0e4;
    long evals0 = px->n_evals;
    if (!nm)
        nm = orc_nelmin;
    memset(res, 0, sizeof(*res));
    memset(best, 0, sizeof(best));
    int accepted = 0;

    if (seed_start && seed_start[0] == seed_start[0]) { /* a NaN seed: no seed start */
        double err;
        int icount = 0, numres = 0, ifault = 0;
        memcpy(start, seed_start, sizeof(double) * (size_t)n);
        memcpy(step, seed_step, sizeof(double) * (size_t)n);
        nm(orc_objective, px, n, start, xmin, &err, cfg->reqmin, step, cfg->konvge,
           cfg->kcount, &icount, &numres, &ifault);
        res->n_starts++;
        if (err < lowest) {
            lowest = err;
            memcpy(best, xmin, sizeof(double) * (size_t)n);
            res->n_iterations = icount;
            res->converged = (ifault == 0);
            if (lowest < cfg->early_exit)
                accepted = 1;
        }
    }
    for (int k = 0; k < cfg->n_hstarts && !accepted; k++) {
        double err;
        int icount = 0, numres = 0, ifault = 0;
        orc_start_step(px, cfg->hstart[k], start, step);
        err = orc_objective(start, px); /* the reference evaluates the start first */
        nm(orc_objective, px, n, start, xmin, &err, cfg->reqmin, step, cfg->konvge,
           cfg->kcount, &icount, &numres, &ifault);
        res->n_starts++;
        if (err < lowest) {
            lowest = err;
            memcpy(best, xmin, sizeof(double) * (size_t)n);
            res->n_iterations = icount;
            res->converged = (ifault == 0);
            if (lowest < cfg->early_exit)
                break;
        }
    }
    orc_objective(best, px); /* recompute every model-derived quantity at the optimum */
    orc_outputs(px, best, res);
    res->n_evals = px->n_evals - evals0;
}

void orc_eval_outputs(orc_pixel *px, const double *x, orc_result *res)
{
    double xv[ORC_MAXN];
    memset(res, 0, sizeof(*res));
    memcpy(xv, x, sizeof(double) * (size_t)orc_n_params(px));
    orc_objective(xv, px);
    orc_outputs(px, xv, res);
}

void orc_optimise(orc_pixel *px, const orc_nm_cfg *cfg, orc_nelmin_fn nm, orc_result *res)
{
    orc_run_starts(px, cfg, nm, NULL, NULL, res);
}

void orc_optimise_seeded(orc_pixel *px, const orc_nm_cfg *cfg, orc_nelmin_fn nm,
                         const double *seed_start, const double *seed_step,
                         orc_result *res)
{
    orc_run_starts(px, cfg, nm, seed_start, seed_step, res);
}

/* ------------------------------------------------------------------------------------ */
/* scene driver: the reference's dynamic-schedule OpenMP pixel loop (model/samodel.c:    */
/* 901-1128) without its look-up table and hot start                                     */
/* ------------------------------------------------------------------------------------ */

long orc_invert_scene(const orc_scene_cfg *cfg, long n_pix, const float *rrs,
                      const float *sec_sun, const uint8_t *valid, int threads,
                      float *depth, float *rrs_error, float *model_error, int *status,
                      int *n_evals, double *x_out, int x_stride)
{
    orc_bands bands;
    orc_band_tables(cfg->wl, cfg->n_bands, &bands);
    const int nb = cfg->n_bands, nd = cfg->n_dates;
    long done = 0;
#ifdef _OPENMP
    if (threads > 0)
        omp_set_num_threads(threads);
#else
    (void)threads;
#endif

#pragma omp parallel for schedule(dynamic, 16) reduction(+ : done)
    for (long p = 0; p < n_pix; p++) {
        double r[ORC_MAXD * ORC_MAXB], ss[ORC_MAXD], sv[ORC_MAXD];
        int k = 0, bad = 0;
        for (int s = 0; s < nd; s++) {
            if (valid && !valid[(long)s * n_pix + p])
                continue;
            for (int b = 0; b < nb; b++) {
                float v = rrs[((long)s * nb + b) * n_pix + p];
                if (orc_approx_equal(v, cfg->nodata, 1.0e-6f) || v < 0.0f)
                    bad = 1;
                r[k * nb + b] = v;
            }
            ss[k] = sec_sun ? (double)sec_sun[(long)s * n_pix + p] : 1.0;
            sv[k] = 1.0;
            k++;
        }
        depth[p] = 0.0f;
        rrs_error[p] = 0.0f;
        if (model_error)
            model_error[p] = 0.0f;
        if (n_evals)
            n_evals[p] = 0;
        if (bad || k == 0) {
            status[p] = -1;
            continue;
        }
        orc_pixel px;
        orc_nm_cfg nm = cfg->nm;
        nm.early_exit = 2.5 * (double)(float)k;
        orc_pixel_init(&px, cfg->mode, cfg->n_bottoms, cfg->bottom_idx, k, &bands, ss, sv, r);
        orc_result res;
        orc_optimise(&px, &nm, NULL, &res);
        depth[p] = (float)res.depth;
        rrs_error[p] = (float)res.Rrs_error;
        if (model_error)
            model_error[p] = (float)res.model_error;
        status[p] = res.converged ? 0 : 2;
        if (n_evals)
            n_evals[p] = (int)res.n_evals;
        if (x_out)
            for (int i = 0; i < res.n_params && i < x_stride; i++)
                x_out[p * x_stride + i] = res.x[i];
        done++;
    }
    return done;
}
