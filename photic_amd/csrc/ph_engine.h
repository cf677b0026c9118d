// photic_amd -- per-pixel semi-analytical inversion on gfx950.
//
// The lane engine: photic's forward model (model/samodel.c:2846-2949), its misfit
// (model/samodel.c:2432-2759, one region) and its modified AS 47 Nelder-Mead
// (model/asa047.c:10-502) driven by the multi-start of model/samodel.c:2119-2427, recast as a
// per-lane state machine in which every pass of the kernel's loop makes exactly one
// objective evaluation per lane.  The evaluation -- the expensive, branch-free part -- then
// runs in lock-step across the 64 lanes of a wavefront, whatever step of the simplex each
// lane is at; the short bookkeeping between evaluations is the only divergent code.
//
// Everything here is __host__ __device__ so the FP64 instantiation can also be run on the
// CPU by the test suite (tests/hostsim), where it must agree bit for bit with the FP64
// oracle; the library itself only ever launches it on the GPU.
//
// Arithmetic in the FP64 path follows the reference's order of operations; the file is
// compiled with -ffp-contract=off so no multiply-add is fused behind our back.  The FP32
// path fuses deliberately (explicit fmaf) and uses the hardware transcendentals.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ph_crmath.h"
#include "ph_launch.h"

// Transcendentals of the exact (FP64) path: correctly rounded, ph_crmath.h.  The host test
// build can substitute the C library's (PH_HOST_LIBM) to check the state machine against the
// oracle bit for bit with identical arithmetic on both sides.
#if defined(PH_HOST_LIBM) && !defined(__HIP_DEVICE_COMPILE__)
#define PH_EXP exp
#define PH_LOG log
#define PH_POW pow
#define PH_EXP_NB(x, slow) exp(x)
#define PH_LOG_NB(x, slow) log(x)
#else
#define PH_EXP ph_exp_cr
#define PH_LOG ph_log_cr
#define PH_POW ph_pow_cr
#define PH_EXP_NB ph_exp_cr_nb // the branch-free forms: *slow set where they do not apply
#define PH_LOG_NB ph_log_cr_nb
#endif

// Every engine function is inlined into the kernel: the lane state must stay in registers,
// and a real call would force it through scratch memory.
#define PH_INLINE __host__ __device__ __attribute__((always_inline)) inline

#include <utility>

// Source-level unrolling.  Loops over the lane's register arrays are written as
// ph_for<N>([&](auto i) {...}): every index is a constant from the first optimisation pass on,
// so the arrays become registers before instruction combining could turn a
// "c ? s.a[i] : s.b[i]" into a load through a selected pointer (which would pin the whole
// lane state in scratch memory).
template <typename F, int... I>
PH_INLINE void ph_for_seq(F&& f, std::integer_sequence<int, I...>)
{
    (f(std::integral_constant<int, I>{}), ...);
}
template <int M, typename F>
PH_INLINE void ph_for(F&& f)
{
    ph_for_seq(f, std::make_integer_sequence<int, M>{});
}
#define PH_BODY __attribute__((always_inline))

// Side quantities of one evaluation that the outputs need (samodel_optimise reads the
// same fields of model_data after its final re-evaluation).
template <int ND, int NB>
struct PhSide {
    double K[ND][NB];
    double rrs_bot[ND][NB];
    double rrs_mod[ND][NB];
    double Rrs_error, K_error, model_error, albedo;
};

// Per-pixel constants held by a lane while it inverts the pixel.
template <typename FT, int ND, int NB>
struct PhPixel {
    double rrs[ND][NB];       // measured Rrs, float input widened
    double pw[ND][NB];        // (440/lambda)^Y, Y from the measured 440/490 ratio
    double ss[ND], sv[ND];    // sec(sun), sec(view), subsurface
    double tot;               // sum of measured Rrs in (date, band) order
    double p0[ND];            // every depth start's P: 100 * 0.072 (Rrs490 / Rrs550)^-1.7
};

// ---------------------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------------------

// The reference's approx_equal(float, float, float) (model/common.c:392-394).
PH_INLINE bool ph_approx_equal(float a, float b, float eps)
{
    float d = a - b;
    double fa = fabs((double)a), fb = fabs((double)b);
    return fabs((double)d) <= ((fa < fb ? fb : fa) * (double)eps);
}

// Rrs at one of the five reference wavelengths (k: 0 = 440, 1 = 490, 2 = 550, 3 = 640,
// 4 = 750 nm) by the reference's linear interpolation / extrapolation (model/common.c:
// 298-335), with the segment and weights resolved on the host.
// c ? a : b on values.  With two lvalue operands C++'s conditional is itself an lvalue and
// would be compiled as a select between addresses, which keeps the arrays involved out of
// registers; taking the operands by value avoids that.
PH_INLINE double ph_pick(bool c, double a, double b) { return c ? a : b; }

PH_INLINE double ph_rrs_at_ptr(const PhConst& c, const double* r, int k)
{
    return r[c.ip_i0[k]] * c.ip_w0[k] + r[c.ip_i1[k]] * c.ip_w1[k];
}

template <int NB>
PH_INLINE double ph_rrs_at(const PhConst& c, const double (&r)[NB], int k)
{
    double y0 = r[0], y1 = r[0];
#pragma unroll
    for (int b = 1; b < NB; b++) {
        y0 = ph_pick(b == c.ip_i0[k], r[b], y0);
        y1 = ph_pick(b == c.ip_i1[k], r[b], y1);
    }
    return y0 * c.ip_w0[k] + y1 * c.ip_w1[k];
}

// (long)(10 * v) as x86-64 evaluates it (cvttsd2si: NaN and out-of-range give INT64_MIN),
// kept in a double -- exact for every value the misfit can take.
PH_INLINE double ph_trunc10(double v)
{
    double t = 10.0 * v;
    if (!(fabs(t) < 9.2233720368547758e18)) return -9.2233720368547758e18;
    return trunc(t);
}

// Dynamic-index access into register arrays without scratch: unrolled selects.
template <int M>
PH_INLINE double ph_sel(const double (&a)[M], int j)
{
    double v = a[0];
    ph_for<M>([&](auto k) PH_BODY { v = ph_pick(j == k, a[k], v); });
    return v;
}

// x / d for a constant d with inv = RN(1/d), in three operations: by Markstein's theorem the
// correction fma(x - q d, inv, q) of q = RN(x inv) is the correctly rounded quotient of a finite
// nonzero x (tested on 1.75e9 quotients, a quarter of them constructed next to rounding
// midpoints, for the divisors used here).  Zeros and non-finite x keep q, which is then exact.
PH_INLINE double ph_div_const(double x, double d, double inv)
{
    double q = x * inv;
    double r = fma(-q, d, x);
    double q1 = fma(r, inv, q);
    return (q1 == q1 && q1 != 0.0) ? q1 : q;
}

// The same for a finite x >= +0 (a bottom reflectance), where the correction is always exact.
PH_INLINE double ph_div_const_nonneg(double x, double d, double inv)
{
    double q = x * inv;
    return fma(fma(-q, d, x), inv, q);
}

// ---------------------------------------------------------------------------------------
// forward model, one date and band
// ---------------------------------------------------------------------------------------

// FP64: model/samodel.c:2886-2944 in the reference's operation order, split at its two
// exponentials: the head gives K, r_dp and the two exponents, the tail the rest from
// exp(exponent).  lnP = log|P| is given (the same value in every band of a date, so a caller
// evaluating several bands forms it once).
PH_INLINE void ph_band64_head(double P, double lnP, double G, double X, double H, double a0,
                              double a1, double aw, double bbw, double eg, double pw, double ss,
                              double sv, double* K_out, double* dp_out, double* argC,
                              double* argB)
{
    double a_phi = (a0 + a1 * lnP) * fabs(P);
    double a_g = fabs(G) * eg;
    double a = aw + a_phi + a_g;
    double b_p = X * pw;
    double bb = bbw + b_p;
    double u = bb / (a + bb);
    double K = a + bb;
    if (K < 0.0) K = 0.0;
    if (K > 2.5) K = 2.5;
    double rrs_dp = (0.084 + 0.170 * u) * u;
    double DuC = 1.03 * sqrt(1.0 + 2.4 * u);
    double DuB = 1.04 * sqrt(1.0 + 5.4 * u);
    double M = ss + DuC * sv;
    *argC = -M * K * H;
    M = ss + DuB * sv;
    *argB = -M * K * H;
    *K_out = K;
    *dp_out = rrs_dp;
}

// eC = exp(argC), eB = exp(argB) of the head
PH_INLINE double ph_band64_tail(double rho, double rrs_dp, double eC, double eB,
                                double* rB_out, double* r_out)
{
    double rC = rrs_dp * (1.0 - eC);
    const double pi = 3.141592653589793;
    double rB = ph_div_const_nonneg(rho, pi, 1.0 / pi) * eB;
    double r = rC + rB;
    *rB_out = rB;
    *r_out = r;
    return 0.5 * r / (1.0 - 1.5 * r) + 0.0;
}

PH_INLINE double ph_band64_l(double P, double lnP, double G, double X, double rho, double H,
                             double a0, double a1, double aw, double bbw, double eg, double pw,
                             double ss, double sv, double* K_out, double* rB_out, double* r_out)
{
    double dp, argC, argB;
    ph_band64_head(P, lnP, G, X, H, a0, a1, aw, bbw, eg, pw, ss, sv, K_out, &dp, &argC, &argB);
    return ph_band64_tail(rho, dp, PH_EXP(argC), PH_EXP(argB), rB_out, r_out);
}

PH_INLINE double ph_band64(double P, double G, double X, double rho, double H, double a0,
                           double a1, double aw, double bbw, double eg, double pw, double ss,
                           double sv, double* K_out, double* rB_out, double* r_out)
{
    return ph_band64_l(P, PH_LOG(fabs(P)), G, X, rho, H, a0, a1, aw, bbw, eg, pw, ss, sv, K_out,
                       rB_out, r_out);
}

// ph_band64 with exp and log in their branch-free forms and one branch at the end, taken only
// when an argument falls outside them: the same value, with the table reads of the log and of
// both exponentials in one straight line.
PH_INLINE double ph_band64_nb(double P, double G, double X, double rho, double H, double a0,
                              double a1, double aw, double bbw, double eg, double pw, double ss,
                              double sv, double* K_out, double* rB_out, double* r_out)
{
    bool slow = false;
    double lnP = PH_LOG_NB(fabs(P), &slow);
    double dp, argC, argB;
    ph_band64_head(P, lnP, G, X, H, a0, a1, aw, bbw, eg, pw, ss, sv, K_out, &dp, &argC, &argB);
    double eC = PH_EXP_NB(argC, &slow);
    double eB = PH_EXP_NB(argB, &slow);
    double m = ph_band64_tail(rho, dp, eC, eB, rB_out, r_out);
    if (slow)
        m = ph_band64(P, G, X, rho, H, a0, a1, aw, bbw, eg, pw, ss, sv, K_out, rB_out, r_out);
    return m;
}

// Fast path: the same model with the arithmetic in FP64 (full rate on gfx950) and the
// transcendentals from the single-precision hardware units, each made accurate to a few
// 1e-8 relative: exp with an exact FP64 range reduction (2^n * 2^f, |f| <= 1/2), and the
// double-precision reciprocal / reciprocal square root refined by one Newton step.
__device__ PH_INLINE double ph_rcp_fast(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

__device__ PH_INLINE double ph_sqrt_fast(double x)
{
    double r = __builtin_amdgcn_rsq(x); // ~ 1/sqrt(x)
    double s = x * r;
    double h = 0.5 * r;
    double e = fma(-s, h, 0.5);
    return fma(s, e, s);
}

__device__ PH_INLINE double ph_exp_fast(double x)
{
    x = fmax(x, -700.0);
    double y = x * 1.4426950408889634;
    double n = rint(y);
    float f = (float)(y - n);
    return __builtin_amdgcn_ldexp((double)__builtin_amdgcn_exp2f(f), (int)n);
}

__device__ PH_INLINE double ph_band_fast(double P, double lnP, double G, double X,
                                         double rho_pi, double H, double a0, double a1,
                                         double aw, double bbw, double eg, double pw,
                                         double ss, double sv, double* K_out, double* rB_out,
                                         double* r_out)
{
    double a = fma(G, eg, fma(fma(a1, lnP, a0), P, aw));
    double bb = fma(X, pw, bbw);
    double apb = a + bb;
    double u = bb * ph_rcp_fast(apb);
    double K = fmin(fmax(apb, 0.0), 2.5);
    double rrs_dp = fma(0.170, u, 0.084) * u;
    double DuC = 1.03 * ph_sqrt_fast(fma(2.4, u, 1.0));
    double DuB = 1.04 * ph_sqrt_fast(fma(5.4, u, 1.0));
    double KH = K * H;
    double eC = ph_exp_fast(-fma(DuC, sv, ss) * KH);
    double eB = ph_exp_fast(-fma(DuB, sv, ss) * KH);
    double rB = rho_pi * eB;
    double r = fma(rrs_dp, 1.0 - eC, rB);
    *K_out = K;
    *rB_out = rB;
    *r_out = r;
    return 0.5 * r * ph_rcp_fast(fma(-1.5, r, 1.0));
}

// ---------------------------------------------------------------------------------------
// objective: model/samodel.c:2432-2759 with one region
// ---------------------------------------------------------------------------------------

// A kernel's only argument T, read through the kernel-argument segment pointer behind an
// empty asm the compiler cannot see through: loads through it stay where they are used
// (scalar loads from the segment) instead of being hoisted out of the persistent loop into
// scalar registers that then spill into vector lanes.  (The host pass of a HIP build parses
// kernel bodies without running them; there it is never called.)
template <typename T>
__device__ __attribute__((always_inline)) inline const T& ph_kernarg()
{
#if defined(__HIP_DEVICE_COMPILE__)
    typedef const __attribute__((address_space(4))) T* P4;
    P4 p = (P4)__builtin_amdgcn_kernarg_segment_ptr();
    asm volatile("" : "+s"(p));
    return *(const T*)p;
#else
    __builtin_unreachable();
#endif
}

// The launch constants of a kernel whose first argument starts with its PhConst.
__device__ __attribute__((always_inline)) inline const PhConst& ph_kernarg_const()
{
    return ph_kernarg<PhConst>();
}

template <int NBOT>
struct PhLayout {
    static constexpr int nbot = NBOT == 0 ? 2 : NBOT;
    static constexpr int off = NBOT == 0 ? 2 : 1 + 2 * NBOT;
    template <int ND>
    static constexpr int n() { return off + 3 * ND; }
};

// KARG: c is the kernel's own launch constants (ph_kernarg_const); each band then reads its
// coefficients afresh from the kernel-argument segment.
template <typename FT, int ND, int NB, int NBOT, bool SIDE, bool KARG = false>
PH_INLINE double ph_objective(const double* x, const PhConst& c,
                                               const PhPixel<FT, ND, NB>& px,
                                               PhSide<ND, NB>* side)
{
    using L = PhLayout<NBOT>;
    constexpr int nbot = L::nbot;
    double B[nbot], q[nbot], qsum = 0.0;
    if (NBOT == 0) {
        B[0] = 0.01 * fabs(100.0);
        B[nbot - 1] = 0.01 * fabs(100.0);
        q[0] = fabs(fabs(x[1]));
        q[nbot - 1] = fabs(fabs(1.0 - x[1]));
    } else {
#pragma unroll
        for (int k = 0; k < nbot; k++) {
            B[k] = 0.01 * fabs(x[1 + k]);
            q[k] = fabs(x[1 + nbot + k]);
        }
    }
#pragma unroll
    for (int k = 0; k < nbot; k++) qsum += q[k];
#pragma unroll
    for (int k = 0; k < nbot; k++) q[k] /= qsum;

    // bottom reflectance per band: rho = sum_i (q_i B_i) R_i(lambda)
    double qb[nbot];
    double alb = 0.0;
#pragma unroll
    for (int k = 0; k < nbot; k++) {
        qb[k] = q[k] * B[k];
        alb += qb[k];
    }
    const double H = fabs(x[0]);
    double err = 0.0;
    double Kd[ND][NB];
#pragma unroll
    for (int s = 0; s < ND; s++) {
        double P = 0.01 * fabs(x[L::off + 3 * s]);
        double G = 0.01 * fabs(x[L::off + 3 * s + 1]);
        double X = 0.01 * fabs(x[L::off + 3 * s + 2]);
        if constexpr (sizeof(FT) == 8) {
            // The date's bands in three passes -- every band's head, the exponentials, every
            // band's tail -- with exp and log in their branch-free forms, so that the table
            // reads of all the date's exponentials can be in flight together.  Each value is
            // computed by the same operations as in ph_band64_l; arguments outside the
            // branch-free range (P zero or not finite; subnormal, overflowing or NaN exps)
            // take the full functions.
            bool slow = false;
            double lnP = PH_LOG_NB(fabs(P), &slow); // a_phi's log, shared by the date's bands
            if (slow) lnP = PH_LOG(fabs(P));
            slow = false;
            double rho[NB], dp[NB], aC[NB], aB[NB], eC[NB], eB[NB];
#pragma unroll
            for (int b = 0; b < NB; b++) {
#if defined(__HIP_DEVICE_COMPILE__)
                const PhConst& cb = KARG ? ph_kernarg_const() : c;
#else
                const PhConst& cb = c;
#endif
                rho[b] = 0.0;
#pragma unroll
                for (int k = 0; k < nbot; k++) {
                    int t = NBOT == 0 ? k : cb.bottom_idx[k];
                    rho[b] += qb[k] * cb.bot[t][b];
                }
                ph_band64_head(P, lnP, G, X, H, cb.a0[b], cb.a1[b], cb.aw[b], cb.bbw[b],
                               cb.eg[b], px.pw[s][b], px.ss[s], px.sv[s], &Kd[s][b], &dp[b],
                               &aC[b], &aB[b]);
            }
#pragma unroll
            for (int b = 0; b < NB; b++) {
                eC[b] = PH_EXP_NB(aC[b], &slow);
                eB[b] = PH_EXP_NB(aB[b], &slow);
            }
            if (slow) {
#pragma unroll
                for (int b = 0; b < NB; b++) {
                    eC[b] = PH_EXP(aC[b]);
                    eB[b] = PH_EXP(aB[b]);
                }
            }
#pragma unroll
            for (int b = 0; b < NB; b++) {
                double rB, r;
                double m = ph_band64_tail(rho[b], dp[b], eC[b], eB[b], &rB, &r);
                double d = m - px.rrs[s][b];
                err += d * d;
                if (SIDE) {
                    side->K[s][b] = Kd[s][b];
                    side->rrs_bot[s][b] = rB;
                    side->rrs_mod[s][b] = r;
                }
            }
        } else {
            const double lnP = (double)__builtin_amdgcn_logf((float)fmax(P, 1.0e-30)) *
                               0.6931471805599453;
#pragma unroll
            for (int b = 0; b < NB; b++) {
                double rho = 0.0;
#pragma unroll
                for (int k = 0; k < nbot; k++) {
                    int t = NBOT == 0 ? k : c.bottom_idx[k];
                    rho = fma(qb[k], c.bot[t][b], rho);
                }
                double K, rB, r;
                double m = ph_band_fast(P, lnP, G, X, rho * 0.3183098861837907, H, c.a0[b],
                                        c.a1[b], c.aw[b], c.bbw[b], c.eg[b], px.pw[s][b],
                                        px.ss[s], px.sv[s], &K, &rB, &r);
                double d = m - px.rrs[s][b];
                err = fma(d, d, err);
                Kd[s][b] = K;
                if (SIDE) {
                    side->K[s][b] = K;
                    side->rrs_bot[s][b] = rB;
                    side->rrs_mod[s][b] = r;
                }
            }
        }
    }
    const double nt = (double)(ND * NB);
    double mean = px.tot / nt;
    double e_rrs = 100.0 * sqrt(err / nt) / mean;

    // K penalty (model/samodel.c:2696-2732)
    const double kmm = 0.275, kmn = 0.185;
    double Kerr = 0.0, Kmin = 1.0e4;
#pragma unroll
    for (int s = 0; s < ND; s++) {
        Kmin = 1.0e4;
#pragma unroll
        for (int b = 0; b < NB; b++) {
            double k = Kd[s][b];
            if (!ph_approx_equal((float)k, 0.0f, 1.0e-6f) && k < Kmin) Kmin = k;
        }
        double thr = -1.0;
        if (H < 1.0 && Kmin < kmm) thr = 0.01 + kmn;
        else if (H < 2.0 && Kmin < 0.5 * (1.5 * kmn + 0.5 * kmm)) thr = 0.01 + 0.5 * (1.5 * kmn + 0.5 * kmm);
        else if (H < 3.0 && Kmin < 0.5 * (1.25 * kmn + 0.75 * kmm)) thr = 0.01 + 0.5 * (1.25 * kmn + 0.75 * kmm);
        else if (H < 4.0 && Kmin < 0.5 * (1.5 * kmn + 0.5 * kmm)) thr = 0.01 + 0.5 * (1.5 * kmn + 0.5 * kmm);
        else if (H < 5.0 && Kmin < 0.5 * (1.75 * kmn + 0.25 * kmm)) thr = 0.01 + 0.5 * (1.75 * kmn + 0.25 * kmm);
        if (thr > 0.0) {
            double t = 1.0 / (0.01 + Kmin) - 1.0 / thr;
            Kerr += 100.0 * (t * t);
        }
    }
    if (Kmin > 0.7) {
        double t = 4.0 * (Kmin - 0.7);
        Kerr += 100.0 * (t * t);
    }
    double e = 80.0 * e_rrs + 15.0 * Kerr;
    e /= 120.0;
    if (SIDE) {
        side->Rrs_error = e_rrs;
        side->K_error = Kerr;
        side->model_error = e;
        side->albedo = alb;
    }
    return e;
}

// ---------------------------------------------------------------------------------------
// pixel set-up: model/samodel.c:1785-1799 (Rrs at 440/490/550/640 nm), :2898-2903 (Y)
// ---------------------------------------------------------------------------------------

// rrs_in[s][b]; returns false when the pixel is nodata on any band of any date
// (model/samodel.c:933-947: within 1e-6 of the nodata value, or negative).
template <typename FT, int ND, int NB>
PH_INLINE bool ph_pixel_init(PhPixel<FT, ND, NB>& px, const PhConst& c,
                                              const float (&rrs_in)[ND][NB], const double* ss,
                                              const double* sv)
{
    bool ok = true;
    double tot = 0.0;
#pragma unroll
    for (int s = 0; s < ND; s++) {
        double r[NB];
#pragma unroll
        for (int b = 0; b < NB; b++) {
            float v = rrs_in[s][b];
            if (!c.prescreened && (ph_approx_equal(v, (float)c.nodata, 1.0e-6f) || v < 0.0f))
                ok = false;
            r[b] = (double)v;
            px.rrs[s][b] = r[b];
            tot += r[b];
        }
        double r440 = ph_rrs_at(c, r, 0);
        double r490 = ph_rrs_at(c, r, 1);
        if (r440 < 0.0) r440 = 0.0001;
        double chi = r440 / r490;
        double Y = 3.44 * (1.0 - 3.17 * PH_EXP(-2.01 * chi));
        if (Y < 0.0) Y = 0.0;
        if (Y > 2.5) Y = 2.5;
#pragma unroll
        for (int b = 0; b < NB; b++) px.pw[s][b] = PH_POW(440.0 / c.wl[b], Y);
        // the start rule's P depends on the pixel alone (model/samodel.c:2342): formed
        // once here rather than at each of the pixel's depth starts
        px.p0[s] = 100.0 * 0.072 * PH_POW(r490 / (double)1 / (ph_rrs_at(c, r, 2) / (double)1),
                                          -1.7);
        px.ss[s] = ss[s];
        px.sv[s] = sv[s];
    }
    px.tot = tot;
    return ok;
}

// ---------------------------------------------------------------------------------------
// the lane state machine
// ---------------------------------------------------------------------------------------

enum PhPhase {
    PH_IDLE = 0,
    PH_START,        // set up the next start (depth start, or the seed)      [no eval]
    PH_BUILD,        // evaluating simplex vertex `sub` (sub == N: the start point)
    PH_REFLECT,
    PH_EXPAND,
    PH_CONTRACT_IN,  // contraction on the worst vertex's side
    PH_CONTRACT_OUT, // contraction on the reflection's side
    PH_SHRINK,       // evaluating shrunk vertex `sub`
    PH_FACT_PLUS,    // local-minimum test, coordinate `sub`, + side
    PH_FACT_MINUS,
    PH_FINAL,        // final re-evaluation at the best point
    PH_DONE          // outputs ready                                          [no eval]
};

// The lane's large arrays -- simplex vertices, steps, best point -- live in LDS (on the GPU)
// or in an ordinary array (host test build), interleaved across the lanes of a block so that
// a wavefront's accesses to one element are contiguous and free of bank conflicts.  The
// element of lane t at slot e is base[e * stride].
template <int N>
struct PhLaneMem {
    static constexpr int n_slots = (N + 1) * N + 2 * N;
    double* base;
    int stride;
    PH_INLINE double p(int k, int i) const { return base[(k * N + i) * stride]; }
    PH_INLINE void set_p(int k, int i, double v) const { base[(k * N + i) * stride] = v; }
    PH_INLINE double step(int i) const { return base[((N + 1) * N + i) * stride]; }
    PH_INLINE void set_step(int i, double v) const { base[((N + 1) * N + i) * stride] = v; }
    PH_INLINE double best(int i) const { return base[((N + 1) * N + N + i) * stride]; }
    PH_INLINE void set_best(int i, double v) const { base[((N + 1) * N + N + i) * stride] = v; }
};

template <int N>
struct PhNM {
    PhLaneMem<N> m;
    double y[N + 1];     // vertex values
    double xs[N];        // start point while building; reflected point while iterating
    double xb[N];        // centroid while iterating; probe point in the local-minimum test
    double xe[N];        // point being evaluated
    double ylo, ystar, ynewlo, del, lowest, yr;
    int ilo, ihi, icount, jcount, numres, ifault, phase, sub, kstart, nstart;
    int best_icount, best_conv, evals;
};

// Start point and steps of depth start k (k < 0: the seed), model/samodel.c:2246-2354.
template <typename FT, int ND, int NB, int NBOT>
PH_INLINE void ph_start_step(PhNM<PhLayout<NBOT>::template n<ND>()>& S, const PhConst& c,
                             const PhPixel<FT, ND, NB>& px, int k, const double* seed,
                             double h_override = -1.0)
{
    using L = PhLayout<NBOT>;
    constexpr int N = L::template n<ND>();
    if (k < 0) {
        ph_for<N>([&](auto i) PH_BODY {
            constexpr int cls = i == 0 ? 0 : (i < L::off ? 1 : 2 + (i - L::off) % 3);
            double st = fabs(seed[i]) * c.seed_rel_step[cls];
            if (st < c.seed_min_step[cls]) st = c.seed_min_step[cls];
            S.xs[i] = seed[i];
            S.m.set_step(i, st);
        });
        return;
    }
    double h0 = c.hstart[0];
    ph_for<PH_MAXH - 1>([&](auto m0) PH_BODY {
        constexpr int m = m0 + 1;
        h0 = ph_pick(k == m, c.hstart[m], h0);
    });
    if (h_override >= 0.0) h0 = h_override; // the pixel's own depth start
    S.xs[0] = h0;
    S.m.set_step(0, 1.25 * S.xs[0]);
    if (NBOT == 0) {
        S.xs[1] = 0.5;
        S.m.set_step(1, 0.5);
    } else {
        double m490 = 0.0;
        ph_for<ND>([&](auto s) PH_BODY { m490 += ph_rrs_at(c, px.rrs[s], 1); });
        m490 /= (double)1 * ND;
        ph_for<L::nbot>([&](auto b) PH_BODY {
            S.xs[1 + b] = 100.0 * 4.0 * m490;
            S.m.set_step(1 + b, 1.5 * S.xs[1 + b]);
        });
        ph_for<L::nbot>([&](auto b) PH_BODY {
            S.xs[1 + L::nbot + b] = 1.0;
            S.m.set_step(1 + L::nbot + b, 0.5 * S.xs[1 + L::nbot + b]);
        });
    }
    ph_for<ND>([&](auto s) PH_BODY {
        constexpr int o = L::off + 3 * s;
        double m640 = ph_rrs_at(c, px.rrs[s], 3) / (double)1;
        S.xs[o] = px.p0[s];
        S.xs[o + 1] = 1.5 * S.xs[o];
        S.xs[o + 2] = 100.0 * 30.0 * c.aw640 * m640;
        S.m.set_step(o, 2.0 * S.xs[o]);
        S.m.set_step(o + 1, 2.0 * S.xs[o + 1]);
        S.m.set_step(o + 2, 2.0 * S.xs[o + 2]);
    });
}

// Begin a Nelder-Mead run from S.xs and the steps (asa047.c:162-170): the first point
// evaluated is the start itself, stored as vertex N.
template <int N>
PH_INLINE void ph_nm_begin(PhNM<N>& S, const PhConst& c)
{
    S.icount = 0;
    S.numres = 0;
    S.ihi = 0;
    S.ilo = 0;
    S.ylo = 0.0;
    S.yr = 0.0;
    S.jcount = S.kstart < 0 ? c.seed_konvge : c.konvge;
    S.del = 1.0;
    S.phase = PH_BUILD;
    S.sub = N;
    ph_for<N>([&](auto i) PH_BODY { S.xe[i] = S.xs[i]; });
}

// One step of the lane state machine: consume the value f of the point S.xe evaluated in
// phase S.phase, and advance to the next point to evaluate (left in S.xe), or to PH_START
// (the kernel then sets up the next start) or PH_DONE (outputs ready).
//
// Written as classify -> at most one simplex-vertex write -> shared follow-up, so that lanes
// at different steps of the simplex run the same instructions as far as possible.
template <int N>
PH_INLINE void ph_nm_consume(PhNM<N>& S, const PhConst& c, double f)
{
    const double dn = (double)N, dnn = (double)(N + 1);
    const double eps = 1.0e-6;
    enum {
        A_NONE,
        A_BUILD_NEXT,  // evaluate simplex vertex S.sub
        A_AFTER_BUILD, // simplex complete (or shrunk): find the best vertex, then iterate
        A_MOVE,        // evaluate xb + coef * (src - xb): expansion / contractions
        A_SHRINK_NEXT, // evaluate shrunk vertex S.sub
        A_ITER_END,
        A_FACT_FAIL,
        A_FACT_MINUS,
        A_FACT_NEXT,
        A_RUN_DONE
    };
    S.evals++;
    const int ph = S.phase;
    if (ph == PH_FINAL) {
        S.phase = PH_DONE;
        return;
    }
    if (ph < PH_BUILD || ph > PH_FACT_MINUS) return;
    S.icount++;

    // ---- classify (asa047.c:176-393, 455-479), as selects: lanes of a wave sit in every
    // phase at once, so branches would all run anyway ----
    const bool bB = ph == PH_BUILD, bR = ph == PH_REFLECT, bE = ph == PH_EXPAND;
    const bool bCI = ph == PH_CONTRACT_IN, bCO = ph == PH_CONTRACT_OUT, bS = ph == PH_SHRINK;
    const bool bF = ph >= PH_FACT_PLUS; // PH_FACT_PLUS or PH_FACT_MINUS
    int l = 0;
    ph_for<N + 1>([&](auto i) PH_BODY { l += (f < S.y[i]) ? 1 : 0; });
    const bool r_lo = f < S.ylo;
    const bool ci_shrink = ph_sel(S.y, S.ihi) < f;
    const bool fact_fail = ph_trunc10(f) < S.yr;
    const int bnext = (S.sub == N) ? 0 : S.sub + 1;
    const int snext = S.sub + 1;
    if (bR) S.ystar = f;
    const bool wr = bB || (bR && !r_lo && l > 1) || bE || (bCI && !ci_shrink) || bCO || bS;
    const int wj = (bB || bS) ? S.sub : S.ihi;
    const bool wr_xs = bR || (bE && S.ystar < f) || (bCO && !(f <= S.ystar));
    const double wf = ph_pick(wr_xs, S.ystar, f);
    int act = A_ITER_END; // bR (l > 1), bE, bCO, bCI without a shrink
    if (bB) act = bnext < N ? A_BUILD_NEXT : A_AFTER_BUILD;
    if (bR && (r_lo || l <= 1)) act = A_MOVE;
    if (bCI && ci_shrink) act = A_SHRINK_NEXT;
    if (bS) act = snext <= N ? A_SHRINK_NEXT : A_AFTER_BUILD;
    if (bF) act = fact_fail ? A_FACT_FAIL : (ph == PH_FACT_PLUS ? A_FACT_MINUS : A_FACT_NEXT);
    const bool mv_phi = !r_lo && l == 0;
    const double mv_coef = r_lo ? 2.0 : 0.5;
    const int mv_phase = r_lo ? PH_EXPAND : (l == 0 ? PH_CONTRACT_IN : PH_CONTRACT_OUT);
    if (bB) S.sub = bnext;
    if (bCI && ci_shrink) S.sub = 0;
    if (bS) S.sub = snext;

    // ---- the vertex write ----
    if (wr) {
        ph_for<N>([&](auto i) PH_BODY { S.m.set_p(wj, i, ph_pick(wr_xs, S.xs[i], S.xe[i])); });
        ph_for<N + 1>([&](auto k) PH_BODY { S.y[k] = ph_pick(wj == (int)k, wf, S.y[k]); });
    }

    // ---- follow-ups that only set the next point (asa047.c:174-194, 263-393, 455-493): one
    // read of at most two rows of the lane's arrays (row N + 1 is the steps), the formulas as
    // selects ----
    const bool fBN = act == A_BUILD_NEXT, fMV = act == A_MOVE, fSN = act == A_SHRINK_NEXT;
    const bool fFF = act == A_FACT_FAIL, fFM = act == A_FACT_MINUS, fFN = act == A_FACT_NEXT;
    if (fBN || fMV || fSN || fFF || fFM || fFN) {
        // BUILD_NEXT: vertex sub = start + step * del on coordinate sub; MOVE: xb + coef (v -
        // xb), v the worst vertex or the reflection; SHRINK_NEXT: (vertex sub + best vertex) /
        // 2; FACT_FAIL: restart from the probe point with a simplex of relative size eps
        // (jcount and icount carry over); FACT_MINUS / FACT_NEXT: the probe point moved by
        // step * eps on coordinates i0 and i1
        const int ra = (fBN || fFM || fFN) ? N + 1 : (fSN ? S.sub : S.ihi);
        const int rb = S.ilo;
        const int i0 = S.sub;
        const int i1 = fFM ? -1 : i0 + 1;
        ph_for<N>([&](auto i) PH_BODY {
            const double a = S.m.p(ra, i), b = S.m.p(rb, i);
            const double xs = S.xs[i], xb = S.xb[i];
            const double bn = (i0 == (int)i) ? xs + a * S.del : +xs;
            const double v = mv_phi ? a : xs;
            const double mv = xb + mv_coef * (v - xb);
            const double sn = (a + b) * 0.5;
            const double d = a * eps;
            double fv = xb;
            if (i0 == (int)i) fv = fFM ? fv - d - d : fv + d;
            if (i1 == (int)i) fv = fv + d;
            S.xe[i] = fBN ? bn : (fMV ? mv : (fSN ? sn : (fFF ? xb : fv)));
            if (fFF) S.xs[i] = xb;
            if (fFM || fFN) S.xb[i] = fv;
        });
        const bool run_done = fFN && i1 >= N;
        int nph = fBN || fFF ? (int)PH_BUILD : (fMV ? mv_phase : (fSN ? (int)PH_SHRINK : (fFM ? (int)PH_FACT_MINUS : (int)PH_FACT_PLUS)));
        if (!run_done) S.phase = nph;
        if (fFF) {
            S.ifault = 2;
            S.del = eps;
            S.numres++;
            S.sub = N;
        }
        if (fFN && !run_done) S.sub = i1;
        if (!run_done) return;
        S.ifault = 0;
        act = A_RUN_DONE;
    }

    bool iter_begin = false, nm_finish = false, give_up = false;
    if (act == A_AFTER_BUILD) {
        S.ylo = S.y[0];
        S.ilo = 0;
        ph_for<N>([&](auto i0) PH_BODY {
            constexpr int i = i0 + 1;
            if (S.y[i] < S.ylo) { S.ylo = S.y[i]; S.ilo = i; }
        });
        iter_begin = true;
    }
    // ---- end of an iteration: asa047.c:397-434 ----
    if (act == A_ITER_END) {
        const double yhi = ph_sel(S.y, S.ihi);
        if (yhi < S.ylo) { S.ylo = yhi; S.ilo = S.ihi; }
        S.jcount--;
        if (S.jcount > 0) {
            iter_begin = true;
        } else if (S.kstart < 0 && S.nstart == 0 && S.numres == 0 && c.seed_abort > 0.0 &&
                   S.ylo > c.seed_abort) {
            // a seed start still above the abort level at its first convergence check is
            // given up, as a run that did not converge; the depth starts then run as usual
            // (the seed start has no counterpart in the reference)
            give_up = true;
            nm_finish = true;
        } else if (S.icount <= c.kcount) {
            S.jcount = S.kstart < 0 ? c.seed_konvge : c.konvge;
            double z = 0.0;
            ph_for<N + 1>([&](auto i) PH_BODY { z = z + S.y[i]; });
            const double xm = ph_div_const(z, dnn, 1.0 / dnn);
            z = 0.0;
            ph_for<N + 1>([&](auto i) PH_BODY { z = z + (S.y[i] - xm) * (S.y[i] - xm); });
            if (z <= c.rq) nm_finish = true;
            else iter_begin = true;
        } else {
            iter_begin = true;
        }
    }
    // ---- start of an iteration: asa047.c:215-253 ----
    if (iter_begin) {
        if (c.kcount <= S.icount) {
            nm_finish = true;
        } else {
            S.ynewlo = S.y[0];
            S.ihi = 0;
            ph_for<N>([&](auto i0) PH_BODY {
                constexpr int i = i0 + 1;
                if (S.ynewlo < S.y[i]) { S.ynewlo = S.y[i]; S.ihi = i; }
            });
            const int hi = S.ihi;
            ph_for<N>([&](auto i) PH_BODY {
                double z = 0.0;
                ph_for<N + 1>([&](auto k) PH_BODY { z = z + S.m.p(k, i); });
                const double phi = S.m.p(hi, i);
                z = z - phi;
                S.xb[i] = ph_div_const(z, dn, 1.0 / dn);
                S.xs[i] = S.xb[i] + 1.0 * (S.xb[i] - phi);
                S.xe[i] = S.xs[i];
            });
            S.phase = PH_REFLECT;
            return;
        }
    }
    // ---- convergence or evaluations exhausted: asa047.c:440-454 ----
    if (nm_finish) {
        const int lo = S.ilo;
        ph_for<N>([&](auto i) PH_BODY { S.xb[i] = S.m.p(lo, i); });
        S.ynewlo = ph_sel(S.y, lo);
        S.yr = ph_trunc10(S.ynewlo);
        if (c.kcount < S.icount || give_up) {
            S.ifault = 2;
            act = A_RUN_DONE;
        } else {
            S.ifault = 0;
            S.sub = 0;
            ph_for<N>([&](auto i) PH_BODY {
                if (i == 0) S.xb[i] = S.xb[i] + S.m.step(i) * eps;
                S.xe[i] = S.xb[i];
            });
            S.phase = PH_FACT_PLUS;
            return;
        }
    }
    // ---- a Nelder-Mead run ended; its result is (xb, ynewlo, icount, ifault) ----
    // multi-start bookkeeping, model/samodel.c:2385-2405
    if (act == A_RUN_DONE) {
        S.nstart++;
        if (S.ynewlo < S.lowest) {
            S.lowest = S.ynewlo;
            ph_for<N>([&](auto i) PH_BODY { S.m.set_best(i, S.xb[i]); });
            S.best_icount = S.icount;
            S.best_conv = (S.ifault == 0) ? 1 : 0;
            if (S.lowest < c.early_exit) S.kstart = c.n_hstarts;
        }
        S.kstart++;
        if (S.kstart < c.n_hstarts) {
            S.phase = PH_START;
            return;
        }
        ph_for<N>([&](auto i) PH_BODY { S.xe[i] = S.m.best(i); });
        S.phase = PH_FINAL;
    }
}
