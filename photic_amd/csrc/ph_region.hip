// photic_amd -- spatial regions on gfx950: the reference's n_spatial > 1 run, in which each
// pixel is inverted jointly with its (2 n_spatial - 1)^2 neighbours (model/samodel.c:
// 2957-3027): one depth, bottom amplitudes and bottom weights per region, one P, G, X per
// scene shared by every region, and a misfit over every (region, scene, band) plus the depth
// and bottom continuity penalties and the K penalty (model/samodel.c:2432-2759).
//
// With up to 25 regions the problem has up to ~100 parameters and 800 band evaluations per
// objective, so one wavefront runs one pixel:
//
//   lane l          owns coordinates l and l + 64 (MC slots): start point, centroid, trial
//                   point, step and best value in registers, its columns of the simplex in
//                   LDS ([vertex][coordinate], stride N);
//   lane l          evaluates pairs l, l + 64, ... (pair j = (region, scene, band) in the
//                   reference's loop order) of every objective evaluation;
//   lane k % 64     holds vertex value k in register slot k / 64; the scans for the best and
//                   worst vertex are wave reductions;
//   every lane      runs the modified AS 47 Nelder-Mead state machine (model/asa047.c:10-502)
//                   and the multi-start (model/samodel.c:2119-2427) on identical values.
//
// The trial point goes through LDS once per evaluation; the pairs' squared residuals, the last
// region's attenuations and the per-region depth and bottom terms come back through LDS, and
// every lane re-forms the misfit from them in the reference's order, so the result is the
// oracle's bit for bit (FP64, correctly rounded transcendentals, no contraction).  Band
// coefficients and scene geometry sit in small LDS tables.  Pixels come from one global
// counter, one per wavefront.
#include <hip/hip_runtime.h>

#include "ph_grpengine.h"
#include "ph_launch.h"

// The region kernel reads its arguments through the kernel-argument segment pointer
// (ph_kernarg): they must be its only argument.
static_assert(offsetof(PhRegionArgs, c) == 0, "PhConst must start the kernel arguments");

namespace {

PH_DEV int ph_uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Lane l's value of v, for a wave-uniform l.
PH_DEV double ph_readlane(double v, int l)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned int)lo);
}

// v[0] + v[1] + ... + v[n - 1] in that order, from LDS: the loads of a batch of eight go out
// together, the additions stay in sequence.
PH_DEV double ph_seq_sum(const double* v, int n)
{
    double s = 0.0;
    int j = 0;
    for (; j + 8 <= n; j += 8) {
        double t[8];
#pragma unroll
        for (int u = 0; u < 8; u++) t[u] = v[j + u];
#pragma unroll
        for (int u = 0; u < 8; u++) s += t[u];
    }
    for (; j < n; j++) s += v[j];
    return s;
}

// j / d for 0 <= j < 2^23 and d >= 1 from the float reciprocal inv = 1 / d, corrected to the
// exact quotient (the estimate is within one of it).
PH_DEV int ph_div_small(int j, int d, float inv)
{
    int r = (int)((float)j * inv);
    if (r * d > j) r--;
    else if ((r + 1) * d <= j) r++;
    return r;
}

template <int MC>
struct PhReg {
    // vertex value k lives in lane k % 64, slot k / 64
    static constexpr int YC = MC + 1;
    double* g; // the wave's simplex, [vertex][coordinate], stride N
    double* t; // the wave's arrays, at the o_* offsets
    int li, R, N, NP, nd, origin;
    int p1;       // band slots of one region: the scenes' band counts summed
    float p1inv;  // 1 / p1
    double rq;    // reqmin * N
    double xs[MC], xb[MC], xe[MC], step[MC], best[MC], xs0[MC], step0[MC];
    double yv[YC];
    double h0, tot;
    double ylo, ystar, ynewlo, del, lowest, yr;
    int ilo, ihi, icount, jcount, numres, ifault, phase, sub, kstart, nstart;
    int best_icount, best_conv, evals;

    PH_DEV int cidx(int m) const { return li + 64 * m; }
    PH_DEV bool co(int m) const { return li + 64 * m < N; }
    PH_DEV double p(int k, int m) const { return g[k * N + li + 64 * m]; }
    PH_DEV void set_p(int k, int m, double v) const { g[k * N + li + 64 * m] = v; }
    // vertex value k (k wave-uniform)
    PH_DEV double y(int k) const
    {
        double v = yv[0];
#pragma unroll
        for (int m = 1; m < YC; m++) v = ph_pick((k >> 6) == m, yv[m], v);
        return ph_readlane(v, k & 63);
    }
    PH_DEV void set_y(int k, double v)
    {
#pragma unroll
        for (int m = 0; m < YC; m++) yv[m] = ph_pick(k == li + 64 * m, v, yv[m]);
    }
    // The scan "v = y_0, at = 0; for i = 1..N: if (v < y_i) v = y_i, at = i"
    // (asa047.c:223-231) as a wave reduction: no comparison with a NaN holds, so a NaN y_0
    // stays; otherwise it ends at the first index of the largest value.  WANT_MAX false: the
    // same with ">" (the lowest vertex, asa047.c:196-204).
    template <bool WANT_MAX>
    PH_DEV int scan(double* val) const
    {
        const double y0 = ph_readlane(yv[0], 0);
        if (y0 != y0) {
            *val = y0;
            return 0;
        }
        double m = y0;
#pragma unroll
        for (int s = 0; s < YC; s++)
            if (li + 64 * s <= N && (WANT_MAX ? yv[s] > m : yv[s] < m)) m = yv[s];
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            const double o = __shfl_xor(m, off);
            if (WANT_MAX ? o > m : o < m) m = o;
        }
        int at = 0;
#pragma unroll
        for (int s = YC - 1; s >= 0; s--) {
            const unsigned long long b = __ballot(li + 64 * s <= N && yv[s] == m);
            if (b) at = 64 * s + __ffsll((long long)b) - 1;
        }
        *val = y(at);
        return at;
    }
};

struct PhRegSide {
    double k_min, iod_sum, nobs, Rrs_error, model_error, albedo;
};

// The region misfit of the point held in the lanes' xe (model/samodel.c:2432-2759).
template <int NBOT, int MC>
PH_DEV double ph_reg_objective(const PhReg<MC>& S, const PhRegionArgs& a, const PhConst& c,
                               bool side_on, PhRegSide* side)
{
    double* g = S.t;
    const int li = S.li, R = S.R, NP = S.NP, nd = S.nd;
    const int off = R + 2 * NBOT * R;
    ph_wave_sync(); // the last evaluation's readers are done with the buffers
#pragma unroll
    for (int m = 0; m < MC; m++)
        if (S.co(m)) g[a.o_x + S.cidx(m)] = S.xe[m];
    ph_wave_sync();
    const double* X = g + a.o_x;
    for (int j = li; j < NP; j += 64) {
        // pair j = (region r, band slot): the slot is band b of scene s
        const int r = ph_div_small(j, S.p1, S.p1inv);
        const int slot = j - r * S.p1;
        const double* bc = g + a.o_bc + PH_REG_SLOT * slot; // a0, a1, aw, bbw, eg, bottoms, ...
        const int s = (int)bc[9] >> 4;
        const double H = fabs(X[r]);
        double Bk[NBOT], q[NBOT], qsum = 0.0;
#pragma unroll
        for (int k = 0; k < NBOT; k++) Bk[k] = 0.01 * fabs(X[R + r * NBOT + k]);
#pragma unroll
        for (int k = 0; k < NBOT; k++) {
            q[k] = fabs(X[R + R * NBOT + r * NBOT + k]);
            qsum += q[k];
        }
#pragma unroll
        for (int k = 0; k < NBOT; k++) q[k] /= qsum;
        double rho = 0.0;
#pragma unroll
        for (int k = 0; k < NBOT; k++) rho += q[k] * Bk[k] * bc[5 + k];
        const double P = 0.01 * fabs(X[off + 3 * s]);
        const double Gc = 0.01 * fabs(X[off + 1 + 3 * s]);
        const double Xc = 0.01 * fabs(X[off + 2 + 3 * s]);
        const double ss = g[a.o_geo + s], sv = g[a.o_geo + nd + s];
        double K, rB, rr;
        const double m = ph_band64_nb(P, Gc, Xc, rho, H, bc[0], bc[1], bc[2], bc[3], bc[4],
                                      g[a.o_pw + j], ss, sv, &K, &rB, &rr);
        const double d = m - g[a.o_rm + j];
        g[a.o_d + j] = d * d;
        if (r == R - 1) g[a.o_k + slot] = K; // the reference keeps the last region's K
        if (side_on) g[a.o_q + j] = rB / rr;
    }
    // lane r < R: region r's bottom fractions |B_k| |q_k| / sum |q| (model/samodel.c:2648-2686),
    // bottom-major in LDS
    double bq[NBOT];
    if (li < R) {
        double qs = 0.0;
#pragma unroll
        for (int kb = 0; kb < NBOT; kb++) qs += fabs(X[R + R * NBOT + li * NBOT + kb]);
#pragma unroll
        for (int k = 0; k < NBOT; k++) {
            bq[k] = fabs(X[R + li * NBOT + k]) * fabs(X[R + R * NBOT + li * NBOT + k]) / qs;
            g[a.o_bqv + k * R + li] = bq[k];
        }
    }
    ph_wave_sync();
    // lane s < nd: scene s's Kmin over the last region's bands and its K-penalty term, at the
    // origin's depth (model/samodel.c:2696-2732)
    double kterm = 0.0, kmin_s = 1.0e4;
    if (li < nd) {
        const double H = fabs(X[S.origin]);
        const double kmm = 0.275, kmn = 0.185;
        const int s0 = (int)g[a.o_sc + 2 * li], nbs = (int)g[a.o_sc + 2 * li + 1];
        for (int b = 0; b < nbs; b++) {
            const double K = g[a.o_k + s0 + b];
            if (!ph_approx_equal((float)K, 0.0f, 1.0e-6f) && K < kmin_s) kmin_s = K;
        }
        double thr = -1.0;
        if (H < 1.0 && kmin_s < kmm) thr = 0.01 + kmn;
        else if (H < 2.0 && kmin_s < 0.5 * (1.5 * kmn + 0.5 * kmm)) thr = 0.01 + 0.5 * (1.5 * kmn + 0.5 * kmm);
        else if (H < 3.0 && kmin_s < 0.5 * (1.25 * kmn + 0.75 * kmm)) thr = 0.01 + 0.5 * (1.25 * kmn + 0.75 * kmm);
        else if (H < 4.0 && kmin_s < 0.5 * (1.5 * kmn + 0.5 * kmm)) thr = 0.01 + 0.5 * (1.5 * kmn + 0.5 * kmm);
        else if (H < 5.0 && kmin_s < 0.5 * (1.75 * kmn + 0.25 * kmm)) thr = 0.01 + 0.5 * (1.75 * kmn + 0.25 * kmm);
        // the sum below adds +0.0 for a scene without a penalty: exact, since the running
        // penalty is never -0.0
        if (thr > 0.0) {
            const double t = 1.0 / (0.01 + kmin_s) - 1.0 / thr;
            kterm = 100.0 * (t * t);
        }
    }
    // depth continuity (model/samodel.c:2599-2630) and the bottoms' means.  An in-band region
    // adds +0.0 to the running sums of squares below: exact, they are never -0.0.
    double dm = 0.0;
    {
        int r = 0;
        for (; r + 8 <= R; r += 8) {
            double t[8];
#pragma unroll
            for (int u = 0; u < 8; u++) t[u] = X[r + u];
#pragma unroll
            for (int u = 0; u < 8; u++) dm += fabs(t[u]);
        }
        for (; r < R; r++) dm += fabs(X[r]);
    }
    dm /= (double)R;
    double thr;
    if (dm < 4.0) thr = 0.4;
    else if (dm < 8.0) thr = 0.2;
    else if (dm < 12.0) thr = 0.1;
    else thr = 0.05;
    double bthr;
    if (dm < 5.0) bthr = 0.25;
    else if (dm < 10.0) bthr = 0.1;
    else if (dm < 15.0) bthr = 0.05;
    else bthr = 0.01;
    double bm[NBOT], btot = 0.0;
#pragma unroll
    for (int k = 0; k < NBOT; k++) {
        bm[k] = ph_seq_sum(g + a.o_bqv + k * R, R) / (double)R;
        btot += bm[k];
    }
    bool hout = false, bout[NBOT];
#pragma unroll
    for (int k = 0; k < NBOT; k++) bout[k] = false;
    if (li < R) {
        const double hl = fabs(X[li]);
        hout = hl < (1.0 - thr) * dm || hl > (1.0 + thr) * dm;
        g[a.o_hd + li] = hout ? (hl - dm) * (hl - dm) : 0.0;
#pragma unroll
        for (int k = 0; k < NBOT; k++) {
            bout[k] = bq[k] < (1.0 - bthr) * bm[k] || bq[k] > (1.0 + bthr) * bm[k];
            g[a.o_bd + k * R + li] = bout[k] ? (bq[k] - bm[k]) * (bq[k] - bm[k]) : 0.0;
        }
    }
    const double hcount = (double)__popcll(__ballot(hout));
    int nb_out = 0;
#pragma unroll
    for (int k = 0; k < NBOT; k++) nb_out += __popcll(__ballot(bout[k]));
    const double bcount = (double)nb_out;
    ph_wave_sync();

    // ---- the misfit in the reference's order, every lane alike ----
    const double err = ph_seq_sum(g + a.o_d, NP);
    const double nt = (double)NP;
    const double mean = S.tot / nt;
    const double e_rrs = 100.0 * sqrt(err / nt) / mean;
    double ed = ph_seq_sum(g + a.o_hd, R);
    if (hcount > 0.5) ed = 100.0 * sqrt(ed / hcount) / dm;
    double eb = ph_seq_sum(g + a.o_bd, NBOT * R);
    if (bcount > 0.5) {
        const double bmean = btot / (double)NBOT;
        eb = 100.0 * sqrt(eb / bcount) / bmean;
    }
    double Kerr = 0.0;
    for (int s = 0; s < nd; s++) Kerr += ph_readlane(kterm, s);
    const double Kmin = ph_readlane(kmin_s, nd - 1);
    if (Kmin > 0.7) {
        const double t = 4.0 * (Kmin - 0.7);
        Kerr += 100.0 * (t * t);
    }
    double e = 80.0 * e_rrs + 15.0 * ed + 10.0 * eb + 15.0 * Kerr;
    e /= 120.0;
    if (side_on) {
        double kmin_all = 1.0e10;
        for (int j = 0; j < S.p1; j++) {
            const double K = g[a.o_k + j];
            if (!ph_approx_equal((float)K, 0.0f, 1.0e-4f) && K < kmin_all) kmin_all = K;
        }
        // IOD in (scene, region, band) order (model/samodel.c:2051-2064)
        double iod = 0.0, nobs = 0.0;
        for (int s = 0; s < nd; s++) {
            const int s0 = (int)g[a.o_sc + 2 * s], nbs = (int)g[a.o_sc + 2 * s + 1];
            for (int r = 0; r < R; r++)
                for (int b = 0; b < nbs; b++) {
                    iod += g[a.o_q + r * S.p1 + s0 + b];
                    nobs += 1.0;
                }
        }
        // the albedo of the last band evaluated: region R - 1's sum q_k B_k
        double Bk[NBOT], q[NBOT], qsum = 0.0, alb = 0.0;
#pragma unroll
        for (int k = 0; k < NBOT; k++) Bk[k] = 0.01 * fabs(X[R + (R - 1) * NBOT + k]);
#pragma unroll
        for (int k = 0; k < NBOT; k++) {
            q[k] = fabs(X[R + R * NBOT + (R - 1) * NBOT + k]);
            qsum += q[k];
        }
#pragma unroll
        for (int k = 0; k < NBOT; k++) q[k] /= qsum;
#pragma unroll
        for (int k = 0; k < NBOT; k++) alb += q[k] * Bk[k];
        side->k_min = kmin_all == 1.0e10 ? 0.0 : kmin_all;
        side->iod_sum = iod;
        side->nobs = nobs;
        side->Rrs_error = e_rrs;
        side->model_error = e;
        side->albedo = alb;
    }
    return e;
}

// Start point and steps of depth start k (model/samodel.c:2246-2354): every region's depth at
// the start depth (the pixel's own when it has one), step 1.25 H; the rest from xs0 / step0.
template <int MC>
PH_DEV void ph_reg_start_step(PhReg<MC>& S, const PhConst& c, int k)
{
    double h0 = c.hstart[0];
    ph_for<PH_MAXH - 1>([&](auto m0) PH_BODY {
        constexpr int m = m0 + 1;
        h0 = ph_pick(k == m, c.hstart[m], h0);
    });
    if (S.h0 >= 0.0) h0 = S.h0;
#pragma unroll
    for (int m = 0; m < MC; m++) {
        const bool hc = S.cidx(m) < S.R;
        S.xs[m] = hc ? h0 : S.xs0[m];
        S.step[m] = hc ? 1.25 * h0 : S.step0[m];
    }
}

template <int MC>
PH_DEV void ph_reg_nm_begin(PhReg<MC>& S, const PhConst& c)
{
    S.icount = 0;
    S.numres = 0;
    S.jcount = c.konvge;
    S.del = 1.0;
    S.phase = PH_BUILD;
    S.sub = S.N;
#pragma unroll
    for (int m = 0; m < MC; m++) S.xe[m] = S.xs[m];
}

// One step of the state machine (the group engine's, with runtime N and MC coordinates per
// lane): consume f, the value of the point in xe in phase S.phase, and leave the next point to
// evaluate in xe (or PH_START / PH_DONE).
template <int MC>
PH_DEV void ph_reg_consume(PhReg<MC>& S, const PhConst& c, double f)
{
    const int N = S.N;
    const double dn = (double)N, dnn = (double)(N + 1);
    const double eps = 1.0e-6;
    enum {
        A_NONE,
        A_BUILD_NEXT,
        A_AFTER_BUILD,
        A_MOVE,
        A_SHRINK_NEXT,
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

    // ---- classify (asa047.c:176-393, 455-479) ----
    bool wr = false, wr_xs = false;
    int wj = 0;
    double wf = f;
    int act = A_NONE;
    bool mv_phi = false;
    double mv_coef = 0.5;
    int mv_phase = PH_EXPAND;
    if (ph == PH_BUILD) {
        wr = true;
        wj = S.sub;
        const int next = (S.sub == N) ? 0 : S.sub + 1;
        S.sub = next;
        act = (next < N) ? A_BUILD_NEXT : A_AFTER_BUILD;
    } else if (ph == PH_REFLECT) {
        S.ystar = f;
        if (f < S.ylo) {
            act = A_MOVE;
            mv_coef = 2.0;
            mv_phase = PH_EXPAND;
        } else {
            int l = 0;
#pragma unroll
            for (int m = 0; m < PhReg<MC>::YC; m++)
                l += __popcll(__ballot(S.li + 64 * m <= N && f < S.yv[m]));
            if (l > 1) {
                wr = true;
                wj = S.ihi;
                wr_xs = true;
                wf = f;
                act = A_ITER_END;
            } else if (l == 0) {
                act = A_MOVE;
                mv_phi = true;
                mv_phase = PH_CONTRACT_IN;
            } else {
                act = A_MOVE;
                mv_phase = PH_CONTRACT_OUT;
            }
        }
    } else if (ph == PH_EXPAND) {
        wr = true;
        wj = S.ihi;
        wr_xs = S.ystar < f;
        wf = ph_pick(wr_xs, S.ystar, f);
        act = A_ITER_END;
    } else if (ph == PH_CONTRACT_IN) {
        if (S.y(S.ihi) < f) {
            S.sub = 0;
            act = A_SHRINK_NEXT;
        } else {
            wr = true;
            wj = S.ihi;
            act = A_ITER_END;
        }
    } else if (ph == PH_CONTRACT_OUT) {
        wr = true;
        wj = S.ihi;
        wr_xs = !(f <= S.ystar);
        wf = ph_pick(wr_xs, S.ystar, f);
        act = A_ITER_END;
    } else if (ph == PH_SHRINK) {
        wr = true;
        wj = S.sub;
        S.sub = S.sub + 1;
        act = (S.sub <= N) ? A_SHRINK_NEXT : A_AFTER_BUILD;
    } else if (ph == PH_FACT_PLUS) {
        act = (ph_trunc10(f) < S.yr) ? A_FACT_FAIL : A_FACT_MINUS;
    } else { // PH_FACT_MINUS
        act = (ph_trunc10(f) < S.yr) ? A_FACT_FAIL : A_FACT_NEXT;
    }

    // ---- the vertex write ----
    if (wr) {
#pragma unroll
        for (int m = 0; m < MC; m++)
            if (S.co(m)) S.set_p(wj, m, ph_pick(wr_xs, S.xs[m], S.xe[m]));
        S.set_y(wj, wf);
    }

    // ---- follow-up ----
    if (act == A_BUILD_NEXT) {
        const int j = S.sub;
#pragma unroll
        for (int m = 0; m < MC; m++)
            S.xe[m] = (j == S.cidx(m)) ? S.xs[m] + S.step[m] * S.del : +S.xs[m];
        S.phase = PH_BUILD;
        return;
    }
    if (act == A_MOVE) {
#pragma unroll
        for (int m = 0; m < MC; m++)
            if (S.co(m)) {
                const double v = mv_phi ? S.p(S.ihi, m) : +S.xs[m];
                const double b = S.xb[m];
                S.xe[m] = b + mv_coef * (v - b);
            }
        S.phase = mv_phase;
        return;
    }
    if (act == A_SHRINK_NEXT) {
#pragma unroll
        for (int m = 0; m < MC; m++)
            if (S.co(m)) S.xe[m] = (S.p(S.sub, m) + S.p(S.ilo, m)) * 0.5;
        S.phase = PH_SHRINK;
        return;
    }
    if (act == A_FACT_FAIL) {
        // asa047.c:488-493: restart from the probe point with a simplex of relative size eps
        S.ifault = 2;
#pragma unroll
        for (int m = 0; m < MC; m++) {
            S.xs[m] = S.xb[m];
            S.xe[m] = S.xb[m];
        }
        S.del = eps;
        S.numres++;
        S.phase = PH_BUILD;
        S.sub = N;
        return;
    }
    if (act == A_FACT_MINUS || act == A_FACT_NEXT) {
        const int i0 = S.sub;
        const bool minus = (act == A_FACT_MINUS);
        const int i1 = minus ? -1 : i0 + 1;
#pragma unroll
        for (int m = 0; m < MC; m++) {
            const double d = S.step[m] * eps;
            double v = S.xb[m];
            if (i0 == S.cidx(m)) v = minus ? v - d - d : v + d;
            if (i1 == S.cidx(m)) v = v + d;
            S.xb[m] = v;
            S.xe[m] = v;
        }
        if (minus) {
            S.phase = PH_FACT_MINUS;
            return;
        }
        if (i1 < N) {
            S.sub = i1;
            S.phase = PH_FACT_PLUS;
            return;
        }
        S.ifault = 0;
        act = A_RUN_DONE;
    }

    bool iter_begin = false, nm_finish = false;
    if (act == A_AFTER_BUILD) {
        S.ilo = S.template scan<false>(&S.ylo);
        iter_begin = true;
    }
    // ---- end of an iteration: asa047.c:397-434 ----
    if (act == A_ITER_END) {
        const double yhi = S.y(S.ihi);
        if (yhi < S.ylo) {
            S.ylo = yhi;
            S.ilo = S.ihi;
        }
        S.jcount--;
        if (S.jcount > 0) {
            iter_begin = true;
        } else if (S.icount <= c.kcount) {
            S.jcount = c.konvge;
            double z = 0.0;
            for (int i = 0; i <= N; i++) z = z + S.y(i);
            const double xm = z / dnn;
            z = 0.0;
            for (int i = 0; i <= N; i++) {
                const double yi = S.y(i);
                z = z + (yi - xm) * (yi - xm);
            }
            if (z <= S.rq) nm_finish = true;
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
            S.ihi = S.template scan<true>(&S.ynewlo);
            // centroid without the worst vertex: the sum over vertices 0..N in order, less
            // the worst vertex's coordinate
            const int hi = S.ihi;
#pragma unroll
            for (int m = 0; m < MC; m++)
                if (S.co(m)) {
                    // the column's loads go out eight at a time, the sum stays in order
                    double z = 0.0;
                    int k = 0;
                    for (; k + 8 <= N + 1; k += 8) {
                        double t[8];
#pragma unroll
                        for (int u = 0; u < 8; u++) t[u] = S.p(k + u, m);
#pragma unroll
                        for (int u = 0; u < 8; u++) z = z + t[u];
                    }
                    for (; k <= N; k++) z = z + S.p(k, m);
                    const double phi = S.p(hi, m);
                    const double b = (z - phi) / dn;
                    const double r = b + 1.0 * (b - phi);
                    S.xb[m] = b;
                    S.xs[m] = r;
                    S.xe[m] = r;
                }
            S.phase = PH_REFLECT;
            return;
        }
    }
    // ---- convergence or evaluations exhausted: asa047.c:440-454 ----
    if (nm_finish) {
        const int lo = S.ilo;
#pragma unroll
        for (int m = 0; m < MC; m++)
            if (S.co(m)) S.xb[m] = S.p(lo, m);
        S.ynewlo = S.y(lo);
        S.yr = ph_trunc10(S.ynewlo);
        if (c.kcount < S.icount) {
            S.ifault = 2;
            act = A_RUN_DONE;
        } else {
            S.ifault = 0;
            S.sub = 0;
            if (S.li == 0) S.xb[0] = S.xb[0] + S.step[0] * eps;
#pragma unroll
            for (int m = 0; m < MC; m++) S.xe[m] = S.xb[m];
            S.phase = PH_FACT_PLUS;
            return;
        }
    }
    // ---- a run ended; the multi-start bookkeeping, model/samodel.c:2385-2405 ----
    if (act == A_RUN_DONE) {
        S.nstart++;
        if (S.ynewlo < S.lowest) {
            S.lowest = S.ynewlo;
#pragma unroll
            for (int m = 0; m < MC; m++) S.best[m] = S.xb[m];
            S.best_icount = S.icount;
            S.best_conv = (S.ifault == 0) ? 1 : 0;
            if (S.lowest < c.early_exit) S.kstart = c.n_hstarts;
        }
        S.kstart++;
        if (S.kstart < c.n_hstarts) {
            S.phase = PH_START;
            return;
        }
#pragma unroll
        for (int m = 0; m < MC; m++) S.xe[m] = S.best[m];
        S.phase = PH_FINAL;
    }
}

__device__ void ph_reg_write_nodata(const PhRegionArgs& a, long q, int n_regions)
{
    // the reference's defaults for pixels it does not model (model/samodel.c:819-841)
    if (a.depth) a.depth[q] = 0.0f;
    if (a.rrs_error) a.rrs_error[q] = 0.0f;
    if (a.model_error) a.model_error[q] = 0.0f;
    if (a.albedo) a.albedo[q] = 0.0f;
    if (a.k_min) a.k_min[q] = 0.0f;
    if (a.iod) a.iod[q] = 0.0f;
    if (a.b_pct)
        for (int m = 0; m < 3; m++) a.b_pct[(long)m * a.n + q] = -9999.0f;
    if (a.bottom_type) a.bottom_type[q] = -9999.0f;
    if (a.status) a.status[q] = -1;
    if (a.n_evals) a.n_evals[q] = 0;
    if (a.n_iter) a.n_iter[q] = 0;
    if (a.n_starts) a.n_starts[q] = 0;
    if (a.n_regions) a.n_regions[q] = n_regions;
}

} // namespace

// One wavefront per pixel (block = 64 lanes).  GM false: the wave's block of a.lds_doubles
// doubles is its dynamic LDS; GM true: its simplex is at a.work + blockIdx.x * a.work_doubles
// in device memory, read and written by this wave alone, and its arrays in LDS (TL 1, the
// launch's a.tables_in_lds) or after the simplex in device memory (TL 2).  The choice is made
// at compile time so that the arrays' accesses compile to LDS instructions with 32-bit
// addresses: through a pointer that may address either they would be flat ones.
template <int NBOT, int MC, bool GM, int TL>
__device__ __forceinline__ void ph_region_body(const PhRegionArgs& a)
{
    static_assert(GM ? (TL == 1 || TL == 2) : TL == 0, "TL: 1 or 2 with GM, else 0");
    extern __shared__ double ph_reg_lds[];
    const PhConst& c = a.c;
    PhReg<MC> S;
    if constexpr (GM) {
        S.g = a.work + (long)blockIdx.x * a.work_doubles;
        if constexpr (TL == 1) S.t = ph_reg_lds;
        else S.t = S.g;
    } else {
        S.g = ph_reg_lds;
        S.t = ph_reg_lds;
    }
    S.li = (int)__lane_id();
    const int li = S.li;
    const int nd = a.n_scenes;
    S.nd = nd;
    const int P1 = a.n_slots;
    S.p1 = P1;
    S.p1inv = 1.0f / (float)P1;
    const long npix = (long)a.n_rows * a.n_cols;
    // watchdog: the wall clock at the wave's last progress (a pixel taken); a wave that
    // finishes no pixel within max_ticks gives up and raises the flag
    unsigned long long t0 = wall_clock64();
    double* g = S.t;
    // the per-call tables (ph_launch.h, PhRegionArgs::scn): each band slot's coefficients,
    // each scene's geometry and its band slots, once per wave
    const double* scn_rsig = a.scn + 2 * nd;                              // [slot]
    const double* scn_nodata = scn_rsig + P1;                             // [slot], floats
    const double* scn_slot = scn_nodata + P1;                             // [slot][PH_REG_SLOT]
    const double* scn_scene = scn_slot + PH_REG_SLOT * P1;                // [scene][PH_REG_SCENE]
    for (int t = li; t < PH_REG_SLOT * P1; t += 64) g[a.o_bc + t] = scn_slot[t];
    for (int t = li; t < 2 * nd; t += 64) {
        g[a.o_geo + t] = a.scn[t];
        g[a.o_sc + t] = scn_scene[PH_REG_SCENE * (t >> 1) + (t & 1)]; // first slot, band count
    }

    for (;;) {
        if (wall_clock64() - t0 > a.max_ticks) {
            if (li == 0) atomicOr(a.queue + 1, 1ull);
            break;
        }
        unsigned long long qq = 0;
        if (li == 0) qq = atomicAdd(a.queue, 1ull);
        const long q = (long)__shfl(qq, 0);
        if (q >= a.n) break;
        t0 = wall_clock64();
        const long p = a.idx ? a.idx[q] : q;
        if (p < 0 || p >= npix) { // a listed index outside the scene: not modelled
            if (li == 0) ph_reg_write_nodata(a, q, 0);
            if (a.x_out)
                for (int i = li; i < a.n_max; i += 64) a.x_out[(long)i * a.n + q] = 0.0;
            continue;
        }
        const int pi = (int)(p / a.n_cols), pj = (int)(p % a.n_cols);

        // without a pixel list, the pixel's own screen: any band of any scene no-data or
        // negative (model/samodel.c:933-947)
        if (!a.idx) {
            bool bad = false;
            for (int t = li; t < P1; t += 64) {
                const float v = a.rrs[(long)t * npix + p];
                bad |= ph_approx_equal(v, (float)scn_nodata[t], 1.0e-6f) || v < 0.0f;
            }
            if (__ballot(bad) != 0ull) {
                if (li == 0) ph_reg_write_nodata(a, q, 0);
                if (a.x_out)
                    for (int i = li; i < a.n_max; i += 64) a.x_out[(long)i * a.n + q] = 0.0;
                continue;
            }
        }

        // ---- the regions (model/samodel.c:2957-3027): candidate li of the (2 ns - 1)^2
        // block, row-major, clamped; dropped when any band is no-data ----
        const int ns = a.n_spatial == 0 ? 1 : a.n_spatial;
        const int w = 2 * ns - 1;
        bool ok = false, centre = false;
        long cpix = 0;
        if (li < w * w) {
            const int ii = li / w - (ns - 1), jj = li % w - (ns - 1);
            const int iii = pi + ii < 0 ? 0 : (pi + ii > a.n_rows - 1 ? a.n_rows - 1 : pi + ii);
            const int jjj = pj + jj < 0 ? 0 : (pj + jj > a.n_cols - 1 ? a.n_cols - 1 : pj + jj);
            cpix = (long)iii * a.n_cols + jjj;
            centre = (iii == pi && jjj == pj);
            ok = true;
            for (int t = 0; t < P1; t++)
                if (ph_approx_equal(a.rrs[(long)t * npix + cpix], (float)scn_nodata[t], 1.0e-6f))
                    ok = false;
        }
        const unsigned long long okm = __ballot(ok), cm = __ballot(centre);
        const int R = ph_uni(__popcll(okm));
        const int tc = 63 - __clzll((long long)cm); // the last visit of the centre
        const int origin = ph_uni(__popcll(okm & ((1ull << tc) - 1ull)));
        if (R == 0) {
            if (li == 0) ph_reg_write_nodata(a, q, 0);
            if (a.x_out)
                for (int i = li; i < a.n_max; i += 64) a.x_out[(long)i * a.n + q] = 0.0;
            continue;
        }
        const int N = ph_uni(R * (1 + 2 * NBOT) + 3 * nd);
        const int NP = ph_uni(R * P1); // (region, band slot) pairs
        S.R = R;
        S.N = N;
        S.NP = NP;
        S.origin = origin;
        S.rq = c.reqmin * (double)N;

        ph_wave_sync(); // the last pixel's readers are done with the buffers
        if (ok) g[a.o_pix + __popcll(okm & ((1ull << li) - 1ull))] = (double)cpix;
        ph_wave_sync();

        // ---- pairs: measured Rrs and (440 / lambda)^Y of the region's scene (Y from the
        // region's 440/490 ratio, model/samodel.c:2898-2903); per (region, scene) Rrs at
        // 490, 550 and 640 nm for the starts ----
        for (int j = li; j < NP; j += 64) {
            // pair j = (region r, band slot qs), slot qs band b of scene s
            const int r = ph_div_small(j, P1, S.p1inv);
            const int qs = j - r * P1;
            const int sb = (int)g[a.o_bc + PH_REG_SLOT * qs + 9];
            const int s = sb >> 4, b = sb & 15;
            const double* sc = scn_scene + PH_REG_SCENE * s;
            const int row0 = (int)sc[0], nbs = (int)sc[1];
            const long rp = (long)g[a.o_pix + r];
            double rv[PH_MAXB];
#pragma unroll
            for (int bb = 0; bb < PH_MAXB; bb++)
                rv[bb] = bb < nbs ? (double)a.rrs[(long)(row0 + bb) * npix + rp] : 0.0;
            // a perturbed run adds n_sigma R_sigma to every region's values after the no-data
            // test (model/samodel.c:3003-3007)
            if (a.pert && !ph_approx_equal(a.pert[q], 0.0f, 1.0e-6f)) {
                const double ns_ = (double)a.pert[q];
#pragma unroll
                for (int bb = 0; bb < PH_MAXB; bb++)
                    if (bb < nbs) rv[bb] += ns_ * scn_rsig[row0 + bb];
            }
            // Rrs at 440 / 490 / 550 / 640 nm by the scene's own interpolation weights
            // (model/common.c:298-335): ph_rrs_at's operations
            auto rrs_at = [&](int k) {
                const double* w = sc + 2 + 4 * k; // i0, i1, w0, w1
                return ph_sel(rv, (int)w[0]) * w[2] + ph_sel(rv, (int)w[1]) * w[3];
            };
            g[a.o_rm + j] = ph_sel(rv, b);
            double r440 = rrs_at(0);
            const double r490 = rrs_at(1);
            if (r440 < 0.0) r440 = 0.0001;
            const double chi = r440 / r490;
            double Y = 3.44 * (1.0 - 3.17 * PH_EXP(-2.01 * chi));
            if (Y < 0.0) Y = 0.0;
            if (Y > 2.5) Y = 2.5;
            g[a.o_pw + j] = PH_POW(g[a.o_bc + PH_REG_SLOT * qs + 8], Y); // (440 / lambda)^Y
            if (b == 0) {
                const int t = r * nd + s;
                g[a.o_m + 3 * t] = r490;
                g[a.o_m + 3 * t + 1] = rrs_at(2);
                g[a.o_m + 3 * t + 2] = rrs_at(3);
            }
        }
        ph_wave_sync();
        double tot = 0.0;
        for (int j = 0; j < NP; j++) tot += g[a.o_rm + j];
        S.tot = tot;

        // ---- coordinate starts other than depth (model/samodel.c:2246-2354) ----
        double m490 = 0.0;
        for (int t = 0; t < R * nd; t++) m490 += g[a.o_m + 3 * t];
        m490 /= (double)R * nd;
        const int off = R + 2 * NBOT * R;
#pragma unroll
        for (int m = 0; m < MC; m++) {
            const int cc = S.cidx(m);
            double v = 0.0, st = 0.0;
            if (cc >= R && cc < R + R * NBOT) {
                v = 100.0 * 4.0 * m490;
                st = 1.5 * v;
            } else if (cc >= R + R * NBOT && cc < off) {
                v = 1.0;
                st = 0.5 * v;
            } else if (cc >= off && cc < N) {
                const int s = (cc - off) / 3, comp = (cc - off) % 3;
                double a4 = 0.0, b5 = 0.0, c6 = 0.0;
                for (int r = 0; r < R; r++) {
                    a4 += g[a.o_m + 3 * (r * nd + s)];
                    b5 += g[a.o_m + 3 * (r * nd + s) + 1];
                    c6 += g[a.o_m + 3 * (r * nd + s) + 2];
                }
                a4 /= (double)R;
                b5 /= (double)R;
                c6 /= (double)R;
                const double P0 = 100.0 * 0.072 * PH_POW(a4 / b5, -1.7);
                const double G0 = 1.5 * P0;
                const double X0 = 100.0 * 30.0 * c.aw640 * c6;
                v = comp == 0 ? P0 : (comp == 1 ? G0 : X0);
                st = 2.0 * v;
            }
            S.xs0[m] = v;
            S.step0[m] = st;
            S.best[m] = 0.0;
        }
        S.h0 = a.hstart ? (double)a.hstart[q] : -1.0;
        S.kstart = 0;
        S.lowest = 1.0e4;
        S.nstart = 0;
        S.evals = 0;
        S.best_icount = 0;
        S.best_conv = 0;
        S.phase = PH_START;
#pragma unroll
        for (int m = 0; m < PhReg<MC>::YC; m++) S.yv[m] = 0.0;

        // ---- the multi-start Nelder-Mead, one evaluation per pass ----
        PhRegSide side;
        bool timed_out = false;
        for (;;) {
            // the launch arguments read afresh on every pass, as in the other inversion
            // kernels: none of them is held in scalar registers across the loop
            const PhRegionArgs& ka = ph_kernarg<PhRegionArgs>();
            const PhConst& kc = ka.c;
            if (S.phase == PH_START) {
                ph_reg_start_step<MC>(S, kc, S.kstart);
                ph_reg_nm_begin<MC>(S, kc);
            }
            const bool fin = S.phase == PH_FINAL;
            const double f = ph_reg_objective<NBOT, MC>(S, ka, kc, fin, &side);
            ph_reg_consume<MC>(S, kc, f);
            if (S.phase == PH_DONE) break;
            if ((S.evals & 255) == 0 && wall_clock64() - t0 > ka.max_ticks) {
                timed_out = true;
                break;
            }
        }
        if (timed_out) {
            if (li == 0) atomicOr(a.queue + 1, 1ull);
            break;
        }

        // ---- outputs (model/samodel.c:1992-2079) ----
        if (a.x_out) {
#pragma unroll
            for (int m = 0; m < MC; m++)
                if (S.co(m)) a.x_out[(long)S.cidx(m) * a.n + q] = fabs(S.best[m]);
            for (int i = N + li; i < a.n_max; i += 64) a.x_out[(long)i * a.n + q] = 0.0;
        }
        ph_wave_sync();
#pragma unroll
        for (int m = 0; m < MC; m++)
            if (S.co(m)) g[a.o_x + S.cidx(m)] = fabs(S.best[m]);
        ph_wave_sync();
        if (li == 0) {
            const double* X = g + a.o_x;
            // depth: regions weighted 1, the origin sqrt(N_r) (model/samodel.c:2000-2017)
            const double ow = sqrt((double)R);
            double d = 0.0;
            for (int r = 0; r < R; r++) {
                const double H = X[r];
                if (r == origin) d += ow * H;
                else d += H;
            }
            if (a.depth) a.depth[q] = (float)(d / (ow + ((double)R) - 1.0));
            if (a.depth64) a.depth64[q] = d / (ow + ((double)R) - 1.0);
            if (a.rrs_error) a.rrs_error[q] = (float)side.Rrs_error;
            if (a.model_error) a.model_error[q] = (float)side.model_error;
            if (a.albedo) a.albedo[q] = (float)side.albedo;
            if (a.k_min) a.k_min[q] = (float)side.k_min;
            if (a.iod) a.iod[q] = (float)(100.0 * side.iod_sum / side.nobs);
            if (a.b_pct || a.bottom_type) {
                // the origin's bottom percentages, dominant bottom 1-based
                double qsum = 0.0, pct[PH_NBOT];
                for (int k = 0; k < NBOT; k++) qsum += X[R + R * NBOT + origin * NBOT + k];
                for (int m = 0; m < PH_NBOT; m++) pct[m] = 0.0;
                for (int k = 0; k < NBOT; k++) {
                    const int t = c.bottom_idx[k];
                    const double v = 100.0 * X[R + R * NBOT + origin * NBOT + k] / qsum;
                    for (int m = 0; m < PH_NBOT; m++) pct[m] = ph_pick(m == t, v, pct[m]);
                }
                double largest = 0.0, bt = 0.0;
                for (int k = 0; k < NBOT; k++) {
                    const int t = c.bottom_idx[k];
                    double v = 0.0;
                    for (int m = 0; m < PH_NBOT; m++) v = ph_pick(m == t, pct[m], v);
                    if (v > largest) {
                        largest = v;
                        bt = 1.0 + t;
                    }
                }
                if (a.b_pct)
                    for (int m = 0; m < 3; m++) a.b_pct[(long)m * a.n + q] = (float)pct[m];
                if (a.bottom_type) a.bottom_type[q] = (float)bt;
            }
            if (a.status) a.status[q] = S.best_conv ? 0 : 2;
            if (a.n_evals) a.n_evals[q] = S.evals;
            if (a.n_iter) a.n_iter[q] = S.best_icount;
            if (a.n_starts) a.n_starts[q] = S.nstart;
            if (a.n_regions) a.n_regions[q] = R;
        }
    }
}

// The kernels.  Blocks in LDS: one coordinate slot compiled for three waves per SIMD, used
// where the LDS holds three blocks per SIMD (30 parameters: 101 -> 128 k px/s), and two slots
// at one wave per SIMD, where the simplexes fill the LDS.  Simplexes in device memory with up
// to four slots, compiled for several waves per SIMD: the LDS then holds only the small
// arrays, and the other waves hide each one's latency.  Two slots run at three waves (168
// registers, 107 spilled: 25.9 k px/s at two waves, 31.1 k at three, 27.1 k at four on a
// 512^2 scene with n_spatial 2 and 3 bottoms), four slots at two (their spills at three cost
// more than the wave gains); six slots keep their registers unspilled at one.  All take the
// launch arguments as their only argument (ph_kernarg).
template <int NBOT, int MC, bool GM, int TL>
__global__ void __launch_bounds__(64) ph_region_kernel(PhRegionArgs a)
{
    ph_region_body<NBOT, MC, GM, TL>(a);
}

// one coordinate slot, blocks in LDS, compiled for three waves per SIMD
template <int NBOT>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3, 3)))
ph_region_kernel_lds3(PhRegionArgs a)
{
    ph_region_body<NBOT, 1, false, 0>(a);
}

// TL: 1 arrays in LDS, 2 arrays in device memory after the simplex
constexpr int ph_reg_dm_waves(int mc) { return mc <= 2 ? 3 : 2; }
template <int NBOT, int MC, int TL>
__global__ void __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu(ph_reg_dm_waves(MC), ph_reg_dm_waves(MC))))
ph_region_kernel_dm2(PhRegionArgs a)
{
    ph_region_body<NBOT, MC, true, TL>(a);
}

template <int NBOT, int MC, bool GM>
static const void* ph_region_fn(const PhRegionArgs* a)
{
    if constexpr (GM && MC <= 4) {
        if (a->tables_in_lds) return (const void*)ph_region_kernel_dm2<NBOT, MC, 1>;
        return (const void*)ph_region_kernel_dm2<NBOT, MC, 2>;
    } else if constexpr (GM) {
        if (a->tables_in_lds) return (const void*)ph_region_kernel<NBOT, MC, true, 1>;
        return (const void*)ph_region_kernel<NBOT, MC, true, 2>;
    } else if constexpr (MC == 1) {
        return (const void*)ph_region_kernel_lds3<NBOT>;
    } else {
        return (const void*)ph_region_kernel<NBOT, MC, false, 0>;
    }
}

template <int NBOT, int MC, bool GM>
static int ph_region_launch_t(const PhRegionArgs* a, int grid, void* stream)
{
    const size_t lds = GM && !a->tables_in_lds ? 0 : sizeof(double) * (size_t)a->lds_doubles;
    if (lds > 65536) {
        const hipError_t e = hipFuncSetAttribute(ph_region_fn<NBOT, MC, GM>(a),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                                 (int)lds);
        if (e != hipSuccess) return (int)e;
    }
    if constexpr (GM && MC <= 4) {
        if (a->tables_in_lds)
            ph_region_kernel_dm2<NBOT, MC, 1><<<dim3(grid), dim3(64), lds, (hipStream_t)stream>>>(*a);
        else
            ph_region_kernel_dm2<NBOT, MC, 2><<<dim3(grid), dim3(64), lds, (hipStream_t)stream>>>(*a);
    } else if constexpr (GM) {
        if (a->tables_in_lds)
            ph_region_kernel<NBOT, MC, true, 1><<<dim3(grid), dim3(64), lds, (hipStream_t)stream>>>(*a);
        else
            ph_region_kernel<NBOT, MC, true, 2><<<dim3(grid), dim3(64), lds, (hipStream_t)stream>>>(*a);
    } else if constexpr (MC == 1) {
        ph_region_kernel_lds3<NBOT><<<dim3(grid), dim3(64), lds, (hipStream_t)stream>>>(*a);
    } else {
        ph_region_kernel<NBOT, MC, false, 0><<<dim3(grid), dim3(64), lds, (hipStream_t)stream>>>(*a);
    }
    return (int)hipGetLastError();
}

template <int NBOT, int MC, bool GM>
static int ph_region_occ_t(const PhRegionArgs* a, int* bpc)
{
    const size_t lds = GM && !a->tables_in_lds ? 0 : sizeof(double) * (size_t)a->lds_doubles;
    return (int)hipOccupancyMaxActiveBlocksPerMultiprocessor(bpc, ph_region_fn<NBOT, MC, GM>(a),
                                                             64, lds);
}

template <int NBOT, typename Fn>
static int ph_region_dispatch_mc(int mc, bool gm, Fn&& fn)
{
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, 2>;
    using I4 = std::integral_constant<int, 4>;
    using I6 = std::integral_constant<int, 6>;
    using NB = std::integral_constant<int, NBOT>;
    using T = std::true_type;
    using F = std::false_type;
    if (!gm) {
        if (mc == 1) return fn(NB{}, I1{}, F{});
        if (mc == 2) return fn(NB{}, I2{}, F{});
        return PH_LAUNCH_UNSUPPORTED;
    }
    if (mc <= 2) return fn(NB{}, I2{}, T{});
    if (mc <= 4) return fn(NB{}, I4{}, T{});
    if (mc <= 6) return fn(NB{}, I6{}, T{});
    return PH_LAUNCH_UNSUPPORTED;
}

// The instance for a launch: coordinate slots per lane from n_max (64 coordinates a slot), the
// block in device memory when gm is set.
template <typename Fn>
static int ph_region_dispatch(const PhRegionArgs* a, bool gm, Fn&& fn)
{
    const int mc = (a->n_max + 63) / 64;
    switch (a->c.n_bottoms) {
    case 1: return ph_region_dispatch_mc<1>(mc, gm, fn);
    case 2: return ph_region_dispatch_mc<2>(mc, gm, fn);
    case 3: return ph_region_dispatch_mc<3>(mc, gm, fn);
    default: return PH_LAUNCH_UNSUPPORTED;
    }
}

extern "C" int ph_launch_region(const PhRegionArgs* a, int grid, void* stream)
{
    return ph_region_dispatch(a, a->work != nullptr, [&](auto nb, auto m, auto gm) {
        return ph_region_launch_t<decltype(nb)::value, decltype(m)::value, decltype(gm)::value>(
            a, grid, stream);
    });
}

extern "C" int ph_region_occupancy(const PhRegionArgs* a, int device_blocks, int* blocks_per_cu)
{
    return ph_region_dispatch(a, device_blocks != 0, [&](auto nb, auto m, auto gm) {
        return ph_region_occ_t<decltype(nb)::value, decltype(m)::value, decltype(gm)::value>(
            a, blocks_per_cu);
    });
}
