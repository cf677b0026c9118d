// Test infrastructure: the per-lane engine for multi-date stacks, on the host.
//
// The state machine of the product's ph_engine.h (one objective evaluation per step; the
// modified AS 47 Nelder-Mead of model/asa047.c:10-502 and the multi-start of
// model/samodel.c:2119-2427) for the mix layout over ND dates, [H, B, (P, G, X) per date],
// N = 2 + 3 ND parameters, one stack at a time.  Built by build.sh with the host C library's
// exp/log/pow it must match the FP64 oracle bit for bit; with the correctly rounded functions
// of ph_crmath.h it is the reference the GPU's group engine (photic_amd/csrc/ph_grpengine.h)
// is held to bit for bit.  The simplex, steps, best point and trial points live in a
// caller-supplied array (element e at base[e * stride]); the vertex values and the scalars of
// the state machine are members.  Never built into the library.
#pragma once

#include "ph_engine.h"

// Element e of the stack's arrays lives at base[e * stride].
template <int N>
struct PhStackNM {
    static constexpr int P0 = 0;                 // vertices, [N + 1][N]
    static constexpr int STEP0 = (N + 1) * N;    // steps
    static constexpr int BEST0 = STEP0 + N;      // best point over the starts
    static constexpr int XS0 = BEST0 + N;        // start point / reflected point
    static constexpr int XB0 = XS0 + N;          // centroid / local-minimum probe
    static constexpr int XE0 = XB0 + N;          // point being evaluated
    static constexpr int ND = (N - 2) / 3;       // dates
    static constexpr int RRS0 = XE0 + N;         // measured Rrs, [ND][4]
    static constexpr int PW0 = RRS0 + 4 * ND;    // (440 / lambda)^Y, [ND][4]
    static constexpr int SS0 = PW0 + 4 * ND;     // sec(sun), [ND]
    static constexpr int SV0 = SS0 + ND;         // sec(view), [ND]
    static constexpr int n_slots = SV0 + ND;

    double* base;
    long stride;
    PH_INLINE double ld(int e) const { return base[(long)e * stride]; }
    PH_INLINE void st(int e, double v) const { base[(long)e * stride] = v; }
    PH_INLINE double p(int k, int i) const { return ld(P0 + k * N + i); }
    PH_INLINE void set_p(int k, int i, double v) const { st(P0 + k * N + i, v); }
    PH_INLINE double step(int i) const { return ld(STEP0 + i); }
    PH_INLINE void set_step(int i, double v) const { st(STEP0 + i, v); }
    PH_INLINE double best(int i) const { return ld(BEST0 + i); }
    PH_INLINE void set_best(int i, double v) const { st(BEST0 + i, v); }
    PH_INLINE double xs(int i) const { return ld(XS0 + i); }
    PH_INLINE void set_xs(int i, double v) const { st(XS0 + i, v); }
    PH_INLINE double xb(int i) const { return ld(XB0 + i); }
    PH_INLINE void set_xb(int i, double v) const { st(XB0 + i, v); }
    PH_INLINE double xe(int i) const { return ld(XE0 + i); }
    PH_INLINE void set_xe(int i, double v) const { st(XE0 + i, v); }
    PH_INLINE double rrs(int s, int b) const { return ld(RRS0 + 4 * s + b); }
    PH_INLINE double pw(int s, int b) const { return ld(PW0 + 4 * s + b); }
    PH_INLINE double ss(int s) const { return ld(SS0 + s); }
    PH_INLINE double sv(int s) const { return ld(SV0 + s); }

    double yv[N + 1]; // vertex values (registers: every iteration scans them)
    double tot;       // sum of measured Rrs in (date, band) order
    double ylo, ystar, ynewlo, del, lowest, yr;
    int ilo, ihi, icount, jcount, numres, ifault, phase, sub, kstart, nstart;
    int best_icount, best_conv, evals;
};

// Rrs of date s at reference wavelength k (0 = 440, 1 = 490, 2 = 550, 3 = 640 nm), the
// reference's interpolation with host-resolved segment and weights (model/common.c:298-335).
template <int N>
PH_INLINE double ph_stack_rrs_at(const PhStackNM<N>& S, const PhConst& c, int s, int k)
{
    return S.rrs(s, c.ip_i0[k]) * c.ip_w0[k] + S.rrs(s, c.ip_i1[k]) * c.ip_w1[k];
}

// Pixel set-up (model/samodel.c:1785-1799, 2898-2903) into the lane's memory: rrs_in [ND][4]
// (float), the dates' sec(sun), sec(view).  False when any value is no-data or negative
// (model/samodel.c:933-947).
template <int ND>
PH_INLINE bool ph_stack_pixel_init(PhStackNM<2 + 3 * ND>& S, const PhConst& c,
                                   const float (&rrs_in)[ND][4], const double* ss,
                                   const double* sv)
{
    bool ok = true;
    double tot = 0.0;
    for (int s = 0; s < ND; s++) {
        double r[4];
#pragma unroll
        for (int b = 0; b < 4; b++) {
            float v = rrs_in[s][b];
            if (ph_approx_equal(v, (float)c.nodata, 1.0e-6f) || v < 0.0f) ok = false;
            r[b] = (double)v;
            S.st(PhStackNM<2 + 3 * ND>::RRS0 + 4 * s + b, r[b]);
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
        for (int b = 0; b < 4; b++)
            S.st(PhStackNM<2 + 3 * ND>::PW0 + 4 * s + b, PH_POW(440.0 / c.wl[b], Y));
        S.st(PhStackNM<2 + 3 * ND>::SS0 + s, ss[s]);
        S.st(PhStackNM<2 + 3 * ND>::SV0 + s, sv[s]);
    }
    S.tot = tot;
    return ok;
}

// Side quantities of an evaluation, aggregated as the outputs need them.
struct PhStackSide {
    double k_min;            // min K over dates and bands, zeros skipped (samodel.c:1994)
    double iod_sum, nobs;    // sum of rrs_B / rrs in (date, band) order
    double Rrs_error, K_error, model_error, albedo;
};

// The misfit of the point in the lane's xe slots (model/samodel.c:2432-2759, one region, mix
// layout, ND dates), with the arithmetic and order of ph_objective.
template <int ND, bool SIDE>
PH_INLINE double ph_stack_objective(const PhStackNM<2 + 3 * ND>& S, const PhConst& c, int xbase,
                                    PhStackSide* side)
{
    constexpr int NB = 4;
    const double x0 = S.ld(xbase), x1 = S.ld(xbase + 1);
    double B0 = 0.01 * fabs(100.0), B1 = 0.01 * fabs(100.0);
    double q0 = fabs(fabs(x1)), q1 = fabs(fabs(1.0 - x1));
    double qsum = 0.0;
    qsum += q0;
    qsum += q1;
    q0 /= qsum;
    q1 /= qsum;
    const double qb0 = q0 * B0, qb1 = q1 * B1;
    double alb = 0.0;
    alb += qb0;
    alb += qb1;
    const double H = fabs(x0);
    double err = 0.0;
    const double kmm = 0.275, kmn = 0.185;
    double Kerr = 0.0, Kmin = 1.0e4;
    double kmin_all = 1.0e10, iod = 0.0, nobs = 0.0;
    for (int s = 0; s < ND; s++) {
        const double P = 0.01 * fabs(S.ld(xbase + 2 + 3 * s));
        const double G = 0.01 * fabs(S.ld(xbase + 3 + 3 * s));
        const double X = 0.01 * fabs(S.ld(xbase + 4 + 3 * s));
        const double ss = S.ss(s), sv = S.sv(s);
        Kmin = 1.0e4;
#pragma unroll
        for (int b = 0; b < NB; b++) {
            double rho = 0.0;
            rho += qb0 * c.bot[0][b];
            rho += qb1 * c.bot[1][b];
            double K, rB, r;
            double m = ph_band64(P, G, X, rho, H, c.a0[b], c.a1[b], c.aw[b], c.bbw[b], c.eg[b],
                                 S.pw(s, b), ss, sv, &K, &rB, &r);
            double d = m - S.rrs(s, b);
            err += d * d;
            if (!ph_approx_equal((float)K, 0.0f, 1.0e-6f) && K < Kmin) Kmin = K;
            if (SIDE) {
                if (!ph_approx_equal((float)K, 0.0f, 1.0e-4f) && K < kmin_all) kmin_all = K;
                iod += rB / r;
                nobs += 1.0;
            }
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
    const double nt = (double)(ND * NB);
    double mean = S.tot / nt;
    double e_rrs = 100.0 * sqrt(err / nt) / mean;
    if (Kmin > 0.7) {
        double t = 4.0 * (Kmin - 0.7);
        Kerr += 100.0 * (t * t);
    }
    double e = 80.0 * e_rrs + 15.0 * Kerr;
    e /= 120.0;
    if (SIDE) {
        side->k_min = kmin_all == 1.0e10 ? 0.0 : kmin_all;
        side->iod_sum = iod;
        side->nobs = nobs;
        side->Rrs_error = e_rrs;
        side->K_error = Kerr;
        side->model_error = e;
        side->albedo = alb;
    }
    return e;
}

// Start point and steps of depth start k (k < 0: the seed), model/samodel.c:2246-2354, mix
// layout: H = the start depth (step 1.25 H), B = 0.5 (step 0.5), and per date P from the
// 490/550 ratio, G = 1.5 P, X from a_w(640) Rrs(640), steps twice the start.
template <int ND>
PH_INLINE void ph_stack_start_step(PhStackNM<2 + 3 * ND>& S, const PhConst& c, int k,
                                   const double* seed)
{
    constexpr int N = 2 + 3 * ND;
    if (k < 0) {
        for (int i = 0; i < N; i++) {
            const int cls = i == 0 ? 0 : (i < 2 ? 1 : 2 + (i - 2) % 3);
            double st = fabs(seed[i]) * c.seed_rel_step[cls];
            if (st < c.seed_min_step[cls]) st = c.seed_min_step[cls];
            S.set_xs(i, seed[i]);
            S.set_step(i, st);
        }
        return;
    }
    double h0 = c.hstart[0];
    ph_for<PH_MAXH - 1>([&](auto m0) PH_BODY {
        constexpr int m = m0 + 1;
        h0 = ph_pick(k == m, c.hstart[m], h0);
    });
    S.set_xs(0, h0);
    S.set_step(0, 1.25 * h0);
    S.set_xs(1, 0.5);
    S.set_step(1, 0.5);
    for (int s = 0; s < ND; s++) {
        const int o = 2 + 3 * s;
        double m490 = ph_stack_rrs_at(S, c, s, 1) / (double)1;
        double m550 = ph_stack_rrs_at(S, c, s, 2) / (double)1;
        double m640 = ph_stack_rrs_at(S, c, s, 3) / (double)1;
        const double P0 = 100.0 * 0.072 * PH_POW(m490 / m550, -1.7);
        const double G0 = 1.5 * P0;
        const double X0 = 100.0 * 30.0 * c.aw640 * m640;
        S.set_xs(o, P0);
        S.set_xs(o + 1, G0);
        S.set_xs(o + 2, X0);
        S.set_step(o, 2.0 * P0);
        S.set_step(o + 1, 2.0 * G0);
        S.set_step(o + 2, 2.0 * X0);
    }
}

// Begin a Nelder-Mead run from xs and the steps (asa047.c:162-170).
template <int N>
PH_INLINE void ph_stack_nm_begin(PhStackNM<N>& S, const PhConst& c)
{
    S.icount = 0;
    S.numres = 0;
    S.jcount = c.konvge;
    S.del = 1.0;
    S.phase = PH_BUILD;
    S.sub = N;
    for (int i = 0; i < N; i++) S.set_xe(i, S.xs(i));
}

// One step of the state machine: consume f, the value of xe in phase S.phase, and leave the
// next point to evaluate in xe (or PH_START / PH_DONE).  The logic, step by step, of
// ph_nm_consume in ph_engine.h.
template <int N>
PH_INLINE void ph_stack_nm_consume(PhStackNM<N>& S, const PhConst& c, double f)
{
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
            ph_for<N + 1>([&](auto i) PH_BODY { l += (f < S.yv[i]) ? 1 : 0; });
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
        if (ph_sel(S.yv, S.ihi) < f) {
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
        for (int i = 0; i < N; i++) S.set_p(wj, i, wr_xs ? S.xs(i) : S.xe(i));
        ph_for<N + 1>([&](auto k) PH_BODY { S.yv[k] = ph_pick(wj == (int)k, wf, S.yv[k]); });
    }

    // ---- follow-up ----
    if (act == A_BUILD_NEXT) {
        const int j = S.sub;
        for (int i = 0; i < N; i++) S.set_xe(i, (j == i) ? S.xs(i) + S.step(i) * S.del : +S.xs(i));
        S.phase = PH_BUILD;
        return;
    }
    if (act == A_MOVE) {
        for (int i = 0; i < N; i++) {
            const double v = mv_phi ? S.p(S.ihi, i) : +S.xs(i);
            const double b = S.xb(i);
            S.set_xe(i, b + mv_coef * (v - b));
        }
        S.phase = mv_phase;
        return;
    }
    if (act == A_SHRINK_NEXT) {
        const int j = S.sub, lo = S.ilo;
        for (int i = 0; i < N; i++) S.set_xe(i, (S.p(j, i) + S.p(lo, i)) * 0.5);
        S.phase = PH_SHRINK;
        return;
    }
    if (act == A_FACT_FAIL) {
        // asa047.c:488-493: restart from the probe point with a simplex of relative size eps
        S.ifault = 2;
        for (int i = 0; i < N; i++) {
            const double v = S.xb(i);
            S.set_xs(i, v);
            S.set_xe(i, v);
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
        for (int i = 0; i < N; i++) {
            const double d = S.step(i) * eps;
            double v = S.xb(i);
            if (i0 == i) v = minus ? v - d - d : v + d;
            if (i1 == i) v = v + d;
            S.set_xb(i, v);
            S.set_xe(i, v);
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
        S.ylo = S.yv[0];
        S.ilo = 0;
        ph_for<N>([&](auto i0) PH_BODY {
            constexpr int i = i0 + 1;
            if (S.yv[i] < S.ylo) {
                S.ylo = S.yv[i];
                S.ilo = i;
            }
        });
        iter_begin = true;
    }
    // ---- end of an iteration: asa047.c:397-434 ----
    if (act == A_ITER_END) {
        const double yhi = ph_sel(S.yv, S.ihi);
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
            ph_for<N + 1>([&](auto i) PH_BODY { z = z + S.yv[i]; });
            const double xm = ph_div_const(z, dnn, 1.0 / dnn);
            z = 0.0;
            ph_for<N + 1>([&](auto i) PH_BODY { z = z + (S.yv[i] - xm) * (S.yv[i] - xm); });
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
            S.ynewlo = S.yv[0];
            S.ihi = 0;
            ph_for<N>([&](auto i0) PH_BODY {
                constexpr int i = i0 + 1;
                if (S.ynewlo < S.yv[i]) {
                    S.ynewlo = S.yv[i];
                    S.ihi = i;
                }
            });
            const int hi = S.ihi;
            // centroid: z_i = sum over vertices k = 0..N of p(k, i), in k order, accumulated
            // a row at a time; phi_i = p(hi, i) taken on the way
            double z[N], phi[N];
            ph_for<N>([&](auto i) PH_BODY {
                z[i] = 0.0;
                phi[i] = 0.0;
            });
#pragma unroll 2
            for (int k = 0; k <= N; k++) {
                ph_for<N>([&](auto i) PH_BODY {
                    const double v = S.p(k, i);
                    z[i] = z[i] + v;
                    phi[i] = ph_pick(k == hi, v, phi[i]);
                });
            }
            ph_for<N>([&](auto i) PH_BODY {
                const double zi = z[i] - phi[i];
                const double b = ph_div_const(zi, dn, 1.0 / dn);
                const double r = b + 1.0 * (b - phi[i]);
                S.set_xb(i, b);
                S.set_xs(i, r);
                S.set_xe(i, r);
            });
            S.phase = PH_REFLECT;
            return;
        }
    }
    // ---- convergence or evaluations exhausted: asa047.c:440-454 ----
    if (nm_finish) {
        const int lo = S.ilo;
        for (int i = 0; i < N; i++) S.set_xb(i, S.p(lo, i));
        S.ynewlo = ph_sel(S.yv, lo);
        S.yr = ph_trunc10(S.ynewlo);
        if (c.kcount < S.icount) {
            S.ifault = 2;
            act = A_RUN_DONE;
        } else {
            S.ifault = 0;
            S.sub = 0;
            S.set_xb(0, S.xb(0) + S.step(0) * eps);
            for (int i = 0; i < N; i++) S.set_xe(i, S.xb(i));
            S.phase = PH_FACT_PLUS;
            return;
        }
    }
    // ---- a run ended; the multi-start bookkeeping, model/samodel.c:2385-2405 ----
    if (act == A_RUN_DONE) {
        S.nstart++;
        if (S.ynewlo < S.lowest) {
            S.lowest = S.ynewlo;
            for (int i = 0; i < N; i++) S.set_best(i, S.xb(i));
            S.best_icount = S.icount;
            S.best_conv = (S.ifault == 0) ? 1 : 0;
            if (S.lowest < c.early_exit) S.kstart = c.n_hstarts;
        }
        S.kstart++;
        if (S.kstart < c.n_hstarts) {
            S.phase = PH_START;
            return;
        }
        for (int i = 0; i < N; i++) S.set_xe(i, S.best(i));
        S.phase = PH_FINAL;
    }
}
