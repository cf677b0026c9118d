// photic_amd -- the group engine for multi-date stacks (gfx950, device only).
//
// A stack with k valid dates is the reference's N_s = k problem (model/samodel.c:1808-1815
// layout, one region): depth H and the bottom shared, P, G, X per date, and a misfit over the
// 4k (date, band) pairs.  The bottom is either the sand/seagrass mix B (NBOT = 0, N = 2 + 3k)
// or the reference's native amplitudes and weights of 1-3 bottoms (N = 1 + 2 NBOT + 3k).  With
// N up to 31 one lane cannot hold a stack's simplex, so a group of G lanes (G = 8, 16 or 32:
// the smallest power of two >= max(N, 4k)) runs one stack:
//
//   lane i < N      owns coordinate i: its start point, centroid, trial point, step and best
//                   value in registers, its column of the simplex in LDS;
//   lane j < 4k     owns pair j = (date j / 4, band j % 4): its measured Rrs, (440 / lambda)^Y,
//                   geometry and band coefficients in registers, and evaluates the forward
//                   model for that pair in every objective evaluation;
//   every lane      runs the scalar state machine of the modified AS 47 Nelder-Mead
//                   (model/asa047.c:10-502) and the multi-start (model/samodel.c:2119-2427)
//                   on identical values, so the group stays convergent without votes.
//
// Cross-lane traffic is two exchanges per evaluation: the trial point's H and bottom and the
// pair's date's P, G, X come from their owner lanes by ds_bpermute, and the pairs' squared residuals
// and attenuations go through a per-stack LDS buffer from which every lane re-forms the misfit
// in the reference's (date, band) order.  Every floating-point operation, and its order, is the
// one of the per-lane stack engine (tests/hostsim), so the results agree bit for bit.
#pragma once

#include "ph_engine.h"

#define PH_DEV __device__ __attribute__((always_inline)) inline

// Lanes per stack: the smallest of 8, 16, 32 that holds the n + 1 vertex values (one per
// lane) and np pairs.
__host__ __device__ constexpr int ph_grp_lanes(int n, int np)
{
    return (n + 1 <= 8 && np <= 8) ? 8 : ((n + 1 <= 16 && np <= 16) ? 16 : 32);
}

// A stack's LDS block, in doubles.  Sub-arrays start at even offsets (16-byte aligned).
template <int ND, int NBOT>
struct PhGrpLayout {
    static constexpr int N = PhLayout<NBOT>::template n<ND>();
    static constexpr int OFF = PhLayout<NBOT>::off;   // first P coordinate
    static constexpr int NBT = PhLayout<NBOT>::nbot;  // bottoms in rho
    static constexpr int NP = 4 * ND;
    static constexpr int G = ph_grp_lanes(N, NP);
    static constexpr int P0 = 0;                        // simplex, [N + 1][N] (lane i: column i)
    static constexpr int Y0 = (N + 1) * N;              // vertex values at a convergence test, [N + 1]
    static constexpr int D0 = Y0 + ((N + 2) & ~1);      // squared residual per pair, [NP]
    static constexpr int K0 = D0 + NP;                  // attenuation K per pair, [NP]
    static constexpr int Q0 = K0 + NP;                  // rB / r per pair, [NP]
    static constexpr int T0 = Q0 + NP;                  // K-penalty term per date, [ND]
    static constexpr int KL = T0 + ND;                  // the last date's Kmin
    static constexpr int n = (KL + 2) & ~1;
    static_assert(N + 1 <= G && NP <= G, "group too small");
};

// Orders the group's LDS accesses around a cross-lane exchange (one wavefront: in-order LDS,
// so only the compiler needs fencing).
PH_DEV void ph_wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// v[b] for a lane-varying band b, by selects (no dynamic indexing of the kernel arguments).
PH_DEV double ph_band_sel(const double* v, int b)
{
    double r = v[0];
    r = ph_pick(b == 1, v[1], r);
    r = ph_pick(b == 2, v[2], r);
    r = ph_pick(b == 3, v[3], r);
    return r;
}

template <int ND, int NBOT>
struct PhGrp {
    using L = PhGrpLayout<ND, NBOT>;
    double* g; // the stack's LDS block
    int li;    // lane index in the group
    int gb;    // the group's first lane in the wavefront
    double yv; // vertex value li (li <= N)

    // coordinate li (li < N)
    double xs, xb, xe, step, best;
    double xs0, step0; // start value and step of the depth starts (li >= 1)
    double h0;         // the stack's own depth start (< 0: the configured list)
    // pair li (li < NP): its band's reflectance of each bottom in rho
    double rrs, pw, ss, sv, a0, a1, aw, bbw, eg;
    double bt[L::NBT];

    // the state machine, identical in every lane of the group
    double tot; // sum of the measured Rrs in (date, band) order
    double ylo, ystar, ynewlo, del, lowest, yr;
    int ilo, ihi, icount, jcount, numres, ifault, phase, sub, kstart, nstart;
    int best_icount, best_conv, evals;

    PH_DEV double p(int k) const { return g[L::P0 + k * L::N + li]; }
    PH_DEV void set_p(int k, double v) const { g[L::P0 + k * L::N + li] = v; }
    // vertex value k (k the same in every lane of the group)
    PH_DEV double y(int k) const { return __shfl(yv, k, L::G); }
    PH_DEV void set_y(int k, double v) { yv = ph_pick(li == k, v, yv); }
    // the group's bits of a wavefront ballot, lane li at bit li
    PH_DEV unsigned long long bits(bool pred) const
    {
        constexpr unsigned long long low = L::G == 64 ? ~0ull : ((1ull << L::G) - 1ull);
        return (__ballot(pred) >> gb) & low;
    }
    // The scan "v = y_0, at = 0; for i = 1..N: if (v < y_i) v = y_i, at = i" (asa047.c:
    // 223-231) as a reduction over the group: no comparison with a NaN holds, so a NaN y_0
    // stays; otherwise it ends at the first index of the largest value.  WANT_MAX false: the
    // same with ">" (the lowest vertex, asa047.c:196-204).  *val = y_at.
    template <bool WANT_MAX>
    PH_DEV int scan(double* val) const
    {
        const double y0 = y(0);
        if (y0 != y0) {
            *val = y0;
            return 0;
        }
        double m = y0;
        if (li <= L::N && (WANT_MAX ? yv > m : yv < m)) m = yv;
#pragma unroll
        for (int off = L::G / 2; off >= 1; off >>= 1) {
            const double o = __shfl_xor(m, off, L::G);
            if (WANT_MAX ? o > m : o < m) m = o;
        }
        const int at = __ffsll((long long)bits(li <= L::N && yv == m)) - 1;
        *val = y(at);
        return at;
    }
};

// Side quantities of the final evaluation (PhStackSide of the per-lane engine).
struct PhGrpSide {
    double k_min, iod_sum, nobs, Rrs_error, K_error, model_error, albedo;
};

// The misfit of the point whose coordinate li is xv (model/samodel.c:2432-2759, one region,
// mix layout, ND dates).  Called by all G lanes of the group together.
template <int ND, int NBOT>
PH_DEV double ph_grp_objective(const PhGrp<ND, NBOT>& S, const PhConst& c, double xv,
                               bool side_on, PhGrpSide* side)
{
    using L = PhGrpLayout<ND, NBOT>;
    constexpr int G = L::G, NBT = L::NBT;
    const int li = S.li;
    const int sd = li < L::NP ? (li >> 2) : 0;
    const double x0 = __shfl(xv, 0, G);
    const double xp = __shfl(xv, L::OFF + 3 * sd, G);
    const double xg = __shfl(xv, L::OFF + 1 + 3 * sd, G);
    const double xx = __shfl(xv, L::OFF + 2 + 3 * sd, G);
    // the bottom: B_k and q_k (the mix: both amplitudes 100, q = (|B|, |1 - B|)), q normalised
    double Bk[NBT], q[NBT], qsum = 0.0;
    if constexpr (NBOT == 0) {
        const double x1 = __shfl(xv, 1, G);
        Bk[0] = 0.01 * fabs(100.0);
        Bk[1] = 0.01 * fabs(100.0);
        q[0] = fabs(fabs(x1));
        q[1] = fabs(fabs(1.0 - x1));
    } else {
#pragma unroll
        for (int k = 0; k < NBT; k++) {
            Bk[k] = 0.01 * fabs(__shfl(xv, 1 + k, G));
            q[k] = fabs(__shfl(xv, 1 + NBT + k, G));
        }
    }
#pragma unroll
    for (int k = 0; k < NBT; k++) qsum += q[k];
#pragma unroll
    for (int k = 0; k < NBT; k++) q[k] /= qsum;
    double qb[NBT], alb = 0.0;
#pragma unroll
    for (int k = 0; k < NBT; k++) {
        qb[k] = q[k] * Bk[k];
        alb += qb[k];
    }
    const double H = fabs(x0);
    const double kmm = 0.275, kmn = 0.185;

    ph_wave_sync(); // the last evaluation's readers are done with the buffers
    if (li < L::NP) {
        const double P = 0.01 * fabs(xp);
        const double Gc = 0.01 * fabs(xg);
        const double X = 0.01 * fabs(xx);
        double rho = 0.0;
#pragma unroll
        for (int k = 0; k < NBT; k++) rho += qb[k] * S.bt[k];
        double K, rB, r;
        const double m = ph_band64_nb(P, Gc, X, rho, H, S.a0, S.a1, S.aw, S.bbw, S.eg, S.pw,
                                      S.ss, S.sv, &K, &rB, &r);
        const double d = m - S.rrs;
        S.g[L::D0 + li] = d * d;
        S.g[L::K0 + li] = K;
        if (side_on) S.g[L::Q0 + li] = rB / r;
    }
    ph_wave_sync();
    // lane s < ND: date s's Kmin over its bands (zeros skipped) and its K-penalty term; the
    // per-date Kmin restarts at 1e4 and the last date's is the one tested after the loop
    if (li < ND) {
        double Kmin = 1.0e4;
#pragma unroll
        for (int b = 0; b < 4; b++) {
            const double K = S.g[L::K0 + 4 * li + b];
            if (!ph_approx_equal((float)K, 0.0f, 1.0e-6f) && K < Kmin) Kmin = K;
        }
        double thr = -1.0;
        if (H < 1.0 && Kmin < kmm) thr = 0.01 + kmn;
        else if (H < 2.0 && Kmin < 0.5 * (1.5 * kmn + 0.5 * kmm)) thr = 0.01 + 0.5 * (1.5 * kmn + 0.5 * kmm);
        else if (H < 3.0 && Kmin < 0.5 * (1.25 * kmn + 0.75 * kmm)) thr = 0.01 + 0.5 * (1.25 * kmn + 0.75 * kmm);
        else if (H < 4.0 && Kmin < 0.5 * (1.5 * kmn + 0.5 * kmm)) thr = 0.01 + 0.5 * (1.5 * kmn + 0.5 * kmm);
        else if (H < 5.0 && Kmin < 0.5 * (1.75 * kmn + 0.25 * kmm)) thr = 0.01 + 0.5 * (1.75 * kmn + 0.25 * kmm);
        // the sum below adds +0.0 for a date without a penalty: exact, since the running
        // penalty is never -0.0
        double term = 0.0;
        if (thr > 0.0) {
            const double t = 1.0 / (0.01 + Kmin) - 1.0 / thr;
            term = 100.0 * (t * t);
        }
        S.g[L::T0 + li] = term;
        if (li == ND - 1) S.g[L::KL] = Kmin;
    }
    ph_wave_sync();
    double err = 0.0;
#pragma unroll
    for (int j = 0; j < L::NP; j++) err += S.g[L::D0 + j];
    double Kerr = 0.0;
#pragma unroll
    for (int s = 0; s < ND; s++) Kerr += S.g[L::T0 + s];
    const double Kmin = S.g[L::KL];
    const double nt = (double)L::NP;
    const double mean = S.tot / nt;
    const double e_rrs = 100.0 * sqrt(err / nt) / mean;
    if (Kmin > 0.7) {
        const double t = 4.0 * (Kmin - 0.7);
        Kerr += 100.0 * (t * t);
    }
    double e = 80.0 * e_rrs + 15.0 * Kerr;
    e /= 120.0;
    if (side_on) {
        double kmin_all = 1.0e10, iod = 0.0, nobs = 0.0;
        for (int j = 0; j < L::NP; j++) {
            const double K = S.g[L::K0 + j];
            if (!ph_approx_equal((float)K, 0.0f, 1.0e-4f) && K < kmin_all) kmin_all = K;
            iod += S.g[L::Q0 + j];
            nobs += 1.0;
        }
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

// Start point and step of coordinate li for depth start k (k < 0: seed_i), model/samodel.c:
// 2246-2354: H = the start depth (step 1.25 H) -- the stack's own when it has one; the bottom,
// P, G, X from xs0 / step0.
template <int ND, int NBOT>
PH_DEV void ph_grp_start_step(PhGrp<ND, NBOT>& S, const PhConst& c, int k, double seed_i)
{
    constexpr int OFF = PhGrpLayout<ND, NBOT>::OFF;
    const int li = S.li;
    if (k < 0) {
        const int cls = li == 0 ? 0 : (li < OFF ? 1 : 2 + (li - OFF) % 3);
        double rel = c.seed_rel_step[0], mn = c.seed_min_step[0];
        ph_for<4>([&](auto m0) PH_BODY {
            constexpr int m = m0 + 1;
            rel = ph_pick(cls == m, c.seed_rel_step[m], rel);
            mn = ph_pick(cls == m, c.seed_min_step[m], mn);
        });
        double st = fabs(seed_i) * rel;
        if (st < mn) st = mn;
        S.xs = seed_i;
        S.step = st;
        return;
    }
    double h0 = c.hstart[0];
    ph_for<PH_MAXH - 1>([&](auto m0) PH_BODY {
        constexpr int m = m0 + 1;
        h0 = ph_pick(k == m, c.hstart[m], h0);
    });
    if (S.h0 >= 0.0) h0 = S.h0;
    S.xs = li == 0 ? h0 : S.xs0;
    S.step = li == 0 ? 1.25 * h0 : S.step0;
}

// Begin a Nelder-Mead run from xs and the steps (asa047.c:162-170).
template <int ND, int NBOT>
PH_DEV void ph_grp_nm_begin(PhGrp<ND, NBOT>& S, const PhConst& c)
{
    S.icount = 0;
    S.numres = 0;
    S.ihi = 0;
    S.ilo = 0;
    S.ylo = 0.0;
    S.yr = 0.0;
    S.jcount = c.konvge;
    S.del = 1.0;
    S.phase = PH_BUILD;
    S.sub = PhGrpLayout<ND, NBOT>::N;
    S.xe = S.xs;
}

// One step of the state machine: consume f, the value of the point in xe in phase S.phase,
// and leave the next point to evaluate in xe (or PH_START / PH_DONE) -- the per-lane stack
// engine's step, with each loop over coordinates done by the coordinate's lane.
template <int ND, int NBOT>
PH_DEV void ph_grp_consume(PhGrp<ND, NBOT>& S, const PhConst& c, double f)
{
    constexpr int N = PhGrpLayout<ND, NBOT>::N;
    const int li = S.li;
    const bool co = li < N;
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

    // Two forms of the same step, bit for bit alike.  With 32 lanes per stack (two stacks per
    // wave) the classification and the follow-ups that only set the next point run as
    // selects; with 8 or 16 lanes the branched form keeps fewer registers live and measured
    // faster (2048^2 x 8 dates: 3-4 date stacks 3-7% slower as selects, 7-8 dates 6-7% faster).
    int act;
    if constexpr (PhGrpLayout<ND, NBOT>::G >= 32) {
        // ---- classify (asa047.c:176-393, 455-479), as selects: the groups of a wave sit in
        // different phases, so branches would all run anyway ----
        const bool bB = ph == PH_BUILD, bR = ph == PH_REFLECT, bE = ph == PH_EXPAND;
        const bool bCI = ph == PH_CONTRACT_IN, bCO = ph == PH_CONTRACT_OUT, bS = ph == PH_SHRINK;
        const bool bF = ph >= PH_FACT_PLUS; // PH_FACT_PLUS or PH_FACT_MINUS
        const int l = __popcll(S.bits(li <= N && f < S.yv));
        const bool r_lo = f < S.ylo;
        const bool ci_shrink = S.y(S.ihi) < f;
        const bool fact_fail = ph_trunc10(f) < S.yr;
        const int bnext = (S.sub == N) ? 0 : S.sub + 1;
        const int snext = S.sub + 1;
        if (bR) S.ystar = f;
        const bool wr = bB || (bR && !r_lo && l > 1) || bE || (bCI && !ci_shrink) || bCO || bS;
        const int wj = (bB || bS) ? S.sub : S.ihi;
        const bool wr_xs = bR || (bE && S.ystar < f) || (bCO && !(f <= S.ystar));
        const double wf = ph_pick(wr_xs, S.ystar, f);
        act = A_ITER_END; // bR (l > 1), bE, bCO, bCI without a shrink
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

        // ---- the vertex write (every lane stores the same vertex value) ----
        if (wr) {
            if (co) S.set_p(wj, ph_pick(wr_xs, S.xs, S.xe));
            S.set_y(wj, wf);
        }

        // ---- follow-ups that only set the next point (asa047.c:174-194, 263-393, 455-493), as
        // selects: BUILD_NEXT, the start plus step * del on coordinate sub; MOVE, xb + coef (v -
        // xb), v the worst vertex or the reflection; SHRINK_NEXT, (vertex sub + best vertex) / 2;
        // FACT_FAIL, a restart from the probe point with a simplex of relative size eps;
        // FACT_MINUS / FACT_NEXT, the probe point moved by step * eps on coordinates i0 and i1 ----
        const bool fBN = act == A_BUILD_NEXT, fMV = act == A_MOVE, fSN = act == A_SHRINK_NEXT;
        const bool fFF = act == A_FACT_FAIL, fFM = act == A_FACT_MINUS, fFN = act == A_FACT_NEXT;
        if (fBN || fMV || fSN || fFF || fFM || fFN) {
            double a = 0.0, b = 0.0;
            if (co && (fMV || fSN)) {
                a = S.p(fSN ? S.sub : S.ihi);
                b = S.p(S.ilo);
            }
            const int i0 = S.sub;
            const int i1 = fFM ? -1 : i0 + 1;
            const double xs = S.xs, xb = S.xb;
            const double bn = (i0 == li) ? xs + S.step * S.del : +xs;
            const double v = mv_phi ? a : xs;
            const double mv = xb + mv_coef * (v - xb);
            const double sn = (a + b) * 0.5;
            const double d = S.step * eps;
            double fv = xb;
            if (i0 == li) fv = fFM ? fv - d - d : fv + d;
            if (i1 == li) fv = fv + d;
            double xe = fBN ? bn : (fMV ? mv : (fSN ? sn : (fFF ? xb : fv)));
            if ((fMV || fSN) && !co) xe = S.xe; // those two set coordinate lanes only
            S.xe = xe;
            if (fFF) S.xs = xb;
            if (fFM || fFN) S.xb = fv;
            const bool run_done = fFN && i1 >= N;
            const int nph = fBN || fFF ? (int)PH_BUILD
                                       : (fMV ? mv_phase
                                              : (fSN ? (int)PH_SHRINK
                                                     : (fFM ? (int)PH_FACT_MINUS : (int)PH_FACT_PLUS)));
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
    } else {
        // ---- classify (asa047.c:176-393, 455-479) ----
        bool wr = false, wr_xs = false;
        int wj = 0;
        double wf = f;
        act = A_NONE;
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
                const int l = __popcll(S.bits(li <= N && f < S.yv));
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

        // ---- the vertex write (every lane stores the same vertex value) ----
        if (wr) {
            if (co) S.set_p(wj, ph_pick(wr_xs, S.xs, S.xe));
            S.set_y(wj, wf);
        }

        // ---- follow-up ----
        if (act == A_BUILD_NEXT) {
            const int j = S.sub;
            S.xe = (j == li) ? S.xs + S.step * S.del : +S.xs;
            S.phase = PH_BUILD;
            return;
        }
        if (act == A_MOVE) {
            if (co) {
                const double v = mv_phi ? S.p(S.ihi) : +S.xs;
                const double b = S.xb;
                S.xe = b + mv_coef * (v - b);
            }
            S.phase = mv_phase;
            return;
        }
        if (act == A_SHRINK_NEXT) {
            if (co) S.xe = (S.p(S.sub) + S.p(S.ilo)) * 0.5;
            S.phase = PH_SHRINK;
            return;
        }
        if (act == A_FACT_FAIL) {
            // asa047.c:488-493: restart from the probe point with a simplex of relative size eps
            S.ifault = 2;
            S.xs = S.xb;
            S.xe = S.xb;
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
            const double d = S.step * eps;
            double v = S.xb;
            if (i0 == li) v = minus ? v - d - d : v + d;
            if (i1 == li) v = v + d;
            S.xb = v;
            S.xe = v;
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
            // the sums in vertex order, every lane alike, from the group's values in LDS
            using L = PhGrpLayout<ND, NBOT>;
            ph_wave_sync();
            if (li <= N) S.g[L::Y0 + li] = S.yv;
            ph_wave_sync();
            double z = 0.0;
            ph_for<N + 1>([&](auto i) PH_BODY { z = z + S.g[L::Y0 + i]; });
            const double xm = ph_div_const(z, dnn, 1.0 / dnn);
            z = 0.0;
            ph_for<N + 1>([&](auto i) PH_BODY {
                const double yi = S.g[L::Y0 + i];
                z = z + (yi - xm) * (yi - xm);
            });
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
            S.ihi = S.template scan<true>(&S.ynewlo);
            if (co) {
                // centroid without the worst vertex: z = sum over vertices k = 0..N in k
                // order, less the worst vertex's phi
                const double phi = S.p(S.ihi);
                double z = 0.0;
                ph_for<N + 1>([&](auto k) PH_BODY { z = z + S.p(k); });
                const double zi = z - phi;
                const double b = ph_div_const(zi, dn, 1.0 / dn);
                const double r = b + 1.0 * (b - phi);
                S.xb = b;
                S.xs = r;
                S.xe = r;
            }
            S.phase = PH_REFLECT;
            return;
        }
    }
    // ---- convergence or evaluations exhausted: asa047.c:440-454 ----
    if (nm_finish) {
        const int lo = S.ilo;
        if (co) S.xb = S.p(lo);
        S.ynewlo = S.y(lo);
        S.yr = ph_trunc10(S.ynewlo);
        if (c.kcount < S.icount) {
            S.ifault = 2;
            act = A_RUN_DONE;
        } else {
            S.ifault = 0;
            S.sub = 0;
            if (li == 0) S.xb = S.xb + S.step * eps;
            S.xe = S.xb;
            S.phase = PH_FACT_PLUS;
            return;
        }
    }
    // ---- a run ended; the multi-start bookkeeping, model/samodel.c:2385-2405 ----
    if (act == A_RUN_DONE) {
        S.nstart++;
        if (S.ynewlo < S.lowest) {
            S.lowest = S.ynewlo;
            S.best = S.xb;
            S.best_icount = S.icount;
            S.best_conv = (S.ifault == 0) ? 1 : 0;
            if (S.lowest < c.early_exit) S.kstart = c.n_hstarts;
        }
        S.kstart++;
        if (S.kstart < c.n_hstarts) {
            S.phase = PH_START;
            return;
        }
        S.xe = S.best;
        S.phase = PH_FINAL;
    }
}
