// photic_amd -- the per-pixel outputs of a single-date inversion (model/samodel.c:1992-2079)
// and the defaults of a pixel that is not modelled (:819-841), written at pixel `pix` of a
// launch's output arrays (any of them may be NULL).  Shared by the inversion kernel and the
// ordered inversion's kernels.
#pragma once

#include "ph_engine.h"

template <typename FT, int NB, int NBOT>
__device__ __attribute__((always_inline)) inline void ph_write_outputs(const PhInvertArgs& a, const PhConst& c, long pix,
                                        const PhNM<PhLayout<NBOT>::template n<1>()>& S,
                                        const PhSide<1, NB>& side)
{
    using L = PhLayout<NBOT>;
    constexpr int N = L::template n<1>();
    double x[N];
#pragma unroll
    for (int i = 0; i < N; i++) x[i] = fabs(S.m.best(i));
    if (a.x_out)
#pragma unroll
        for (int i = 0; i < N; i++) a.x_out[(long)i * a.n_pix + pix] = x[i];
    // one region: origin weight sqrt(1), depth = H (model/samodel.c:2000-2017)
    if (a.depth) a.depth[pix] = (float)((0.0 + 1.0 * x[0]) / (1.0 + 1.0 - 1.0));
    if (a.rrs_error) a.rrs_error[pix] = (float)side.Rrs_error;
    if (a.model_error) a.model_error[pix] = (float)side.model_error;
    if (a.albedo) a.albedo[pix] = (float)side.albedo;
    if (a.k_min) {
        double kmin = 1.0e10;
#pragma unroll
        for (int b = 0; b < NB; b++) {
            double k = side.K[0][b];
            if (!ph_approx_equal((float)k, 0.0f, 1.0e-4f) && k < kmin) kmin = k;
        }
        a.k_min[pix] = (float)(kmin == 1.0e10 ? 0.0 : kmin);
    }
    if (a.iod) {
        double iod = 0.0, nobs = 0.0;
#pragma unroll
        for (int b = 0; b < NB; b++) {
            iod += side.rrs_bot[0][b] / side.rrs_mod[0][b];
            nobs += 1.0;
        }
        a.iod[pix] = (float)(100.0 * iod / nobs);
    }
    if (a.b_pct || a.bottom_type) {
        constexpr int nb = L::nbot;
        double qv[nb], qsum = 0.0, pct[PH_NBOT];
        if (NBOT == 0) {
            qv[0] = fabs(S.m.best(1));
            qv[nb - 1] = fabs(1.0 - S.m.best(1));
        } else {
#pragma unroll
            for (int k = 0; k < nb; k++) qv[k] = x[1 + nb + k];
        }
#pragma unroll
        for (int k = 0; k < nb; k++) qsum += qv[k];
#pragma unroll
        for (int k = 0; k < PH_NBOT; k++) pct[k] = 0.0;
        double largest = 0.0, btype = 0.0;
#pragma unroll
        for (int k = 0; k < nb; k++) {
            int t = NBOT == 0 ? k : c.bottom_idx[k];
            double v = 100.0 * qv[k] / qsum;
#pragma unroll
            for (int m = 0; m < PH_NBOT; m++) pct[m] = ph_pick(m == t, v, pct[m]);
        }
#pragma unroll
        for (int k = 0; k < nb; k++) {
            int t = NBOT == 0 ? k : c.bottom_idx[k];
            double v = 0.0;
#pragma unroll
            for (int m = 0; m < PH_NBOT; m++) v = ph_pick(m == t, pct[m], v);
            if (v > largest) {
                largest = v;
                btype = 1.0 + t;
            }
        }
        if (a.b_pct)
#pragma unroll
            for (int m = 0; m < 3; m++) a.b_pct[(long)m * a.n_pix + pix] = (float)pct[m];
        if (a.bottom_type) a.bottom_type[pix] = (float)btype;
    }
    if (a.rrs_model)
#pragma unroll
        for (int b = 0; b < NB; b++) {
            // the above-surface Rrs the objective formed from r (ph_band64_tail)
            const double r = side.rrs_mod[0][b];
            a.rrs_model[(long)b * a.n_pix + pix] = 0.5 * r / (1.0 - 1.5 * r) + 0.0;
        }
    if (a.status) a.status[pix] = S.best_conv ? 0 : 2;
    if (a.n_evals) a.n_evals[pix] = S.evals;
    if (a.n_iter) a.n_iter[pix] = S.best_icount;
    if (a.n_starts) a.n_starts[pix] = S.nstart;
}

__device__ inline void ph_write_nodata(const PhInvertArgs& a, long pix, int n_params)
{
    // the reference's defaults for pixels it does not model (model/samodel.c:819-841)
    if (a.depth) a.depth[pix] = 0.0f;
    if (a.rrs_error) a.rrs_error[pix] = 0.0f;
    if (a.model_error) a.model_error[pix] = 0.0f;
    if (a.albedo) a.albedo[pix] = 0.0f;
    if (a.k_min) a.k_min[pix] = 0.0f;
    if (a.iod) a.iod[pix] = 0.0f;
    if (a.b_pct)
        for (int m = 0; m < 3; m++) a.b_pct[(long)m * a.n_pix + pix] = -9999.0f;
    if (a.bottom_type) a.bottom_type[pix] = -9999.0f;
    if (a.x_out)
        for (int i = 0; i < n_params; i++) a.x_out[(long)i * a.n_pix + pix] = 0.0;
    if (a.rrs_model)
        for (int b = 0; b < a.c.n_bands; b++) a.rrs_model[(long)b * a.n_pix + pix] = 0.0;
    if (a.status) a.status[pix] = -1;
    if (a.n_evals) a.n_evals[pix] = 0;
    if (a.n_iter) a.n_iter[pix] = 0;
    if (a.n_starts) a.n_starts[pix] = 0;
}
