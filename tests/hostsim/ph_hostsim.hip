// Test infrastructure: the product's lane engine (photic_amd/csrc/ph_engine.h) compiled for
// the host and run one lane at a time, so the CPU test suite can check the exact state
// machine that the GPU kernel runs -- in its FP64 instantiation, with the host's libm, where
// it must match the FP64 oracle bit for bit.  Never loaded by the library.
#include <hip/hip_runtime.h>
#include <vector>

#include "ph_engine.h"
#include "ph_stackengine.h"

extern "C" int phs_invert(const PhConst* c, int n_bottoms, long n_pix, const float* rrs,
                          const float* sec_sun, const float* seed, double* x_out, float* depth,
                          float* rrs_error, float* model_error, int* status, int* n_evals,
                          int* n_iter, int* n_starts);

template <int NBOT>
static int run(const PhConst& c, long n_pix, const float* rrs, const float* sec_sun,
               const float* seed, double* x_out, float* depth, float* rrs_error,
               float* model_error, int* status, int* n_evals, int* n_iter, int* n_starts)
{
    constexpr int NB = 4;
    constexpr int N = PhLayout<NBOT>::template n<1>();
    std::vector<double> mem(PhLaneMem<N>::n_slots);
    for (long p = 0; p < n_pix; p++) {
        PhNM<N> S;
        S.m.base = mem.data();
        S.m.stride = 1;
        for (int k = 0; k <= N; k++) S.y[k] = 0.0;
        PhPixel<double, 1, NB> px;
        float r[1][NB];
        for (int b = 0; b < NB; b++) r[0][b] = rrs[(long)b * n_pix + p];
        double ss = sec_sun ? (double)sec_sun[p] : c.sec_sun0, sv = c.sec_view0;
        if (!ph_pixel_init<double, 1, NB>(px, c, r, &ss, &sv)) {
            status[p] = -1;
            depth[p] = 0.0f;
            continue;
        }
        S.kstart = c.has_seed ? -1 : 0;
        S.lowest = 1.0e4;
        S.nstart = 0;
        S.evals = 0;
        S.best_icount = 0;
        S.best_conv = 0;
        for (int i = 0; i < N; i++) S.m.set_best(i, 0.0);
        S.phase = PH_START;
        PhSide<1, NB> side;
        for (;;) {
            if (S.phase == PH_START) {
                double sd[N];
                if (S.kstart < 0) {
                    for (int i = 0; i < N; i++) sd[i] = (double)seed[(long)i * n_pix + p];
                    if (sd[0] != sd[0]) S.kstart = 0; // a NaN seed: no seed start
                }
                ph_start_step<double, 1, NB, NBOT>(S, c, px, S.kstart, sd);
                ph_nm_begin(S, c);
            }
            double f = ph_objective<double, 1, NB, NBOT, true>(S.xe, c, px, &side);
            ph_nm_consume(S, c, f);
            if (S.phase == PH_DONE) break;
        }
        for (int i = 0; i < N; i++) x_out[(long)i * n_pix + p] = fabs(S.m.best(i));
        depth[p] = (float)fabs(S.m.best(0));
        rrs_error[p] = (float)side.Rrs_error;
        model_error[p] = (float)side.model_error;
        status[p] = S.best_conv ? 0 : 2;
        n_evals[p] = S.evals;
        n_iter[p] = S.best_icount;
        n_starts[p] = S.nstart;
    }
    return 0;
}

extern "C" int phs_invert(const PhConst* c, int n_bottoms, long n_pix, const float* rrs,
                          const float* sec_sun, const float* seed, double* x_out, float* depth,
                          float* rrs_error, float* model_error, int* status, int* n_evals,
                          int* n_iter, int* n_starts)
{
    if (c->n_bands != 4) return 2;
    switch (n_bottoms) {
    case 0: return run<0>(*c, n_pix, rrs, sec_sun, seed, x_out, depth, rrs_error, model_error, status, n_evals, n_iter, n_starts);
    case 1: return run<1>(*c, n_pix, rrs, sec_sun, seed, x_out, depth, rrs_error, model_error, status, n_evals, n_iter, n_starts);
    case 2: return run<2>(*c, n_pix, rrs, sec_sun, seed, x_out, depth, rrs_error, model_error, status, n_evals, n_iter, n_starts);
    case 3: return run<3>(*c, n_pix, rrs, sec_sun, seed, x_out, depth, rrs_error, model_error, status, n_evals, n_iter, n_starts);
    default: return 2;
    }
}

// Multi-date stacks (ph_stackengine.h beside this file): the pixels' valid dates in order,
// k = their count, the instance for k; x [2 + 3 n_dates][n] at the dates' slots.
template <int K>
static void run_stack_pixel(PhConst c, long n_pix, int n_dates, long p, const int* slot,
                            const float* rrs, const float* sec_sun, const float* seed,
                            double* x_out, float* depth, float* rrs_error, float* model_error,
                            int* status, int* n_evals, int* n_iter, int* n_starts)
{
    constexpr int N = 2 + 3 * K;
    c.n_dates = K;
    c.rq = c.reqmin * (double)N;
    c.early_exit = 2.5 * (double)(float)K;
    std::vector<double> mem(PhStackNM<N>::n_slots);
    PhStackNM<N> S;
    S.base = mem.data();
    S.stride = 1;
    for (int k = 0; k <= N; k++) S.yv[k] = 0.0;
    float r[K][4];
    double ss[K], sv[K];
    for (int s = 0; s < K; s++) {
        for (int b = 0; b < 4; b++) r[s][b] = rrs[((long)slot[s] * 4 + b) * n_pix + p];
        ss[s] = (double)sec_sun[(long)slot[s] * n_pix + p];
        sv[s] = 1.0;
    }
    for (int i = 0; i < 2 + 3 * n_dates; i++) x_out[(long)i * n_pix + p] = 0.0;
    if (!ph_stack_pixel_init<K>(S, c, r, ss, sv)) {
        status[p] = -1;
        depth[p] = 0.0f;
        return;
    }
    S.kstart = seed ? -1 : 0;
    S.lowest = 1.0e4;
    S.nstart = 0;
    S.evals = 0;
    S.best_icount = 0;
    S.best_conv = 0;
    for (int i = 0; i < N; i++) S.set_best(i, 0.0);
    S.phase = PH_START;
    for (;;) {
        if (S.phase == PH_START) {
            double sd[N];
            if (S.kstart < 0)
                for (int i = 0; i < N; i++) {
                    const int src = i < 2 ? i : 2 + 3 * slot[(i - 2) / 3] + (i - 2) % 3;
                    sd[i] = (double)seed[(long)src * n_pix + p];
                }
            ph_stack_start_step<K>(S, c, S.kstart, sd);
            ph_stack_nm_begin(S, c);
        }
        double f = ph_stack_objective<K, false>(S, c, PhStackNM<N>::XE0, nullptr);
        ph_stack_nm_consume(S, c, f);
        if (S.phase == PH_DONE) break;
    }
    PhStackSide side;
    ph_stack_objective<K, true>(S, c, PhStackNM<N>::BEST0, &side);
    for (int i = 0; i < N; i++) {
        const int dst = i < 2 ? i : 2 + 3 * slot[(i - 2) / 3] + (i - 2) % 3;
        x_out[(long)dst * n_pix + p] = fabs(S.best(i));
    }
    depth[p] = (float)fabs(S.best(0));
    rrs_error[p] = (float)side.Rrs_error;
    model_error[p] = (float)side.model_error;
    status[p] = S.best_conv ? 0 : 2;
    n_evals[p] = S.evals;
    n_iter[p] = S.best_icount;
    n_starts[p] = S.nstart;
}

extern "C" int phs_invert_stack(const PhConst* c, int n_dates, long n_pix, const float* rrs,
                                const float* sec_sun, const unsigned char* valid,
                                const float* seed, double* x_out, float* depth,
                                float* rrs_error, float* model_error, int* status, int* n_evals,
                                int* n_iter, int* n_starts)
{
    if (c->n_bands != 4 || n_dates < 1 || n_dates > 8) return 2;
    for (long p = 0; p < n_pix; p++) {
        int slot[8], k = 0;
        for (int d = 0; d < n_dates; d++)
            if (valid[(long)d * n_pix + p]) slot[k++] = d;
        auto go = [&](auto kk) {
            run_stack_pixel<decltype(kk)::value>(*c, n_pix, n_dates, p, slot, rrs, sec_sun, seed,
                                                 x_out, depth, rrs_error, model_error, status,
                                                 n_evals, n_iter, n_starts);
        };
        switch (k) {
        case 1: go(std::integral_constant<int, 1>{}); break;
        case 2: go(std::integral_constant<int, 2>{}); break;
        case 3: go(std::integral_constant<int, 3>{}); break;
        case 4: go(std::integral_constant<int, 4>{}); break;
        case 5: go(std::integral_constant<int, 5>{}); break;
        case 6: go(std::integral_constant<int, 6>{}); break;
        case 7: go(std::integral_constant<int, 7>{}); break;
        case 8: go(std::integral_constant<int, 8>{}); break;
        default:
            for (int i = 0; i < 2 + 3 * n_dates; i++) x_out[(long)i * n_pix + p] = 0.0;
            status[p] = -1;
            depth[p] = 0.0f;
        }
    }
    return 0;
}
