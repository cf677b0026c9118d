// photic_amd -- the stack kernel template (ph_stack_kernel<ND, NBOT, SEED>), its loader and its
// output writer, shared by ph_stack.hip (mix instances) and ph_stack_native.hip (1-3 native
// bottoms), which compile side by side.
#pragma once

#include <hip/hip_runtime.h>

#include "ph_grpengine.h"
#include "ph_launch.h"

__device__ inline void ph_stack_write_nodata(const PhStackArgs& a, long p, int n_params)
{
    // the reference's defaults for pixels it does not model (model/samodel.c:819-841)
    if (a.depth) a.depth[p] = 0.0f;
    if (a.rrs_error) a.rrs_error[p] = 0.0f;
    if (a.model_error) a.model_error[p] = 0.0f;
    if (a.albedo) a.albedo[p] = 0.0f;
    if (a.k_min) a.k_min[p] = 0.0f;
    if (a.iod) a.iod[p] = 0.0f;
    if (a.b_pct)
        for (int m = 0; m < 3; m++) a.b_pct[(long)m * a.n_pix + p] = -9999.0f;
    if (a.bottom_type) a.bottom_type[p] = -9999.0f;
    if (a.x_out)
        for (int i = 0; i < n_params; i++) a.x_out[(long)i * a.n_pix + p] = 0.0;
    if (a.status) a.status[p] = -1;
    if (a.n_evals) a.n_evals[p] = 0;
    if (a.n_iter) a.n_iter[p] = 0;
    if (a.n_starts) a.n_starts[p] = 0;
}

// Outputs of a finished stack: lane li < N writes its coordinate at its date slot (zeros at the
// slots of clouded dates), lane 0 the scalars (model/samodel.c:1992-2079).  qv: |q_k| of the
// best point (the mix: |B|, |1 - B|), gathered by the caller.
template <int ND, int NBOT>
__device__ __attribute__((always_inline)) inline void ph_stack_write_outputs(
    const PhStackArgs& a, const PhConst& c, long p, unsigned int vmask, int dst,
    const PhGrp<ND, NBOT>& S, const double (&qv)[PhGrpLayout<ND, NBOT>::NBT],
    const PhGrpSide& side)
{
    using L = PhGrpLayout<ND, NBOT>;
    constexpr int N = L::N, G = L::G, OFF = L::OFF, NBT = L::NBT;
    const int li = S.li;
    if (a.x_out) {
        if (li < N) a.x_out[(long)dst * a.n_pix + p] = fabs(S.best);
        for (int i = OFF + li; i < OFF + 3 * a.n_dates; i += G)
            if (!((vmask >> ((i - OFF) / 3)) & 1u)) a.x_out[(long)i * a.n_pix + p] = 0.0;
    }
    if (li != 0) return;
    const double H = fabs(S.best);
    // one region: origin weight sqrt(1), depth = H (model/samodel.c:2000-2017)
    if (a.depth) a.depth[p] = (float)((0.0 + 1.0 * H) / (1.0 + 1.0 - 1.0));
    if (a.rrs_error) a.rrs_error[p] = (float)side.Rrs_error;
    if (a.model_error) a.model_error[p] = (float)side.model_error;
    if (a.albedo) a.albedo[p] = (float)side.albedo;
    if (a.k_min) a.k_min[p] = (float)side.k_min; // min over dates and bands (samodel.c:1994)
    if (a.iod) a.iod[p] = (float)(100.0 * side.iod_sum / side.nobs);
    if (a.b_pct || a.bottom_type) {
        // bottom percentages from q, dominant bottom 1-based (model/samodel.c:2028-2047)
        double qsum = 0.0, pct[PH_NBOT];
#pragma unroll
        for (int k = 0; k < NBT; k++) qsum += qv[k];
#pragma unroll
        for (int m = 0; m < PH_NBOT; m++) pct[m] = 0.0;
#pragma unroll
        for (int k = 0; k < NBT; k++) {
            const int t = NBOT == 0 ? k : c.bottom_idx[k];
            const double v = 100.0 * qv[k] / qsum;
#pragma unroll
            for (int m = 0; m < PH_NBOT; m++) pct[m] = ph_pick(m == t, v, pct[m]);
        }
        double largest = 0.0, bt = 0.0;
#pragma unroll
        for (int k = 0; k < NBT; k++) {
            const int t = NBOT == 0 ? k : c.bottom_idx[k];
            double v = 0.0;
#pragma unroll
            for (int m = 0; m < PH_NBOT; m++) v = ph_pick(m == t, pct[m], v);
            if (v > largest) {
                largest = v;
                bt = 1.0 + t;
            }
        }
        if (a.b_pct)
            for (int m = 0; m < 3; m++) a.b_pct[(long)m * a.n_pix + p] = (float)pct[m];
        if (a.bottom_type) a.bottom_type[p] = (float)bt;
    }
    if (a.status) a.status[p] = S.best_conv ? 0 : 2;
    if (a.n_evals) a.n_evals[p] = S.evals;
    if (a.n_iter) a.n_iter[p] = S.best_icount;
    if (a.n_starts) a.n_starts[p] = S.nstart;
}

// Load stack p into the group (model/samodel.c:1785-1799, 2898-2903, per date): pair data,
// the start values of the bottom, P, G, X, the sum of the measured Rrs, the stack's own depth
// start.  False when any value is no-data or negative (model/samodel.c:933-947).  *dst:
// coordinate li's row of x_out and of the seed.
template <int ND, int NBOT>
__device__ __attribute__((always_inline)) inline bool ph_stack_load(
    PhGrp<ND, NBOT>& S, const PhStackArgs& a, long p, unsigned long long gmask, int* dst,
    unsigned int* vmask_out)
{
    using L = PhGrpLayout<ND, NBOT>;
    constexpr int N = L::N, NP = L::NP, OFF = L::OFF, NBT = L::NBT;
    const PhConst& c = a.c;
    const int li = S.li;
    unsigned int slots = 0, vmask = 0; // slot s's date in bits 4s..4s+3
    int s = 0;
    for (int d = 0; d < a.n_dates; d++) {
        if (!a.valid[(long)d * a.n_pix + p]) continue;
        vmask |= 1u << d;
        if (s < ND) slots |= (unsigned int)d << (4 * s);
        s++;
    }
    *vmask_out = vmask;
    const int b = li & 3;
    const int sp = li < NP ? (li >> 2) : 0;                     // the pair's date slot
    const int sc = (li >= OFF && li < N) ? (li - OFF) / 3 : 0; // the coordinate's date slot
    const long dp = (long)((slots >> (4 * sp)) & 15u), dc = (long)((slots >> (4 * sc)) & 15u);
    double rp[4], rc[4];
#pragma unroll
    for (int bb = 0; bb < 4; bb++) {
        rp[bb] = (double)a.rrs[(dp * 4 + bb) * a.n_pix + p];
        rc[bb] = (double)a.rrs[(dc * 4 + bb) * a.n_pix + p];
    }
    const double rv = ph_band_sel(rp, b);
    const float v = (float)rv;
    const bool bad = li < NP && !c.prescreened &&
                     (ph_approx_equal(v, (float)c.nodata, 1.0e-6f) || v < 0.0f);
    if (__ballot(bad) & gmask) return false;

    // the sum of the measured Rrs, in (date, band) order
    ph_wave_sync();
    if (li < NP) S.g[L::D0 + li] = rv;
    ph_wave_sync();
    double tot = 0.0;
#pragma unroll
    for (int j = 0; j < NP; j++) tot += S.g[L::D0 + j];
    S.tot = tot;

    // pair li: its Rrs, (440 / lambda)^Y of its date, geometry, band coefficients, and its
    // band's reflectance of each bottom in rho
    S.rrs = rv;
    double r440 = ph_rrs_at(c, rp, 0);
    const double r490 = ph_rrs_at(c, rp, 1);
    if (r440 < 0.0) r440 = 0.0001;
    const double chi = r440 / r490;
    double Y = 3.44 * (1.0 - 3.17 * PH_EXP(-2.01 * chi));
    if (Y < 0.0) Y = 0.0;
    if (Y > 2.5) Y = 2.5;
    S.pw = PH_POW(440.0 / ph_band_sel(c.wl, b), Y);
    S.ss = (double)a.sec_sun[dp * a.n_pix + p];
    S.sv = a.sec_view ? (double)a.sec_view[dp * a.n_pix + p] : c.sec_view0;
    S.a0 = ph_band_sel(c.a0, b);
    S.a1 = ph_band_sel(c.a1, b);
    S.aw = ph_band_sel(c.aw, b);
    S.bbw = ph_band_sel(c.bbw, b);
    S.eg = ph_band_sel(c.eg, b);
#pragma unroll
    for (int k = 0; k < NBT; k++) S.bt[k] = ph_band_sel(c.bot[NBOT == 0 ? k : c.bottom_idx[k]], b);

    // coordinate li's start and step for the depth starts (model/samodel.c:2246-2354): the
    // mix B = 0.5 (step 0.5); native amplitudes 100 * 4 * mean Rrs(490) over the dates (step
    // 1.5x) and weights 1 (step 0.5); per date P from the 490/550 ratio, G = 1.5 P, X from
    // a_w(640) Rrs(640), steps twice the start
    S.xs0 = 0.0;
    S.step0 = 0.0;
    if (NBOT == 0 && li == 1) {
        S.xs0 = 0.5;
        S.step0 = 0.5;
    } else if (NBOT != 0 && li >= 1 && li < OFF) {
        if (li <= NBT) {
            double m490 = 0.0;
            for (int t = 0; t < ND; t++) {
                const long d = (long)((slots >> (4 * t)) & 15u);
                double r[4];
#pragma unroll
                for (int bb = 0; bb < 4; bb++) r[bb] = (double)a.rrs[(d * 4 + bb) * a.n_pix + p];
                m490 += ph_rrs_at(c, r, 1);
            }
            m490 /= (double)1 * ND;
            S.xs0 = 100.0 * 4.0 * m490;
            S.step0 = 1.5 * S.xs0;
        } else {
            S.xs0 = 1.0;
            S.step0 = 0.5 * S.xs0;
        }
    } else if (li >= OFF && li < N) {
        const double m490 = ph_rrs_at(c, rc, 1) / (double)1;
        const double m550 = ph_rrs_at(c, rc, 2) / (double)1;
        const double m640 = ph_rrs_at(c, rc, 3) / (double)1;
        const double P0 = 100.0 * 0.072 * PH_POW(m490 / m550, -1.7);
        const double G0 = 1.5 * P0;
        const double X0 = 100.0 * 30.0 * c.aw640 * m640;
        const int comp = (li - OFF) % 3;
        S.xs0 = comp == 0 ? P0 : (comp == 1 ? G0 : X0);
        S.step0 = 2.0 * S.xs0;
    }
    S.h0 = a.hstart ? (double)a.hstart[p] : -1.0;
    *dst = li < OFF ? li : OFF + 3 * (int)dc + (li - OFF) % 3;
    return true;
}

// SEED: every stack's first start is its seed (a.seed at the stack's date slots).  Needs
// blockDim.x a multiple of 64 and (blockDim.x / G) * PhGrpLayout<ND, NBOT>::n doubles of
// dynamic LDS.  Compiled for three waves per SIMD (168 VGPRs): the registers it then spills to
// scratch cost less than the extra wave gains (measured: +5% at 8 dates, +19% at 4; four waves
// spill so much that they lose 40%).
template <int ND, int NBOT, bool SEED>
__global__ void __launch_bounds__(256, 3) ph_stack_kernel(PhStackArgs a_in)
{
    using L = PhGrpLayout<ND, NBOT>;
    constexpr int N = L::N, G = L::G, NBT = L::NBT;
    extern __shared__ double ph_stack_lds[];
    const int lane = (int)__lane_id();
    PhGrp<ND, NBOT> S;
    S.li = lane & (G - 1);
    S.gb = lane & ~(G - 1);
    S.yv = 0.0;
    S.g = ph_stack_lds + (long)(threadIdx.x / G) * L::n;
    S.phase = PH_IDLE;
    const int gbase = lane & ~(G - 1);
    const unsigned long long gmask = (G == 64 ? ~0ull : ((1ull << G) - 1ull)) << gbase;
    long p = -1;
    int dst = 0;
    unsigned int vmask = 0;

    const long m = (long)*a_in.count;
    long q_next = 0, q_end = 0;
    bool exhausted = false;
    // watchdog: the wall clock at the wave's last progress (a chunk taken, a stack retired); a
    // wave that makes none for max_ticks stops and raises the flag
    unsigned long long t0 = wall_clock64();

    for (;;) {
        // The launch arguments, read afresh through the kernel-argument segment pointer on
        // every pass: the compiler cannot hoist them out of the loop into scalar registers,
        // which it would then spill into vector lanes (a VALU instruction per reload).
        const PhStackArgs& a = ph_kernarg<PhStackArgs>();
        const PhConst& c = a.c;
        if (wall_clock64() - t0 > a.max_ticks) {
            if (lane == 0) atomicOr(a.queue + 1, 1ull);
            break;
        }
        // ---- refill idle groups from the bucket's queue (one stack per group) ----
        const unsigned long long idle = __ballot(S.li == 0 && S.phase == PH_IDLE);
        if (idle != 0ull && !exhausted) {
            if (q_next >= q_end) {
                const int ld = __ffsll((long long)idle) - 1;
                unsigned long long base = 0;
                if (lane == ld) base = atomicAdd(a.queue, (unsigned long long)a.chunk);
                base = __shfl(base, ld);
                if ((long)base >= m) {
                    exhausted = true;
                } else {
                    q_next = (long)base;
                    q_end = q_next + a.chunk;
                    if (q_end > m) q_end = m;
                    t0 = wall_clock64();
                }
            }
            if (!exhausted || q_next < q_end) {
                const int rank = __popcll(idle & ((1ull << gbase) - 1ull));
                const long avail = q_end - q_next;
                if (S.phase == PH_IDLE && rank < avail) {
                    p = (long)a.idx[q_next + rank];
                    if (ph_stack_load<ND, NBOT>(S, a, p, gmask, &dst, &vmask)) {
                        S.kstart = SEED ? -1 : 0;
                        S.lowest = 1.0e4;
                        S.nstart = 0;
                        S.evals = 0;
                        S.best_icount = 0;
                        S.best_conv = 0;
                        S.best = 0.0;
                        S.phase = PH_START;
                    } else if (S.li == 0) {
                        ph_stack_write_nodata(a, p, a.x_off + 3 * a.n_dates);
                    }
                }
                const int taken = __popcll(idle);
                q_next += (taken < avail) ? taken : avail;
            }
        }
        if (__ballot(S.phase != PH_IDLE) == 0ull) {
            if (exhausted) break;
            continue;
        }

        // ---- set up a new start ----
        if (S.phase == PH_START) {
            double sd = 0.0;
            if constexpr (SEED) {
                if (S.kstart < 0 && S.li < N) sd = (double)a.seed[(long)dst * a.n_pix + p];
            }
            ph_grp_start_step<ND, NBOT>(S, c, S.kstart, sd);
            ph_grp_nm_begin<ND, NBOT>(S, c);
        }

        // ---- one evaluation per group, in lock-step ----
        bool retired = false;
        if (S.phase != PH_IDLE) {
            PhGrpSide side;
            const bool fin = S.phase == PH_FINAL; // the final re-evaluation at the best point
            const double f = ph_grp_objective<ND, NBOT>(S, c, S.xe, fin, &side);
            ph_grp_consume<ND, NBOT>(S, c, f);
            if (S.phase == PH_DONE) {
                double qv[NBT];
                if constexpr (NBOT == 0) {
                    const double B = __shfl(S.best, 1, G);
                    qv[0] = fabs(B);
                    qv[1] = fabs(1.0 - B);
                } else {
#pragma unroll
                    for (int k = 0; k < NBT; k++) qv[k] = fabs(__shfl(S.best, 1 + NBT + k, G));
                }
                ph_stack_write_outputs<ND, NBOT>(a, c, p, vmask, dst, S, qv, side);
                S.phase = PH_IDLE;
                retired = true;
            }
        }
        if (__ballot(retired) != 0ull) t0 = wall_clock64();
    }
}

template <int ND, int NBOT>
static size_t ph_stack_lds_bytes(int block)
{
    return sizeof(double) * (size_t)(block / PhGrpLayout<ND, NBOT>::G) * PhGrpLayout<ND, NBOT>::n;
}

// fn(std::integral_constant<int, K>) for K = k in 1..PH_MAXD.
template <typename Fn>
static int ph_stack_dates(int k, Fn&& fn)
{
    switch (k) {
    case 1: return fn(std::integral_constant<int, 1>{});
    case 2: return fn(std::integral_constant<int, 2>{});
    case 3: return fn(std::integral_constant<int, 3>{});
    case 4: return fn(std::integral_constant<int, 4>{});
    case 5: return fn(std::integral_constant<int, 5>{});
    case 6: return fn(std::integral_constant<int, 6>{});
    case 7: return fn(std::integral_constant<int, 7>{});
    case 8: return fn(std::integral_constant<int, 8>{});
    default: return PH_LAUNCH_UNSUPPORTED;
    }
}

template <int ND, int NBOT, bool SEED>
static int ph_stack_occupancy_t(int block, int* bpc)
{
    return (int)hipOccupancyMaxActiveBlocksPerMultiprocessor(
        bpc, (const void*)ph_stack_kernel<ND, NBOT, SEED>, block,
        ph_stack_lds_bytes<ND, NBOT>(block));
}

template <int ND, int NBOT, bool SEED>
static int ph_launch_stack_t(const PhStackArgs* a, int grid, int block, void* stream)
{
    ph_stack_kernel<ND, NBOT, SEED><<<dim3(grid), dim3(block), ph_stack_lds_bytes<ND, NBOT>(block),
                                      (hipStream_t)stream>>>(*a);
    return (int)hipGetLastError();
}
