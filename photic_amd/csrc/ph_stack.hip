// photic_amd -- multi-date stacks on gfx950.
//
// A pixel seen on D dates, some of them clouded: its problem is the reference's N_s = k run
// over its k valid dates (model/samodel.c:1808-1815 layout, one region, the sand/seagrass mix
// of the single-date path): depth H and mix B shared, P, G, X per date, 2 + 3k parameters,
// misfit over every (date, band) in date order, early exit at 2.5 k.
//
//   ph_stack_bucket_kernel  sorts the pixels into buckets by k (wave-aggregated atomics) and
//                           writes the no-data outputs of pixels with no valid date
//   ph_stack_kernel<K, S>   the group engine of ph_grpengine.h with K dates compiled in: one
//                           persistent launch per bucket pulls its stacks from a work queue,
//                           one stack per group of 8, 16 or 32 lanes, the simplex in LDS
//   ph_gen_stack_kernel     the synthetic stack scene
#include <hip/hip_runtime.h>

#include "ph_gen.h"
#include "ph_launch.h"
#include "ph_grpengine.h"

__device__ inline void ph_stack_write_nodata(const PhStackArgs& a, long p, int n_params)
{
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

// Buckets: lists[k * n_pix + i] is the i-th pixel with k valid dates, counts[k] their number.
__global__ void ph_stack_bucket_kernel(PhStackArgs a, unsigned int* counts, int* lists)
{
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = p < a.n_pix;
    int k = -1;
    if (live) {
        k = 0;
        for (int d = 0; d < a.n_dates; d++) k += a.valid[(long)d * a.n_pix + p] ? 1 : 0;
    }
    const int lane = (int)__lane_id();
    for (int kk = 0; kk <= a.n_dates; kk++) {
        const unsigned long long m = __ballot(k == kk);
        if (m == 0ull) continue;
        const int leader = __ffsll((long long)m) - 1;
        unsigned int base = 0;
        if (lane == leader) base = atomicAdd(counts + kk, (unsigned int)__popcll(m));
        base = __shfl(base, leader);
        if (k == kk) {
            const int rank = __popcll(m & ((1ull << lane) - 1ull));
            lists[(long)kk * a.n_pix + base + rank] = (int)p;
        }
    }
    if (live && k == 0) ph_stack_write_nodata(a, p, 2 + 3 * a.n_dates);
}

// Outputs of a finished stack: lane li < N writes its coordinate at its date slot (zeros at the
// slots of clouded dates), lane 0 the scalars.
template <int ND>
__device__ __attribute__((always_inline)) inline void ph_stack_write_outputs(
    const PhStackArgs& a, long p, unsigned int vmask, int dst, const PhGrp<ND>& S, double B,
    const PhGrpSide& side)
{
    constexpr int N = PhGrpLayout<ND>::N, G = PhGrpLayout<ND>::G;
    const int li = S.li;
    if (a.x_out) {
        if (li < N) a.x_out[(long)dst * a.n_pix + p] = fabs(S.best);
        for (int i = 2 + li; i < 2 + 3 * a.n_dates; i += G)
            if (!((vmask >> ((i - 2) / 3)) & 1u)) a.x_out[(long)i * a.n_pix + p] = 0.0;
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
        // the mix's two bottoms: q = (|B|, |1 - B|) (model/samodel.c:2028-2047)
        const double q0 = fabs(B), q1 = fabs(1.0 - B);
        const double qs = q0 + q1;
        const double p0 = 100.0 * q0 / qs, p1 = 100.0 * q1 / qs;
        double largest = 0.0, bt = 0.0;
        if (p0 > largest) { largest = p0; bt = 1.0; }
        if (p1 > largest) { largest = p1; bt = 2.0; }
        if (a.b_pct) {
            a.b_pct[p] = (float)p0;
            a.b_pct[a.n_pix + p] = (float)p1;
            a.b_pct[2 * a.n_pix + p] = 0.0f;
        }
        if (a.bottom_type) a.bottom_type[p] = (float)bt;
    }
    if (a.status) a.status[p] = S.best_conv ? 0 : 2;
    if (a.n_evals) a.n_evals[p] = S.evals;
    if (a.n_iter) a.n_iter[p] = S.best_icount;
    if (a.n_starts) a.n_starts[p] = S.nstart;
}

// Load stack p into the group (model/samodel.c:1785-1799, 2898-2903, per date): pair data,
// the start values of B, P, G, X, the sum of the measured Rrs.  False when any value is no-data
// or negative (model/samodel.c:933-947).  *dst: coordinate li's row of x_out and of the seed.
template <int ND>
__device__ __attribute__((always_inline)) inline bool ph_stack_load(
    PhGrp<ND>& S, const PhStackArgs& a, long p, unsigned long long gmask, int* dst,
    unsigned int* vmask_out)
{
    using L = PhGrpLayout<ND>;
    constexpr int N = L::N, NP = L::NP;
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
    const int sp = li < NP ? (li >> 2) : 0;                 // the pair's date slot
    const int sc = (li >= 2 && li < N) ? (li - 2) / 3 : 0; // the coordinate's date slot
    const long dp = (long)((slots >> (4 * sp)) & 15u), dc = (long)((slots >> (4 * sc)) & 15u);
    double rp[4], rc[4];
#pragma unroll
    for (int bb = 0; bb < 4; bb++) {
        rp[bb] = (double)a.rrs[(dp * 4 + bb) * a.n_pix + p];
        rc[bb] = (double)a.rrs[(dc * 4 + bb) * a.n_pix + p];
    }
    const double rv = ph_band_sel(rp, b);
    const float v = (float)rv;
    const bool bad = li < NP && (ph_approx_equal(v, (float)c.nodata, 1.0e-6f) || v < 0.0f);
    if (__ballot(bad) & gmask) return false;

    // the sum of the measured Rrs, in (date, band) order
    ph_wave_sync();
    if (li < NP) S.g[L::D0 + li] = rv;
    ph_wave_sync();
    double tot = 0.0;
#pragma unroll
    for (int j = 0; j < NP; j++) tot += S.g[L::D0 + j];
    S.tot = tot;

    // pair li: its Rrs, (440 / lambda)^Y of its date, geometry, band coefficients
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
    S.sv = 1.0;
    S.a0 = ph_band_sel(c.a0, b);
    S.a1 = ph_band_sel(c.a1, b);
    S.aw = ph_band_sel(c.aw, b);
    S.bbw = ph_band_sel(c.bbw, b);
    S.eg = ph_band_sel(c.eg, b);
    S.bot0 = ph_band_sel(c.bot[0], b);
    S.bot1 = ph_band_sel(c.bot[1], b);

    // coordinate li: B = 0.5 (step 0.5); per date P from the 490/550 ratio, G = 1.5 P, X from
    // a_w(640) Rrs(640), steps twice the start (model/samodel.c:2246-2354)
    S.xs0 = 0.0;
    S.step0 = 0.0;
    if (li == 1) {
        S.xs0 = 0.5;
        S.step0 = 0.5;
    } else if (li >= 2 && li < N) {
        const double m490 = ph_rrs_at(c, rc, 1) / (double)1;
        const double m550 = ph_rrs_at(c, rc, 2) / (double)1;
        const double m640 = ph_rrs_at(c, rc, 3) / (double)1;
        const double P0 = 100.0 * 0.072 * PH_POW(m490 / m550, -1.7);
        const double G0 = 1.5 * P0;
        const double X0 = 100.0 * 30.0 * c.aw640 * m640;
        const int comp = (li - 2) % 3;
        S.xs0 = comp == 0 ? P0 : (comp == 1 ? G0 : X0);
        S.step0 = 2.0 * S.xs0;
    }
    *dst = li < 2 ? li : 2 + 3 * (int)dc + (li - 2) % 3;
    return true;
}

// SEED: every stack's first start is its seed (a.seed at the stack's date slots).  Needs
// blockDim.x a multiple of 64 and (blockDim.x / G) * PhGrpLayout<ND>::n doubles of dynamic LDS.
template <int ND, bool SEED>
__global__ void __launch_bounds__(256) ph_stack_kernel(PhStackArgs a)
{
    using L = PhGrpLayout<ND>;
    constexpr int N = L::N, G = L::G;
    extern __shared__ double ph_stack_lds[];
    const PhConst& c = a.c;
    const int lane = (int)__lane_id();
    PhGrp<ND> S;
    S.li = lane & (G - 1);
    S.g = ph_stack_lds + (long)(threadIdx.x / G) * L::n;
    S.phase = PH_IDLE;
    const int gbase = lane & ~(G - 1);
    const unsigned long long gmask = (G == 64 ? ~0ull : ((1ull << G) - 1ull)) << gbase;
    long p = -1;
    int dst = 0;
    unsigned int vmask = 0;

    const long m = (long)*a.count;
    long q_next = 0, q_end = 0;
    bool exhausted = false;
    const unsigned long long t0 = wall_clock64();

    for (;;) {
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
                }
            }
            if (!exhausted || q_next < q_end) {
                const int rank = __popcll(idle & ((1ull << gbase) - 1ull));
                const long avail = q_end - q_next;
                if (S.phase == PH_IDLE && rank < avail) {
                    p = (long)a.idx[q_next + rank];
                    if (ph_stack_load<ND>(S, a, p, gmask, &dst, &vmask)) {
                        S.kstart = SEED ? -1 : 0;
                        S.lowest = 1.0e4;
                        S.nstart = 0;
                        S.evals = 0;
                        S.best_icount = 0;
                        S.best_conv = 0;
                        S.best = 0.0;
                        S.phase = PH_START;
                    } else if (S.li == 0) {
                        ph_stack_write_nodata(a, p, 2 + 3 * a.n_dates);
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
            ph_grp_start_step<ND>(S, c, S.kstart, sd);
            ph_grp_nm_begin<ND>(S, c);
        }

        // ---- one evaluation per group, in lock-step ----
        if (S.phase != PH_IDLE) {
            PhGrpSide side;
            const bool fin = S.phase == PH_FINAL; // the final re-evaluation at the best point
            const double f = ph_grp_objective<ND>(S, c, S.xe, fin, &side);
            ph_grp_consume<ND>(S, c, f);
            if (S.phase == PH_DONE) {
                const double B = __shfl(S.best, 1, G);
                ph_stack_write_outputs<ND>(a, p, vmask, dst, S, B, side);
                S.phase = PH_IDLE;
            }
        }
    }
}

__global__ void ph_gen_stack_kernel(const PhConst c, uint64_t seed, long pix0, long n,
                                    int n_dates, double sigma, double p_cloud, int k_min,
                                    float* rrs, float* sec_sun, unsigned char* valid,
                                    float* truth)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double t[2 + 4 * PH_MAXD];
    float r[PH_MAXD * PH_MAXB];
    unsigned char v[PH_MAXD];
    ph_gen_stack_pixel(seed, (uint64_t)(pix0 + i), c, c.n_bands, sigma, n_dates, p_cloud, k_min,
                       t, r, v);
    for (int d = 0; d < n_dates; d++) {
        for (int b = 0; b < c.n_bands; b++) rrs[((long)d * c.n_bands + b) * n + i] = r[d * c.n_bands + b];
        double th = t[5 + 4 * d];
        th *= 3.141592653589793 / 180.0;
        sec_sun[(long)d * n + i] = (float)(1.0 / cos(th));
        valid[(long)d * n + i] = v[d];
    }
    if (truth)
        for (int k = 0; k < 2 + 4 * n_dates; k++) truth[(long)k * n + i] = (float)t[k];
}

// ---------------------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------------------

template <typename Fn>
static int ph_stack_dispatch(int k, int seeded, Fn&& fn)
{
    auto go = [&](auto kd) {
        constexpr int K = decltype(kd)::value;
        return seeded ? fn(std::integral_constant<int, K>{}, std::true_type{})
                      : fn(std::integral_constant<int, K>{}, std::false_type{});
    };
    switch (k) {
    case 1: return go(std::integral_constant<int, 1>{});
    case 2: return go(std::integral_constant<int, 2>{});
    case 3: return go(std::integral_constant<int, 3>{});
    case 4: return go(std::integral_constant<int, 4>{});
    case 5: return go(std::integral_constant<int, 5>{});
    case 6: return go(std::integral_constant<int, 6>{});
    case 7: return go(std::integral_constant<int, 7>{});
    case 8: return go(std::integral_constant<int, 8>{});
    default: return PH_LAUNCH_UNSUPPORTED;
    }
}

template <int ND>
static size_t ph_stack_lds_bytes(int block)
{
    return sizeof(double) * (size_t)(block / PhGrpLayout<ND>::G) * PhGrpLayout<ND>::n;
}

extern "C" int ph_stack_occupancy(int k, int seeded, int block, int* bpc)
{
    return ph_stack_dispatch(k, seeded, [&](auto kd, auto sd) {
        constexpr int K = decltype(kd)::value;
        return (int)hipOccupancyMaxActiveBlocksPerMultiprocessor(
            bpc, (const void*)ph_stack_kernel<K, decltype(sd)::value>, block,
            ph_stack_lds_bytes<K>(block));
    });
}

extern "C" int ph_launch_stack(const PhStackArgs* a, int k, int seeded, int grid, int block,
                               void* stream)
{
    return ph_stack_dispatch(k, seeded, [&](auto kd, auto sd) {
        constexpr int K = decltype(kd)::value;
        hipLaunchKernelGGL((ph_stack_kernel<K, decltype(sd)::value>), dim3(grid), dim3(block),
                           ph_stack_lds_bytes<K>(block), (hipStream_t)stream, *a);
        return (int)hipGetLastError();
    });
}

extern "C" int ph_launch_stack_bucket(const PhStackArgs* a, unsigned int* counts, int* lists,
                                      void* stream)
{
    if (a->n_pix <= 0) return 0;
    hipLaunchKernelGGL(ph_stack_bucket_kernel, dim3((unsigned)((a->n_pix + 255) / 256)),
                       dim3(256), 0, (hipStream_t)stream, *a, counts, lists);
    return (int)hipGetLastError();
}

extern "C" int ph_launch_gen_stack(const PhConst* c, uint64_t seed, long pix0, long n,
                                   int n_dates, double sigma, double p_cloud, int k_min,
                                   float* rrs, float* sec_sun, unsigned char* valid,
                                   float* truth, void* stream)
{
    if (n <= 0) return 0;
    hipLaunchKernelGGL(ph_gen_stack_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, *c, seed, pix0, n, n_dates, sigma, p_cloud, k_min,
                       rrs, sec_sun, valid, truth);
    return (int)hipGetLastError();
}
