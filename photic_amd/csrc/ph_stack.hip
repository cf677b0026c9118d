// photic_amd -- multi-date stacks on gfx950.
//
// A pixel seen on D dates, some of them clouded: its problem is the reference's N_s = k run
// over its k valid dates (model/samodel.c:1808-1815 layout, one region, the sand/seagrass mix
// of the single-date path): depth H and mix B shared, P, G, X per date, 2 + 3k parameters,
// misfit over every (date, band) in date order, early exit at 2.5 k.
//
//   ph_stack_bucket_kernel  sorts the pixels into buckets by k (wave-aggregated atomics) and
//                           writes the no-data outputs of pixels with no valid date
//   ph_stack_kernel<K, S>   the lane engine of ph_stackengine.h with K dates compiled in: one
//                           persistent launch per bucket pulls its pixels from a work queue.
//                           A simplex of up to 27 x 26 doubles does not fit LDS at useful
//                           occupancy, so the Nelder-Mead arrays live in a per-lane slice of
//                           global memory, each wavefront's slice contiguous ([slot][64
//                           lanes]): one slot of a wave is one 512-byte run
//   ph_gen_stack_kernel     the synthetic stack scene
#include <hip/hip_runtime.h>

#include "ph_gen.h"
#include "ph_launch.h"
#include "ph_stackengine.h"

static constexpr int PH_STACK_NB = 4;

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

template <int ND>
__device__ __attribute__((always_inline)) inline void ph_stack_write_outputs(
    const PhStackArgs& a, long p, unsigned int slots, const PhStackNM<2 + 3 * ND>& S,
    const PhStackSide& side)
{
    constexpr int N = 2 + 3 * ND;
    if (a.x_out) {
        for (int i = 0; i < 2 + 3 * a.n_dates; i++) a.x_out[(long)i * a.n_pix + p] = 0.0;
#pragma unroll
        for (int i = 0; i < N; i++) {
            const int s = i < 2 ? 0 : (i - 2) / 3;
            const int d = (int)((slots >> (4 * s)) & 15u);
            const int dst = i < 2 ? i : 2 + 3 * d + (i - 2) % 3;
            a.x_out[(long)dst * a.n_pix + p] = fabs(S.best(i));
        }
    }
    const double H = fabs(S.best(0));
    // one region: origin weight sqrt(1), depth = H (model/samodel.c:2000-2017)
    if (a.depth) a.depth[p] = (float)((0.0 + 1.0 * H) / (1.0 + 1.0 - 1.0));
    if (a.rrs_error) a.rrs_error[p] = (float)side.Rrs_error;
    if (a.model_error) a.model_error[p] = (float)side.model_error;
    if (a.albedo) a.albedo[p] = (float)side.albedo;
    if (a.k_min) a.k_min[p] = (float)side.k_min; // min over dates and bands (samodel.c:1994)
    if (a.iod) a.iod[p] = (float)(100.0 * side.iod_sum / side.nobs);
    if (a.b_pct || a.bottom_type) {
        // the mix's two bottoms: q = (|B|, |1 - B|) (model/samodel.c:2028-2047)
        const double q0 = fabs(S.best(1)), q1 = fabs(1.0 - S.best(1));
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

// SEED: every pixel's first start is its seed (a.seed at the pixel's date slots).
template <int ND, bool SEED>
__global__ void __launch_bounds__(256) ph_stack_kernel(PhStackArgs a)
{
    constexpr int NB = PH_STACK_NB;
    constexpr int N = PhLayout<0>::n<ND>();
    const PhConst& c = a.c;
    const int lane = (int)__lane_id();
    const long gl = (long)blockIdx.x * blockDim.x + threadIdx.x;

    PhStackNM<N> S;
    S.base = a.work + (gl >> 6) * (long)PhStackNM<N>::n_slots * 64 + lane;
    S.stride = 64;
#pragma unroll
    for (int k = 0; k <= N; k++) S.yv[k] = 0.0;
    S.phase = PH_IDLE;
    long p = -1;
    unsigned int slots = 0; // date index of valid date s in bits 4s..4s+3

    const long m = (long)*a.count;
    long q_next = 0, q_end = 0;
    bool exhausted = false;
    const unsigned long long t0 = wall_clock64();

    for (;;) {
        if (wall_clock64() - t0 > a.max_ticks) {
            if (lane == 0) atomicOr(a.queue + 1, 1ull);
            break;
        }
        // ---- refill idle lanes from the bucket's queue ----
        const unsigned long long idle = __ballot(S.phase == PH_IDLE);
        if (idle != 0ull && !exhausted) {
            if (q_next >= q_end) {
                unsigned long long base = 0;
                if (lane == __ffsll((long long)idle) - 1)
                    base = atomicAdd(a.queue, (unsigned long long)a.chunk);
                base = __shfl(base, __ffsll((long long)idle) - 1);
                if ((long)base >= m) {
                    exhausted = true;
                } else {
                    q_next = (long)base;
                    q_end = q_next + a.chunk;
                    if (q_end > m) q_end = m;
                }
            }
            if (!exhausted || q_next < q_end) {
                const int rank = __popcll(idle & ((1ull << lane) - 1ull));
                const long avail = q_end - q_next;
                if (S.phase == PH_IDLE && rank < avail) {
                    p = (long)a.idx[q_next + rank];
                    float r[ND][NB];
                    double ss[ND], sv[ND];
                    slots = 0;
                    int s = 0;
                    for (int d = 0; d < a.n_dates && s < ND; d++) {
                        if (!a.valid[(long)d * a.n_pix + p]) continue;
                        slots |= (unsigned int)d << (4 * s);
                        s++;
                    }
#pragma unroll
                    for (int t = 0; t < ND; t++) {
                        const long d = (long)((slots >> (4 * t)) & 15u);
#pragma unroll
                        for (int b = 0; b < NB; b++)
                            r[t][b] = a.rrs[(d * NB + b) * a.n_pix + p];
                        ss[t] = (double)a.sec_sun[d * a.n_pix + p];
                        sv[t] = 1.0;
                    }
                    if (ph_stack_pixel_init<ND>(S, c, r, ss, sv)) {
                        S.kstart = SEED ? -1 : 0;
                        S.lowest = 1.0e4;
                        S.nstart = 0;
                        S.evals = 0;
                        S.best_icount = 0;
                        S.best_conv = 0;
                        for (int i = 0; i < N; i++) S.set_best(i, 0.0);
                        S.phase = PH_START;
                    } else {
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
            if constexpr (SEED) {
                double sd[N];
                if (S.kstart < 0) {
                    for (int i = 0; i < N; i++) {
                        const int s = i < 2 ? 0 : (i - 2) / 3;
                        const long d = (long)((slots >> (4 * s)) & 15u);
                        const long src = i < 2 ? i : 2 + 3 * d + (i - 2) % 3;
                        sd[i] = (double)a.seed[src * a.n_pix + p];
                    }
                }
                ph_stack_start_step<ND>(S, c, S.kstart, sd);
            } else {
                ph_stack_start_step<ND>(S, c, S.kstart, nullptr);
            }
            ph_stack_nm_begin(S, c);
        }

        // ---- one evaluation per lane, in lock-step ----
        if (S.phase != PH_IDLE) {
            const double f = ph_stack_objective<ND, false>(S, c, PhStackNM<N>::XE0, nullptr);
            ph_stack_nm_consume(S, c, f);
            if (S.phase == PH_DONE) {
                // the final re-evaluation's side quantities, at the best point
                PhStackSide side;
                ph_stack_objective<ND, true>(S, c, PhStackNM<N>::BEST0, &side);
                ph_stack_write_outputs<ND>(a, p, slots, S, side);
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

extern "C" int ph_stack_slots(int k)
{
    switch (k) {
    case 1: return PhStackNM<5>::n_slots;
    case 2: return PhStackNM<8>::n_slots;
    case 3: return PhStackNM<11>::n_slots;
    case 4: return PhStackNM<14>::n_slots;
    case 5: return PhStackNM<17>::n_slots;
    case 6: return PhStackNM<20>::n_slots;
    case 7: return PhStackNM<23>::n_slots;
    case 8: return PhStackNM<26>::n_slots;
    default: return -1;
    }
}

extern "C" int ph_stack_occupancy(int k, int seeded, int block, int* bpc)
{
    return ph_stack_dispatch(k, seeded, [&](auto kd, auto sd) {
        return (int)hipOccupancyMaxActiveBlocksPerMultiprocessor(
            bpc, (const void*)ph_stack_kernel<decltype(kd)::value, decltype(sd)::value>, block,
            0);
    });
}

extern "C" int ph_launch_stack(const PhStackArgs* a, int k, int seeded, int grid, int block,
                               void* stream)
{
    return ph_stack_dispatch(k, seeded, [&](auto kd, auto sd) {
        hipLaunchKernelGGL((ph_stack_kernel<decltype(kd)::value, decltype(sd)::value>),
                           dim3(grid), dim3(block), 0, (hipStream_t)stream, *a);
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
