// photic_amd -- gfx950 kernels and their C-ABI launchers.
//
//   ph_invert_kernel   persistent work-queue inversion: one pixel per lane, the lane engine
//                      of ph_engine.h, wave-level refill from a global chunk counter
//   ph_fwd_kernel      forward model only (parity of the FP32/FP64 forward model)
//   ph_objective_kernel misfit of given parameter vectors (parity of the objective)
//   ph_gen_kernel      Philox synthetic scene
//
// Launchers return hipError_t as int; photic_capi.c turns them into the public API.
#include <hip/hip_runtime.h>

#include <mutex>

#include "ph_engine.h"
#include "ph_gen.h"
#include "ph_launch.h"
#include "ph_outputs.h"

// ---------------------------------------------------------------------------------------
// inversion kernel
// ---------------------------------------------------------------------------------------

// The objective reads the band coefficients through the kernel-argument segment pointer
// (ph_kernarg_const): the launch constants must open the kernel's only argument.
static_assert(offsetof(PhInvertArgs, c) == 0, "PhConst must start the kernel arguments");

// Waves per SIMD the kernel is compiled for: two for the five-parameter mix (its lane state
// fits 256 registers and its simplexes 80 KB of LDS per block, so two blocks share a CU), one
// for the native layouts, whose larger simplexes fill the LDS with one block.
template <int NBOT>
constexpr int ph_invert_waves() { return NBOT == 0 ? 2 : 1; }

// SEED: every pixel's first start is its seed point (a.seed), before the depth starts; a pixel
// whose seed depth is NaN has no seed start.
template <typename FT, int NB, int NBOT, bool SEED>
__global__ void __launch_bounds__(256, ph_invert_waves<NBOT>()) ph_invert_kernel(PhInvertArgs a_in)
{
    (void)a_in; // read through ph_kernarg<PhInvertArgs>() in the loop
    using L = PhLayout<NBOT>;
    constexpr int N = L::template n<1>();
    const int lane = (int)__lane_id();

    extern __shared__ double ph_lds[];
    PhNM<N> S;
    S.m.base = ph_lds + threadIdx.x;
    S.m.stride = (int)blockDim.x;
    PhPixel<FT, 1, NB> px;
#pragma unroll
    for (int k = 0; k <= N; k++) S.y[k] = 0.0;
    S.phase = PH_IDLE;
    long pix = -1;

    // wave-local slice of the global work queue: [q_next, q_end)
    long q_next = 0, q_end = 0;
    bool exhausted = false;
    // watchdog: the wall clock at the wave's last progress (a chunk taken, a pixel retired);
    // a wave that makes none for max_ticks stops and raises the flag
    unsigned long long t0 = wall_clock64();

    for (;;) {
        // The launch arguments, read afresh through the kernel-argument segment pointer on
        // every pass: the compiler cannot hoist them out of the loop into scalar registers,
        // which it would then spill into vector lanes (a VALU instruction per reload).
        const PhInvertArgs& a = ph_kernarg<PhInvertArgs>();
        const PhConst& c = a.c;
        if (wall_clock64() - t0 > a.max_ticks) {
            if (lane == 0) atomicOr(a.queue + 1, 1ull);
            break;
        }
        // ---- refill idle lanes ----
        unsigned long long idle = __ballot(S.phase == PH_IDLE);
        const int n_idle = __popcll(idle);
        if (idle != 0ull && !exhausted && (n_idle >= a.refill_min || n_idle == 64)) {
            if (q_next >= q_end) {
                unsigned long long base = 0;
                if (lane == __ffsll((long long)idle) - 1)
                    base = atomicAdd(a.queue, (unsigned long long)a.chunk);
                base = __shfl(base, __ffsll((long long)idle) - 1);
                if ((long)base >= a.n_pix) {
                    exhausted = true;
                } else {
                    q_next = (long)base;
                    q_end = q_next + a.chunk;
                    if (q_end > a.n_pix) q_end = a.n_pix;
                    t0 = wall_clock64();
                }
            }
            if (!exhausted || q_next < q_end) {
                int rank = __popcll(idle & ((1ull << lane) - 1ull));
                long avail = q_end - q_next;
                if (S.phase == PH_IDLE && rank < avail) {
                    pix = q_next + rank;
                    float r[1][NB];
#pragma unroll
                    for (int b = 0; b < NB; b++) r[0][b] = a.rrs[(long)b * a.n_pix + pix];
                    double ss = a.sec_sun ? (double)a.sec_sun[pix] : c.sec_sun0;
                    double sv = a.sec_view ? (double)a.sec_view[pix] : c.sec_view0;
                    bool ok = ph_pixel_init<FT, 1, NB>(px, c, r, &ss, &sv);
                    if (ok) {
                        S.kstart = SEED ? -1 : 0;
                        S.lowest = 1.0e4;
                        S.nstart = 0;
                        S.evals = 0;
                        S.best_icount = 0;
                        S.best_conv = 0;
#pragma unroll
                        for (int i = 0; i < N; i++) S.m.set_best(i, 0.0);
                        S.phase = PH_START;
                    } else {
                        ph_write_nodata(a, pix, N);
                    }
                }
                int taken = __popcll(idle);
                q_next += (taken < avail) ? taken : avail;
            }
        }
        if (__ballot(S.phase != PH_IDLE) == 0ull) {
            if (exhausted) break;
            continue;
        }

        // ---- set up a new start ----
        if (S.phase == PH_START) {
            const double hov = a.hstart ? (double)a.hstart[pix] : -1.0;
            if constexpr (SEED) {
                double sd[N];
                if (S.kstart < 0) {
#pragma unroll
                    for (int i = 0; i < N; i++) sd[i] = (double)a.seed[(long)i * a.n_pix + pix];
                    if (sd[0] != sd[0]) S.kstart = 0; // a NaN seed: no seed start
                }
                ph_start_step<FT, 1, NB, NBOT>(S, c, px, S.kstart, sd, hov);
            } else {
                ph_start_step<FT, 1, NB, NBOT>(S, c, px, S.kstart, nullptr, hov);
            }
            ph_nm_begin(S, c);
        }

        // ---- one evaluation per lane, in lock-step ----
        // The side quantities the outputs need are kept on every evaluation: one inlined copy
        // of the objective keeps the kernel within 256 registers.
        {
            PhSide<1, NB> side;
            double f = 0.0;
            bool retired = false;
            if (S.phase != PH_IDLE)
                f = ph_objective<FT, 1, NB, NBOT, true, true>(S.xe, c, px, &side);
            if (S.phase != PH_IDLE) {
                ph_nm_consume(S, c, f);
                if (S.phase == PH_DONE) {
                    ph_write_outputs<FT, NB, NBOT>(a, c, pix, S, side);
                    S.phase = PH_IDLE;
                    retired = true;
                }
            }
            if (__ballot(retired) != 0ull) t0 = wall_clock64();
        }
    }
}

// ---------------------------------------------------------------------------------------
// forward model / objective kernels (parity tests)
// ---------------------------------------------------------------------------------------

// params [n][5] = (H, P, G, X, B) per sample; sec_sun [n]; Y [n]; out Rrs [n][NB].
template <typename FT>
__global__ void ph_fwd_kernel(const PhConst c, long n, const double* params,
                              const double* sec_sun, const double* Yv, double* out)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double H = params[i * 5 + 0], P = params[i * 5 + 1], G = params[i * 5 + 2];
    double X = params[i * 5 + 3], B = params[i * 5 + 4];
    double ss = sec_sun[i], Y = Yv[i];
    for (int b = 0; b < c.n_bands; b++) {
        double rho = B * c.bot[0][b] + (1.0 - B) * c.bot[1][b];
        double pw = PH_POW(440.0 / c.wl[b], Y);
        if constexpr (sizeof(FT) == 8) {
            double K, rB, r;
            out[i * c.n_bands + b] = ph_band64(P, G, X, rho, H, c.a0[b], c.a1[b], c.aw[b],
                                               c.bbw[b], c.eg[b], pw, ss, 1.0, &K, &rB, &r);
        } else {
            double K, rB, r;
            double lnP = (double)__builtin_amdgcn_logf((float)fmax(P, 1.0e-30)) *
                         0.6931471805599453;
            out[i * c.n_bands + b] =
                ph_band_fast(P, lnP, G, X, rho * 0.3183098861837907, H, c.a0[b], c.a1[b],
                             c.aw[b], c.bbw[b], c.eg[b], pw, ss, 1.0, &K, &rB, &r);
        }
    }
}

// x [n][N] parameter vectors (mix layout); rrs [n][NB] float measured; sec_sun [n].
template <typename FT, int NB>
__global__ void ph_objective_kernel(const PhConst c, long n, const double* x,
                                    const float* rrs, const float* sec_sun, double* out,
                                    double* rrs_error)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    PhPixel<FT, 1, NB> px;
    float r[1][NB];
    for (int b = 0; b < NB; b++) r[0][b] = rrs[i * NB + b];
    double ss = sec_sun[i], sv = 1.0;
    ph_pixel_init<FT, 1, NB>(px, c, r, &ss, &sv);
    PhSide<1, NB> side;
    double xv[5];
    for (int k = 0; k < 5; k++) xv[k] = x[i * 5 + k];
    out[i] = ph_objective<FT, 1, NB, 0, true>(xv, c, px, &side);
    if (rrs_error) rrs_error[i] = side.Rrs_error;
}

// ---------------------------------------------------------------------------------------
// synthetic scene
// ---------------------------------------------------------------------------------------

__global__ void ph_gen_kernel(const PhConst c, uint64_t seed, long pix0, long n,
                              double sigma, double deep_frac, float* rrs /*[NB][n]*/,
                              float* sec_sun, float* truth /*[6][n] or null*/)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    PhTruth t;
    float r[PH_MAXB];
    ph_gen_pixel(seed, (uint64_t)(pix0 + i), c, c.n_bands, sigma, deep_frac, &t, r);
    for (int b = 0; b < c.n_bands; b++) rrs[(long)b * n + i] = r[b];
    double th = t.theta;
    th *= 3.141592653589793 / 180.0;
    sec_sun[i] = (float)(1.0 / cos(th));
    if (truth) {
        truth[0 * n + i] = (float)t.H;
        truth[1 * n + i] = (float)t.P;
        truth[2 * n + i] = (float)t.G;
        truth[3 * n + i] = (float)t.X;
        truth[4 * n + i] = (float)t.B;
        truth[5 * n + i] = (float)t.theta;
    }
}

// Plain device copy, 16 bytes per lane: the reference point that scales the HBM counters of
// the profiler (a launch of known traffic).
__global__ void ph_copy16_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, long n)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long stride = (long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[i];
}

__global__ void ph_zero_u64_kernel(unsigned long long* p, int n)
{
    int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) p[i] = 0ull;
}

// ---------------------------------------------------------------------------------------
// the no-allocation entry point's scratch
// ---------------------------------------------------------------------------------------

// photic_invert_noalloc's device memory beyond its inputs and outputs: the launch's work
// counter and watchdog flag (words 0 and 1).  Each device holds its own copy for the life of
// the process; every call on a device shares it, so calls that overlap on one device race.
// In host code `scratch` names a host-side stand-in, never the device array: kernels receive
// the address ph_scratch_words looks up.
static __device__ __align__(256) unsigned long long scratch[PH_SCRATCH_WORDS];

#define PH_SCRATCH_MAX_DEVICES 64
static unsigned long long* ph_scratch_addr[PH_SCRATCH_MAX_DEVICES];
static std::mutex ph_scratch_lock;

// The address of the current device's scratch array, looked up once per device with that
// device current and kept.
extern "C" int ph_scratch_words(unsigned long long** out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return (int)e;
    if (dev < 0 || dev >= PH_SCRATCH_MAX_DEVICES) return (int)hipErrorInvalidDevice;
    std::lock_guard<std::mutex> g(ph_scratch_lock);
    if (!ph_scratch_addr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return (int)e;
        ph_scratch_addr[dev] = (unsigned long long*)p;
    }
    *out = ph_scratch_addr[dev];
    return 0;
}

// ---------------------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------------------

extern "C" int ph_launch_copy16(const void* src, void* dst, long n16, void* stream)
{
    if (n16 <= 0) return 0;
    long blocks = (n16 + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(ph_copy16_kernel, dim3((unsigned)blocks), dim3(256), 0,
                       (hipStream_t)stream, (const uint4*)src, (uint4*)dst, n16);
    return (int)hipGetLastError();
}

extern "C" int ph_launch_zero_u64(unsigned long long* p, int n, void* stream)
{
    hipLaunchKernelGGL(ph_zero_u64_kernel, dim3((n + 63) / 64), dim3(64), 0,
                       (hipStream_t)stream, p, n);
    return (int)hipGetLastError();
}

// The compiled instances of the inversion kernel: FP64 for the mix and native layouts with
// 1..3 bottoms, FP32 for the mix layout; each plain and seeded.  fn receives a PhKind tag.
template <typename FT, int NBOT, bool SEED>
struct PhKind {
    using ft = FT;
    static constexpr int nbot = NBOT;
    static constexpr bool seed = SEED;
};

template <typename FT, int NBOT, typename Fn>
static int ph_dispatch_seed(int seeded, Fn&& fn)
{
    return seeded ? fn(PhKind<FT, NBOT, true>{}) : fn(PhKind<FT, NBOT, false>{});
}

template <typename Fn>
static int ph_dispatch(int fp32, int n_bands, int n_bottoms, int seeded, Fn&& fn)
{
    if (n_bands != 4) return PH_LAUNCH_UNSUPPORTED;
    if (fp32) return n_bottoms == 0 ? ph_dispatch_seed<float, 0>(seeded, fn) : PH_LAUNCH_UNSUPPORTED;
    switch (n_bottoms) {
    case 0: return ph_dispatch_seed<double, 0>(seeded, fn);
    case 1: return ph_dispatch_seed<double, 1>(seeded, fn);
    case 2: return ph_dispatch_seed<double, 2>(seeded, fn);
    case 3: return ph_dispatch_seed<double, 3>(seeded, fn);
    default: return PH_LAUNCH_UNSUPPORTED;
    }
}

template <typename K>
static const void* ph_invert_fn(K)
{
    return (const void*)ph_invert_kernel<typename K::ft, 4, K::nbot, K::seed>;
}

template <typename K>
static size_t ph_invert_lds(K, int block)
{
    constexpr int N = PhLayout<K::nbot>::template n<1>();
    return sizeof(double) * (size_t)PhLaneMem<N>::n_slots * (size_t)block;
}

extern "C" int ph_lds_bytes(int n_bottoms, int block)
{
    switch (n_bottoms) {
    case 0: return (int)ph_invert_lds(PhKind<double, 0, false>{}, block);
    case 1: return (int)ph_invert_lds(PhKind<double, 1, false>{}, block);
    case 2: return (int)ph_invert_lds(PhKind<double, 2, false>{}, block);
    case 3: return (int)ph_invert_lds(PhKind<double, 3, false>{}, block);
    default: return -1;
    }
}

extern "C" int ph_kernel_occupancy(int fp32, int n_bands, int n_bottoms, int seeded, int block,
                                   int* bpc)
{
    return ph_dispatch(fp32, n_bands, n_bottoms, seeded, [&](auto k) {
        return (int)hipOccupancyMaxActiveBlocksPerMultiprocessor(bpc, ph_invert_fn(k), block,
                                                                 ph_invert_lds(k, block));
    });
}

extern "C" int ph_kernel_attrs(int fp32, int n_bands, int n_bottoms, int* vgprs, int* sgprs,
                               int* scratch, int* maxt)
{
    return ph_dispatch(fp32, n_bands, n_bottoms, 0, [&](auto k) {
        hipFuncAttributes fa;
        hipError_t e = hipFuncGetAttributes(&fa, ph_invert_fn(k));
        if (e != hipSuccess) return (int)e;
        *vgprs = fa.numRegs;
        *sgprs = 0;
        *scratch = (int)fa.localSizeBytes;
        *maxt = fa.maxThreadsPerBlock;
        return 0;
    });
}

extern "C" int ph_launch_invert(const PhInvertArgs* a, int fp32, int n_bands, int n_bottoms,
                                int grid, int block, void* stream)
{
    hipStream_t st = (hipStream_t)stream;
    return ph_dispatch(fp32, n_bands, n_bottoms, a->seed != nullptr, [&](auto k) {
        using K = decltype(k);
        const size_t lds = ph_invert_lds(k, block);
        hipError_t e = hipFuncSetAttribute(ph_invert_fn(k),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return (int)e;
        hipLaunchKernelGGL((ph_invert_kernel<typename K::ft, 4, K::nbot, K::seed>), dim3(grid),
                           dim3(block), lds, st, *a);
        return (int)hipGetLastError();
    });
}

extern "C" int ph_launch_fwd(const PhConst* c, long n, const double* params, const double* sec_sun,
                             const double* Y, double* out, int fp32, void* stream)
{
    int blocks = (int)((n + 255) / 256);
    if (blocks == 0) return 0;
    if (fp32)
        hipLaunchKernelGGL(ph_fwd_kernel<float>, dim3(blocks), dim3(256), 0,
                           (hipStream_t)stream, *c, n, params, sec_sun, Y, out);
    else
        hipLaunchKernelGGL(ph_fwd_kernel<double>, dim3(blocks), dim3(256), 0,
                           (hipStream_t)stream, *c, n, params, sec_sun, Y, out);
    return (int)hipGetLastError();
}

extern "C" int ph_launch_objective(const PhConst* c, long n, const double* x, const float* rrs,
                                   const float* sec_sun, double* out, double* rrs_error,
                                   int fp32, int n_bands, void* stream)
{
    int blocks = (int)((n + 255) / 256);
    if (n_bands != 4) return PH_LAUNCH_UNSUPPORTED;
    if (blocks == 0) return 0;
    if (fp32)
        hipLaunchKernelGGL((ph_objective_kernel<float, 4>), dim3(blocks), dim3(256), 0,
                           (hipStream_t)stream, *c, n, x, rrs, sec_sun, out,
                           rrs_error);
    else
        hipLaunchKernelGGL((ph_objective_kernel<double, 4>), dim3(blocks), dim3(256), 0,
                           (hipStream_t)stream, *c, n, x, rrs, sec_sun, out,
                           rrs_error);
    return (int)hipGetLastError();
}

// Sets every band of the pixels inside the scene's no-data discs to the no-data value.
__global__ void ph_nodata_discs_kernel(uint64_t seed, long n_rows, long n_cols, int n_discs,
                                       double r_min, double r_max, int n_bands, float nodata,
                                       float* rrs)
{
    const long n = n_rows * n_cols;
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    if (ph_in_nodata_disc(seed, p / n_cols, p % n_cols, n_rows, n_cols, n_discs, r_min, r_max))
        for (int b = 0; b < n_bands; b++) rrs[(long)b * n + p] = nodata;
}

extern "C" int ph_launch_nodata_discs(uint64_t seed, long n_rows, long n_cols, int n_discs,
                                      double r_min, double r_max, int n_bands, float nodata,
                                      float* rrs, void* stream)
{
    const long n = n_rows * n_cols;
    if (n <= 0) return 0;
    hipLaunchKernelGGL(ph_nodata_discs_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, seed, n_rows, n_cols, n_discs, r_min, r_max, n_bands,
                       nodata, rrs);
    return (int)hipGetLastError();
}

__global__ void ph_gen_field_kernel(const PhConst c, uint64_t seed, long n_rows, long n_cols,
                                    double theta, double sigma, float* rrs /*[NB][n]*/,
                                    float* truth /*[6][n] or null*/)
{
    const long n = n_rows * n_cols;
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    PhTruth t;
    float r[PH_MAXB];
    ph_gen_field_pixel(seed, q / n_cols, q % n_cols, n_cols, theta, c, c.n_bands, sigma, &t, r);
    for (int b = 0; b < c.n_bands; b++) rrs[(long)b * n + q] = r[b];
    if (truth) {
        truth[0 * n + q] = (float)t.H;
        truth[1 * n + q] = (float)t.P;
        truth[2 * n + q] = (float)t.G;
        truth[3 * n + q] = (float)t.X;
        truth[4 * n + q] = (float)t.B;
        truth[5 * n + q] = (float)t.theta;
    }
}

extern "C" int ph_launch_gen_field(const PhConst* c, uint64_t seed, long n_rows, long n_cols,
                                   double theta, double sigma, float* rrs, float* truth,
                                   void* stream)
{
    const long n = n_rows * n_cols;
    int blocks = (int)((n + 255) / 256);
    if (blocks == 0) return 0;
    hipLaunchKernelGGL(ph_gen_field_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream,
                       *c, seed, n_rows, n_cols, theta, sigma, rrs, truth);
    return (int)hipGetLastError();
}

extern "C" int ph_launch_gen(const PhConst* c, uint64_t seed, long pix0, long n, double sigma,
                             double deep_frac, float* rrs, float* sec_sun, float* truth,
                             void* stream)
{
    int blocks = (int)((n + 255) / 256);
    if (blocks == 0) return 0;
    hipLaunchKernelGGL(ph_gen_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream,
                       *c, seed, pix0, n, sigma, deep_frac, rrs, sec_sun, truth);
    return (int)hipGetLastError();
}
