// photic_amd -- multi-date stacks on gfx950.
//
// A pixel seen on D dates, some of them clouded: its problem is the reference's N_s = k run
// over its k valid dates (model/samodel.c:1808-1815 layout, one region): depth H and the
// bottom shared, P, G, X per date, the misfit over every (date, band) in date order, early
// exit at 2.5 k.
//
//   ph_stack_bucket_kernel     sorts the pixels into buckets by k (wave-aggregated atomics)
//                              and writes the no-data outputs of pixels with no valid date
//   ph_stack_kernel<K, NB, S>  the group engine of ph_grpengine.h with K dates and bottom
//                              layout NB compiled in (ph_stackkern.h): one persistent launch
//                              per bucket pulls its stacks from a work queue, one stack per
//                              group of 8, 16 or 32 lanes, the simplex in LDS.  The mix
//                              instances are here, the native ones in ph_stack_native.hip
//   ph_gen_stack_kernel        the synthetic stack scene
#include <hip/hip_runtime.h>

#include "ph_gen.h"
#include "ph_launch.h"
#include "ph_stackkern.h"

// The stack kernel reads its arguments through the kernel-argument segment pointer
// (ph_kernarg): they must be its only argument.
static_assert(offsetof(PhStackArgs, c) == 0, "PhConst must start the kernel arguments");

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
    if (live && k == 0) ph_stack_write_nodata(a, p, a.x_off + 3 * a.n_dates);
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

extern "C" int ph_stack_occupancy_native(int k, int nbot, int block, int* bpc);
extern "C" int ph_launch_stack_native(const PhStackArgs* a, int k, int nbot, int grid,
                                      int block, void* stream);

extern "C" int ph_stack_occupancy(int k, int nbot, int seeded, int block, int* bpc)
{
    if (nbot != 0) return seeded ? PH_LAUNCH_UNSUPPORTED : ph_stack_occupancy_native(k, nbot, block, bpc);
    return ph_stack_dates(k, [&](auto kd) {
        constexpr int K = decltype(kd)::value;
        return seeded ? ph_stack_occupancy_t<K, 0, true>(block, bpc)
                      : ph_stack_occupancy_t<K, 0, false>(block, bpc);
    });
}

extern "C" int ph_launch_stack(const PhStackArgs* a, int k, int nbot, int seeded, int grid,
                               int block, void* stream)
{
    if (nbot != 0)
        return seeded ? PH_LAUNCH_UNSUPPORTED
                      : ph_launch_stack_native(a, k, nbot, grid, block, stream);
    return ph_stack_dates(k, [&](auto kd) {
        constexpr int K = decltype(kd)::value;
        return seeded ? ph_launch_stack_t<K, 0, true>(a, grid, block, stream)
                      : ph_launch_stack_t<K, 0, false>(a, grid, block, stream);
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
