// photic_amd -- the reference's pixel ordering, look-up table and hot start (model/samodel.c:
// 649-660, 701-814, 980-1109, 2762-2842) recast for gfx950: the kernels of
// photic_invert_ordered (photic_order.c drives them).
//
// The reference inverts pixels one after another in the order of a spectral index, so that
// neighbours in that order look alike.  It keeps recent solutions in a look-up table and
// copies one whenever its modelled reflectance matches the pixel's (model_spectral_diff_mean
// at most 1.5 %); otherwise it starts the pixel's single Nelder-Mead run from the previous
// pixel's depth (start_at_previous) and restarts with the full depth list when the misfit stays
// above 5.  Each pixel depends on the one before.  Here the sorted sequence is cut into runs;
// each run's first pixel (its leader) is inverted with the full depth list, all leaders at
// once; every other pixel (a follower) then looks for the best match among the modelled
// reflectances of its own leader and the leaders of the runs before it (the table of recent
// solutions), and adopts that leader's parameter vector on a match, or is inverted with one
// start from its own leader's depth, and again with the full list if its misfit stays above the
// restart level.  Every follower is decided at once, and each batch is one persistent launch.
//
//   ph_order_key_kernel      spectral index of every pixel (+inf where the pixel is not
//                            modelled, so it sorts last) and its no-data outputs
//   ph_order_sort            stable ascending radix sort of (index, pixel) pairs (hipCUB)
//   ph_order_decide_kernel   each sorted position's role: leader, LUT match, hot start
//   ph_order_list_kernel     the pixels of one role as a compact list (wave-aggregated)
//   ph_order_roles_kernel    every pixel's role, in pixel order
//   ph_order_gather_kernel   a list's Rrs, sec(sun) and depth starts as planar arrays
//   ph_order_adopt_kernel    a LUT match's outputs: the leader's vector, the pixel's own misfit
//   ph_order_scatter_kernel  a compact list's outputs back to the pixels
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include "ph_engine.h"
#include "ph_launch.h"
#include "ph_outputs.h"
#include "photic.h"

// ---------------------------------------------------------------------------------------
// the spectral index (model/samodel.c:701-768), one scene
// ---------------------------------------------------------------------------------------

// The reference's blue, green and red are the scene's bands 1, 2 and 3 (model/samodel.c:
// 361-363).  Their sum over scenes is a float; the norm is formed in double from the squares
// (pow(x, 2) = x * x exactly) and stored as a float.  With a noise spectrum (every R_sigma >=
// 1e-6) the index is multiplied by the spectral angle between the pixel and the noise,
// acos(<b, n> / (|b| |n|)) with |b| the stored float.  acos is the device library's: a
// last-bit difference from the host's can only move the float index by one ulp, and only the
// order of equal-looking pixels depends on it.
__global__ void ph_order_key_kernel(PhInvertArgs a, int noise, double nb, double ng, double nr,
                                    float* key, int* pix, unsigned int* n_valid, int* role_px)
{
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = p < a.n_pix;
    bool ok = false;
    float k = __builtin_huge_valf();
    if (live) {
        float r[4];
        ok = true;
        for (int b = 0; b < 4; b++) {
            r[b] = a.rrs[(long)b * a.n_pix + p];
            if (ph_approx_equal(r[b], (float)a.c.nodata, 1.0e-6f) || r[b] < 0.0f) ok = false;
        }
        if (ok) {
            const float bs = r[1], gs = r[2], rs = r[3];
            const double sq = (double)bs * (double)bs + (double)gs * (double)gs +
                              (double)rs * (double)rs;
            k = (float)sqrt(sq);
            if (noise) {
                double ang = (double)bs * nb + (double)gs * ng + (double)rs * nr;
                ang /= (double)k * sqrt(nb * nb + ng * ng + nr * nr);
                ang = acos(ang);
                k = (float)((double)k * ang);
            }
        } else {
            ph_write_nodata(a, p, a.c.mode ? 5 : 1 + 2 * a.c.n_bottoms + 3);
            if (role_px) role_px[p] = -1;
        }
        key[p] = k;
        pix[p] = (int)p;
    }
    const unsigned long long m = __ballot(live && ok);
    if ((int)__lane_id() == 0 && m) atomicAdd(n_valid, (unsigned int)__popcll(m));
}

// ---------------------------------------------------------------------------------------
// roles
// ---------------------------------------------------------------------------------------

// model_spectral_diff_mean (model/samodel.c:2795-2842) for one region and one scene: over
// bands 1..3, the relative difference of the measured (m) and modelled (l) Rrs, a band whose
// either value is zero (as a float) giving 0/0; 100 sqrt(mean of their squares), as a float.
__device__ inline float ph_spectral_diff_mean(const double* m, const double* l)
{
    double err = 0.0;
    int nbands = 0;
    for (int b = 1; b <= 3; b++) {
        double cur = 0.0, prev = 0.0, nobs = 0.0;
        if (!ph_approx_equal((float)m[b], 0.0f, 1.0e-6f) &&
            !ph_approx_equal((float)l[b], 0.0f, 1.0e-6f)) {
            cur += m[b];
            prev += l[b];
            nobs += 1.0;
        }
        if (nobs > 0.5) {
            cur /= nobs;
            prev /= nobs;
        }
        const double t = (cur - prev) / cur;
        err += t * t;
        nbands++;
    }
    return (float)(100.0 * sqrt(err / (double)nbands));
}

// Role of sorted position t < n_valid (run k = t / run): 0 the run's leader; 1 a LUT match
// (src: the matching leader); 2 a hot start from leader k's depth (h; 1 m, the reference's
// start for very shallow water, should that depth not be positive).  A leader is a LUT entry
// when it was modelled, its objective is below entry_max and its depth below 32 m
// (model/samodel.c:1034); a follower matches the entry of smallest spectral difference among
// leaders k - window .. k, oldest first, when it is at most match_max (model/samodel.c:1007).
__global__ void ph_order_decide_kernel(long n_valid, int run, int window, int use_lut,
                                       double match_max, double entry_max, const float* rrs,
                                       long n_pix, const int* order, const float* l_depth,
                                       const float* l_objective, const int* l_status,
                                       const double* l_rrs, long n_lead, int* role, int* src,
                                       float* hstart)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_valid) return;
    const long k = t / run;
    if (t % run == 0) {
        role[t] = 0;
        src[t] = (int)k;
        hstart[t] = 0.0f;
        return;
    }
    const long p = order[t];
    double m[4];
    for (int b = 0; b < 4; b++) m[b] = (double)rrs[(long)b * n_pix + p];
    double best = 1.0e4;
    long bk = -1;
    if (use_lut) {
        const long k0 = k - window < 0 ? 0 : k - window;
        for (long kk = k0; kk <= k; kk++) {
            if (!(l_status[kk] >= 0 && (double)l_objective[kk] < entry_max &&
                  l_depth[kk] < 32.0f))
                continue;
            double l[4];
            for (int b = 0; b < 4; b++) l[b] = l_rrs[(long)b * n_lead + kk];
            const double d = (double)ph_spectral_diff_mean(m, l);
            if (d < best) {
                best = d;
                bk = kk;
            }
        }
    }
    if (bk >= 0 && !(best > match_max)) {
        role[t] = 1;
        src[t] = (int)bk;
        hstart[t] = 0.0f;
    } else {
        role[t] = 2;
        src[t] = (int)k;
        hstart[t] = l_depth[k] > 0.0f ? l_depth[k] : 1.0f;
    }
}

// The entries t < n whose role is `want` (or, want < 0, whose flag is set), appended to list
// as map[t] (map NULL: t) in no particular order; *count their number.
__global__ void ph_order_list_kernel(long n, const int* role, int want, const int* flag,
                                     const int* map, int* list, unsigned int* count)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool take = t < n && (want >= 0 ? role[t] == want : flag[t] != 0);
    const unsigned long long m = __ballot(take);
    if (m == 0ull) return;
    const int lane = (int)__lane_id();
    const int first = __ffsll((long long)m) - 1;
    unsigned int base = 0;
    if (lane == first) base = atomicAdd(count, (unsigned int)__popcll(m));
    base = __shfl(base, first);
    if (take) list[base + __popcll(m & ((1ull << lane) - 1ull))] = map ? map[t] : (int)t;
}

// Each pixel's role: role_s[t] at pixel order[t] for t < n (list NULL), or 3 (restarted) at
// the pixels of the listed sorted positions.
__global__ void ph_order_roles_kernel(long n, const int* list, const int* role_s,
                                      const int* order, int* role_px)
{
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    if (list) role_px[order[list[q]]] = 3;
    else role_px[order[q]] = role_s[q];
}

// For every entry q (the sorted position t = list[q], or q * stride without a list, of pixel
// order[t]): its Rrs and sec(sun) as planar arrays, and its depth start.
__global__ void ph_order_gather_kernel(long m, const int* list, long stride, const int* order,
                                       const float* rrs, const float* sec_sun, long n_pix,
                                       const float* h_in, float* rrs_out, float* ss_out,
                                       float* h_out)
{
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= m) return;
    const long t = list ? list[q] : q * stride;
    const long p = order[t];
    for (int b = 0; b < 4; b++) rrs_out[(long)b * m + q] = rrs[(long)b * n_pix + p];
    if (sec_sun) ss_out[q] = sec_sun[p];
    if (h_out) h_out[q] = h_in[t];
}

// A LUT match: the pixel takes its leader's parameter vector as its solution; its outputs are
// those of that vector on its own reflectances (one evaluation: its own misfit, IOD, K).
template <int NBOT>
__global__ void ph_order_adopt_kernel(PhInvertArgs a, long m, const int* list, const int* order,
                                      const int* src, const double* l_x, const int* l_status,
                                      long n_lead)
{
    using L = PhLayout<NBOT>;
    constexpr int N = L::template n<1>();
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= m) return;
    const long t = list[q];
    const long p = order[t];
    const long k = src[t];
    const PhConst& c = a.c;
    PhPixel<double, 1, 4> px;
    float r[1][4];
    for (int b = 0; b < 4; b++) r[0][b] = a.rrs[(long)b * a.n_pix + p];
    const double ss = a.sec_sun ? (double)a.sec_sun[p] : c.sec_sun0;
    const double sv = c.sec_view0;
    ph_pixel_init<double, 1, 4>(px, c, r, &ss, &sv);
    double mem[PhLaneMem<N>::n_slots];
    PhNM<N> S;
    S.m.base = mem;
    S.m.stride = 1;
    double x[N];
#pragma unroll
    for (int i = 0; i < N; i++) {
        x[i] = l_x[(long)i * n_lead + k];
        S.m.set_best(i, x[i]);
    }
    PhSide<1, 4> side;
    ph_objective<double, 1, 4, NBOT, true>(x, c, px, &side);
    S.best_conv = l_status[k] == 0 ? 1 : 0;
    S.evals = 1;
    S.best_icount = 0;
    S.nstart = 0;
    ph_write_outputs<double, 4, NBOT>(a, c, p, S, side);
}

// Outputs of a compact list (src arrays [rows][m]) written at the pixels order[list[q]] (or
// order[q * stride] without a list) of the destination arrays ([rows][n_pix]); accumulate: the
// evaluation and start counts add to what the pixel holds (a restart after a hot start).
__global__ void ph_order_scatter_kernel(photic_outputs dst, photic_outputs from, long m,
                                        const int* list, long stride, const int* order,
                                        long n_pix, int n_params, int accumulate)
{
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= m) return;
    const long p = order[list ? list[q] : q * stride];
    if (dst.depth) dst.depth[p] = from.depth[q];
    if (dst.rrs_error) dst.rrs_error[p] = from.rrs_error[q];
    if (dst.model_error) dst.model_error[p] = from.model_error[q];
    if (dst.albedo) dst.albedo[p] = from.albedo[q];
    if (dst.k_min) dst.k_min[p] = from.k_min[q];
    if (dst.iod) dst.iod[p] = from.iod[q];
    if (dst.bottom_type) dst.bottom_type[p] = from.bottom_type[q];
    if (dst.b_pct)
        for (int r = 0; r < 3; r++) dst.b_pct[(long)r * n_pix + p] = from.b_pct[(long)r * m + q];
    if (dst.x)
        for (int i = 0; i < n_params; i++) dst.x[(long)i * n_pix + p] = from.x[(long)i * m + q];
    if (dst.rrs_model)
        for (int b = 0; b < 4; b++)
            dst.rrs_model[(long)b * n_pix + p] = from.rrs_model[(long)b * m + q];
    if (dst.status) dst.status[p] = from.status[q];
    if (dst.n_iter) dst.n_iter[p] = from.n_iter[q];
    if (dst.n_evals) dst.n_evals[p] = from.n_evals[q] + (accumulate ? dst.n_evals[p] : 0);
    if (dst.n_starts) dst.n_starts[p] = from.n_starts[q] + (accumulate ? dst.n_starts[p] : 0);
}

// Restart flags of a hot-started list: its objective above the level (model/samodel.c:1015).
__global__ void ph_order_restart_kernel(long m, const float* objective, double level,
                                        int* flag)
{
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= m) return;
    flag[q] = (double)objective[q] > level ? 1 : 0;
}

// ---------------------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------------------

static unsigned ph_blocks(long n) { return (unsigned)((n + 255) / 256); }

extern "C" int ph_launch_order_key(const PhInvertArgs* a, int noise, const double* nsig,
                                   float* key, int* pix, unsigned int* n_valid, int* role_px,
                                   void* stream)
{
    if (a->n_pix <= 0) return 0;
    hipLaunchKernelGGL(ph_order_key_kernel, dim3(ph_blocks(a->n_pix)), dim3(256), 0,
                       (hipStream_t)stream, *a, noise, nsig[1], nsig[2], nsig[3], key, pix,
                       n_valid, role_px);
    return (int)hipGetLastError();
}

// Stable ascending sort of n (key, value) pairs; with tmp NULL, *tmp_bytes receives the
// scratch size the sort needs.
extern "C" int ph_order_sort(const float* key_in, float* key_out, const int* val_in,
                             int* val_out, long n, void* tmp, size_t* tmp_bytes, void* stream)
{
    return (int)hipcub::DeviceRadixSort::SortPairs(tmp, *tmp_bytes, key_in, key_out, val_in,
                                                   val_out, (int)n, 0, 32, (hipStream_t)stream);
}

extern "C" int ph_launch_order_decide(long n_valid, int run, int window, int use_lut,
                                      double match_max, double entry_max, const float* rrs,
                                      long n_pix, const int* order, const float* l_depth,
                                      const float* l_objective, const int* l_status,
                                      const double* l_rrs, long n_lead, int* role, int* src,
                                      float* hstart, void* stream)
{
    if (n_valid <= 0) return 0;
    hipLaunchKernelGGL(ph_order_decide_kernel, dim3(ph_blocks(n_valid)), dim3(256), 0,
                       (hipStream_t)stream, n_valid, run, window, use_lut, match_max, entry_max,
                       rrs, n_pix, order, l_depth, l_objective, l_status, l_rrs, n_lead, role,
                       src, hstart);
    return (int)hipGetLastError();
}

extern "C" int ph_launch_order_list(long n, const int* role, int want, const int* flag,
                                    const int* map, int* list, unsigned int* count, void* stream)
{
    if (n <= 0) return 0;
    hipLaunchKernelGGL(ph_order_list_kernel, dim3(ph_blocks(n)), dim3(256), 0,
                       (hipStream_t)stream, n, role, want, flag, map, list, count);
    return (int)hipGetLastError();
}

extern "C" int ph_launch_order_roles(long n, const int* list, const int* role_s,
                                     const int* order, int* role_px, void* stream)
{
    if (n <= 0) return 0;
    hipLaunchKernelGGL(ph_order_roles_kernel, dim3(ph_blocks(n)), dim3(256), 0,
                       (hipStream_t)stream, n, list, role_s, order, role_px);
    return (int)hipGetLastError();
}

extern "C" int ph_launch_order_gather(long m, const int* list, long stride, const int* order,
                                      const float* rrs, const float* sec_sun, long n_pix,
                                      const float* h_in, float* rrs_out, float* ss_out,
                                      float* h_out, void* stream)
{
    if (m <= 0) return 0;
    hipLaunchKernelGGL(ph_order_gather_kernel, dim3(ph_blocks(m)), dim3(256), 0,
                       (hipStream_t)stream, m, list, stride, order, rrs, sec_sun, n_pix, h_in,
                       rrs_out, ss_out, h_out);
    return (int)hipGetLastError();
}

extern "C" int ph_launch_order_adopt(const PhInvertArgs* a, int n_bottoms, long m,
                                     const int* list, const int* order, const int* src,
                                     const double* l_x, const int* l_status, long n_lead,
                                     void* stream)
{
    if (m <= 0) return 0;
    hipStream_t st = (hipStream_t)stream;
    switch (n_bottoms) {
    case 0: hipLaunchKernelGGL(ph_order_adopt_kernel<0>, dim3(ph_blocks(m)), dim3(256), 0, st, *a, m, list, order, src, l_x, l_status, n_lead); break;
    case 1: hipLaunchKernelGGL(ph_order_adopt_kernel<1>, dim3(ph_blocks(m)), dim3(256), 0, st, *a, m, list, order, src, l_x, l_status, n_lead); break;
    case 2: hipLaunchKernelGGL(ph_order_adopt_kernel<2>, dim3(ph_blocks(m)), dim3(256), 0, st, *a, m, list, order, src, l_x, l_status, n_lead); break;
    case 3: hipLaunchKernelGGL(ph_order_adopt_kernel<3>, dim3(ph_blocks(m)), dim3(256), 0, st, *a, m, list, order, src, l_x, l_status, n_lead); break;
    default: return PH_LAUNCH_UNSUPPORTED;
    }
    return (int)hipGetLastError();
}

extern "C" int ph_launch_order_scatter(const photic_outputs* dst, const photic_outputs* from,
                                       long m, const int* list, long stride, const int* order,
                                       long n_pix, int n_params, int accumulate, void* stream)
{
    if (m <= 0) return 0;
    hipLaunchKernelGGL(ph_order_scatter_kernel, dim3(ph_blocks(m)), dim3(256), 0,
                       (hipStream_t)stream, *dst, *from, m, list, stride, order, n_pix,
                       n_params, accumulate);
    return (int)hipGetLastError();
}

extern "C" int ph_launch_order_restart(long m, const float* objective, double level, int* flag,
                                       void* stream)
{
    if (m <= 0) return 0;
    hipLaunchKernelGGL(ph_order_restart_kernel, dim3(ph_blocks(m)), dim3(256), 0,
                       (hipStream_t)stream, m, objective, level, flag);
    return (int)hipGetLastError();
}
