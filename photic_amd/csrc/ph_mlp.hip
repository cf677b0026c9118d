// photic_amd -- the learned warm start: a 5-128-128-128-5 GELU network mapping band Rrs and
// sun zenith to a starting point (H, B, P, G, X) for the simplex, trained on the GPU on
// samples the forward model generates.
//
// Kernels (BF16 operands, FP32 accumulation, v_mfma_f32_16x16x16_bf16):
//   ph_mlp_fwdbwd_kernel  forward pass of 16-sample tiles; with targets, also the loss and the
//                         backward pass to every layer's pre-activation gradient dZ, written
//                         with the layer inputs H as [feature][batch] bf16 for the GEMM below
//   ph_mlp_wgrad_kernel   dW = dZ . H^T and db = sum dZ, one 16x16 tile of dW per wave over
//                         the batch (no atomics)
//   ph_mlp_adam_kernel    Adam on the FP32 master weights, then the packed BF16 copy
//   ph_mlp_trainset_kernel / ph_mlp_seed_kernel: normalised training pairs from the scene
//                         generator; inference straight into simplex seeds.
//
// Orientation: activations are [features x samples].  With the 16x16x16 MFMA a 16x16 output
// tile (lane l holds column l%16, rows 4(l/16)..+3) is exactly the B operand (lane l holds
// column l%16, K rows 4(l/16)..+3) of the next layer's product, so layers chain in registers
// with no shuffles or LDS round trips.  Weights stay in LDS as BF16.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ph_gen.h"
#include "ph_launch.h"

typedef short ph_bf16x4 __attribute__((ext_vector_type(4)));
typedef float ph_f32x4 __attribute__((ext_vector_type(4)));
typedef float ph_f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 ph_bf16x2 __attribute__((ext_vector_type(2)));

#define MLP_IN 5
#define MLP_HID 128
#define MLP_OUT 5
// master FP32 parameter layout (row-major out x in):
// W1[128][5] b1[128] W2[128][128] b2[128] W3[128][128] b3[128] W4[5][128] b4[5]
#define MO_W1 0
#define MO_B1 (MO_W1 + MLP_HID * MLP_IN)
#define MO_W2 (MO_B1 + MLP_HID)
#define MO_B2 (MO_W2 + MLP_HID * MLP_HID)
#define MO_W3 (MO_B2 + MLP_HID)
#define MO_B3 (MO_W3 + MLP_HID * MLP_HID)
#define MO_W4 (MO_B3 + MLP_HID)
#define MO_B4 (MO_W4 + MLP_OUT * MLP_HID)
#define MLP_NPARAM (MO_B4 + MLP_OUT)
// packed BF16 weights (zero padded): W1p[128][16] W2[128][128] W3[128][128] W4p[16][128]
#define PK_W1 0
#define PK_W2 (PK_W1 + MLP_HID * 16)
#define PK_W3 (PK_W2 + MLP_HID * MLP_HID)
#define PK_W4 (PK_W3 + MLP_HID * MLP_HID)
#define PK_N (PK_W4 + 16 * MLP_HID)
// packed FP32 biases: b1[128] b2[128] b3[128] b4p[16]
#define PB_B1 0
#define PB_B2 128
#define PB_B3 256
#define PB_B4 384
#define PB_N 400

__device__ __forceinline__ unsigned short ph_f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u); // round to nearest even (finite inputs)
    return (unsigned short)(u >> 16);
}

__device__ __forceinline__ float ph_bf2f(unsigned short h)
{
    return __uint_as_float((uint32_t)h << 16);
}

// Four floats rounded to BF16 (to nearest, ties to even, as ph_f2bf for finite inputs), two
// at a time by the conversion instruction gfx950 has for it.
__device__ __forceinline__ ph_bf16x4 ph_pack4(float a, float b, float c, float d)
{
    const ph_bf16x2 lo = __builtin_convertvector((ph_f32x2){a, b}, ph_bf16x2);
    const ph_bf16x2 hi = __builtin_convertvector((ph_f32x2){c, d}, ph_bf16x2);
    const unsigned int ul = __builtin_bit_cast(unsigned int, lo);
    const unsigned int uh = __builtin_bit_cast(unsigned int, hi);
    ph_bf16x4 v;
    v[0] = (short)(ul & 0xffffu);
    v[1] = (short)(ul >> 16);
    v[2] = (short)(uh & 0xffffu);
    v[3] = (short)(uh >> 16);
    return v;
}

// GELU, tanh form, and its derivative.  tanh u = 1 - 2 / (e^2u + 1) with the hardware
// reciprocal (about 1 ulp): the activations are rounded to BF16 next, 2^-8 relative.
__device__ __forceinline__ void ph_gelu(float z, float* g, float* dg)
{
    const float c = 0.7978845608028654f, a = 0.044715f;
    float u = c * fmaf(a * z * z, z, z);
    float e = __expf(2.0f * u);
    float t = 1.0f - 2.0f * __builtin_amdgcn_rcpf(e + 1.0f);
    *g = 0.5f * z * (1.0f + t);
    *dg = 0.5f * (1.0f + t) + 0.5f * z * (1.0f - t * t) * c * fmaf(3.0f * a, z * z, 1.0f);
}

__device__ __forceinline__ ph_f32x4 ph_mfma(ph_bf16x4 a, ph_bf16x4 b, ph_f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// A operand from an LDS row-major matrix M[rows][ld]: rows r0..r0+15, columns k0..k0+15.
__device__ __forceinline__ ph_bf16x4 ph_ldsA(const unsigned short* M, int ld, int r0, int k0,
                                            int lane)
{
    const unsigned short* p = M + (r0 + (lane & 15)) * ld + k0 + 4 * (lane >> 4);
    return *(const ph_bf16x4*)p;
}

// A operand of M^T: rows of M^T are columns of M.  A[m][k] = M[k0 + k][r0 + m].
__device__ __forceinline__ ph_bf16x4 ph_ldsAT(const unsigned short* M, int ld, int r0, int k0,
                                             int lane)
{
    int m = r0 + (lane & 15), k = k0 + 4 * (lane >> 4);
    ph_bf16x4 v;
    v[0] = (short)M[(k + 0) * ld + m];
    v[1] = (short)M[(k + 1) * ld + m];
    v[2] = (short)M[(k + 2) * ld + m];
    v[3] = (short)M[(k + 3) * ld + m];
    return v;
}

struct __align__(16) PhMlpLds {
    unsigned short w1[MLP_HID * 16];
    unsigned short w2[MLP_HID * MLP_HID];
    unsigned short w3[MLP_HID * MLP_HID];
    unsigned short w4[16 * MLP_HID];
    float b[PB_N];
};

__device__ __forceinline__ void ph_mlp_load_lds(PhMlpLds& L, const unsigned short* __restrict__ pk,
                                                const float* __restrict__ pb)
{
    const uint4* src = (const uint4*)pk;
    uint4* dst = (uint4*)L.w1; // w1..w4 are contiguous in the struct
    for (int i = threadIdx.x; i < PK_N / 8; i += blockDim.x) dst[i] = src[i];
    for (int i = threadIdx.x; i < PB_N; i += blockDim.x) L.b[i] = pb[i];
    __syncthreads();
}

// Forward through one hidden 128x128 layer: z[t] = W[16t.., :] . h + b, t = 0..7.
__device__ __forceinline__ void ph_layer128(const unsigned short* W, const float* b,
                                            const ph_bf16x4 (&h)[8], ph_f32x4 (&z)[8], int lane)
{
    const int g = lane >> 4;
#pragma unroll
    for (int t = 0; t < 8; t++) {
        ph_f32x4 acc;
        acc[0] = b[16 * t + 4 * g + 0];
        acc[1] = b[16 * t + 4 * g + 1];
        acc[2] = b[16 * t + 4 * g + 2];
        acc[3] = b[16 * t + 4 * g + 3];
#pragma unroll
        for (int kk = 0; kk < 8; kk++) acc = ph_mfma(ph_ldsA(W, MLP_HID, 16 * t, 16 * kk, lane), h[kk], acc);
        z[t] = acc;
    }
}

__device__ __forceinline__ void ph_act(const ph_f32x4 (&z)[8], ph_bf16x4 (&h)[8],
                                       ph_f32x4 (&dg)[8])
{
#pragma unroll
    for (int t = 0; t < 8; t++) {
        float g0, g1, g2, g3, e0, e1, e2, e3;
        ph_gelu(z[t][0], &g0, &e0);
        ph_gelu(z[t][1], &g1, &e1);
        ph_gelu(z[t][2], &g2, &e2);
        ph_gelu(z[t][3], &g3, &e3);
        dg[t][0] = e0;
        dg[t][1] = e1;
        dg[t][2] = e2;
        dg[t][3] = e3;
        h[t] = ph_pack4(g0, g1, g2, g3);
    }
}

// Write a [128 x 16-sample] C-layout tile set as bf16 rows of a [128][batch] array.
__device__ __forceinline__ void ph_store_h(unsigned short* out, long batch, long s0,
                                           const ph_bf16x4 (&h)[8], int lane)
{
    const int g = lane >> 4;
    const long s = s0 + (lane & 15);
#pragma unroll
    for (int t = 0; t < 8; t++)
#pragma unroll
        for (int r = 0; r < 4; r++) out[(long)(16 * t + 4 * g + r) * batch + s] = (unsigned short)h[t][r];
}

// One 16-sample tile per wave.  x: [n][5] normalised inputs; t: [n][5] normalised targets
// (NULL: inference); y: [n][5] outputs (may be NULL).  With targets: loss (sum of squared
// errors) into *loss, and per layer l the inputs H_l and pre-activation gradients dZ_l as
// [feature][batch] bf16 (gbuf), for samples s0 .. s0 + n - 1 of the batch.
template <bool TRAIN>
__global__ void __launch_bounds__(256) ph_mlp_fwdbwd_kernel(
    const unsigned short* __restrict__ pk, const float* __restrict__ pb, const float* __restrict__ x,
    const float* __restrict__ tgt, float* __restrict__ y, long n, float gscale,
    unsigned short* __restrict__ gbuf, float* __restrict__ loss)
{
    __shared__ PhMlpLds L;
    ph_mlp_load_lds(L, pk, pb);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, g = lane >> 4;
    const long n_tiles = (n + 15) / 16;
    // gradient buffer layout: H1in [16][n] (inputs, rows 5..15 zero), dZ1 [128][n], H2in
    // [128][n], dZ2, H3in, dZ3, H4in [128][n], dZ4 [16][n]
    unsigned short* gH1 = gbuf;
    unsigned short* gZ1 = gH1 + 16 * n;
    unsigned short* gH2 = gZ1 + 128 * n;
    unsigned short* gZ2 = gH2 + 128 * n;
    unsigned short* gH3 = gZ2 + 128 * n;
    unsigned short* gZ3 = gH3 + 128 * n;
    unsigned short* gH4 = gZ3 + 128 * n;
    unsigned short* gZ4 = gH4 + 128 * n;

    for (long tile = (long)blockIdx.x * 4 + wave; tile < n_tiles; tile += (long)gridDim.x * 4) {
        const long s0 = tile * 16;
        const long s = s0 + (lane & 15);
        const bool live = s < n;
        // input B operand: K rows 0..15 = features 0..4, zero padded
        ph_bf16x4 hx;
        {
            float v[4] = {0.f, 0.f, 0.f, 0.f};
            if (live && g == 0) {
                v[0] = x[s * 5 + 0]; v[1] = x[s * 5 + 1]; v[2] = x[s * 5 + 2]; v[3] = x[s * 5 + 3];
            } else if (live && g == 1) {
                v[0] = x[s * 5 + 4];
            }
            hx = ph_pack4(v[0], v[1], v[2], v[3]);
        }
        // layer 1: 128 x 16 (K padded)
        ph_f32x4 d1[8], d2[8], d3[8];
        ph_bf16x4 h1[8], h2[8], h3[8];
        {
            ph_f32x4 z1[8];
#pragma unroll
            for (int t = 0; t < 8; t++) {
                ph_f32x4 acc;
                acc[0] = L.b[PB_B1 + 16 * t + 4 * g + 0];
                acc[1] = L.b[PB_B1 + 16 * t + 4 * g + 1];
                acc[2] = L.b[PB_B1 + 16 * t + 4 * g + 2];
                acc[3] = L.b[PB_B1 + 16 * t + 4 * g + 3];
                z1[t] = ph_mfma(ph_ldsA(L.w1, 16, 16 * t, 0, lane), hx, acc);
            }
            ph_act(z1, h1, d1);
        }
        {
            ph_f32x4 z2[8];
            ph_layer128(L.w2, L.b + PB_B2, h1, z2, lane);
            ph_act(z2, h2, d2);
        }
        {
            ph_f32x4 z3[8];
            ph_layer128(L.w3, L.b + PB_B3, h2, z3, lane);
            ph_act(z3, h3, d3);
        }
        // output layer: 16 (5 valid) x 16
        ph_f32x4 yo;
        yo[0] = L.b[PB_B4 + 4 * g + 0];
        yo[1] = L.b[PB_B4 + 4 * g + 1];
        yo[2] = L.b[PB_B4 + 4 * g + 2];
        yo[3] = L.b[PB_B4 + 4 * g + 3];
#pragma unroll
        for (int kk = 0; kk < 8; kk++) yo = ph_mfma(ph_ldsA(L.w4, MLP_HID, 0, 16 * kk, lane), h3[kk], yo);
        if (y && live) {
#pragma unroll
            for (int r = 0; r < 4; r++)
                if (4 * g + r < 5) y[s * 5 + 4 * g + r] = yo[r];
        }
        if (!TRAIN) continue;

        // ---- loss and backward ----
        ph_f32x4 dy;
        float se = 0.0f;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            int o = 4 * g + r;
            float e = (live && o < 5) ? yo[r] - tgt[s * 5 + o] : 0.0f;
            se += e * e;
            dy[r] = gscale * e;
        }
        // wave sum of squared errors -> one atomic
        for (int off = 32; off > 0; off >>= 1) se += __shfl_xor(se, off);
        if (lane == 0) atomicAdd(loss, se);
        ph_bf16x4 dyb = ph_pack4(dy[0], dy[1], dy[2], dy[3]);
        // dZ4 rows 0..15 (5 valid), H4 = h3
        if (live) {
#pragma unroll
            for (int r = 0; r < 4; r++) gZ4[(long)(4 * g + r) * n + s] = (unsigned short)dyb[r];
        }
        if (live) ph_store_h(gH4, n, s0, h3, lane);
        // dH3 = W4^T dY ; dZ3 = dH3 * gelu'(z3)
        ph_bf16x4 dz[8];
#pragma unroll
        for (int t = 0; t < 8; t++) {
            ph_f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            acc = ph_mfma(ph_ldsAT(L.w4, MLP_HID, 16 * t, 0, lane), dyb, acc);
            dz[t] = ph_pack4(acc[0] * d3[t][0], acc[1] * d3[t][1], acc[2] * d3[t][2], acc[3] * d3[t][3]);
        }
        if (live) {
            ph_store_h(gZ3, n, s0, dz, lane);
            ph_store_h(gH3, n, s0, h2, lane);
        }
        // dH2 = W3^T dZ3
        ph_bf16x4 dz2[8];
#pragma unroll
        for (int t = 0; t < 8; t++) {
            ph_f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kk = 0; kk < 8; kk++) acc = ph_mfma(ph_ldsAT(L.w3, MLP_HID, 16 * t, 16 * kk, lane), dz[kk], acc);
            dz2[t] = ph_pack4(acc[0] * d2[t][0], acc[1] * d2[t][1], acc[2] * d2[t][2], acc[3] * d2[t][3]);
        }
        if (live) {
            ph_store_h(gZ2, n, s0, dz2, lane);
            ph_store_h(gH2, n, s0, h1, lane);
        }
        // dH1 = W2^T dZ2
#pragma unroll
        for (int t = 0; t < 8; t++) {
            ph_f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kk = 0; kk < 8; kk++) acc = ph_mfma(ph_ldsAT(L.w2, MLP_HID, 16 * t, 16 * kk, lane), dz2[kk], acc);
            dz[t] = ph_pack4(acc[0] * d1[t][0], acc[1] * d1[t][1], acc[2] * d1[t][2], acc[3] * d1[t][3]);
        }
        if (live) {
            ph_store_h(gZ1, n, s0, dz, lane);
#pragma unroll
            for (int r = 0; r < 4; r++) gH1[(long)(4 * g + r) * n + s] = (unsigned short)hx[r];
        }
    }
}

// dW (rows x cols) = dZ . H^T over the batch for one layer: one 16x16 tile per wave and one
// of PH_MLP_KSPLIT slices of the batch per blockIdx.y, the slice's partial sums written to
// its own copy of the gradient in the master layout (part + slice * nparam; the bias from the
// waves of the first column tile).  ph_mlp_gsum_kernel then adds the slices in order, so the
// result does not depend on scheduling.  dZ: [rows_pad][n], H: [cols_pad][n] bf16.
__global__ void __launch_bounds__(64) ph_mlp_wgrad_kernel(
    const unsigned short* __restrict__ dZ, const unsigned short* __restrict__ H, long n,
    int rows, int cols, int col_tiles, float* __restrict__ part, int nparam, int w_off,
    int b_off)
{
    const int lane = threadIdx.x, g = lane >> 4;
    const int tr = blockIdx.x / col_tiles, tc = blockIdx.x % col_tiles;
    // this slice's samples: [k_lo, k_hi), whole 16-sample steps
    const long steps = (n + 15) / 16;
    const long per = (steps + PH_MLP_KSPLIT - 1) / PH_MLP_KSPLIT;
    const long k_lo = 16 * per * (long)blockIdx.y;
    const long k_hi = k_lo + 16 * per < n ? k_lo + 16 * per : n;
    float* grad = part + (long)blockIdx.y * nparam;
    ph_f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    float bsum = 0.0f;
    const unsigned short* zr = dZ + (long)(16 * tr + (lane & 15)) * n;
    const unsigned short* hr = H + (long)(16 * tc + (lane & 15)) * n;
    for (long k0 = k_lo; k0 < k_hi; k0 += 16) {
        const long k = k0 + 4 * g;
        ph_bf16x4 a, b;
        if (k + 3 < n) {
            a = *(const ph_bf16x4*)(zr + k);
            b = *(const ph_bf16x4*)(hr + k);
        } else {
#pragma unroll
            for (int j = 0; j < 4; j++) {
                a[j] = (k + j < n) ? (short)zr[k + j] : (short)0;
                b[j] = (k + j < n) ? (short)hr[k + j] : (short)0;
            }
        }
        if (tc == 0)
#pragma unroll
            for (int j = 0; j < 4; j++) bsum += ph_bf2f((unsigned short)a[j]);
        acc = ph_mfma(a, b, acc);
    }
    // acc: lane holds dW[16 tr + 4g + r][16 tc + lane%16]
    const int c = 16 * tc + (lane & 15);
#pragma unroll
    for (int r = 0; r < 4; r++) {
        int row = 16 * tr + 4 * g + r;
        if (row < rows && c < cols) grad[w_off + row * cols + c] = acc[r];
    }
    if (tc == 0) {
        // bias: lanes with the same row (lane%16) hold partial sums over disjoint samples
        bsum += __shfl_xor(bsum, 16);
        bsum += __shfl_xor(bsum, 32);
        int row = 16 * tr + (lane & 15);
        if (lane < 16 && row < rows) grad[b_off + row] = bsum;
    }
}

// grad[i] = the slices' partial sums added in slice order.
__global__ void ph_mlp_gsum_kernel(const float* __restrict__ part, int nparam,
                                   float* __restrict__ grad)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nparam) return;
    float s = part[i];
#pragma unroll
    for (int k = 1; k < PH_MLP_KSPLIT; k++) s += part[(long)k * nparam + i];
    grad[i] = s;
}

__global__ void ph_mlp_adam_kernel(float* __restrict__ w, float* __restrict__ m,
                                   float* __restrict__ v, const float* __restrict__ g, float lr,
                                   float b1, float b2, float eps, float c1, float c2,
                                   unsigned short* __restrict__ pk, float* __restrict__ pb)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= MLP_NPARAM) return;
    float gi = g ? g[i] : 0.0f;
    float wi = w[i];
    if (g) {
        float mi = b1 * m[i] + (1.0f - b1) * gi;
        float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        wi -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        w[i] = wi;
    }
    // pack
    if (i < MO_B1) { int r = i / 5, k = i % 5; pk[PK_W1 + r * 16 + k] = ph_f2bf(wi); }
    else if (i < MO_W2) pb[PB_B1 + (i - MO_B1)] = wi;
    else if (i < MO_B2) pk[PK_W2 + (i - MO_W2)] = ph_f2bf(wi);
    else if (i < MO_W3) pb[PB_B2 + (i - MO_B2)] = wi;
    else if (i < MO_B3) pk[PK_W3 + (i - MO_W3)] = ph_f2bf(wi);
    else if (i < MO_W4) pb[PB_B3 + (i - MO_B3)] = wi;
    else if (i < MO_B4) pk[PK_W4 + (i - MO_W4)] = ph_f2bf(wi);
    else pb[PB_B4 + (i - MO_B4)] = wi;
}

// ---- normalisation: the model's fixed input and target scalings ----
__device__ __forceinline__ void ph_mlp_inputs(const float* r4, float sec_sun, float* x)
{
    for (int b = 0; b < 4; b++) x[b] = (logf(fmaxf(r4[b], 1.0e-4f)) + 4.0f) * (1.0f / 1.5f);
    float th = acosf(fminf(1.0f / fmaxf(sec_sun, 1.0f), 1.0f)) * 57.29577951308232f;
    x[4] = (th - 35.0f) * (1.0f / 15.0f);
}

__global__ void ph_mlp_trainset_kernel(PhConst c, uint64_t seed, long pix0, long n, double sigma,
                                       float* __restrict__ x, float* __restrict__ t)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    PhTruth tr;
    float r[PH_MAXB];
    ph_gen_pixel(seed, (uint64_t)(pix0 + i), c, 4, sigma, 0.0, &tr, r);
    for (int b = 0; b < 4; b++)
        if (r[b] < 0.0f) r[b] = 1.0e-4f;
    double th = tr.theta;
    th *= 3.141592653589793 / 180.0;
    float ss = (float)(1.0 / cos(th));
    ph_mlp_inputs(r, ss, x + i * 5);
    t[i * 5 + 0] = (float)((log(tr.H) - 1.25) / 1.0);
    t[i * 5 + 1] = (float)((log(tr.P) + 3.45) / 1.85);
    t[i * 5 + 2] = (float)((log(tr.G) + 3.91) / 2.30);
    t[i * 5 + 3] = (float)((log(tr.X) + 4.95) / 1.96);
    t[i * 5 + 4] = (float)(2.0 * tr.B - 1.0);
}

// Network inputs from the scene arrays (rrs [4][n], sec_sun [n] or the uniform value).
__global__ void ph_mlp_prep_kernel(const float* __restrict__ rrs, const float* __restrict__ ss,
                                   float ss0, long n, float* __restrict__ x)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float r[4];
    for (int b = 0; b < 4; b++) r[b] = rrs[(long)b * n + i];
    ph_mlp_inputs(r, ss ? ss[i] : ss0, x + i * 5);
}

// One sample's network outputs (the training targets' normalisation undone) -> a seed in the
// MIX layout: s = (H, B, 100 P, 100 G, 100 X).
__device__ __forceinline__ void ph_mlp_decode(const float* o, float* s)
{
    float H = __expf(fminf(fmaxf(o[0], -4.0f), 4.0f) * 1.0f + 1.25f);
    float P = __expf(fminf(fmaxf(o[1], -4.0f), 4.0f) * 1.85f - 3.45f);
    float G = __expf(fminf(fmaxf(o[2], -4.0f), 4.0f) * 2.30f - 3.91f);
    float X = __expf(fminf(fmaxf(o[3], -4.0f), 4.0f) * 1.96f - 4.95f);
    s[0] = fminf(fmaxf(H, 0.1f), 60.0f);
    s[1] = fminf(fmaxf(0.5f * (o[4] + 1.0f), 0.0f), 1.0f);
    s[2] = 100.0f * P;
    s[3] = 100.0f * G;
    s[4] = 100.0f * X;
}

// Network outputs -> simplex seeds in the MIX layout [H, B, 100 P, 100 G, 100 X] ([5][n]); a
// predicted depth below skip_h gives a NaN depth: no seed start for that pixel.
__global__ void ph_mlp_seed_kernel(const float* __restrict__ y, long n, float skip_h,
                                   float* __restrict__ seed)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s[5];
    ph_mlp_decode(y + i * 5, s);
    if (s[0] < skip_h) s[0] = __builtin_nanf("");
    for (int k = 0; k < 5; k++) seed[k * n + i] = s[k];
}

// Date d's network outputs -> its part of a stack seed [2 + 3 D][n] (H, B, then P, G, X per
// date): P, G, X at rows 2 + 3d .. 4 + 3d; H and B summed over the pixel's valid dates into
// rows 0 and 1 (date 0 writes them, later dates add).
__global__ void ph_mlp_stack_seed_kernel(const float* __restrict__ y, long n, int d,
                                         const unsigned char* __restrict__ valid,
                                         float* __restrict__ seed)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s[5];
    ph_mlp_decode(y + i * 5, s);
    for (int k = 0; k < 3; k++) seed[(long)(2 + 3 * d + k) * n + i] = s[2 + k];
    const bool v = valid[(long)d * n + i] != 0;
    const float h = v ? s[0] : 0.0f, b = v ? s[1] : 0.0f;
    seed[i] = d == 0 ? h : seed[i] + h;
    seed[n + i] = d == 0 ? b : seed[n + i] + b;
}

// Rows 0 and 1 of a stack seed: the sums over valid dates -> their means (zero without one).
__global__ void ph_mlp_stack_mean_kernel(long n, int n_dates,
                                         const unsigned char* __restrict__ valid,
                                         float* __restrict__ seed)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int k = 0;
    for (int d = 0; d < n_dates; d++) k += valid[(long)d * n + i] ? 1 : 0;
    if (k > 0) {
        seed[i] = seed[i] / (float)k;
        seed[n + i] = seed[n + i] / (float)k;
    }
}

// Deterministic initialisation (Philox): uniform +-sqrt(6 / (fan_in + fan_out)), biases 0.
__global__ void ph_mlp_init_kernel(uint64_t seed, float* __restrict__ w)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= MLP_NPARAM) return;
    float lim = 0.0f;
    if (i < MO_B1) lim = sqrtf(6.0f / (MLP_IN + MLP_HID));
    else if (i >= MO_W2 && i < MO_B2) lim = sqrtf(6.0f / (2 * MLP_HID));
    else if (i >= MO_W3 && i < MO_B3) lim = sqrtf(6.0f / (2 * MLP_HID));
    else if (i >= MO_W4 && i < MO_B4) lim = sqrtf(6.0f / (MLP_HID + MLP_OUT));
    double u0, u1;
    ph_uniforms(seed, (uint64_t)i, 7, 0, &u0, &u1);
    w[i] = lim * (float)(2.0 * u0 - 1.0);
}

// ---------------------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------------------

extern "C" int ph_mlp_nparam(void) { return MLP_NPARAM; }
extern "C" int ph_mlp_packed_sizes(int* n_bf16, int* n_f32) { *n_bf16 = PK_N; *n_f32 = PB_N; return 0; }
extern "C" long ph_mlp_gbuf_elems(long batch) { return (16 + 128 * 6 + 16) * batch; }

extern "C" int ph_launch_mlp_init(uint64_t seed, float* w, void* stream)
{
    hipLaunchKernelGGL(ph_mlp_init_kernel, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0,
                       (hipStream_t)stream, seed, w);
    return (int)hipGetLastError();
}

extern "C" int ph_launch_mlp_adam(float* w, float* m, float* v, const float* g, float lr,
                                  float b1, float b2, float eps, int step, unsigned short* pk,
                                  float* pb, void* stream)
{
    float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    hipLaunchKernelGGL(ph_mlp_adam_kernel, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0,
                       (hipStream_t)stream, w, m, v, g, lr, b1, b2, eps, c1, c2, pk, pb);
    return (int)hipGetLastError();
}

extern "C" int ph_launch_mlp_fwdbwd(const unsigned short* pk, const float* pb, const float* x,
                                    const float* t, float* y, long n, float gscale,
                                    unsigned short* gbuf, float* loss, int grid, void* stream)
{
    if (n <= 0) return 0;
    long tiles = (n + 15) / 16;
    long need = (tiles + 3) / 4;
    if (grid <= 0 || grid > need) grid = (int)(need < 2048 ? need : 2048);
    if (t)
        hipLaunchKernelGGL(ph_mlp_fwdbwd_kernel<true>, dim3(grid), dim3(256), 0,
                           (hipStream_t)stream, pk, pb, x, t, y, n, gscale, gbuf, loss);
    else
        hipLaunchKernelGGL(ph_mlp_fwdbwd_kernel<false>, dim3(grid), dim3(256), 0,
                           (hipStream_t)stream, pk, pb, x, t, y, n, gscale, gbuf, loss);
    return (int)hipGetLastError();
}

extern "C" int ph_launch_mlp_wgrad(const unsigned short* gbuf, long n, float* grad, float* part,
                                   void* stream)
{
    const unsigned short* gH1 = gbuf;
    const unsigned short* gZ1 = gH1 + 16 * n;
    const unsigned short* gH2 = gZ1 + 128 * n;
    const unsigned short* gZ2 = gH2 + 128 * n;
    const unsigned short* gH3 = gZ2 + 128 * n;
    const unsigned short* gZ3 = gH3 + 128 * n;
    const unsigned short* gH4 = gZ3 + 128 * n;
    const unsigned short* gZ4 = gH4 + 128 * n;
    hipStream_t st = (hipStream_t)stream;
    const int np = ph_mlp_nparam();
    const int ks = PH_MLP_KSPLIT;
    hipLaunchKernelGGL(ph_mlp_wgrad_kernel, dim3(8 * 1, ks), dim3(64), 0, st, gZ1, gH1, n, 128, 5, 1, part, np, MO_W1, MO_B1);
    hipLaunchKernelGGL(ph_mlp_wgrad_kernel, dim3(8 * 8, ks), dim3(64), 0, st, gZ2, gH2, n, 128, 128, 8, part, np, MO_W2, MO_B2);
    hipLaunchKernelGGL(ph_mlp_wgrad_kernel, dim3(8 * 8, ks), dim3(64), 0, st, gZ3, gH3, n, 128, 128, 8, part, np, MO_W3, MO_B3);
    hipLaunchKernelGGL(ph_mlp_wgrad_kernel, dim3(1 * 8, ks), dim3(64), 0, st, gZ4, gH4, n, 5, 128, 8, part, np, MO_W4, MO_B4);
    hipLaunchKernelGGL(ph_mlp_gsum_kernel, dim3((np + 255) / 256), dim3(256), 0, st, part, np, grad);
    return (int)hipGetLastError();
}

extern "C" int ph_launch_mlp_trainset(const PhConst* c, uint64_t seed, long pix0, long n,
                                      double sigma, float* x, float* t, void* stream)
{
    if (n <= 0) return 0;
    hipLaunchKernelGGL(ph_mlp_trainset_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, *c, seed, pix0, n, sigma, x, t);
    return (int)hipGetLastError();
}

extern "C" int ph_launch_mlp_prep(const float* rrs, const float* ss, float ss0, long n, float* x,
                                  void* stream)
{
    if (n <= 0) return 0;
    hipLaunchKernelGGL(ph_mlp_prep_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, rrs, ss, ss0, n, x);
    return (int)hipGetLastError();
}

extern "C" int ph_launch_mlp_seed(const float* y, long n, float skip_h, float* seed,
                                  void* stream)
{
    if (n <= 0) return 0;
    hipLaunchKernelGGL(ph_mlp_seed_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, y, n, skip_h, seed);
    return (int)hipGetLastError();
}

extern "C" int ph_launch_mlp_stack_seed(const float* y, long n, int d, const unsigned char* valid,
                                        float* seed, void* stream)
{
    if (n <= 0) return 0;
    hipLaunchKernelGGL(ph_mlp_stack_seed_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, y, n, d, valid, seed);
    return (int)hipGetLastError();
}

extern "C" int ph_launch_mlp_stack_mean(long n, int n_dates, const unsigned char* valid,
                                        float* seed, void* stream)
{
    if (n <= 0) return 0;
    hipLaunchKernelGGL(ph_mlp_stack_mean_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, n, n_dates, valid, seed);
    return (int)hipGetLastError();
}
