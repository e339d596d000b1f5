/*
 * Station-beam surrogate network on gfx950 matrix cores.
 *
 * No counterpart in the reference, which evaluates the station beam exactly
 * (oskar/telescope/station/src/oskar_evaluate_station_beam_aperture_array.c
 * :31-288).  The surrogate is a multilayer perceptron E(l, m, f) -> 2x2
 * complex Jones matrix (8 reals):
 *   x = encode(l, m, f)             128 features: sin/cos of 56 spatial
 *                                   frequencies of the scaled offsets
 *                                   u = (f/f_max)(l - l0), v = (f/f_max)(m - m0),
 *                                   and 16 smooth terms of (l, m, n, f)
 *   a_i = silu(W_i a_{i-1} + b_i)   four hidden layers of 256
 *   y = W_4 a_3 + b_4               32 rows (8 used: E), rows 8..31 zero
 *
 * Training (forward, backward, Adam) and inference run on the matrix cores
 * with BF16 operands (round to nearest even) and FP32 accumulation; FP32
 * master weights, Adam moments and gradients.  Kernels:
 *  - surr_encode_kernel: features of a batch, stored [b][k] and [k][b];
 *  - gemm_nt_kernel<EPI>: C[m][n] = sum_k A[m][k] B[n][k] (both operands
 *    k-contiguous) on v_mfma_f32_32x32x16_bf16, a 128 x 128 tile per
 *    workgroup (2 x 2 waves of 64 x 64), k split over grid z; epilogues fuse
 *    bias + SiLU (storing z in FP32 and silu(z) in BF16, both layouts), the
 *    SiLU derivative in the backward pass, or a plain FP32 partial store;
 *  - surr_loss_kernel: mean-square loss over the 8 outputs and dL/dy;
 *  - surr_adam_kernel: Adam on every parameter, refreshing the BF16 copies
 *    of the weights in both layouts;
 *  - surr_infer_kernel: the whole network for 128 directions per
 *    workgroup, activations resident in LDS (two 64 KB buffers), weights
 *    streamed from L2, layers computed transposed (D[n][m] = W[n][:] a[m][:])
 *    so each lane's results are 4 consecutive features of one direction and
 *    store to the next layer's LDS planes as 8-byte writes.
 * The exact-beam fallback near nulls is in surrogate.c.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <vector>

#include "device_common.h"
#include "private.h"

using namespace oskar_amd;

namespace {

constexpr int D_IN = 128;
constexpr int NF = 56;
constexpr int WIDTH = 256;
constexpr int DEPTH = 4;
constexpr int NUM_LAYERS = DEPTH + 1;
constexpr int N_OUT = 8;
constexpr int N_OUT_PAD = 32;
constexpr int LAYER_N[NUM_LAYERS] = {WIDTH, WIDTH, WIDTH, WIDTH, N_OUT_PAD};
constexpr int LAYER_K[NUM_LAYERS] = {D_IN, WIDTH, WIDTH, WIDTH, WIDTH};

constexpr int layer_param_offset(int i)
{
    int o = 0;
    for (int j = 0; j < i; ++j) o += LAYER_N[j] * LAYER_K[j] + LAYER_N[j];
    return o;
}
constexpr int layer_weight_offset(int i)
{
    int o = 0;
    for (int j = 0; j < i; ++j) o += LAYER_N[j] * LAYER_K[j];
    return o;
}
constexpr int NUM_PARAMS = layer_param_offset(NUM_LAYERS);
constexpr int NUM_WEIGHTS = layer_weight_offset(NUM_LAYERS);

typedef __bf16 b8 __attribute__((ext_vector_type(8)));
typedef float f16v __attribute__((ext_vector_type(16)));

#define GLOBAL_AS __attribute__((address_space(1)))
typedef unsigned int u4v __attribute__((ext_vector_type(4)));

/* 16-byte global load through an address-space-1 pointer. */
__device__ __forceinline__ uint4 gld16(const void* p)
{
    const u4v v = *(const GLOBAL_AS u4v*) p;
    return make_uint4(v.x, v.y, v.z, v.w);
}

struct Enc
{
    float l0, m0, f_min, f_max;
    float fourier[NF][2];        /* cycles per unit of (u, v) */
};

struct Layers
{
    int param_off[NUM_LAYERS];   /* W_i then b_i in the flat FP32 vector */
    int weight_off[NUM_LAYERS];  /* W_i in the BF16 copies */
};

Layers layer_table()
{
    Layers t;
    for (int i = 0; i < NUM_LAYERS; ++i)
    {
        t.param_off[i] = layer_param_offset(i);
        t.weight_off[i] = layer_weight_offset(i);
    }
    return t;
}

__device__ __forceinline__ f16v mfma_bf16(const uint4& a, const uint4& b,
        const f16v& c)
{
    b8 av, bv;
    __builtin_memcpy(&av, &a, 16);
    __builtin_memcpy(&bv, &b, 16);
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, c, 0, 0, 0);
}

/* SiLU and its derivative with the hardware reciprocal (about 1 ulp; the
 * activations are rounded to BF16, 2^-9, right after): the IEEE division
 * sequence would cost more than the layer's MFMAs. */
__device__ __forceinline__ float sigmoid(float z)
{
    return __builtin_amdgcn_rcpf(1.0f + __expf(-z));
}

__device__ __forceinline__ float silu(float z)
{
    return z * sigmoid(z);
}

__device__ __forceinline__ float dsilu(float z)
{
    const float s = sigmoid(z);
    return s * (1.0f + z * (1.0f - s));
}

/* sin and cos of 2 pi r: reduction to |r| <= 1/2 in FP32, then the hardware
 * sine and cosine (which take revolutions). */
__device__ __forceinline__ void sincos_rev(float r, float* s, float* c)
{
    const float rr = r - __builtin_rintf(r);
    *s = __builtin_amdgcn_sinf(rr);
    *c = __builtin_amdgcn_cosf(rr);
}

/* Features K0 .. K0 + COUNT - 1 of direction (l, m) at frequency f:
 * k < NF: sin of frequency k; NF <= k < 2 NF: cos of frequency k - NF; then
 * the 16 smooth terms.  The feature range is a compile-time constant, so
 * every index into the encoding table is too (uniform scalar loads). */
template <int K0, int COUNT>
__device__ __forceinline__ void encode_features(const Enc& e, float l,
        float m, float f, float* out)
{
    const float n = sqrtf(fmaxf(0.0f, 1.0f - l * l - m * m));
    const float s = f / e.f_max;
    const float u = s * (l - e.l0), v = s * (m - e.m0);
    const float fh = (2.0f * f - e.f_min - e.f_max) / (e.f_max - e.f_min);
#pragma unroll
    for (int i = 0; i < COUNT; ++i)
    {
        const int k = K0 + i;
        float val;
        if (k < 2 * NF)
        {
            const int j = k < NF ? k : k - NF;
            float sn, cs;
            sincos_rev(e.fourier[j][0] * u + e.fourier[j][1] * v, &sn, &cs);
            val = k < NF ? sn : cs;
        }
        else
        {
            float sp, cp;
            sincos_rev(0.5f * fh, &sp, &cp);
            switch (k - 2 * NF)
            {
            case 0: val = l; break;
            case 1: val = m; break;
            case 2: val = n; break;
            case 3: val = fh; break;
            case 4: val = fh * fh; break;
            case 5: val = l * fh; break;
            case 6: val = m * fh; break;
            case 7: val = n * fh; break;
            case 8: val = l * l; break;
            case 9: val = m * m; break;
            case 10: val = l * m; break;
            case 11: val = n * n; break;
            case 12: val = sp; break;
            case 13: val = cp; break;
            case 14: val = l * n; break;
            default: val = m * n; break;
            }
        }
        out[i] = val;
    }
}

/* Writes the 8 features K0 .. K0 + 7 of one sample, as BF16, to X[k] and
 * XT[k * B]. */
template <int K0>
__device__ __forceinline__ void encode_store8(const Enc& e, float l, float m,
        float f, __bf16* X, __bf16* XT, int B)
{
    float feat[8];
    encode_features<K0, 8>(e, l, m, f, feat);
    __bf16 h[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) h[j] = (__bf16) feat[j];
    uint4 v;
    __builtin_memcpy(&v, h, 16);
    *reinterpret_cast<uint4*>(X + K0) = v;
#pragma unroll
    for (int j = 0; j < 8; ++j) XT[(size_t) (K0 + j) * B] = h[j];
}

template <int K0>
__device__ __forceinline__ void encode_all(const Enc& e, float l, float m,
        float f, __bf16* X, __bf16* XT, int B)
{
    if constexpr (K0 < D_IN)
    {
        encode_store8<K0>(e, l, m, f, X, XT, B);
        encode_all<K0 + 8>(e, l, m, f, X, XT, B);
    }
}

/* ---------------------------------------------------------------------- */
/* Training kernels.                                                       */

/* Features of samples idx[b] (or b), stored as X[b][k] and XT[k][b]. */
__global__ __launch_bounds__(256)
void surr_encode_kernel(int B, const float* l, const float* m, const float* f,
        const int* idx, Enc e, __bf16* X, __bf16* XT)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const int s = idx ? idx[b] : b;
    encode_all<0>(e, l[s], m[s], f[s], X + (size_t) b * D_IN, XT + b, B);
}

/* Given FP32 features [B][128] (the gradient check's input path). */
__global__ __launch_bounds__(256)
void surr_features_kernel(int B, const float* feat, __bf16* X, __bf16* XT)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B * D_IN) return;
    const int b = i / D_IN, k = i - b * D_IN;
    const __bf16 h = (__bf16) feat[i];
    X[i] = h;
    XT[(size_t) k * B + b] = h;
}

/* Batch of training samples: indices from a counter-based hash of (seed,
 * step, b), targets gathered as FP32 [b][8]. */
__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ __launch_bounds__(256)
void surr_gather_kernel(int B, int num_samples, uint64_t seed, int step,
        float inv_scale, const float* E, int* idx, float* T)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const uint64_t h = mix64(seed ^ mix64(((uint64_t) step << 32) |
            (uint32_t) b));
    const int s = (int) (h % (uint64_t) num_samples);
    idx[b] = s;
    for (int j = 0; j < N_OUT; ++j)
        T[b * N_OUT + j] = E[(size_t) s * N_OUT + j] * inv_scale;
}

enum { EPI_F32 = 0, EPI_FWD_HIDDEN = 1, EPI_FWD_OUT = 2, EPI_BWD_HIDDEN = 3 };

struct GemmArgs
{
    const __bf16* A;             /* [M][lda] */
    const __bf16* Bm;            /* [N][ldb] */
    int lda, ldb, M, N, K, k_split;
    const float* bias;           /* [N] */
    float* Z;                    /* FWD_HIDDEN out / BWD in: [M][N] */
    __bf16* act;                 /* [M][N] */
    __bf16* actT;                /* [N][M] */
    float* out;                  /* FWD_OUT [M][N]; F32: [split][M][N] */
};

template <int EPI>
__global__ __launch_bounds__(256)
void gemm_nt_kernel(GemmArgs g)
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int lr = lane & 31, lh = lane >> 5;
    const int m0 = blockIdx.y * 128 + (wave & 1) * 64;
    const int n0 = blockIdx.x * 128 + (wave >> 1) * 64;
    const int k_begin = blockIdx.z * g.k_split;
    const int k_end = min(g.K, k_begin + g.k_split);
    f16v acc[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.0f;
    const uint4 zero = make_uint4(0, 0, 0, 0);
    const __bf16* pa[2];
    const __bf16* pb[2];
    bool ok_a[2], ok_b[2];
    for (int i = 0; i < 2; ++i)
    {
        const int r = m0 + i * 32 + lr, c = n0 + i * 32 + lr;
        ok_a[i] = r < g.M;
        ok_b[i] = c < g.N;
        pa[i] = g.A + (size_t) min(r, g.M - 1) * g.lda + 8 * lh;
        pb[i] = g.Bm + (size_t) min(c, g.N - 1) * g.ldb + 8 * lh;
    }
    for (int k = k_begin; k < k_end; k += 16)
    {
        uint4 a[2], b[2];
        for (int i = 0; i < 2; ++i)
        {
            const uint4 va = *reinterpret_cast<const uint4*>(pa[i] + k);
            const uint4 vb = *reinterpret_cast<const uint4*>(pb[i] + k);
            a[i] = ok_a[i] ? va : zero;
            b[i] = ok_b[i] ? vb : zero;
        }
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j)
                acc[i][j] = mfma_bf16(a[i], b[j], acc[i][j]);
    }
    /* Element (row i, col j) of a 32x32 block: lane j + 32 ((i / 4) % 2),
     * register 4 (i / 8) + i % 4. */
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
        {
            const int col = n0 + j * 32 + lr;
            if (col >= g.N) continue;
            for (int grp = 0; grp < 4; ++grp)
            {
                const int row0 = m0 + i * 32 + 8 * grp + 4 * lh;
                if (row0 >= g.M) continue;
                float val[4];
                for (int r = 0; r < 4; ++r) val[r] = acc[i][j][4 * grp + r];
                if (EPI == EPI_F32)
                {
                    float* o = g.out + (size_t) blockIdx.z * g.M * g.N;
                    for (int r = 0; r < 4; ++r)
                        o[(size_t) (row0 + r) * g.N + col] = val[r];
                }
                else if (EPI == EPI_FWD_OUT)
                {
                    const float bb = g.bias[col];
                    for (int r = 0; r < 4; ++r)
                        g.out[(size_t) (row0 + r) * g.N + col] = val[r] + bb;
                }
                else
                {
                    __bf16 h[4];
                    for (int r = 0; r < 4; ++r)
                    {
                        const size_t o = (size_t) (row0 + r) * g.N + col;
                        float a;
                        if (EPI == EPI_FWD_HIDDEN)
                        {
                            const float z = val[r] + g.bias[col];
                            g.Z[o] = z;
                            a = silu(z);
                        }
                        else
                        {
                            a = val[r] * dsilu(g.Z[o]);
                        }
                        h[r] = (__bf16) a;
                        g.act[o] = h[r];
                    }
                    uint2 pk;
                    __builtin_memcpy(&pk, h, 8);
                    *reinterpret_cast<uint2*>(g.actT + (size_t) col * g.M +
                            row0) = pk;
                }
            }
        }
}

/* Mean-square loss over the 8 outputs: dL/dy = 2 (y - t) / (8 B) stored in
 * BF16 as dY[b][32] and dYT[32][b] (columns 8..31 zero); per-workgroup
 * partial sums of (y - t)^2 in FP64. */
__global__ __launch_bounds__(256)
void surr_loss_kernel(int B, const float* Y, const float* T, __bf16* dY,
        __bf16* dYT, double* partial)
{
    __shared__ double red[256];
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    double s = 0.0;
    if (b < B)
    {
        const float scale = 2.0f / (float) (B * N_OUT);
        for (int j = 0; j < N_OUT_PAD; ++j)
        {
            float d = 0.0f;
            if (j < N_OUT)
            {
                d = Y[(size_t) b * N_OUT_PAD + j] - T[(size_t) b * N_OUT + j];
                s += (double) d * (double) d;
            }
            const __bf16 h = (__bf16) (d * scale);
            dY[(size_t) b * N_OUT_PAD + j] = h;
            dYT[(size_t) j * B + b] = h;
        }
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1)
    {
        if ((int) threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

/* Loss = sum of partials / (8 B), in a fixed order. */
__global__ void surr_loss_sum_kernel(int num_partials, int B,
        const double* partial, double* loss)
{
    __shared__ double red[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < num_partials; i += blockDim.x)
        s += partial[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1)
    {
        if ((int) threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *loss = red[0] / (double) (B * N_OUT);
}

/* Weight gradient: sum of the split-k partials, in split order. */
__global__ __launch_bounds__(256)
void surr_wgrad_reduce_kernel(int count, int splits, const float* part,
        float* grad)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    float s = 0.0f;
    for (int k = 0; k < splits; ++k) s += part[(size_t) k * count + i];
    grad[i] = s;
}

/* Bias gradient: row sums of dZT[n][b]. */
__global__ __launch_bounds__(256)
void surr_bgrad_kernel(int B, const __bf16* dZT, float* grad)
{
    __shared__ float red[256];
    const int n = blockIdx.x;
    float s = 0.0f;
    for (int b = threadIdx.x; b < B; b += blockDim.x)
        s += (float) dZT[(size_t) n * B + b];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1)
    {
        if ((int) threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) grad[n] = red[0];
}

/* Writes the BF16 copy of weight p (flat index) in both layouts; biases
 * have no copy. */
__device__ __forceinline__ void pack_weight(const Layers& t, int p, float w,
        __bf16* wb, __bf16* wbT)
{
    for (int i = NUM_LAYERS - 1; i >= 0; --i)
    {
        if (p < t.param_off[i]) continue;
        const int N = LAYER_N[i], K = LAYER_K[i];
        const int q = p - t.param_off[i];
        if (q < N * K)
        {
            const int r = q / K, c = q - r * K;
            const __bf16 h = (__bf16) w;
            wb[t.weight_off[i] + q] = h;
            wbT[t.weight_off[i] + c * N + r] = h;
        }
        return;
    }
}

__global__ __launch_bounds__(256)
void surr_pack_kernel(const float* w, Layers t, __bf16* wb, __bf16* wbT)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < NUM_PARAMS) pack_weight(t, p, w[p], wb, wbT);
}

/* Adam on the FP32 master parameters, with bias corrections bc1, bc2 of
 * this step; refreshes the BF16 weight copies. */
__global__ __launch_bounds__(256)
void surr_adam_kernel(float* w, float* mom, float* vel, const float* g,
        float lr, float beta1, float beta2, float eps, float bc1, float bc2,
        Layers t, __bf16* wb, __bf16* wbT)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= NUM_PARAMS) return;
    const float gp = g[p];
    const float mp = beta1 * mom[p] + (1.0f - beta1) * gp;
    const float vp = beta2 * vel[p] + (1.0f - beta2) * gp * gp;
    mom[p] = mp;
    vel[p] = vp;
    const float wn = w[p] - lr * (mp / bc1) / (sqrtf(vp / bc2) + eps);
    w[p] = wn;
    pack_weight(t, p, wn, wb, wbT);
}

/* ---------------------------------------------------------------------- */
/* Fused inference.                                                        */

constexpr int INF_ROWS = 128;            /* directions per workgroup */

__device__ __forceinline__ float gld_bias(const float* p, int i)
{
    return *(const GLOBAL_AS float*) (p + i);
}
constexpr int ACT_PLANES = WIDTH / 8;    /* uint4 planes of one activation */

/* One hidden layer's operands for this wave: the BF16 weight fragments of
 * its 64 rows n (two 32-row blocks) for every k-step, and the biases of the
 * rows its lanes hold in the output layout (n_base + 32 i + 8 grp + 4 lh + r).
 * Loaded into registers ahead of use, so the MFMAs never wait on L2. */
struct LayerRegs
{
    uint4 w[WIDTH / 16][2];
    float b[2][4][4];
};

template <int K>
__device__ __forceinline__ void load_weights(const __bf16* W, int n_base,
        LayerRegs& r)
{
    const int lane = threadIdx.x & 63, lr = lane & 31, lh = lane >> 5;
    const __bf16* w0 = W + (size_t) (n_base + lr) * K + 8 * lh;
    const __bf16* w1 = w0 + (size_t) 32 * K;
#pragma unroll
    for (int kk = 0; kk < K / 16; ++kk)
    {
        r.w[kk][0] = gld16(w0 + 16 * kk);
        r.w[kk][1] = gld16(w1 + 16 * kk);
    }
}

__device__ __forceinline__ void load_biases(const float* bias, int n_base,
        LayerRegs& r)
{
    const int lh = (threadIdx.x & 63) >> 5;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int grp = 0; grp < 4; ++grp)
#pragma unroll
            for (int q = 0; q < 4; ++q)
                r.b[i][grp][q] = gld_bias(bias,
                        n_base + 32 * i + 8 * grp + 4 * lh + q);
}

/* Next-layer weights that one layer's k-loop prefetches into the weight
 * registers it has finished with: NEXT_HIDDEN for a hidden layer of WIDTH
 * inputs (both 32-row blocks of this wave's rows), NEXT_OUT for the output
 * layer (row lr of its 32 rows, one block), NEXT_NONE for nothing. */
enum NextKind { NEXT_NONE = 0, NEXT_HIDDEN = 1, NEXT_OUT = 2 };

/* One hidden layer, transposed: D[n][m] = b[n] + sum_k W[n][k] a[m][k] for
 * this wave's 64 rows n (2 blocks) and all 128 directions m (4 blocks of
 * 32); K features in LDS planes act_in[k / 8][m].  The blocks of directions
 * are done one after the other, each finished by its SiLU epilogue
 * (BF16, into the next layer's planes: lane (m = 32 j + lr) holds rows
 * n = n_base + 32 i + 8 grp + 4 lh + q, i.e. 4 consecutive features, one
 * 8-byte write per (i, grp)), so one block's epilogue can run beside the
 * next block's MFMAs.  In the last block, after the MFMAs of k-step kk have
 * read r.w[kk], the same registers receive k-step kk of the next layer's
 * weights (Wn, of the kind NEXT), so the next layer finds its operands
 * loaded. */
template <int K, int NEXT>
__device__ __forceinline__ void infer_layer(LayerRegs& r,
        const uint4* act_in, uint4* act_out, const __bf16* Wn,
        const float* bn, int n_base)
{
    const int lane = threadIdx.x & 63, lr = lane & 31, lh = lane >> 5;
    const __bf16* w0 = Wn + (NEXT == NEXT_OUT ? (size_t) lr * WIDTH :
            (size_t) (n_base + lr) * WIDTH) + 8 * lh;
#pragma unroll
    for (int j = 0; j < 4; ++j)
    {
        /* The accumulators start at the rows' biases. */
        f16v acc[2];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[i][v] = r.b[i][v / 4][v % 4];
        /* The last block's accumulators hold the biases now: load the next
         * hidden layer's into the registers. */
        if constexpr (NEXT == NEXT_HIDDEN)
            if (j == 3) load_biases(bn, n_base, r);
#pragma unroll
        for (int kk = 0; kk < K / 16; ++kk)
        {
            const uint4 b = act_in[(2 * kk + lh) * INF_ROWS + 32 * j + lr];
            acc[0] = mfma_bf16(r.w[kk][0], b, acc[0]);
            acc[1] = mfma_bf16(r.w[kk][1], b, acc[1]);
            if (j == 3)
            {
                if constexpr (NEXT == NEXT_HIDDEN)
                {
                    r.w[kk][0] = gld16(w0 + 16 * kk);
                    r.w[kk][1] = gld16(w0 + (size_t) 32 * WIDTH + 16 * kk);
                }
                else if constexpr (NEXT == NEXT_OUT)
                    r.w[kk][0] = gld16(w0 + 16 * kk);
            }
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int grp = 0; grp < 4; ++grp)
            {
                const int n = n_base + 32 * i + 8 * grp + 4 * lh;
                __bf16 h[4];
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    h[q] = (__bf16) silu(acc[i][4 * grp + q]);
                uint2 pk;
                __builtin_memcpy(&pk, h, 8);
                reinterpret_cast<uint2*>(
                        &act_out[(n / 8) * INF_ROWS + j * 32 + lr])[lh] = pk;
            }
    }
}

template <typename FP>
__global__ __launch_bounds__(256, 1)
void surr_infer_kernel(int num_points, const FP* l, const FP* m,
        float freq, Enc e, const __bf16* wb, const float* params, Layers t,
        float out_scale, float flag_level, Cplx<FP>* beam, int* flags)
{
    __shared__ uint4 act[2][ACT_PLANES * INF_ROWS];
    __shared__ float out_e[INF_ROWS][N_OUT];
    const int tid = threadIdx.x, wave = tid >> 6;
    const int lane = tid & 63, lr = lane & 31, lh = lane >> 5;
    const int d0 = blockIdx.x * INF_ROWS;
    const int n_base = wave * 64;
    LayerRegs r;
    /* Layer 0's operands load while the encoding is computed, and so do the
     * k-steps of layer 1 whose registers layer 0 does not use (its K is
     * half the width); layer 0's k-loop loads the rest of layer 1. */
    load_weights<D_IN>(wb + t.weight_off[0], n_base, r);
    {
        const __bf16* w1 = wb + t.weight_off[1] +
                (size_t) (n_base + lr) * WIDTH + 8 * lh;
#pragma unroll
        for (int kk = D_IN / 16; kk < WIDTH / 16; ++kk)
        {
            r.w[kk][0] = gld16(w1 + 16 * kk);
            r.w[kk][1] = gld16(w1 + (size_t) 32 * WIDTH + 16 * kk);
        }
    }
    load_biases(params + t.param_off[0] + WIDTH * D_IN, n_base, r);

    /* Encoding: two threads per direction, 64 features each (the half is
     * uniform over a wave). */
    {
        const int row = tid % INF_ROWS;
        const int half = __builtin_amdgcn_readfirstlane(tid / INF_ROWS);
        const int d = min(d0 + row, num_points - 1);
        float feat[64];
        if (half == 0)
            encode_features<0, 64>(e, (float) l[d], (float) m[d], freq, feat);
        else
            encode_features<64, 64>(e, (float) l[d], (float) m[d], freq, feat);
#pragma unroll
        for (int p = 0; p < 8; ++p)
        {
            __bf16 h[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) h[q] = (__bf16) feat[8 * p + q];
            uint4 v;
            __builtin_memcpy(&v, h, 16);
            act[0][(half * 8 + p) * INF_ROWS + row] = v;
        }
    }
    __syncthreads();
#pragma unroll
    for (int layer = 0; layer < DEPTH; ++layer)
    {
        const uint4* in = act[layer & 1];
        uint4* out = act[(layer + 1) & 1];
        /* Each layer's k-loop loads the next layer's weights (the output
         * layer's after the last hidden layer). */
        const __bf16* wn = wb + t.weight_off[layer + 1];
        const float* bn = params + t.param_off[layer + 1] + WIDTH * WIDTH;
        if (layer == 0)
            infer_layer<D_IN, NEXT_HIDDEN>(r, in, out, wn, bn, n_base);
        else if (layer + 1 < DEPTH)
            infer_layer<WIDTH, NEXT_HIDDEN>(r, in, out, wn, bn, n_base);
        else
            infer_layer<WIDTH, NEXT_OUT>(r, in, out, wn, bn, n_base);
        __syncthreads();
    }
    /* Output layer: 32 rows (8 used) for this wave's 32 directions; its
     * weights were loaded by the last hidden layer's k-loop. */
    {
        const uint4* in = act[DEPTH & 1];
        f16v o;
#pragma unroll
        for (int v = 0; v < 16; ++v) o[v] = 0.0f;
#pragma unroll
        for (int kk = 0; kk < WIDTH / 16; ++kk)
        {
            const uint4 b = in[(2 * kk + lh) * INF_ROWS + wave * 32 + lr];
            o = mfma_bf16(r.w[kk][0], b, o);
        }
        /* Rows n = 4 lh + q (q < 4) are the 8 outputs. */
        const float* bias = params + t.param_off[DEPTH] + N_OUT_PAD * WIDTH;
#pragma unroll
        for (int q = 0; q < 4; ++q)
            out_e[wave * 32 + lr][4 * lh + q] = out_scale * (o[q] +
                    gld_bias(bias, 4 * lh + q));
    }
    __syncthreads();
    if (tid < INF_ROWS && d0 + tid < num_points)
    {
        const float* v = out_e[tid];
        Cplx<FP>* o = beam + 4 * (size_t) (d0 + tid);
        float mx = 0.0f;
        for (int j = 0; j < 4; ++j)
        {
            o[j].x = (FP) v[2 * j];
            o[j].y = (FP) v[2 * j + 1];
            mx = fmaxf(mx, v[2 * j] * v[2 * j] + v[2 * j + 1] * v[2 * j + 1]);
        }
        if (flags) flags[d0 + tid] = (sqrtf(mx) < flag_level) ? 1 : 0;
    }
}

} /* namespace */

/* ---------------------------------------------------------------------- */
/* Host side.                                                             */

namespace {

struct Surrogate
{
    Enc enc;
    Layers t;
    float* params = nullptr;     /* FP32 master, NUM_PARAMS */
    float* mom = nullptr;
    float* vel = nullptr;
    float* grad = nullptr;
    __bf16* wb = nullptr;        /* W_i [N][K], BF16 */
    __bf16* wbT = nullptr;       /* W_i^T [K][N], BF16 */
    float out_scale = 1.0f;      /* network output units -> beam units */
    int adam_step = 0;
    /* Batch workspace (sized for ws_B samples). */
    int ws_B = 0;
    void* ws = nullptr;
    __bf16 *X, *XT, *A[DEPTH], *AT[DEPTH], *dZ[2], *dZT[2], *dY, *dYT;
    float *Z[DEPTH], *Y, *T, *wpart;
    double *lpart, *loss;
    int* idx;
};

constexpr int WGRAD_SPLIT = 512;         /* samples per weight-grad split */

hipError_t ws_ensure(Surrogate* s, int B)
{
    if (s->ws_B >= B) return hipSuccess;
    if (s->ws) (void) hipFree(s->ws);
    s->ws = nullptr;
    s->ws_B = 0;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off += (bytes + 255) / 256 * 256;
        return o;
    };
    const size_t bf = sizeof(__bf16);
    const size_t oX = take((size_t) B * D_IN * bf), oXT = take((size_t) B * D_IN * bf);
    size_t oA[DEPTH], oAT[DEPTH], oZ[DEPTH];
    for (int i = 0; i < DEPTH; ++i)
    {
        oA[i] = take((size_t) B * WIDTH * bf);
        oAT[i] = take((size_t) B * WIDTH * bf);
        oZ[i] = take((size_t) B * WIDTH * 4);
    }
    const size_t odZ0 = take((size_t) B * WIDTH * bf), odZ1 = take((size_t) B * WIDTH * bf);
    const size_t odZT0 = take((size_t) B * WIDTH * bf), odZT1 = take((size_t) B * WIDTH * bf);
    const size_t odY = take((size_t) B * N_OUT_PAD * bf), odYT = take((size_t) B * N_OUT_PAD * bf);
    const size_t oY = take((size_t) B * N_OUT_PAD * 4), oT = take((size_t) B * N_OUT * 4);
    const int splits = (B + WGRAD_SPLIT - 1) / WGRAD_SPLIT;
    const size_t owp = take((size_t) splits * WIDTH * WIDTH * 4);
    const size_t olp = take((size_t) ((B + 255) / 256) * 8);
    const size_t oloss = take(8);
    const size_t oidx = take((size_t) B * 4);
    const hipError_t err = hipMalloc(&s->ws, off);
    if (err != hipSuccess) { s->ws = nullptr; return err; }
    char* w = (char*) s->ws;
    s->X = (__bf16*) (w + oX);
    s->XT = (__bf16*) (w + oXT);
    for (int i = 0; i < DEPTH; ++i)
    {
        s->A[i] = (__bf16*) (w + oA[i]);
        s->AT[i] = (__bf16*) (w + oAT[i]);
        s->Z[i] = (float*) (w + oZ[i]);
    }
    s->dZ[0] = (__bf16*) (w + odZ0);
    s->dZ[1] = (__bf16*) (w + odZ1);
    s->dZT[0] = (__bf16*) (w + odZT0);
    s->dZT[1] = (__bf16*) (w + odZT1);
    s->dY = (__bf16*) (w + odY);
    s->dYT = (__bf16*) (w + odYT);
    s->Y = (float*) (w + oY);
    s->T = (float*) (w + oT);
    s->wpart = (float*) (w + owp);
    s->lpart = (double*) (w + olp);
    s->loss = (double*) (w + oloss);
    s->idx = (int*) (w + oidx);
    s->ws_B = B;
    return hipSuccess;
}

template <int EPI>
void gemm(const GemmArgs& g, int splits, hipStream_t st)
{
    const dim3 grid((g.N + 127) / 128, (g.M + 127) / 128, splits);
    hipLaunchKernelGGL(gemm_nt_kernel<EPI>, grid, dim3(256), 0, st, g);
}

/* Forward pass of the batch already encoded in s->X (B samples). */
void forward(Surrogate* s, int B, hipStream_t st)
{
    const __bf16* in = s->X;
    for (int i = 0; i < NUM_LAYERS; ++i)
    {
        GemmArgs g = {};
        g.A = in;
        g.lda = LAYER_K[i];
        g.Bm = s->wb + s->t.weight_off[i];
        g.ldb = LAYER_K[i];
        g.M = B;
        g.N = LAYER_N[i];
        g.K = LAYER_K[i];
        g.k_split = LAYER_K[i];
        g.bias = s->params + s->t.param_off[i] + LAYER_N[i] * LAYER_K[i];
        if (i < DEPTH)
        {
            g.Z = s->Z[i];
            g.act = s->A[i];
            g.actT = s->AT[i];
            gemm<EPI_FWD_HIDDEN>(g, 1, st);
            in = s->A[i];
        }
        else
        {
            g.out = s->Y;
            gemm<EPI_FWD_OUT>(g, 1, st);
        }
    }
}

/* Loss of the batch (targets in s->T) into *s->loss, and the gradient of
 * every parameter into s->grad. */
void backward(Surrogate* s, int B, hipStream_t st)
{
    const int nb = (B + 255) / 256;
    hipLaunchKernelGGL(surr_loss_kernel, dim3(nb), dim3(256), 0, st, B, s->Y,
            s->T, s->dY, s->dYT, s->lpart);
    hipLaunchKernelGGL(surr_loss_sum_kernel, dim3(1), dim3(256), 0, st, nb, B,
            s->lpart, s->loss);
    const __bf16* dz = s->dY;
    const __bf16* dzT = s->dYT;
    const int splits = (B + WGRAD_SPLIT - 1) / WGRAD_SPLIT;
    int buf = 0;
    for (int i = NUM_LAYERS - 1; i >= 0; --i)
    {
        const int N = LAYER_N[i], K = LAYER_K[i];
        const __bf16* inT = (i == 0) ? s->XT : s->AT[i - 1];
        GemmArgs g = {};
        g.A = dzT;
        g.lda = B;
        g.Bm = inT;
        g.ldb = B;
        g.M = N;
        g.N = K;
        g.K = B;
        g.k_split = WGRAD_SPLIT;
        g.out = s->wpart;
        gemm<EPI_F32>(g, splits, st);
        float* gw = s->grad + s->t.param_off[i];
        hipLaunchKernelGGL(surr_wgrad_reduce_kernel, dim3((N * K + 255) / 256),
                dim3(256), 0, st, N * K, splits, s->wpart, gw);
        hipLaunchKernelGGL(surr_bgrad_kernel, dim3(N), dim3(256), 0, st, B,
                dzT, gw + N * K);
        if (i == 0) break;
        GemmArgs h = {};
        h.A = dz;
        h.lda = N;
        h.Bm = s->wbT + s->t.weight_off[i];
        h.ldb = N;
        h.M = B;
        h.N = K;
        h.K = N;
        h.k_split = N;
        h.Z = s->Z[i - 1];
        h.act = s->dZ[buf];
        h.actT = s->dZT[buf];
        gemm<EPI_BWD_HIDDEN>(h, 1, st);
        dz = s->dZ[buf];
        dzT = s->dZT[buf];
        buf ^= 1;
    }
}

void pack(Surrogate* s, hipStream_t st)
{
    hipLaunchKernelGGL(surr_pack_kernel, dim3((NUM_PARAMS + 255) / 256),
            dim3(256), 0, st, s->params, s->t, s->wb, s->wbT);
}

void release(Surrogate* s)
{
    if (s->params) (void) hipFree(s->params);
    if (s->wb) (void) hipFree(s->wb);
    if (s->ws) (void) hipFree(s->ws);
    delete s;
}

} /* namespace */

extern "C" int oskar_amd_surr_num_params(void) { return NUM_PARAMS; }

/* Scale from network outputs to beam values: training targets are divided
 * by it, inference outputs multiplied. */
extern "C" void oskar_amd_surr_set_out_scale(void* h, double scale)
{
    ((Surrogate*) h)->out_scale = (float) scale;
}
extern "C" int oskar_amd_surr_num_fourier(void) { return NF; }

/* fourier: NF x 2 floats; params: NUM_PARAMS floats (host). */
extern "C" void* oskar_amd_surr_create(const float* fourier, double l0,
        double m0, double f_min, double f_max, const float* params, int* err)
{
    Surrogate* s = new Surrogate;
    s->enc.l0 = (float) l0;
    s->enc.m0 = (float) m0;
    s->enc.f_min = (float) f_min;
    s->enc.f_max = (float) f_max;
    memcpy(s->enc.fourier, fourier, sizeof(s->enc.fourier));
    s->t = layer_table();
    /* One allocation: params, mom, vel, grad (FP32), then wb, wbT. */
    const size_t fp = (size_t) NUM_PARAMS * 4;
    hipError_t e = hipMalloc((void**) &s->params, 4 * fp);
    if (e == hipSuccess)
        e = hipMalloc((void**) &s->wb, 2 * (size_t) NUM_WEIGHTS * 2);
    if (e != hipSuccess)
    {
        release(s);
        *err = (int) e;
        return nullptr;
    }
    s->mom = s->params + NUM_PARAMS;
    s->vel = s->mom + NUM_PARAMS;
    s->grad = s->vel + NUM_PARAMS;
    s->wbT = s->wb + NUM_WEIGHTS;
    e = hipMemcpy(s->params, params, fp, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(s->mom, 0, 3 * fp);
    if (e == hipSuccess) e = hipMemset(s->wb, 0, 2 * (size_t) NUM_WEIGHTS * 2);
    if (e == hipSuccess)
    {
        pack(s, 0);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess)
    {
        release(s);
        *err = (int) e;
        return nullptr;
    }
    *err = 0;
    return s;
}

extern "C" void oskar_amd_surr_free(void* h)
{
    if (!h) return;
    (void) hipDeviceSynchronize();
    release((Surrogate*) h);
}

extern "C" int oskar_amd_surr_get_params(void* h, float* params)
{
    Surrogate* s = (Surrogate*) h;
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess)
        e = hipMemcpy(params, s->params, (size_t) NUM_PARAMS * 4,
                hipMemcpyDeviceToHost);
    return (int) e;
}

extern "C" int oskar_amd_surr_set_params(void* h, const float* params)
{
    Surrogate* s = (Surrogate*) h;
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess)
        e = hipMemcpy(s->params, params, (size_t) NUM_PARAMS * 4,
                hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(s->mom, 0, (size_t) NUM_PARAMS * 12);
    if (e == hipSuccess)
    {
        s->adam_step = 0;
        pack(s, 0);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    return (int) e;
}

/* Loss and gradient on B samples given as FP32 features [B][128] and
 * targets [B][8] (device); gradient (device, NUM_PARAMS floats) may be
 * NULL. */
extern "C" int oskar_amd_surr_loss_grad(void* h, int B, const float* feat,
        const float* target, double* loss, float* grad_out, void* stream_v)
{
    Surrogate* s = (Surrogate*) h;
    const hipStream_t st = (hipStream_t) stream_v;
    if (B <= 0 || B % 256) return (int) hipErrorInvalidValue;
    hipError_t e = ws_ensure(s, B);
    if (e != hipSuccess) return (int) e;
    hipLaunchKernelGGL(surr_features_kernel, dim3((B * D_IN + 255) / 256),
            dim3(256), 0, st, B, feat, s->X, s->XT);
    e = hipMemcpyAsync(s->T, target, (size_t) B * N_OUT * 4,
            hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess) return (int) e;
    forward(s, B, st);
    backward(s, B, st);
    if (grad_out)
        e = hipMemcpyAsync(grad_out, s->grad, (size_t) NUM_PARAMS * 4,
                hipMemcpyDeviceToDevice, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(loss, s->loss, 8, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e == hipSuccess) e = hipGetLastError();
    return (int) e;
}

/* Adam training on num_samples samples (device arrays l, m, f [n], E
 * [n][8], beam units): num_steps steps of B random samples; learning rate
 * cosine-decayed from lr0 to lr1.  loss_hist (host, num_steps doubles, in
 * network units) may be NULL. */
extern "C" int oskar_amd_surr_train(void* h, int num_samples, const float* l,
        const float* m, const float* f, const float* E, int B, int num_steps,
        double lr0, double lr1, unsigned long long seed, double* loss_hist,
        void* stream_v)
{
    Surrogate* s = (Surrogate*) h;
    const hipStream_t st = (hipStream_t) stream_v;
    if (B <= 0 || B % 256 || num_samples <= 0 || num_steps < 0)
        return (int) hipErrorInvalidValue;
    hipError_t e = ws_ensure(s, B);
    if (e != hipSuccess) return (int) e;
    double* d_hist = nullptr;
    if (loss_hist && num_steps > 0)
    {
        e = hipMalloc((void**) &d_hist, (size_t) num_steps * 8);
        if (e != hipSuccess) return (int) e;
    }
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int k = 0; k < num_steps && e == hipSuccess; ++k)
    {
        hipLaunchKernelGGL(surr_gather_kernel, dim3((B + 255) / 256),
                dim3(256), 0, st, B, num_samples, (uint64_t) seed,
                s->adam_step, 1.0f / s->out_scale, E, s->idx, s->T);
        hipLaunchKernelGGL(surr_encode_kernel, dim3((B + 255) / 256),
                dim3(256), 0, st, B, l, m, f, s->idx, s->enc, s->X, s->XT);
        forward(s, B, st);
        backward(s, B, st);
        const double c = 0.5 * (1.0 + cos(M_PI * (double) k /
                (double) (num_steps > 1 ? num_steps - 1 : 1)));
        const float lr = (float) (lr1 + (lr0 - lr1) * c);
        s->adam_step += 1;
        const float bc1 = 1.0f - powf(b1, (float) s->adam_step);
        const float bc2 = 1.0f - powf(b2, (float) s->adam_step);
        hipLaunchKernelGGL(surr_adam_kernel, dim3((NUM_PARAMS + 255) / 256),
                dim3(256), 0, st, s->params, s->mom, s->vel, s->grad, lr, b1,
                b2, eps, bc1, bc2, s->t, s->wb, s->wbT);
        if (d_hist)
            e = hipMemcpyAsync(d_hist + k, s->loss, 8,
                    hipMemcpyDeviceToDevice, st);
        if (e == hipSuccess) e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e == hipSuccess && d_hist)
        e = hipMemcpy(loss_hist, d_hist, (size_t) num_steps * 8,
                hipMemcpyDeviceToHost);
    if (d_hist) (void) hipFree(d_hist);
    return (int) e;
}

/* Fused inference of num_points directions at one frequency into beam
 * (complex matrices of the given precision); flags[i] = 1 where the
 * largest |E_ij| is below flag_level (flags may be NULL). */
extern "C" int oskar_amd_surr_infer(void* h, int precision, int num_points,
        const void* l, const void* m, double freq_hz, double flag_level,
        void* beam, int* flags, void* stream_v)
{
    Surrogate* s = (Surrogate*) h;
    const hipStream_t st = (hipStream_t) stream_v;
    if (num_points <= 0) return 0;
    const dim3 grid((num_points + INF_ROWS - 1) / INF_ROWS);
    if (precision == OSKAR_DOUBLE)
        hipLaunchKernelGGL(surr_infer_kernel<double>, grid, dim3(256), 0, st,
                num_points, (const double*) l, (const double*) m,
                (float) freq_hz, s->enc, s->wb, s->params, s->t,
                s->out_scale, (float) flag_level, (Cplx<double>*) beam,
                flags);
    else
        hipLaunchKernelGGL(surr_infer_kernel<float>, grid, dim3(256), 0, st,
                num_points, (const float*) l, (const float*) m,
                (float) freq_hz, s->enc, s->wb, s->params, s->t,
                s->out_scale, (float) flag_level, (Cplx<float>*) beam,
                flags);
    return (int) hipGetLastError();
}

/* ---------------------------------------------------------------------- */
/* Fallback helpers: peak level, compaction of flagged directions, scatter */
/* of their exact beam values.                                             */

namespace {

__global__ __launch_bounds__(256)
void surr_maxabs_kernel(long long n, const float* v, unsigned int* out)
{
    __shared__ float red[256];
    float mx = 0.0f;
    for (long long i = blockIdx.x * (long long) blockDim.x + threadIdx.x;
            i < n; i += (long long) gridDim.x * blockDim.x)
        mx = fmaxf(mx, fabsf(v[i]));
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1)
    {
        if ((int) threadIdx.x < o)
            red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0)
        (void) __hip_atomic_fetch_max((GLOBAL_AS unsigned int*) out,
                __float_as_uint(red[0]), __ATOMIC_RELAXED,
                __HIP_MEMORY_SCOPE_AGENT);
}

/* Flagged directions appended to (xc, yc, zc) with their index; the order
 * of the list is arbitrary (each entry is evaluated independently). */
template <typename FP>
__global__ __launch_bounds__(256)
void surr_compact_kernel(int n, const int* flags, const FP* x, const FP* y,
        const FP* z, int* idx, FP* xc, FP* yc, FP* zc, int* counter)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !flags[i]) return;
    const int k = atomicAdd(counter, 1);
    idx[k] = i;
    xc[k] = x[i];
    yc[k] = y[i];
    zc[k] = z[i];
}

template <typename FP>
__global__ __launch_bounds__(256)
void surr_scatter_kernel(int count, const int* idx, const Cplx<FP>* src,
        Cplx<FP>* dst)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const int i = idx[k];
    for (int j = 0; j < 4; ++j) dst[4 * (size_t) i + j] = src[4 * (size_t) k + j];
}

} /* namespace */

extern "C" int oskar_amd_surr_max_abs(const float* v, long long n,
        double* out)
{
    unsigned int* d = nullptr;
    hipError_t e = hipMalloc((void**) &d, 4);
    if (e != hipSuccess) return (int) e;
    e = hipMemset(d, 0, 4);
    if (e == hipSuccess)
    {
        hipLaunchKernelGGL(surr_maxabs_kernel, dim3(1024), dim3(256), 0, 0,
                n, v, d);
        e = hipGetLastError();
    }
    unsigned int h = 0;
    if (e == hipSuccess) e = hipMemcpy(&h, d, 4, hipMemcpyDeviceToHost);
    (void) hipFree(d);
    float f;
    memcpy(&f, &h, 4);
    *out = (double) f;
    return (int) e;
}

/* work: device ints, n + 1 (the last is the counter). */
extern "C" int oskar_amd_surr_compact(int precision, int n, const int* flags,
        const void* x, const void* y, const void* z, int* work, void* xc,
        void* yc, void* zc, int* count, void* stream_v)
{
    const hipStream_t st = (hipStream_t) stream_v;
    int* counter = work + n;
    hipError_t e = hipMemsetAsync(counter, 0, sizeof(int), st);
    if (e != hipSuccess) return (int) e;
    const dim3 grid((n + 255) / 256);
    if (precision == OSKAR_DOUBLE)
        hipLaunchKernelGGL(surr_compact_kernel<double>, grid, dim3(256), 0,
                st, n, flags, (const double*) x, (const double*) y,
                (const double*) z, work, (double*) xc, (double*) yc,
                (double*) zc, counter);
    else
        hipLaunchKernelGGL(surr_compact_kernel<float>, grid, dim3(256), 0,
                st, n, flags, (const float*) x, (const float*) y,
                (const float*) z, work, (float*) xc, (float*) yc,
                (float*) zc, counter);
    e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(count, counter, sizeof(int), hipMemcpyDeviceToHost,
                st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    return (int) e;
}

extern "C" int oskar_amd_surr_scatter(int precision, int count,
        const int* idx, const void* src, void* dst, void* stream_v)
{
    const hipStream_t st = (hipStream_t) stream_v;
    if (count <= 0) return 0;
    const dim3 grid((count + 255) / 256);
    if (precision == OSKAR_DOUBLE)
        hipLaunchKernelGGL(surr_scatter_kernel<double>, grid, dim3(256), 0,
                st, count, idx, (const Cplx<double>*) src,
                (Cplx<double>*) dst);
    else
        hipLaunchKernelGGL(surr_scatter_kernel<float>, grid, dim3(256), 0,
                st, count, idx, (const Cplx<float>*) src,
                (Cplx<float>*) dst);
    return (int) hipGetLastError();
}
