/*
 * Auto-correlation kernels (oskar_auto_correlate) for gfx950.
 *
 * Reference: oskar/correlate/define_auto_correlate.h:7-58 (matrix) and
 * :60-104 (scalar); launcher oskar/correlate/src/oskar_auto_correlate.c:22-151
 * (one 128-thread workgroup per station, a strided sum over sources and a
 * serial sum of the 128 partials by thread 0).
 *
 *   matrix: vis[offset_out + s] += sum_i J_s(i) B_i J_s(i)^H, with the
 *           imaginary parts of the diagonal (a, d) left unchanged;
 *   scalar: vis[offset_out + s].re += sum_i |J_s(i)|^2 I_i.
 *
 * The sum reads every Jones value once (32 B per (station, source) for FP32
 * matrices, 64 B for FP64, 8 / 16 B for scalars) and does ~70 flops with it,
 * so it is bound by HBM bandwidth.  Layout: jones[station][source], source
 * fastest, as oskar_Jones.
 *
 * Design:
 *  - workgroup (chunk c, station group g), 256 threads: the threads stride
 *    the chunk's sources; each loads one source's I, Q, U, V once and the
 *    Jones values of that source for the G stations of the group, so a wave
 *    reads G contiguous runs of 64 Jones values with 2G (FP32) independent
 *    16-byte loads in flight per lane, and the fluxes come from HBM once per
 *    group instead of once per station;
 *  - per thread: FP sums of the six real terms of J B J^H that reach the
 *    output (a.re, b.re, b.im, c.re, c.im, d.re);
 *  - the 64 lanes of a wave are summed with cross-lane shuffles in FP, the
 *    four waves in FP64 through LDS, into partial[c][station][term] (FP64);
 *  - a second kernel sums each (station, term) over the chunks in FP64 in a
 *    fixed order and adds it to vis, so results do not depend on timing.
 * The launch splits sources into enough chunks for about 2048 workgroups.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "private.h"

namespace {

constexpr int THREADS = 256;

template <typename FP, bool MATRIX> struct AcorrCfg;
template <> struct AcorrCfg<float, true>   { static constexpr int G = 8; };
template <> struct AcorrCfg<double, true>  { static constexpr int G = 4; };
template <> struct AcorrCfg<float, false>  { static constexpr int G = 8; };
template <> struct AcorrCfg<double, false> { static constexpr int G = 8; };

template <bool MATRIX> struct Terms { static constexpr int NT = MATRIX ? 6 : 1; };

/* Vector type of the Jones loads: one value is NR reals, loaded as NV
 * aligned vectors (float4 x 2, double2 x 4, float2, double2). */
template <typename FP, bool MATRIX> struct JonesVec;
template <> struct JonesVec<float, true>   { typedef float4 T;  static constexpr int NV = 2; };
template <> struct JonesVec<double, true>  { typedef double2 T; static constexpr int NV = 4; };
template <> struct JonesVec<float, false>  { typedef float2 T;  static constexpr int NV = 1; };
template <> struct JonesVec<double, false> { typedef double2 T; static constexpr int NV = 1; };

/* The six output terms of J B J^H for one Jones matrix A (8 reals
 * a.re a.im b.re b.im c.re c.im d.re d.im) and B = [[I+Q, U+iV],
 * [U-iV, I-Q]]: M = A B, then M A^H (the products of
 * oskar/math/define_multiply.h:54-87). */
template <typename FP>
__device__ __forceinline__ void jbjh_terms(const FP* A, FP i_, FP q_, FP u_,
        FP v_, FP* acc)
{
    const FP ba = i_ + q_, bd = i_ - q_;
    FP M[8];
    M[0] = A[0] * ba + (A[2] * u_ + A[3] * v_);
    M[1] = A[1] * ba + (A[3] * u_ - A[2] * v_);
    M[2] = A[2] * bd + (A[0] * u_ - A[1] * v_);
    M[3] = A[3] * bd + (A[0] * v_ + A[1] * u_);
    M[4] = A[4] * ba + (A[6] * u_ + A[7] * v_);
    M[5] = A[5] * ba + (A[7] * u_ - A[6] * v_);
    M[6] = A[6] * bd + (A[4] * u_ - A[5] * v_);
    M[7] = A[7] * bd + (A[4] * v_ + A[5] * u_);
    acc[0] += (M[0] * A[0] + M[1] * A[1]) + (M[2] * A[2] + M[3] * A[3]);
    acc[1] += (M[0] * A[4] + M[1] * A[5]) + (M[2] * A[6] + M[3] * A[7]);
    acc[2] += (M[1] * A[4] - M[0] * A[5]) + (M[3] * A[6] - M[2] * A[7]);
    acc[3] += (M[4] * A[0] + M[5] * A[1]) + (M[6] * A[2] + M[7] * A[3]);
    acc[4] += (M[5] * A[0] - M[4] * A[1]) + (M[7] * A[2] - M[6] * A[3]);
    acc[5] += (M[4] * A[4] + M[5] * A[5]) + (M[6] * A[6] + M[7] * A[7]);
}

template <typename FP, bool MATRIX>
__global__ __launch_bounds__(THREADS)
void acorr_partial_kernel(int num_sources, int num_stations,
        int src_per_chunk, const FP* __restrict__ jones,
        const FP* __restrict__ I, const FP* __restrict__ Q,
        const FP* __restrict__ U, const FP* __restrict__ V,
        double* __restrict__ partial)
{
    constexpr int G = AcorrCfg<FP, MATRIX>::G;
    constexpr int NT = Terms<MATRIX>::NT;
    constexpr int NR = MATRIX ? 8 : 2;
    typedef typename JonesVec<FP, MATRIX>::T VT;
    constexpr int NV = JonesVec<FP, MATRIX>::NV;
    static_assert(NV * sizeof(VT) == NR * sizeof(FP), "Jones load size");
    __shared__ double red[THREADS / 64][G * NT];

    const int c = blockIdx.x;
    const int st0 = blockIdx.y * G;
    const int ng = min(G, num_stations - st0);
    const int s_begin = c * src_per_chunk;
    const int s_end = min(num_sources, s_begin + src_per_chunk);

    FP acc[G][NT];
#pragma unroll
    for (int k = 0; k < G; ++k)
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[k][t] = (FP) 0;

    for (int s = s_begin + (int) threadIdx.x; s < s_end; s += THREADS)
    {
        const FP i_ = I[s];
        FP q_ = 0, u_ = 0, v_ = 0;
        if constexpr (MATRIX) { q_ = Q[s]; u_ = U[s]; v_ = V[s]; }
        FP A[G][NR];
#pragma unroll
        for (int k = 0; k < G; ++k)
        {
            if (k < ng)
            {
                const FP* src = jones + ((size_t) (st0 + k) *
                        (size_t) num_sources + (size_t) s) * NR;
#pragma unroll
                for (int j = 0; j < NV; ++j)
                    *reinterpret_cast<VT*>(&A[k][j * (NR / NV)]) =
                            reinterpret_cast<const VT*>(src)[j];
            }
            else
            {
#pragma unroll
                for (int j = 0; j < NR; ++j) A[k][j] = (FP) 0;
            }
        }
#pragma unroll
        for (int k = 0; k < G; ++k)
        {
            if constexpr (MATRIX)
                jbjh_terms<FP>(A[k], i_, q_, u_, v_, acc[k]);
            else
                acc[k][0] += (A[k][0] * A[k][0] + A[k][1] * A[k][1]) * i_;
        }
    }

    /* Lanes of a wave, then waves. */
    const int lane = (int) threadIdx.x & 63, wave = (int) threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < G; ++k)
#pragma unroll
        for (int t = 0; t < NT; ++t)
        {
            FP x = acc[k][t];
#pragma unroll
            for (int m = 32; m >= 1; m >>= 1) x += __shfl_xor(x, m);
            if (lane == 0) red[wave][k * NT + t] = (double) x;
        }
    __syncthreads();
    if ((int) threadIdx.x < ng * NT)
    {
        const int e = (int) threadIdx.x;
        double sum = 0.0;
#pragma unroll
        for (int w = 0; w < THREADS / 64; ++w) sum += red[w][e];
        partial[((size_t) c * num_stations + st0) * NT + e] = sum;
    }
}

template <typename FP, bool MATRIX>
__global__ __launch_bounds__(THREADS)
void acorr_finish_kernel(int num_stations, int num_chunks, int offset_out,
        const double* __restrict__ partial, FP* __restrict__ vis)
{
    constexpr int NT = Terms<MATRIX>::NT;
    constexpr int NR = MATRIX ? 8 : 2;
    const int e = (int) (blockIdx.x * blockDim.x + threadIdx.x);
    if (e >= num_stations * NT) return;
    double sum = 0.0;
    for (int c = 0; c < num_chunks; ++c)
        sum += partial[(size_t) c * num_stations * NT + e];
    const int st = e / NT, t = e - st * NT;
    /* Terms a.re, b.re, b.im, c.re, c.im, d.re sit at reals 0, 2..6. */
    const int r = (t == 0) ? 0 : t + 1;
    vis[((size_t) offset_out + st) * NR + r] += (FP) sum;
}

/* Per-device FP64 partial-sum buffer, grown on demand and kept for the life
 * of the process.  Growing frees the old buffer (hipFree waits for the
 * device), so earlier launches that still read it have finished. */
void* acorr_workspace(size_t bytes, hipError_t* err)
{
    static void* ptr[64];
    static size_t size[64];
    int dev = 0;
    *err = hipGetDevice(&dev);
    if (*err != hipSuccess) return nullptr;
    if (dev < 0 || dev >= 64) { *err = hipErrorInvalidDevice; return nullptr; }
    if (size[dev] < bytes)
    {
        if (ptr[dev]) (void) hipFree(ptr[dev]);
        ptr[dev] = nullptr;
        size[dev] = 0;
        *err = hipMalloc(&ptr[dev], bytes);
        if (*err != hipSuccess) { ptr[dev] = nullptr; return nullptr; }
        size[dev] = bytes;
    }
    return ptr[dev];
}

template <typename FP, bool MATRIX>
hipError_t launch_t(int num_sources, int num_stations, int offset_out,
        const void* jones, const void* I, const void* Q, const void* U,
        const void* V, void* vis, hipStream_t stream)
{
    constexpr int G = AcorrCfg<FP, MATRIX>::G;
    constexpr int NT = Terms<MATRIX>::NT;
    const int groups = (num_stations + G - 1) / G;
    /* About 2048 workgroups, each with at least 4 sources per thread. */
    const int max_chunks = max(1, num_sources / (4 * THREADS));
    int chunks = max(1, min((2048 + groups - 1) / groups, max_chunks));
    int per = (num_sources + chunks - 1) / chunks;
    per = ((per + THREADS - 1) / THREADS) * THREADS;
    chunks = (num_sources + per - 1) / per;
    hipError_t err = hipSuccess;
    double* partial = (double*) acorr_workspace(
            (size_t) chunks * num_stations * NT * sizeof(double), &err);
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL((acorr_partial_kernel<FP, MATRIX>),
            dim3(chunks, groups), dim3(THREADS), 0, stream, num_sources,
            num_stations, per, (const FP*) jones, (const FP*) I,
            (const FP*) Q, (const FP*) U, (const FP*) V, partial);
    err = hipGetLastError();
    if (err != hipSuccess) return err;
    const int n = num_stations * NT;
    hipLaunchKernelGGL((acorr_finish_kernel<FP, MATRIX>),
            dim3((n + THREADS - 1) / THREADS), dim3(THREADS), 0, stream,
            num_stations, chunks, offset_out, (const double*) partial,
            (FP*) vis);
    return hipGetLastError();
}

} /* namespace */

extern "C" int oskar_amd_acorr_launch(int precision, int matrix,
        int num_sources, int num_stations, int offset_out, const void* jones,
        const void* I, const void* Q, const void* U, const void* V,
        void* vis, void* stream)
{
    if (num_sources <= 0 || num_stations <= 0) return 0;
    hipStream_t s = (hipStream_t) stream;
    hipError_t e;
    if (precision == OSKAR_SINGLE)
        e = matrix ? launch_t<float, true>(num_sources, num_stations,
                        offset_out, jones, I, Q, U, V, vis, s)
                   : launch_t<float, false>(num_sources, num_stations,
                        offset_out, jones, I, Q, U, V, vis, s);
    else
        e = matrix ? launch_t<double, true>(num_sources, num_stations,
                        offset_out, jones, I, Q, U, V, vis, s)
                   : launch_t<double, false>(num_sources, num_stations,
                        offset_out, jones, I, Q, U, V, vis, s);
    return oskar_amd_hip_status((int) e);
}
