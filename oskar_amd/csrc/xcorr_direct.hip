/*
 * Direct cross-correlation kernel for gfx950 (the "plain" visibility path).
 *
 * Computes, for every baseline (p > q) outside the uv filter,
 *     V_pq += sum_s smear_s(pq) * J_first(s) B_s J_second(s)^H
 * with first/second = (p, q) in the non-CASA convention and (q, p) in the
 * CASA convention (oskar/correlate/define_cross_correlate.h:34-77 and
 * oskar/correlate/src/oskar_cross_correlate_omp.cpp:33-156), and the scalar
 * (Stokes-I) form V_pq += sum_s smear_s I_s J_first conj(J_second)
 * (define_cross_correlate.h:113-155).
 *
 * Work decomposition (MI355X-specific, not the reference's):
 *  - A workgroup of 256 threads (four wave64s) owns a block of 32 "first"
 *    stations x 16 "second" stations.  Thread t owns first station t/8 and
 *    second stations t%8 and t%8 + 8.  In the matrix form M = J_first * B
 *    is formed once per (source, first station) of a staged tile, in LDS
 *    in place of J_first, and read by the 16 baselines that need it.
 *  - Sources are streamed in tiles of TILE sources.  The Jones values of the
 *    48 stations are loaded as 16-byte vectors, one station row of the tile
 *    per wave and instruction (64 lanes on consecutive bytes), and stored
 *    transposed in LDS as [source][station], so the inner loop reads a
 *    first-station value that is broadcast to 8 lanes and 16 consecutive
 *    second-station values.  Source parameters are stored as 4-value
 *    vectors, one LDS read each.
 *  - The second-station Jones is scaled by the smearing factor, and every
 *    term of M J^H is one fused multiply-add into the tile sum.
 *  - Single precision: each tile is summed in FP32 registers, then folded
 *    into FP64 accumulators (cross-tile sums in FP64).  Double precision
 *    accumulates in FP64 throughout.
 *  - The source range is split over gridDim.z workgroups (split-K) so a
 *    call fills all 256 CUs; partial sums go to an FP64 workspace and a
 *    second kernel adds them to vis in a fixed order.  No atomics are used,
 *    so results are identical from run to run.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>

#include "device_common.h"
#include "private.h"
#include "workspace.h"

using namespace oskar_amd;

/* Split partials of the allocation-free entry point: a fixed 64 MiB array
 * in every device's code object (8 splits of 512 stations' 130,816
 * baselines x 8 FP64 values fit).  Host code never passes this name to a
 * kernel: it is a host-side stand-in, not a device address; the device
 * address comes from hipGetSymbolAddress (scratch_address below). */
constexpr size_t XCORR_SCRATCH_BYTES = (size_t) 64 << 20;
static __device__ __align__(256)
        double scratch[XCORR_SCRATCH_BYTES / sizeof(double)];

namespace {

constexpr int BLOCK_A = 32;    /* "first" stations per workgroup */
constexpr int BLOCK_B = 16;    /* "second" stations per workgroup */
constexpr int THREADS = 256;
constexpr double OMEGA_EARTH = 7.272205217e-5; /* rad/s, define_correlate_utils.h:3 */

template <typename FP>
struct XcorrArgs
{
    int num_sources, num_stations, offset_out, use_casa;
    int src_per_split, num_splits;
    const void* jones;
    const FP *I, *Q, *U, *V, *l, *m, *n, *a, *b, *c;
    const FP *st_u, *st_v, *st_w, *st_x, *st_y;
    FP uv_min, uv_max, inv_wl, frac_bw, time_int, gha0, dec0;
    void* vis;
    double* partial;
};

/* Baseline-dependent terms, written exactly as OSKAR_BASELINE_TERMS and
 * OSKAR_BASELINE_DELTAS (define_correlate_utils.h:5-26) with p = SP, q = SQ. */
template <typename FP, bool TS>
struct BaselineTerms
{
    FP uu, vv, ww, uu2, vv2, uuvv, uv_len, du, dv, dw;

    __device__ __forceinline__ void init(const XcorrArgs<FP>& A, int sp,
            int sq)
    {
        uu = (A.st_u[sp] - A.st_u[sq]) * A.inv_wl;
        vv = (A.st_v[sp] - A.st_v[sq]) * A.inv_wl;
        ww = (A.st_w[sp] - A.st_w[sq]) * A.inv_wl;
        uu2 = uu * uu;
        vv2 = vv * vv;
        uv_len = sqrt(uu2 + vv2);
        uuvv = 2 * uu * vv;
        const FP f = pi_v<FP>() * A.frac_bw;
        uu *= f; vv *= f; ww *= f;
        du = dv = dw = (FP) 0;
        if (TS)
        {
            FP sin_ha, cos_ha, sin_dec, cos_dec;
            sincos_v<FP>(A.gha0, &sin_ha, &cos_ha);
            sincos_v<FP>(A.dec0, &sin_dec, &cos_dec);
            FP temp = pi_v<FP>() * A.inv_wl;
            const FP xx = (A.st_x[sp] - A.st_x[sq]) * temp;
            const FP yy = (A.st_y[sp] - A.st_y[sq]) * temp;
            const FP rot = ((FP) OMEGA_EARTH) * A.time_int;
            temp = (xx * sin_ha + yy * cos_ha) * rot;
            du = (xx * cos_ha - yy * sin_ha) * rot;
            dv = temp * sin_dec;
            dw = -temp * cos_dec;
        }
    }
};

template <typename FP, bool GAUSS, bool BS, bool TS>
__device__ __forceinline__ FP smearing(const BaselineTerms<FP, TS>& t,
        FP l, FP m, FP n1, FP sa, FP sb, FP sc)
{
    FP s = (FP) 1;
    if (GAUSS) s = exp_v<FP>(-(sa * t.uu2 + sb * t.uuvv + sc * t.vv2));
    if (BS) s *= sinc_v<FP>(t.uu * l + t.vv * m + t.ww * n1);
    if (TS) s *= sinc_v<FP>(t.du * l + t.dv * m + t.dw * n1);
    return s;
}

/* Number of reals per Jones value and the source-tile length.  A Jones
 * value is moved as VPE vectors of VR reals (16 bytes, or 8 for a float
 * complex scalar); TILE * VPE = 64, so one wave loads one station row of
 * a tile with consecutive lanes on consecutive bytes. */
/* Matrix kernels stage 16 sources per tile: with the FP32 inner loop
 * taking two sources per iteration for point sources without smearing
 * and one otherwise, every FP32 matrix variant fits 128 VGPRs and 26 KB of
 * LDS, four waves per SIMD. */
template <typename FP, bool MATRIX> struct Geometry
{
    static constexpr int NR = MATRIX ? 8 : 2;
    static constexpr int TILE = MATRIX ? 16 : 64;
    static constexpr int VPE = MATRIX ? (sizeof(FP) == 4 ? 2 : 4) : 1;
    static constexpr int VR = NR / VPE;
    /* Station rows of a tile moved by one wave instruction. */
    static constexpr int RPI = 64 / (TILE * VPE);
    static_assert(RPI * TILE * VPE == 64, "whole station rows per wave");
};

template <typename FP, int VR> struct alignas(VR * sizeof(FP)) Vec
{
    FP v[VR];
};

/* Four source values, read from LDS as one vector. */
template <typename FP> struct alignas(4 * sizeof(FP)) Quad
{
    FP x, y, z, w;
};

template <typename FP, bool MATRIX, bool GAUSS, bool BS, bool TS>
__global__ __launch_bounds__(THREADS)
void xcorr_direct_kernel(XcorrArgs<FP> A)
{
    using G = Geometry<FP, MATRIX>;
    constexpr int NR = G::NR;
    constexpr int TILE = G::TILE;
    constexpr int VPE = G::VPE;
    constexpr int VR = G::VR;
    using V = Vec<FP, VR>;
    /* Padding of one station row spreads the transposed stores over banks. */
    __shared__ FP sJa[TILE][BLOCK_A + 1][NR];
    __shared__ FP sJb[TILE][BLOCK_B + 1][NR];
    __shared__ Quad<FP> sFlux[TILE];      /* I, Q, U, V */
    __shared__ Quad<FP> sDir[TILE];       /* l, m, n - 1 */
    __shared__ Quad<FP> sExt[TILE];       /* a, b, c */

    const int N = A.num_stations;
    const int a0 = blockIdx.x * BLOCK_A;
    const int b0 = blockIdx.y * BLOCK_B;
    /* Skip blocks that hold no baseline with the required index order. */
    {
        const int a_last = min(a0 + BLOCK_A, N) - 1;
        const int b_last = min(b0 + BLOCK_B, N) - 1;
        const bool valid = A.use_casa ? (a0 < b_last) : (a_last > b0);
        if (!valid || a0 >= N || b0 >= N) return;
    }
    const int tid = threadIdx.x;
    const int la = tid >> 3;
    /* Second stations lb and lb + 8: the eight lanes of a wave that share
     * a first station read eight consecutive second stations, which fall
     * in distinct LDS banks. */
    const int lb = tid & 7;
    const int sa = a0 + la;
    const int sb[2] = { b0 + lb, b0 + lb + 8 };

    /* Baseline terms and validity for the two baselines of this thread:
     * in[k] = a real baseline in this block, ok[k] = also inside the uv
     * filter.  With split sources every real baseline writes its partial
     * sum and the reduction alone applies the filter. */
    BaselineTerms<FP, TS> bt[2];
    bool ok[2], in[2];
#pragma unroll
    for (int k = 0; k < 2; ++k)
    {
        in[k] = sa < N && sb[k] < N &&
                (A.use_casa ? (sa < sb[k]) : (sa > sb[k]));
        /* Out-of-range slots use the dummy pair (1, 0); N >= 2 here. */
        const int sp = in[k] ? max(sa, sb[k]) : 1;
        const int sq = in[k] ? min(sa, sb[k]) : 0;
        bt[k].init(A, sp, sq);
        ok[k] = in[k] &&
                !(bt[k].uv_len < A.uv_min || bt[k].uv_len > A.uv_max);
    }

    const int s_begin = blockIdx.z * A.src_per_split;
    const int s_end = min(A.num_sources, s_begin + A.src_per_split);

    double acc[2][NR];
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
        for (int r = 0; r < NR; ++r) acc[k][r] = 0.0;

    const FP* J = (const FP*) A.jones;
    /* Tile loads: instruction i of wave w moves station rows
     * (w + 4 i) RPI + r, r < RPI, of the block (rows 0-31 first stations,
     * 32-47 second stations); lane (r, s, q) moves vector q of source s of
     * row r.  Whole-vector loads need the Jones array aligned to a vector,
     * which element-aligned arrays are; otherwise the same mapping moves
     * one real at a time. */
    constexpr int RPI = G::RPI;
    const int lane = tid & 63, wave = tid >> 6;
    const int lq = lane % VPE, ls = (lane / VPE) % TILE;
    const int lr = lane / (TILE * VPE);
    const bool vec_ok = (((uintptr_t) J) % sizeof(V)) == 0;
    for (int s0 = s_begin; s0 < s_end; s0 += TILE)
    {
        const int ns = min(TILE, s_end - s0);
        __syncthreads();
        const bool s_ok = ls < ns;
#pragma unroll
        for (int i = 0; i < (BLOCK_A + BLOCK_B) / (4 * RPI); ++i)
        {
            const int st = (wave + 4 * i) * RPI + lr;
            const bool first = st < BLOCK_A;
            const int station = first ? a0 + st : b0 + (st - BLOCK_A);
            FP* dst = first ? &sJa[ls][st][lq * VR] :
                    &sJb[ls][st - BLOCK_A][lq * VR];
            const FP* src = J + ((size_t) station * A.num_sources + s0 +
                    ls) * NR + lq * VR;
            V x;
#pragma unroll
            for (int r = 0; r < VR; ++r) x.v[r] = (FP) 0;
            if (station < N && s_ok)
            {
                if (vec_ok)
                    x = *(const V*) src;
                else
                {
#pragma unroll
                    for (int r = 0; r < VR; ++r) x.v[r] = src[r];
                }
            }
            *(V*) dst = x;
        }
        for (int s = tid; s < TILE; s += THREADS)
        {
            const bool in = s < ns;
            const int g = s0 + s;
            const FP z = (FP) 0;
            sFlux[s] = Quad<FP>{in ? A.I[g] : z,
                    (MATRIX && in) ? A.Q[g] : z,
                    (MATRIX && in) ? A.U[g] : z,
                    (MATRIX && in) ? A.V[g] : z};
            if (BS || TS)
                sDir[s] = Quad<FP>{in ? A.l[g] : z, in ? A.m[g] : z,
                        in ? A.n[g] - (FP) 1 : z, z};
            if (GAUSS)
                sExt[s] = Quad<FP>{in ? A.a[g] : z, in ? A.b[g] : z,
                        in ? A.c[g] : z, z};
        }
        __syncthreads();
        if constexpr (MATRIX)
        {
            /* M = J_first * B once per (source, first station) of the tile,
             * in place of J_first: B = [[I+Q, U+iV], [U-iV, I-Q]]
             * (define_correlate_utils.h:36-39), M = J B
             * (define_multiply.h:69-87).  Each entry is read and written
             * by one thread. */
            for (int i = tid; i < TILE * BLOCK_A; i += THREADS)
            {
                const int s = i / BLOCK_A, a = i - s * BLOCK_A;
                FP* e = sJa[s][a];
                const Quad<FP> fl = sFlux[s];
                FP ja[8];
#pragma unroll
                for (int r = 0; r < 8; ++r) ja[r] = e[r];
                const FP Ii = fl.x, Qi = fl.y, Ui = fl.z, Vi = fl.w;
                const FP ba = Ii + Qi, bd = Ii - Qi;
                /* M.a = a*ba + b*conj(U+iV) */
                e[0] = ja[0] * ba + (ja[2] * Ui + ja[3] * Vi);
                e[1] = ja[1] * ba + (ja[3] * Ui - ja[2] * Vi);
                /* M.b = b*bd + a*(U+iV) */
                e[2] = ja[2] * bd + (ja[0] * Ui - ja[1] * Vi);
                e[3] = ja[3] * bd + (ja[0] * Vi + ja[1] * Ui);
                /* M.c = c*ba + d*conj(U+iV) */
                e[4] = ja[4] * ba + (ja[6] * Ui + ja[7] * Vi);
                e[5] = ja[5] * ba + (ja[7] * Ui - ja[6] * Vi);
                /* M.d = d*bd + c*(U+iV) */
                e[6] = ja[6] * bd + (ja[4] * Ui - ja[5] * Vi);
                e[7] = ja[7] * bd + (ja[4] * Vi + ja[5] * Ui);
            }
            __syncthreads();
        }

        FP t[2][NR];
#pragma unroll
        for (int k = 0; k < 2; ++k)
#pragma unroll
            for (int r = 0; r < NR; ++r) t[k][r] = (FP) 0;

        /* Two sources per iteration in FP32, except in the matrix form
         * with a smearing or Gaussian factor: one there (the factor's
         * registers would leave three waves per SIMD); one in FP64. */
        constexpr int UNROLL = (sizeof(FP) == 4 &&
                (!MATRIX || !(GAUSS || BS || TS))) ? 2 : 1;
#pragma unroll UNROLL
        for (int s = 0; s < TILE; ++s)
        {
            const Quad<FP> fl = sFlux[s];
            FP sm[2];
            if (GAUSS || BS || TS)
            {
                const Quad<FP> d = (BS || TS) ? sDir[s] : Quad<FP>{};
                const Quad<FP> e = GAUSS ? sExt[s] : Quad<FP>{};
#pragma unroll
                for (int k = 0; k < 2; ++k)
                    sm[k] = smearing<FP, GAUSS, BS, TS>(bt[k], d.x, d.y, d.z,
                            e.x, e.y, e.z);
            }
            const FP* ja = sJa[s][la];
            if (MATRIX)
            {
                /* The tile holds M = J_first * B (formed after staging). */
                FP M[8];
#pragma unroll
                for (int r = 0; r < 8; ++r) M[r] = ja[r];
#pragma unroll
                for (int k = 0; k < 2; ++k)
                {
                    /* t += smear * M * J_second^H (define_multiply.h:54-67):
                     * (M J^H).a = M.a conj(J.a) + M.b conj(J.b), etc.  The
                     * smearing factor scales J_second; each term is one
                     * fused multiply-add into the tile sum. */
                    const FP* jb = sJb[s][lb + 8 * k];
                    FP w[8];
#pragma unroll
                    for (int r = 0; r < 8; ++r)
                        w[r] = (GAUSS || BS || TS) ? jb[r] * sm[k] : jb[r];
                    FP* T = t[k];
                    T[0] += M[0] * w[0]; T[0] += M[1] * w[1];
                    T[0] += M[2] * w[2]; T[0] += M[3] * w[3];
                    T[1] += M[1] * w[0]; T[1] -= M[0] * w[1];
                    T[1] += M[3] * w[2]; T[1] -= M[2] * w[3];
                    T[2] += M[0] * w[4]; T[2] += M[1] * w[5];
                    T[2] += M[2] * w[6]; T[2] += M[3] * w[7];
                    T[3] += M[1] * w[4]; T[3] -= M[0] * w[5];
                    T[3] += M[3] * w[6]; T[3] -= M[2] * w[7];
                    T[4] += M[4] * w[0]; T[4] += M[5] * w[1];
                    T[4] += M[6] * w[2]; T[4] += M[7] * w[3];
                    T[5] += M[5] * w[0]; T[5] -= M[4] * w[1];
                    T[5] += M[7] * w[2]; T[5] -= M[6] * w[3];
                    T[6] += M[4] * w[4]; T[6] += M[5] * w[5];
                    T[6] += M[6] * w[6]; T[6] += M[7] * w[7];
                    T[7] += M[5] * w[4]; T[7] -= M[4] * w[5];
                    T[7] += M[7] * w[6]; T[7] -= M[6] * w[7];
                }
            }
            else
            {
                const FP Ii = fl.x;
#pragma unroll
                for (int k = 0; k < 2; ++k)
                {
                    const FP* jb = sJb[s][lb + 8 * k];
                    /* J_first * conj(J_second) * (smear * I) */
                    const FP w = (GAUSS || BS || TS) ? sm[k] * Ii : Ii;
                    const FP re = ja[0] * jb[0] + ja[1] * jb[1];
                    const FP im = ja[1] * jb[0] - ja[0] * jb[1];
                    t[k][0] += re * w;
                    t[k][1] += im * w;
                }
            }
        }
#pragma unroll
        for (int k = 0; k < 2; ++k)
#pragma unroll
            for (int r = 0; r < NR; ++r) acc[k][r] += (double) t[k][r];
    }

    /* Output: direct accumulation into vis, or per-split partials. */
#pragma unroll
    for (int k = 0; k < 2; ++k)
    {
        if (!(A.num_splits == 1 ? ok[k] : in[k])) continue;
        const int sp = max(sa, sb[k]), sq = min(sa, sb[k]);
        const int bl = sq * (N - 1) - (sq - 1) * sq / 2 + sp - sq - 1;
        if (A.num_splits == 1)
        {
            FP* v = (FP*) A.vis + (size_t) (A.offset_out + bl) * NR;
#pragma unroll
            for (int r = 0; r < NR; ++r) v[r] += (FP) acc[k][r];
        }
        else
        {
            const size_t nb = (size_t) N * (N - 1) / 2;
            double* w = A.partial + ((size_t) blockIdx.z * nb + bl) * NR;
#pragma unroll
            for (int r = 0; r < NR; ++r) w[r] = acc[k][r];
        }
    }
}

/* Adds the split partial sums to vis in split order, skipping baselines
 * outside the uv filter.  Grid: x over the higher station index p, y = the
 * lower index q. */
template <typename FP, int NR>
__global__ __launch_bounds__(256)
void xcorr_reduce_kernel(XcorrArgs<FP> A)
{
    const int N = A.num_stations;
    const size_t nb = (size_t) N * (N - 1) / 2;
    const int q = blockIdx.y;
    const int p = q + 1 + blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N) return;
    BaselineTerms<FP, false> t;
    t.init(A, p, q);
    if (t.uv_len < A.uv_min || t.uv_len > A.uv_max) return;
    const int bl = q * (N - 1) - (q - 1) * q / 2 + p - q - 1;
    double sum[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) sum[r] = 0.0;
    for (int k = 0; k < A.num_splits; ++k)
    {
        const double* w = A.partial + ((size_t) k * nb + bl) * NR;
#pragma unroll
        for (int r = 0; r < NR; ++r) sum[r] += w[r];
    }
    FP* v = (FP*) A.vis + (size_t) (A.offset_out + bl) * NR;
#pragma unroll
    for (int r = 0; r < NR; ++r) v[r] += (FP) sum[r];
}

/* FP64 workspace for the split partial sums, one per (device, stream),
 * grown on demand and kept for the life of the process. */
void* workspace_get(size_t bytes, hipStream_t stream, hipError_t* err)
{
    int dev = 0;
    *err = hipGetDevice(&dev);
    if (*err != hipSuccess) return nullptr;
    return ws_acquire(WS_XCORR_SPLIT, dev, stream, bytes, err);
}

/* Device address of `scratch` on the current device, looked up once per
 * device (with that device current) and kept. */
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_lock;
void* g_scratch_ptr[MAX_DEVICES];

hipError_t scratch_address(double** out)
{
    int dev = 0;
    hipError_t err = hipGetDevice(&dev);
    if (err != hipSuccess) return err;
    if (dev < 0 || dev >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> guard(g_scratch_lock);
    if (!g_scratch_ptr[dev])
    {
        void* p = nullptr;
        err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (err != hipSuccess) return err;
        if (!p) return hipErrorInvalidSymbol;
        g_scratch_ptr[dev] = p;
    }
    *out = (double*) g_scratch_ptr[dev];
    return hipSuccess;
}

/* fixed: the device's scratch array (allocation-free entry point), whose
 * capacity caps the split count; NULL: the per-stream workspace. */
template <typename FP, bool MATRIX, bool GAUSS, bool BS, bool TS>
hipError_t launch_t(XcorrArgs<FP> A, hipStream_t stream, double* fixed)
{
    using G = Geometry<FP, MATRIX>;
    const int N = A.num_stations;
    const int na = (N + BLOCK_A - 1) / BLOCK_A;
    const int nbk = (N + BLOCK_B - 1) / BLOCK_B;
    /* Count the blocks that do work, then split the sources so the launch
     * holds about 2048 working workgroups (8 per CU) with at least 4 tiles
     * of sources each. */
    int valid = 0;
    for (int ia = 0; ia < na; ++ia)
        for (int ib = 0; ib < nbk; ++ib)
        {
            const int a0 = ia * BLOCK_A, b0 = ib * BLOCK_B;
            const int a_last = min(a0 + BLOCK_A, N) - 1;
            const int b_last = min(b0 + BLOCK_B, N) - 1;
            if (A.use_casa ? (a0 < b_last) : (a_last > b0)) ++valid;
        }
    if (valid == 0 || A.num_sources <= 0) return hipSuccess;
    const int min_per_split = 4 * G::TILE;
    int splits = (2048 + valid - 1) / valid;
    const int max_splits = max(1, A.num_sources / min_per_split);
    splits = max(1, min(splits, min(max_splits, 256)));
    const size_t nb = (size_t) N * (N - 1) / 2;
    if (fixed)
    {
        /* As many splits as the scratch array holds (one split writes vis
         * directly and needs none). */
        const size_t cap = XCORR_SCRATCH_BYTES / (nb * G::NR * sizeof(double));
        splits = max(1, (int) min((size_t) splits, cap));
    }
    int per = (A.num_sources + splits - 1) / splits;
    per = ((per + G::TILE - 1) / G::TILE) * G::TILE;
    splits = (A.num_sources + per - 1) / per;
    A.src_per_split = per;
    A.num_splits = splits;
    A.partial = nullptr;
    if (splits > 1 && fixed) A.partial = fixed;
    else if (splits > 1)
    {
        hipError_t err = hipSuccess;
        A.partial = (double*) workspace_get(
                (size_t) splits * nb * G::NR * sizeof(double), stream, &err);
        if (err != hipSuccess) return err;
    }
    dim3 grid(na, nbk, splits);
    hipLaunchKernelGGL((xcorr_direct_kernel<FP, MATRIX, GAUSS, BS, TS>),
            grid, dim3(THREADS), 0, stream, A);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess || splits == 1) return err;
    hipLaunchKernelGGL((xcorr_reduce_kernel<FP, G::NR>),
            dim3((unsigned) ((N + 255) / 256), (unsigned) (N - 1)), dim3(256),
            0, stream, A);
    return hipGetLastError();
}

template <typename FP, bool MATRIX>
hipError_t dispatch(const XcorrArgs<FP>& A, bool gauss, bool bs, bool ts,
        hipStream_t s, double* fixed)
{
#define XC_CASE(G_, B_, T_) if (gauss == G_ && bs == B_ && ts == T_) \
        return launch_t<FP, MATRIX, G_, B_, T_>(A, s, fixed);
    XC_CASE(false, false, false) XC_CASE(false, false, true)
    XC_CASE(false, true, false)  XC_CASE(false, true, true)
    XC_CASE(true, false, false)  XC_CASE(true, false, true)
    XC_CASE(true, true, false)   XC_CASE(true, true, true)
#undef XC_CASE
    return hipErrorInvalidValue;
}

template <typename FP>
hipError_t build_and_launch(int matrix, const oskar_XcorrParams* p,
        const void* jones, const void* I, const void* Q, const void* U,
        const void* V, const void* l, const void* m, const void* n,
        const void* a, const void* b, const void* c, const void* st_u,
        const void* st_v, const void* st_w, const void* st_x,
        const void* st_y, void* vis, hipStream_t stream, bool no_alloc)
{
    XcorrArgs<FP> A;
    A.num_sources = p->num_sources;
    A.num_stations = p->num_stations;
    A.offset_out = p->offset_out;
    A.use_casa = p->use_casa;
    A.src_per_split = 0;
    A.num_splits = 1;
    A.jones = jones;
    A.I = (const FP*) I; A.Q = (const FP*) Q;
    A.U = (const FP*) U; A.V = (const FP*) V;
    A.l = (const FP*) l; A.m = (const FP*) m; A.n = (const FP*) n;
    A.a = (const FP*) a; A.b = (const FP*) b; A.c = (const FP*) c;
    A.st_u = (const FP*) st_u; A.st_v = (const FP*) st_v;
    A.st_w = (const FP*) st_w;
    A.st_x = (const FP*) st_x; A.st_y = (const FP*) st_y;
    A.uv_min = (FP) p->uv_min_lambda;
    A.uv_max = (FP) p->uv_max_lambda;
    A.inv_wl = (FP) p->inv_wavelength;
    A.frac_bw = (FP) p->frac_bandwidth;
    A.time_int = (FP) p->time_int_sec;
    A.gha0 = (FP) p->gha0_rad;
    A.dec0 = (FP) p->dec0_rad;
    A.vis = vis;
    A.partial = nullptr;
    if (A.num_stations < 2) return hipSuccess;
    double* fixed = nullptr;
    if (no_alloc)
    {
        const hipError_t err = scratch_address(&fixed);
        if (err != hipSuccess) return err;
    }
    if (matrix)
        return dispatch<FP, true>(A, p->gaussian != 0,
                p->bandwidth_smearing != 0, p->time_smearing != 0, stream,
                fixed);
    return dispatch<FP, false>(A, p->gaussian != 0,
            p->bandwidth_smearing != 0, p->time_smearing != 0, stream, fixed);
}

} /* namespace */

extern "C" int oskar_amd_xcorr_launch(int precision, int matrix,
        const oskar_XcorrParams* p, const void* jones,
        const void* I, const void* Q, const void* U, const void* V,
        const void* l, const void* m, const void* n,
        const void* a, const void* b, const void* c,
        const void* st_u, const void* st_v, const void* st_w,
        const void* st_x, const void* st_y, void* vis, void* stream)
{
    if (precision == OSKAR_DOUBLE)
        return (int) build_and_launch<double>(matrix, p, jones, I, Q, U, V,
                l, m, n, a, b, c, st_u, st_v, st_w, st_x, st_y, vis,
                (hipStream_t) stream, false);
    return (int) build_and_launch<float>(matrix, p, jones, I, Q, U, V,
            l, m, n, a, b, c, st_u, st_v, st_w, st_x, st_y, vis,
            (hipStream_t) stream, false);
}

extern "C" int oskar_amd_xcorr_launch_no_alloc(int precision, int matrix,
        const oskar_XcorrParams* p, const void* jones,
        const void* I, const void* Q, const void* U, const void* V,
        const void* l, const void* m, const void* n,
        const void* a, const void* b, const void* c,
        const void* st_u, const void* st_v, const void* st_w,
        const void* st_x, const void* st_y, void* vis, void* stream)
{
    if (precision == OSKAR_DOUBLE)
        return (int) build_and_launch<double>(matrix, p, jones, I, Q, U, V,
                l, m, n, a, b, c, st_u, st_v, st_w, st_x, st_y, vis,
                (hipStream_t) stream, true);
    return (int) build_and_launch<float>(matrix, p, jones, I, Q, U, V,
            l, m, n, a, b, c, st_u, st_v, st_w, st_x, st_y, vis,
            (hipStream_t) stream, true);
}
