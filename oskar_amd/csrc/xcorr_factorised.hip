/*
 * Station-factorised cross-correlation on gfx950 matrix cores.
 *
 * For point sources without smearing and a station beam E(s) shared by all
 * stations, the reference's
 *     V_pq = sum_s J_p(s) B_s J_q(s)^H,   J_p(s) = K_p(s) E(s)
 * (oskar/correlate/define_cross_correlate.h:34-77 with J from
 * oskar/interferometer/src/oskar_interferometer_run_block.c:251-287) factorises
 * into scalar phases K_p(s) = exp(+2 pi i (U_p l + V_p m + W_p (n - 1)))
 * (oskar/interferometer/define_evaluate_jones_K.h:56-75; U = u f / c) and
 * one Hermitian 2x2 matrix per source, C(s) = E(s) B_s E(s)^H:
 *     V_pq[i][k] = sum_s K_p(s) conj(K_q(s)) C_ik(s).
 * That is a single complex GEMM  V[p, (q,ik)] = sum_s P[p, s] conj(W[(q,ik), s])
 * with P = K and W[(q,ik), s] = K_q(s) conj(C_ik(s)): 32 real flops per
 * (baseline, source) instead of the 64 of the A A^H form.
 *
 * Precision (the kernel's reason to exist):
 *  - phase arguments, their range reduction and the direction cosines are
 *    FP64; the reduced phase (in revolutions, |r| <= 1/2) goes to the FP32
 *    hardware sine and cosine;
 *  - P and W are split into hi + lo 16-bit parts (FP16 or BF16); each
 *    product is formed as hi*hi + hi*lo + lo*hi on the MFMA units with FP32
 *    accumulation ("x3");
 *  - K is scaled by 2^12 and C by a global power of two (from a device-side
 *    maximum) so the FP16 parts stay in range and the lo parts of small
 *    components stay normal (the matrix cores do not keep FP16 subnormal
 *    inputs); both scales are removed exactly at the end;
 *  - each workgroup accumulates at most SRC_PER_PART sources in FP32, and
 *    the per-part partial sums are added in FP64 by the reduction kernel
 *    (cross-tile sums in FP64), in a fixed order (no atomics on the data).
 *
 * Real form: with X_P = [Re K, Im K] and Y_P = [Im K, -Re K] interleaved per
 * source, X_W = [Re W, Im W],  Re V = X_P X_W^T  and  Im V = Y_P X_W^T.
 *
 * Tiling: a workgroup of 256 threads (one wave64 per SIMD) owns 128
 * P-stations x 64 Q-stations (256 W rows) = 65536 FP32 accumulators held in
 * AGPRs; each wave owns 64 P rows x 128 W rows (2 x 4 MFMA 32x32 blocks,
 * Re and Im).  A k-step is 8 sources (16 reals of K): every thread builds
 * 4 P values and 2 x 4 W values into LDS (double-buffered, 64 KB), then each
 * wave issues 48 v_mfma_f32_32x32x16 instructions.  LDS operand planes are
 * laid out [k-half][row][8] so a wave's 16-byte fragment reads are
 * contiguous; W rows are ordered (ik, q) so the builders' writes are too.
 * Source data for step k+2 is loaded into registers during step k.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_common.h"
#include "private.h"

using namespace oskar_amd;

namespace {

constexpr int TP = 128;          /* P stations per tile */
constexpr int TQ = 64;           /* Q stations per tile */
constexpr int TW = TQ * 4;       /* W rows per tile */
constexpr int KSRC = 8;          /* sources per k-step */
constexpr int THREADS = 256;
constexpr int P_ENTRIES = 4 * 2 * TP;    /* uint4 per stage, P planes */
constexpr int W_ENTRIES = 2 * 2 * TW;    /* uint4 per stage, W planes */
constexpr int STAGE = P_ENTRIES + W_ENTRIES;

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));
typedef float f16v __attribute__((ext_vector_type(16)));

struct Src { double l, m, n1; };        /* n1 = n - 1 */

/* Splits x into hi + lo parts of a 16-bit type (round to nearest). */
template <typename T16> struct Split16;
template <> struct Split16<_Float16>
{
    static __device__ __forceinline__ void split(float x, _Float16* hi,
            _Float16* lo)
    {
        const _Float16 h = (_Float16) x;
        *hi = h;
        *lo = (_Float16) (x - (float) h);
    }
};
template <> struct Split16<__bf16>
{
    static __device__ __forceinline__ void split(float x, __bf16* hi,
            __bf16* lo)
    {
        const __bf16 h = (__bf16) x;
        *hi = h;
        *lo = (__bf16) (x - (float) h);
    }
};

template <typename T16> __device__ __forceinline__ f16v mfma(
        const uint4& a, const uint4& b, const f16v& c);
template <> __device__ __forceinline__ f16v mfma<_Float16>(const uint4& a,
        const uint4& b, const f16v& c)
{
    h8 av, bv;
    __builtin_memcpy(&av, &a, 16);
    __builtin_memcpy(&bv, &b, 16);
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(av, bv, c, 0, 0, 0);
}
template <> __device__ __forceinline__ f16v mfma<__bf16>(const uint4& a,
        const uint4& b, const f16v& c)
{
    b8 av, bv;
    __builtin_memcpy(&av, &a, 16);
    __builtin_memcpy(&bv, &b, 16);
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, c, 0, 0, 0);
}

/* exp(2 pi i r) for an FP64 phase in revolutions: reduce to |r| <= 1/2 in
 * FP64, then the hardware sine/cosine (which take revolutions). */
__device__ __forceinline__ void phasor(double r, float* re, float* im)
{
    const float rr = (float) (r - __builtin_rint(r));
    *re = __builtin_amdgcn_cosf(rr);
    *im = __builtin_amdgcn_sinf(rr);
}

/* Number of (P-block, Q-block) tiles holding at least one p > q pair, and
 * the mapping of a linear tile index to the block pair. */
__host__ __device__ inline int tiles_in_row(int ib, int N, int nq)
{
    const int p_max = min(TP * ib + TP - 1, N - 1);
    if (p_max < 1) return 0;
    return min(nq, (p_max - 1) / TQ + 1);
}

__host__ __device__ inline void tile_coords(int t, int N, int* ib, int* jb)
{
    const int np = (N + TP - 1) / TP, nq = (N + TQ - 1) / TQ;
    int i = 0;
    for (; i < np; ++i)
    {
        const int c = tiles_in_row(i, N, nq);
        if (t < c) break;
        t -= c;
    }
    *ib = i;
    *jb = t;
}

__host__ __device__ inline int tile_index(int ib, int jb, int N)
{
    const int nq = (N + TQ - 1) / TQ;
    int t = 0;
    for (int i = 0; i < ib; ++i) t += tiles_in_row(i, N, nq);
    return t + jb;
}

struct FactArgs
{
    int num_sources, num_stations, num_tiles, src_per_part;
    const Src* src;              /* per source: l, m, n - 1 (FP64) */
    const float4* cmat;          /* per source: C00, Re C01, Im C01, C11 */
    const double* st_uvw;        /* per station: U, V, W (revolutions) */
    const unsigned int* cmax;    /* bit pattern of max |C| (>= 0 float) */
    float* partial;              /* [part][tile][re/im][TP][TW] */
};

/* Scale of the K operand: |K| <= 1 -> <= 2^12, so lo parts of components
 * down to 2^-15 stay normal FP16 numbers. */
constexpr float K_SCALE = 4096.0f;

/* Power-of-two scale that puts max |C| just below 2^14. */
__device__ __forceinline__ float c_scale(const unsigned int* cmax)
{
    const float m = __uint_as_float(*cmax);
    if (!(m > 0.0f) || !isfinite(m)) return 1.0f;
    int e = 0;
    (void) frexpf(m, &e);          /* m = f * 2^e, 0.5 <= f < 1 */
    return ldexpf(1.0f, 14 - e);
}

/* Source data one thread needs for one k-step: 4 sources for its P row,
 * 2 sources (with C) for its W rows.  Loaded two k-steps ahead so the
 * global-memory latency hides behind the MFMAs of the current step. */
struct StepSrc
{
    Src p[4];
    Src w[2];
    float4 c[2];
};

__device__ __forceinline__ void load_step(const FactArgs& A, int sb,
        int s_last, int pg, int qg, StepSrc* o)
{
#pragma unroll
    for (int j = 0; j < 4; ++j)
        o->p[j] = A.src[min(sb + pg * 4 + j, s_last)];
#pragma unroll
    for (int j = 0; j < 2; ++j)
    {
        const int s = min(sb + qg * 2 + j, s_last);
        o->w[j] = A.src[s];
        o->c[j] = A.cmat[s];
    }
}

template <typename T16>
__global__ __launch_bounds__(THREADS, 1)
void fact_xcorr_kernel(FactArgs A)
{
    __shared__ uint4 lds[2][STAGE];
    const int N = A.num_stations;
    const int tile = blockIdx.x;
    const int part = blockIdx.y;
    int ib, jb;
    tile_coords(tile, N, &ib, &jb);
    const int p0 = ib * TP, q0 = jb * TQ;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const float scale = c_scale(A.cmax);

    /* Station constants for this thread's P and W builds; stations past N
     * get a zero mask (their rows build as zeros). */
    const int pb = tid & (TP - 1), pg = tid >> 7;     /* P: station, half */
    const int qb = tid & (TQ - 1), qg = tid >> 6;     /* W: station, group */
    const int ps = min(p0 + pb, N - 1), qs = min(q0 + qb, N - 1);
    const double pu = A.st_uvw[3 * ps], pv = A.st_uvw[3 * ps + 1],
            pw = A.st_uvw[3 * ps + 2];
    const double qu = A.st_uvw[3 * qs], qv = A.st_uvw[3 * qs + 1],
            qw = A.st_uvw[3 * qs + 2];
    const float p_mask = (p0 + pb < N) ? K_SCALE : 0.0f;
    const float q_mask = (q0 + qb < N) ? scale : 0.0f;

    const int s_begin = part * A.src_per_part;
    const int s_end = min(A.num_sources, s_begin + A.src_per_part);
    const int s_last = s_end - 1;
    const int nsteps = (s_end - s_begin + KSRC - 1) / KSRC;

    /* Builds the P and W operand planes of the k-step starting at source sb
     * from the prefetched source data, into one stage buffer.  Branch-free:
     * sources past the end of this part are masked to zero. */
    auto build = [&](int sb, const StepSrc& d, uint4* st) {
        {
            T16 xh[8], xl[8], yh[8], yl[8];
#pragma unroll
            for (int j = 0; j < 4; ++j)
            {
                const float mk = (sb + pg * 4 + j <= s_last) ? p_mask : 0.0f;
                float re, im;
                phasor(pu * d.p[j].l + pv * d.p[j].m + pw * d.p[j].n1, &re,
                        &im);
                re *= mk;
                im *= mk;
                Split16<T16>::split(re, &xh[2 * j], &xl[2 * j]);
                Split16<T16>::split(im, &xh[2 * j + 1], &xl[2 * j + 1]);
                /* Y = (Im, -Re): negation is exact in the split parts. */
                yh[2 * j] = xh[2 * j + 1];
                yl[2 * j] = xl[2 * j + 1];
                yh[2 * j + 1] = -xh[2 * j];
                yl[2 * j + 1] = -xl[2 * j];
            }
            uint4 v;
            __builtin_memcpy(&v, xh, 16); st[(0 * 2 + pg) * TP + pb] = v;
            __builtin_memcpy(&v, xl, 16); st[(1 * 2 + pg) * TP + pb] = v;
            __builtin_memcpy(&v, yh, 16); st[(2 * 2 + pg) * TP + pb] = v;
            __builtin_memcpy(&v, yl, 16); st[(3 * 2 + pg) * TP + pb] = v;
        }
        {
            /* W rows are ordered ik-major (row = ik * TQ + q), so the 64
             * threads of a wave write 64 consecutive rows. */
            T16 wh[4][4], wl[4][4];
#pragma unroll
            for (int j = 0; j < 2; ++j)
            {
                const float mk = (sb + qg * 2 + j <= s_last) ? q_mask : 0.0f;
                float kr, ki;
                phasor(qu * d.w[j].l + qv * d.w[j].m + qw * d.w[j].n1, &kr,
                        &ki);
                const float4 c = make_float4(d.c[j].x * mk, d.c[j].y * mk,
                        d.c[j].z * mk, d.c[j].w * mk);
                /* W_ik = K conj(C_ik): C00, C11 real; C10 = conj(C01). */
                float w[4][2];
                w[0][0] = kr * c.x;  w[0][1] = ki * c.x;
                w[1][0] = kr * c.y + ki * c.z;  w[1][1] = ki * c.y - kr * c.z;
                w[2][0] = kr * c.y - ki * c.z;  w[2][1] = ki * c.y + kr * c.z;
                w[3][0] = kr * c.w;  w[3][1] = ki * c.w;
#pragma unroll
                for (int ik = 0; ik < 4; ++ik)
                {
                    Split16<T16>::split(w[ik][0], &wh[ik][2 * j],
                            &wl[ik][2 * j]);
                    Split16<T16>::split(w[ik][1], &wh[ik][2 * j + 1],
                            &wl[ik][2 * j + 1]);
                }
            }
            const int half = qg >> 1, off = (qg & 1);
#pragma unroll
            for (int ik = 0; ik < 4; ++ik)
            {
                const int row = ik * TQ + qb;
                uint2 v;
                __builtin_memcpy(&v, wh[ik], 8);
                reinterpret_cast<uint2*>(
                        &st[P_ENTRIES + (0 * 2 + half) * TW + row])[off] = v;
                __builtin_memcpy(&v, wl[ik], 8);
                reinterpret_cast<uint2*>(
                        &st[P_ENTRIES + (1 * 2 + half) * TW + row])[off] = v;
            }
        }
    };

    f16v re[2][4], im[2][4];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int v = 0; v < 16; ++v) { re[a][b][v] = 0.0f; im[a][b][v] = 0.0f; }

    const int wp = wave & 1, ww = wave >> 1;
    const int lr = lane & 31, lh = lane >> 5;
    StepSrc cur;
    if (nsteps > 0)
    {
        load_step(A, s_begin, s_last, pg, qg, &cur);
        build(s_begin, cur, lds[0]);
        load_step(A, s_begin + KSRC, s_last, pg, qg, &cur);
    }
    __syncthreads();
    for (int ks = 0; ks < nsteps; ++ks)
    {
        const uint4* st = lds[ks & 1];
        /* Operand fragments of step ks first: the builder's LDS writes to
         * the other stage come after them in program order, so the MFMAs
         * need not wait for the build. */
        uint4 xh[2], xl[2], yh[2], yl[2], wh[4], wl[4];
#pragma unroll
        for (int rf = 0; rf < 2; ++rf)
        {
            const int row = wp * 64 + rf * 32 + lr;
            xh[rf] = st[(0 * 2 + lh) * TP + row];
            xl[rf] = st[(1 * 2 + lh) * TP + row];
            yh[rf] = st[(2 * 2 + lh) * TP + row];
            yl[rf] = st[(3 * 2 + lh) * TP + row];
        }
#pragma unroll
        for (int cf = 0; cf < 4; ++cf)
        {
            const int row = ww * 128 + cf * 32 + lr;
            wh[cf] = st[P_ENTRIES + (0 * 2 + lh) * TW + row];
            wl[cf] = st[P_ENTRIES + (1 * 2 + lh) * TW + row];
        }
        /* hi*hi, then hi*lo, then lo*hi: 16 independent MFMAs between
         * successive updates of one accumulator. */
#pragma unroll
        for (int rf = 0; rf < 2; ++rf)
#pragma unroll
            for (int cf = 0; cf < 4; ++cf)
            {
                re[rf][cf] = mfma<T16>(xh[rf], wh[cf], re[rf][cf]);
                im[rf][cf] = mfma<T16>(yh[rf], wh[cf], im[rf][cf]);
            }
#pragma unroll
        for (int rf = 0; rf < 2; ++rf)
#pragma unroll
            for (int cf = 0; cf < 4; ++cf)
            {
                re[rf][cf] = mfma<T16>(xh[rf], wl[cf], re[rf][cf]);
                im[rf][cf] = mfma<T16>(yh[rf], wl[cf], im[rf][cf]);
            }
#pragma unroll
        for (int rf = 0; rf < 2; ++rf)
#pragma unroll
            for (int cf = 0; cf < 4; ++cf)
            {
                re[rf][cf] = mfma<T16>(xl[rf], wh[cf], re[rf][cf]);
                im[rf][cf] = mfma<T16>(yl[rf], wh[cf], im[rf][cf]);
            }
        /* Build step ks+1 (a masked dummy after the last step) into the
         * other stage, then start the loads for step ks+2. */
        build(s_begin + (ks + 1) * KSRC, cur, lds[(ks + 1) & 1]);
        load_step(A, s_begin + (ks + 2) * KSRC, s_last, pg, qg, &cur);
        /* Schedule: operand reads, then each MFMA followed by a few VALU
         * instructions of the build, then the builder's LDS writes and the
         * next loads.  The matrix pipe runs while the VALU builds. */
        __builtin_amdgcn_sched_group_barrier(0x100, 16, 0);
#pragma unroll
        for (int i = 0; i < 48; ++i)
        {
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
            __builtin_amdgcn_sched_group_barrier(0x002, 4, 0);
        }
        __builtin_amdgcn_sched_group_barrier(0x200, 8, 0);
        __builtin_amdgcn_sched_group_barrier(0x020, 14, 0);
        __syncthreads();
    }

    /* Store this part's FP32 tile: element (row i, col j) of a 32x32 block
     * sits in lane (j + 32 * ((i / 4) % 2)), register 4 * (i / 8) + i % 4.
     * Columns are W rows, ordered ik * TQ + q. */
    float* base = A.partial + ((size_t) part * A.num_tiles + tile) *
            (size_t) (2 * TP * TW);
#pragma unroll
    for (int rf = 0; rf < 2; ++rf)
#pragma unroll
        for (int cf = 0; cf < 4; ++cf)
#pragma unroll
            for (int v = 0; v < 16; ++v)
            {
                const int i = 8 * (v / 4) + 4 * lh + (v % 4);
                const int row = wp * 64 + rf * 32 + i;
                const int col = ww * 128 + cf * 32 + lr;
                base[(size_t) row * TW + col] = re[rf][cf][v];
                base[(size_t) (TP * TW) + (size_t) row * TW + col] =
                        im[rf][cf][v];
            }
}

/* Per-source preparation: direction cosines in FP64 and C = E B E^H, with
 * the running maximum of |C| entries (order-independent atomic max on the
 * non-negative float bit pattern). */
template <typename FP>
__global__ __launch_bounds__(256)
void fact_prepare_kernel(int num_sources, const FP* I, const FP* Q,
        const FP* U, const FP* V, const FP* l, const FP* m, const FP* n,
        const Cplx<FP>* beam, Src* src, float4* cmat, unsigned int* cmax)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    float mx = 0.0f;
    if (s < num_sources)
    {
        Src o;
        o.l = (double) l[s];
        o.m = (double) m[s];
        o.n1 = (double) n[s] - 1.0;
        src[s] = o;
        /* B = [[I+Q, U+iV], [U-iV, I-Q]] */
        const double Ii = I[s], Qi = Q ? (double) Q[s] : 0.0;
        const double Ui = U ? (double) U[s] : 0.0;
        const double Vi = V ? (double) V[s] : 0.0;
        double c00, c01r, c01i, c11;
        if (beam)
        {
            const Cplx<FP>* e = beam + 4 * (size_t) s;
            const double ar = e[0].x, ai = e[0].y, br = e[1].x, bi = e[1].y;
            const double cr = e[2].x, ci = e[2].y, dr = e[3].x, di = e[3].y;
            const double ba = Ii + Qi, bd = Ii - Qi;
            /* M = E B: row 0 = (a, b), row 1 = (c, d). */
            const double m00r = ar * ba + (br * Ui + bi * Vi);
            const double m00i = ai * ba + (bi * Ui - br * Vi);
            const double m01r = br * bd + (ar * Ui - ai * Vi);
            const double m01i = bi * bd + (ar * Vi + ai * Ui);
            const double m10r = cr * ba + (dr * Ui + di * Vi);
            const double m10i = ci * ba + (di * Ui - dr * Vi);
            const double m11r = dr * bd + (cr * Ui - ci * Vi);
            const double m11i = di * bd + (cr * Vi + ci * Ui);
            /* C = M E^H: C_ik = M_i0 conj(E_k0) + M_i1 conj(E_k1). */
            c00 = m00r * ar + m00i * ai + m01r * br + m01i * bi;
            c01r = m00r * cr + m00i * ci + m01r * dr + m01i * di;
            c01i = m00i * cr - m00r * ci + m01i * dr - m01r * di;
            c11 = m10r * cr + m10i * ci + m11r * dr + m11i * di;
        }
        else
        {
            c00 = Ii + Qi; c01r = Ui; c01i = Vi; c11 = Ii - Qi;
        }
        cmat[s] = make_float4((float) c00, (float) c01r, (float) c01i,
                (float) c11);
        mx = fmaxf(fmaxf(fabsf((float) c00), fabsf((float) c11)),
                fmaxf(fabsf((float) c01r), fabsf((float) c01i)));
    }
    /* Workgroup max, then one atomic per workgroup. */
    __shared__ float red[256];
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1)
    {
        if ((int) threadIdx.x < o)
            red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0 && red[0] > 0.0f)
        atomicMax(cmax, __float_as_uint(red[0]));
}

/* Station phase constants U, V, W = (u, v, w) f / c (FP64). */
template <typename FP>
__global__ void fact_stations_kernel(int n, const FP* u, const FP* v,
        const FP* w, double f_over_c, double* out, unsigned int* cmax)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) *cmax = 0u;   /* reset the maximum for this call */
    if (i < n)
    {
        out[3 * i] = (double) u[i] * f_over_c;
        out[3 * i + 1] = (double) v[i] * f_over_c;
        out[3 * i + 2] = (double) w[i] * f_over_c;
    }
}

/* Sum of the per-part tiles in FP64, scale removal, and accumulation into
 * vis[offset_out + b] for p > q outside the uv filter.  Thread per (p, q):
 * grid y = p, x over q, so consecutive threads read consecutive partials. */
template <typename FP>
__global__ __launch_bounds__(256)
void fact_reduce_kernel(int N, int num_tiles, int num_parts, int use_casa,
        const float* partial, const unsigned int* cmax, const double* st_uvw,
        double uv_min, double uv_max, int offset_out, Cplx<FP>* vis)
{
    const int p = blockIdx.y;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= p) return;
    /* uv filter on the baseline length in wavelengths. */
    {
        const double du = st_uvw[3 * p] - st_uvw[3 * q];
        const double dv = st_uvw[3 * p + 1] - st_uvw[3 * q + 1];
        const double len = sqrt(du * du + dv * dv);
        if (len < uv_min || len > uv_max) return;
    }
    const int ib = p / TP, jb = q / TQ;
    const int t = tile_index(ib, jb, N);
    const int pl = p - ib * TP, ql = q - jb * TQ;
    double vr[4] = {0, 0, 0, 0}, vi[4] = {0, 0, 0, 0};
    for (int k = 0; k < num_parts; ++k)
    {
        const float* b = partial + ((size_t) k * num_tiles + t) *
                (size_t) (2 * TP * TW) + (size_t) pl * TW + ql;
#pragma unroll
        for (int ik = 0; ik < 4; ++ik)
        {
            vr[ik] += (double) b[ik * TQ];
            vi[ik] += (double) b[TP * TW + ik * TQ];
        }
    }
    const double inv = 1.0 / ((double) c_scale(cmax) * (double) K_SCALE);
    const int bl = q * (N - 1) - (q - 1) * q / 2 + p - q - 1;
    Cplx<FP>* o = vis + 4 * (size_t) (offset_out + bl);
#pragma unroll
    for (int ik = 0; ik < 4; ++ik)
    {
        /* CASA: V = (J_p B J_q^H)^H -> element [i][k] = conj(V[k][i]). */
        const int src = use_casa ? ((ik & 1) * 2 + (ik >> 1)) : ik;
        const double r = vr[src] * inv;
        const double i = (use_casa ? -vi[src] : vi[src]) * inv;
        o[ik].x += (FP) r;
        o[ik].y += (FP) i;
    }
}

struct Workspace
{
    void* ptr = nullptr;
    size_t size = 0;
};

void* ws_get(Workspace* w, size_t bytes, hipError_t* err)
{
    if (w->size < bytes)
    {
        if (w->ptr) (void) hipFree(w->ptr);
        w->ptr = nullptr;
        w->size = 0;
        *err = hipMalloc(&w->ptr, bytes);
        if (*err != hipSuccess) { w->ptr = nullptr; return nullptr; }
        w->size = bytes;
    }
    *err = hipSuccess;
    return w->ptr;
}

Workspace g_ws[64][2];   /* per device: [0] sources/stations, [1] partials */

/* Optional HIP-event timing of the stages of the last call (per device):
 * [0] before preparation, [1] before the MFMA product, [2] before the
 * reduction, [3] after it.  Recorded on the launch stream. */
bool g_profile = false;
hipEvent_t g_ev[64][4];
bool g_ev_ok[64];

inline void mark(int dev, int i, hipStream_t stream)
{
    if (!g_profile) return;
    if (!g_ev_ok[dev])
    {
        for (int k = 0; k < 4; ++k)
            if (hipEventCreate(&g_ev[dev][k]) != hipSuccess) return;
        g_ev_ok[dev] = true;
    }
    (void) hipEventRecord(g_ev[dev][i], stream);
}

} /* namespace */

/* Number of tiles for N stations (host). */
extern "C" int oskar_amd_fact_num_tiles(int N)
{
    const int np = (N + TP - 1) / TP, nq = (N + TQ - 1) / TQ;
    int t = 0;
    for (int i = 0; i < np; ++i) t += tiles_in_row(i, N, nq);
    return t;
}

extern "C" int oskar_amd_fact_launch(int precision, int mode, int use_casa,
        int num_sources, int num_stations, const void* beam,
        const void* I, const void* Q, const void* U, const void* V,
        const void* l, const void* m, const void* n, const void* st_u,
        const void* st_v, const void* st_w, double frequency_hz,
        double uv_min_lambda, double uv_max_lambda, int offset_out,
        void* vis, int src_per_part, void* stream_v)
{
    const hipStream_t stream = (hipStream_t) stream_v;
    const int N = num_stations, S = num_sources;
    if (N < 2 || S <= 0) return 0;
    int dev = 0;
    hipError_t err = hipGetDevice(&dev);
    if (err != hipSuccess) return (int) err;
    if (dev < 0 || dev >= 64) return (int) hipErrorInvalidDevice;
    if (src_per_part <= 0) src_per_part = 8192;
    src_per_part = ((src_per_part + KSRC - 1) / KSRC) * KSRC;
    const int parts = (S + src_per_part - 1) / src_per_part;
    const int tiles = oskar_amd_fact_num_tiles(N);

    /* Workspace 0: [cmax (256 B)][stations 3N doubles][src S x 24 B]
     * [cmat S x 16 B]; workspace 1: partials. */
    const size_t off_st = 256;
    const size_t off_src = off_st + ((3 * (size_t) N * 8 + 255) / 256) * 256;
    const size_t off_c = off_src + ((sizeof(Src) * (size_t) S + 255) / 256)
            * 256;
    const size_t bytes0 = off_c + 16 * (size_t) S;
    char* w0 = (char*) ws_get(&g_ws[dev][0], bytes0, &err);
    if (err != hipSuccess) return (int) err;
    const size_t bytes1 = (size_t) parts * tiles * 2 * TP * TW * sizeof(float);
    float* partial = (float*) ws_get(&g_ws[dev][1], bytes1, &err);
    if (err != hipSuccess) return (int) err;
    unsigned int* cmax = (unsigned int*) w0;
    double* st = (double*) (w0 + off_st);
    Src* src = (Src*) (w0 + off_src);
    float4* cmat = (float4*) (w0 + off_c);
    const double f_over_c = frequency_hz / 299792458.0;

    mark(dev, 0, stream);
    if (precision == OSKAR_DOUBLE)
    {
        hipLaunchKernelGGL(fact_stations_kernel<double>,
                dim3((N + 255) / 256), dim3(256), 0, stream, N,
                (const double*) st_u, (const double*) st_v,
                (const double*) st_w, f_over_c, st, cmax);
        hipLaunchKernelGGL(fact_prepare_kernel<double>,
                dim3((S + 255) / 256), dim3(256), 0, stream, S,
                (const double*) I, (const double*) Q, (const double*) U,
                (const double*) V, (const double*) l, (const double*) m,
                (const double*) n, (const Cplx<double>*) beam, src, cmat,
                cmax);
    }
    else
    {
        hipLaunchKernelGGL(fact_stations_kernel<float>,
                dim3((N + 255) / 256), dim3(256), 0, stream, N,
                (const float*) st_u, (const float*) st_v,
                (const float*) st_w, f_over_c, st, cmax);
        hipLaunchKernelGGL(fact_prepare_kernel<float>,
                dim3((S + 255) / 256), dim3(256), 0, stream, S,
                (const float*) I, (const float*) Q, (const float*) U,
                (const float*) V, (const float*) l, (const float*) m,
                (const float*) n, (const Cplx<float>*) beam, src, cmat, cmax);
    }
    err = hipGetLastError();
    if (err != hipSuccess) return (int) err;

    FactArgs A;
    A.num_sources = S;
    A.num_stations = N;
    A.num_tiles = tiles;
    A.src_per_part = src_per_part;
    A.src = src;
    A.cmat = cmat;
    A.st_uvw = st;
    A.cmax = cmax;
    A.partial = partial;
    mark(dev, 1, stream);
    if (mode == 1)
        hipLaunchKernelGGL(fact_xcorr_kernel<__bf16>, dim3(tiles, parts),
                dim3(THREADS), 0, stream, A);
    else
        hipLaunchKernelGGL(fact_xcorr_kernel<_Float16>, dim3(tiles, parts),
                dim3(THREADS), 0, stream, A);
    err = hipGetLastError();
    if (err != hipSuccess) return (int) err;

    const double uv_min = uv_min_lambda, uv_max = uv_max_lambda;
    mark(dev, 2, stream);
    if (precision == OSKAR_DOUBLE)
        hipLaunchKernelGGL(fact_reduce_kernel<double>,
                dim3((N + 255) / 256, N), dim3(256), 0, stream, N, tiles,
                parts, use_casa, partial, cmax, st, uv_min, uv_max,
                offset_out, (Cplx<double>*) vis);
    else
        hipLaunchKernelGGL(fact_reduce_kernel<float>,
                dim3((N + 255) / 256, N), dim3(256), 0, stream, N, tiles,
                parts, use_casa, partial, cmax, st, uv_min, uv_max,
                offset_out, (Cplx<float>*) vis);
    mark(dev, 3, stream);
    return (int) hipGetLastError();
}

extern "C" void oskar_amd_fact_set_profiling(int on) { g_profile = on != 0; }

/* Stage times (ms) of the last profiled call on the current device:
 * preparation, MFMA product, reduction.  Waits for the last event. */
extern "C" int oskar_amd_fact_last_times(double* ms)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64 ||
            !g_ev_ok[dev]) return -1;
    if (hipEventSynchronize(g_ev[dev][3]) != hipSuccess) return -1;
    for (int k = 0; k < 3; ++k)
    {
        float t = 0.0f;
        if (hipEventElapsedTime(&t, g_ev[dev][k], g_ev[dev][k + 1]) !=
                hipSuccess) return -1;
        ms[k] = (double) t;
    }
    return 0;
}
