/*
 * Station-factorised cross-correlation with one station beam per station,
 * on gfx950 matrix cores.
 *
 * For point sources without smearing the reference computes
 *     V_pq = sum_s J_p(s) B_s J_q(s)^H,   J_p(s) = K_p(s) E_p(s)
 * (oskar/correlate/define_cross_correlate.h:34-77, with J from
 * oskar/interferometer/src/oskar_interferometer_run_block.c:251-287 and E
 * from oskar_evaluate_jones_E, one beam per station).  Each source's
 * brightness B_s = [[I+Q, U+iV], [U-iV, I-Q]] (define_correlate_utils.h:36-39)
 * is written B_s = L_s D_s L_s^H with D_s = diag(+-1 or 0) (an LDL^H, UDU^H
 * or, for a zero diagonal, eigen factorisation of the 2x2 Hermitian matrix,
 * chosen per source; a physical source, |pol| <= I, has D = I), so that with
 *     A_p(s) = K_p(s) E_p(s) L_s                     (2 x 2 complex)
 *     V_pq   = sum_s A_p(s) D_s A_q(s)^H,
 * i.e. V = A D A^H with A in C^{2N x 2S}: a Hermitian rank-2S update, 64
 * real flops per (baseline, source).
 *
 * Real form: rows r = (station, i), k = (source, j, re/im).  With
 * X = [Re A, Im A] and Y = [Im A, -Re A] interleaved per (s, j),
 *     Re V = X_P X_Q^T,   Im V = Y_P X_Q^T
 * (D's signs applied to the P side).
 *
 * Three kernels plus the FP64 reduction:
 *  - sources: per source l, m, n - 1 in FP64, L_s (FP32) and D_s's signs;
 *  - operands: per (station, part of SP sources) A = K E L in FP32 (the
 *    phase U l + V m + W (n - 1) in FP64, its fractional part in FP64, then
 *    the hardware sine and cosine), a power-of-two scale that puts the part's
 *    largest component of A near 2^14, and the split into hi + lo 16-bit
 *    planes, laid out so that one station's k-step (4 sources) is one
 *    contiguous 128-byte record;
 *  - product: a workgroup owns 64 P stations x 64 Q stations of one part;
 *    each pair of k-steps stages both sides' records in LDS (double-
 *    buffered, loaded two pairs ahead, rows swizzled against bank
 *    conflicts) and issues hi*hi, hi*lo and lo*hi v_mfma_f32_32x32x16
 *    instructions into FP32 accumulators; no operand is built in the loop
 *    beyond Y (a rotate and a sign flip of X); the tiles of one part run
 *    on one XCD, sharing its L2;
 *  - reduction: per baseline the parts' FP32 tiles, each divided by its two
 *    stations' scales, added in FP64 in a fixed order.
 * Because every k-step reads precomputed operands, the station phases and
 * beams are evaluated once per (station, source), not once per tile.
 */
#include <hip/hip_runtime.h>
#include <climits>
#include <stdint.h>

#include "device_common.h"
#include "private.h"
#include "workspace.h"

using namespace oskar_amd;

namespace {

constexpr int KS = 4;            /* sources per k-step */
constexpr int TS = 64;           /* stations per tile side */
constexpr int TR = 2 * TS;       /* operand rows per tile side */
constexpr int MAX_SP = 2048;     /* sources per part (operand LDS) */
constexpr int NT = 256;
static_assert(MAX_SP % NT == 0 && NT % 4 == 0, "operand sources per thread");

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));
typedef float f16v __attribute__((ext_vector_type(16)));

#define GLOBAL_AS __attribute__((address_space(1)))

struct Src { double l, m, n1; };

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

/* Tiles holding at least one p > q pair: P block ib takes Q blocks
 * 0 .. min(ib, last) (64-station blocks on both sides). */
__host__ __device__ inline int fs_num_tiles(int N)
{
    const int nb = (N + TS - 1) / TS;
    return nb * (nb + 1) / 2;
}

__host__ __device__ inline void fs_tile_coords(int t, int* ib, int* jb)
{
    int i = 0;
    while (t > i) { t -= i + 1; ++i; }
    *ib = i;
    *jb = t;
}

__host__ __device__ inline int pad4(int n) { return (n + KS - 1) / KS * KS; }

/* ---------------------------------------------------------------------- */
/* Sources: direction in FP64, B = L D L^H.                               */

/* L (column-major: L00, L10, L01, L11 as complex) and the signs of D
 * (bit j set: D_jj = -1) of one 2x2 Hermitian B = [[a, b], [conj b, d]]. */
__device__ inline void factor_b(double a, double br, double bi, double d,
        double L[8], unsigned int* sgn)
{
    for (int k = 0; k < 8; ++k) L[k] = 0.0;
    *sgn = 0u;
    const double bb = br * br + bi * bi;
    if (a == 0.0 && d == 0.0)
    {
        if (bb == 0.0) return;
        /* Eigenvectors (1, +-conj(b)/|b|)/sqrt2, eigenvalues +-|b|. */
        const double mb = sqrt(bb), s = sqrt(0.5 * mb);
        const double cr = br / mb, ci = -bi / mb;       /* conj(b)/|b| */
        L[0] = s;  L[2] = s * cr;  L[3] = s * ci;       /* column 0 */
        L[4] = s;  L[6] = -s * cr; L[7] = -s * ci;      /* column 1 */
        *sgn = 2u;
        return;
    }
    if (fabs(a) >= fabs(d))
    {
        /* LDL^H: L = [[1, 0], [conj(b)/a, 1]] sqrt|D|, D = (a, d - |b|^2/a). */
        const double e = d - bb / a;
        const double sa = sqrt(fabs(a)), se = sqrt(fabs(e));
        L[0] = sa;
        L[2] = br / a * sa;  L[3] = -bi / a * sa;
        L[6] = se;
        *sgn = (a < 0.0 ? 1u : 0u) | (e < 0.0 ? 2u : 0u);
    }
    else
    {
        /* UDU^H: U = [[1, b/d], [0, 1]] sqrt|D|, D = (a - |b|^2/d, d). */
        const double f = a - bb / d;
        const double sf = sqrt(fabs(f)), sd = sqrt(fabs(d));
        L[0] = sf;
        L[4] = br / d * sd;  L[5] = bi / d * sd;
        L[6] = sd;
        *sgn = (f < 0.0 ? 1u : 0u) | (d < 0.0 ? 2u : 0u);
    }
}

template <typename FP>
__global__ __launch_bounds__(NT)
void fs_sources_kernel(int S, int S4, int SP, const FP* I, const FP* Q,
        const FP* U, const FP* V, const FP* l, const FP* m, const FP* n,
        Src* src, float4* lmat, unsigned char* sgn, int* part_flag)
{
    const int s = blockIdx.x * NT + threadIdx.x;
    if (s >= S4) return;
    Src o = {0.0, 0.0, 0.0};
    double L[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    unsigned int sg = 0u;
    if (s < S)
    {
        o.l = (double) l[s];
        o.m = (double) m[s];
        o.n1 = (double) n[s] - 1.0;
        const double Ii = I[s], Qi = Q ? (double) Q[s] : 0.0;
        const double Ui = U ? (double) U[s] : 0.0;
        const double Vi = V ? (double) V[s] : 0.0;
        factor_b(Ii + Qi, Ui, Vi, Ii - Qi, L, &sg);
    }
    src[s] = o;
    lmat[2 * s] = make_float4((float) L[0], (float) L[1], (float) L[2],
            (float) L[3]);
    lmat[2 * s + 1] = make_float4((float) L[4], (float) L[5], (float) L[6],
            (float) L[7]);
    sgn[s] = (unsigned char) sg;
    if (sg) atomicOr(part_flag + s / SP, 1);
}

/* Station phase constants U, V, W = (u, v, w) f / c (FP64); clears the
 * per-part sign flags for the sources kernel that follows. */
template <typename FP>
__global__ void fs_stations_kernel(int N, const FP* u, const FP* v,
        const FP* w, double f_over_c, double* st, int parts, int* part_flag)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N)
    {
        st[3 * i] = (double) u[i] * f_over_c;
        st[3 * i + 1] = (double) v[i] * f_over_c;
        st[3 * i + 2] = (double) w[i] * f_over_c;
    }
    for (int k = i; k < parts; k += gridDim.x * blockDim.x) part_flag[k] = 0;
}

/* ---------------------------------------------------------------------- */
/* Operands: A = K E L per (station, source), scaled and split.           */

template <typename T16>
__device__ __forceinline__ unsigned int pack2(T16 a, T16 b)
{
    unsigned short x, y;
    __builtin_memcpy(&x, &a, 2);
    __builtin_memcpy(&y, &b, 2);
    return (unsigned int) x | ((unsigned int) y << 16);
}

/* Largest value over the workgroup (non-negative inputs). */
__device__ __forceinline__ float block_max(float v, float* red)
{
    red[threadIdx.x] = v;
    __syncthreads();
    for (int o = NT / 2; o > 0; o >>= 1)
    {
        if ((int) threadIdx.x < o)
            red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    return red[0];
}

/* Workgroup (station p = blockIdx.x, part = blockIdx.y).  Operand record
 * of (part, station, k-step t): 8 entries of 16 bytes, entry
 * (g * 2 + hl) * 2 + i for k-group g (sources 4t + 2g, 4t + 2g + 1), hi or
 * lo part hl and operand row i; an entry holds, for its two sources,
 * (Re A_i0, Im A_i0, Re A_i1, Im A_i1).  Thread x computes A of sources
 * x + k NT (k < SPT), so that its reads of E, L and the directions are
 * coalesced, and keeps A in registers until the part's scale is known.
 * Then lane h of each group of 4 lanes (one k-step) takes its partner's
 * A (lane h ^ 1, the other source of its k-group) and writes entries
 * 2h and 2h + 1: k-group h / 2, hi (h even) or lo (h odd). */
constexpr int SPT = MAX_SP / NT;                 /* sources per thread */

template <typename FP, typename T16>
__global__ __launch_bounds__(NT)
void fs_operands_kernel(int N, int S4, int SP, const FP* E, size_t ld_e,
        const double* st, const Src* src, const float4* lmat, uint4* planes,
        float* scales)
{
    __shared__ float red[NT];
    const int p = blockIdx.x, part = blockIdx.y;
    const int s0 = part * SP;
    const int ns = min(SP, S4 - s0);
    const double U = st[3 * p], V = st[3 * p + 1], W = st[3 * p + 2];
    const FP* Ep = E + 8 * ((size_t) p * ld_e);
    float a[SPT][8];
    float mx = 0.0f;
#pragma unroll
    for (int k = 0; k < SPT; ++k)
    {
#pragma unroll
        for (int c = 0; c < 8; ++c) a[k][c] = 0.0f;
        const int i = (int) threadIdx.x + k * NT;
        if (i >= ns) continue;
        const int s = s0 + i;
        const float4 L0 = lmat[2 * s], L1 = lmat[2 * s + 1];
        /* Zero L (padding past the last source, zero brightness): A = 0
         * and E is not read. */
        if (L0.x == 0.0f && L0.y == 0.0f && L0.z == 0.0f && L0.w == 0.0f &&
                L1.x == 0.0f && L1.y == 0.0f && L1.z == 0.0f &&
                L1.w == 0.0f)
            continue;
        const Src d = src[s];
        const double r = U * d.l + V * d.m + W * d.n1;
        const float rr = (float) __builtin_amdgcn_fract(r);
        const float kr = __builtin_amdgcn_cosf(rr);
        const float ki = __builtin_amdgcn_sinf(rr);
        const FP* e = Ep + 8 * (size_t) s;
        float er[4], ei[4];
#pragma unroll
        for (int q = 0; q < 4; ++q)
        {
            er[q] = (float) e[2 * q];
            ei[q] = (float) e[2 * q + 1];
        }
        /* L columns: c0 = (L00, L10), c1 = (L01, L11). */
        const float lr[4] = {L0.x, L0.z, L1.x, L1.z};
        const float li[4] = {L0.y, L0.w, L1.y, L1.w};
#pragma unroll
        for (int row = 0; row < 2; ++row)
#pragma unroll
            for (int col = 0; col < 2; ++col)
            {
                /* M = (E L)[row][col] = E[row][0] L[0][col]
                 *                      + E[row][1] L[1][col] */
                const int e0 = 2 * row, e1 = 2 * row + 1;
                const int l0 = 2 * col, l1 = 2 * col + 1;
                const float mr = er[e0] * lr[l0] - ei[e0] * li[l0] +
                        er[e1] * lr[l1] - ei[e1] * li[l1];
                const float mi = er[e0] * li[l0] + ei[e0] * lr[l0] +
                        er[e1] * li[l1] + ei[e1] * lr[l1];
                /* A = K M */
                a[k][4 * row + 2 * col] = kr * mr - ki * mi;
                a[k][4 * row + 2 * col + 1] = kr * mi + ki * mr;
            }
#pragma unroll
        for (int c = 0; c < 8; ++c) mx = fmaxf(mx, fabsf(a[k][c]));
    }
    const float m = block_max(mx, red);
    float scale = 1.0f;
    if (m > 0.0f && isfinite(m))
    {
        int ex = 0;
        (void) frexpf(m, &ex);           /* m = f 2^ex, 0.5 <= f < 1 */
        scale = ldexpf(1.0f, 14 - ex);
    }
    if (threadIdx.x == 0) scales[(size_t) part * N + p] = scale;
    uint4* out = planes + 8 * ((size_t) part * N + p) * (size_t) (SP / KS);
    const int h = (int) threadIdx.x & 3;
    const bool odd = (h & 1) != 0;
#pragma unroll
    for (int k = 0; k < SPT; ++k)
    {
        /* The k-group's two sources: this lane's and its partner's, in
         * source order (first, second). */
        float f[8], g2[8];
#pragma unroll
        for (int c = 0; c < 8; ++c)
        {
            const float mine = a[k][c] * scale;
            const float other = __shfl_xor(mine, 1);
            f[c] = odd ? other : mine;
            g2[c] = odd ? mine : other;
        }
        const int i = (int) threadIdx.x + k * NT;
        if (i >= ns) continue;
        const int t = i / KS;
#pragma unroll
        for (int row = 0; row < 2; ++row)
        {
            unsigned int w[4];
#pragma unroll
            for (int hh = 0; hh < 2; ++hh)       /* source 2g + hh */
#pragma unroll
                for (int j = 0; j < 2; ++j)
                {
                    const float* v = hh ? g2 : f;
                    float xr = v[4 * row + 2 * j];
                    float xi = v[4 * row + 2 * j + 1];
                    /* hi and lo from the one FP32 value (opaque, so that
                     * no conversion rounds an unrounded product). */
                    asm("" : "+v"(xr), "+v"(xi));
                    const T16 hr = (T16) xr, hi_ = (T16) xi;
                    if (!odd)
                        w[2 * hh + j] = pack2<T16>(hr, hi_);
                    else
                        w[2 * hh + j] = pack2<T16>(
                                (T16) (xr - (float) hr),
                                (T16) (xi - (float) hi_));
                }
            out[8 * (size_t) t + 2 * h + row] =
                    make_uint4(w[0], w[1], w[2], w[3]);
        }
    }
}

/* ---------------------------------------------------------------------- */
/* Product: V tiles of one part in FP32.                                   */

/* Y = (Im, -Re) per complex pair of halves: rotate by 16 bits, flip the
 * sign bit of the upper half. */
template <typename T16>
__device__ __forceinline__ uint4 to_y(const uint4& x)
{
    uint4 y;
    y.x = __builtin_amdgcn_alignbit(x.x, x.x, 16) ^ 0x80000000u;
    y.y = __builtin_amdgcn_alignbit(x.y, x.y, 16) ^ 0x80000000u;
    y.z = __builtin_amdgcn_alignbit(x.z, x.z, 16) ^ 0x80000000u;
    y.w = __builtin_amdgcn_alignbit(x.w, x.w, 16) ^ 0x80000000u;
    return y;
}

/* FP16: one packed multiply per word (exact): the low result is the high
 * half times 1, the high result the low half times -1 (1.0 in both halves
 * of the second source, negated for the high result). */
__device__ __forceinline__ unsigned int y_word_f16(unsigned int w)
{
    unsigned int r;
    asm("v_pk_mul_f16 %0, %1, %2 op_sel:[1,0] op_sel_hi:[0,1] neg_hi:[0,1]"
        : "=v"(r) : "v"(w), "s"(0x3C003C00u));
    return r;
}
template <>
__device__ __forceinline__ uint4 to_y<_Float16>(const uint4& x)
{
    uint4 y;
    y.x = y_word_f16(x.x);
    y.y = y_word_f16(x.y);
    y.z = y_word_f16(x.z);
    y.w = y_word_f16(x.w);
    return y;
}

__device__ __forceinline__ uint4 xor4(const uint4& x, const uint4& m)
{
    return make_uint4(x.x ^ m.x, x.y ^ m.y, x.z ^ m.z, x.w ^ m.w);
}

/* LDS stage: two k-steps, each [side P/Q][g][hl][row] of 16-byte entries
 * (row = 2 station + i within the side's 64 stations), then one entry
 * whose first two words hold the two k-steps' sign bytes. */
constexpr int SUB = 2 * 2 * 2 * TR;
constexpr int STAGE = 2 * SUB + 1;
constexpr int SIGN_WORD = 4 * (STAGE - 1);

/* Position of operand row `row` in LDS plane P (P = k-group * 2 + hi/lo
 * within a side): bits 1-3 of the row XORed with 2 P + 8 (bit 1 of the
 * row).  The 8 lanes that stage one station's entries (4 planes x 2 rows)
 * then write 8 distinct bank groups, and 16 lanes (two stations) 16;
 * a fragment read of 32 consecutive rows stays a permutation of those
 * rows. */
__device__ __forceinline__ int swz(int P, int row)
{
    return row ^ (2 * P + 8 * ((row >> 1) & 1));
}

/* Staging: each thread moves 4 of a k-step's 1024 entries; entry
 * x = side * 512 + station * 8 + c with c = (g * 2 + hl) * 2 + i.  Its
 * source (step 0) and its LDS slot. */
__device__ __forceinline__ void stage_entry(int x, int N, int p0, int q0,
        int SPQ, const uint4* base_p, const uint4* base_q, const uint4** src,
        int* slot)
{
    const int side = x >> 9, j = (x >> 3) & 63, c = x & 7;
    const int g = c >> 2, hl = (c >> 1) & 1, i = c & 1;
    const int st = min((side ? q0 : p0) + j, N - 1) - (side ? q0 : p0);
    *src = (side ? base_q : base_p) + 8 * (size_t) st * SPQ + c;
    *slot = ((side * 2 + g) * 2 + hl) * TR + swz(g * 2 + hl, 2 * j + i);
}

/* The 4 sign bytes of k-step k (one per source; the part's base is a
 * multiple of 4 bytes). */
__device__ __forceinline__ unsigned int step_signs(
        const unsigned char* sgn_part, int k)
{
    return *(const unsigned int*) (sgn_part + KS * (size_t) k);
}

/* The MFMAs of one k-step whose operands sit at stg. */
template <typename T16, bool SIGNS>
__device__ __forceinline__ void step_mfma(const uint4* stg,
        unsigned int sign_word, int lr, int lh, int wp, int wq,
        f16v (&re)[2][2], f16v (&im)[2][2])
{
    uint4 xh[2], xl[2], wh[2], wl[2];
#pragma unroll
    for (int rf = 0; rf < 2; ++rf)
    {
        const int row = 64 * wp + 32 * rf + lr;
        xh[rf] = stg[((0 * 2 + lh) * 2 + 0) * TR + swz(lh * 2 + 0, row)];
        xl[rf] = stg[((0 * 2 + lh) * 2 + 1) * TR + swz(lh * 2 + 1, row)];
    }
#pragma unroll
    for (int cf = 0; cf < 2; ++cf)
    {
        const int col = 64 * wq + 32 * cf + lr;
        wh[cf] = stg[((1 * 2 + lh) * 2 + 0) * TR + swz(lh * 2 + 0, col)];
        wl[cf] = stg[((1 * 2 + lh) * 2 + 1) * TR + swz(lh * 2 + 1, col)];
    }
    if constexpr (SIGNS)
    {
        /* D = -1 for (source, j): negate that pair on the P side
         * (bytes 2 lh and 2 lh + 1 of the step's sign word). */
        const unsigned int b = (sign_word >> (16 * lh)) & 0xffffu;
        const uint4 msk = make_uint4(
                (b & 1u) ? 0x80008000u : 0u, (b & 2u) ? 0x80008000u : 0u,
                (b & 256u) ? 0x80008000u : 0u,
                (b & 512u) ? 0x80008000u : 0u);
#pragma unroll
        for (int rf = 0; rf < 2; ++rf)
        {
            xh[rf] = xor4(xh[rf], msk);
            xl[rf] = xor4(xl[rf], msk);
        }
    }
    uint4 yh[2], yl[2];
#pragma unroll
    for (int rf = 0; rf < 2; ++rf)
    {
        yh[rf] = to_y<T16>(xh[rf]);
        yl[rf] = to_y<T16>(xl[rf]);
    }
#pragma unroll
    for (int rf = 0; rf < 2; ++rf)
#pragma unroll
        for (int cf = 0; cf < 2; ++cf)
        {
            re[rf][cf] = mfma<T16>(xh[rf], wh[cf], re[rf][cf]);
            im[rf][cf] = mfma<T16>(yh[rf], wh[cf], im[rf][cf]);
        }
#pragma unroll
    for (int rf = 0; rf < 2; ++rf)
#pragma unroll
        for (int cf = 0; cf < 2; ++cf)
        {
            re[rf][cf] = mfma<T16>(xh[rf], wl[cf], re[rf][cf]);
            im[rf][cf] = mfma<T16>(yh[rf], wl[cf], im[rf][cf]);
        }
#pragma unroll
    for (int rf = 0; rf < 2; ++rf)
#pragma unroll
        for (int cf = 0; cf < 2; ++cf)
        {
            re[rf][cf] = mfma<T16>(xl[rf], wh[cf], re[rf][cf]);
            im[rf][cf] = mfma<T16>(yl[rf], wh[cf], im[rf][cf]);
        }
}

template <typename T16, bool SIGNS>
__device__ __forceinline__ void product_loop(int N, int p0, int q0,
        int steps, int SPQ, const uint4* base_p, const uint4* base_q,
        const unsigned char* sgn_part, uint4 (*lds)[STAGE], int tid,
        f16v (&re)[2][2], f16v (&im)[2][2])
{
    const int lane = tid & 63, wave = tid >> 6;
    const int lr = lane & 31, lh = lane >> 5;
    const int wp = wave & 1, wq = wave >> 1;
    /* Named registers rather than arrays: with arrays the compiler kept
     * the prefetched entries in scratch memory. */
    const uint4 *g0, *g1, *g2, *g3;
    int d0, d1, d2, d3;
    stage_entry(tid + 0 * NT, N, p0, q0, SPQ, base_p, base_q, &g0, &d0);
    stage_entry(tid + 1 * NT, N, p0, q0, SPQ, base_p, base_q, &g1, &d1);
    stage_entry(tid + 2 * NT, N, p0, q0, SPQ, base_p, base_q, &g2, &d2);
    stage_entry(tid + 3 * NT, N, p0, q0, SPQ, base_p, base_q, &g3, &d3);
    /* Pairs of k-steps per stage.  Loads and stages are unconditional,
     * the k-step index clamped to the last one (the second k-step of a
     * pair past the end is staged but not multiplied), so that the loads
     * of pair m + 2, issued during pair m, are waited for one pair later. */
    if (steps <= 0) return;
    const int last = steps - 1, pairs = (steps + 1) / 2;
    size_t ka = 0, kb = 8 * (size_t) min(1, last);
    lds[0][d0] = g0[ka];
    lds[0][d1] = g1[ka];
    lds[0][d2] = g2[ka];
    lds[0][d3] = g3[ka];
    lds[0][SUB + d0] = g0[kb];
    lds[0][SUB + d1] = g1[kb];
    lds[0][SUB + d2] = g2[kb];
    lds[0][SUB + d3] = g3[kb];
    ka = 8 * (size_t) min(2, last);
    kb = 8 * (size_t) min(3, last);
    uint4 pa0 = g0[ka], pa1 = g1[ka], pa2 = g2[ka], pa3 = g3[ka];
    uint4 pb0 = g0[kb], pb1 = g1[kb], pb2 = g2[kb], pb3 = g3[kb];
    /* The sign bytes travel with the entries, staged by thread 0. */
    unsigned int psa = 0u, psb = 0u;
    if constexpr (SIGNS)
    {
        if (tid == 0)
        {
            ((unsigned int*) lds[0])[SIGN_WORD] = step_signs(sgn_part, 0);
            ((unsigned int*) lds[0])[SIGN_WORD + 1] =
                    step_signs(sgn_part, min(1, last));
        }
        psa = step_signs(sgn_part, min(2, last));
        psb = step_signs(sgn_part, min(3, last));
    }
    __syncthreads();
    for (int m = 0; m < pairs; ++m)
    {
        const uint4* stg = lds[m & 1];
        const unsigned int* sw = (const unsigned int*) stg;
        __builtin_amdgcn_s_setprio(2);
        step_mfma<T16, SIGNS>(stg, SIGNS ? sw[SIGN_WORD] : 0u, lr, lh, wp,
                wq, re, im);
        if (2 * m + 1 < steps)
            step_mfma<T16, SIGNS>(stg + SUB,
                    SIGNS ? sw[SIGN_WORD + 1] : 0u, lr, lh, wp, wq, re, im);
        __builtin_amdgcn_s_setprio(0);
        /* Stage pair m + 1 from registers, load pair m + 2. */
        uint4* nxt = lds[(m + 1) & 1];
        nxt[d0] = pa0;
        nxt[d1] = pa1;
        nxt[d2] = pa2;
        nxt[d3] = pa3;
        nxt[SUB + d0] = pb0;
        nxt[SUB + d1] = pb1;
        nxt[SUB + d2] = pb2;
        nxt[SUB + d3] = pb3;
        ka = 8 * (size_t) min(2 * m + 4, last);
        kb = 8 * (size_t) min(2 * m + 5, last);
        pa0 = g0[ka];
        pa1 = g1[ka];
        pa2 = g2[ka];
        pa3 = g3[ka];
        pb0 = g0[kb];
        pb1 = g1[kb];
        pb2 = g2[kb];
        pb3 = g3[kb];
        if constexpr (SIGNS)
        {
            if (tid == 0)
            {
                ((unsigned int*) nxt)[SIGN_WORD] = psa;
                ((unsigned int*) nxt)[SIGN_WORD + 1] = psb;
            }
            psa = step_signs(sgn_part, min(2 * m + 4, last));
            psb = step_signs(sgn_part, min(2 * m + 5, last));
        }
        __syncthreads();
    }
}

template <typename T16>
__global__ __launch_bounds__(NT, 2)
void fs_product_kernel(int N, int S4, int SP, int tiles, const uint4* planes,
        const unsigned char* sgn, const int* part_flag, float* partial)
{
    __shared__ uint4 lds[2][STAGE];
    /* Workgroup w runs on XCD w % 8 (round-robin dispatch); each XCD
     * takes a contiguous range of (part, tile) items, so the tiles of one
     * part, which read the same operand records, run together and share
     * that XCD's L2 (measured 28.5 -> 25.7 ms at 512 stations x 1M
     * sources).  The mapping is a bijection for any grid size. */
    const int total = (int) gridDim.x, w = (int) blockIdx.x;
    const int xq = total / 8, xr = total % 8, xcd = w % 8;
    const int item = xcd * xq + min(xcd, xr) + w / 8;
    const int tile = item % tiles, part = item / tiles;
    int ib, jb;
    fs_tile_coords(tile, &ib, &jb);
    const int p0 = TS * ib, q0 = TS * jb;
    const int SPQ = SP / KS;
    const int steps = min(SP, S4 - part * SP) / KS;
    const uint4* base = planes + 8 * (size_t) part * N * SPQ;
    const uint4* base_p = base + 8 * (size_t) p0 * SPQ;
    const uint4* base_q = base + 8 * (size_t) q0 * SPQ;
    const unsigned char* sgn_part = sgn + (size_t) part * SP;
    const int tid = threadIdx.x;
    f16v re[2][2], im[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int v = 0; v < 16; ++v) { re[a][b][v] = 0.0f; im[a][b][v] = 0.0f; }
    if (part_flag[part])
        product_loop<T16, true>(N, p0, q0, steps, SPQ, base_p, base_q,
                sgn_part, lds, tid, re, im);
    else
        product_loop<T16, false>(N, p0, q0, steps, SPQ, base_p, base_q,
                sgn_part, lds, tid, re, im);
    /* Element (row i, col j) of a 32x32 block: lane j + 32 ((i / 4) % 2),
     * register 4 (i / 8) + i % 4.  Tile layout [re/im][128 P rows][128 Q
     * rows]. */
    const int lane = tid & 63, wave = tid >> 6;
    const int lr = lane & 31, lh = lane >> 5;
    const int wp = wave & 1, wq = wave >> 1;
    float* out = partial + ((size_t) part * tiles + tile) * (2 * TR * TR);
#pragma unroll
    for (int rf = 0; rf < 2; ++rf)
#pragma unroll
        for (int cf = 0; cf < 2; ++cf)
#pragma unroll
            for (int v = 0; v < 16; ++v)
            {
                const int row = 64 * wp + 32 * rf + 8 * (v / 4) + 4 * lh +
                        (v % 4);
                const int col = 64 * wq + 32 * cf + lr;
                out[(size_t) row * TR + col] = re[rf][cf][v];
                out[(size_t) (TR * TR) + (size_t) row * TR + col] =
                        im[rf][cf][v];
            }
}

/* ---------------------------------------------------------------------- */
/* Reduction: FP64 sum over parts of each baseline's 2x2 block.            */

template <typename FP>
__global__ __launch_bounds__(NT)
void fs_reduce_kernel(int N, int tiles, int parts, int use_casa,
        const float* partial, const float* scales, const double* st,
        double uv_min, double uv_max, int offset_out, Cplx<FP>* vis)
{
    const int p = blockIdx.x;
    const int ib = p / TS, pl = p - ib * TS;
    const int t0 = ib * (ib + 1) / 2;
    for (int q = threadIdx.x; q < p; q += blockDim.x)
    {
        const double du = st[3 * p] - st[3 * q];
        const double dv = st[3 * p + 1] - st[3 * q + 1];
        const double len = sqrt(du * du + dv * dv);
        if (len < uv_min || len > uv_max) continue;
        const int jb = q / TS, ql = q - jb * TS;
        const int t = t0 + jb;
        double vr[4] = {0, 0, 0, 0}, vi[4] = {0, 0, 0, 0};
        for (int k = 0; k < parts; ++k)
        {
            const double inv = 1.0 / ((double) scales[(size_t) k * N + p] *
                    (double) scales[(size_t) k * N + q]);
            const float* b = partial + ((size_t) k * tiles + t) *
                    (2 * TR * TR) + (size_t) (2 * pl) * TR + 2 * ql;
#pragma unroll
            for (int i = 0; i < 2; ++i)
            {
                const float2 r = *(const float2*) (b + (size_t) i * TR);
                const float2 m = *(const float2*) (b + TR * TR +
                        (size_t) i * TR);
                vr[2 * i] += (double) r.x * inv;
                vr[2 * i + 1] += (double) r.y * inv;
                vi[2 * i] += (double) m.x * inv;
                vi[2 * i + 1] += (double) m.y * inv;
            }
        }
        const int bl = q * (N - 1) - (q - 1) * q / 2 + p - q - 1;
        FP* o = (FP*) (vis + 4 * (size_t) (offset_out + bl));
#pragma unroll
        for (int ik = 0; ik < 4; ++ik)
        {
            /* CASA: element [i][k] = conj(V[k][i]). */
            const int sidx = use_casa ? ((ik & 1) * 2 + (ik >> 1)) : ik;
            o[2 * ik] += (FP) vr[sidx];
            o[2 * ik + 1] += (FP) (use_casa ? -vi[sidx] : vi[sidx]);
        }
    }
}

/* HIP events around the stages of the last call (per device) when
 * profiling is on: sources + operands, product, reduction. */
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

inline size_t a256(size_t x) { return (x + 255) / 256 * 256; }

template <typename FP, typename T16>
int fs_launch_t(int dev, int use_casa, int S, int N, const void* E,
        size_t ld_e, const void* I, const void* Q, const void* U,
        const void* V, const void* l, const void* m, const void* n,
        const void* su, const void* sv, const void* sw, double freq,
        double uv_min, double uv_max, int offset_out, void* vis, int SP,
        hipStream_t stream)
{
    const int S4 = pad4(S);
    const int parts = (S4 + SP - 1) / SP;
    const int tiles = fs_num_tiles(N);
    const int SPQ = SP / KS;
    /* Grid limits: parts along y of the operand grid, (part, tile) items
     * along x of the product grid. */
    if (parts > 65535 || (size_t) tiles * parts > (size_t) INT_MAX)
        return (int) hipErrorInvalidConfiguration;
    /* Workspace: [stations 3N doubles][part flags][scales parts x N]
     * [src S4 x 24 B][L S4 x 32 B][signs S4 (padded to parts x SP)]
     * [operand planes parts x N x SP/4 x 128 B][partials]. */
    const size_t o_flag = a256(3 * (size_t) N * 8);
    const size_t o_scale = o_flag + a256(4 * (size_t) parts);
    const size_t o_src = o_scale + a256(4 * (size_t) parts * N);
    const size_t o_l = o_src + a256(sizeof(Src) * (size_t) S4);
    const size_t o_sgn = o_l + a256(32 * (size_t) S4);
    const size_t o_pl = o_sgn + a256((size_t) parts * SP);
    const size_t o_part = o_pl + a256(128 * (size_t) parts * N * SPQ);
    const size_t bytes = o_part + 4 * (size_t) parts * tiles * 2 * TR * TR;
    hipError_t err = hipSuccess;
    char* w = (char*) ws_acquire(WS_FACT_STATIONS, dev, stream, bytes, &err);
    if (err != hipSuccess) return (int) err;
    double* st = (double*) w;
    int* flag = (int*) (w + o_flag);
    float* scales = (float*) (w + o_scale);
    Src* src = (Src*) (w + o_src);
    float4* lmat = (float4*) (w + o_l);
    unsigned char* sgn = (unsigned char*) (w + o_sgn);
    uint4* planes = (uint4*) (w + o_pl);
    float* partial = (float*) (w + o_part);

    mark(dev, 0, stream);
    hipLaunchKernelGGL(fs_stations_kernel<FP>, dim3((N + 255) / 256),
            dim3(256), 0, stream, N, (const FP*) su, (const FP*) sv,
            (const FP*) sw, freq / 299792458.0, st, parts, flag);
    hipLaunchKernelGGL(fs_sources_kernel<FP>, dim3((S4 + NT - 1) / NT),
            dim3(NT), 0, stream, S, S4, SP, (const FP*) I, (const FP*) Q,
            (const FP*) U, (const FP*) V, (const FP*) l, (const FP*) m,
            (const FP*) n, src, lmat, sgn, flag);
    hipLaunchKernelGGL((fs_operands_kernel<FP, T16>), dim3(N, parts),
            dim3(NT), 0, stream, N, S4, SP, (const FP*) E, ld_e, st, src,
            lmat, planes, scales);
    err = hipGetLastError();
    if (err != hipSuccess) return (int) err;
    mark(dev, 1, stream);
    hipLaunchKernelGGL((fs_product_kernel<T16>), dim3(tiles * parts),
            dim3(NT), 0, stream, N, S4, SP, tiles, planes, sgn, flag,
            partial);
    err = hipGetLastError();
    if (err != hipSuccess) return (int) err;
    mark(dev, 2, stream);
    hipLaunchKernelGGL(fs_reduce_kernel<FP>, dim3(N), dim3(NT), 0, stream,
            N, tiles, parts, use_casa, partial, scales, st, uv_min, uv_max,
            offset_out, (Cplx<FP>*) vis);
    mark(dev, 3, stream);
    return (int) hipGetLastError();
}

} /* namespace */

/* V_pq += sum_s K_p E_p B E_q^H conj(K_q) for every p > q (see top).
 * E: num_stations rows of ld_e complex 2x2 matrices (precision of vis);
 * the first num_sources of each row are used. */
extern "C" int oskar_amd_fact_stations_launch(int precision, int mode,
        int use_casa, int num_sources, int num_stations, const void* E,
        size_t ld_e, const void* I, const void* Q, const void* U,
        const void* V, const void* l, const void* m, const void* n,
        const void* st_u, const void* st_v, const void* st_w,
        double frequency_hz, double uv_min_lambda, double uv_max_lambda,
        int offset_out, void* vis, int src_per_part, void* stream_v)
{
    const hipStream_t stream = (hipStream_t) stream_v;
    const int N = num_stations, S = num_sources;
    if (N < 2 || S <= 0) return 0;
    int dev = 0;
    hipError_t err = hipGetDevice(&dev);
    if (err != hipSuccess) return (int) err;
    if (dev < 0 || dev >= 64) return (int) hipErrorInvalidDevice;
    int SP = src_per_part <= 0 ? MAX_SP : src_per_part;
    SP = std::min(MAX_SP, std::max(KS, SP / KS * KS));
#define FS_ARGS dev, use_casa, S, N, E, ld_e, I, Q, U, V, l, m, n, st_u, \
        st_v, st_w, frequency_hz, uv_min_lambda, uv_max_lambda, offset_out, \
        vis, SP, stream
    if (precision == OSKAR_DOUBLE)
        return mode == 1 ? fs_launch_t<double, __bf16>(FS_ARGS) :
                fs_launch_t<double, _Float16>(FS_ARGS);
    return mode == 1 ? fs_launch_t<float, __bf16>(FS_ARGS) :
            fs_launch_t<float, _Float16>(FS_ARGS);
#undef FS_ARGS
}

extern "C" void oskar_amd_fact_stations_set_profiling(int on)
{
    g_profile = on != 0;
}

/* Stage times (ms) of the last profiled call on the current device:
 * sources and operands, MFMA product, reduction. */
extern "C" int oskar_amd_fact_stations_last_times(double* ms)
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
