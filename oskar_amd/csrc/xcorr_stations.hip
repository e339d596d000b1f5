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
 *    phase U l + V m + W (n - 1) in FP64, reduced in FP64, then the
 *    hardware sine and cosine), a power-of-two scale that puts the part's
 *    largest component of A near 2^14, and the split into hi + lo 16-bit
 *    planes, laid out so that one station's k-step (4 sources) is one
 *    contiguous 128-byte record;
 *  - product: a workgroup owns 64 P stations x 64 Q stations of one part;
 *    each k-step stages both stations' records in LDS (double-buffered,
 *    loaded two steps ahead) and issues hi*hi, hi*lo and lo*hi
 *    v_mfma_f32_32x32x16 instructions into FP32 accumulators; no operand
 *    is built in the loop beyond Y (a rotate and a sign flip of X);
 *  - reduction: per baseline the parts' FP32 tiles, each divided by its two
 *    stations' scales, added in FP64 in a fixed order.
 * Because every k-step reads precomputed operands, the station phases and
 * beams are evaluated once per (station, source), not once per tile.
 */
#include <hip/hip_runtime.h>
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
 * (Re A_i0, Im A_i0, Re A_i1, Im A_i1). */
template <typename FP, typename T16>
__global__ __launch_bounds__(NT)
void fs_operands_kernel(int N, int S4, int SP, const FP* E, size_t ld_e,
        const double* st, const Src* src, const float4* lmat, uint4* planes,
        float* scales)
{
    __shared__ float a_lds[8][MAX_SP];    /* A components, [comp][source] */
    __shared__ float red[NT];
    const int p = blockIdx.x, part = blockIdx.y;
    const int s0 = part * SP;
    const int ns = min(SP, S4 - s0);
    const double U = st[3 * p], V = st[3 * p + 1], W = st[3 * p + 2];
    const FP* Ep = E + 8 * ((size_t) p * ld_e);
    float mx = 0.0f;
    for (int i = threadIdx.x; i < ns; i += NT)
    {
        const int s = s0 + i;
        const float4 L0 = lmat[2 * s], L1 = lmat[2 * s + 1];
        float a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        if (L0.x != 0.0f || L0.y != 0.0f || L0.z != 0.0f || L0.w != 0.0f ||
                L1.x != 0.0f || L1.y != 0.0f || L1.z != 0.0f ||
                L1.w != 0.0f)
        {
            const Src d = src[s];
            const double r = U * d.l + V * d.m + W * d.n1;
            const float rr = (float) (r - __builtin_rint(r));
            const float kr = __builtin_amdgcn_cosf(rr);
            const float ki = __builtin_amdgcn_sinf(rr);
            const FP* e = Ep + 8 * (size_t) s;
            float er[4], ei[4];
#pragma unroll
            for (int k = 0; k < 4; ++k)
            {
                er[k] = (float) e[2 * k];
                ei[k] = (float) e[2 * k + 1];
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
                    a[4 * row + 2 * col] = kr * mr - ki * mi;
                    a[4 * row + 2 * col + 1] = kr * mi + ki * mr;
                }
#pragma unroll
            for (int k = 0; k < 8; ++k) mx = fmaxf(mx, fabsf(a[k]));
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) a_lds[k][i] = a[k];
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
    /* block_max ended with a barrier: a_lds is complete. */
    const int steps = ns / KS;
    uint4* out = planes + 8 * ((size_t) part * N + p) * (size_t) (SP / KS);
    for (int t = threadIdx.x; t < steps; t += NT)
    {
        float v[8][4];                    /* [comp][source of the step] */
#pragma unroll
        for (int k = 0; k < 8; ++k)
        {
            const float4 q = *(const float4*) &a_lds[k][KS * t];
            v[k][0] = q.x * scale;
            v[k][1] = q.y * scale;
            v[k][2] = q.z * scale;
            v[k][3] = q.w * scale;
        }
#pragma unroll
        for (int g = 0; g < 2; ++g)
#pragma unroll
            for (int i = 0; i < 2; ++i)
            {
                unsigned int hi[4], lo[4];
#pragma unroll
                for (int h = 0; h < 2; ++h)      /* source 2g + h */
#pragma unroll
                    for (int j = 0; j < 2; ++j)
                    {
                        const float xr = v[4 * i + 2 * j][2 * g + h];
                        const float xi = v[4 * i + 2 * j + 1][2 * g + h];
                        const T16 hr = (T16) xr, hi_ = (T16) xi;
                        const T16 lr_ = (T16) (xr - (float) hr);
                        const T16 li_ = (T16) (xi - (float) hi_);
                        hi[2 * h + j] = pack2<T16>(hr, hi_);
                        lo[2 * h + j] = pack2<T16>(lr_, li_);
                    }
                out[8 * (size_t) t + (g * 2 + 0) * 2 + i] =
                        make_uint4(hi[0], hi[1], hi[2], hi[3]);
                out[8 * (size_t) t + (g * 2 + 1) * 2 + i] =
                        make_uint4(lo[0], lo[1], lo[2], lo[3]);
            }
    }
}

/* ---------------------------------------------------------------------- */
/* Product: V tiles of one part in FP32.                                   */

/* Y = (Im, -Re) per complex pair of halves: rotate by 16 bits, flip the
 * sign bit of the upper half. */
__device__ __forceinline__ uint4 to_y(const uint4& x)
{
    uint4 y;
    y.x = __builtin_amdgcn_alignbit(x.x, x.x, 16) ^ 0x80000000u;
    y.y = __builtin_amdgcn_alignbit(x.y, x.y, 16) ^ 0x80000000u;
    y.z = __builtin_amdgcn_alignbit(x.z, x.z, 16) ^ 0x80000000u;
    y.w = __builtin_amdgcn_alignbit(x.w, x.w, 16) ^ 0x80000000u;
    return y;
}

__device__ __forceinline__ uint4 xor4(const uint4& x, const uint4& m)
{
    return make_uint4(x.x ^ m.x, x.y ^ m.y, x.z ^ m.z, x.w ^ m.w);
}

/* LDS stage: [side P/Q][g][hl][row] of 16-byte entries (row = 2 station
 * + i within the side's 64 stations). */
constexpr int STAGE = 2 * 2 * 2 * TR;

template <typename T16, bool SIGNS>
__device__ __forceinline__ void product_loop(int N, int p0, int q0,
        int steps, int SPQ, const uint4* base_p, const uint4* base_q,
        const unsigned char* sgn_part, uint4 (*lds)[STAGE], int tid,
        f16v (&re)[2][2], f16v (&im)[2][2])
{
    const int lane = tid & 63, wave = tid >> 6;
    const int lr = lane & 31, lh = lane >> 5;
    const int wp = wave & 1, wq = wave >> 1;
    /* Staging: each thread moves 4 of the step's 1024 entries; entry
     * x = side * 512 + station * 8 + c with c = (g * 2 + hl) * 2 + i. */
    const uint4* gsrc[4];
    int ldst[4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
    {
        const int x = tid + NT * r;
        const int side = x >> 9, j = (x >> 3) & 63, c = x & 7;
        const int g = c >> 2, hl = (c >> 1) & 1, i = c & 1;
        const int st = min((side ? q0 : p0) + j, N - 1) -
                (side ? q0 : p0);
        gsrc[r] = (side ? base_q : base_p) + 8 * (size_t) st * SPQ + c;
        ldst[r] = ((side * 2 + g) * 2 + hl) * TR + 2 * j + i;
    }
    uint4 pre[4];
    if (steps > 0)
    {
#pragma unroll
        for (int r = 0; r < 4; ++r) lds[0][ldst[r]] = gsrc[r][0];
    }
    if (steps > 1)
    {
#pragma unroll
        for (int r = 0; r < 4; ++r) pre[r] = gsrc[r][8];
    }
    __syncthreads();
    for (int ks = 0; ks < steps; ++ks)
    {
        const uint4* stg = lds[ks & 1];
        uint4 xh[2], xl[2], wh[2], wl[2];
#pragma unroll
        for (int rf = 0; rf < 2; ++rf)
        {
            const int row = 64 * wp + 32 * rf + lr;
            xh[rf] = stg[((0 * 2 + lh) * 2 + 0) * TR + row];
            xl[rf] = stg[((0 * 2 + lh) * 2 + 1) * TR + row];
        }
#pragma unroll
        for (int cf = 0; cf < 2; ++cf)
        {
            const int col = 64 * wq + 32 * cf + lr;
            wh[cf] = stg[((1 * 2 + lh) * 2 + 0) * TR + col];
            wl[cf] = stg[((1 * 2 + lh) * 2 + 1) * TR + col];
        }
        if constexpr (SIGNS)
        {
            /* D = -1 for (source, j): negate that pair on the P side. */
            const unsigned int b = (unsigned int) sgn_part[KS * ks + 2 * lh]
                    | ((unsigned int) sgn_part[KS * ks + 2 * lh + 1] << 8);
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
            yh[rf] = to_y(xh[rf]);
            yl[rf] = to_y(xl[rf]);
        }
        __builtin_amdgcn_s_setprio(2);
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
        __builtin_amdgcn_s_setprio(0);
        /* Stage step ks + 1 from registers, load step ks + 2. */
        if (ks + 1 < steps)
        {
#pragma unroll
            for (int r = 0; r < 4; ++r) lds[(ks + 1) & 1][ldst[r]] = pre[r];
        }
        if (ks + 2 < steps)
        {
#pragma unroll
            for (int r = 0; r < 4; ++r)
                pre[r] = gsrc[r][8 * (size_t) (ks + 2)];
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
    const int tile = blockIdx.x, part = blockIdx.y;
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
    hipLaunchKernelGGL((fs_product_kernel<T16>), dim3(tiles, parts),
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
