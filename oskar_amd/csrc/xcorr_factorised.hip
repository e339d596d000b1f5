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
 *  - K is scaled by 2^12 and C by a power of two (from a device-side
 *    maximum) so the FP16 parts stay in range and the lo parts of small
 *    components stay normal (the matrix cores do not keep FP16 subnormal
 *    inputs); both scales are removed exactly at the end;
 *  - each workgroup accumulates at most src_per_part sources in FP32, and
 *    the per-part partial sums are added in FP64 by the reduction kernel
 *    (cross-tile sums in FP64), in a fixed order (no atomics on the data).
 *
 * Real form: with X_P = [Re K, Im K] and Y_P = [Im K, -Re K] interleaved per
 * source, X_W = [Re W, Im W],  Re V = X_P X_W^T  and  Im V = Y_P X_W^T.
 *
 * Tiling (template Tile<TP, TQ, OCC>): a workgroup of 256 threads (four wave64s,
 * 2 x 2) owns TP P-stations x TQ Q-stations (TW = 4 TQ W rows) of FP32
 * accumulators held in AGPRs; each wave owns TP/2 P rows x TW/2 W rows as
 * RF x CF MFMA 32x32 blocks, Re and Im.  A k-step is 8 sources (16 reals of
 * K): the threads build the P and W values of the step into LDS
 * (double-buffered), then each wave issues 6 RF CF v_mfma_f32_32x32x16
 * instructions.  LDS operand planes are laid out [k-half][row][8] so a wave's
 * 16-byte fragment reads are contiguous; W rows are ordered (ik, q) so the
 * builders' writes are too.  Source data for step k+2 is loaded into
 * registers during step k.
 *
 * Launch forms sharing the tile body (tile_part):
 *  - one call: grid (tiles, parts) over one (channel, time) unit;
 *  - a batch: a persistent grid pulls (unit, part, tile) items from a device
 *    counter over many units with different source counts; every size and
 *    pointer is read from device-side unit descriptors, so the four launches
 *    of a batch are captured once in a hipGraph and replayed.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <type_traits>
#include <vector>

#include "device_common.h"
#include "private.h"
#include "workspace.h"

using namespace oskar_amd;

namespace {

constexpr int KSRC = 8;          /* sources per k-step */
constexpr int THREADS = 256;

template <int TP_, int TQ_, int OCC_, int THREADS_ = THREADS>
struct Tile
{
    static constexpr int TP = TP_;               /* P stations per tile */
    static constexpr int TQ = TQ_;               /* Q stations per tile */
    static constexpr int TW = 4 * TQ_;           /* W rows per tile */
    static constexpr int NT = THREADS_;          /* threads per workgroup */
    static constexpr int PG = NT / TP;           /* P builder groups */
    static constexpr int PS = KSRC / PG;         /* sources per P builder */
    static constexpr int QG = NT / TQ;           /* W builder groups */
    static constexpr int QS = KSRC / QG;         /* sources per W builder */
    static constexpr int P_ENTRIES = 8 * TP;     /* uint4 per stage, P */
    static constexpr int W_ENTRIES = 4 * TW;     /* uint4 per stage, W */
    static constexpr int STAGE = P_ENTRIES + W_ENTRIES;
    static constexpr int WR = NT / 128;          /* waves along P */
    static constexpr int WC = 2;                 /* waves along W */
    static constexpr int RF = TP / (32 * WR);    /* MFMA rows per wave */
    static constexpr int CF = TW / (32 * WC);    /* MFMA columns per wave */
    static constexpr int NMFMA = 6 * RF * CF;    /* per wave per k-step */
    static constexpr int OCC = OCC_;             /* workgroups per CU */
    static_assert(PS >= 1 && PS <= 4 && QS >= 1 && QS <= 4, "tile shape");
    static_assert(RF >= 1 && CF >= 1, "tile shape");
};

/* The tile shapes offered to the autotuner (index = shape code). */
/* 256 FP32 accumulators per lane fill the AGPRs: one workgroup per CU;
 * the 128-accumulator shapes fit two per CU when their other registers
 * allow it without spilling. */
using Tile0 = Tile<128, 64, 1>;
using Tile1 = Tile<64, 64, 2>;
using Tile2 = Tile<128, 32, 1>;
/* Shape 3 is the 64 x 64 geometry of Tile1 run by the wave-specialised
 * body (512 threads).  Shape 4 is the 128 x 64 tile with 512 threads: two
 * waves per SIMD (each 32 P x 128 W rows, 128 accumulators), so one wave's
 * build overlaps the other's MFMAs, with the smaller build work per MFMA of
 * the larger tile. */
using Tile4 = Tile<128, 64, 1, 512>;
constexpr int NUM_SHAPES = 5;
constexpr int SHAPE_TP[NUM_SHAPES] = {128, 64, 128, 64, 128};
constexpr int SHAPE_TQ[NUM_SHAPES] = {64, 64, 32, 64, 64};

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));
typedef float f16v __attribute__((ext_vector_type(16)));

struct Src { double l, m, n1; };        /* n1 = n - 1 */

/* Global-memory access through pointers the compiler cannot see are
 * global (those read from device-side descriptors): explicit address
 * space 1, so the loads are global_* rather than flat_* instructions (flat
 * loads also count against the LDS wait counter). */
#define GLOBAL_AS __attribute__((address_space(1)))
typedef float fv4 __attribute__((ext_vector_type(4)));

template <class T>
__device__ __forceinline__ T gld(const T* p)
{
    return *(const GLOBAL_AS T*) p;
}
template <class T>
__device__ __forceinline__ void gst(T* p, T v)
{
    *(GLOBAL_AS T*) p = v;
}
__device__ __forceinline__ Src gld_src(const Src* p)
{
    const GLOBAL_AS double* q = (const GLOBAL_AS double*) p;
    Src s;
    s.l = q[0];
    s.m = q[1];
    s.n1 = q[2];
    return s;
}
__device__ __forceinline__ void gst_src(Src* p, const Src& s)
{
    GLOBAL_AS double* q = (GLOBAL_AS double*) p;
    q[0] = s.l;
    q[1] = s.m;
    q[2] = s.n1;
}
__device__ __forceinline__ float4 gld_f4(const float4* p)
{
    const fv4 v = *(const GLOBAL_AS fv4*) p;
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void gst_f4(float4* p, float x, float y, float z,
        float w)
{
    fv4 v;
    v.x = x; v.y = y; v.z = z; v.w = w;
    *(GLOBAL_AS fv4*) p = v;
}

/* Splits a pair (x0, x1) into hi + lo parts of a 16-bit type (round to
 * nearest), so that hi + lo = x to the 16-bit type's precision of lo.
 * Both parts must come from the one rounded FP32 value x.  Where x is a
 * product, the compiler would otherwise round the exact product to 16
 * bits in a mixed-precision FMA for one use of hi and round the FP32 x for
 * another (it does so even with contraction switched off); where x sits on
 * a 16-bit rounding tie those differ by one unit and hi + lo misses x by
 * it (seen as a 2^-11 error on the brightest source of one station).  The
 * empty asm makes x an opaque FP32 register value.  The two hi parts come
 * from one packed conversion, and lo = x - hi is written as an FMA by an
 * opaque 1, so that each is one mixed-precision instruction reading its hi
 * by halves of the packed register (x - hi is exact in FP32). */
template <typename T16>
__device__ __forceinline__ void split16x2(float x0, float x1, T16* hi,
        T16* lo)
{
    typedef T16 t2 __attribute__((ext_vector_type(2)));
    typedef float f2 __attribute__((ext_vector_type(2)));
    float one = 1.0f;
    asm("" : "+v"(x0), "+v"(x1));
    asm("" : "+s"(one));
    const f2 x = {x0, x1};
    const t2 h = __builtin_convertvector(x, t2);
    hi[0] = h.x;
    hi[1] = h.y;
    lo[0] = (T16) __builtin_fmaf(x0, one, -(float) h.x);
    lo[1] = (T16) __builtin_fmaf(x1, one, -(float) h.y);
}

/* FP16: the same split in three instructions written out, so that no
 * other instruction selection can arise: the packed conversion of the
 * pair, then lo of each half as one mixed-precision FMA x * 1 - hi into
 * the low and the high half of one register. */
template <>
__device__ __forceinline__ void split16x2<_Float16>(float x0, float x1,
        _Float16* hi, _Float16* lo)
{
    unsigned int h, l;
    asm("v_cvt_pk_f16_f32 %0, %1, %2" : "=v"(h) : "v"(x0), "v"(x1));
    asm("v_fma_mixlo_f16 %0, %1, 1.0, -%2 op_sel_hi:[0,0,1]\n\t"
        "v_fma_mixhi_f16 %0, %3, 1.0, -%2 op_sel:[0,0,1] op_sel_hi:[0,0,1]"
        : "=&v"(l) : "v"(x0), "v"(h), "v"(x1));
    __builtin_memcpy(hi, &h, 4);
    __builtin_memcpy(lo, &l, 4);
}

/* split16x2 of (s x0, s x1) for a power of two s (the products are exact
 * in FP32). */
template <typename T16>
__device__ __forceinline__ void split16x2_scaled(float x0, float x1,
        float s, T16* hi, T16* lo)
{
    split16x2<T16>(x0 * s, x1 * s, hi, lo);
}

/* FP16: four mixed-precision FMAs, each rounding the exact product x * s
 * (or x * s - hi) once: the same parts as the split of the scaled FP32
 * values, without the two multiplications. */
template <>
__device__ __forceinline__ void split16x2_scaled<_Float16>(float x0,
        float x1, float s, _Float16* hi, _Float16* lo)
{
    unsigned int h, l;
    asm("v_fma_mixlo_f16 %0, %1, %2, 0\n\t"
        "v_fma_mixhi_f16 %0, %3, %2, 0"
        : "=&v"(h) : "v"(x0), "s"(s), "v"(x1));
    asm("v_fma_mixlo_f16 %0, %1, %2, -%3 op_sel_hi:[0,0,1]\n\t"
        "v_fma_mixhi_f16 %0, %4, %2, -%3 op_sel:[0,0,1] op_sel_hi:[0,0,1]"
        : "=&v"(l) : "v"(x0), "s"(s), "v"(h), "v"(x1));
    __builtin_memcpy(hi, &h, 4);
    __builtin_memcpy(lo, &l, 4);
}

/* y = (x1, -x0) of a pair x = (x0, x1) (exact). */
template <typename T16>
__device__ __forceinline__ void y_pair(const T16* x, T16* y)
{
    y[0] = x[1];
    y[1] = -x[0];
}

/* FP16: one packed multiply of the 32-bit word: the low result is the
 * high half times 1, the high result the low half times -1 (1.0 in both
 * halves of the second source, negated for the high result). */
template <>
__device__ __forceinline__ void y_pair<_Float16>(const _Float16* x,
        _Float16* y)
{
    unsigned int w, r;
    __builtin_memcpy(&w, x, 4);
    asm("v_pk_mul_f16 %0, %1, %2 op_sel:[1,0] op_sel_hi:[0,1] neg_hi:[0,1]"
        : "=v"(r) : "v"(w), "s"(0x3C003C00u));
    __builtin_memcpy(y, &r, 4);
}

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

/* Phase of station (U, V, W) for source (l, m, n - 1) in revolutions,
 * with an explicit fused multiply-add order so that every kernel that
 * forms it rounds it alike. */
__device__ __forceinline__ double phase_rev(double U, double V, double W,
        double l, double m, double n1)
{
    return __builtin_fma(W, n1, __builtin_fma(V, m, U * l));
}

/* exp(2 pi i r) for an FP64 phase in revolutions: its fractional part in
 * [0, 1), exact in FP64 (one v_fract_f64), then the hardware sine/cosine
 * (which take revolutions). */
__device__ __forceinline__ void phasor(double r, float* re, float* im)
{
    const float rr = (float) __builtin_amdgcn_fract(r);
    *re = __builtin_amdgcn_cosf(rr);
    *im = __builtin_amdgcn_sinf(rr);
}

/* Stores CNT 16-bit values (CNT = 2, 4 or 8) at element offset off (a
 * multiple of CNT) of the 16-byte LDS entry e. */
template <int CNT, typename T16>
__device__ __forceinline__ void put(uint4* e, int off, const T16* v)
{
    if constexpr (CNT == 8)
    {
        uint4 x;
        __builtin_memcpy(&x, v, 16);
        *e = x;
    }
    else if constexpr (CNT == 4)
    {
        uint2 x;
        __builtin_memcpy(&x, v, 8);
        reinterpret_cast<uint2*>(e)[off / 4] = x;
    }
    else
    {
        unsigned int x;
        __builtin_memcpy(&x, v, 4);
        reinterpret_cast<unsigned int*>(e)[off / 2] = x;
    }
}

/* Number of (P-block, Q-block) tiles holding at least one p > q pair, and
 * the mapping of a linear tile index to the block pair. */
__host__ __device__ inline int tiles_in_row(int ib, int N, int TP, int TQ)
{
    const int nq = (N + TQ - 1) / TQ;
    const int p_max = min(TP * ib + TP - 1, N - 1);
    if (p_max < 1) return 0;
    return min(nq, (p_max - 1) / TQ + 1);
}

__host__ __device__ inline int num_tiles(int N, int TP, int TQ)
{
    const int np = (N + TP - 1) / TP;
    int t = 0;
    for (int i = 0; i < np; ++i) t += tiles_in_row(i, N, TP, TQ);
    return t;
}

__host__ __device__ inline void tile_coords(int t, int N, int TP, int TQ,
        int* ib, int* jb)
{
    const int np = (N + TP - 1) / TP;
    int i = 0;
    for (; i < np; ++i)
    {
        const int c = tiles_in_row(i, N, TP, TQ);
        if (t < c) break;
        t -= c;
    }
    *ib = i;
    *jb = t;
}

__host__ __device__ inline int row_first_tile(int ib, int N, int TP, int TQ)
{
    int t = 0;
    for (int i = 0; i < ib; ++i) t += tiles_in_row(i, N, TP, TQ);
    return t;
}

/* Scale of the K operand: |K| <= 1 -> <= 2^12, so lo parts of components
 * down to 2^-15 stay normal FP16 numbers. */
constexpr float K_SCALE = 4096.0f;

/* Power-of-two scale that puts max |C| just below 2^14. */
__device__ __forceinline__ float c_scale(const unsigned int* cmax)
{
    const float m = __uint_as_float(gld(cmax));
    if (!(m > 0.0f) || !isfinite(m)) return 1.0f;
    int e = 0;
    (void) frexpf(m, &e);          /* m = f * 2^e, 0.5 <= f < 1 */
    return ldexpf(1.0f, 14 - e);
}

/* Source data one thread needs for one k-step: PS sources for its P row,
 * QS sources (with C) for its W rows.  Loaded two k-steps ahead so the
 * global-memory latency hides behind the MFMAs of the current step. */
template <int PS, int QS>
struct StepSrc
{
    Src p[PS];
    Src w[QS];
    float4 c[QS];
};

/* One (tile, source part) of the product: the FP32 sums over sources
 * [s_begin, s_end) for P stations p0.. and Q stations q0.., stored to out as
 * [re/im][TP][TW] (W rows ordered ik * TQ + q). */
struct TileJob
{
    int N, p0, q0, s_begin, s_end;
    int s_clamp;                 /* last valid source slot (padded count - 1) */
    const Src* src;              /* per source: l, m, n - 1 (FP64) */
    const float4* cmat;          /* per source: C00, Re C01, Im C01, C11 */
    const double* st_uvw;        /* per station: U, V, W (revolutions) */
    float scale;                 /* power-of-two scale of C */
    float* out;
};

/* sb is a multiple of the k-step and the padded source count is too, so
 * the slots of a step are either all valid or all past the end (a step
 * after the last, whose build is never read): one clamp of the step's
 * first slot to the last whole step keeps every load in bounds. */
template <int PS, int QS>
__device__ __forceinline__ void load_step(const TileJob& J, int sb,
        int s_last, int pg, int qg, StepSrc<PS, QS>* o)
{
    const int s0 = min(sb, s_last - (KSRC - 1));
#pragma unroll
    for (int j = 0; j < PS; ++j)
        o->p[j] = gld_src(J.src + s0 + pg * PS + j);
#pragma unroll
    for (int j = 0; j < QS; ++j)
    {
        const int s = s0 + qg * QS + j;
        o->w[j] = gld_src(J.src + s);
        o->c[j] = gld_f4(J.cmat + s);
    }
}

/* The operand builder of one thread (bt: its index among 256 builder
 * threads): the station constants of its P and W rows, and the build of one
 * k-step into a stage buffer.  Stations past N get a zero mask (their rows
 * build as zeros). */
template <typename T16, class TL>
struct Builder
{
    static constexpr int TP = TL::TP, TQ = TL::TQ, TW = TL::TW;
    static constexpr int PS = TL::PS, QS = TL::QS;
    int pb, pg, qb, qg, p_half, p_off, q_half, q_off, s_last;
    int ps, qs;
    double pu, pv, pw, qu, qv, qw;
    float q_scale;

    __device__ __forceinline__ Builder(const TileJob& J, int bt)
    {
        const int N = J.N;
        pb = bt % TP;
        pg = bt / TP;
        qb = bt % TQ;
        qg = bt / TQ;
        /* A group of at least 64 threads is whole waves: the group index
         * is wave-uniform, and so are the source addresses of a k-step,
         * which are then formed in scalar registers (14 fewer VALU
         * instructions per k-step at 64 x 64). */
        if constexpr (TP % 64 == 0) pg = __builtin_amdgcn_readfirstlane(pg);
        if constexpr (TQ % 64 == 0) qg = __builtin_amdgcn_readfirstlane(qg);
        ps = min(J.p0 + pb, N - 1);
        qs = min(J.q0 + qb, N - 1);
        pu = gld(J.st_uvw + 3 * ps);
        pv = gld(J.st_uvw + 3 * ps + 1);
        pw = gld(J.st_uvw + 3 * ps + 2);
        qu = gld(J.st_uvw + 3 * qs);
        qv = gld(J.st_uvw + 3 * qs + 1);
        qw = gld(J.st_uvw + 3 * qs + 2);
        q_scale = J.scale;
        /* Position of this thread's sources in the k-step: plane half and
         * 16-bit element offset within the 16-byte entry. */
        p_half = (pg * PS) / 4;
        p_off = ((pg * PS) % 4) * 2;
        q_half = (qg * QS) / 4;
        q_off = ((qg * QS) % 4) * 2;
        s_last = J.s_clamp;
    }

    __device__ __forceinline__ void load(const TileJob& J, int sb,
            StepSrc<PS, QS>* o) const
    {
        load_step(J, sb, s_last, pg, qg, o);
    }

    __device__ __forceinline__ void k_p(const StepSrc<PS, QS>& d, int j,
            float* re, float* im) const
    {
        phasor(phase_rev(pu, pv, pw, d.p[j].l, d.p[j].m, d.p[j].n1), re, im);
    }

    __device__ __forceinline__ void k_w(const StepSrc<PS, QS>& d, int j,
            float* re, float* im) const
    {
        phasor(phase_rev(qu, qv, qw, d.w[j].l, d.w[j].m, d.w[j].n1), re, im);
    }

    /* Builds the P and W operand planes of the k-step starting at source sb
     * from the prefetched source data, into one stage buffer.  No masks:
     * the source slots past the last source (up to a multiple of 8) hold
     * C = 0, so their products are exactly zero, and the rows of stations
     * past N (built from the last station's constants) are never read. */
    __device__ __forceinline__ void build(int sb,
            const StepSrc<PS, QS>& d, uint4* st) const
    {
        {
            T16 xh[2 * PS], xl[2 * PS], yh[2 * PS], yl[2 * PS];
#pragma unroll
            for (int j = 0; j < PS; ++j)
            {
                float re, im;
                k_p(d, j, &re, &im);
                split16x2_scaled<T16>(re, im, K_SCALE, &xh[2 * j],
                        &xl[2 * j]);
                /* Y = (Im, -Re): negation is exact in the split parts. */
                y_pair<T16>(&xh[2 * j], &yh[2 * j]);
                y_pair<T16>(&xl[2 * j], &yl[2 * j]);
            }
            put<2 * PS>(&st[(0 * 2 + p_half) * TP + pb], p_off, xh);
            put<2 * PS>(&st[(1 * 2 + p_half) * TP + pb], p_off, xl);
            put<2 * PS>(&st[(2 * 2 + p_half) * TP + pb], p_off, yh);
            put<2 * PS>(&st[(3 * 2 + p_half) * TP + pb], p_off, yl);
        }
        {
            /* W rows are ordered ik-major (row = ik * TQ + q), so the
             * threads of a wave write consecutive rows. */
            T16 wh[4][2 * QS], wl[4][2 * QS];
#pragma unroll
            for (int j = 0; j < QS; ++j)
            {
                float kr, ki;
                k_w(d, j, &kr, &ki);
                kr *= q_scale;
                ki *= q_scale;
                const float4 c = d.c[j];
                /* W_ik = K conj(C_ik): C00, C11 real; C10 = conj(C01). */
                float w[4][2];
                w[0][0] = kr * c.x;
                w[0][1] = ki * c.x;
                w[1][0] = __builtin_fmaf(kr, c.y, ki * c.z);
                w[1][1] = __builtin_fmaf(ki, c.y, -(kr * c.z));
                w[2][0] = __builtin_fmaf(kr, c.y, -(ki * c.z));
                w[2][1] = __builtin_fmaf(ki, c.y, kr * c.z);
                w[3][0] = kr * c.w;
                w[3][1] = ki * c.w;
#pragma unroll
                for (int ik = 0; ik < 4; ++ik)
                {
                    split16x2<T16>(w[ik][0], w[ik][1], &wh[ik][2 * j],
                            &wl[ik][2 * j]);
                }
            }
#pragma unroll
            for (int ik = 0; ik < 4; ++ik)
            {
                const int row = ik * TQ + qb;
                put<2 * QS>(&st[TL::P_ENTRIES + (0 * 2 + q_half) * TW + row],
                        q_off, wh[ik]);
                put<2 * QS>(&st[TL::P_ENTRIES + (1 * 2 + q_half) * TW + row],
                        q_off, wl[ik]);
            }
        }
    }
};

/* Stores a wave's FP32 accumulators of one part: element (row i, col j) of
 * a 32x32 block sits in lane (j + 32 * ((i / 4) % 2)), register
 * 4 * (i / 8) + i % 4.  Columns are W rows, ordered ik * TQ + q. */
template <int TP, int TW>
__device__ __forceinline__ void store_block(float* base, int row0, int col0,
        const f16v& re, const f16v& im, int tid)
{
    const int lane = tid & 63, lr = lane & 31, lh = lane >> 5;
#pragma unroll
    for (int v = 0; v < 16; ++v)
    {
        const int row = row0 + 8 * (v / 4) + 4 * lh + (v % 4);
        const int col = col0 + lr;
        gst(base + (size_t) row * TW + col, (float) re[v]);
        gst(base + (size_t) (TP * TW) + (size_t) row * TW + col,
                (float) im[v]);
    }
}

/* tid: the thread's index in the workgroup (a value the persistent batch
 * kernel makes opaque per work item, so that nothing derived from it is
 * hoisted out of its item loop and held in registers across it). */
template <typename T16, class TL, bool DIAG = false>
__device__ __forceinline__ void tile_part(const TileJob& J,
        uint4 (*lds)[TL::STAGE], int tid)
{
    constexpr int TP = TL::TP, TW = TL::TW;
    constexpr int PS = TL::PS, QS = TL::QS, RF = TL::RF, CF = TL::CF;
    constexpr int P_ENTRIES = TL::P_ENTRIES;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const Builder<T16, TL> bld(J, tid);
    const int s_begin = J.s_begin;
    const int nsteps = (J.s_end - s_begin + KSRC - 1) / KSRC;

    f16v re[RF][CF], im[RF][CF];
#pragma unroll
    for (int a = 0; a < RF; ++a)
#pragma unroll
        for (int b = 0; b < CF; ++b)
#pragma unroll
            for (int v = 0; v < 16; ++v) { re[a][b][v] = 0.0f; im[a][b][v] = 0.0f; }

    const int wp = wave % TL::WR, ww = wave / TL::WR;
    const int lr = lane & 31, lh = lane >> 5;
    /* On a diagonal tile (P and Q the same stations) a wave's 32 x 32 block
     * whose P stations all precede its Q stations holds no baseline: its
     * MFMAs are skipped (the wave-uniform test below; its partial sums stay
     * zero and are never read). */
    auto live = [&](int rf, int cf) {
        if constexpr (!DIAG) return true;
        else
        {
            const int p_last = wp * (TP / TL::WR) + rf * 32 + 31;
            const int q_first = (ww * (TW / TL::WC) + cf * 32) % TL::TQ;
            return p_last >= q_first;
        }
    };
    StepSrc<PS, QS> cur;
    if (nsteps > 0)
    {
        bld.load(J, s_begin, &cur);
        bld.build(s_begin, cur, lds[0]);
        bld.load(J, s_begin + KSRC, &cur);
    }
    __syncthreads();
    /* One k-step: multiply stage st, build step ks+1 into stage nxt.  The
     * loop below runs steps in pairs so that each stage's address is a
     * constant of its copy of the body. */
    auto kstep = [&](int ks, const uint4* st, uint4* nxt)
            __attribute__((always_inline))
    {
        /* Operand fragments of step ks first: the builder's LDS writes to
         * the other stage come after them in program order, so the MFMAs
         * need not wait for the build. */
        uint4 xh[RF], xl[RF], yh[RF], yl[RF], wh[CF], wl[CF];
#pragma unroll
        for (int rf = 0; rf < RF; ++rf)
        {
            const int row = wp * (TP / TL::WR) + rf * 32 + lr;
            xh[rf] = st[(0 * 2 + lh) * TP + row];
            xl[rf] = st[(1 * 2 + lh) * TP + row];
            yh[rf] = st[(2 * 2 + lh) * TP + row];
            yl[rf] = st[(3 * 2 + lh) * TP + row];
        }
#pragma unroll
        for (int cf = 0; cf < CF; ++cf)
        {
            const int row = ww * (TW / TL::WC) + cf * 32 + lr;
            wh[cf] = st[P_ENTRIES + (0 * 2 + lh) * TW + row];
            wl[cf] = st[P_ENTRIES + (1 * 2 + lh) * TW + row];
        }
        /* Two workgroups per CU: raise this wave's priority while its
         * MFMAs issue (see the schedule below). */
        if constexpr (TL::OCC >= 2) __builtin_amdgcn_s_setprio(2);
        /* hi*hi, then hi*lo, then lo*hi. */
#pragma unroll
        for (int rf = 0; rf < RF; ++rf)
#pragma unroll
            for (int cf = 0; cf < CF; ++cf)
            {
                if (!live(rf, cf)) continue;
                re[rf][cf] = mfma<T16>(xh[rf], wh[cf], re[rf][cf]);
                im[rf][cf] = mfma<T16>(yh[rf], wh[cf], im[rf][cf]);
            }
#pragma unroll
        for (int rf = 0; rf < RF; ++rf)
#pragma unroll
            for (int cf = 0; cf < CF; ++cf)
            {
                if (!live(rf, cf)) continue;
                re[rf][cf] = mfma<T16>(xh[rf], wl[cf], re[rf][cf]);
                im[rf][cf] = mfma<T16>(yh[rf], wl[cf], im[rf][cf]);
            }
#pragma unroll
        for (int rf = 0; rf < RF; ++rf)
#pragma unroll
            for (int cf = 0; cf < CF; ++cf)
            {
                if (!live(rf, cf)) continue;
                re[rf][cf] = mfma<T16>(xl[rf], wh[cf], re[rf][cf]);
                im[rf][cf] = mfma<T16>(yl[rf], wh[cf], im[rf][cf]);
            }
        if constexpr (TL::OCC >= 2) __builtin_amdgcn_s_setprio(0);
        /* Build step ks+1 (a masked dummy after the last step) into the
         * other stage, then start the loads for step ks+2. */
        bld.build(s_begin + (ks + 1) * KSRC, cur, nxt);
        bld.load(J, s_begin + (ks + 2) * KSRC, &cur);
        if constexpr (TL::OCC >= 2)
        {
            /* Two workgroups per CU (two waves per SIMD from independent
             * workgroups): operand reads, all MFMAs, then the whole build,
             * its LDS writes and the next loads.  The two workgroups drift
             * into opposite phases, so one's MFMAs run beside the other's
             * build (measured 6 % faster than interleaving at 64 x 64). */
            __builtin_amdgcn_sched_group_barrier(0x100, 4 * RF + 2 * CF, 0);
            __builtin_amdgcn_sched_group_barrier(0x008, TL::NMFMA, 0);
            __builtin_amdgcn_sched_group_barrier(0x002, 400, 0);
            __builtin_amdgcn_sched_group_barrier(0x200, 12, 0);
            __builtin_amdgcn_sched_group_barrier(0x020, 2 * PS + 3 * QS, 0);
        }
        else
        {
            /* One workgroup per CU: operand reads, then each MFMA followed
             * by a few VALU instructions of the build, then the builder's
             * LDS writes and the next loads. */
            __builtin_amdgcn_sched_group_barrier(0x100, 4 * RF + 2 * CF, 0);
#pragma unroll
            for (int i = 0; i < TL::NMFMA; ++i)
            {
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
                __builtin_amdgcn_sched_group_barrier(0x002,
                        192 / TL::NMFMA, 0);
            }
            __builtin_amdgcn_sched_group_barrier(0x200, 12, 0);
            __builtin_amdgcn_sched_group_barrier(0x020, 2 * PS + 3 * QS, 0);
        }
        __syncthreads();
    };
    int ks = 0;
    for (; ks + 2 <= nsteps; ks += 2)
    {
        kstep(ks, lds[0], lds[1]);
        kstep(ks + 1, lds[1], lds[0]);
    }
    if (ks < nsteps) kstep(ks, lds[0], lds[1]);
#pragma unroll
    for (int rf = 0; rf < RF; ++rf)
#pragma unroll
        for (int cf = 0; cf < CF; ++cf)
            store_block<TP, TW>(J.out, wp * (TP / TL::WR) + rf * 32,
                    ww * (TW / TL::WC) + cf * 32, re[rf][cf], im[rf][cf],
                    tid);
}

/* Wave-specialised tile body: the 64 x 64 tile of Tile1 with 512 threads.
 * Waves 0-3 (one per SIMD) only read operands and issue MFMAs; waves 4-7
 * (one per SIMD) only build operands.  Each SIMD interleaves the two
 * instruction streams, so the builder's VALU work runs while the matrix
 * pipe works, and a stall of one stream does not stop the other.  Three LDS
 * stages: during step k the MFMA waves multiply operands already in
 * registers and read step k+1's, while the builder waves write step k+2.
 * Both roles pass the same number of barriers (1 + steps). */
constexpr int WS_THREADS = 512;
constexpr int WS_STAGES = 3;

template <typename T16>
__device__ __forceinline__ void tile_part_ws(const TileJob& J,
        uint4 (*lds)[Tile1::STAGE], int tid)
{
    using TL = Tile1;
    constexpr int TP = TL::TP, TW = TL::TW, P_ENTRIES = TL::P_ENTRIES;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int s0 = J.s_begin;
    const int nsteps = (J.s_end - s0 + KSRC - 1) / KSRC;
    if (wave >= 4)
    {
        const Builder<T16, TL> bld(J, tid - 256);
        StepSrc<TL::PS, TL::QS> cur;
        if (nsteps > 0)
        {
            bld.load(J, s0, &cur);
            bld.build(s0, cur, lds[0]);
        }
        if (nsteps > 1)
        {
            bld.load(J, s0 + KSRC, &cur);
            bld.build(s0 + KSRC, cur, lds[1]);
        }
        bld.load(J, s0 + 2 * KSRC, &cur);
        __syncthreads();
        for (int ks = 0; ks < nsteps; ++ks)
        {
            if (ks + 2 < nsteps)
            {
                bld.build(s0 + (ks + 2) * KSRC, cur, lds[(ks + 2) % WS_STAGES]);
                bld.load(J, s0 + (ks + 3) * KSRC, &cur);
            }
            __syncthreads();
        }
        return;
    }
    const int lane = tid & 63, lr = lane & 31, lh = lane >> 5;
    const int prow = (wave & 1) * 32 + lr;
    const int wcol = (wave >> 1) * 128 + lr;
    f16v re[4], im[4];
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int v = 0; v < 16; ++v) { re[b][v] = 0.0f; im[b][v] = 0.0f; }
    /* Operand sets: [0..3] xh, xl, yh, yl; [4..7] wh; [8..11] wl. */
    uint4 opa[12], opb[12];
    auto read = [&](const uint4* st, uint4* o) {
        o[0] = st[(0 * 2 + lh) * TP + prow];
        o[1] = st[(1 * 2 + lh) * TP + prow];
        o[2] = st[(2 * 2 + lh) * TP + prow];
        o[3] = st[(3 * 2 + lh) * TP + prow];
#pragma unroll
        for (int cf = 0; cf < 4; ++cf)
        {
            o[4 + cf] = st[P_ENTRIES + (0 * 2 + lh) * TW + wcol + cf * 32];
            o[8 + cf] = st[P_ENTRIES + (1 * 2 + lh) * TW + wcol + cf * 32];
        }
    };
    /* hi*hi, hi*lo, lo*hi: 8 independent MFMAs between successive updates
     * of one accumulator. */
    auto multiply = [&](const uint4* o) {
#pragma unroll
        for (int cf = 0; cf < 4; ++cf)
        {
            re[cf] = mfma<T16>(o[0], o[4 + cf], re[cf]);
            im[cf] = mfma<T16>(o[2], o[4 + cf], im[cf]);
        }
#pragma unroll
        for (int cf = 0; cf < 4; ++cf)
        {
            re[cf] = mfma<T16>(o[0], o[8 + cf], re[cf]);
            im[cf] = mfma<T16>(o[2], o[8 + cf], im[cf]);
        }
#pragma unroll
        for (int cf = 0; cf < 4; ++cf)
        {
            re[cf] = mfma<T16>(o[1], o[4 + cf], re[cf]);
            im[cf] = mfma<T16>(o[3], o[4 + cf], im[cf]);
        }
    };
    /* The next step's reads are issued before this step's MFMAs (the
     * scheduling barriers keep them there), so they complete under the
     * matrix work rather than at the next barrier. */
    __syncthreads();
    if (nsteps > 0) read(lds[0], opa);
    /* lgkmcnt(0): the first operands are in registers before the loop, so
     * the loop body waits only for what it reads itself. */
    __builtin_amdgcn_s_waitcnt(0xC07F);
    for (int ks = 0; ks < nsteps; ks += 2)
    {
        if (ks + 1 < nsteps) read(lds[(ks + 1) % WS_STAGES], opb);
        __builtin_amdgcn_sched_barrier(0);
        multiply(opa);
        __builtin_amdgcn_sched_barrier(0);
        __syncthreads();
        if (ks + 1 < nsteps)
        {
            if (ks + 2 < nsteps) read(lds[(ks + 2) % WS_STAGES], opa);
            __builtin_amdgcn_sched_barrier(0);
            multiply(opb);
            __builtin_amdgcn_sched_barrier(0);
            __syncthreads();
        }
    }
    const int row0 = (wave & 1) * 32, col0 = (wave >> 1) * 128;
#pragma unroll
    for (int cf = 0; cf < 4; ++cf)
        store_block<TP, TW>(J.out, row0, col0 + cf * 32, re[cf], im[cf],
                tid);
}

/* ---------------------------------------------------------------------- */
/* Per-source and per-station preparation (shared by both launch forms).  */

/* Writes src[s] and cmat[s] = C = E B E^H (packed as C00, Re C01, Im C01,
 * C11) for one source; returns max |C entry|. */
template <typename FP>
__device__ __forceinline__ float prepare_source(int s, const FP* I,
        const FP* Q, const FP* U, const FP* V, const FP* l, const FP* m,
        const FP* n, const Cplx<FP>* beam, Src* src, float4* cmat)
{
    Src o;
    o.l = (double) gld(l + s);
    o.m = (double) gld(m + s);
    o.n1 = (double) gld(n + s) - 1.0;
    gst_src(src + s, o);
    /* B = [[I+Q, U+iV], [U-iV, I-Q]] */
    const double Ii = gld(I + s), Qi = Q ? (double) gld(Q + s) : 0.0;
    const double Ui = U ? (double) gld(U + s) : 0.0;
    const double Vi = V ? (double) gld(V + s) : 0.0;
    double c00, c01r, c01i, c11;
    if (beam)
    {
        const FP* e = (const FP*) (beam + 4 * (size_t) s);
        const double ar = gld(e), ai = gld(e + 1), br = gld(e + 2),
                bi = gld(e + 3);
        const double cr = gld(e + 4), ci = gld(e + 5), dr = gld(e + 6),
                di = gld(e + 7);
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
    gst_f4(cmat + s, (float) c00, (float) c01r, (float) c01i, (float) c11);
    return fmaxf(fmaxf(fabsf((float) c00), fabsf((float) c11)),
            fmaxf(fabsf((float) c01r), fabsf((float) c01i)));
}

/* Source slots are padded to a multiple of the k-step with C = 0. */
__host__ __device__ inline int pad8(int n) { return (n + KSRC - 1) / KSRC * KSRC; }

__device__ __forceinline__ void zero_source(int s, Src* src, float4* cmat)
{
    Src o;
    o.l = o.m = o.n1 = 0.0;
    gst_src(src + s, o);
    gst_f4(cmat + s, 0.0f, 0.0f, 0.0f, 0.0f);
}

/* Workgroup maximum of non-negative values, then one atomic per workgroup
 * (order-independent: atomic max on the float bit pattern). */
__device__ __forceinline__ void block_max_to(float mx, unsigned int* cmax,
        float* red)
{
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1)
    {
        if ((int) threadIdx.x < o)
            red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0 && red[0] > 0.0f)
        (void) __hip_atomic_fetch_max((GLOBAL_AS unsigned int*) cmax,
                __float_as_uint(red[0]), __ATOMIC_RELAXED,
                __HIP_MEMORY_SCOPE_AGENT);
}

/* Sum over parts in FP64 of one station row p (q from thread, stepping by
 * the workgroup size), scale removal, and accumulation into
 * vis[offset_out + b] for baselines inside the uv filter.  Consecutive
 * threads read consecutive partials. */
template <typename FP, class TL>
__device__ __forceinline__ void reduce_row(int N, int p, int tiles,
        int parts, int use_casa, const float* partial,
        const unsigned int* cmax, const double* st_uvw, double uv_min,
        double uv_max, int offset_out, Cplx<FP>* vis)
{
    constexpr int TP = TL::TP, TQ = TL::TQ, TW = TL::TW;
    const int ib = p / TP, pl = p - ib * TP;
    const int t0 = row_first_tile(ib, N, TP, TQ);
    const double inv = 1.0 / ((double) c_scale(cmax) * (double) K_SCALE);
    for (int q = threadIdx.x; q < p; q += blockDim.x)
    {
        /* uv filter on the baseline length in wavelengths. */
        const double du = gld(st_uvw + 3 * p) - gld(st_uvw + 3 * q);
        const double dv = gld(st_uvw + 3 * p + 1) - gld(st_uvw + 3 * q + 1);
        const double len = sqrt(du * du + dv * dv);
        if (len < uv_min || len > uv_max) continue;
        const int jb = q / TQ, ql = q - jb * TQ;
        const int t = t0 + jb;
        double vr[4] = {0, 0, 0, 0}, vi[4] = {0, 0, 0, 0};
        for (int k = 0; k < parts; ++k)
        {
            const float* b = partial + ((size_t) k * tiles + t) *
                    (size_t) (2 * TP * TW) + (size_t) pl * TW + ql;
#pragma unroll
            for (int ik = 0; ik < 4; ++ik)
            {
                vr[ik] += (double) gld(b + ik * TQ);
                vi[ik] += (double) gld(b + TP * TW + ik * TQ);
            }
        }
        const int bl = q * (N - 1) - (q - 1) * q / 2 + p - q - 1;
        FP* o = (FP*) (vis + 4 * (size_t) (offset_out + bl));
#pragma unroll
        for (int ik = 0; ik < 4; ++ik)
        {
            /* CASA: V = (J_p B J_q^H)^H -> element [i][k] = conj(V[k][i]). */
            const int sidx = use_casa ? ((ik & 1) * 2 + (ik >> 1)) : ik;
            const double r = vr[sidx] * inv;
            const double i = (use_casa ? -vi[sidx] : vi[sidx]) * inv;
            gst(o + 2 * ik, (FP) (gld(o + 2 * ik) + (FP) r));
            gst(o + 2 * ik + 1, (FP) (gld(o + 2 * ik + 1) + (FP) i));
        }
    }
}

/* ---------------------------------------------------------------------- */
/* One call: one (channel, time) unit.                                     */

struct FactArgs
{
    int num_sources, num_stations, num_tiles, src_per_part, num_parts;
    const Src* src;
    const float4* cmat;
    const double* st_uvw;
    const unsigned int* cmax;    /* bit pattern of max |C| (>= 0 float) */
    float* partial;              /* [part][tile][re/im][TP][TW] */
};

/* Tile bodies: geometry (TL), workgroup size, occupancy and LDS of the
 * one-role body (tile_part, 256 threads) and the wave-specialised body
 * (tile_part_ws, 512 threads, 64 x 64 geometry). */
template <typename T16, class TL_>
struct BodyOneRole
{
    using TL = TL_;
    static constexpr int THREADS_ = TL::NT, OCC = TL::OCC, STAGES = 2;
    static __device__ __forceinline__ void run(const TileJob& J,
            uint4 (*lds)[TL::STAGE], int tid)
    {
        if (TL::TP == TL::TQ && J.p0 == J.q0)
            tile_part<T16, TL, true>(J, lds, tid);
        else
            tile_part<T16, TL, false>(J, lds, tid);
    }
};

template <typename T16>
struct BodyWS
{
    using TL = Tile1;
    static constexpr int THREADS_ = WS_THREADS, OCC = 1, STAGES = WS_STAGES;
    static __device__ __forceinline__ void run(const TileJob& J,
            uint4 (*lds)[TL::STAGE], int tid)
    {
        tile_part_ws<T16>(J, lds, tid);
    }
};

template <class Body>
__global__ __launch_bounds__(Body::THREADS_, Body::OCC)
void fact_xcorr_kernel(FactArgs A)
{
    using TL = typename Body::TL;
    __shared__ uint4 lds[Body::STAGES][TL::STAGE];
    const int tile = blockIdx.x, part = blockIdx.y;
    int ib, jb;
    tile_coords(tile, A.num_stations, TL::TP, TL::TQ, &ib, &jb);
    TileJob J;
    J.N = A.num_stations;
    J.p0 = ib * TL::TP;
    J.q0 = jb * TL::TQ;
    J.s_begin = part * A.src_per_part;
    J.s_end = min(A.num_sources, J.s_begin + A.src_per_part);
    J.s_clamp = pad8(A.num_sources) - 1;
    J.src = A.src;
    J.cmat = A.cmat;
    J.st_uvw = A.st_uvw;
    J.scale = c_scale(A.cmax);
    J.out = A.partial + ((size_t) part * A.num_tiles + tile) *
            (size_t) (2 * TL::TP * TL::TW);
    Body::run(J, lds, threadIdx.x);
}

template <typename FP>
__global__ __launch_bounds__(256)
void fact_prepare_kernel(int num_sources, const FP* I, const FP* Q,
        const FP* U, const FP* V, const FP* l, const FP* m, const FP* n,
        const Cplx<FP>* beam, Src* src, float4* cmat, unsigned int* cmax)
{
    __shared__ float red[256];
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    float mx = 0.0f;
    if (s < num_sources)
        mx = prepare_source<FP>(s, I, Q, U, V, l, m, n, beam, src, cmat);
    else if (s < pad8(num_sources))
        zero_source(s, src, cmat);
    block_max_to(mx, cmax, red);
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

template <typename FP, class TL>
__global__ __launch_bounds__(256)
void fact_reduce_kernel(int N, int num_tiles, int num_parts, int use_casa,
        const float* partial, const unsigned int* cmax, const double* st_uvw,
        double uv_min, double uv_max, int offset_out, Cplx<FP>* vis)
{
    reduce_row<FP, TL>(N, blockIdx.x, num_tiles, num_parts, use_casa,
            partial, cmax, st_uvw, uv_min, uv_max, offset_out, vis);
}

/* ---------------------------------------------------------------------- */
/* Batches: many units, one persistent launch, descriptors in HBM.        */

struct BatchHeader
{
    int num_units, num_stations, num_tiles, src_per_part;
    int total_items, prep_total, use_casa, pad0;
};

struct UnitDesc
{
    const void *I, *Q, *U, *V, *l, *m, *n, *beam, *st_u, *st_v, *st_w;
    void* vis;
    Src* src;
    float4* cmat;
    double* st;
    unsigned int* cmax;
    float* partial;
    double f_over_c;
    double uv_min, uv_max;       /* uv filter (wavelengths) */
    int num_sources, offset_out;
    int item_start;              /* first work item (prefix over units) */
    int num_parts;
    int prep_start;              /* first prepare slot (256-aligned) */
    int pad0;
};

/* x, through an empty asm the compiler must assume changes it: values
 * derived from the result are computed where they are used. */
__device__ __forceinline__ int opaque(int x)
{
    asm volatile("" : "+v"(x));
    return x;
}

/* Last unit u with key(u) <= x (keys non-decreasing; key(0) = 0 <= x).
 * Uniform across the workgroup: the loads are scalar. */
template <class F>
__device__ __forceinline__ int find_unit(int nu, int x, F key)
{
    int lo = 0, hi = nu - 1;
    while (lo < hi)
    {
        const int mid = (lo + hi + 1) >> 1;
        if (key(mid) <= x) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

/* Station constants of every unit; resets each unit's |C| maximum and the
 * work-item counter (before the product kernel, in stream order). */
template <typename FP>
__global__ __launch_bounds__(256)
void fact_batch_stations_kernel(const BatchHeader* H, const UnitDesc* D,
        int* counter)
{
    const int N = H->num_stations, total = H->num_units * N;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < total;
            i += gridDim.x * blockDim.x)
    {
        const int u = i / N, k = i - u * N;
        const UnitDesc& d = D[u];
        const double f = d.f_over_c;
        gst(d.st + 3 * k, (double) gld((const FP*) d.st_u + k) * f);
        gst(d.st + 3 * k + 1, (double) gld((const FP*) d.st_v + k) * f);
        gst(d.st + 3 * k + 2, (double) gld((const FP*) d.st_w + k) * f);
        if (k == 0) gst(d.cmax, 0u);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *counter = 0;
}

/* Sources of every unit: slots of 256 per workgroup; a unit's slots start
 * 256-aligned, so one workgroup never straddles two units. */
template <typename FP>
__global__ __launch_bounds__(256)
void fact_batch_prepare_kernel(const BatchHeader* H, const UnitDesc* D)
{
    __shared__ float red[256];
    const int nu = H->num_units, total = H->prep_total;
    for (int c = blockIdx.x * 256; c < total; c += gridDim.x * 256)
    {
        const int u = find_unit(nu, c,
                [&](int k) { return D[k].prep_start; });
        const UnitDesc& d = D[u];
        const int s = c - d.prep_start + (int) threadIdx.x;
        float mx = 0.0f;
        if (s < d.num_sources)
            mx = prepare_source<FP>(s, (const FP*) d.I, (const FP*) d.Q,
                    (const FP*) d.U, (const FP*) d.V, (const FP*) d.l,
                    (const FP*) d.m, (const FP*) d.n,
                    (const Cplx<FP>*) d.beam, d.src, d.cmat);
        else if (s < pad8(d.num_sources))
            zero_source(s, d.src, d.cmat);
        block_max_to(mx, d.cmax, red);
        __syncthreads();
    }
}

/* Persistent product: each workgroup takes (unit, part, tile) items from
 * the device counter until all are done (every workgroup exits when the
 * counter passes total_items).  Items of one part are consecutive, so
 * concurrently running workgroups share source data in L2. */
template <class Body>
__global__ __launch_bounds__(Body::THREADS_, Body::OCC)
void fact_batch_kernel(const BatchHeader* H, const UnitDesc* D, int* counter)
{
    using TL = typename Body::TL;
    __shared__ uint4 lds[Body::STAGES][TL::STAGE];
    __shared__ int s_item;
    const int total = H->total_items, nu = H->num_units;
    const int N = H->num_stations, tiles = H->num_tiles;
    const int spp = H->src_per_part;
    for (;;)
    {
        __syncthreads();          /* every thread has read s_item */
        if (threadIdx.x == 0) s_item = atomicAdd(counter, 1);
        __syncthreads();
        /* Uniform over the workgroup: scalar descriptor loads. */
        const int item = __builtin_amdgcn_readfirstlane(s_item);
        if (item >= total) break;
        const int u = find_unit(nu, item,
                [&](int k) { return D[k].item_start; });
        const UnitDesc& d = D[u];
        const int local = item - d.item_start;
        const int part = local / tiles, tile = local - part * tiles;
        int ib, jb;
        tile_coords(tile, N, TL::TP, TL::TQ, &ib, &jb);
        TileJob J;
        J.N = N;
        J.p0 = ib * TL::TP;
        J.q0 = jb * TL::TQ;
        J.s_begin = part * spp;
        J.s_end = min(d.num_sources, J.s_begin + spp);
        J.s_clamp = pad8(d.num_sources) - 1;
        J.src = d.src;
        J.cmat = d.cmat;
        J.st_uvw = d.st;
        J.scale = c_scale(d.cmax);
        J.out = d.partial + ((size_t) part * tiles + tile) *
                (size_t) (2 * TL::TP * TL::TW);
        Body::run(J, lds, opaque(threadIdx.x));
    }
}

template <typename FP, class TL>
__global__ __launch_bounds__(256)
void fact_batch_reduce_kernel(const BatchHeader* H, const UnitDesc* D)
{
    const int N = H->num_stations, rows = H->num_units * N;
    for (int r = blockIdx.x; r < rows; r += gridDim.x)
    {
        const int u = r / N, p = r - u * N;
        const UnitDesc& d = D[u];
        if (d.num_parts == 0) continue;      /* no sources: untouched */
        reduce_row<FP, TL>(N, p, H->num_tiles, d.num_parts, H->use_casa,
                d.partial, d.cmax, d.st, d.uv_min, d.uv_max, d.offset_out,
                (Cplx<FP>*) d.vis);
    }
}

/* ---------------------------------------------------------------------- */
/* Host side.                                                             */

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

inline size_t align256(size_t x) { return (x + 255) / 256 * 256; }


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

/* The body of each tile shape, for each 16-bit type. */
template <int SHAPE, typename T16> struct ShapeBody;
template <typename T16> struct ShapeBody<0, T16>
{ using type = BodyOneRole<T16, Tile0>; };
template <typename T16> struct ShapeBody<1, T16>
{ using type = BodyOneRole<T16, Tile1>; };
template <typename T16> struct ShapeBody<2, T16>
{ using type = BodyOneRole<T16, Tile2>; };
template <typename T16> struct ShapeBody<3, T16>
{ using type = BodyWS<T16>; };
template <typename T16> struct ShapeBody<4, T16>
{ using type = BodyOneRole<T16, Tile4>; };

int occupancy(int shape)
{
    switch (shape)
    {
    case 1: return Tile1::OCC;
    case 2: return Tile2::OCC;
    case 3: return BodyWS<_Float16>::OCC;
    case 4: return Tile4::OCC;
    default: return Tile0::OCC;
    }
}

} /* namespace */

/* Number of tiles for N stations with the default tile shape (host). */
extern "C" int oskar_amd_fact_num_tiles(int N)
{
    return num_tiles(N, Tile0::TP, Tile0::TQ);
}

namespace {

template <int SHAPE>
int fact_launch_shape(int dev, int precision, int mode, int use_casa,
        int S, int N, const void* beam, const void* I, const void* Q,
        const void* U, const void* V, const void* l, const void* m,
        const void* n, const void* st_u, const void* st_v, const void* st_w,
        double frequency_hz, double uv_min, double uv_max, int offset_out,
        void* vis, int src_per_part, hipStream_t stream)
{
    using BodyH = typename ShapeBody<SHAPE, _Float16>::type;
    using BodyB = typename ShapeBody<SHAPE, __bf16>::type;
    using TL = typename BodyH::TL;
    hipError_t err = hipSuccess;
    const int parts = (S + src_per_part - 1) / src_per_part;
    const int tiles = num_tiles(N, TL::TP, TL::TQ);

    /* Workspace 0: [cmax (256 B)][stations 3N doubles][src S x 24 B]
     * [cmat S x 16 B]; workspace 1: partials. */
    const size_t off_st = 256;
    const size_t off_src = off_st + align256(3 * (size_t) N * 8);
    const int S8 = pad8(S);
    const size_t off_c = off_src + align256(sizeof(Src) * (size_t) S8);
    const size_t bytes0 = off_c + align256(16 * (size_t) S8);
    char* w0 = (char*) ws_acquire(WS_FACT_SOURCES, dev, stream, bytes0,
            &err);
    if (err != hipSuccess) return (int) err;
    const size_t bytes1 = (size_t) parts * tiles * 2 * TL::TP * TL::TW *
            sizeof(float);
    float* partial = (float*) ws_acquire(WS_FACT_PARTIALS, dev, stream,
            bytes1, &err);
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
                dim3((S8 + 255) / 256), dim3(256), 0, stream, S,
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
                dim3((S8 + 255) / 256), dim3(256), 0, stream, S,
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
    A.num_parts = parts;
    A.src = src;
    A.cmat = cmat;
    A.st_uvw = st;
    A.cmax = cmax;
    A.partial = partial;
    const dim3 grid(tiles, parts);
    mark(dev, 1, stream);
    if (mode == 1)
        hipLaunchKernelGGL((fact_xcorr_kernel<BodyB>), grid,
                dim3(BodyB::THREADS_), 0, stream, A);
    else
        hipLaunchKernelGGL((fact_xcorr_kernel<BodyH>), grid,
                dim3(BodyH::THREADS_), 0, stream, A);
    err = hipGetLastError();
    if (err != hipSuccess) return (int) err;

    mark(dev, 2, stream);
    if (precision == OSKAR_DOUBLE)
        hipLaunchKernelGGL((fact_reduce_kernel<double, TL>), dim3(N),
                dim3(256), 0, stream, N, tiles, parts, use_casa, partial,
                cmax, st, uv_min, uv_max, offset_out, (Cplx<double>*) vis);
    else
        hipLaunchKernelGGL((fact_reduce_kernel<float, TL>), dim3(N),
                dim3(256), 0, stream, N, tiles, parts, use_casa, partial,
                cmax, st, uv_min, uv_max, offset_out, (Cplx<float>*) vis);
    mark(dev, 3, stream);
    return (int) hipGetLastError();
}


/* Tile shape of the one-call path: 64 x 64 with two workgroups per CU
 * measured fastest at 512 stations x 1M sources (14.3 ms against 15.7 ms
 * for 128 x 64 and 16.9 ms for the wave-specialised body). */
int g_shape = 1;

} /* namespace */

/* Tile shape of oskar_amd_fact_launch (0..3); returns the previous one.  A
 * negative shape only returns the current one; a shape past the last
 * returns -1 and changes nothing. */
extern "C" int oskar_amd_fact_set_tile_shape(int shape)
{
    if (shape < 0) return g_shape;
    if (shape >= NUM_SHAPES) return -1;
    const int old = g_shape;
    g_shape = shape;
    return old;
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
#define FACT_ARGS dev, precision, mode, use_casa, S, N, beam, I, Q, U, V, \
        l, m, n, st_u, st_v, st_w, frequency_hz, uv_min_lambda, \
        uv_max_lambda, offset_out, vis, src_per_part, stream
    switch (g_shape)
    {
    case 1: return fact_launch_shape<1>(FACT_ARGS);
    case 2: return fact_launch_shape<2>(FACT_ARGS);
    case 3: return fact_launch_shape<3>(FACT_ARGS);
    case 4: return fact_launch_shape<4>(FACT_ARGS);
    default: return fact_launch_shape<0>(FACT_ARGS);
    }
#undef FACT_ARGS
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

/* ---------------------------------------------------------------------- */
/* Batch object.                                                          */

namespace {

struct BatchConfig
{
    int shape = 1;               /* index into SHAPE_TP / SHAPE_TQ */
    int src_per_part = 8192;
    int workgroups = 0;          /* persistent grid; 0 = CUs x occupancy */
    bool operator==(const BatchConfig& o) const
    {
        return shape == o.shape && src_per_part == o.src_per_part &&
                workgroups == o.workgroups;
    }
};

struct Batch
{
    int precision = 0, mode = 0, N = 0, max_units = 0;
    int dev = 0, num_cus = 0, use_graph = 1;
    BatchConfig cfg;
    std::vector<UnitDesc> units;         /* host copy, inputs only */
    std::vector<char> unit_set;
    BatchHeader* h_hdr = nullptr;        /* pinned staging */
    UnitDesc* h_desc = nullptr;
    BatchHeader* d_hdr = nullptr;        /* device */
    UnitDesc* d_desc = nullptr;
    int* d_counter = nullptr;
    Workspace ws;
    hipEvent_t ev_upload = nullptr, ev_start = nullptr, ev_end = nullptr;
    bool upload_pending = false, timed = false;
    hipStream_t cap_stream = nullptr;
    hipGraphExec_t exec = nullptr;
    BatchConfig exec_cfg;
};

template <typename FP, int SHAPE>
hipError_t batch_launch_t(Batch* b, hipStream_t s)
{
    using BodyH = typename ShapeBody<SHAPE, _Float16>::type;
    using BodyB = typename ShapeBody<SHAPE, __bf16>::type;
    using TL = typename BodyH::TL;
    const int wgs = b->cfg.workgroups > 0 ? b->cfg.workgroups :
            b->num_cus * BodyH::OCC;
    const int rows_grid = std::min(4096, b->max_units * b->N);
    hipLaunchKernelGGL(fact_batch_stations_kernel<FP>, dim3(1024),
            dim3(256), 0, s, b->d_hdr, b->d_desc, b->d_counter);
    hipLaunchKernelGGL(fact_batch_prepare_kernel<FP>, dim3(4096), dim3(256),
            0, s, b->d_hdr, b->d_desc);
    if (b->mode == 1)
        hipLaunchKernelGGL((fact_batch_kernel<BodyB>), dim3(wgs),
                dim3(BodyB::THREADS_), 0, s, b->d_hdr, b->d_desc,
                b->d_counter);
    else
        hipLaunchKernelGGL((fact_batch_kernel<BodyH>), dim3(wgs),
                dim3(BodyH::THREADS_), 0, s, b->d_hdr, b->d_desc,
                b->d_counter);
    hipLaunchKernelGGL((fact_batch_reduce_kernel<FP, TL>), dim3(rows_grid),
            dim3(256), 0, s, b->d_hdr, b->d_desc);
    return hipGetLastError();
}

template <typename FP>
hipError_t batch_launch_fp(Batch* b, hipStream_t s)
{
    switch (b->cfg.shape)
    {
    case 1: return batch_launch_t<FP, 1>(b, s);
    case 2: return batch_launch_t<FP, 2>(b, s);
    case 3: return batch_launch_t<FP, 3>(b, s);
    case 4: return batch_launch_t<FP, 4>(b, s);
    default: return batch_launch_t<FP, 0>(b, s);
    }
}

hipError_t batch_launch(Batch* b, hipStream_t s)
{
    return b->precision == OSKAR_DOUBLE ? batch_launch_fp<double>(b, s) :
            batch_launch_fp<float>(b, s);
}

/* Fills the descriptors of units [0, nu), sizes the workspace and uploads
 * header and descriptors on stream s.  vis_override (if not NULL) replaces
 * every unit's output with consecutive blocks of num_baselines values. */
hipError_t batch_upload(Batch* b, int nu, int use_casa,
        const double* uv_min, const double* uv_max, void* vis_override,
        hipStream_t s)
{
    const int N = b->N, TP = SHAPE_TP[b->cfg.shape],
            TQ = SHAPE_TQ[b->cfg.shape];
    const int tiles = num_tiles(N, TP, TQ);
    const int spp = b->cfg.src_per_part;
    const size_t tile_bytes = 2 * (size_t) TP * 4 * TQ * sizeof(float);
    const size_t nb = (size_t) N * (N - 1) / 2;
    const size_t vis_elem = (b->precision == OSKAR_DOUBLE ? 64 : 32);

    /* Workspace layout per unit: [cmax 256 B][stations][src][cmat][partial]. */
    size_t bytes = 0;
    long long items = 0, prep = 0;
    std::vector<size_t> off(nu);
    for (int u = 0; u < nu; ++u)
    {
        const int S = b->units[u].num_sources;
        const int parts = (S + spp - 1) / spp;
        off[u] = bytes;
        bytes += 256 + align256(3 * (size_t) N * 8) +
                align256(sizeof(Src) * (size_t) pad8(S)) +
                align256(16 * (size_t) pad8(S))
                + (size_t) parts * tiles * tile_bytes;
        items += (long long) parts * tiles;
        prep += (S + 255) / 256 * 256;
    }
    if (items > INT32_MAX || prep > INT32_MAX) return hipErrorInvalidValue;
    hipError_t err = hipSuccess;
    if (b->ws.size < bytes)
    {
        /* The previous batch may still read the old workspace. */
        err = hipDeviceSynchronize();
        if (err != hipSuccess) return err;
    }
    char* base = (char*) ws_get(&b->ws, bytes, &err);
    if (err != hipSuccess) return err;

    /* The previous upload must have left the staging buffers. */
    if (b->upload_pending)
    {
        err = hipEventSynchronize(b->ev_upload);
        if (err != hipSuccess) return err;
        b->upload_pending = false;
    }
    BatchHeader* h = b->h_hdr;
    memset(h, 0, sizeof(*h));
    h->num_units = nu;
    h->num_stations = N;
    h->num_tiles = tiles;
    h->src_per_part = spp;
    h->total_items = (int) items;
    h->prep_total = (int) prep;
    h->use_casa = use_casa;
    int item_start = 0, prep_start = 0;
    for (int u = 0; u < nu; ++u)
    {
        UnitDesc d = b->units[u];
        const int S = d.num_sources;
        const int parts = (S + spp - 1) / spp;
        char* w = base + off[u];
        d.cmax = (unsigned int*) w;
        w += 256;
        d.st = (double*) w;
        w += align256(3 * (size_t) N * 8);
        d.src = (Src*) w;
        w += align256(sizeof(Src) * (size_t) pad8(S));
        d.cmat = (float4*) w;
        w += align256(16 * (size_t) pad8(S));
        d.partial = (float*) w;
        d.num_parts = parts;
        d.item_start = item_start;
        d.prep_start = prep_start;
        item_start += parts * tiles;
        prep_start += (S + 255) / 256 * 256;
        d.uv_min = uv_min ? uv_min[u] : 0.0;
        d.uv_max = uv_max ? uv_max[u] : 1e300;
        if (vis_override)
        {
            d.vis = (char*) vis_override + (size_t) u * nb * vis_elem;
            d.offset_out = 0;
        }
        b->h_desc[u] = d;
    }
    err = hipMemcpyAsync(b->d_hdr, h, sizeof(BatchHeader),
            hipMemcpyHostToDevice, s);
    if (err == hipSuccess)
        err = hipMemcpyAsync(b->d_desc, b->h_desc, nu * sizeof(UnitDesc),
                hipMemcpyHostToDevice, s);
    if (err == hipSuccess) err = hipEventRecord(b->ev_upload, s);
    if (err == hipSuccess) b->upload_pending = true;
    return err;
}

/* Runs the four kernels of a batch on stream s: replays the captured graph
 * of the current configuration (capturing it first if needed), or launches
 * directly when graphs are off. */
hipError_t batch_execute(Batch* b, hipStream_t s)
{
    hipError_t err;
    if (!b->use_graph) return batch_launch(b, s);
    if (!b->exec || !(b->exec_cfg == b->cfg))
    {
        if (b->exec)
        {
            err = hipStreamSynchronize(s);
            if (err != hipSuccess) return err;
            (void) hipGraphExecDestroy(b->exec);
            b->exec = nullptr;
        }
        hipGraph_t g = nullptr;
        err = hipStreamBeginCapture(b->cap_stream,
                hipStreamCaptureModeThreadLocal);
        if (err != hipSuccess) return err;
        const hipError_t le = batch_launch(b, b->cap_stream);
        err = hipStreamEndCapture(b->cap_stream, &g);
        if (le != hipSuccess) { if (g) (void) hipGraphDestroy(g); return le; }
        if (err != hipSuccess) return err;
        err = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
        (void) hipGraphDestroy(g);
        if (err != hipSuccess) { b->exec = nullptr; return err; }
        b->exec_cfg = b->cfg;
    }
    return hipGraphLaunch(b->exec, s);
}

void batch_release(Batch* b)
{
    if (b->exec) (void) hipGraphExecDestroy(b->exec);
    if (b->cap_stream) (void) hipStreamDestroy(b->cap_stream);
    if (b->ev_upload) (void) hipEventDestroy(b->ev_upload);
    if (b->ev_start) (void) hipEventDestroy(b->ev_start);
    if (b->ev_end) (void) hipEventDestroy(b->ev_end);
    if (b->ws.ptr) (void) hipFree(b->ws.ptr);
    if (b->d_hdr) (void) hipFree(b->d_hdr);
    if (b->d_desc) (void) hipFree(b->d_desc);
    if (b->d_counter) (void) hipFree(b->d_counter);
    if (b->h_hdr) (void) hipHostFree(b->h_hdr);
    if (b->h_desc) (void) hipHostFree(b->h_desc);
    delete b;
}

} /* namespace */

extern "C" void* oskar_amd_fact_batch_create(int precision, int mode,
        int num_stations, int max_units, int* err_out)
{
    Batch* b = new Batch;
    b->precision = precision;
    b->mode = mode;
    b->N = num_stations;
    b->max_units = max_units;
    b->units.assign(max_units, UnitDesc());
    b->unit_set.assign(max_units, 0);
    hipError_t err = hipGetDevice(&b->dev);
    if (err == hipSuccess)
        err = hipDeviceGetAttribute(&b->num_cus,
                hipDeviceAttributeMultiprocessorCount, b->dev);
    if (err == hipSuccess)
        err = hipHostMalloc((void**) &b->h_hdr, sizeof(BatchHeader), 0);
    if (err == hipSuccess)
        err = hipHostMalloc((void**) &b->h_desc,
                max_units * sizeof(UnitDesc), 0);
    if (err == hipSuccess)
        err = hipMalloc((void**) &b->d_hdr, sizeof(BatchHeader));
    if (err == hipSuccess)
        err = hipMalloc((void**) &b->d_desc, max_units * sizeof(UnitDesc));
    if (err == hipSuccess) err = hipMalloc((void**) &b->d_counter, 256);
    if (err == hipSuccess) err = hipEventCreate(&b->ev_upload);
    if (err == hipSuccess) err = hipEventCreate(&b->ev_start);
    if (err == hipSuccess) err = hipEventCreate(&b->ev_end);
    if (err == hipSuccess)
        err = hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking);
    if (err != hipSuccess)
    {
        batch_release(b);
        *err_out = (int) err;
        return nullptr;
    }
    *err_out = 0;
    return b;
}

extern "C" void oskar_amd_fact_batch_free(void* h)
{
    Batch* b = (Batch*) h;
    if (!b) return;
    (void) hipDeviceSynchronize();
    batch_release(b);
}

extern "C" void oskar_amd_fact_batch_set_unit(void* h, int unit,
        int num_sources, const void* beam, const void* I, const void* Q,
        const void* U, const void* V, const void* l, const void* m,
        const void* n, const void* st_u, const void* st_v, const void* st_w,
        double frequency_hz, int offset_out, void* vis)
{
    Batch* b = (Batch*) h;
    UnitDesc d;
    memset(&d, 0, sizeof(d));
    d.I = I; d.Q = Q; d.U = U; d.V = V;
    d.l = l; d.m = m; d.n = n;
    d.beam = beam;
    d.st_u = st_u; d.st_v = st_v; d.st_w = st_w;
    d.vis = vis;
    d.f_over_c = frequency_hz / 299792458.0;
    d.num_sources = num_sources;
    d.offset_out = offset_out;
    b->units[unit] = d;
    b->unit_set[unit] = 1;
}

extern "C" int oskar_amd_fact_batch_unit_is_set(const void* h, int unit)
{
    return ((const Batch*) h)->unit_set[unit];
}

extern "C" int oskar_amd_fact_batch_set_config(void* h, int shape,
        int src_per_part, int workgroups, int use_graph)
{
    Batch* b = (Batch*) h;
    if (shape < 0 || shape >= NUM_SHAPES || src_per_part < KSRC ||
            workgroups < 0) return (int) hipErrorInvalidValue;
    b->cfg.shape = shape;
    b->cfg.src_per_part = (src_per_part + KSRC - 1) / KSRC * KSRC;
    b->cfg.workgroups = workgroups;
    b->use_graph = use_graph != 0;
    return 0;
}

extern "C" void oskar_amd_fact_batch_get_config(const void* h, int* out)
{
    const Batch* b = (const Batch*) h;
    out[0] = b->cfg.shape;
    out[1] = b->cfg.src_per_part;
    out[2] = b->cfg.workgroups > 0 ? b->cfg.workgroups :
            b->num_cus * occupancy(b->cfg.shape);
    out[3] = b->use_graph;
}

extern "C" int oskar_amd_fact_batch_run(void* h, int num_units,
        int use_casa, const double* uv_min, const double* uv_max,
        void* stream_v)
{
    Batch* b = (Batch*) h;
    const hipStream_t s = (hipStream_t) stream_v;
    if (num_units <= 0) return 0;
    hipError_t err = batch_upload(b, num_units, use_casa, uv_min, uv_max,
            nullptr, s);
    if (err == hipSuccess) err = hipEventRecord(b->ev_start, s);
    if (err == hipSuccess) err = batch_execute(b, s);
    if (err == hipSuccess) err = hipEventRecord(b->ev_end, s);
    b->timed = (err == hipSuccess);
    return (int) err;
}

/* Elapsed time (ms) of the last run's kernels; waits for them. */
extern "C" int oskar_amd_fact_batch_last_ms(void* h, double* ms)
{
    Batch* b = (Batch*) h;
    if (!b->timed) return -1;
    float t = 0.0f;
    if (hipEventSynchronize(b->ev_end) != hipSuccess ||
            hipEventElapsedTime(&t, b->ev_start, b->ev_end) != hipSuccess)
        return -1;
    *ms = (double) t;
    return 0;
}

/* Times every candidate configuration (tile shape x part size x grid) on
 * the first num_units units, with outputs sent to a scratch buffer (the
 * callers' visibilities are not touched), and keeps the fastest.  Direct
 * launches, one warm-up and `reps` timed runs each (median).  Part sizes
 * stop at 8192 sources: the FP32 accumulation error grows about as the
 * square root of the part size (measured 3e-6 at 8192, up to 9e-6 at
 * 16384), and 8192 keeps a factor of two below 1e-5. */
extern "C" int oskar_amd_fact_batch_autotune(void* h, int num_units,
        int reps, void* stream_v, double* best_ms)
{
    Batch* b = (Batch*) h;
    const hipStream_t s = (hipStream_t) stream_v;
    if (num_units <= 0) return (int) hipErrorInvalidValue;
    if (reps < 1) reps = 1;
    const size_t nb = (size_t) b->N * (b->N - 1) / 2;
    const size_t vis_bytes = (size_t) num_units * nb *
            (b->precision == OSKAR_DOUBLE ? 64 : 32);
    void* scratch = nullptr;
    hipError_t err = hipMalloc(&scratch, vis_bytes);
    if (err != hipSuccess) return (int) err;
    const BatchConfig saved = b->cfg;
    const int saved_graph = b->use_graph;
    b->use_graph = 0;
    BatchConfig best = saved;
    double best_t = 1e300;
    static const int parts[] = {4096, 8192};
    for (int shape = 0; shape < NUM_SHAPES && err == hipSuccess; ++shape)
    {
        for (int pi = 0; pi < 2 && err == hipSuccess; ++pi)
        {
            for (int occ = 1; occ <= occupancy(shape) && err == hipSuccess;
                    ++occ)
            {
                BatchConfig c;
                c.shape = shape;
                c.src_per_part = parts[pi];
                c.workgroups = b->num_cus * occ;
                b->cfg = c;
                std::vector<double> t;
                for (int r = 0; r <= reps && err == hipSuccess; ++r)
                {
                    err = batch_upload(b, num_units, 0, nullptr, nullptr,
                            scratch, s);
                    if (err == hipSuccess) err = hipEventRecord(b->ev_start, s);
                    if (err == hipSuccess) err = batch_launch(b, s);
                    if (err == hipSuccess) err = hipEventRecord(b->ev_end, s);
                    if (err == hipSuccess) err = hipEventSynchronize(b->ev_end);
                    float ms = 0.0f;
                    if (err == hipSuccess)
                        err = hipEventElapsedTime(&ms, b->ev_start, b->ev_end);
                    if (err == hipSuccess && r > 0) t.push_back(ms);
                }
                if (err != hipSuccess) break;
                std::sort(t.begin(), t.end());
                const double med = t[t.size() / 2];
                if (med < best_t) { best_t = med; best = c; }
            }
        }
    }
    const hipError_t e2 = hipDeviceSynchronize();
    (void) hipFree(scratch);
    b->use_graph = saved_graph;
    b->cfg = (err == hipSuccess) ? best : saved;
    b->timed = false;
    if (err == hipSuccess) err = e2;
    if (err == hipSuccess && best_ms) *best_ms = best_t;
    return (int) err;
}
