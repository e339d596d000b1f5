/*
 * Weighted DFT (array factor x element data) for gfx950.
 *
 * Computes, for each output direction j,
 *   out[j] = norm * sum_i data[idx_i][j] * w_i * exp(i k (x_j x_i + y_j y_i
 *            [+ z_j z_i]))
 * (oskar/math/define_dftw_c2c.h:24-77, define_dftw_m2m.h:24-107).
 *
 * Design:
 *  - One thread per output direction, 256 per workgroup.  Element weights,
 *    positions and data indices are staged through LDS in chunks of 256 and
 *    read as wavefront-uniform (broadcast) values.
 *  - The data index of element i is the same for every thread, so the sum
 *    is formed as a run-length product: consecutive elements that share a
 *    data index are summed into one complex array factor, which multiplies
 *    the (coalesced) data value once per run.  With one element type this
 *    reads the element pattern once per direction instead of once per
 *    element; with data_idx absent each element is its own run and the
 *    kernel reads data per element as the reference does.
 *  - Elements are staged with position and weight in one LDS record (a
 *    broadcast read per element).  A chunk whose elements all use one
 *    data row (the usual case: one element type, or long runs of one)
 *    skips the per-element run check and unrolls the element loop by 4.
 *  - Single precision carries the phase in revolutions (the wavenumber
 *    times 1/(2 pi) folded into the direction once per thread), reduces it
 *    to [-1/2, 1/2] and uses the hardware sine and cosine, which take
 *    revolutions; each term is four fused multiply-adds into the array
 *    factor.  Double precision uses radians and the full-accuracy sincos.
 */
#include <hip/hip_runtime.h>
#include <type_traits>

#include "device_common.h"
#include "dipole_device.h"
#include "private.h"

using namespace oskar_amd;

namespace {

constexpr int THREADS = 256;
constexpr int CHUNK = 256;

/* exp(i t) of the phase in the unit the kernel carries it in: FP32 phases
 * are in revolutions for the hardware sine and cosine (which take
 * revolutions), reduced to [-1/2, 1/2] first unless DIRECT; FP64 phases
 * are in radians, with the full-accuracy sincos.  The hardware reduces a
 * phase of up to 256 revolutions exactly (the same errors against the FP64
 * DFT with and without the reduction at up to 250 revolutions, measured),
 * so DIRECT serves when |t| is bounded below that: two fewer instructions
 * per term. */
template <bool DIRECT, typename FP>
__device__ __forceinline__ void phase_sincos(FP t, FP* s, FP* c)
{
    if constexpr (sizeof(FP) == 8)
        sincos(t, s, c);
    else
    {
        const float r = DIRECT ? t : t - __builtin_rintf(t);
        *s = __builtin_amdgcn_sinf(r);
        *c = __builtin_amdgcn_cosf(r);
    }
}

/* Factor from k (x_out x_in + ...) to the kernel's phase unit. */
template <typename FP> __device__ __forceinline__ FP phase_unit();
template <> __device__ __forceinline__ float phase_unit<float>()
{ return 0.159154943091895336f; }           /* 1 / (2 pi) */
template <> __device__ __forceinline__ double phase_unit<double>()
{ return 1.0; }

/* One element of the chunk staged in LDS: position and weight. */
template <typename FP> struct Elem
{
    FP x, y, z, pad;
    Cplx<FP> w;
};

/* Largest |coordinate| of the chunk's elements, gathered in LDS as float
 * bits (non-negative floats order as unsigned integers). */
template <typename FP>
__device__ __forceinline__ void chunk_extent(unsigned int* bits,
        const Elem<FP>& e)
{
    if constexpr (sizeof(FP) == 4)
    {
        const float a = fmaxf(fabsf(e.x), fmaxf(fabsf(e.y), fabsf(e.z)));
        atomicMax(bits, __float_as_uint(a));
    }
}

/* Whether every phase of a thread's direction (xo, yo, zo, in revolutions
 * per metre) over elements of extent at most ext stays within the
 * hardware's exact reduction: |t| <= (|xo| + |yo| + |zo|) ext + 1/2 (the
 * weight's phase) < 255. */
template <typename FP>
__device__ __forceinline__ bool direct_phase(FP xo, FP yo, FP zo,
        unsigned int ext_bits)
{
    if constexpr (sizeof(FP) == 8) return false;
    else
        return (fabsf(xo) + fabsf(yo) + fabsf(zo)) *
                __uint_as_float(ext_bits) < 254.0f;
}

/* An element's weight in the form the terms use.  FP32: amplitude and
 * phase in revolutions, the phase then added into the element's phase by
 * the first fused multiply-add, so that a term takes two fused
 * multiply-adds into the array factor instead of the four of a complex
 * product (the weight's phase in FP32 revolutions errs by under 6e-8 of a
 * revolution, well under the FP32 phase of the element itself).  FP64:
 * the complex weight as it is. */
template <typename FP>
__device__ __forceinline__ Cplx<FP> term_weight(const Cplx<FP>& w)
{
    if constexpr (sizeof(FP) == 4)
    {
        Cplx<FP> p;
        p.x = sqrtf(w.x * w.x + w.y * w.y);
        p.y = atan2f(w.y, w.x) * 0.159154943091895336f;    /* 1 / (2 pi) */
        return p;
    }
    else
        return w;
}

/* Phase of a term: the direction (xo, yo, zo, already in phase units per
 * metre) dotted with the element's position, plus the FP32 weight's
 * phase. */
template <typename FP, bool IS_3D>
__device__ __forceinline__ FP term_phase(FP xo, FP yo, FP zo, FP ex, FP ey,
        FP ez, const Cplx<FP>& w)
{
    if constexpr (sizeof(FP) == 4)
    {
        float t = __builtin_fmaf(xo, ex, w.y);
        t = __builtin_fmaf(yo, ey, t);
        if (IS_3D) t = __builtin_fmaf(zo, ez, t);
        return t;
    }
    else
    {
        FP t = xo * ex + yo * ey;
        if (IS_3D) t += zo * ez;
        return t;
    }
}

/* af += w exp(i t), with w in term_weight's form and (s, c) the sine and
 * cosine of the term's phase. */
template <typename FP>
__device__ __forceinline__ void term_add(Cplx<FP>& af, const Cplx<FP>& w,
        FP s, FP c)
{
    if constexpr (sizeof(FP) == 4)
    {
        af.x = __builtin_fmaf(w.x, c, af.x);
        af.y = __builtin_fmaf(w.x, s, af.y);
    }
    else
    {
        af.x += w.x * c;
        af.x -= w.y * s;
        af.y += w.x * s;
        af.y += w.y * c;
    }
}

template <typename FP, bool MATRIX, bool IS_3D>
__global__ __launch_bounds__(THREADS)
void dftw_kernel(const int num_in, const FP wavenumber,
        const Cplx<FP>* __restrict__ weights, const FP* __restrict__ x_in,
        const FP* __restrict__ y_in, const FP* __restrict__ z_in,
        const int offset_coord_out, const int num_out,
        const FP* __restrict__ x_out, const FP* __restrict__ y_out,
        const FP* __restrict__ z_out, const int* __restrict__ data_idx,
        const Cplx<FP>* __restrict__ data, const int num_rows,
        const int eval_x, const int eval_y,
        const int offset_out, Cplx<FP>* __restrict__ output,
        const FP norm_factor)
{
    __shared__ Elem<FP> c_e[CHUNK];
    __shared__ int c_idx[CHUNK];
    __shared__ int c_uniform;
    __shared__ unsigned int c_ext;

    const int i_out = blockIdx.x * THREADS + threadIdx.x;
    const bool active = i_out < num_out;
    FP xo = (FP) 0, yo = (FP) 0, zo = (FP) 0;
    if (active)
    {
        const FP kk = wavenumber * phase_unit<FP>();
        xo = kk * x_out[i_out + offset_coord_out];
        yo = kk * y_out[i_out + offset_coord_out];
        if (IS_3D) zo = kk * z_out[i_out + offset_coord_out];
    }
    constexpr int NC = MATRIX ? 4 : 1;
    Cplx<FP> out[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) out[k].x = out[k].y = (FP) 0;
    Cplx<FP> af; af.x = af.y = (FP) 0;
    int run_idx = -1;

    /* Adds data[run_idx][i_out] * af to the outputs.  Indices outside the
     * num_rows rows held by data contribute nothing. */
    auto flush = [&]() {
        if (run_idx >= 0 && run_idx < num_rows && active)
        {
            const size_t base = (size_t) run_idx * num_out + i_out;
#pragma unroll
            for (int k = 0; k < NC; ++k)
            {
                if (MATRIX && ((k < 2 && !eval_x) || (k >= 2 && !eval_y)))
                    continue;
                const Cplx<FP> d = data[base * NC + k];
                out[k].x += d.x * af.x - d.y * af.y;
                out[k].y += d.y * af.x + d.x * af.y;
            }
        }
        af.x = af.y = (FP) 0;
    };
    /* af += w exp(i t) for element e. */
    auto term = [&](const Elem<FP>& e, auto direct) {
        const FP t = term_phase<FP, IS_3D>(xo, yo, zo, e.x, e.y, e.z, e.w);
        FP s, c;
        phase_sincos<decltype(direct)::value, FP>(t, &s, &c);
        term_add<FP>(af, e.w, s, c);
    };
    const std::true_type direct_t;
    const std::false_type reduce_t;

    for (int j = 0; j < num_in; j += CHUNK)
    {
        const int chunk = min(CHUNK, num_in - j);
        __syncthreads();
        if (threadIdx.x == 0)
        {
            c_uniform = 1;
            c_ext = 0u;
        }
        __syncthreads();
        for (int t = threadIdx.x; t < chunk; t += THREADS)
        {
            const int g = j + t;
            Elem<FP> e;
            e.x = x_in[g];
            e.y = y_in[g];
            e.z = IS_3D ? z_in[g] : (FP) 0;
            e.pad = (FP) 0;
            e.w = term_weight<FP>(weights[g]);
            c_e[t] = e;
            chunk_extent<FP>(&c_ext, e);
            const int idx = data_idx ? data_idx[g] : g;
            c_idx[t] = idx;
            /* Any element whose data row differs from the chunk's first
             * clears the flag (every writer stores the same value). */
            if (idx != (data_idx ? data_idx[j] : j)) c_uniform = 0;
        }
        __syncthreads();
        const bool direct = direct_phase<FP>(xo, yo, zo, c_ext);
        if (c_uniform)
        {
            /* One data row for the whole chunk: no run checks, and the
             * element loop unrolled. */
            const int idx = c_idx[0];
            if (idx != run_idx)
            {
                flush();
                run_idx = idx;
            }
            int i = 0;
            if (direct)
            {
                for (; i + 4 <= chunk; i += 4)
                {
#pragma unroll
                    for (int u = 0; u < 4; ++u) term(c_e[i + u], direct_t);
                }
                for (; i < chunk; ++i) term(c_e[i], direct_t);
            }
            else
            {
                for (; i + 4 <= chunk; i += 4)
                {
#pragma unroll
                    for (int u = 0; u < 4; ++u) term(c_e[i + u], reduce_t);
                }
                for (; i < chunk; ++i) term(c_e[i], reduce_t);
            }
        }
        else
        {
            for (int i = 0; i < chunk; ++i)
            {
                const int idx = c_idx[i];
                if (idx != run_idx)
                {
                    flush();
                    run_idx = idx;
                }
                if (direct) term(c_e[i], direct_t);
                else term(c_e[i], reduce_t);
            }
        }
    }
    flush();
    if (active)
    {
        const size_t o = (size_t) (i_out + offset_out) * NC;
#pragma unroll
        for (int k = 0; k < NC; ++k)
        {
            if (MATRIX && ((k < 2 && !eval_x) || (k >= 2 && !eval_y)))
                continue;
            Cplx<FP> v;
            v.x = out[k].x * norm_factor;
            v.y = out[k].y * norm_factor;
            output[o + k] = v;
        }
    }
}

/* The beam of a flat aperture-array station whose elements are all of one
 * dipole type with a common orientation, in one pass per direction: the
 * array factor (dftw_kernel's sum, in its order), times the direction's
 * dipole pattern (computed here with the pattern kernels' device
 * functions instead of read from a stored pattern), times norm_factor,
 * zero below the horizon (blank_kernel's rule).  With NORM the result is
 * divided as oskar_mem_normalise divides by the beam in the normalisation
 * direction (nx, ny, nz), which every workgroup forms itself: each thread
 * sums the array factor over a stride of the elements, a fixed-order tree
 * in LDS adds the 256 partial sums (so every workgroup gets the same
 * value), and thread 0 applies the pattern, the blank and the scale rule.
 * Only the order of that one sum differs from the separate kernels. */
template <typename FP, bool MATRIX, bool IS_3D, bool NORM>
__global__ __launch_bounds__(THREADS)
void aa_fused_kernel(const int num_in, const FP wavenumber,
        const Cplx<FP>* __restrict__ weights, const FP* __restrict__ x_in,
        const FP* __restrict__ y_in, const FP* __restrict__ z_in,
        const int num_out, const FP* __restrict__ x_out,
        const FP* __restrict__ y_out, const FP* __restrict__ z_out,
        const FP nx, const FP ny, const FP nz, const FP kL, const FP cos_kL,
        const FP dphi_x, const FP dphi_y, const int swap_xy,
        const FP norm_factor, Cplx<FP>* __restrict__ output)
{
    constexpr int NC = MATRIX ? 4 : 1;
    __shared__ Elem<FP> c_e[CHUNK];
    __shared__ Cplx<FP> red[THREADS];
    __shared__ FP s_scal;
    __shared__ unsigned int c_ext;
    const int tid = (int) threadIdx.x;
    const FP kk = wavenumber * phase_unit<FP>();

    /* The array factor plus w exp(i t) of one element, as in dftw_kernel
     * (w in term_weight's form). */
    auto add_term = [](Cplx<FP>& af, FP t, const Cplx<FP>& w, auto direct) {
        FP s, c;
        phase_sincos<decltype(direct)::value, FP>(t, &s, &c);
        term_add<FP>(af, w, s, c);
    };
    /* The beam of one direction from its array factor: pattern x af,
     * scaled, blanked (dftw_kernel's flush and output, blank_kernel). */
    auto beam_of = [&](FP x, FP y, FP z, const Cplx<FP>& af, FP* v) {
        Cplx<FP> d[NC];
        FP theta, p1, p2;
        theta_phi_point<FP>(x, y, z, dphi_x, dphi_y, &theta, &p1, &p2);
        dipole_point<FP, MATRIX>(theta, p1, p2, kL, cos_kL, swap_xy, d, 0);
#pragma unroll
        for (int k = 0; k < NC; ++k)
        {
            Cplx<FP> o;
            o.x = o.y = (FP) 0;
            o.x += d[k].x * af.x - d[k].y * af.y;
            o.y += d[k].y * af.x + d[k].x * af.y;
            v[2 * k] = z < (FP) 0 ? (FP) 0 : o.x * norm_factor;
            v[2 * k + 1] = z < (FP) 0 ? (FP) 0 : o.y * norm_factor;
        }
    };

    FP scal = (FP) 1;
    if constexpr (NORM)
    {
        const FP xn = kk * nx, yn = kk * ny;
        const FP zn = IS_3D ? kk * nz : (FP) 0;
        Cplx<FP> a;
        a.x = a.y = (FP) 0;
        for (int e = tid; e < num_in; e += THREADS)
        {
            const Cplx<FP> w = term_weight<FP>(weights[e]);
            const FP t = term_phase<FP, IS_3D>(xn, yn, zn, x_in[e], y_in[e],
                    IS_3D ? z_in[e] : (FP) 0, w);
            add_term(a, t, w, std::false_type());
        }
        red[tid] = a;
        __syncthreads();
        for (int o = THREADS / 2; o > 0; o >>= 1)
        {
            if (tid < o)
            {
                red[tid].x += red[tid + o].x;
                red[tid].y += red[tid + o].y;
            }
            __syncthreads();
        }
        if (tid == 0)
        {
            FP v[2 * NC];
            beam_of(nx, ny, nz, red[0], v);
            FP amp = (FP) 0;
#pragma unroll
            for (int r = 0; r < 2 * NC; ++r) amp += v[r] * v[r];
            if (MATRIX) amp /= (FP) 2;
            s_scal = (FP) 1 / sqrt(amp);
        }
        __syncthreads();
        scal = s_scal;
    }

    const int i_out = blockIdx.x * THREADS + tid;
    const bool active = i_out < num_out;
    FP xd = (FP) 0, yd = (FP) 0, zd = (FP) 0;
    if (active)
    {
        xd = x_out[i_out];
        yd = y_out[i_out];
        zd = z_out[i_out];
    }
    const FP xo = kk * xd, yo = kk * yd, zo = IS_3D ? kk * zd : (FP) 0;
    Cplx<FP> af;
    af.x = af.y = (FP) 0;
    auto term = [&](const Elem<FP>& e, auto direct) {
        add_term(af, term_phase<FP, IS_3D>(xo, yo, zo, e.x, e.y, e.z, e.w),
                e.w, direct);
    };
    const std::true_type direct_t;
    const std::false_type reduce_t;
    for (int j = 0; j < num_in; j += CHUNK)
    {
        const int chunk = min(CHUNK, num_in - j);
        __syncthreads();
        if (tid == 0) c_ext = 0u;
        __syncthreads();
        for (int t = tid; t < chunk; t += THREADS)
        {
            const int g = j + t;
            Elem<FP> e;
            e.x = x_in[g];
            e.y = y_in[g];
            e.z = IS_3D ? z_in[g] : (FP) 0;
            e.pad = (FP) 0;
            e.w = term_weight<FP>(weights[g]);
            c_e[t] = e;
            chunk_extent<FP>(&c_ext, e);
        }
        __syncthreads();
        int i = 0;
        if (direct_phase<FP>(xo, yo, zo, c_ext))
        {
            for (; i + 4 <= chunk; i += 4)
            {
#pragma unroll
                for (int u = 0; u < 4; ++u) term(c_e[i + u], direct_t);
            }
            for (; i < chunk; ++i) term(c_e[i], direct_t);
        }
        else
        {
            for (; i + 4 <= chunk; i += 4)
            {
#pragma unroll
                for (int u = 0; u < 4; ++u) term(c_e[i + u], reduce_t);
            }
            for (; i < chunk; ++i) term(c_e[i], reduce_t);
        }
    }
    if (!active) return;
    FP v[2 * NC];
    beam_of(xd, yd, zd, af, v);
    Cplx<FP>* o = output + (size_t) i_out * NC;
#pragma unroll
    for (int k = 0; k < NC; ++k)
    {
        Cplx<FP> r;
        r.x = v[2 * k];
        r.y = v[2 * k + 1];
        if constexpr (NORM)
        {
            r.x = r.x * scal;
            r.y = r.y * scal;
        }
        o[k] = r;
    }
}

template <typename FP, bool MATRIX, bool IS_3D>
hipError_t aa_fused_t(int normalise, int num_in, double wavenumber,
        const void* weights, const void* x_in, const void* y_in,
        const void* z_in, int num_out, const void* x_out, const void* y_out,
        const void* z_out, double nx, double ny, double nz, double kL,
        double cos_kL, double dphi_x, double dphi_y, int swap_xy,
        double norm_factor, void* output, hipStream_t stream)
{
    const dim3 grid((num_out + THREADS - 1) / THREADS);
#define AF_ARGS num_in, (FP) wavenumber, (const Cplx<FP>*) weights,       \
        (const FP*) x_in, (const FP*) y_in, (const FP*) z_in, num_out,      \
        (const FP*) x_out, (const FP*) y_out, (const FP*) z_out, (FP) nx,   \
        (FP) ny, (FP) nz, (FP) kL, (FP) cos_kL, (FP) dphi_x, (FP) dphi_y,   \
        swap_xy, (FP) norm_factor, (Cplx<FP>*) output
    if (normalise)
        hipLaunchKernelGGL((aa_fused_kernel<FP, MATRIX, IS_3D, true>), grid,
                dim3(THREADS), 0, stream, AF_ARGS);
    else
        hipLaunchKernelGGL((aa_fused_kernel<FP, MATRIX, IS_3D, false>), grid,
                dim3(THREADS), 0, stream, AF_ARGS);
#undef AF_ARGS
    return hipGetLastError();
}

template <typename FP>
hipError_t launch_fp(int matrix, int is_3d, int num_in, double wavenumber,
        const void* weights, const void* x_in, const void* y_in,
        const void* z_in, int offset_coord_out, int num_out,
        const void* x_out, const void* y_out, const void* z_out,
        const int* data_idx, const void* data, int num_rows, int eval_x,
        int eval_y, int offset_out, void* output, double norm_factor,
        hipStream_t stream)
{
    if (num_out <= 0) return hipSuccess;
    const dim3 grid((num_out + THREADS - 1) / THREADS);
#define DFTW_ARGS num_in, (FP) wavenumber, (const Cplx<FP>*) weights,       \
        (const FP*) x_in, (const FP*) y_in, (const FP*) z_in,               \
        offset_coord_out, num_out, (const FP*) x_out, (const FP*) y_out,    \
        (const FP*) z_out, data_idx, (const Cplx<FP>*) data, num_rows,      \
        eval_x, eval_y,                                                     \
        offset_out, (Cplx<FP>*) output, (FP) norm_factor
    if (matrix && is_3d)
        hipLaunchKernelGGL((dftw_kernel<FP, true, true>), grid, dim3(THREADS),
                0, stream, DFTW_ARGS);
    else if (matrix)
        hipLaunchKernelGGL((dftw_kernel<FP, true, false>), grid,
                dim3(THREADS), 0, stream, DFTW_ARGS);
    else if (is_3d)
        hipLaunchKernelGGL((dftw_kernel<FP, false, true>), grid,
                dim3(THREADS), 0, stream, DFTW_ARGS);
    else
        hipLaunchKernelGGL((dftw_kernel<FP, false, false>), grid,
                dim3(THREADS), 0, stream, DFTW_ARGS);
#undef DFTW_ARGS
    return hipGetLastError();
}

} /* namespace */

extern "C" int oskar_amd_dftw_launch(int precision, int matrix, int is_3d,
        int num_in, double wavenumber, const void* weights,
        const void* x_in, const void* y_in, const void* z_in,
        int offset_coord_out, int num_out, const void* x_out,
        const void* y_out, const void* z_out, const int* data_idx,
        const void* data, int num_rows, int eval_x, int eval_y,
        int offset_out, void* output, double norm_factor, void* stream)
{
    if (precision == OSKAR_DOUBLE)
        return (int) launch_fp<double>(matrix, is_3d, num_in, wavenumber,
                weights, x_in, y_in, z_in, offset_coord_out, num_out, x_out,
                y_out, z_out, data_idx, data, num_rows, eval_x, eval_y,
                offset_out, output, norm_factor, (hipStream_t) stream);
    return (int) launch_fp<float>(matrix, is_3d, num_in, wavenumber,
            weights, x_in, y_in, z_in, offset_coord_out, num_out, x_out,
            y_out, z_out, data_idx, data, num_rows, eval_x, eval_y,
            offset_out, output, norm_factor, (hipStream_t) stream);
}

extern "C" int oskar_amd_aa_fused_launch(int precision, int matrix,
        int is_3d, int normalise, int num_in, double wavenumber,
        const void* weights, const void* x_in, const void* y_in,
        const void* z_in, int num_out, const void* x_out, const void* y_out,
        const void* z_out, double nx, double ny, double nz, double kL,
        double cos_kL, double dphi_x, double dphi_y, int swap_xy,
        double norm_factor, void* output, void* stream)
{
    if (num_out <= 0) return 0;
    const hipStream_t s = (hipStream_t) stream;
#define AF(FP, M, D) return (int) aa_fused_t<FP, M, D>(normalise, num_in, \
        wavenumber, weights, x_in, y_in, z_in, num_out, x_out, y_out, z_out, \
        nx, ny, nz, kL, cos_kL, dphi_x, dphi_y, swap_xy, norm_factor, output, \
        s)
    if (precision == OSKAR_DOUBLE)
    {
        if (matrix) { if (is_3d) AF(double, true, true); AF(double, true, false); }
        if (is_3d) AF(double, false, true);
        AF(double, false, false);
    }
    if (matrix) { if (is_3d) AF(float, true, true); AF(float, true, false); }
    if (is_3d) AF(float, false, true);
    AF(float, false, false);
#undef AF
}
