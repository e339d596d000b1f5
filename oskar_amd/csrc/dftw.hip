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
 *  - Single precision uses the hardware sine/cosine (__sincosf) of the
 *    FP32 phase; double precision uses the full-accuracy sincos.
 */
#include <hip/hip_runtime.h>

#include "device_common.h"
#include "private.h"

using namespace oskar_amd;

namespace {

constexpr int THREADS = 256;
constexpr int CHUNK = 256;

template <typename FP> __device__ __forceinline__ void fast_sincos(FP t,
        FP* s, FP* c);
template <> __device__ __forceinline__ void fast_sincos<float>(float t,
        float* s, float* c) { __sincosf(t, s, c); }
template <> __device__ __forceinline__ void fast_sincos<double>(double t,
        double* s, double* c) { sincos(t, s, c); }

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
    __shared__ Cplx<FP> c_w[CHUNK];
    __shared__ FP c_x[CHUNK], c_y[CHUNK], c_z[CHUNK];
    __shared__ int c_idx[CHUNK];

    const int i_out = blockIdx.x * THREADS + threadIdx.x;
    const bool active = i_out < num_out;
    FP xo = (FP) 0, yo = (FP) 0, zo = (FP) 0;
    if (active)
    {
        xo = wavenumber * x_out[i_out + offset_coord_out];
        yo = wavenumber * y_out[i_out + offset_coord_out];
        if (IS_3D) zo = wavenumber * z_out[i_out + offset_coord_out];
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

    for (int j = 0; j < num_in; j += CHUNK)
    {
        const int chunk = min(CHUNK, num_in - j);
        __syncthreads();
        for (int t = threadIdx.x; t < chunk; t += THREADS)
        {
            const int g = j + t;
            c_w[t] = weights[g];
            c_x[t] = x_in[g];
            c_y[t] = y_in[g];
            if (IS_3D) c_z[t] = z_in[g];
            c_idx[t] = data_idx ? data_idx[g] : g;
        }
        __syncthreads();
        for (int i = 0; i < chunk; ++i)
        {
            const int idx = c_idx[i];
            if (idx != run_idx)
            {
                flush();
                run_idx = idx;
            }
            FP t = xo * c_x[i] + yo * c_y[i];
            if (IS_3D) t += zo * c_z[i];
            FP s, c;
            fast_sincos<FP>(t, &s, &c);
            const Cplx<FP> w = c_w[i];
            /* af += w * exp(i t) */
            af.x += w.x * c - w.y * s;
            af.y += w.x * s + w.y * c;
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
