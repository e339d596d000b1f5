/*
 * Horizon clip of the sky model: the horizon mask, the exclusive prefix sum
 * and the masked row copy, separately (the reference's three entry points)
 * and fused into one clip.
 *
 * Reference: oskar/sky/define_update_horizon_mask.h:3-12 (mask[i] |=
 * l ll + m mm + n nn > 0), oskar/math/define_prefix_sum.h (one workgroup of
 * at most 512 threads sweeping the whole array), oskar/sky/
 * define_sky_copy_source_data.h:3-25 (row copy to indices[i] where mask[i]),
 * oskar/sky/src/oskar_sky_horizon_clip.c:20-106 (clear, one mask update per
 * station model, prefix sum, copy).
 *
 * All of it moves integers and table rows, so it is bound by HBM traffic:
 * per source, 3 coordinates, 4 B of mask, 4 B of index, the row (num_columns
 * values) and 80 B of valid-column counts when kept.  Design:
 *  - the mask takes the OR over up to 16 station models in one pass, with
 *    the dot product evaluated ((l ll + m mm) + n nn) with every product and
 *    sum rounded (no fused multiply-add), as the reference's CPU body, so
 *    the mask agrees with it bit for bit;
 *  - tiles of 2048 sources per 256-thread workgroup, read in rows of 256
 *    (thread t takes source base + 256 j + t, so every load is coalesced);
 *    a 0/1 mask is scanned with wave ballots and population counts;
 *  - block counts -> one workgroup scans them into block offsets and
 *    writes the total after the last index -> each block scans its tile
 *    again, writes the indices of all its sources and scatters its kept
 *    rows;
 *  - the general prefix sum (any int input) gives each thread 8 consecutive
 *    values (two 16-byte loads), scans them serially and the thread totals
 *    across the workgroup.
 * No atomics and no spin waits: results do not depend on timing.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "private.h"

namespace {

constexpr int THREADS = 256;
constexpr int WAVES = THREADS / 64;
constexpr int ROWS = 8;                   /* rows of 256 sources per tile */
constexpr int TILE = THREADS * ROWS;      /* 2048 */
constexpr int NFIX = (int) OSKAR_SKY_NUM_FIXED_COLUMN_TYPES;
static_assert(NFIX % 4 == 0, "valid-column rows are copied as int4");

template <typename FP> struct Models
{
    int num;
    FP lmn[SKY_CLIP_MAX_MODELS][3];
};

/* Sine of the elevation of (l, m, n) for one station model, with rounded
 * products and sums in the order of define_update_horizon_mask.h:7. */
template <typename FP>
__device__ __forceinline__ bool above(FP l, FP m, FP n, const FP* f)
{
#pragma clang fp contract(off)
    const FP a = l * f[0];
    const FP b = m * f[1];
    const FP c = n * f[2];
    const FP s = (a + b) + c;
    return s > (FP) 0;
}

template <typename FP>
__device__ __forceinline__ int mask_of(const Models<FP>& M, FP l, FP m, FP n)
{
    int bit = 0;
    for (int k = 0; k < M.num; ++k) bit |= above<FP>(l, m, n, M.lmn[k]);
    return bit;
}

__device__ __forceinline__ int lane_id() { return (int) threadIdx.x & 63; }

/* Inclusive scan of v over the 64 lanes of a wave. */
__device__ __forceinline__ int wave_scan_incl(int v)
{
    const int lane = lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1)
    {
        const int y = __shfl_up(v, d);
        if (lane >= d) v += y;
    }
    return v;
}

/* Exclusive scan of v over the workgroup (NT threads); *total = the sum.
 * Uses lds[NT / 64 + 1]. */
template <int NT>
__device__ __forceinline__ int block_scan_excl(int v, int* lds, int* total)
{
    constexpr int NW = NT / 64;
    const int lane = lane_id(), wave = (int) threadIdx.x >> 6;
    const int incl = wave_scan_incl(v);
    if (lane == 63) lds[wave] = incl;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        int run = 0;
        for (int w = 0; w < NW; ++w)
        {
            const int x = lds[w];
            lds[w] = run;
            run += x;
        }
        lds[NW] = run;
    }
    __syncthreads();
    const int out = lds[wave] + incl - v;
    *total = lds[NW];
    __syncthreads();
    return out;
}

/* ---- standalone horizon mask ------------------------------------------ */
template <typename FP>
__global__ __launch_bounds__(THREADS)
void horizon_mask_kernel(int num, const FP* __restrict__ l,
        const FP* __restrict__ m, const FP* __restrict__ n, Models<FP> M,
        int* __restrict__ mask)
{
    const int stride = (int) (gridDim.x * blockDim.x);
    for (int i = (int) (blockIdx.x * blockDim.x + threadIdx.x); i < num;
            i += stride)
    {
        if (mask_of<FP>(M, l[i], m[i], n[i])) mask[i] = 1;
    }
}

/* ---- block counts -------------------------------------------------------
 * Fused clip: with M.num > 0 the mask is computed (and written), else the
 * 0/1 mask already in memory is read.  Writes the tile's count of set
 * entries to counts[blockIdx.x]. */
template <typename FP>
__global__ __launch_bounds__(THREADS)
void clip_count_kernel(int num, const FP* __restrict__ l,
        const FP* __restrict__ m, const FP* __restrict__ n, Models<FP> M,
        int* __restrict__ mask, int* __restrict__ counts)
{
    __shared__ int lds[WAVES];
    const int base = (int) blockIdx.x * TILE;
    int cnt = 0;
#pragma unroll
    for (int j = 0; j < ROWS; ++j)
    {
        const int i = base + j * THREADS + (int) threadIdx.x;
        int bit = 0;
        if (i < num)
        {
            if (M.num > 0)
            {
                bit = mask_of<FP>(M, l[i], m[i], n[i]);
                mask[i] = bit;
            }
            else bit = (mask[i] != 0);
        }
        cnt += __popcll(__ballot(bit));
    }
    if (lane_id() == 0) lds[threadIdx.x >> 6] = cnt;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        int s = 0;
        for (int w = 0; w < WAVES; ++w) s += lds[w];
        counts[blockIdx.x] = s;
    }
}

/* General prefix sum, pass 1: the sum of each tile of 2048 ints. */
__global__ __launch_bounds__(THREADS)
void scan_reduce_kernel(int num, int vec, const int* __restrict__ in,
        int* __restrict__ counts)
{
    __shared__ int lds[WAVES + 1];
    const int i0 = (int) blockIdx.x * TILE + (int) threadIdx.x * ROWS;
    int s = 0;
    if (vec && i0 + ROWS <= num)
    {
        const int4 a = *reinterpret_cast<const int4*>(in + i0);
        const int4 b = *reinterpret_cast<const int4*>(in + i0 + 4);
        s = ((a.x + a.y) + (a.z + a.w)) + ((b.x + b.y) + (b.z + b.w));
    }
    else
    {
        for (int k = 0; k < ROWS; ++k)
            if (i0 + k < num) s += in[i0 + k];
    }
    int total = 0;
    (void) block_scan_excl<THREADS>(s, lds, &total);
    if (threadIdx.x == 0) counts[blockIdx.x] = total;
}

/* Block offsets: exclusive scan of counts[0, nb) in one workgroup of 1024
 * threads, and the grand total to *total_out. */
__global__ __launch_bounds__(1024)
void scan_offsets_kernel(int nb, int* __restrict__ counts,
        int* __restrict__ total_out)
{
    __shared__ int lds[1024 / 64 + 1];
    int carry = 0;
    for (int c0 = 0; c0 < nb; c0 += 1024)
    {
        const int i = c0 + (int) threadIdx.x;
        const int v = (i < nb) ? counts[i] : 0;
        int total = 0;
        const int ex = block_scan_excl<1024>(v, lds, &total);
        if (i < nb) counts[i] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0) *total_out = carry;
}

/* General prefix sum, pass 2: each thread's 8 values scanned serially on
 * top of the thread's offset within the tile and the tile's offset. */
__global__ __launch_bounds__(THREADS)
void scan_final_kernel(int num, int vec, const int* __restrict__ in,
        const int* __restrict__ offsets, int* __restrict__ out)
{
    __shared__ int lds[WAVES + 1];
    const int i0 = (int) blockIdx.x * TILE + (int) threadIdx.x * ROWS;
    int v[ROWS];
    const bool full = vec && (i0 + ROWS <= num);
    if (full)
    {
        const int4 a = *reinterpret_cast<const int4*>(in + i0);
        const int4 b = *reinterpret_cast<const int4*>(in + i0 + 4);
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
        v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
    }
    else
    {
#pragma unroll
        for (int k = 0; k < ROWS; ++k) v[k] = (i0 + k < num) ? in[i0 + k] : 0;
    }
    int s = 0;
#pragma unroll
    for (int k = 0; k < ROWS; ++k) s += v[k];
    int total = 0;
    int run = block_scan_excl<THREADS>(s, lds, &total) + offsets[blockIdx.x];
    int o[ROWS];
#pragma unroll
    for (int k = 0; k < ROWS; ++k)
    {
        o[k] = run;
        run += v[k];
    }
    if (full)
    {
        *reinterpret_cast<int4*>(out + i0) = make_int4(o[0], o[1], o[2], o[3]);
        *reinterpret_cast<int4*>(out + i0 + 4) =
                make_int4(o[4], o[5], o[6], o[7]);
    }
    else
    {
#pragma unroll
        for (int k = 0; k < ROWS; ++k)
            if (i0 + k < num) out[i0 + k] = o[k];
    }
}

/* Copies row i of the input table to row r of the output table, with its
 * valid-column counts. */
template <typename FP>
__device__ __forceinline__ void copy_row(int i, int r, int cap_in,
        int cap_out, int num_columns, const FP* __restrict__ tin,
        FP* __restrict__ tout, const int* __restrict__ vin,
        int* __restrict__ vout)
{
    for (int c = 0; c < num_columns; ++c)
        tout[(size_t) c * cap_out + r] = tin[(size_t) c * cap_in + i];
    const int4* src = reinterpret_cast<const int4*>(vin + (size_t) NFIX * i);
    int4* dst = reinterpret_cast<int4*>(vout + (size_t) NFIX * r);
#pragma unroll
    for (int k = 0; k < NFIX / 4; ++k) dst[k] = src[k];
}

/* Fused clip, last pass: indices of every source of the tile (exclusive
 * scan of the 0/1 mask in source order) and the kept rows scattered. */
template <typename FP>
__global__ __launch_bounds__(THREADS)
void clip_scatter_kernel(int num, const int* __restrict__ mask,
        const int* __restrict__ offsets, int* __restrict__ indices,
        int cap_in, int cap_out, int num_columns, const FP* __restrict__ tin,
        FP* __restrict__ tout, const int* __restrict__ vin,
        int* __restrict__ vout)
{
    __shared__ int pre[ROWS * WAVES];
    const int base = (int) blockIdx.x * TILE;
    const int lane = lane_id(), wave = (int) threadIdx.x >> 6;
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    uint64_t bal[ROWS];
#pragma unroll
    for (int j = 0; j < ROWS; ++j)
    {
        const int i = base + j * THREADS + (int) threadIdx.x;
        const int bit = (i < num) ? (mask[i] != 0) : 0;
        bal[j] = __ballot(bit);
        if (lane == 0) pre[j * WAVES + wave] = __popcll(bal[j]);
    }
    __syncthreads();
    /* Exclusive scan of the (row, wave) counts in source order. */
    if (threadIdx.x < 64)
    {
        const int v = (lane < ROWS * WAVES) ? pre[lane] : 0;
        const int ex = wave_scan_incl(v) - v;
        if (lane < ROWS * WAVES) pre[lane] = ex;
    }
    __syncthreads();
    const int off = offsets[blockIdx.x];
#pragma unroll
    for (int j = 0; j < ROWS; ++j)
    {
        const int i = base + j * THREADS + (int) threadIdx.x;
        if (i >= num) continue;
        const int r = off + pre[j * WAVES + wave] + __popcll(bal[j] & below);
        indices[i] = r;
        if ((bal[j] >> lane) & 1ull)
            copy_row<FP>(i, r, cap_in, cap_out, num_columns, tin, tout, vin,
                    vout);
    }
}

/* Standalone copy with caller-given indices. */
template <typename FP>
__global__ __launch_bounds__(THREADS)
void sky_copy_kernel(int num, int cap_in, int cap_out,
        const int* __restrict__ mask, const int* __restrict__ indices,
        int num_columns, const FP* __restrict__ tin, FP* __restrict__ tout,
        const int* __restrict__ vin, int* __restrict__ vout)
{
    const int stride = (int) (gridDim.x * blockDim.x);
    for (int i = (int) (blockIdx.x * blockDim.x + threadIdx.x); i < num;
            i += stride)
    {
        if (mask[i])
            copy_row<FP>(i, indices[i], cap_in, cap_out, num_columns, tin,
                    tout, vin, vout);
    }
}

/* Per-device block-count buffer, grown on demand (hipFree waits for the
 * device, so launches that still read the old one have finished). */
int* counts_buffer(size_t n, hipError_t* err)
{
    static void* ptr[64];
    static size_t size[64];
    int dev = 0;
    *err = hipGetDevice(&dev);
    if (*err != hipSuccess) return nullptr;
    if (dev < 0 || dev >= 64) { *err = hipErrorInvalidDevice; return nullptr; }
    const size_t bytes = (n > 0 ? n : 1) * sizeof(int);
    if (size[dev] < bytes)
    {
        if (ptr[dev]) (void) hipFree(ptr[dev]);
        ptr[dev] = nullptr;
        size[dev] = 0;
        *err = hipMalloc(&ptr[dev], bytes);
        if (*err != hipSuccess) { ptr[dev] = nullptr; return nullptr; }
        size[dev] = bytes;
    }
    return (int*) ptr[dev];
}

template <typename FP>
Models<FP> to_models(const oskar_HorizonModels* h)
{
    Models<FP> M;
    M.num = h ? h->num : 0;
    for (int k = 0; k < SKY_CLIP_MAX_MODELS; ++k)
        for (int d = 0; d < 3; ++d)
            M.lmn[k][d] = (h && k < h->num) ? (FP) h->lmn[k][d] : (FP) 0;
    return M;
}

int grid_for(int num)
{
    const int g = (num + THREADS - 1) / THREADS;
    return g < 1 ? 1 : (g > 8192 ? 8192 : g);
}

template <typename FP>
hipError_t clip_t(int num_in, const void* l, const void* m, const void* n,
        const oskar_HorizonModels* models, int* mask, int* indices,
        int cap_in, int cap_out, int num_columns, const void* tin, void* tout,
        const int* vin, int* vout, hipStream_t s)
{
    const int nb = (num_in + TILE - 1) / TILE;
    hipError_t err = hipSuccess;
    int* counts = counts_buffer((size_t) nb, &err);
    if (err != hipSuccess) return err;
    if (nb > 0)
    {
        hipLaunchKernelGGL((clip_count_kernel<FP>), dim3(nb), dim3(THREADS),
                0, s, num_in, (const FP*) l, (const FP*) m, (const FP*) n,
                to_models<FP>(models), mask, counts);
        err = hipGetLastError();
        if (err != hipSuccess) return err;
    }
    hipLaunchKernelGGL(scan_offsets_kernel, dim3(1), dim3(1024), 0, s, nb,
            counts, indices + num_in);
    err = hipGetLastError();
    if (err != hipSuccess || nb == 0) return err;
    hipLaunchKernelGGL((clip_scatter_kernel<FP>), dim3(nb), dim3(THREADS), 0,
            s, num_in, (const int*) mask, (const int*) counts, indices,
            cap_in, cap_out, num_columns, (const FP*) tin, (FP*) tout, vin,
            vout);
    return hipGetLastError();
}

} /* namespace */

extern "C" int oskar_amd_horizon_mask_launch(int precision, int num,
        const void* l, const void* m, const void* n,
        const oskar_HorizonModels* models, int* mask, void* stream)
{
    if (num <= 0 || !models || models->num <= 0) return 0;
    hipStream_t s = (hipStream_t) stream;
    if (precision == OSKAR_SINGLE)
        hipLaunchKernelGGL((horizon_mask_kernel<float>), dim3(grid_for(num)),
                dim3(THREADS), 0, s, num, (const float*) l, (const float*) m,
                (const float*) n, to_models<float>(models), mask);
    else
        hipLaunchKernelGGL((horizon_mask_kernel<double>),
                dim3(grid_for(num)), dim3(THREADS), 0, s, num,
                (const double*) l, (const double*) m, (const double*) n,
                to_models<double>(models), mask);
    return oskar_amd_hip_status((int) hipGetLastError());
}

extern "C" int oskar_amd_prefix_sum_launch(int num, const int* in, int* out,
        void* stream)
{
    hipStream_t s = (hipStream_t) stream;
    const int nb = (num + TILE - 1) / TILE;
    hipError_t err = hipSuccess;
    int* counts = counts_buffer((size_t) nb, &err);
    if (err != hipSuccess) return oskar_amd_hip_status((int) err);
    /* 16-byte loads and stores need 16-byte aligned rows of 8 values. */
    const int vec = !(((uintptr_t) in & 15) || ((uintptr_t) out & 15));
    if (nb > 0)
    {
        hipLaunchKernelGGL(scan_reduce_kernel, dim3(nb), dim3(THREADS), 0, s,
                num, vec, in, counts);
        err = hipGetLastError();
        if (err != hipSuccess) return oskar_amd_hip_status((int) err);
    }
    hipLaunchKernelGGL(scan_offsets_kernel, dim3(1), dim3(1024), 0, s, nb,
            counts, out + num);
    err = hipGetLastError();
    if (err != hipSuccess || nb == 0) return oskar_amd_hip_status((int) err);
    hipLaunchKernelGGL(scan_final_kernel, dim3(nb), dim3(THREADS), 0, s, num,
            vec, in, (const int*) counts, out);
    return oskar_amd_hip_status((int) hipGetLastError());
}

extern "C" int oskar_amd_sky_copy_launch(int precision, int num_in,
        int capacity_in, int capacity_out, const int* mask,
        const int* indices, int num_columns, const void* table_in,
        void* table_out, const int* valid_in, int* valid_out, void* stream)
{
    if (num_in <= 0) return 0;
    hipStream_t s = (hipStream_t) stream;
    if (precision == OSKAR_SINGLE)
        hipLaunchKernelGGL((sky_copy_kernel<float>), dim3(grid_for(num_in)),
                dim3(THREADS), 0, s, num_in, capacity_in, capacity_out, mask,
                indices, num_columns, (const float*) table_in,
                (float*) table_out, valid_in, valid_out);
    else
        hipLaunchKernelGGL((sky_copy_kernel<double>), dim3(grid_for(num_in)),
                dim3(THREADS), 0, s, num_in, capacity_in, capacity_out, mask,
                indices, num_columns, (const double*) table_in,
                (double*) table_out, valid_in, valid_out);
    return oskar_amd_hip_status((int) hipGetLastError());
}

extern "C" int oskar_amd_horizon_clip_launch(int precision, int num_in,
        const void* l, const void* m, const void* n,
        const oskar_HorizonModels* models, int* mask, int* indices,
        int capacity_in, int capacity_out, int num_columns,
        const void* table_in, void* table_out, const int* valid_in,
        int* valid_out, void* stream)
{
    hipStream_t s = (hipStream_t) stream;
    const hipError_t e = (precision == OSKAR_SINGLE) ?
            clip_t<float>(num_in, l, m, n, models, mask, indices, capacity_in,
                    capacity_out, num_columns, table_in, table_out, valid_in,
                    valid_out, s) :
            clip_t<double>(num_in, l, m, n, models, mask, indices,
                    capacity_in, capacity_out, num_columns, table_in,
                    table_out, valid_in, valid_out, s);
    return oskar_amd_hip_status((int) e);
}
