/*
 * Element-wise kernels of the aperture-array station beam and of oskar_Mem
 * utilities, for gfx950.
 *
 *  - DFT beamforming weights w_i = exp(-i k (x_i x_b + y_i y_b + z_i z_b
 *    + cable_i)) (oskar/telescope/station/define_evaluate_element_weights_dft.h).
 *  - ENU direction cosines -> (theta, phi_x, phi_y), with an optional extra
 *    point at the pole (oskar/convert/define_convert_enu_directions_to_theta_phi.h).
 *  - Half-wave dipole pattern E_theta, E_phi for the X and Y dipoles
 *    (oskar/telescope/station/element/define_evaluate_dipole_pattern.h).
 *  - Zeroing below the horizon (oskar/telescope/station/define_blank_below_horizon.h).
 *  - Normalisation by one element (oskar/mem/define_mem_normalise.h).
 *  - Set real value (oskar/mem/define_mem_set_value_real.h).
 *  - Jones join (element-wise 2x2 complex products, oskar/mem/define_mem_multiply.h).
 * All are one thread per element, grid-stride, 256 threads per workgroup.
 */
#include <hip/hip_runtime.h>

#include "device_common.h"
#include "dipole_device.h"
#include "private.h"

using namespace oskar_amd;

namespace {

constexpr int THREADS = 256;

inline dim3 grid_for(size_t n)
{
    size_t g = (n + THREADS - 1) / THREADS;
    if (g > 65535u * 8u) g = 65535u * 8u;
    if (g == 0) g = 1;
    return dim3((unsigned) g);
}

#define GRID_STRIDE(I, N) \
    for (size_t I = (size_t) blockIdx.x * blockDim.x + threadIdx.x; I < (N); \
            I += (size_t) gridDim.x * blockDim.x)

template <typename FP>
__global__ void weights_kernel(int n, const FP* x, const FP* y, const FP* z,
        const FP* cable, FP wavenumber, FP x1, FP y1, FP z1,
        Cplx<FP>* weights)
{
    GRID_STRIDE(i, (size_t) n)
    {
        const FP p = wavenumber * (x[i] * x1 + y[i] * y1 + z[i] * z1 +
                (cable ? cable[i] : (FP) 0));
        FP s, c;
        sincos_v<FP>(-p, &s, &c);
        weights[i].x = c;
        weights[i].y = s;
    }
}

/* Philox4x32-10 (Salmon et al., SC 2011), the reference's Random123
 * generator (oskar/math/private_random_generators.h). */
__device__ __forceinline__ uint4 philox4x32_10(uint4 c, unsigned int k0,
        unsigned int k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r)
    {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        const unsigned int hi0 = __umulhi(0xD2511F53u, c.x);
        const unsigned int lo0 = 0xD2511F53u * c.x;
        const unsigned int hi1 = __umulhi(0xCD9E8D57u, c.z);
        const unsigned int lo1 = 0xCD9E8D57u * c.z;
        c = make_uint4(hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0);
    }
    return c;
}

/* Box-Muller of two 32-bit integers in the reference's CPU form
 * (oskar/math/private_random_helpers.h): angle pi (u0 / 2^31 + half step)
 * (u0 read as a 64-bit long), radius sqrt(-2 ln((u1 + 1/2) / 2^32)), in the
 * precision of FP. */
template <typename FP>
__device__ __forceinline__ void box_muller(unsigned int u0, unsigned int u1,
        FP* f0, FP* f1)
{
#pragma clang fp contract(off)
    const FP fa = (FP) 1 / ((FP) 1 + (FP) 0x7FFFFFFFu);
    const FP fb = (FP) 1 / ((FP) 1 + (FP) 0xFFFFFFFFu);
    FP t = (FP) 3.14159265358979323846;
    t *= ((FP) (long long) u0 * fa) + (FP) 0.5 * fa;
    FP s, c;
    sincos_v<FP>(t, &s, &c);
    const FP r = sqrt((FP) -2 * log(((FP) u1 * fb) + (FP) 0.5 * fb));
    *f0 = s * r;
    *f1 = c * r;
}

/* Element errors and apodisation on the DFT weights
 * (oskar_station_evaluate_element_weights.c:34-64):
 *  - error draws (r.x, r.y) for element i are values 2i and 2i+1 of
 *    oskar_mem_random_gaussian(seed, time_index, station_id, 0x12345678)
 *    (Philox counter i / 2, its outputs 0-1 or 2-3), or zero when no error
 *    scale is set;
 *  - amplitude r.x gain_error + gain, phase r.y phase_error + phase, factor
 *    amplitude (cos, sin)(phase) (define_evaluate_element_weights_errors.h
 *    :8-12);
 *  - weights *= factor, then *= apodisation, each a complex product
 *    (a.x b.x - a.y b.y, a.x b.y + a.y b.x) with every product and sum
 *    rounded. */
template <typename FP>
struct ErrArgs
{
    const FP *gain, *gain_error, *phase, *phase_error;
    const Cplx<FP>* apod;
    unsigned int seed, time_index, station_id;
};

template <typename FP>
__global__ void weights_scale_kernel(int n, Cplx<FP>* w, ErrArgs<FP> e)
{
#pragma clang fp contract(off)
    GRID_STRIDE(i, (size_t) n)
    {
        Cplx<FP> x = w[i];
        if (e.gain)
        {
            FP rx = (FP) 0, ry = (FP) 0;
            if (e.gain_error)
            {
                const uint4 u = philox4x32_10(make_uint4((unsigned int) i / 2,
                        e.time_index, e.station_id, 0x12345678u), e.seed,
                        0xCAFEF00Du);
                if (i % 2 == 0) box_muller<FP>(u.x, u.y, &rx, &ry);
                else box_muller<FP>(u.z, u.w, &rx, &ry);
                rx *= e.gain_error[i];
                ry *= e.phase_error[i];
            }
            rx += e.gain[i];
            ry += e.phase[i];
            FP s, c;
            sincos_v<FP>(ry, &s, &c);
            const FP ex = c * rx, ey = s * rx;
            const Cplx<FP> y = {x.x * ex - x.y * ey, x.x * ey + x.y * ex};
            x = y;
        }
        if (e.apod)
        {
            const Cplx<FP> a = e.apod[i];
            const Cplx<FP> y = {x.x * a.x - x.y * a.y, x.x * a.y + x.y * a.x};
            x = y;
        }
        w[i] = x;
    }
}

template <typename FP>
__global__ void theta_phi_kernel(int off_in, int num, const FP* x,
        const FP* y, const FP* z, int extra_point_at_pole, FP delta_phi1,
        FP delta_phi2, FP* theta, FP* phi1, FP* phi2)
{
    GRID_STRIDE(i, (size_t) num)
    {
        theta_phi_point<FP>(x[i + off_in], y[i + off_in], z[i + off_in],
                delta_phi1, delta_phi2, &theta[i], &phi1[i], &phi2[i]);
    }
    if (extra_point_at_pole && blockIdx.x == 0 && threadIdx.x == 0)
    {
        theta[num] = (FP) 0;
        phi1[num] = delta_phi1;
        phi2[num] = delta_phi2;
    }
}

template <typename FP, bool MATRIX>
__global__ void dipole_kernel(int n, const FP* theta, const FP* phi_x,
        const FP* phi_y, FP kL, FP cos_kL, int swap_xy, int offset_out,
        Cplx<FP>* pattern)
{
    GRID_STRIDE(i, (size_t) n)
    {
        dipole_point<FP, MATRIX>(theta[i], phi_x[i], phi_y[i], kL, cos_kL,
                swap_xy, pattern, i + offset_out);
    }
}

/* theta_phi_kernel and dipole_kernel in one pass, theta and phi kept in
 * registers: the pattern of directions [off_in, off_in + num), then, with
 * extra_point_at_pole, the pole (theta 0, phi = the deltas) at index num. */
template <typename FP, bool MATRIX>
__global__ void dipole_enu_kernel(int off_in, int num, const FP* x,
        const FP* y, const FP* z, int extra_point_at_pole, FP delta_phi1,
        FP delta_phi2, FP kL, FP cos_kL, int swap_xy, int offset_out,
        Cplx<FP>* pattern)
{
    const size_t total = (size_t) num + (extra_point_at_pole ? 1 : 0);
    GRID_STRIDE(i, total)
    {
        FP theta = (FP) 0, p1 = delta_phi1, p2 = delta_phi2;
        if (i < (size_t) num)
            theta_phi_point<FP>(x[i + off_in], y[i + off_in], z[i + off_in],
                    delta_phi1, delta_phi2, &theta, &p1, &p2);
        dipole_point<FP, MATRIX>(theta, p1, p2, kL, cos_kL, swap_xy,
                pattern, i + offset_out);
    }
}

template <typename FP, int NR>
__global__ void blank_kernel(int offset_mask, int n, const FP* mask,
        int offset_out, FP* jones)
{
    GRID_STRIDE(i, (size_t) n)
    {
        if (mask[i + offset_mask] < (FP) 0)
        {
            FP* o = jones + (size_t) (offset_out + i) * NR;
#pragma unroll
            for (int r = 0; r < NR; ++r) o[r] = (FP) 0;
        }
    }
}

/* NR reals per element: 1 real, 2 complex, 8 matrix. */
template <typename FP, int NR>
__global__ void normalise_kernel(unsigned int offset, unsigned int n,
        FP* a, unsigned int idx)
{
    const FP* v = a + (size_t) idx * NR;
    FP scal;
    if (NR == 1) scal = (FP) 1 / v[0];
    else
    {
        FP amp = (FP) 0;
#pragma unroll
        for (int r = 0; r < NR; ++r) amp += v[r] * v[r];
        if (NR == 8) amp /= (FP) 2;
        scal = (FP) 1 / sqrt(amp);
    }
    GRID_STRIDE(i, (size_t) n)
    {
        const size_t j = offset + i;
        if (j != idx)
        {
            FP* o = a + j * NR;
#pragma unroll
            for (int r = 0; r < NR; ++r) o[r] *= scal;
        }
    }
}

/* The last steps of a station beam, in one pass: the horizon blank of
 * blank_kernel (elements whose mask is negative become zeros) and
 * normalise_kernel's scaling by element idx (blanked alike), written to
 * out[0 .. n) (n <= idx).  Without NORM the blanked values are copied.
 * Elements move as NV vectors of type V (16 bytes when the pointers allow
 * it); the products are normalise_kernel's. */
template <typename FP, int NR, typename V, bool NORM, bool BLANK>
__global__ void beam_finish_kernel(unsigned int n, const FP* a,
        unsigned int idx, const FP* mask, FP* out)
{
    constexpr int NV = NR * (int) sizeof(FP) / (int) sizeof(V);
    static_assert(NV * sizeof(V) == NR * sizeof(FP), "vector size");
    FP scal = (FP) 1;
    if constexpr (NORM)
    {
        FP v[NR];
        const bool zero = BLANK && mask[idx] < (FP) 0;
#pragma unroll
        for (int r = 0; r < NR; ++r)
            v[r] = zero ? (FP) 0 : a[(size_t) idx * NR + r];
        if (NR == 1) scal = (FP) 1 / v[0];
        else
        {
            FP amp = (FP) 0;
#pragma unroll
            for (int r = 0; r < NR; ++r) amp += v[r] * v[r];
            if (NR == 8) amp /= (FP) 2;
            scal = (FP) 1 / sqrt(amp);
        }
    }
    GRID_STRIDE(i, (size_t) n)
    {
        V buf[NV];
        const V* src = (const V*) a + i * NV;
#pragma unroll
        for (int k = 0; k < NV; ++k) buf[k] = src[k];
        FP x[NR];
        __builtin_memcpy(x, buf, sizeof(x));
        const bool zero = BLANK && mask[i] < (FP) 0;
#pragma unroll
        for (int r = 0; r < NR; ++r)
        {
            FP y = zero ? (FP) 0 : x[r];
            if constexpr (NORM) y = y * scal;
            x[r] = y;
        }
        __builtin_memcpy(buf, x, sizeof(x));
        V* dst = (V*) out + i * NV;
#pragma unroll
        for (int k = 0; k < NV; ++k) dst[k] = buf[k];
    }
}

template <typename FP, int NR, typename V>
int beam_finish_t(unsigned int n, const void* data, unsigned int idx,
        int normalise, const void* mask, void* out, hipStream_t s)
{
    const FP* a = (const FP*) data;
    const FP* m = (const FP*) mask;
    FP* o = (FP*) out;
#define BF(NORM, BLANK) hipLaunchKernelGGL( \
        (beam_finish_kernel<FP, NR, V, NORM, BLANK>), grid_for(n), \
        dim3(THREADS), 0, s, n, a, idx, m, o)
    if (normalise)
    {
        if (mask) BF(true, true);
        else BF(true, false);
    }
    else
    {
        if (mask) BF(false, true);
        else BF(false, false);
    }
#undef BF
    return (int) hipGetLastError();
}

/* Three direction arrays of n values copied, and the point (px, py, pz)
 * written after them at index n when `extra` is set. */
template <typename FP>
__global__ void copy3_point_kernel(unsigned int n, const FP* x, const FP* y,
        const FP* z, FP* xo, FP* yo, FP* zo, int extra, double px, double py,
        double pz)
{
    GRID_STRIDE(i, (size_t) n)
    {
        xo[i] = x[i];
        yo[i] = y[i];
        zo[i] = z[i];
    }
    if (extra && blockIdx.x == 0 && threadIdx.x == 0)
    {
        xo[n] = (FP) px;
        yo[n] = (FP) py;
        zo[n] = (FP) pz;
    }
}

template <typename FP, int NR>
__global__ void set_value_kernel(size_t offset, size_t n, FP value, FP* a)
{
    GRID_STRIDE(i, n)
    {
        FP* o = a + (offset + i) * NR;
#pragma unroll
        for (int r = 0; r < NR; ++r) o[r] = (FP) 0;
        o[0] = value;
        if (NR == 8) o[6] = value;
    }
}

/* Complex product with every product and sum rounded on its own (no fused
 * multiply-add), so the join's result does not depend on how the compiler
 * schedules a kernel: both join kernels below give the same bits, which are
 * those of plain IEEE arithmetic in this term order. */
template <typename FP>
__device__ __forceinline__ Cplx<FP> cmul_rn(Cplx<FP> a, Cplx<FP> b)
{
#pragma clang fp contract(off)
    Cplx<FP> c;
    c.x = a.x * b.x - a.y * b.y;
    c.y = a.x * b.y + a.y * b.x;
    return c;
}

/* out = a * b per element; a scalar operand is promoted to a*I. */
template <typename FP, bool AM, bool BM>
__global__ void join_kernel(size_t n, Cplx<FP>* out, const Cplx<FP>* a,
        const Cplx<FP>* b)
{
    GRID_STRIDE(i, n)
    {
        if (!AM && !BM)
        {
            out[i] = cmul_rn(a[i], b[i]);
        }
        else if (!AM)
        {
            const Cplx<FP> s = a[i];
            Cplx<FP> r[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) r[k] = cmul_rn(s, b[4 * i + k]);
#pragma unroll
            for (int k = 0; k < 4; ++k) out[4 * i + k] = r[k];
        }
        else if (!BM)
        {
            const Cplx<FP> s = b[i];
            Cplx<FP> r[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) r[k] = cmul_rn(a[4 * i + k], s);
#pragma unroll
            for (int k = 0; k < 4; ++k) out[4 * i + k] = r[k];
        }
        else
        {
            const Cplx<FP>* A = a + 4 * i;
            const Cplx<FP>* B = b + 4 * i;
            Cplx<FP> r[4];
            /* [[a b],[c d]] x [[e f],[g h]] */
            Cplx<FP> t0 = cmul_rn(A[0], B[0]), t1 = cmul_rn(A[1], B[2]);
            r[0].x = t0.x + t1.x; r[0].y = t0.y + t1.y;
            t0 = cmul_rn(A[0], B[1]); t1 = cmul_rn(A[1], B[3]);
            r[1].x = t0.x + t1.x; r[1].y = t0.y + t1.y;
            t0 = cmul_rn(A[2], B[0]); t1 = cmul_rn(A[3], B[2]);
            r[2].x = t0.x + t1.x; r[2].y = t0.y + t1.y;
            t0 = cmul_rn(A[2], B[1]); t1 = cmul_rn(A[3], B[3]);
            r[3].x = t0.x + t1.x; r[3].y = t0.y + t1.y;
#pragma unroll
            for (int k = 0; k < 4; ++k) out[4 * i + k] = r[k];
        }
    }
}

/* The join on whole 2x2 matrices loaded and stored as aligned blocks
 * (two 16-byte accesses per FP32 matrix, four per FP64), one element per
 * thread and iteration; arithmetic as join_kernel, term for term.  Used
 * when every pointer is aligned to a matrix. */
template <typename FP> struct alignas(4 * sizeof(Cplx<FP>)) JonesMat
{
    Cplx<FP> v[4];
};

template <typename FP, bool AM, bool BM>
__global__ __launch_bounds__(THREADS)
void join_vec_kernel(size_t n, JonesMat<FP>* out, const void* a, const void* b)
{
    GRID_STRIDE(i, n)
    {
        JonesMat<FP> r;
        if (AM && BM)
        {
            const JonesMat<FP> A = ((const JonesMat<FP>*) a)[i];
            const JonesMat<FP> B = ((const JonesMat<FP>*) b)[i];
            Cplx<FP> t0 = cmul_rn(A.v[0], B.v[0]);
            Cplx<FP> t1 = cmul_rn(A.v[1], B.v[2]);
            r.v[0].x = t0.x + t1.x; r.v[0].y = t0.y + t1.y;
            t0 = cmul_rn(A.v[0], B.v[1]); t1 = cmul_rn(A.v[1], B.v[3]);
            r.v[1].x = t0.x + t1.x; r.v[1].y = t0.y + t1.y;
            t0 = cmul_rn(A.v[2], B.v[0]); t1 = cmul_rn(A.v[3], B.v[2]);
            r.v[2].x = t0.x + t1.x; r.v[2].y = t0.y + t1.y;
            t0 = cmul_rn(A.v[2], B.v[1]); t1 = cmul_rn(A.v[3], B.v[3]);
            r.v[3].x = t0.x + t1.x; r.v[3].y = t0.y + t1.y;
        }
        else if (AM)
        {
            const JonesMat<FP> A = ((const JonesMat<FP>*) a)[i];
            const Cplx<FP> sc = ((const Cplx<FP>*) b)[i];
#pragma unroll
            for (int k = 0; k < 4; ++k) r.v[k] = cmul_rn(A.v[k], sc);
        }
        else
        {
            const Cplx<FP> sc = ((const Cplx<FP>*) a)[i];
            const JonesMat<FP> B = ((const JonesMat<FP>*) b)[i];
#pragma unroll
            for (int k = 0; k < 4; ++k) r.v[k] = cmul_rn(sc, B.v[k]);
        }
        out[i] = r;
    }
}

} /* namespace */

extern "C" int oskar_amd_weights_scale_launch(int precision, int num,
        void* weights, const void* gain, const void* gain_error,
        const void* phase, const void* phase_error, const void* apod,
        unsigned int seed, int time_index, int station_id, void* stream)
{
    if (num <= 0) return 0;
    const hipStream_t s = (hipStream_t) stream;
    if (precision == OSKAR_DOUBLE)
    {
        const ErrArgs<double> e = {(const double*) gain,
                (const double*) gain_error, (const double*) phase,
                (const double*) phase_error, (const Cplx<double>*) apod, seed,
                (unsigned int) time_index, (unsigned int) station_id};
        hipLaunchKernelGGL(weights_scale_kernel<double>, grid_for(num),
                dim3(THREADS), 0, s, num, (Cplx<double>*) weights, e);
    }
    else
    {
        const ErrArgs<float> e = {(const float*) gain,
                (const float*) gain_error, (const float*) phase,
                (const float*) phase_error, (const Cplx<float>*) apod, seed,
                (unsigned int) time_index, (unsigned int) station_id};
        hipLaunchKernelGGL(weights_scale_kernel<float>, grid_for(num),
                dim3(THREADS), 0, s, num, (Cplx<float>*) weights, e);
    }
    return (int) hipGetLastError();
}

extern "C" int oskar_amd_element_weights_dft_launch(int precision,
        int num_elements, const void* x, const void* y, const void* z,
        const void* cable, double wavenumber, double x1, double y1, double z1,
        void* weights, void* stream)
{
    if (num_elements <= 0) return 0;
    const hipStream_t s = (hipStream_t) stream;
    if (precision == OSKAR_DOUBLE)
        hipLaunchKernelGGL(weights_kernel<double>, grid_for(num_elements),
                dim3(THREADS), 0, s, num_elements, (const double*) x,
                (const double*) y, (const double*) z, (const double*) cable,
                wavenumber, x1, y1, z1, (Cplx<double>*) weights);
    else
        hipLaunchKernelGGL(weights_kernel<float>, grid_for(num_elements),
                dim3(THREADS), 0, s, num_elements, (const float*) x,
                (const float*) y, (const float*) z, (const float*) cable,
                (float) wavenumber, (float) x1, (float) y1, (float) z1,
                (Cplx<float>*) weights);
    return (int) hipGetLastError();
}

extern "C" int oskar_amd_theta_phi_launch(int precision, int offset_in,
        int num, const void* x, const void* y, const void* z,
        int extra_point_at_pole, double delta_phi1, double delta_phi2,
        void* theta, void* phi1, void* phi2, void* stream)
{
    const hipStream_t s = (hipStream_t) stream;
    if (precision == OSKAR_DOUBLE)
        hipLaunchKernelGGL(theta_phi_kernel<double>, grid_for(num),
                dim3(THREADS), 0, s, offset_in, num, (const double*) x,
                (const double*) y, (const double*) z, extra_point_at_pole,
                delta_phi1, delta_phi2, (double*) theta, (double*) phi1,
                (double*) phi2);
    else
        hipLaunchKernelGGL(theta_phi_kernel<float>, grid_for(num),
                dim3(THREADS), 0, s, offset_in, num, (const float*) x,
                (const float*) y, (const float*) z, extra_point_at_pole,
                (float) delta_phi1, (float) delta_phi2, (float*) theta,
                (float*) phi1, (float*) phi2);
    return (int) hipGetLastError();
}

extern "C" int oskar_amd_dipole_enu_launch(int precision, int matrix,
        int offset_in, int num, const void* x, const void* y, const void* z,
        int extra_point_at_pole, double delta_phi1, double delta_phi2,
        double kL, double cos_kL, int swap_xy, int offset_out, void* pattern,
        void* stream)
{
    const int total = num + (extra_point_at_pole ? 1 : 0);
    if (total <= 0) return 0;
    const hipStream_t s = (hipStream_t) stream;
#define DE(FP, M) hipLaunchKernelGGL((dipole_enu_kernel<FP, M>), \
        grid_for(total), dim3(THREADS), 0, s, offset_in, num, (const FP*) x, \
        (const FP*) y, (const FP*) z, extra_point_at_pole, (FP) delta_phi1, \
        (FP) delta_phi2, (FP) kL, (FP) cos_kL, swap_xy, offset_out, \
        (Cplx<FP>*) pattern)
    if (precision == OSKAR_DOUBLE)
    {
        if (matrix) DE(double, true);
        else DE(double, false);
    }
    else
    {
        if (matrix) DE(float, true);
        else DE(float, false);
    }
#undef DE
    return (int) hipGetLastError();
}

extern "C" int oskar_amd_dipole_launch(int precision, int matrix,
        int num_points, const void* theta, const void* phi_x,
        const void* phi_y, double kL, double cos_kL, int swap_xy,
        int offset_out, void* pattern, void* stream)
{
    if (num_points <= 0) return 0;
    const hipStream_t s = (hipStream_t) stream;
    if (precision == OSKAR_DOUBLE)
    {
        if (matrix)
            hipLaunchKernelGGL((dipole_kernel<double, true>),
                    grid_for(num_points), dim3(THREADS), 0, s, num_points,
                    (const double*) theta, (const double*) phi_x,
                    (const double*) phi_y, kL, cos_kL, swap_xy, offset_out,
                    (Cplx<double>*) pattern);
        else
            hipLaunchKernelGGL((dipole_kernel<double, false>),
                    grid_for(num_points), dim3(THREADS), 0, s, num_points,
                    (const double*) theta, (const double*) phi_x,
                    (const double*) phi_y, kL, cos_kL, swap_xy, offset_out,
                    (Cplx<double>*) pattern);
    }
    else
    {
        if (matrix)
            hipLaunchKernelGGL((dipole_kernel<float, true>),
                    grid_for(num_points), dim3(THREADS), 0, s, num_points,
                    (const float*) theta, (const float*) phi_x,
                    (const float*) phi_y, (float) kL, (float) cos_kL, swap_xy,
                    offset_out, (Cplx<float>*) pattern);
        else
            hipLaunchKernelGGL((dipole_kernel<float, false>),
                    grid_for(num_points), dim3(THREADS), 0, s, num_points,
                    (const float*) theta, (const float*) phi_x,
                    (const float*) phi_y, (float) kL, (float) cos_kL, swap_xy,
                    offset_out, (Cplx<float>*) pattern);
    }
    return (int) hipGetLastError();
}

/* One point (x, y, z) written at index i of three device arrays, in stream
 * order (a host copy would wait for the device). */
template <typename FP>
__global__ void set_point3_kernel(FP* x, FP* y, FP* z, size_t i, double vx,
        double vy, double vz)
{
    if (threadIdx.x == 0)
    {
        x[i] = (FP) vx;
        y[i] = (FP) vy;
        z[i] = (FP) vz;
    }
}

extern "C" int oskar_amd_set_point3_launch(int precision, void* x, void* y,
        void* z, size_t index, double vx, double vy, double vz, void* stream)
{
    const hipStream_t s = (hipStream_t) stream;
    if (precision == OSKAR_DOUBLE)
        hipLaunchKernelGGL(set_point3_kernel<double>, dim3(1), dim3(64), 0, s,
                (double*) x, (double*) y, (double*) z, index, vx, vy, vz);
    else
        hipLaunchKernelGGL(set_point3_kernel<float>, dim3(1), dim3(64), 0, s,
                (float*) x, (float*) y, (float*) z, index, vx, vy, vz);
    return (int) hipGetLastError();
}

extern "C" int oskar_amd_blank_below_horizon_launch(int precision,
        int matrix, int offset_mask, int num, const void* mask,
        int offset_out, void* jones, void* stream)
{
    if (num <= 0) return 0;
    const hipStream_t s = (hipStream_t) stream;
    if (precision == OSKAR_DOUBLE)
    {
        if (matrix)
            hipLaunchKernelGGL((blank_kernel<double, 8>), grid_for(num),
                    dim3(THREADS), 0, s, offset_mask, num,
                    (const double*) mask, offset_out, (double*) jones);
        else
            hipLaunchKernelGGL((blank_kernel<double, 2>), grid_for(num),
                    dim3(THREADS), 0, s, offset_mask, num,
                    (const double*) mask, offset_out, (double*) jones);
    }
    else
    {
        if (matrix)
            hipLaunchKernelGGL((blank_kernel<float, 8>), grid_for(num),
                    dim3(THREADS), 0, s, offset_mask, num,
                    (const float*) mask, offset_out, (float*) jones);
        else
            hipLaunchKernelGGL((blank_kernel<float, 2>), grid_for(num),
                    dim3(THREADS), 0, s, offset_mask, num,
                    (const float*) mask, offset_out, (float*) jones);
    }
    return (int) hipGetLastError();
}

extern "C" int oskar_amd_mem_normalise_launch(int type, unsigned int offset,
        unsigned int n, void* data, unsigned int idx, void* stream)
{
    if (n == 0) return 0;
    const hipStream_t s = (hipStream_t) stream;
    switch (type)
    {
    case OSKAR_SINGLE:
        hipLaunchKernelGGL((normalise_kernel<float, 1>), grid_for(n),
                dim3(THREADS), 0, s, offset, n, (float*) data, idx); break;
    case OSKAR_DOUBLE:
        hipLaunchKernelGGL((normalise_kernel<double, 1>), grid_for(n),
                dim3(THREADS), 0, s, offset, n, (double*) data, idx); break;
    case OSKAR_SINGLE_COMPLEX:
        hipLaunchKernelGGL((normalise_kernel<float, 2>), grid_for(n),
                dim3(THREADS), 0, s, offset, n, (float*) data, idx); break;
    case OSKAR_DOUBLE_COMPLEX:
        hipLaunchKernelGGL((normalise_kernel<double, 2>), grid_for(n),
                dim3(THREADS), 0, s, offset, n, (double*) data, idx); break;
    case OSKAR_SINGLE_COMPLEX_MATRIX:
        hipLaunchKernelGGL((normalise_kernel<float, 8>), grid_for(n),
                dim3(THREADS), 0, s, offset, n, (float*) data, idx); break;
    case OSKAR_DOUBLE_COMPLEX_MATRIX:
        hipLaunchKernelGGL((normalise_kernel<double, 8>), grid_for(n),
                dim3(THREADS), 0, s, offset, n, (double*) data, idx); break;
    default:
        return OSKAR_ERR_BAD_DATA_TYPE;
    }
    return (int) hipGetLastError();
}

extern "C" int oskar_amd_beam_finish_launch(int type, unsigned int n,
        const void* data, unsigned int idx, int normalise, const void* mask,
        void* out, void* stream)
{
    if (n == 0) return 0;
    const hipStream_t s = (hipStream_t) stream;
    /* 16-byte vectors when both arrays start on 16-byte boundaries (every
     * element then does, for 16- and 32-byte elements). */
    const bool v16 = (((uintptr_t) data | (uintptr_t) out) & 15u) == 0;
    switch (type)
    {
    case OSKAR_SINGLE_COMPLEX:
        return beam_finish_t<float, 2, float2>(n, data, idx, normalise, mask,
                out, s);
    case OSKAR_DOUBLE_COMPLEX:
        return v16 ? beam_finish_t<double, 2, double2>(n, data, idx,
                normalise, mask, out, s) :
                beam_finish_t<double, 2, double>(n, data, idx, normalise,
                mask, out, s);
    case OSKAR_SINGLE_COMPLEX_MATRIX:
        return v16 ? beam_finish_t<float, 8, float4>(n, data, idx,
                normalise, mask, out, s) :
                beam_finish_t<float, 8, float2>(n, data, idx, normalise,
                mask, out, s);
    case OSKAR_DOUBLE_COMPLEX_MATRIX:
        return v16 ? beam_finish_t<double, 8, double2>(n, data, idx,
                normalise, mask, out, s) :
                beam_finish_t<double, 8, double>(n, data, idx, normalise,
                mask, out, s);
    default:
        return OSKAR_ERR_BAD_DATA_TYPE;
    }
}

extern "C" int oskar_amd_copy3_point_launch(int precision, unsigned int n,
        const void* x, const void* y, const void* z, void* xo, void* yo,
        void* zo, int extra, double px, double py, double pz, void* stream)
{
    const hipStream_t s = (hipStream_t) stream;
    const dim3 g = grid_for(n > 0 ? n : 1);
    if (precision == OSKAR_DOUBLE)
        hipLaunchKernelGGL(copy3_point_kernel<double>, g, dim3(THREADS), 0, s,
                n, (const double*) x, (const double*) y, (const double*) z,
                (double*) xo, (double*) yo, (double*) zo, extra, px, py, pz);
    else
        hipLaunchKernelGGL(copy3_point_kernel<float>, g, dim3(THREADS), 0, s,
                n, (const float*) x, (const float*) y, (const float*) z,
                (float*) xo, (float*) yo, (float*) zo, extra, px, py, pz);
    return (int) hipGetLastError();
}

extern "C" int oskar_amd_mem_set_value_real_launch(int type, size_t offset,
        size_t n, double value, void* data, void* stream)
{
    if (n == 0) return 0;
    const hipStream_t s = (hipStream_t) stream;
    switch (type)
    {
    case OSKAR_SINGLE:
        hipLaunchKernelGGL((set_value_kernel<float, 1>), grid_for(n),
                dim3(THREADS), 0, s, offset, n, (float) value, (float*) data);
        break;
    case OSKAR_DOUBLE:
        hipLaunchKernelGGL((set_value_kernel<double, 1>), grid_for(n),
                dim3(THREADS), 0, s, offset, n, value, (double*) data); break;
    case OSKAR_SINGLE_COMPLEX:
        hipLaunchKernelGGL((set_value_kernel<float, 2>), grid_for(n),
                dim3(THREADS), 0, s, offset, n, (float) value, (float*) data);
        break;
    case OSKAR_DOUBLE_COMPLEX:
        hipLaunchKernelGGL((set_value_kernel<double, 2>), grid_for(n),
                dim3(THREADS), 0, s, offset, n, value, (double*) data); break;
    case OSKAR_SINGLE_COMPLEX_MATRIX:
        hipLaunchKernelGGL((set_value_kernel<float, 8>), grid_for(n),
                dim3(THREADS), 0, s, offset, n, (float) value, (float*) data);
        break;
    case OSKAR_DOUBLE_COMPLEX_MATRIX:
        hipLaunchKernelGGL((set_value_kernel<double, 8>), grid_for(n),
                dim3(THREADS), 0, s, offset, n, value, (double*) data); break;
    default:
        return OSKAR_ERR_BAD_DATA_TYPE;
    }
    return (int) hipGetLastError();
}

extern "C" int oskar_amd_jones_join_launch(int precision, int a_matrix,
        int b_matrix, size_t n, void* out, const void* a, const void* b,
        void* stream)
{
    if (n == 0) return 0;
    const hipStream_t s = (hipStream_t) stream;
    const size_t mat = (precision == OSKAR_DOUBLE ? 16 : 8) * 4;
    const bool aligned = ((size_t) out % mat == 0) &&
            (!a_matrix || (size_t) a % mat == 0) &&
            (!b_matrix || (size_t) b % mat == 0);
#define JOIN(FP, AM, BM)                                                     \
    if ((AM || BM) && aligned)                                               \
        hipLaunchKernelGGL((join_vec_kernel<FP, AM, BM>), grid_for(n),       \
                dim3(THREADS), 0, s, n, (JonesMat<FP>*) out, a, b);              \
    else                                                                     \
        hipLaunchKernelGGL((join_kernel<FP, AM, BM>), grid_for(n),           \
                dim3(THREADS), 0, s, n, (Cplx<FP>*) out,                     \
                (const Cplx<FP>*) a, (const Cplx<FP>*) b)
    if (precision == OSKAR_DOUBLE)
    {
        if (a_matrix && b_matrix) JOIN(double, true, true);
        else if (a_matrix) JOIN(double, true, false);
        else if (b_matrix) JOIN(double, false, true);
        else JOIN(double, false, false);
    }
    else
    {
        if (a_matrix && b_matrix) JOIN(float, true, true);
        else if (a_matrix) JOIN(float, true, false);
        else if (b_matrix) JOIN(float, false, true);
        else JOIN(float, false, false);
    }
#undef JOIN
    return (int) hipGetLastError();
}

namespace {

/* K[a][s] (define_evaluate_jones_K.h:56-75, CPU evaluation order).  A
 * thread holds one source's direction and filter value and walks a group of
 * K_GROUP stations (wave-uniform station values), so the source inputs are
 * read once per group and each store of a wave covers 64 consecutive
 * sources of one station row. */
constexpr int K_GROUP = 16;

template <typename FP>
__global__ __launch_bounds__(256)
void jones_K_kernel(int num_sources, const FP* l, const FP* m, const FP* n,
        int num_stations, const FP* u, const FP* v, const FP* w,
        FP wavenumber, const FP* filter, FP fmin, FP fmax, int ignore_w,
        Cplx<FP>* jones)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    const int a0 = blockIdx.y * K_GROUP;
    if (s >= num_sources) return;
    const FP f = filter[s];
    const bool keep = f > fmin && f <= fmax;
    const FP ls = l[s], ms = m[s], ns1 = n[s] - (FP) 1;
    const int a1 = min(a0 + K_GROUP, num_stations);
    for (int a = a0; a < a1; ++a)
    {
        Cplx<FP> k;
        k.x = k.y = (FP) 0;
        if (keep)
        {
            FP phase = u[a] * ls + v[a] * ms;
            if (!ignore_w) phase += w[a] * ns1;
            phase *= wavenumber;
            sincos_v<FP>(phase, &k.y, &k.x);
        }
        jones[(size_t) s + (size_t) num_sources * a] = k;
    }
}

} /* namespace */

extern "C" int oskar_amd_jones_K_launch(int precision, int num_sources,
        const void* l, const void* m, const void* n, int num_stations,
        const void* u, const void* v, const void* w, double wavenumber,
        const void* filter, double fmin, double fmax, int ignore_w,
        void* jones, void* stream)
{
    if (num_sources <= 0 || num_stations <= 0) return 0;
    const dim3 grid((num_sources + 255) / 256,
            (num_stations + K_GROUP - 1) / K_GROUP);
    const hipStream_t s = (hipStream_t) stream;
    if (precision == OSKAR_DOUBLE)
        hipLaunchKernelGGL(jones_K_kernel<double>, grid, dim3(256), 0, s,
                num_sources, (const double*) l, (const double*) m,
                (const double*) n, num_stations, (const double*) u,
                (const double*) v, (const double*) w, wavenumber,
                (const double*) filter, fmin, fmax, ignore_w,
                (Cplx<double>*) jones);
    else
        hipLaunchKernelGGL(jones_K_kernel<float>, grid, dim3(256), 0, s,
                num_sources, (const float*) l, (const float*) m,
                (const float*) n, num_stations, (const float*) u,
                (const float*) v, (const float*) w, (float) wavenumber,
                (const float*) filter, (float) fmin, (float) fmax, ignore_w,
                (Cplx<float>*) jones);
    return (int) hipGetLastError();
}

/* ---- Direction conversions to ENU direction cosines -------------------- */
/* oskar/convert/define_convert_apparent_ra_dec_to_enu_directions.h,
 * define_convert_az_el_to_enu_directions.h and
 * define_convert_relative_directions_to_enu_directions.h (the rotation of
 * relative direction cosines about the reference point into the horizon
 * frame).  kind: OSKAR_COORDS_RADEC (a = RA, b = Dec; p = lst, sin lat,
 * cos lat), OSKAR_COORDS_AZEL (a = az, b = el), OSKAR_COORDS_REL_DIR
 * (a, b, c = l, m, n; p = sin ha0, cos ha0, sin dec0, cos dec0, sin lat,
 * cos lat). */
namespace {

struct EnuParams { double p[6]; };

template <typename FP>
__global__ __launch_bounds__(THREADS)
void enu_convert_kernel(int kind, int num, const FP* a, const FP* b,
        const FP* c, EnuParams P, FP* x, FP* y, FP* z)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= num) return;
    if (kind == OSKAR_COORDS_RADEC)
    {
        FP sin_ha, cos_ha, sin_dec, cos_dec;
        sincos_v<FP>((FP) P.p[0] - a[i], &sin_ha, &cos_ha);
        sincos_v<FP>(b[i], &sin_dec, &cos_dec);
        const FP sin_lat = (FP) P.p[1], cos_lat = (FP) P.p[2];
        x[i] = -cos_dec * sin_ha;
        y[i] = cos_lat * sin_dec - sin_lat * cos_dec * cos_ha;
        z[i] = sin_lat * sin_dec + cos_lat * cos_dec * cos_ha;
    }
    else if (kind == OSKAR_COORDS_AZEL)
    {
        FP sin_az, cos_az, sin_el, cos_el;
        sincos_v<FP>(a[i], &sin_az, &cos_az);
        sincos_v<FP>(b[i], &sin_el, &cos_el);
        x[i] = cos_el * sin_az;
        y[i] = cos_el * cos_az;
        z[i] = sin_el;
    }
    else
    {
        const FP l = a[i], m = b[i], n = c[i];
        const FP sin_ha0 = (FP) P.p[0], cos_ha0 = (FP) P.p[1];
        const FP sin_dec0 = (FP) P.p[2], cos_dec0 = (FP) P.p[3];
        const FP sin_lat = (FP) P.p[4], cos_lat = (FP) P.p[5];
        x[i] = l * cos_ha0 + m * sin_ha0 * sin_dec0 - n * sin_ha0 * cos_dec0;
        FP t = sin_lat * cos_ha0;
        y[i] = -l * sin_lat * sin_ha0 + m * (cos_lat * cos_dec0 + t * sin_dec0)
                + n * (cos_lat * sin_dec0 - t * cos_dec0);
        t = cos_lat * cos_ha0;
        z[i] = l * cos_lat * sin_ha0 + m * (sin_lat * cos_dec0 - t * sin_dec0)
                + n * (sin_lat * sin_dec0 + t * cos_dec0);
    }
}

} /* namespace */

extern "C" int oskar_amd_enu_convert_launch(int precision, int kind, int num,
        const void* a, const void* b, const void* c, const double* params,
        void* x, void* y, void* z, void* stream)
{
    if (num <= 0) return 0;
    const hipStream_t s = (hipStream_t) stream;
    EnuParams P;
    for (int k = 0; k < 6; ++k) P.p[k] = params[k];
    if (precision == OSKAR_DOUBLE)
        hipLaunchKernelGGL(enu_convert_kernel<double>, grid_for(num),
                dim3(THREADS), 0, s, kind, num, (const double*) a,
                (const double*) b, (const double*) c, P, (double*) x,
                (double*) y, (double*) z);
    else
        hipLaunchKernelGGL(enu_convert_kernel<float>, grid_for(num),
                dim3(THREADS), 0, s, kind, num, (const float*) a,
                (const float*) b, (const float*) c, P, (float*) x,
                (float*) y, (float*) z);
    return (int) hipGetLastError();
}

namespace {

/* Station (u, v, w) to baseline (u, v, w)
 * (oskar/convert/define_convert_station_uvw_to_baseline_uvw.h): workgroup
 * (p, t) takes the baselines (p, q > p) of time t, uu[b] = sign (u[q] -
 * u[p]), sign -1 for the CASA convention.  Time t reads stations from
 * offset_in + t num and writes baselines from offset_out + t num (num-1)/2. */
template <typename FP>
__global__ __launch_bounds__(THREADS)
void station_to_baseline_kernel(int casa, int num, int offset_in,
        const FP* __restrict__ u, const FP* __restrict__ v,
        const FP* __restrict__ w, int offset_out, FP* __restrict__ uu,
        FP* __restrict__ vv, FP* __restrict__ ww)
{
    const int p = (int) blockIdx.x, t = (int) blockIdx.y;
    const FP sign = casa ? (FP) -1 : (FP) 1;
    const long long in0 = offset_in + (long long) t * num;
    const FP up = u[in0 + p], vp = v[in0 + p], wp = w[in0 + p];
    const long long b0 = offset_out + (long long) t * num * (num - 1) / 2 +
            (long long) p * (num - 1) - (long long) (p - 1) * p / 2 - p - 1;
    for (int q = p + 1 + (int) threadIdx.x; q < num; q += THREADS)
    {
        uu[b0 + q] = sign * (u[in0 + q] - up);
        vv[b0 + q] = sign * (v[in0 + q] - vp);
        ww[b0 + q] = sign * (w[in0 + q] - wp);
    }
}

} /* namespace */

extern "C" int oskar_amd_station_to_baseline_launch(int precision, int casa,
        int num_times, int num_stations, int offset_in, const void* u,
        const void* v, const void* w, int offset_out, void* uu, void* vv,
        void* ww, void* stream)
{
    if (num_stations < 2 || num_times < 1) return 0;
    const hipStream_t s = (hipStream_t) stream;
    const dim3 grid(num_stations - 1, num_times);
    if (precision == OSKAR_DOUBLE)
        hipLaunchKernelGGL(station_to_baseline_kernel<double>, grid,
                dim3(THREADS), 0, s, casa, num_stations, offset_in,
                (const double*) u, (const double*) v, (const double*) w,
                offset_out, (double*) uu, (double*) vv, (double*) ww);
    else
        hipLaunchKernelGGL(station_to_baseline_kernel<float>, grid,
                dim3(THREADS), 0, s, casa, num_stations, offset_in,
                (const float*) u, (const float*) v, (const float*) w,
                offset_out, (float*) uu, (float*) vv, (float*) ww);
    return (int) hipGetLastError();
}
