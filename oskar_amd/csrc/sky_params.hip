/*
 * Per-source parameters of the sky model that the correlator reads: the
 * relative direction cosines (l, m, n) of each source about the phase
 * centre and the Gaussian a, b, c of extended sources.
 *
 * Reference:
 *  - oskar/sky/src/oskar_sky_evaluate_relative_directions.c:20-37 with the
 *    3D body of oskar/convert/define_convert_lon_lat_to_relative_directions.h;
 *  - oskar/sky/src/oskar_sky_evaluate_gaussian_source_parameters.c:32-262,
 *    which calls oskar/math/src/oskar_fit_ellipse.c:17-131 (normal
 *    equations through oskar_matrix_multiply.c, LU through
 *    oskar_lapack_subset.c getrf2/getrs), oskar_rotate.c:31-48 and the
 *    xyz/lon-lat/tangent-plane conversions.
 *
 * The reference runs the Gaussian fit serially on the host (CPU location
 * only).  Here one thread owns one source and does the whole chain in
 * registers: six ellipse points projected through the sphere, the 5x5
 * normal equations, LU with partial pivoting (row swaps as selects over a
 * fully unrolled loop, so nothing is dynamically indexed and nothing goes
 * to scratch memory), the solve, the axis extraction and a, b, c.  It reads
 * 5 values and writes 3 per source; the work is a few dozen transcendental
 * calls, so the kernel is bound by VALU issue (FP64 for double skies).
 *
 * Arithmetic follows the reference's rounding: every product and sum is
 * rounded on its own (no fused multiply-add), the LU applies its rank-1
 * updates in the reference's column order, and in the float instance each
 * expression is evaluated in the precision that the reference's C gives it
 * (double libm calls on promoted values in the sky function and the fit,
 * float libm calls in the per-precision helpers and in the conversion
 * macros, which the reference compiles with <tgmath.h>).  The cosines and
 * sines of the six ellipse angles come from the host's libm, as in the
 * reference.  The remaining difference is the device libm (sin, cos, asin,
 * atan2 within a few ulp of the host's).
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "private.h"

namespace {

constexpr int THREADS = 256;
constexpr double PI = 3.14159265358979323846;
constexpr double PI_2_2_LN_2 = 7.11941466249375271693034; /* pi^2/(2 ln 2) */

/* libm functions of the precision (sinf etc. for float). */
__device__ __forceinline__ float n_sin(float x) { return sinf(x); }
__device__ __forceinline__ float n_cos(float x) { return cosf(x); }
__device__ __forceinline__ float n_asin(float x) { return asinf(x); }
__device__ __forceinline__ float n_atan2(float y, float x)
{
    return atan2f(y, x);
}
__device__ __forceinline__ float n_sqrt(float x) { return sqrtf(x); }
__device__ __forceinline__ double n_sin(double x) { return sin(x); }
__device__ __forceinline__ double n_cos(double x) { return cos(x); }
__device__ __forceinline__ double n_asin(double x) { return asin(x); }
__device__ __forceinline__ double n_atan2(double y, double x)
{
    return atan2(y, x);
}
__device__ __forceinline__ double n_sqrt(double x) { return sqrt(x); }

template <typename FP> struct Limits;
template <> struct Limits<float>
{
    static constexpr float min_normal = 1.17549435e-38f;
};
template <> struct Limits<double>
{
    static constexpr double min_normal = 2.2250738585072014e-308;
};

/* Relative direction cosines (3D), one source per thread. */
template <typename FP>
__global__ void __launch_bounds__(THREADS)
rel_dir_kernel(int num, const FP* __restrict__ lon, const FP* __restrict__ lat,
        FP lon0, FP cos_lat0, FP sin_lat0, FP* __restrict__ l,
        FP* __restrict__ m, FP* __restrict__ n)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * THREADS + threadIdx.x;
    if (i >= num) return;
    const FP dl = lon[i] - lon0, la = lat[i];
    const FP sin_lon = n_sin(dl), cos_lon = n_cos(dl);
    const FP sin_lat = n_sin(la), cos_lat = n_cos(la);
    l[i] = cos_lat * sin_lon;
    m[i] = cos_lat0 * sin_lat - sin_lat0 * cos_lat * cos_lon;
    n[i] = sin_lat0 * sin_lat + cos_lat0 * cos_lat * cos_lon;
}

/* Least-squares ellipse through six points (oskar_fit_ellipse.c:17-131).
 * Returns false where the fit fails (zero pivot, or not an ellipse). */
template <typename FP>
__device__ __forceinline__ bool fit_ellipse(const FP (&x)[6], const FP (&y)[6],
        FP& major, FP& minor, FP& pa_out)
{
#pragma clang fp contract(off)
    bool ok = true;
    FP mean_x = 0, mean_y = 0;
#pragma unroll
    for (int i = 0; i < 6; ++i) { mean_x += x[i]; mean_y += y[i]; }
    mean_x /= (FP) 6;
    mean_y /= (FP) 6;

    /* Rows t = (x^2, xy, y^2, x, y) of the mean-removed points, and the sum
     * of each column (the right-hand side X^T 1). */
    FP X[6][5], s[5] = {0, 0, 0, 0, 0};
#pragma unroll
    for (int i = 0; i < 6; ++i)
    {
        const FP x_ = x[i] - mean_x, y_ = y[i] - mean_y;
        X[i][0] = x_ * x_;
        X[i][1] = x_ * y_;
        X[i][2] = y_ * y_;
        X[i][3] = x_;
        X[i][4] = y_;
#pragma unroll
        for (int j = 0; j < 5; ++j) s[j] += X[i][j];
    }

    /* A(i, j) = sum_k X[k][j] X[k][i], skipping zero X[k][j]
     * (oskar_matrix_multiply.c, transpose_a branch). */
    FP A[5][5];
#pragma unroll
    for (int j = 0; j < 5; ++j)
#pragma unroll
        for (int i = 0; i < 5; ++i)
        {
            FP acc = 0;
#pragma unroll
            for (int k = 0; k < 6; ++k)
                if (X[k][j] != (FP) 0) acc += X[k][j] * X[k][i];
            A[i][j] = acc;
        }

    /* LU with partial pivoting, first largest magnitude wins (getrf2). */
    int piv[5];
#pragma unroll
    for (int k = 0; k < 5; ++k)
    {
        int p = k;
        FP mx = fabs(A[k][k]);
#pragma unroll
        for (int i = k + 1; i < 5; ++i)
        {
            const FP v = fabs(A[i][k]);
            if (v > mx) { mx = v; p = i; }
        }
        piv[k] = p;
        FP pv = A[k][k];
#pragma unroll
        for (int i = k + 1; i < 5; ++i) if (p == i) pv = A[i][k];
        if (pv != (FP) 0)
        {
#pragma unroll
            for (int i = k + 1; i < 5; ++i)
            {
                if (p == i)
                {
#pragma unroll
                    for (int j = 0; j < 5; ++j)
                    {
                        const FP t = A[k][j];
                        A[k][j] = A[i][j];
                        A[i][j] = t;
                    }
                }
            }
            if (fabs(A[k][k]) >= Limits<FP>::min_normal)
            {
                const FP r = (FP) 1 / A[k][k];
#pragma unroll
                for (int i = k + 1; i < 5; ++i) A[i][k] = r * A[i][k];
            }
            else
            {
#pragma unroll
                for (int i = k + 1; i < 5; ++i) A[i][k] /= A[k][k];
            }
        }
        else ok = false;
#pragma unroll
        for (int j = k + 1; j < 5; ++j)
        {
            if (A[k][j] != (FP) 0)
            {
#pragma unroll
                for (int i = k + 1; i < 5; ++i) A[i][j] -= A[k][j] * A[i][k];
            }
        }
    }

    /* getrs: row interchanges, unit-lower solve, upper solve. */
#pragma unroll
    for (int k = 0; k < 5; ++k)
    {
#pragma unroll
        for (int i = k + 1; i < 5; ++i)
        {
            if (piv[k] == i)
            {
                const FP t = s[k];
                s[k] = s[i];
                s[i] = t;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 5; ++k)
    {
        if (s[k] != (FP) 0)
        {
#pragma unroll
            for (int i = k + 1; i < 5; ++i) s[i] -= s[k] * A[i][k];
        }
    }
#pragma unroll
    for (int k = 4; k >= 0; --k)
    {
        if (s[k] != (FP) 0)
        {
            s[k] /= A[k][k];
#pragma unroll
            for (int i = 0; i < k; ++i) s[i] -= s[k] * A[i][k];
        }
    }

    FP a = s[0], b = s[1], c = s[2], d = s[3], e = s[4];
    const double r1 = fabs((double) (b / a)), r2 = fabs((double) (b / c));
    if ((r1 < r2 ? r1 : r2) > 1e-3)
    {
        /* Remove the orientation. */
        pa_out = (FP) (0.5 * atan2((double) b, (double) (c - a)) - PI / 2.0);
        const FP cos_phi = (FP) cos((double) pa_out);
        const FP sin_phi = (FP) sin((double) pa_out);
        const FP a2 = a * cos_phi * cos_phi - b * cos_phi * sin_phi +
                c * sin_phi * sin_phi;
        const FP c2 = a * sin_phi * sin_phi + b * cos_phi * sin_phi +
                c * cos_phi * cos_phi;
        const FP d2 = d * cos_phi - e * sin_phi;
        const FP e2 = d * sin_phi + e * cos_phi;
        a = a2; c = c2; d = d2; e = e2;
    }
    else pa_out = (FP) (((double) (c - a) < 0.0) ? 0.0 : PI / 2.0);
    if ((double) (a * c) <= 0.0) return false;
    if ((double) a < 0.0) { a = -a; c = -c; d = -d; e = -e; }
    const FP F = (FP) (1.0 + (double) (d * d) / (4.0 * (double) a) +
            (double) (e * e) / (4.0 * (double) c));
    a = (FP) sqrt((double) (F / a));
    b = (FP) sqrt((double) (F / c));
    major = (FP) (2.0 * (double) (a > b ? a : b));
    minor = (FP) (2.0 * (double) (a < b ? a : b));
    return ok;
}

template <typename FP> struct GaussArgs
{
    int num;
    const FP* ra;
    const FP* dec;
    const FP* maj;
    const FP* min;
    const FP* pa;
    FP* a;
    FP* b;
    FP* c;
    FP lon0, cos_lat0, sin_lat0;
    double cos_t[6], sin_t[6];     /* of the six ellipse angles (host libm) */
    int* num_failed;
};

template <typename FP>
__global__ void __launch_bounds__(THREADS)
gaussian_params_kernel(const GaussArgs<FP> g)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * THREADS + threadIdx.x;
    if (i >= g.num) return;
    const FP maj_i = g.maj[i], min_i = g.min[i];
    if ((double) maj_i == 0.0 && (double) min_i == 0.0) return;

    /* Ellipse on the tangent plane at (0, 0). */
    const FP ea = (FP) ((double) maj_i / 2.0), eb = (FP) ((double) min_i / 2.0);
    const double pa_i = (double) g.pa[i];
    const FP cos_pa = (FP) cos(pa_i), sin_pa = (FP) sin(pa_i);
    FP x[6], y[6], z[6], lon[6], lat[6];
#pragma unroll
    for (int j = 0; j < 6; ++j)
    {
        const FP l0 = (FP) ((double) ea * g.cos_t[j] * (double) sin_pa +
                (double) eb * g.sin_t[j] * (double) cos_pa);
        const FP m0 = (FP) ((double) ea * g.cos_t[j] * (double) cos_pa -
                (double) eb * g.sin_t[j] * (double) sin_pa);
        /* Tangent plane at (0, 0) -> lon, lat (2D body, lon0 = 0,
         * cos_lat0 = 1, sin_lat0 = 0). */
        const FP n0 = n_sqrt((FP) 1 - l0 * l0 - m0 * m0);
        lat[j] = n_asin(n0 * (FP) 0 + m0 * (FP) 1);
        lon[j] = (FP) 0 + n_atan2(l0, (FP) 1 * n0 - m0 * (FP) 0);
        /* lon, lat -> xyz. */
        const FP cl = n_cos(lon[j]), sl = n_sin(lon[j]);
        const FP ct = n_cos(lat[j]), st = n_sin(lat[j]);
        x[j] = ct * cl;
        y[j] = ct * sl;
        z[j] = st;
    }

    /* Rotate to the source position, back to lon, lat, then to the tangent
     * plane at the phase centre. */
    const FP ra_i = g.ra[i], dec_i = g.dec[i];
    const FP cos_lon = n_cos(ra_i), sin_lon = n_sin(ra_i);
    const FP cos_lat = n_cos(dec_i), sin_lat = n_sin(dec_i);
    FP l[6], m[6];
#pragma unroll
    for (int j = 0; j < 6; ++j)
    {
        const FP x_ = x[j], y_ = y[j], z_ = z[j];
        const FP xr = x_ * cos_lon * cos_lat - y_ * sin_lon -
                z_ * cos_lon * sin_lat;
        const FP yr = x_ * cos_lat * sin_lon + y_ * cos_lon -
                z_ * sin_lon * sin_lat;
        const FP zr = x_ * sin_lat + z_ * cos_lat;
        const FP lo = n_atan2(yr, xr);
        const FP la = n_atan2(zr, n_sqrt(xr * xr + yr * yr));
        const FP dl = lo - g.lon0;
        const FP s_lon = n_sin(dl), c_lon = n_cos(dl);
        const FP s_lat = n_sin(la), c_lat = n_cos(la);
        l[j] = c_lat * s_lon;
        m[j] = g.cos_lat0 * s_lat - g.sin_lat0 * c_lat * c_lon;
    }

    FP maj, min, pa;
    if (!fit_ellipse<FP>(l, m, maj, min, pa))
    {
        atomicAdd(g.num_failed, 1);
        return;
    }

    /* Gaussian parameters in the uv plane. */
    const FP inv_std_maj_2 = (FP) (0.5 * (double) (maj * maj) * PI_2_2_LN_2);
    const FP inv_std_min_2 = (FP) (0.5 * (double) (min * min) * PI_2_2_LN_2);
    const double cp = cos((double) pa), sp = sin((double) pa);
    const FP cos_pa_2 = (FP) (cp * cp);
    const FP sin_pa_2 = (FP) (sp * sp);
    const FP sin_2pa = (FP) sin(2.0 * (double) pa);
    g.a[i] = cos_pa_2 * inv_std_min_2 + sin_pa_2 * inv_std_maj_2;
    g.b[i] = (FP) ((double) (-sin_2pa * inv_std_min_2) * 0.5 +
            (double) (sin_2pa * inv_std_maj_2) * 0.5);
    g.c[i] = sin_pa_2 * inv_std_min_2 + cos_pa_2 * inv_std_maj_2;
}

__global__ void zero_int_kernel(int* p) { *p = 0; }

template <typename FP>
hipError_t gaussian_t(int num, const void* table, size_t capacity,
        const oskar_GaussianColumns* cols, double ra0, double cos_dec0,
        double sin_dec0, int* num_failed, hipStream_t s)
{
    GaussArgs<FP> g;
    FP* t = (FP*) table;
    g.num = num;
    g.ra = t + cols->ra * capacity;
    g.dec = t + cols->dec * capacity;
    g.maj = t + cols->maj * capacity;
    g.min = t + cols->min * capacity;
    g.pa = t + cols->pa * capacity;
    g.a = t + cols->a * capacity;
    g.b = t + cols->b * capacity;
    g.c = t + cols->c * capacity;
    g.lon0 = (FP) ra0;
    g.cos_lat0 = (FP) cos_dec0;
    g.sin_lat0 = (FP) sin_dec0;
    for (int j = 0; j < 6; ++j)
    {
        g.cos_t[j] = cols->cos_t[j];
        g.sin_t[j] = cols->sin_t[j];
    }
    g.num_failed = num_failed;
    hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(1), 0, s, num_failed);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess || num <= 0) return err;
    hipLaunchKernelGGL((gaussian_params_kernel<FP>),
            dim3((num + THREADS - 1) / THREADS), dim3(THREADS), 0, s, g);
    return hipGetLastError();
}

} /* namespace */

extern "C" int oskar_amd_relative_directions_launch(int precision, int num,
        const void* lon, const void* lat, double lon0, double cos_lat0,
        double sin_lat0, void* l, void* m, void* n, void* stream)
{
    if (num <= 0) return 0;
    hipStream_t s = (hipStream_t) stream;
    const dim3 grid((num + THREADS - 1) / THREADS);
    if (precision == OSKAR_SINGLE)
        hipLaunchKernelGGL((rel_dir_kernel<float>), grid, dim3(THREADS), 0, s,
                num, (const float*) lon, (const float*) lat, (float) lon0,
                (float) cos_lat0, (float) sin_lat0, (float*) l, (float*) m,
                (float*) n);
    else
        hipLaunchKernelGGL((rel_dir_kernel<double>), grid, dim3(THREADS), 0,
                s, num, (const double*) lon, (const double*) lat, lon0,
                cos_lat0, sin_lat0, (double*) l, (double*) m, (double*) n);
    return oskar_amd_hip_status((int) hipGetLastError());
}

extern "C" int oskar_amd_gaussian_params_launch(int precision, int num,
        void* table, size_t capacity, const oskar_GaussianColumns* cols,
        double ra0, double cos_dec0, double sin_dec0, int* num_failed_dev,
        void* stream)
{
    hipStream_t s = (hipStream_t) stream;
    const hipError_t e = (precision == OSKAR_SINGLE) ?
            gaussian_t<float>(num, table, capacity, cols, ra0, cos_dec0,
                    sin_dec0, num_failed_dev, s) :
            gaussian_t<double>(num, table, capacity, cols, ra0, cos_dec0,
                    sin_dec0, num_failed_dev, s);
    return oskar_amd_hip_status((int) e);
}
