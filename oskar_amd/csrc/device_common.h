/*
 * Device-side helpers shared by the HIP kernels of this library (gfx950).
 * Complex and 2x2 complex-matrix value types with the memory layout of the
 * reference's float2/double2 and float4c/double4c
 * (oskar/utility/oskar_vector_types.h:85,98), plus the smearing functions.
 */
#ifndef OSKAR_AMD_DEVICE_COMMON_H_
#define OSKAR_AMD_DEVICE_COMMON_H_

#include <hip/hip_runtime.h>

namespace oskar_amd {

template <typename FP> struct Cplx { FP x, y; };
template <typename FP> struct Mat4 { Cplx<FP> a, b, c, d; };

/* pi in each precision; the reference uses ((FP) M_PI). */
template <typename FP> __device__ __forceinline__ FP pi_v();
template <> __device__ __forceinline__ float pi_v<float>()
{ return 3.14159265358979323846f; }
template <> __device__ __forceinline__ double pi_v<double>()
{ return 3.14159265358979323846; }

/*
 * sinc(x) = sin(x)/x with sinc(0) = 1 (oskar/utility/oskar_kernel_macros.h:281).
 *
 * Single precision: |x| < 0.5 uses the Taylor series of sin(x)/x to x^8
 * (truncation error below x^10/11! = 2.5e-11 there), which keeps the full
 * relative precision that a quotient of a rounded sin by a small x would
 * lose; larger |x| use the hardware sine (__sinf).
 * Double precision: |x| < 0.5 uses the series to x^14 (truncation below
 * x^16/17! = 4.3e-20 there; within 1.2e-16 of the exact value over
 * 2e5 random arguments, as sin(x)/x is), which costs 8 fused multiply-adds
 * instead of a double sine and a division; larger |x| use sin(x)/x.
 */
__device__ __forceinline__ float sinc_f(float x)
{
    const float x2 = x * x;
    if (fabsf(x) < 0.5f)
    {
        /* 1 - x^2/3! + x^4/5! - x^6/7! + x^8/9! */
        float p = 2.7557319224e-6f;
        p = fmaf(p, x2, -1.9841269841e-4f);
        p = fmaf(p, x2, 8.3333333333e-3f);
        p = fmaf(p, x2, -1.6666666667e-1f);
        return fmaf(p, x2, 1.0f);
    }
    return __sinf(x) / x;
}

__device__ __forceinline__ double sinc_d(double x)
{
    if (fabs(x) < 0.5)
    {
        /* 1 - x^2/3! + x^4/5! - ... - x^14/15! */
        const double y = x * x;
        double p = -7.647163731819816e-13;
        p = fma(p, y, 1.6059043836821613e-10);
        p = fma(p, y, -2.505210838544172e-8);
        p = fma(p, y, 2.7557319223985893e-6);
        p = fma(p, y, -1.984126984126984e-4);
        p = fma(p, y, 8.333333333333333e-3);
        p = fma(p, y, -1.6666666666666666e-1);
        return fma(p, y, 1.0);
    }
    return sin(x) / x;
}

template <typename FP> __device__ __forceinline__ FP sinc_v(FP x);
template <> __device__ __forceinline__ float sinc_v<float>(float x)
{ return sinc_f(x); }
template <> __device__ __forceinline__ double sinc_v<double>(double x)
{ return sinc_d(x); }

template <typename FP> __device__ __forceinline__ FP exp_v(FP x);
template <> __device__ __forceinline__ float exp_v<float>(float x)
{ return __expf(x); }
template <> __device__ __forceinline__ double exp_v<double>(double x)
{ return exp(x); }

template <typename FP> __device__ __forceinline__ void sincos_v(FP x,
        FP* s, FP* c);
template <> __device__ __forceinline__ void sincos_v<float>(float x,
        float* s, float* c) { sincosf(x, s, c); }
template <> __device__ __forceinline__ void sincos_v<double>(double x,
        double* s, double* c) { sincos(x, s, c); }

} /* namespace oskar_amd */

#endif /* OSKAR_AMD_DEVICE_COMMON_H_ */
