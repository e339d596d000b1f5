/*
 * Plain-C complex and complex-matrix element types.
 *
 * Layout follows the reference (oskar/utility/oskar_vector_types.h:85,98):
 * a 2x2 complex matrix is four complex values {a, b, c, d} read row-major as
 * [[a, b], [c, d]]; float4c is 32-byte aligned, double4c 64-byte aligned.
 * HIP kernels in this library use their own device-side structs of the same
 * layout, so these definitions are only seen by host C code.
 */
#ifndef OSKAR_VECTOR_TYPES_H_
#define OSKAR_VECTOR_TYPES_H_

#if !defined(__HIPCC__) && !defined(__HIP_INCLUDE_HIP_HIP_VECTOR_TYPES_H__) \
    && !defined(HIP_INCLUDE_HIP_AMD_DETAIL_HIP_VECTOR_TYPES_H)
typedef struct float2 { float x, y; } float2;
typedef struct double2 { double x, y; } double2;
#endif

#if defined(__cplusplus)
struct alignas(32) float4c { float2 a, b, c, d; };
struct alignas(64) double4c { double2 a, b, c, d; };
#else
typedef struct float4c { _Alignas(32) float2 a; float2 b, c, d; } float4c;
typedef struct double4c { _Alignas(64) double2 a; double2 b, c, d; } double4c;
#endif

#endif /* OSKAR_VECTOR_TYPES_H_ */
