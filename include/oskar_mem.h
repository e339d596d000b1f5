/*
 * oskar_Mem: typed, reference-counted memory block in host or HIP device
 * memory.
 *
 * Reference: struct oskar/mem/private_mem.h:24-38; type codes
 * oskar/binary/oskar_binary_data_types.h:23-50; location codes
 * oskar/mem/oskar_mem.h:36-42; accessors oskar/mem/oskar_mem_accessors.h.
 * Every function below has the same name, signature and error behaviour as
 * its reference counterpart.  Location OSKAR_GPU means HIP device memory on
 * the current HIP device.
 */
#ifndef OSKAR_MEM_H_
#define OSKAR_MEM_H_

#include "oskar_global.h"
#include "oskar_vector_types.h"

#ifdef __cplusplus
extern "C" {
#endif

struct oskar_Mem;
#ifndef OSKAR_MEM_TYPEDEF_
#define OSKAR_MEM_TYPEDEF_
typedef struct oskar_Mem oskar_Mem;
#endif

enum OSKAR_MEM_LOCATION
{
    OSKAR_CPU = 0,
    OSKAR_GPU = 1,
    OSKAR_CL = 2
};

enum OSKAR_DATA_TYPE
{
    OSKAR_CHAR                   = 0x01,
    OSKAR_INT                    = 0x02,
    OSKAR_SINGLE                 = 0x04,
    OSKAR_DOUBLE                 = 0x08,
    OSKAR_COMPLEX                = 0x20,
    OSKAR_MATRIX                 = 0x40,
    OSKAR_SINGLE_COMPLEX         = OSKAR_SINGLE | OSKAR_COMPLEX,
    OSKAR_DOUBLE_COMPLEX         = OSKAR_DOUBLE | OSKAR_COMPLEX,
    OSKAR_SINGLE_COMPLEX_MATRIX  = OSKAR_SINGLE | OSKAR_COMPLEX | OSKAR_MATRIX,
    OSKAR_DOUBLE_COMPLEX_MATRIX  = OSKAR_DOUBLE | OSKAR_COMPLEX | OSKAR_MATRIX
};

/* Type-code helpers (oskar/binary/oskar_binary_data_types.h:52-120). */
OSKAR_EXPORT int oskar_type_precision(int type);
OSKAR_EXPORT int oskar_type_is_single(int type);
OSKAR_EXPORT int oskar_type_is_double(int type);
OSKAR_EXPORT int oskar_type_is_complex(int type);
OSKAR_EXPORT int oskar_type_is_real(int type);
OSKAR_EXPORT int oskar_type_is_matrix(int type);
OSKAR_EXPORT int oskar_type_is_scalar(int type);

/* oskar/mem/oskar_mem_element_size.h */
OSKAR_EXPORT size_t oskar_mem_element_size(int type);

/* oskar/mem/oskar_mem_create.h */
OSKAR_EXPORT
oskar_Mem* oskar_mem_create(int type, int location, size_t num_elements,
        int* status);

/* oskar/mem/oskar_mem_create_alias_from_raw.h: wraps caller memory; the
 * returned handle never frees it. */
OSKAR_EXPORT
oskar_Mem* oskar_mem_create_alias_from_raw(void* ptr, int type, int location,
        size_t num_elements, int* status);

/* oskar/mem/oskar_mem_create_alias.h: a view of num_elements elements of
 * src from offset (src NULL: an empty view for oskar_mem_set_alias). */
OSKAR_EXPORT
oskar_Mem* oskar_mem_create_alias(const oskar_Mem* src, size_t offset,
        size_t num_elements, int* status);

/* oskar/mem/oskar_mem_set_alias.h: re-points a view (not an owner:
 * OSKAR_ERR_MEMORY_NOT_ALLOCATED) at src; OSKAR_ERR_OUT_OF_RANGE if the
 * range passes the end of src. */
OSKAR_EXPORT
void oskar_mem_set_alias(oskar_Mem* mem, const oskar_Mem* src, size_t offset,
        size_t num_elements, int* status);

/* oskar/mem/oskar_mem_create_copy.h */
OSKAR_EXPORT
oskar_Mem* oskar_mem_create_copy(const oskar_Mem* src, int location,
        int* status);

/* oskar/mem/oskar_mem_free.h: decrements the reference count, frees at 0. */
OSKAR_EXPORT void oskar_mem_free(oskar_Mem* mem, int* status);
OSKAR_EXPORT void oskar_mem_ref_dec(oskar_Mem* mem);
OSKAR_EXPORT oskar_Mem* oskar_mem_ref_inc(oskar_Mem* mem);

/* oskar/mem/oskar_mem_copy.h: resizes dst to src's length, copies all. */
OSKAR_EXPORT void oskar_mem_copy(oskar_Mem* dst, const oskar_Mem* src,
        int* status);

/* oskar/mem/oskar_mem_copy_contents.h */
OSKAR_EXPORT
void oskar_mem_copy_contents(oskar_Mem* dst, const oskar_Mem* src,
        size_t offset_dst, size_t offset_src, size_t num_elements,
        int* status);

/* oskar/mem/oskar_mem_clear_contents.h */
OSKAR_EXPORT void oskar_mem_clear_contents(oskar_Mem* dst, int* status);

/* oskar/mem/oskar_mem_ensure.h: grows (never shrinks) the allocation. */
OSKAR_EXPORT void oskar_mem_ensure(oskar_Mem* mem, size_t num_elements,
        int* status);

/* oskar/mem/oskar_mem_realloc.h */
OSKAR_EXPORT void oskar_mem_realloc(oskar_Mem* mem, size_t num_elements,
        int* status);

/* oskar/mem/oskar_mem_set_value_real.h: sets the real part (or, for
 * matrices, the diagonal real parts) and clears the rest. */
OSKAR_EXPORT
void oskar_mem_set_value_real(oskar_Mem* mem, double value,
        size_t offset, size_t num_elements, int* status);

/* oskar/mem/oskar_mem_normalise.h: scales elements [offset, offset+n) except
 * norm_index by 1/|mem[norm_index]| (matrix: 1/sqrt(sum|.|^2 / 2)). */
OSKAR_EXPORT
void oskar_mem_normalise(oskar_Mem* mem, size_t offset,
        size_t num_elements, size_t norm_index, int* status);

/* oskar/mem/oskar_mem_evaluate_relative_error.h */
OSKAR_EXPORT
void oskar_mem_evaluate_relative_error(const oskar_Mem* val_approx,
        const oskar_Mem* val_accurate, double* min_rel_error,
        double* max_rel_error, double* avg_rel_error, double* std_rel_error,
        int* status);

/* oskar/mem/oskar_mem_get_element.h, oskar_mem_set_element.h */
OSKAR_EXPORT double oskar_mem_get_element(const oskar_Mem* mem, size_t index,
        int* status);
OSKAR_EXPORT void oskar_mem_set_element_real(oskar_Mem* mem, size_t index,
        double val, int* status);

/* oskar/mem/oskar_mem_accessors.h */
OSKAR_EXPORT int oskar_mem_allocated(const oskar_Mem* mem);
OSKAR_EXPORT size_t oskar_mem_length(const oskar_Mem* mem);
OSKAR_EXPORT int oskar_mem_location(const oskar_Mem* mem);
OSKAR_EXPORT int oskar_mem_type(const oskar_Mem* mem);
OSKAR_EXPORT int oskar_mem_precision(const oskar_Mem* mem);
OSKAR_EXPORT int oskar_mem_is_double(const oskar_Mem* mem);
OSKAR_EXPORT int oskar_mem_is_single(const oskar_Mem* mem);
OSKAR_EXPORT int oskar_mem_is_complex(const oskar_Mem* mem);
OSKAR_EXPORT int oskar_mem_is_real(const oskar_Mem* mem);
OSKAR_EXPORT int oskar_mem_is_matrix(const oskar_Mem* mem);
OSKAR_EXPORT int oskar_mem_is_scalar(const oskar_Mem* mem);
OSKAR_EXPORT void* oskar_mem_buffer(oskar_Mem* mem);
OSKAR_EXPORT const void* oskar_mem_buffer_const(const oskar_Mem* mem);
OSKAR_EXPORT void* oskar_mem_void(oskar_Mem* mem);
OSKAR_EXPORT const void* oskar_mem_void_const(const oskar_Mem* mem);
OSKAR_EXPORT char* oskar_mem_char(oskar_Mem* mem);
OSKAR_EXPORT const char* oskar_mem_char_const(const oskar_Mem* mem);
OSKAR_EXPORT int* oskar_mem_int(oskar_Mem* mem, int* status);
OSKAR_EXPORT const int* oskar_mem_int_const(const oskar_Mem* mem, int* status);
OSKAR_EXPORT float* oskar_mem_float(oskar_Mem* mem, int* status);
OSKAR_EXPORT const float* oskar_mem_float_const(const oskar_Mem* mem,
        int* status);
OSKAR_EXPORT float2* oskar_mem_float2(oskar_Mem* mem, int* status);
OSKAR_EXPORT const float2* oskar_mem_float2_const(const oskar_Mem* mem,
        int* status);
OSKAR_EXPORT float4c* oskar_mem_float4c(oskar_Mem* mem, int* status);
OSKAR_EXPORT const float4c* oskar_mem_float4c_const(const oskar_Mem* mem,
        int* status);
OSKAR_EXPORT double* oskar_mem_double(oskar_Mem* mem, int* status);
OSKAR_EXPORT const double* oskar_mem_double_const(const oskar_Mem* mem,
        int* status);
OSKAR_EXPORT double2* oskar_mem_double2(oskar_Mem* mem, int* status);
OSKAR_EXPORT const double2* oskar_mem_double2_const(const oskar_Mem* mem,
        int* status);
OSKAR_EXPORT double4c* oskar_mem_double4c(oskar_Mem* mem, int* status);
OSKAR_EXPORT const double4c* oskar_mem_double4c_const(const oskar_Mem* mem,
        int* status);

#ifdef __cplusplus
}
#endif

#endif /* OSKAR_MEM_H_ */
