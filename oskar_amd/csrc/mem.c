/*
 * oskar_Mem: host (calloc) or HIP device (hipMalloc) typed buffers.
 *
 * Reference semantics: oskar/mem/src/oskar_mem_create.c:31-113,
 * oskar_mem_free.c, oskar_mem_copy.c, oskar_mem_copy_contents.c:66-91,
 * oskar_mem_ensure.c, oskar_mem_realloc.c, oskar_mem_accessors.c:20-260,
 * oskar_mem_set_value_real.c, oskar_mem_normalise.c,
 * oskar_mem_evaluate_relative_error.c:42-176.
 * Device copies and fills are queued on the current HIP stream; copies that
 * touch host memory wait for that stream before returning, as the
 * reference's cudaMemcpy does for stream 0.
 */
#include <hip/hip_runtime_api.h>
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "private.h"

/* ---- type helpers -------------------------------------------------------- */
int oskar_type_precision(int type) { return type & 0x0F; }
int oskar_type_is_single(int type) { return (type & 0x0F) == OSKAR_SINGLE; }
int oskar_type_is_double(int type) { return (type & 0x0F) == OSKAR_DOUBLE; }
int oskar_type_is_complex(int type) { return (type & OSKAR_COMPLEX) != 0; }
int oskar_type_is_real(int type) { return (type & OSKAR_COMPLEX) == 0; }
int oskar_type_is_matrix(int type) { return (type & OSKAR_MATRIX) != 0; }
int oskar_type_is_scalar(int type) { return (type & OSKAR_MATRIX) == 0; }

size_t oskar_mem_element_size(int type)
{
    size_t s = 0;
    switch (type & 0x0F)
    {
    case OSKAR_CHAR:   s = 1; break;
    case OSKAR_INT:    s = sizeof(int); break;
    case OSKAR_SINGLE: s = sizeof(float); break;
    case OSKAR_DOUBLE: s = sizeof(double); break;
    default: return 0;
    }
    if (type & OSKAR_COMPLEX) s *= 2;
    if (type & OSKAR_MATRIX) s *= 4;
    return s;
}

static int valid_type(int type)
{
    switch (type)
    {
    case OSKAR_CHAR: case OSKAR_INT: case OSKAR_SINGLE: case OSKAR_DOUBLE:
    case OSKAR_SINGLE_COMPLEX: case OSKAR_DOUBLE_COMPLEX:
    case OSKAR_SINGLE_COMPLEX_MATRIX: case OSKAR_DOUBLE_COMPLEX_MATRIX:
        return 1;
    default:
        return 0;
    }
}

static void* mem_alloc(int location, size_t bytes, int* status)
{
    void* p = 0;
    if (bytes == 0) return 0;
    if (location == OSKAR_CPU)
    {
        p = calloc(1, bytes);
        if (!p) *status = OSKAR_ERR_MEMORY_ALLOC_FAILURE;
    }
    else if (location == OSKAR_GPU)
    {
        const hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess)
        {
            p = 0;
            *status = OSKAR_ERR_MEMORY_ALLOC_FAILURE;
        }
    }
    else *status = OSKAR_ERR_BAD_LOCATION;
    return p;
}

static void mem_release(int location, void* p)
{
    if (!p) return;
    if (location == OSKAR_CPU) free(p);
    else if (location == OSKAR_GPU) (void) hipFree(p);
}

oskar_Mem* oskar_mem_create(int type, int location, size_t num_elements,
        int* status)
{
    oskar_Mem* mem = 0;
    if (*status) return 0;
    if (!valid_type(type))
    {
        *status = OSKAR_ERR_BAD_DATA_TYPE;
        return 0;
    }
    if (location != OSKAR_CPU && location != OSKAR_GPU)
    {
        *status = OSKAR_ERR_BAD_LOCATION;
        return 0;
    }
    mem = (oskar_Mem*) calloc(1, sizeof(oskar_Mem));
    if (!mem)
    {
        *status = OSKAR_ERR_MEMORY_ALLOC_FAILURE;
        return 0;
    }
    mem->type = type;
    mem->location = location;
    mem->owner = 1;
    mem->ref_count = 1;
    mem->data = mem_alloc(location, num_elements * oskar_mem_element_size(type),
            status);
    if (*status)
    {
        free(mem);
        return 0;
    }
    mem->num_elements = num_elements;
    return mem;
}

oskar_Mem* oskar_mem_create_alias_from_raw(void* ptr, int type, int location,
        size_t num_elements, int* status)
{
    oskar_Mem* mem = 0;
    if (*status) return 0;
    if (!valid_type(type))
    {
        *status = OSKAR_ERR_BAD_DATA_TYPE;
        return 0;
    }
    mem = (oskar_Mem*) calloc(1, sizeof(oskar_Mem));
    if (!mem)
    {
        *status = OSKAR_ERR_MEMORY_ALLOC_FAILURE;
        return 0;
    }
    mem->type = type;
    mem->location = location;
    mem->owner = 0;
    mem->ref_count = 1;
    mem->data = ptr;
    mem->num_elements = num_elements;
    return mem;
}

oskar_Mem* oskar_mem_create_alias(const oskar_Mem* src, size_t offset,
        size_t num_elements, int* status)
{
    oskar_Mem* mem = (oskar_Mem*) calloc(1, sizeof(oskar_Mem));
    if (!mem)
    {
        *status = OSKAR_ERR_MEMORY_ALLOC_FAILURE;
        return 0;
    }
    mem->owner = 0;
    mem->ref_count = 1;
    if (src)
    {
        mem->type = src->type;
        mem->location = src->location;
        mem->num_elements = num_elements;
        mem->data = src->data ? (char*) src->data +
                offset * oskar_mem_element_size(src->type) : 0;
    }
    return mem;
}

void oskar_mem_set_alias(oskar_Mem* mem, const oskar_Mem* src, size_t offset,
        size_t num_elements, int* status)
{
    if (*status) return;
    if (mem->owner)
    {
        *status = OSKAR_ERR_MEMORY_NOT_ALLOCATED;
        return;
    }
    if (offset + num_elements > src->num_elements)
    {
        *status = OSKAR_ERR_OUT_OF_RANGE;
        return;
    }
    mem->type = src->type;
    mem->location = src->location;
    mem->num_elements = num_elements;
    mem->data = src->data ? (char*) src->data +
            offset * oskar_mem_element_size(src->type) : 0;
}

oskar_Mem* oskar_mem_create_copy(const oskar_Mem* src, int location,
        int* status)
{
    oskar_Mem* mem = 0;
    if (*status || !src) return 0;
    mem = oskar_mem_create(src->type, location, src->num_elements, status);
    oskar_mem_copy_contents(mem, src, 0, 0, src->num_elements, status);
    return mem;
}

oskar_Mem* oskar_mem_ref_inc(oskar_Mem* mem)
{
    if (mem) mem->ref_count++;
    return mem;
}

void oskar_mem_ref_dec(oskar_Mem* mem)
{
    int status = 0;
    oskar_mem_free(mem, &status);
}

void oskar_mem_free(oskar_Mem* mem, int* status)
{
    (void) status;
    if (!mem) return;
    if (--mem->ref_count > 0) return;
    if (mem->owner) mem_release(mem->location, mem->data);
    free(mem);
}

void oskar_mem_copy_contents(oskar_Mem* dst, const oskar_Mem* src,
        size_t offset_dst, size_t offset_src, size_t num_elements,
        int* status)
{
    if (*status || !dst || !src || num_elements == 0) return;
    if (src->type != dst->type)
    {
        *status = OSKAR_ERR_TYPE_MISMATCH;
        return;
    }
    if (offset_src + num_elements > src->num_elements ||
            offset_dst + num_elements > dst->num_elements)
    {
        *status = OSKAR_ERR_OUT_OF_RANGE;
        return;
    }
    if (!src->data || !dst->data)
    {
        *status = OSKAR_ERR_MEMORY_NOT_ALLOCATED;
        return;
    }
    {
        const size_t es = oskar_mem_element_size(src->type);
        const size_t bytes = num_elements * es;
        char* d = (char*) dst->data + offset_dst * es;
        const char* s = (const char*) src->data + offset_src * es;
        const hipStream_t stream = (hipStream_t) oskar_amd_current_stream();
        hipError_t e = hipSuccess;
        if (src->location == OSKAR_CPU && dst->location == OSKAR_CPU)
        {
            memmove(d, s, bytes);
            return;
        }
        if (src->location == OSKAR_GPU && dst->location == OSKAR_GPU)
            e = hipMemcpyAsync(d, s, bytes, hipMemcpyDeviceToDevice, stream);
        else if (src->location == OSKAR_CPU && dst->location == OSKAR_GPU)
        {
            e = hipMemcpyAsync(d, s, bytes, hipMemcpyHostToDevice, stream);
            if (e == hipSuccess) e = hipStreamSynchronize(stream);
        }
        else if (src->location == OSKAR_GPU && dst->location == OSKAR_CPU)
        {
            e = hipMemcpyAsync(d, s, bytes, hipMemcpyDeviceToHost, stream);
            if (e == hipSuccess) e = hipStreamSynchronize(stream);
        }
        else
        {
            *status = OSKAR_ERR_BAD_LOCATION;
            return;
        }
        if (e != hipSuccess) *status = (int) e;
    }
}

void oskar_mem_realloc(oskar_Mem* mem, size_t num_elements, int* status)
{
    if (*status || !mem) return;
    if (!mem->owner)
    {
        *status = OSKAR_ERR_MEMORY_NOT_ALLOCATED;
        return;
    }
    if (num_elements == mem->num_elements) return;
    {
        const size_t es = oskar_mem_element_size(mem->type);
        void* p = mem_alloc(mem->location, num_elements * es, status);
        const size_t keep = num_elements < mem->num_elements ?
                num_elements : mem->num_elements;
        if (*status) return;
        if (keep > 0 && mem->data)
        {
            if (mem->location == OSKAR_CPU) memcpy(p, mem->data, keep * es);
            else
            {
                const hipError_t e = hipMemcpyAsync(p, mem->data, keep * es,
                        hipMemcpyDeviceToDevice,
                        (hipStream_t) oskar_amd_current_stream());
                if (e != hipSuccess) *status = (int) e;
                (void) hipStreamSynchronize(
                        (hipStream_t) oskar_amd_current_stream());
            }
        }
        mem_release(mem->location, mem->data);
        mem->data = p;
        mem->num_elements = num_elements;
    }
}

void oskar_mem_ensure(oskar_Mem* mem, size_t num_elements, int* status)
{
    if (*status || !mem) return;
    if (mem->num_elements >= num_elements) return;
    oskar_mem_realloc(mem, num_elements, status);
}

void oskar_mem_copy(oskar_Mem* dst, const oskar_Mem* src, int* status)
{
    if (*status || !dst || !src) return;
    if (src->type != dst->type)
    {
        *status = OSKAR_ERR_TYPE_MISMATCH;
        return;
    }
    oskar_mem_ensure(dst, src->num_elements, status);
    oskar_mem_copy_contents(dst, src, 0, 0, src->num_elements, status);
}

void oskar_mem_clear_contents(oskar_Mem* mem, int* status)
{
    if (*status || !mem || !mem->data || mem->num_elements == 0) return;
    {
        const size_t bytes = mem->num_elements *
                oskar_mem_element_size(mem->type);
        if (mem->location == OSKAR_CPU) memset(mem->data, 0, bytes);
        else if (mem->location == OSKAR_GPU)
        {
            const hipError_t e = hipMemsetAsync(mem->data, 0, bytes,
                    (hipStream_t) oskar_amd_current_stream());
            if (e != hipSuccess) *status = (int) e;
        }
        else *status = OSKAR_ERR_BAD_LOCATION;
    }
}

/* ---- element-wise utilities --------------------------------------------- */
void oskar_mem_set_value_real(oskar_Mem* mem, double value, size_t offset,
        size_t num_elements, int* status)
{
    size_t i = 0;
    if (*status || !mem) return;
    if (offset + num_elements > mem->num_elements)
    {
        *status = OSKAR_ERR_OUT_OF_RANGE;
        return;
    }
    if (mem->location == OSKAR_GPU)
    {
        *status = oskar_amd_hip_status(oskar_amd_mem_set_value_real_launch(
                mem->type, offset, num_elements, value, mem->data,
                oskar_amd_current_stream()));
        return;
    }
    switch (mem->type)
    {
#define SETV(T, STRIDE, DIAG) { T* p = (T*) mem->data;                       \
        for (i = offset; i < offset + num_elements; ++i) {                    \
            size_t k; for (k = 0; k < STRIDE; ++k) p[i * STRIDE + k] = 0;     \
            p[i * STRIDE] = (T) value;                                        \
            if (DIAG) p[i * STRIDE + 6] = (T) value; } }
    case OSKAR_SINGLE: SETV(float, 1, 0) break;
    case OSKAR_DOUBLE: SETV(double, 1, 0) break;
    case OSKAR_SINGLE_COMPLEX: SETV(float, 2, 0) break;
    case OSKAR_DOUBLE_COMPLEX: SETV(double, 2, 0) break;
    case OSKAR_SINGLE_COMPLEX_MATRIX: SETV(float, 8, 1) break;
    case OSKAR_DOUBLE_COMPLEX_MATRIX: SETV(double, 8, 1) break;
#undef SETV
    default:
        *status = OSKAR_ERR_BAD_DATA_TYPE;
    }
}

void oskar_mem_normalise(oskar_Mem* mem, size_t offset, size_t num_elements,
        size_t norm_index, int* status)
{
    size_t i = 0;
    if (*status || !mem) return;
    if (norm_index >= mem->num_elements ||
            offset + num_elements > mem->num_elements)
    {
        *status = OSKAR_ERR_OUT_OF_RANGE;
        return;
    }
    if (mem->location == OSKAR_GPU)
    {
        *status = oskar_amd_hip_status(oskar_amd_mem_normalise_launch(
                mem->type, (unsigned int) offset, (unsigned int) num_elements,
                mem->data, (unsigned int) norm_index,
                oskar_amd_current_stream()));
        return;
    }
#define NORM(T, STRIDE) { T* p = (T*) mem->data; T amp = 0, scal;             \
        size_t k; const T* v = p + norm_index * STRIDE;                       \
        if (STRIDE == 1) scal = (T) 1 / v[0];                                 \
        else { for (k = 0; k < STRIDE; ++k) amp += v[k] * v[k];               \
               if (STRIDE == 8) amp /= (T) 2;                                 \
               scal = (T) 1 / (T) sqrt(amp); }                                \
        for (i = offset; i < offset + num_elements; ++i)                      \
            if (i != norm_index) for (k = 0; k < STRIDE; ++k)                 \
                p[i * STRIDE + k] *= scal; }
    switch (mem->type)
    {
    case OSKAR_SINGLE: NORM(float, 1) break;
    case OSKAR_DOUBLE: NORM(double, 1) break;
    case OSKAR_SINGLE_COMPLEX: NORM(float, 2) break;
    case OSKAR_DOUBLE_COMPLEX: NORM(double, 2) break;
    case OSKAR_SINGLE_COMPLEX_MATRIX: NORM(float, 8) break;
    case OSKAR_DOUBLE_COMPLEX_MATRIX: NORM(double, 8) break;
    default:
        *status = OSKAR_ERR_BAD_DATA_TYPE;
    }
#undef NORM
}

/* Same statistic as oskar_mem_evaluate_relative_error.c:42-176: per real
 * component |(|a| - |b|)| / (|a| + |b|) (or the absolute difference when it
 * is below a tolerance), over the first length (x4 for matrices) reals. */
void oskar_mem_evaluate_relative_error(const oskar_Mem* val_approx,
        const oskar_Mem* val_accurate, double* min_rel_error,
        double* max_rel_error, double* avg_rel_error, double* std_rel_error,
        int* status)
{
    size_t i = 0, n = 0;
    oskar_Mem *ta = 0, *tb = 0;
    const oskar_Mem *pa = val_approx, *pb = val_accurate;
    double mean = 0.0, s = 0.0, tol = 1e-5;
    if (*status) return;
    if (max_rel_error) *max_rel_error = -DBL_MAX;
    if (min_rel_error) *min_rel_error = DBL_MAX;
    if (avg_rel_error) *avg_rel_error = DBL_MAX;
    if (std_rel_error) *std_rel_error = DBL_MAX;
    if ((oskar_mem_is_matrix(pa) && !oskar_mem_is_matrix(pb)) ||
            (oskar_mem_is_complex(pa) && !oskar_mem_is_complex(pb)))
    {
        *status = OSKAR_ERR_TYPE_MISMATCH;
        return;
    }
    {
        const int qa = oskar_mem_precision(pa), qb = oskar_mem_precision(pb);
        if ((qa != OSKAR_SINGLE && qa != OSKAR_DOUBLE) ||
                (qb != OSKAR_SINGLE && qb != OSKAR_DOUBLE))
        {
            *status = OSKAR_ERR_BAD_DATA_TYPE;
            return;
        }
        if (qa == OSKAR_DOUBLE && qb == OSKAR_DOUBLE) tol = 1e-15;
    }
    n = pa->num_elements < pb->num_elements ? pa->num_elements :
            pb->num_elements;
    if (oskar_mem_is_matrix(pa)) n *= 4;
    if (pa->location != OSKAR_CPU)
        pa = ta = oskar_mem_create_copy(val_approx, OSKAR_CPU, status);
    if (pb->location != OSKAR_CPU)
        pb = tb = oskar_mem_create_copy(val_accurate, OSKAR_CPU, status);
    if (*status) goto done;
    for (i = 0; i < n; ++i)
    {
        const double a = oskar_mem_is_double(pa) ?
                ((const double*) pa->data)[i] : ((const float*) pa->data)[i];
        const double b = oskar_mem_is_double(pb) ?
                ((const double*) pb->data)[i] : ((const float*) pb->data)[i];
        const double abs_a = fabs(a), abs_b = fabs(b);
        const double diff = fabs(abs_a - abs_b);
        double r = 0.0;
        if (a == b || (abs_a < FLT_EPSILON && abs_b < FLT_EPSILON)) r = 0.0;
        else if (diff < tol) r = diff;
        else r = diff / (abs_a + abs_b);
        if (max_rel_error && r > *max_rel_error) *max_rel_error = r;
        if (min_rel_error && r < *min_rel_error) *min_rel_error = r;
        if (i == 0) { mean = r; s = 0.0; }
        else
        {
            const double m_new = mean + (r - mean) / (double) (i + 1);
            s += (r - mean) * (r - m_new);
            mean = m_new;
        }
    }
    if (avg_rel_error) *avg_rel_error = mean;
    if (std_rel_error) *std_rel_error = (n > 0) ? sqrt(s / (double) n) : 0.0;
done:
    oskar_mem_free(ta, status);
    oskar_mem_free(tb, status);
}

double oskar_mem_get_element(const oskar_Mem* mem, size_t index, int* status)
{
    double v = 0.0;
    if (*status || !mem) return 0.0;
    if (index >= mem->num_elements * (oskar_mem_is_complex(mem) ? 2 : 1) *
            (oskar_mem_is_matrix(mem) ? 4 : 1))
    {
        *status = OSKAR_ERR_OUT_OF_RANGE;
        return 0.0;
    }
    {
        const size_t es = (oskar_mem_precision(mem) == OSKAR_DOUBLE) ? 8 :
                (oskar_mem_precision(mem) == OSKAR_INT) ? 4 : 4;
        char buf[8];
        if (mem->location == OSKAR_CPU)
            memcpy(buf, (const char*) mem->data + index * es, es);
        else
        {
            const hipError_t e = hipMemcpy(buf, (const char*) mem->data +
                    index * es, es, hipMemcpyDeviceToHost);
            if (e != hipSuccess) { *status = (int) e; return 0.0; }
        }
        switch (oskar_mem_precision(mem))
        {
        case OSKAR_DOUBLE: { double d; memcpy(&d, buf, 8); v = d; break; }
        case OSKAR_SINGLE: { float f; memcpy(&f, buf, 4); v = f; break; }
        case OSKAR_INT: { int k; memcpy(&k, buf, 4); v = k; break; }
        default: *status = OSKAR_ERR_BAD_DATA_TYPE;
        }
    }
    return v;
}

void oskar_mem_set_element_real(oskar_Mem* mem, size_t index, double val,
        int* status)
{
    if (*status || !mem) return;
    if (index >= mem->num_elements)
    {
        *status = OSKAR_ERR_OUT_OF_RANGE;
        return;
    }
    {
        char buf[8];
        size_t es = 0;
        switch (oskar_mem_precision(mem))
        {
        case OSKAR_DOUBLE: { double d = val; memcpy(buf, &d, 8); es = 8;
            break; }
        case OSKAR_SINGLE: { float f = (float) val; memcpy(buf, &f, 4);
            es = 4; break; }
        case OSKAR_INT: { int k = (int) val; memcpy(buf, &k, 4); es = 4;
            break; }
        default: *status = OSKAR_ERR_BAD_DATA_TYPE; return;
        }
        {
            const size_t stride = oskar_mem_element_size(mem->type);
            char* dst = (char*) mem->data + index * stride;
            if (mem->location == OSKAR_CPU) memcpy(dst, buf, es);
            else
            {
                const hipError_t e = hipMemcpy(dst, buf, es,
                        hipMemcpyHostToDevice);
                if (e != hipSuccess) *status = (int) e;
            }
        }
    }
}

/* ---- accessors ------------------------------------------------------------ */
int oskar_mem_allocated(const oskar_Mem* m) { return m && m->data != 0; }
size_t oskar_mem_length(const oskar_Mem* m) { return m ? m->num_elements : 0; }
int oskar_mem_location(const oskar_Mem* m) { return m ? m->location : 0; }
int oskar_mem_type(const oskar_Mem* m) { return m ? m->type : 0; }
int oskar_mem_precision(const oskar_Mem* m)
{ return m ? oskar_type_precision(m->type) : 0; }
int oskar_mem_is_double(const oskar_Mem* m)
{ return m ? oskar_type_is_double(m->type) : 0; }
int oskar_mem_is_single(const oskar_Mem* m)
{ return m ? oskar_type_is_single(m->type) : 0; }
int oskar_mem_is_complex(const oskar_Mem* m)
{ return m ? oskar_type_is_complex(m->type) : 0; }
int oskar_mem_is_real(const oskar_Mem* m)
{ return m ? oskar_type_is_real(m->type) : 0; }
int oskar_mem_is_matrix(const oskar_Mem* m)
{ return m ? oskar_type_is_matrix(m->type) : 0; }
int oskar_mem_is_scalar(const oskar_Mem* m)
{ return m ? oskar_type_is_scalar(m->type) : 0; }
void* oskar_mem_buffer(oskar_Mem* m) { return m ? &m->data : 0; }
const void* oskar_mem_buffer_const(const oskar_Mem* m)
{ return m ? &m->data : 0; }
void* oskar_mem_void(oskar_Mem* m) { return m ? m->data : 0; }
const void* oskar_mem_void_const(const oskar_Mem* m) { return m ? m->data : 0; }
char* oskar_mem_char(oskar_Mem* m) { return m ? (char*) m->data : 0; }
const char* oskar_mem_char_const(const oskar_Mem* m)
{ return m ? (const char*) m->data : 0; }

#define TYPED_ACCESSOR(NAME, CTYPE, CHECK)                                    \
    CTYPE* NAME(oskar_Mem* m, int* status) {                                  \
        if (!m || !(CHECK)) { *status = OSKAR_ERR_TYPE_MISMATCH; return 0; }  \
        return (CTYPE*) m->data; }                                            \
    const CTYPE* NAME##_const(const oskar_Mem* m, int* status) {              \
        if (!m || !(CHECK)) { *status = OSKAR_ERR_TYPE_MISMATCH; return 0; }  \
        return (const CTYPE*) m->data; }
TYPED_ACCESSOR(oskar_mem_int, int, m->type == OSKAR_INT)
TYPED_ACCESSOR(oskar_mem_float, float, oskar_type_precision(m->type) ==
        OSKAR_SINGLE)
TYPED_ACCESSOR(oskar_mem_float2, float2, oskar_type_precision(m->type) ==
        OSKAR_SINGLE && oskar_type_is_complex(m->type))
TYPED_ACCESSOR(oskar_mem_float4c, float4c, m->type ==
        OSKAR_SINGLE_COMPLEX_MATRIX)
TYPED_ACCESSOR(oskar_mem_double, double, oskar_type_precision(m->type) ==
        OSKAR_DOUBLE)
TYPED_ACCESSOR(oskar_mem_double2, double2, oskar_type_precision(m->type) ==
        OSKAR_DOUBLE && oskar_type_is_complex(m->type))
TYPED_ACCESSOR(oskar_mem_double4c, double4c, m->type ==
        OSKAR_DOUBLE_COMPLEX_MATRIX)
#undef TYPED_ACCESSOR
