/*
 * oskar_auto_correlate(): checks and dispatch to the HIP auto-correlation
 * kernels.
 *
 * Reference: oskar/correlate/src/oskar_auto_correlate.c:22-151.  The checks
 * run in the same order: vis location against the Jones location, vis type
 * against the Jones type, the Jones source count against num_sources, then
 * the data type (complex or complex-matrix, single or double).
 */
#include "private.h"
#include "oskar_auto_correlate.h"

void oskar_auto_correlate(
        int num_sources,
        const oskar_Jones* jones,
        const oskar_Mem* const src_flux[4],
        int offset_out,
        oskar_Mem* vis,
        int* status)
{
    if (*status) return;
    const oskar_Mem* J = oskar_jones_mem_const(jones);
    const int num_stations = oskar_jones_num_stations(jones);
    const int location = oskar_mem_location(J);
    if (oskar_mem_location(vis) != location)
    {
        *status = OSKAR_ERR_LOCATION_MISMATCH;
        return;
    }
    const int type = oskar_mem_type(J);
    if (oskar_mem_type(vis) != type)
    {
        *status = OSKAR_ERR_TYPE_MISMATCH;
        return;
    }
    if (oskar_jones_num_sources(jones) < num_sources)
    {
        *status = OSKAR_ERR_DIMENSION_MISMATCH;
        return;
    }
    if (type != OSKAR_SINGLE_COMPLEX_MATRIX &&
            type != OSKAR_DOUBLE_COMPLEX_MATRIX &&
            type != OSKAR_SINGLE_COMPLEX && type != OSKAR_DOUBLE_COMPLEX)
    {
        *status = OSKAR_ERR_BAD_DATA_TYPE;
        return;
    }
    if (location != OSKAR_GPU)
    {
        /* This library provides the HIP path only. */
        *status = OSKAR_ERR_BAD_LOCATION;
        return;
    }
    const int matrix = oskar_type_is_matrix(type);
    const int precision = oskar_type_precision(type);
    const int num_flux = matrix ? 4 : 1;
    int i;
    for (i = 0; i < num_flux; ++i)
    {
        if (!src_flux[i] || oskar_mem_location(src_flux[i]) != location)
        {
            *status = OSKAR_ERR_LOCATION_MISMATCH;
            return;
        }
        if (oskar_mem_type(src_flux[i]) != precision)
        {
            *status = OSKAR_ERR_TYPE_MISMATCH;
            return;
        }
        if ((int) oskar_mem_length(src_flux[i]) < num_sources)
        {
            *status = OSKAR_ERR_DIMENSION_MISMATCH;
            return;
        }
    }
    if (offset_out < 0 || oskar_mem_length(vis) <
            (size_t) offset_out + (size_t) num_stations)
    {
        *status = OSKAR_ERR_DIMENSION_MISMATCH;
        return;
    }
    {
        const int e = oskar_amd_acorr_launch(precision, matrix, num_sources,
                num_stations, offset_out, oskar_mem_void_const(J),
                oskar_mem_void_const(src_flux[0]),
                matrix ? oskar_mem_void_const(src_flux[1]) : 0,
                matrix ? oskar_mem_void_const(src_flux[2]) : 0,
                matrix ? oskar_mem_void_const(src_flux[3]) : 0,
                oskar_mem_void(vis), oskar_amd_current_stream());
        if (e) *status = e;
    }
}
