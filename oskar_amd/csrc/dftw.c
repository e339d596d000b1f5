/*
 * oskar_dftw(): argument checks and dispatch to the HIP DFT kernel.
 *
 * Reference: oskar/math/src/oskar_dftw.c:42-119 (checks, in this order:
 * complex output and non-matrix complex weights; locations; precisions;
 * data type equal to output type; 3D only when both z arrays are given; the
 * output grown to offset_out + num_out; norm = 1/num_in when normalise).
 */
#include "private.h"
#include "oskar_dftw.h"

void oskar_dftw(int normalise, int num_in, double wavenumber,
        const oskar_Mem* weights_in, const oskar_Mem* x_in,
        const oskar_Mem* y_in, const oskar_Mem* z_in, int offset_coord_out,
        int num_out, const oskar_Mem* x_out, const oskar_Mem* y_out,
        const oskar_Mem* z_out, const oskar_Mem* data_idx,
        const oskar_Mem* data, int eval_x, int eval_y, int offset_out,
        oskar_Mem* output, int* status)
{
    if (*status) return;
    const int location = oskar_mem_location(output);
    const int type = oskar_mem_precision(output);
    const int is_3d = (z_in != 0 && z_out != 0);
    const int is_matrix = oskar_mem_is_matrix(output);
    const double norm_factor = normalise ? 1.0 / num_in : 1.0;
    if (!oskar_mem_is_complex(output) || !oskar_mem_is_complex(weights_in) ||
            oskar_mem_is_matrix(weights_in))
    {
        *status = OSKAR_ERR_BAD_DATA_TYPE;
        return;
    }
    if (oskar_mem_location(data) != location ||
            oskar_mem_location(weights_in) != location ||
            oskar_mem_location(x_in) != location ||
            oskar_mem_location(y_in) != location ||
            oskar_mem_location(x_out) != location ||
            oskar_mem_location(y_out) != location)
    {
        *status = OSKAR_ERR_LOCATION_MISMATCH;
        return;
    }
    if (oskar_mem_precision(weights_in) != type ||
            oskar_mem_type(x_in) != type ||
            oskar_mem_type(y_in) != type ||
            oskar_mem_type(x_out) != type ||
            oskar_mem_type(y_out) != type)
    {
        *status = OSKAR_ERR_TYPE_MISMATCH;
        return;
    }
    if (!oskar_mem_is_complex(data) ||
            oskar_mem_type(data) != oskar_mem_type(output) ||
            oskar_mem_precision(data) != type)
    {
        *status = OSKAR_ERR_TYPE_MISMATCH;
        return;
    }
    if (is_3d)
    {
        if (oskar_mem_location(z_in) != location ||
                oskar_mem_location(z_out) != location)
        {
            *status = OSKAR_ERR_LOCATION_MISMATCH;
            return;
        }
        if (oskar_mem_type(z_in) != type || oskar_mem_type(z_out) != type)
        {
            *status = OSKAR_ERR_TYPE_MISMATCH;
            return;
        }
    }
    if (data_idx && (oskar_mem_type(data_idx) != OSKAR_INT ||
            oskar_mem_location(data_idx) != location))
    {
        *status = OSKAR_ERR_TYPE_MISMATCH;
        return;
    }
    oskar_mem_ensure(output, (size_t) offset_out + num_out, status);
    if (*status) return;
    if (location != OSKAR_GPU)
    {
        *status = OSKAR_ERR_BAD_LOCATION;
        return;
    }
    /* Bounds of the inputs the kernel reads. */
    if ((int) oskar_mem_length(weights_in) < num_in ||
            (int) oskar_mem_length(x_in) < num_in ||
            (int) oskar_mem_length(y_in) < num_in ||
            (is_3d && (int) oskar_mem_length(z_in) < num_in) ||
            (data_idx && (int) oskar_mem_length(data_idx) < num_in) ||
            (int) oskar_mem_length(x_out) < offset_coord_out + num_out ||
            (int) oskar_mem_length(y_out) < offset_coord_out + num_out ||
            (is_3d && (int) oskar_mem_length(z_out) <
                    offset_coord_out + num_out) ||
            (!data_idx && oskar_mem_length(data) <
                    (size_t) num_in * (size_t) num_out))
    {
        *status = OSKAR_ERR_DIMENSION_MISMATCH;
        return;
    }
    {
        const int e = oskar_amd_dftw_launch(type, is_matrix, is_3d, num_in,
                wavenumber, oskar_mem_void_const(weights_in),
                oskar_mem_void_const(x_in), oskar_mem_void_const(y_in),
                is_3d ? oskar_mem_void_const(z_in) : 0, offset_coord_out,
                num_out, oskar_mem_void_const(x_out),
                oskar_mem_void_const(y_out),
                is_3d ? oskar_mem_void_const(z_out) : 0,
                data_idx ? (const int*) oskar_mem_void_const(data_idx) : 0,
                oskar_mem_void_const(data),
                num_out > 0 ? (int) (oskar_mem_length(data) /
                        (size_t) num_out) : 0,
                eval_x, eval_y, offset_out,
                oskar_mem_void(output), norm_factor,
                oskar_amd_current_stream());
        if (e) *status = e;
    }
}
