/*
 * oskar_cross_correlate_factorised(): checks and dispatch to the MFMA
 * kernels in xcorr_factorised.hip.  Derived scalars and checks follow
 * oskar/correlate/src/oskar_cross_correlate.c:42-110.
 */
#include <float.h>
#include <math.h>

#include "private.h"
#include "oskar_cross_correlate_factorised.h"

int oskar_amd_fact_launch(int precision, int mode, int use_casa,
        int num_sources, int num_stations, const void* beam,
        const void* I, const void* Q, const void* U, const void* V,
        const void* l, const void* m, const void* n, const void* st_u,
        const void* st_v, const void* st_w, double frequency_hz,
        double uv_min_lambda, double uv_max_lambda, int offset_out,
        void* vis, int src_per_part, void* stream);

void oskar_amd_fact_set_profiling(int on);
int oskar_amd_fact_last_times(double* ms);

static int part_size = 8192;

void oskar_cross_correlate_factorised_set_profiling(int on)
{
    oskar_amd_fact_set_profiling(on);
}

int oskar_cross_correlate_factorised_last_times(double ms[3])
{
    return oskar_amd_fact_last_times(ms);
}

int oskar_cross_correlate_factorised_set_part_size(int sources)
{
    const int old = part_size;
    if (sources > 0) part_size = ((sources + 7) / 8) * 8;
    return old;
}

void oskar_cross_correlate_factorised(int use_casa, int num_sources,
        const oskar_Mem* station_beam, const oskar_Mem* const src_flux[4],
        const oskar_Mem* const src_dir[3], const oskar_Telescope* tel,
        const oskar_Mem* const station_uvw[3], double frequency_hz,
        int mode, int offset_out, oskar_Mem* vis, int* status)
{
    int i = 0;
    if (*status) return;
    const int num_stations = oskar_telescope_num_stations(tel);
    frequency_hz = fabs(frequency_hz);
    const double inv_wavelength = frequency_hz / 299792458.0;
    double uv_min = oskar_telescope_uv_filter_min(tel);
    double uv_max = oskar_telescope_uv_filter_max(tel);
    if (oskar_telescope_uv_filter_units(tel) == OSKAR_METRES)
    {
        uv_min *= inv_wavelength;
        uv_max *= inv_wavelength;
    }
    if (uv_max < 0.0 || uv_max > FLT_MAX) uv_max = FLT_MAX;

    const int location = oskar_mem_location(vis);
    const int prec = oskar_mem_precision(vis);
    if (oskar_telescope_mem_location(tel) != location)
    {
        *status = OSKAR_ERR_LOCATION_MISMATCH;
        return;
    }
    for (i = 0; i < 3; ++i)
    {
        if (oskar_mem_location(station_uvw[i]) != location ||
                oskar_mem_location(src_dir[i]) != location)
        {
            *status = OSKAR_ERR_LOCATION_MISMATCH;
            return;
        }
    }
    for (i = 0; i < 4; ++i)
    {
        if (!src_flux[i] || oskar_mem_location(src_flux[i]) != location)
        {
            *status = OSKAR_ERR_LOCATION_MISMATCH;
            return;
        }
    }
    if (station_beam && oskar_mem_location(station_beam) != location)
    {
        *status = OSKAR_ERR_LOCATION_MISMATCH;
        return;
    }
    if (!oskar_mem_is_matrix(vis) || !oskar_mem_is_complex(vis) ||
            (prec != OSKAR_SINGLE && prec != OSKAR_DOUBLE))
    {
        *status = OSKAR_ERR_BAD_DATA_TYPE;
        return;
    }
    for (i = 0; i < 3; ++i)
    {
        if (oskar_mem_type(station_uvw[i]) != prec ||
                oskar_mem_type(src_dir[i]) != prec)
        {
            *status = OSKAR_ERR_TYPE_MISMATCH;
            return;
        }
    }
    for (i = 0; i < 4; ++i)
    {
        if (oskar_mem_type(src_flux[i]) != prec)
        {
            *status = OSKAR_ERR_TYPE_MISMATCH;
            return;
        }
    }
    if (station_beam && oskar_mem_type(station_beam) != oskar_mem_type(vis))
    {
        *status = OSKAR_ERR_TYPE_MISMATCH;
        return;
    }
    for (i = 0; i < 3; ++i)
    {
        if ((int) oskar_mem_length(station_uvw[i]) != num_stations ||
                (int) oskar_mem_length(src_dir[i]) < num_sources)
        {
            *status = OSKAR_ERR_DIMENSION_MISMATCH;
            return;
        }
    }
    for (i = 0; i < 4; ++i)
    {
        if ((int) oskar_mem_length(src_flux[i]) < num_sources)
        {
            *status = OSKAR_ERR_DIMENSION_MISMATCH;
            return;
        }
    }
    if ((station_beam && (int) oskar_mem_length(station_beam) < num_sources)
            || oskar_mem_length(vis) < (size_t) offset_out +
            (size_t) oskar_telescope_num_baselines(tel))
    {
        *status = OSKAR_ERR_DIMENSION_MISMATCH;
        return;
    }
    if (location != OSKAR_GPU)
    {
        *status = OSKAR_ERR_BAD_LOCATION;
        return;
    }
    if (mode != OSKAR_FACTORISED_FP16X3 && mode != OSKAR_FACTORISED_BF16X3)
    {
        *status = OSKAR_ERR_INVALID_ARGUMENT;
        return;
    }
    /* Smearing factors depend on the baseline and do not factorise. */
    if (oskar_telescope_channel_bandwidth_hz(tel) != 0.0 ||
            (oskar_telescope_time_average_sec(tel) != 0.0 &&
             oskar_telescope_phase_centre_coord_type(tel) !=
             OSKAR_COORDS_AZEL))
    {
        *status = OSKAR_ERR_FUNCTION_NOT_AVAILABLE;
        return;
    }
    {
        const int e = oskar_amd_fact_launch(prec, mode, use_casa,
                num_sources, num_stations,
                station_beam ? oskar_mem_void_const(station_beam) : 0,
                oskar_mem_void_const(src_flux[0]),
                oskar_mem_void_const(src_flux[1]),
                oskar_mem_void_const(src_flux[2]),
                oskar_mem_void_const(src_flux[3]),
                oskar_mem_void_const(src_dir[0]),
                oskar_mem_void_const(src_dir[1]),
                oskar_mem_void_const(src_dir[2]),
                oskar_mem_void_const(station_uvw[0]),
                oskar_mem_void_const(station_uvw[1]),
                oskar_mem_void_const(station_uvw[2]), frequency_hz,
                uv_min, uv_max, offset_out, oskar_mem_void(vis), part_size,
                oskar_amd_current_stream());
        if (e) *status = e;
    }
}
