/*
 * oskar_cross_correlate(): derived scalars, checks and dispatch to the HIP
 * direct-correlation kernel.
 *
 * Reference: oskar/correlate/src/oskar_cross_correlate.c:20-739.  The
 * derived quantities are computed as there (:42-71): inv_wavelength = |f|/c,
 * frac_bandwidth = dnu/|f|; time smearing only when the phase centre is not
 * AZEL, with gha0 = gast - RA0 and dec0 = Dec0; the uv filter converted to
 * wavelengths when given in metres and capped at FLT_MAX.  Checks run in
 * the same order (location, types, dimensions).  Kernel selection follows
 * oskar_cross_correlate_omp.cpp:178-188: bandwidth smearing iff
 * frac_bandwidth != 0, time smearing iff time_average != 0.
 */
#include <float.h>
#include <math.h>

#include "private.h"
#include "oskar_cross_correlate.h"

static void launch(int precision, int matrix, int gaussian,
        int use_casa, int num_sources, int num_stations, int offset_out,
        const void* jones, const void* I, const void* Q, const void* U,
        const void* V, const void* l, const void* m, const void* n,
        const void* a, const void* b, const void* c, const void* st_u,
        const void* st_v, const void* st_w, const void* st_x,
        const void* st_y, double uv_min, double uv_max, double inv_wl,
        double frac_bw, double time_int, double gha0, double dec0,
        void* vis, int no_alloc, int* status)
{
    oskar_XcorrParams p;
    p.use_casa = use_casa;
    p.num_sources = num_sources;
    p.num_stations = num_stations;
    p.offset_out = offset_out;
    p.gaussian = gaussian;
    p.bandwidth_smearing = (frac_bw != 0.0);
    p.time_smearing = (time_int != 0.0);
    p.uv_min_lambda = uv_min;
    p.uv_max_lambda = uv_max;
    p.inv_wavelength = inv_wl;
    p.frac_bandwidth = frac_bw;
    p.time_int_sec = time_int;
    p.gha0_rad = gha0;
    p.dec0_rad = dec0;
    if (precision == OSKAR_SINGLE)
    {
        /* The single-precision kernels select smearing on the rounded
         * values, as the reference's float templates do. */
        p.bandwidth_smearing = ((float) frac_bw != 0.0f);
        p.time_smearing = ((float) time_int != 0.0f);
    }
    {
        const int e = (no_alloc ? oskar_amd_xcorr_launch_no_alloc :
                oskar_amd_xcorr_launch)(precision, matrix, &p, jones,
                I, Q, U, V, l, m, n, a, b, c, st_u, st_v, st_w, st_x, st_y,
                vis, oskar_amd_current_stream());
        if (e && !*status) *status = e;
    }
}

static void correlate(
        int no_alloc,
        int use_casa_phase_convention,
        int source_type,
        int num_sources,
        const oskar_Jones* jones,
        const oskar_Mem* const src_flux[4],
        const oskar_Mem* const src_dir[3],
        const oskar_Mem* const src_ext[3],
        const oskar_Telescope* tel,
        const oskar_Mem* const station_uvw[3],
        double gast,
        double frequency_hz,
        int offset_out,
        oskar_Mem* vis,
        int* status)
{
    double uv_filter_min = 0.0, uv_filter_max = 0.0;
    double time_avg = 0.0, gha0 = 0.0, dec0 = 0.0;
    int i = 0;
    if (*status) return;

    const int num_stations = oskar_telescope_num_stations(tel);
    const int use_extended = (source_type == 1);

    /* Bandwidth-smearing terms. */
    frequency_hz = fabs(frequency_hz);
    const double inv_wavelength = frequency_hz / 299792458.0;
    const double channel_bandwidth = oskar_telescope_channel_bandwidth_hz(tel);
    const double frac_bandwidth = channel_bandwidth / frequency_hz;

    /* Time-average smearing terms (not for drift scans). */
    if (oskar_telescope_phase_centre_coord_type(tel) != OSKAR_COORDS_AZEL)
    {
        time_avg = oskar_telescope_time_average_sec(tel);
        gha0 = gast - oskar_telescope_phase_centre_longitude_rad(tel);
        dec0 = oskar_telescope_phase_centre_latitude_rad(tel);
    }

    /* UV filter in wavelengths. */
    uv_filter_min = oskar_telescope_uv_filter_min(tel);
    uv_filter_max = oskar_telescope_uv_filter_max(tel);
    if (oskar_telescope_uv_filter_units(tel) == OSKAR_METRES)
    {
        uv_filter_min *= inv_wavelength;
        uv_filter_max *= inv_wavelength;
    }
    if (uv_filter_max < 0.0 || uv_filter_max > FLT_MAX)
        uv_filter_max = FLT_MAX;

    /* Locations. */
    const int location = oskar_jones_mem_location(jones);
    if (oskar_telescope_mem_location(tel) != location ||
            oskar_mem_location(vis) != location ||
            oskar_mem_location(station_uvw[0]) != location ||
            oskar_mem_location(station_uvw[1]) != location ||
            oskar_mem_location(station_uvw[2]) != location)
    {
        *status = OSKAR_ERR_LOCATION_MISMATCH;
        return;
    }

    /* Types. */
    const int jones_type = oskar_jones_type(jones);
    const int base_type = oskar_type_precision(jones_type);
    if (oskar_mem_precision(vis) != base_type ||
            oskar_mem_type(station_uvw[0]) != base_type ||
            oskar_mem_type(station_uvw[1]) != base_type ||
            oskar_mem_type(station_uvw[2]) != base_type)
    {
        *status = OSKAR_ERR_TYPE_MISMATCH;
        return;
    }
    if (oskar_mem_type(vis) != jones_type)
    {
        *status = OSKAR_ERR_TYPE_MISMATCH;
        return;
    }

    /* Dimensions. */
    if (oskar_jones_num_sources(jones) < num_sources ||
            (int) oskar_mem_length(station_uvw[0]) != num_stations ||
            (int) oskar_mem_length(station_uvw[1]) != num_stations ||
            (int) oskar_mem_length(station_uvw[2]) != num_stations)
    {
        *status = OSKAR_ERR_DIMENSION_MISMATCH;
        return;
    }

    if (location != OSKAR_GPU)
    {
        /* This library provides the HIP path only. */
        *status = OSKAR_ERR_BAD_LOCATION;
        return;
    }
    if (jones_type != OSKAR_SINGLE_COMPLEX_MATRIX &&
            jones_type != OSKAR_DOUBLE_COMPLEX_MATRIX &&
            jones_type != OSKAR_SINGLE_COMPLEX &&
            jones_type != OSKAR_DOUBLE_COMPLEX)
    {
        *status = OSKAR_ERR_BAD_DATA_TYPE;
        return;
    }
    /* Source arrays must be on the device, of the base type, and hold at
     * least num_sources values (the reference reads them unchecked). */
    {
        const int matrix = oskar_type_is_matrix(jones_type);
        const oskar_Mem* need[10];
        int num_need = 0;
        need[num_need++] = src_flux[0];
        if (matrix)
        {
            need[num_need++] = src_flux[1];
            need[num_need++] = src_flux[2];
            need[num_need++] = src_flux[3];
        }
        for (i = 0; i < 3; ++i) need[num_need++] = src_dir[i];
        if (use_extended)
            for (i = 0; i < 3; ++i) need[num_need++] = src_ext[i];
        for (i = 0; i < num_need; ++i)
        {
            if (!need[i] || oskar_mem_location(need[i]) != location)
            {
                *status = OSKAR_ERR_LOCATION_MISMATCH;
                return;
            }
            if (oskar_mem_type(need[i]) != base_type)
            {
                *status = OSKAR_ERR_TYPE_MISMATCH;
                return;
            }
            if ((int) oskar_mem_length(need[i]) < num_sources)
            {
                *status = OSKAR_ERR_DIMENSION_MISMATCH;
                return;
            }
        }
        if (oskar_mem_length(vis) < (size_t) offset_out +
                (size_t) oskar_telescope_num_baselines(tel))
        {
            *status = OSKAR_ERR_DIMENSION_MISMATCH;
            return;
        }
        {
            const oskar_Mem* x =
                    oskar_telescope_station_true_offset_ecef_metres_const(tel, 0);
            const oskar_Mem* y =
                    oskar_telescope_station_true_offset_ecef_metres_const(tel, 1);
            const oskar_Mem* J = oskar_jones_mem_const(jones);
            launch(base_type, matrix, use_extended, use_casa_phase_convention,
                    num_sources, num_stations, offset_out,
                    oskar_mem_void_const(J),
                    oskar_mem_void_const(src_flux[0]),
                    matrix ? oskar_mem_void_const(src_flux[1]) : 0,
                    matrix ? oskar_mem_void_const(src_flux[2]) : 0,
                    matrix ? oskar_mem_void_const(src_flux[3]) : 0,
                    oskar_mem_void_const(src_dir[0]),
                    oskar_mem_void_const(src_dir[1]),
                    oskar_mem_void_const(src_dir[2]),
                    use_extended ? oskar_mem_void_const(src_ext[0]) : 0,
                    use_extended ? oskar_mem_void_const(src_ext[1]) : 0,
                    use_extended ? oskar_mem_void_const(src_ext[2]) : 0,
                    oskar_mem_void_const(station_uvw[0]),
                    oskar_mem_void_const(station_uvw[1]),
                    oskar_mem_void_const(station_uvw[2]),
                    oskar_mem_void_const(x), oskar_mem_void_const(y),
                    uv_filter_min, uv_filter_max, inv_wavelength,
                    frac_bandwidth, time_avg, gha0, dec0,
                    oskar_mem_void(vis), no_alloc, status);
        }
    }
}

void oskar_cross_correlate(
        int use_casa_phase_convention,
        int source_type,
        int num_sources,
        const oskar_Jones* jones,
        const oskar_Mem* const src_flux[4],
        const oskar_Mem* const src_dir[3],
        const oskar_Mem* const src_ext[3],
        const oskar_Telescope* tel,
        const oskar_Mem* const station_uvw[3],
        double gast,
        double frequency_hz,
        int offset_out,
        oskar_Mem* vis,
        int* status)
{
    correlate(0, use_casa_phase_convention, source_type, num_sources, jones,
            src_flux, src_dir, src_ext, tel, station_uvw, gast, frequency_hz,
            offset_out, vis, status);
}

void oskar_cross_correlate_no_alloc(
        int use_casa_phase_convention,
        int source_type,
        int num_sources,
        const oskar_Jones* jones,
        const oskar_Mem* const src_flux[4],
        const oskar_Mem* const src_dir[3],
        const oskar_Mem* const src_ext[3],
        const oskar_Telescope* tel,
        const oskar_Mem* const station_uvw[3],
        double gast,
        double frequency_hz,
        int offset_out,
        oskar_Mem* vis,
        int* status)
{
    correlate(1, use_casa_phase_convention, source_type, num_sources, jones,
            src_flux, src_dir, src_ext, tel, station_uvw, gast, frequency_hz,
            offset_out, vis, status);
}

/* ---- Backend seam: raw-pointer entry points ------------------------------ */
/* The seam functions return void like their reference counterparts; a
 * launch error is kept for oskar_device_check_error_hip(). */
#define SEAM_STATUS int status = 0;
#define SEAM_END if (status) oskar_amd_set_last_error(status);

void oskar_cross_correlate_point_hip_f(int use_casa, int num_sources,
        int num_stations, int offset_out, const float4c* d_jones,
        const float* d_I, const float* d_Q, const float* d_U,
        const float* d_V, const float* d_l, const float* d_m,
        const float* d_n, const float* d_station_u, const float* d_station_v,
        const float* d_station_w, const float* d_station_x,
        const float* d_station_y, float uv_min_lambda, float uv_max_lambda,
        float inv_wavelength, float frac_bandwidth, float time_int_sec,
        float gha0_rad, float dec0_rad, float4c* d_vis)
{
    SEAM_STATUS
    launch(OSKAR_SINGLE, 1, 0, use_casa, num_sources, num_stations,
            offset_out, d_jones, d_I, d_Q, d_U, d_V, d_l, d_m, d_n, 0, 0, 0,
            d_station_u, d_station_v, d_station_w, d_station_x, d_station_y,
            uv_min_lambda, uv_max_lambda, inv_wavelength, frac_bandwidth,
            time_int_sec, gha0_rad, dec0_rad, d_vis, 0, &status);
    SEAM_END
}

void oskar_cross_correlate_point_hip_d(int use_casa, int num_sources,
        int num_stations, int offset_out, const double4c* d_jones,
        const double* d_I, const double* d_Q, const double* d_U,
        const double* d_V, const double* d_l, const double* d_m,
        const double* d_n, const double* d_station_u,
        const double* d_station_v, const double* d_station_w,
        const double* d_station_x, const double* d_station_y,
        double uv_min_lambda, double uv_max_lambda, double inv_wavelength,
        double frac_bandwidth, double time_int_sec, double gha0_rad,
        double dec0_rad, double4c* d_vis)
{
    SEAM_STATUS
    launch(OSKAR_DOUBLE, 1, 0, use_casa, num_sources, num_stations,
            offset_out, d_jones, d_I, d_Q, d_U, d_V, d_l, d_m, d_n, 0, 0, 0,
            d_station_u, d_station_v, d_station_w, d_station_x, d_station_y,
            uv_min_lambda, uv_max_lambda, inv_wavelength, frac_bandwidth,
            time_int_sec, gha0_rad, dec0_rad, d_vis, 0, &status);
    SEAM_END
}

void oskar_cross_correlate_gaussian_hip_f(int use_casa, int num_sources,
        int num_stations, int offset_out, const float4c* d_jones,
        const float* d_I, const float* d_Q, const float* d_U,
        const float* d_V, const float* d_l, const float* d_m,
        const float* d_n, const float* d_a, const float* d_b,
        const float* d_c, const float* d_station_u, const float* d_station_v,
        const float* d_station_w, const float* d_station_x,
        const float* d_station_y, float uv_min_lambda, float uv_max_lambda,
        float inv_wavelength, float frac_bandwidth, float time_int_sec,
        float gha0_rad, float dec0_rad, float4c* d_vis)
{
    SEAM_STATUS
    launch(OSKAR_SINGLE, 1, 1, use_casa, num_sources, num_stations,
            offset_out, d_jones, d_I, d_Q, d_U, d_V, d_l, d_m, d_n, d_a, d_b,
            d_c, d_station_u, d_station_v, d_station_w, d_station_x,
            d_station_y, uv_min_lambda, uv_max_lambda, inv_wavelength,
            frac_bandwidth, time_int_sec, gha0_rad, dec0_rad, d_vis, 0, &status);
    SEAM_END
}

void oskar_cross_correlate_gaussian_hip_d(int use_casa, int num_sources,
        int num_stations, int offset_out, const double4c* d_jones,
        const double* d_I, const double* d_Q, const double* d_U,
        const double* d_V, const double* d_l, const double* d_m,
        const double* d_n, const double* d_a, const double* d_b,
        const double* d_c, const double* d_station_u,
        const double* d_station_v, const double* d_station_w,
        const double* d_station_x, const double* d_station_y,
        double uv_min_lambda, double uv_max_lambda, double inv_wavelength,
        double frac_bandwidth, double time_int_sec, double gha0_rad,
        double dec0_rad, double4c* d_vis)
{
    SEAM_STATUS
    launch(OSKAR_DOUBLE, 1, 1, use_casa, num_sources, num_stations,
            offset_out, d_jones, d_I, d_Q, d_U, d_V, d_l, d_m, d_n, d_a, d_b,
            d_c, d_station_u, d_station_v, d_station_w, d_station_x,
            d_station_y, uv_min_lambda, uv_max_lambda, inv_wavelength,
            frac_bandwidth, time_int_sec, gha0_rad, dec0_rad, d_vis, 0, &status);
    SEAM_END
}

void oskar_cross_correlate_scalar_point_hip_f(int use_casa, int num_sources,
        int num_stations, int offset_out, const float2* d_jones,
        const float* d_I, const float* d_l, const float* d_m,
        const float* d_n, const float* d_station_u, const float* d_station_v,
        const float* d_station_w, const float* d_station_x,
        const float* d_station_y, float uv_min_lambda, float uv_max_lambda,
        float inv_wavelength, float frac_bandwidth, float time_int_sec,
        float gha0_rad, float dec0_rad, float2* d_vis)
{
    SEAM_STATUS
    launch(OSKAR_SINGLE, 0, 0, use_casa, num_sources, num_stations,
            offset_out, d_jones, d_I, 0, 0, 0, d_l, d_m, d_n, 0, 0, 0,
            d_station_u, d_station_v, d_station_w, d_station_x, d_station_y,
            uv_min_lambda, uv_max_lambda, inv_wavelength, frac_bandwidth,
            time_int_sec, gha0_rad, dec0_rad, d_vis, 0, &status);
    SEAM_END
}

void oskar_cross_correlate_scalar_point_hip_d(int use_casa, int num_sources,
        int num_stations, int offset_out, const double2* d_jones,
        const double* d_I, const double* d_l, const double* d_m,
        const double* d_n, const double* d_station_u,
        const double* d_station_v, const double* d_station_w,
        const double* d_station_x, const double* d_station_y,
        double uv_min_lambda, double uv_max_lambda, double inv_wavelength,
        double frac_bandwidth, double time_int_sec, double gha0_rad,
        double dec0_rad, double2* d_vis)
{
    SEAM_STATUS
    launch(OSKAR_DOUBLE, 0, 0, use_casa, num_sources, num_stations,
            offset_out, d_jones, d_I, 0, 0, 0, d_l, d_m, d_n, 0, 0, 0,
            d_station_u, d_station_v, d_station_w, d_station_x, d_station_y,
            uv_min_lambda, uv_max_lambda, inv_wavelength, frac_bandwidth,
            time_int_sec, gha0_rad, dec0_rad, d_vis, 0, &status);
    SEAM_END
}

void oskar_cross_correlate_scalar_gaussian_hip_f(int use_casa,
        int num_sources, int num_stations, int offset_out,
        const float2* d_jones, const float* d_I, const float* d_l,
        const float* d_m, const float* d_n, const float* d_a,
        const float* d_b, const float* d_c, const float* d_station_u,
        const float* d_station_v, const float* d_station_w,
        const float* d_station_x, const float* d_station_y,
        float uv_min_lambda, float uv_max_lambda, float inv_wavelength,
        float frac_bandwidth, float time_int_sec, float gha0_rad,
        float dec0_rad, float2* d_vis)
{
    SEAM_STATUS
    launch(OSKAR_SINGLE, 0, 1, use_casa, num_sources, num_stations,
            offset_out, d_jones, d_I, 0, 0, 0, d_l, d_m, d_n, d_a, d_b, d_c,
            d_station_u, d_station_v, d_station_w, d_station_x, d_station_y,
            uv_min_lambda, uv_max_lambda, inv_wavelength, frac_bandwidth,
            time_int_sec, gha0_rad, dec0_rad, d_vis, 0, &status);
    SEAM_END
}

void oskar_cross_correlate_scalar_gaussian_hip_d(int use_casa,
        int num_sources, int num_stations, int offset_out,
        const double2* d_jones, const double* d_I, const double* d_l,
        const double* d_m, const double* d_n, const double* d_a,
        const double* d_b, const double* d_c, const double* d_station_u,
        const double* d_station_v, const double* d_station_w,
        const double* d_station_x, const double* d_station_y,
        double uv_min_lambda, double uv_max_lambda, double inv_wavelength,
        double frac_bandwidth, double time_int_sec, double gha0_rad,
        double dec0_rad, double2* d_vis)
{
    SEAM_STATUS
    launch(OSKAR_DOUBLE, 0, 1, use_casa, num_sources, num_stations,
            offset_out, d_jones, d_I, 0, 0, 0, d_l, d_m, d_n, d_a, d_b, d_c,
            d_station_u, d_station_v, d_station_w, d_station_x, d_station_y,
            uv_min_lambda, uv_max_lambda, inv_wavelength, frac_bandwidth,
            time_int_sec, gha0_rad, dec0_rad, d_vis, 0, &status);
    SEAM_END
}
