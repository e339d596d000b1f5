/*
 * oskar_station_beam() and the coordinate conversions it needs.
 *
 * oskar/telescope/station/src/oskar_station_beam.c:37-251: source
 * directions in any supported frame are converted to ENU direction cosines
 * at the station (local sidereal time from GAST and the station's
 * longitude), the beam is evaluated for them and, when the station
 * normalises its final beam, for one more direction (the normalisation
 * point), by which all values are then divided; the result is copied to
 * beam[offset_out ..].  Station types: aperture array (the exact beam of
 * oskar_evaluate_station_beam_aperture_array) and isotropic (1 on the
 * diagonal).  The Gaussian and VLA primary-beam types and the ionospheric
 * screen are not on the visibility hot path and return
 * OSKAR_ERR_FUNCTION_NOT_AVAILABLE.
 *
 * Sidereal time: the IAU 1982 expression for GMST from UT1
 * (oskar/convert/src/oskar_convert_mjd_to_gmst.c) plus the low-precision
 * equation of the equinoxes of the Astronomical Almanac
 * (oskar/convert/src/oskar_equation_of_equinoxes_fast.c).
 */
#include <math.h>
#include <stdio.h>

#include "private.h"
#include "oskar_convert.h"
#include "oskar_station.h"

#define TWO_PI 6.283185307179586476925
#define DEG_TO_RAD 0.017453292519943295769
#define HOUR_TO_RAD 0.26179938779914943654
#define SEC_TO_RAD 7.2722052166430399038e-5   /* 2 pi / 86400 */

double oskar_convert_mjd_to_gmst(double mjd)
{
    const double t = (mjd - 51544.5) / 36525.0;    /* centuries since J2000 */
    /* GMST at 0h UT1 (seconds) + the rotation since 0h UT1. */
    const double g0 = 24110.54841 + t * (8640184.812866 +
            t * (0.093104 - 6.2e-6 * t));
    double x = fmod(fmod(mjd, 1.0) * TWO_PI + g0 * SEC_TO_RAD, TWO_PI);
    if (x < 0.0) x += TWO_PI;
    return x;
}

double oskar_equation_of_equinoxes_fast(double mjd)
{
    const double d = mjd - 51544.5;                 /* days since J2000 */
    const double omega = (125.04 - 0.052954 * d) * DEG_TO_RAD;
    const double sun_l = (280.47 + 0.98565 * d) * DEG_TO_RAD;
    const double eps = (23.4393 - 0.0000004 * d) * DEG_TO_RAD;
    /* Nutation in longitude (hours), times cos(obliquity). */
    const double dpsi = -0.000319 * sin(omega) - 0.000024 * sin(2.0 * sun_l);
    return dpsi * cos(eps) * HOUR_TO_RAD;
}

double oskar_convert_mjd_to_gast_fast(double mjd)
{
    return oskar_convert_mjd_to_gmst(mjd) +
            oskar_equation_of_equinoxes_fast(mjd);
}

void oskar_convert_any_to_enu_directions(int coord_type, int num_points,
        const oskar_Mem* const coords[3], double ref_lon_rad,
        double ref_lat_rad, double lst_rad, double lat_rad, oskar_Mem* enu[3],
        int* status)
{
    int i = 0;
    double p[6] = {0, 0, 0, 0, 0, 0};
    if (*status || num_points <= 0) return;
    for (i = 0; i < 3; ++i)
    {
        oskar_mem_ensure(enu[i], (size_t) num_points, status);
        if (!*status && oskar_mem_location(enu[i]) != OSKAR_GPU)
            *status = OSKAR_ERR_BAD_LOCATION;
    }
    if (*status) return;
    if (coord_type == OSKAR_COORDS_ENU_DIR)
    {
        for (i = 0; i < 3; ++i)
            oskar_mem_copy_contents(enu[i], coords[i], 0, 0,
                    (size_t) num_points, status);
        return;
    }
    if (coord_type != OSKAR_COORDS_RADEC && coord_type != OSKAR_COORDS_AZEL
            && coord_type != OSKAR_COORDS_REL_DIR)
    {
        *status = OSKAR_ERR_FUNCTION_NOT_AVAILABLE;
        return;
    }
    for (i = 0; i < (coord_type == OSKAR_COORDS_REL_DIR ? 3 : 2); ++i)
    {
        if (oskar_mem_location(coords[i]) != OSKAR_GPU)
        {
            *status = OSKAR_ERR_BAD_LOCATION;
            return;
        }
        if (oskar_mem_type(coords[i]) != oskar_mem_type(enu[0]))
        {
            *status = OSKAR_ERR_TYPE_MISMATCH;
            return;
        }
        if ((int) oskar_mem_length(coords[i]) < num_points)
        {
            *status = OSKAR_ERR_DIMENSION_MISMATCH;
            return;
        }
    }
    if (coord_type == OSKAR_COORDS_RADEC)
    {
        p[0] = lst_rad;
        p[1] = sin(lat_rad);
        p[2] = cos(lat_rad);
    }
    else if (coord_type == OSKAR_COORDS_REL_DIR)
    {
        const double ha0 = lst_rad - ref_lon_rad;
        p[0] = sin(ha0);
        p[1] = cos(ha0);
        p[2] = sin(ref_lat_rad);
        p[3] = cos(ref_lat_rad);
        p[4] = sin(lat_rad);
        p[5] = cos(lat_rad);
    }
    *status = oskar_amd_enu_convert_launch(oskar_mem_precision(enu[0]),
            coord_type, num_points, oskar_mem_void_const(coords[0]),
            oskar_mem_void_const(coords[1]),
            coord_type == OSKAR_COORDS_REL_DIR ?
            oskar_mem_void_const(coords[2]) : 0, p,
            oskar_mem_void(enu[0]), oskar_mem_void(enu[1]),
            oskar_mem_void(enu[2]), oskar_amd_current_stream());
}

/* One direction (lon, lat) of a frame to ENU on the host. */
static void one_direction_to_enu(int coord_type, double lon, double lat,
        double lst_rad, double lat_rad, double* x, double* y, double* z,
        int* status)
{
    if (coord_type == OSKAR_COORDS_RADEC)
    {
        const double ha = lst_rad - lon;
        *x = -cos(lat) * sin(ha);
        *y = cos(lat_rad) * sin(lat) - sin(lat_rad) * cos(lat) * cos(ha);
        *z = sin(lat_rad) * sin(lat) + cos(lat_rad) * cos(lat) * cos(ha);
    }
    else if (coord_type == OSKAR_COORDS_AZEL)
    {
        *x = cos(lat) * sin(lon);
        *y = cos(lat) * cos(lon);
        *z = sin(lat);
    }
    else *status = OSKAR_ERR_FUNCTION_NOT_AVAILABLE;
}

void oskar_station_beam(oskar_Station* station, oskar_StationWork* work,
        int source_coord_type, int num_points,
        const oskar_Mem* const source_coords[3], double ref_lon_rad,
        double ref_lat_rad, int norm_coord_type, double norm_lon_rad,
        double norm_lat_rad, int time_index, double time_start_mjd_utc,
        double time_mjd_utc, double frequency_hz, int offset_out,
        oskar_Mem* beam, int* status)
{
    double bx = 0, by = 0, bz = 0;
    int i = 0, n = num_points;
    (void) time_start_mjd_utc;
    if (*status) return;
    if (num_points < 0 || offset_out < 0)
    {
        *status = OSKAR_ERR_INVALID_ARGUMENT;
        return;
    }
    if (oskar_mem_location(beam) != OSKAR_GPU ||
            oskar_station_mem_location(station) != OSKAR_GPU)
    {
        *status = OSKAR_ERR_BAD_LOCATION;
        return;
    }
    if (!oskar_mem_is_complex(beam) ||
            oskar_mem_precision(beam) != oskar_station_precision(station) ||
            oskar_station_precision(station) != work->precision)
    {
        *status = OSKAR_ERR_TYPE_MISMATCH;
        return;
    }
    const int type = oskar_station_type(station);
    if (type != OSKAR_STATION_TYPE_AA && type != OSKAR_STATION_TYPE_ISOTROPIC)
    {
        *status = OSKAR_ERR_FUNCTION_NOT_AVAILABLE;
        return;
    }
    const double gast_rad = oskar_convert_mjd_to_gast_fast(time_mjd_utc);
    const double lat_rad = oskar_station_lat_rad(station);
    const double lst_rad = gast_rad + oskar_station_lon_rad(station);

    /* The beam below the horizon gives zeros
     * (oskar_station_beam.c:85-93). */
    one_direction_to_enu(oskar_station_beam_coord_type(station),
            oskar_station_beam_lon_rad(station),
            oskar_station_beam_lat_rad(station), lst_rad, lat_rad,
            &bx, &by, &bz, status);
    if (*status) return;
    oskar_mem_ensure(beam, (size_t) offset_out + (size_t) num_points, status);
    if (*status) return;
    if (atan2(bz, sqrt(bx * bx + by * by)) < 0.0)
    {
        oskar_mem_set_value_real(beam, 0.0, (size_t) offset_out,
                (size_t) num_points, status);
        fprintf(stderr, "oskar_station_beam: beam below horizon at time "
                "index %d\n", time_index);
        return;
    }
    if (num_points == 0) return;

    /* Source directions, then the normalisation point at index n (for
     * directions given as ENU already, one kernel copies them and writes
     * the point). */
    const int normalise = oskar_station_normalise_final_beam(station) &&
            type != OSKAR_STATION_TYPE_ISOTROPIC;
    double nx = 0, ny = 0, nz = 0;
    if (normalise)
        one_direction_to_enu(norm_coord_type, norm_lon_rad, norm_lat_rad,
                lst_rad, lat_rad, &nx, &ny, &nz, status);
    int enu_on_device = (source_coord_type == OSKAR_COORDS_ENU_DIR);
    for (i = 0; i < 3 && enu_on_device; ++i)
        enu_on_device = oskar_mem_location(source_coords[i]) == OSKAR_GPU &&
                oskar_mem_type(source_coords[i]) ==
                oskar_mem_type(work->enu[i]) &&
                oskar_mem_length(source_coords[i]) >= (size_t) n;
    if (*status) return;
    if (type == OSKAR_STATION_TYPE_AA &&
            oskar_amd_station_beam_aa_fused_covers(station, beam))
    {
        /* A flat station of one dipole type: one kernel forms the array
         * factor, the element pattern, the blank and the normalisation and
         * writes beam (directions read where they are when they are ENU
         * on the device, with no normalisation point appended). */
        const oskar_Mem* d[3];
        for (i = 0; i < 3; ++i) d[i] = source_coords[i];
        if (!enu_on_device)
        {
            oskar_convert_any_to_enu_directions(source_coord_type,
                    num_points, source_coords, ref_lon_rad, ref_lat_rad,
                    lst_rad, lat_rad, work->enu, status);
            for (i = 0; i < 3; ++i) d[i] = work->enu[i];
        }
        oskar_amd_station_beam_aa_fused(station, work, num_points, d[0],
                d[1], d[2], normalise, nx, ny, nz, time_index, gast_rad,
                frequency_hz, offset_out, beam, status);
        return;
    }
    if (enu_on_device)
    {
        for (i = 0; i < 3; ++i)
            oskar_mem_ensure(work->enu[i], (size_t) n + 1, status);
        if (!*status)
            *status = oskar_amd_hip_status(oskar_amd_copy3_point_launch(
                    oskar_mem_precision(work->enu[0]), (unsigned int) n,
                    oskar_mem_void_const(source_coords[0]),
                    oskar_mem_void_const(source_coords[1]),
                    oskar_mem_void_const(source_coords[2]),
                    oskar_mem_void(work->enu[0]),
                    oskar_mem_void(work->enu[1]),
                    oskar_mem_void(work->enu[2]), normalise, nx, ny, nz,
                    oskar_amd_current_stream()));
    }
    else
    {
        oskar_convert_any_to_enu_directions(source_coord_type, num_points,
                source_coords, ref_lon_rad, ref_lat_rad, lst_rad, lat_rad,
                work->enu, status);
        if (normalise)
        {
            for (i = 0; i < 3; ++i)
                oskar_mem_ensure(work->enu[i], (size_t) n + 1, status);
            if (!*status)
                *status = oskar_amd_hip_status(oskar_amd_set_point3_launch(
                        oskar_mem_precision(work->enu[0]),
                        oskar_mem_void(work->enu[0]),
                        oskar_mem_void(work->enu[1]),
                        oskar_mem_void(work->enu[2]), (size_t) n, nx, ny,
                        nz, oskar_amd_current_stream()));
        }
    }
    if (normalise) n++;
    if (*status) return;
    if (!work->beam_out ||
            oskar_mem_type(work->beam_out) != oskar_mem_type(beam))
    {
        oskar_mem_free(work->beam_out, status);
        work->beam_out = oskar_mem_create(oskar_mem_type(beam), OSKAR_GPU,
                (size_t) n, status);
    }
    oskar_mem_ensure(work->beam_out, (size_t) n, status);
    if (*status) return;
    /* The horizon blank of a flat aperture-array station (a hierarchical
     * station's top level has none: its children blank their beams) is
     * done by the kernel that copies the beam out, with the normalisation
     * by the beam at the normalisation point (the last one) as
     * oskar_mem_normalise does it. */
    const oskar_Mem* mask = 0;
    if (type == OSKAR_STATION_TYPE_ISOTROPIC)
        oskar_mem_set_value_real(work->beam_out, 1.0, 0, (size_t) n, status);
    else
    {
        oskar_amd_station_beam_aa_unblanked(station, work, n, work->enu[0],
                work->enu[1], work->enu[2], time_index, gast_rad,
                frequency_hz, work->beam_out, status);
        if (!station->child) mask = work->enu[2];
    }
    if (*status) return;
    {
        const size_t es = oskar_mem_element_size(oskar_mem_type(beam));
        *status = oskar_amd_hip_status(oskar_amd_beam_finish_launch(
                oskar_mem_type(beam), (unsigned int) num_points,
                oskar_mem_void_const(work->beam_out), (unsigned int) (n - 1),
                normalise, mask ? oskar_mem_void_const(mask) : 0,
                (char*) oskar_mem_void(beam) + (size_t) offset_out * es,
                oskar_amd_current_stream()));
    }
}
