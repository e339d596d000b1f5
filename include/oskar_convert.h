/*
 * Coordinate conversions used by the station beam.
 *
 * oskar/convert/oskar_convert_mjd_to_gmst.h, oskar_equation_of_equinoxes_fast.h,
 * oskar_convert_mjd_to_gast_fast.h (sidereal time from a UT1 MJD: the IAU
 * 1982 GMST expression plus the low-precision equation of the equinoxes)
 * and oskar_convert_any_to_enu_directions.h (source directions in one of
 * the supported frames to direction cosines in the local horizon frame,
 * x east, y north, z up).
 */
#ifndef OSKAR_CONVERT_H_
#define OSKAR_CONVERT_H_

#include "oskar_mem.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Radians, in [0, 2 pi). */
OSKAR_EXPORT double oskar_convert_mjd_to_gmst(double mjd);

/* Radians (apparent minus mean sidereal time). */
OSKAR_EXPORT double oskar_equation_of_equinoxes_fast(double mjd);

/* GMST + equation of the equinoxes, radians. */
OSKAR_EXPORT double oskar_convert_mjd_to_gast_fast(double mjd);

/* coord_type OSKAR_COORDS_RADEC (coords = apparent RA, Dec), _AZEL (az, el),
 * _ENU_DIR (copied), _REL_DIR (l, m, n about (ref_lon_rad, ref_lat_rad));
 * lst_rad and lat_rad: local sidereal time and latitude.  enu (GPU, grown
 * as needed) receives num_points direction cosines.  Other frames return
 * OSKAR_ERR_FUNCTION_NOT_AVAILABLE. */
OSKAR_EXPORT
void oskar_convert_any_to_enu_directions(int coord_type, int num_points,
        const oskar_Mem* const coords[3], double ref_lon_rad,
        double ref_lat_rad, double lst_rad, double lat_rad,
        oskar_Mem* enu[3], int* status);

#ifdef __cplusplus
}
#endif

#endif /* OSKAR_CONVERT_H_ */
