/*
 * Station beams of every station: oskar/interferometer/oskar_evaluate_jones_E.h
 * (src/oskar_evaluate_jones_E.c:88-192).
 */
#ifndef OSKAR_EVALUATE_JONES_E_H_
#define OSKAR_EVALUATE_JONES_E_H_

#include "oskar_jones.h"
#include "oskar_station.h"
#include "oskar_telescope.h"

#ifdef __cplusplus
extern "C" {
#endif

/* E[station i][source s] = the beam of station i for source s
 * (oskar_station_beam of the station's model, normalised to its phase
 * centre when the model says so).  Without station beam duplication,
 * station i uses model i; with it, station i uses model type_map[i] and
 * each model is evaluated once, later stations of the model receiving a
 * copy.  Sources: num_points of source_coords in coord_type (RADEC, AZEL,
 * ENU_DIR, REL_DIR about ref_lon/lat).  Errors as the reference:
 * OSKAR_ERR_MEMORY_NOT_ALLOCATED with no stations,
 * OSKAR_ERR_DIMENSION_MISMATCH if E's station count differs from the
 * telescope's; here also OSKAR_ERR_OUT_OF_RANGE when a station has no
 * model (the reference reads past its array) and
 * OSKAR_ERR_DIMENSION_MISMATCH when num_points exceeds E's sources. */
OSKAR_EXPORT
void oskar_evaluate_jones_E(oskar_Jones* E, int coord_type, int num_points,
        const oskar_Mem* const source_coords[3], double ref_lon_rad,
        double ref_lat_rad, oskar_Telescope* tel, int time_index,
        double time_start_mjd_utc, double time_mjd_utc, double frequency_hz,
        oskar_StationWork* work, int* status);

#ifdef __cplusplus
}
#endif

#endif /* OSKAR_EVALUATE_JONES_E_H_ */
