/*
 * oskar_Telescope: the subset of the telescope model that the correlator
 * reads.
 *
 * Reference: struct oskar/telescope/private_telescope.h:13-68 (fields
 * phase_centre_coord_type, phase_centre_rad, channel_bandwidth_hz :28,
 * time_average_sec :29, uv_filter_* :30-32, num_stations,
 * station_true_offset_ecef_metres :58, station models :54-55); accessors
 * oskar/telescope/oskar_telescope_accessors.h.  Defaults follow
 * oskar/telescope/src/oskar_telescope_create.c (uv_filter_max = FLT_MAX,
 * units metres, phase centre type 0).
 */
#ifndef OSKAR_TELESCOPE_H_
#define OSKAR_TELESCOPE_H_

#include "oskar_mem.h"

#ifdef __cplusplus
extern "C" {
#endif

struct oskar_Telescope;
#ifndef OSKAR_TELESCOPE_TYPEDEF_
#define OSKAR_TELESCOPE_TYPEDEF_
typedef struct oskar_Telescope oskar_Telescope;
#endif
#ifndef OSKAR_STATION_TYPEDEF_
#define OSKAR_STATION_TYPEDEF_
typedef struct oskar_Station oskar_Station;
#endif

OSKAR_EXPORT
oskar_Telescope* oskar_telescope_create(int type, int location,
        int num_stations, int* status);
OSKAR_EXPORT void oskar_telescope_free(oskar_Telescope* telescope,
        int* status);

OSKAR_EXPORT int oskar_telescope_precision(const oskar_Telescope* model);
OSKAR_EXPORT int oskar_telescope_mem_location(const oskar_Telescope* model);
OSKAR_EXPORT int oskar_telescope_num_stations(const oskar_Telescope* model);
OSKAR_EXPORT int oskar_telescope_num_baselines(const oskar_Telescope* model);
OSKAR_EXPORT
int oskar_telescope_phase_centre_coord_type(const oskar_Telescope* model);
OSKAR_EXPORT
double oskar_telescope_phase_centre_longitude_rad(const oskar_Telescope* m);
OSKAR_EXPORT
double oskar_telescope_phase_centre_latitude_rad(const oskar_Telescope* m);
OSKAR_EXPORT
double oskar_telescope_channel_bandwidth_hz(const oskar_Telescope* model);
OSKAR_EXPORT
double oskar_telescope_time_average_sec(const oskar_Telescope* model);
OSKAR_EXPORT double oskar_telescope_uv_filter_min(const oskar_Telescope* m);
OSKAR_EXPORT double oskar_telescope_uv_filter_max(const oskar_Telescope* m);
OSKAR_EXPORT int oskar_telescope_uv_filter_units(const oskar_Telescope* m);
OSKAR_EXPORT
oskar_Mem* oskar_telescope_station_true_offset_ecef_metres(
        oskar_Telescope* model, int dim);
OSKAR_EXPORT
const oskar_Mem* oskar_telescope_station_true_offset_ecef_metres_const(
        const oskar_Telescope* model, int dim);

OSKAR_EXPORT void oskar_telescope_set_channel_bandwidth(
        oskar_Telescope* model, double bandwidth_hz);
OSKAR_EXPORT void oskar_telescope_set_time_average(oskar_Telescope* model,
        double time_average_sec);
OSKAR_EXPORT void oskar_telescope_set_phase_centre(oskar_Telescope* model,
        int coord_type, double longitude_rad, double latitude_rad);
/* Station models (oskar/telescope/src/oskar_telescope_resize_station_array.c,
 * oskar_telescope_accessors.c:132-185): resizing creates empty models in
 * the telescope's precision and location, or frees the surplus ones;
 * oskar_telescope_station() returns NULL past the last model.  The
 * horizon clip reads each model's longitude and latitude. */
OSKAR_EXPORT void oskar_telescope_resize_station_array(
        oskar_Telescope* model, int size, int* status);
OSKAR_EXPORT int oskar_telescope_num_station_models(
        const oskar_Telescope* model);
OSKAR_EXPORT oskar_Station* oskar_telescope_station(oskar_Telescope* model,
        int i);
OSKAR_EXPORT const oskar_Station* oskar_telescope_station_const(
        const oskar_Telescope* model, int i);
/* Station beam duplication (oskar/telescope/src/oskar_telescope_accessors.c
 * :137-140, 187-196, 300-304, 609-621): with duplication allowed, station i
 * uses model type_map[i] and stations of one model share one evaluated
 * beam; otherwise station i uses model i.  The type map (host int array of
 * num_stations entries) starts at 0 for every station;
 * oskar_telescope_set_unique_stations(1) sets type_map[i] = i, (0) clears
 * it. */
OSKAR_EXPORT int oskar_telescope_allow_station_beam_duplication(
        const oskar_Telescope* model);
OSKAR_EXPORT void oskar_telescope_set_allow_station_beam_duplication(
        oskar_Telescope* model, int value);
OSKAR_EXPORT oskar_Mem* oskar_telescope_station_type_map(
        oskar_Telescope* model);
OSKAR_EXPORT const oskar_Mem* oskar_telescope_station_type_map_const(
        const oskar_Telescope* model);
OSKAR_EXPORT void oskar_telescope_set_unique_stations(oskar_Telescope* model,
        int value, int* status);
/* units: "Metres" or "Wavelengths" (case-insensitive on the first letter),
 * oskar/telescope/src/oskar_telescope_accessors.c. */
OSKAR_EXPORT void oskar_telescope_set_uv_filter(oskar_Telescope* model,
        double uv_filter_min, double uv_filter_max, const char* units,
        int* status);

#ifdef __cplusplus
}
#endif

#endif /* OSKAR_TELESCOPE_H_ */
