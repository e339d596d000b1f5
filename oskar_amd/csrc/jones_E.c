/*
 * oskar_evaluate_jones_E(): the station beam of every station into one
 * Jones array (oskar/interferometer/src/oskar_evaluate_jones_E.c:88-192):
 * one oskar_station_beam() per station model, or, with station beam
 * duplication, one per distinct model of the type map and a device copy for
 * every later station of that model.
 */
#include <stdlib.h>

#include "private.h"
#include "oskar_evaluate_jones_E.h"

void oskar_evaluate_jones_E(oskar_Jones* E, int coord_type, int num_points,
        const oskar_Mem* const source_coords[3], double ref_lon_rad,
        double ref_lat_rad, oskar_Telescope* tel, int time_index,
        double time_start_mjd_utc, double time_mjd_utc, double frequency_hz,
        oskar_StationWork* work, int* status)
{
    int i = 0, j = 0;
    if (*status) return;
    const int num_stations = oskar_telescope_num_stations(tel);
    const int num_sources = oskar_jones_num_sources(E);
    if (num_stations == 0)
    {
        *status = OSKAR_ERR_MEMORY_NOT_ALLOCATED;
        return;
    }
    if (num_stations != oskar_jones_num_stations(E))
    {
        *status = OSKAR_ERR_DIMENSION_MISMATCH;
        return;
    }
    if (num_points > num_sources || num_points < 0)
    {
        *status = OSKAR_ERR_DIMENSION_MISMATCH;
        return;
    }
    oskar_Mem* out = oskar_jones_mem(E);
    const int norm_type = oskar_telescope_phase_centre_coord_type(tel);
    const double norm_lon = oskar_telescope_phase_centre_longitude_rad(tel);
    const double norm_lat = oskar_telescope_phase_centre_latitude_rad(tel);
    if (!oskar_telescope_allow_station_beam_duplication(tel))
    {
        if (oskar_telescope_num_station_models(tel) < num_stations)
        {
            *status = OSKAR_ERR_OUT_OF_RANGE;
            return;
        }
        for (i = 0; i < num_stations && !*status; ++i)
            oskar_station_beam(oskar_telescope_station(tel, i), work,
                    coord_type, num_points, source_coords, ref_lon_rad,
                    ref_lat_rad, norm_type, norm_lon, norm_lat, time_index,
                    time_start_mjd_utc, time_mjd_utc, frequency_hz,
                    i * num_sources, out, status);
        return;
    }
    {
        /* Models evaluated so far and the first station that holds each. */
        const int* type_map = (const int*) oskar_mem_void_const(
                oskar_telescope_station_type_map_const(tel));
        int* model_done = (int*) malloc((size_t) num_stations * sizeof(int));
        int* model_row = (int*) malloc((size_t) num_stations * sizeof(int));
        int num_done = 0;
        if (!model_done || !model_row)
        {
            free(model_done);
            free(model_row);
            *status = OSKAR_ERR_MEMORY_ALLOC_FAILURE;
            return;
        }
        for (i = 0; i < num_stations && !*status; ++i)
        {
            const int model = type_map[i];
            int row = -1;
            if (model < 0 || model >= oskar_telescope_num_station_models(tel))
            {
                *status = OSKAR_ERR_OUT_OF_RANGE;
                break;
            }
            for (j = 0; j < num_done; ++j)
                if (model_done[j] == model) { row = model_row[j]; break; }
            if (row >= 0)
                oskar_mem_copy_contents(out, out,
                        (size_t) i * (size_t) num_sources,
                        (size_t) row * (size_t) num_sources,
                        (size_t) num_sources, status);
            else
            {
                oskar_station_beam(oskar_telescope_station(tel, model), work,
                        coord_type, num_points, source_coords, ref_lon_rad,
                        ref_lat_rad, norm_type, norm_lon, norm_lat,
                        time_index, time_start_mjd_utc, time_mjd_utc,
                        frequency_hz, i * num_sources, out, status);
                model_done[num_done] = model;
                model_row[num_done++] = i;
            }
        }
        free(model_done);
        free(model_row);
    }
}
