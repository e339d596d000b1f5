/*
 * oskar_Jones and oskar_Telescope (correlator subset).
 *
 * Reference: oskar/interferometer/src/oskar_jones_create.c,
 * oskar_jones_set_size.c, oskar_jones_accessors.c, oskar_jones_join.c:13-41;
 * oskar/telescope/src/oskar_telescope_create.c, oskar_telescope_accessors.c.
 */
#include <float.h>
#include <stdlib.h>

#include "private.h"
#include "oskar_evaluate_jones_K.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ---- Jones ---------------------------------------------------------------- */
oskar_Jones* oskar_jones_create(int type, int location, int num_stations,
        int num_sources, int* status)
{
    oskar_Jones* j = 0;
    if (*status) return 0;
    if (!oskar_type_is_complex(type) ||
            (oskar_type_precision(type) != OSKAR_SINGLE &&
             oskar_type_precision(type) != OSKAR_DOUBLE))
    {
        *status = OSKAR_ERR_BAD_DATA_TYPE;
        return 0;
    }
    j = (oskar_Jones*) calloc(1, sizeof(oskar_Jones));
    if (!j)
    {
        *status = OSKAR_ERR_MEMORY_ALLOC_FAILURE;
        return 0;
    }
    j->num_stations = j->cap_stations = num_stations;
    j->num_sources = j->cap_sources = num_sources;
    j->data = oskar_mem_create(type, location,
            (size_t) num_stations * (size_t) num_sources, status);
    return j;
}

oskar_Jones* oskar_jones_create_copy(const oskar_Jones* src, int location,
        int* status)
{
    oskar_Jones* j = 0;
    if (*status || !src) return 0;
    j = oskar_jones_create(src->data->type, location, src->num_stations,
            src->num_sources, status);
    if (j) oskar_mem_copy_contents(j->data, src->data, 0, 0,
            (size_t) src->num_stations * (size_t) src->num_sources, status);
    return j;
}

void oskar_jones_free(oskar_Jones* jones, int* status)
{
    if (!jones) return;
    oskar_mem_free(jones->data, status);
    free(jones);
}

void oskar_jones_set_size(oskar_Jones* jones, int num_stations,
        int num_sources, int* status)
{
    if (*status || !jones) return;
    {
        const size_t need = (size_t) num_stations * (size_t) num_sources;
        if (need > (size_t) jones->cap_stations * (size_t) jones->cap_sources)
        {
            oskar_mem_realloc(jones->data, need, status);
            jones->cap_stations = num_stations;
            jones->cap_sources = num_sources;
        }
    }
    jones->num_stations = num_stations;
    jones->num_sources = num_sources;
}

int oskar_jones_num_sources(const oskar_Jones* j) { return j->num_sources; }
int oskar_jones_num_stations(const oskar_Jones* j) { return j->num_stations; }
int oskar_jones_type(const oskar_Jones* j) { return oskar_mem_type(j->data); }
int oskar_jones_mem_location(const oskar_Jones* j)
{ return oskar_mem_location(j->data); }
oskar_Mem* oskar_jones_mem(oskar_Jones* j) { return j->data; }
const oskar_Mem* oskar_jones_mem_const(const oskar_Jones* j) { return j->data; }

void oskar_jones_join(oskar_Jones* j3, oskar_Jones* j1, const oskar_Jones* j2,
        int* status)
{
    if (*status) return;
    if (!j3) j3 = j1;
    if (j1->num_sources != j2->num_sources ||
            j1->num_stations != j2->num_stations ||
            j3->num_sources != j1->num_sources ||
            j3->num_stations != j1->num_stations)
    {
        *status = OSKAR_ERR_DIMENSION_MISMATCH;
        return;
    }
    if (oskar_jones_mem_location(j1) != OSKAR_GPU ||
            oskar_jones_mem_location(j2) != OSKAR_GPU ||
            oskar_jones_mem_location(j3) != OSKAR_GPU)
    {
        *status = OSKAR_ERR_BAD_LOCATION;
        return;
    }
    if (oskar_type_precision(oskar_jones_type(j1)) !=
            oskar_type_precision(oskar_jones_type(j2)))
    {
        *status = OSKAR_ERR_TYPE_MISMATCH;
        return;
    }
    {
        const int t1 = oskar_jones_type(j1), t2 = oskar_jones_type(j2);
        const int a_matrix = oskar_type_is_matrix(t1);
        const int b_matrix = oskar_type_is_matrix(t2);
        const int out_type = oskar_type_precision(t1) | OSKAR_COMPLEX |
                ((a_matrix || b_matrix) ? OSKAR_MATRIX : 0);
        if (oskar_jones_type(j3) != out_type)
        {
            *status = OSKAR_ERR_TYPE_MISMATCH;
            return;
        }
        *status = oskar_amd_hip_status(oskar_amd_jones_join_launch(
                oskar_type_precision(t1), a_matrix, b_matrix,
                (size_t) j1->num_stations * (size_t) j1->num_sources,
                j3->data->data, j1->data->data, j2->data->data,
                oskar_amd_current_stream()));
    }
}

/* oskar/interferometer/src/oskar_evaluate_jones_K.c:19-157 */
void oskar_evaluate_jones_K(oskar_Jones* K, int num_sources,
        const oskar_Mem* l, const oskar_Mem* m, const oskar_Mem* n,
        const oskar_Mem* u, const oskar_Mem* v, const oskar_Mem* w,
        double frequency_hz, const oskar_Mem* source_filter,
        double source_filter_min, double source_filter_max,
        int ignore_w_components, int* status)
{
    if (*status) return;
    const int type = oskar_jones_type(K);
    const int precision = oskar_type_precision(type);
    const int location = oskar_jones_mem_location(K);
    const int num_stations = oskar_jones_num_stations(K);
    const double wavenumber = 2.0 * M_PI * frequency_hz / 299792458.0;
    if (oskar_mem_location(l) != location ||
            oskar_mem_location(m) != location ||
            oskar_mem_location(n) != location ||
            oskar_mem_location(source_filter) != location ||
            oskar_mem_location(u) != location ||
            oskar_mem_location(v) != location ||
            oskar_mem_location(w) != location)
    {
        *status = OSKAR_ERR_LOCATION_MISMATCH;
        return;
    }
    if (precision != oskar_mem_type(l) || precision != oskar_mem_type(m) ||
            precision != oskar_mem_type(n) || precision != oskar_mem_type(u) ||
            precision != oskar_mem_type(v) || precision != oskar_mem_type(w) ||
            precision != oskar_mem_type(source_filter))
    {
        *status = OSKAR_ERR_TYPE_MISMATCH;
        return;
    }
    if (type != OSKAR_SINGLE_COMPLEX && type != OSKAR_DOUBLE_COMPLEX)
    {
        *status = OSKAR_ERR_BAD_DATA_TYPE;
        return;
    }
    if (oskar_jones_num_sources(K) < num_sources ||
            (int) oskar_mem_length(l) < num_sources ||
            (int) oskar_mem_length(m) < num_sources ||
            (int) oskar_mem_length(n) < num_sources ||
            (int) oskar_mem_length(source_filter) < num_sources ||
            (int) oskar_mem_length(u) < num_stations ||
            (int) oskar_mem_length(v) < num_stations ||
            (int) oskar_mem_length(w) < num_stations)
    {
        *status = OSKAR_ERR_DIMENSION_MISMATCH;
        return;
    }
    if (location != OSKAR_GPU)
    {
        *status = OSKAR_ERR_BAD_LOCATION;
        return;
    }
    /* The kernel writes with the Jones source stride. */
    if (oskar_jones_num_sources(K) != num_sources)
    {
        *status = OSKAR_ERR_DIMENSION_MISMATCH;
        return;
    }
    *status = oskar_amd_hip_status(oskar_amd_jones_K_launch(precision,
            num_sources, oskar_mem_void_const(l), oskar_mem_void_const(m),
            oskar_mem_void_const(n), num_stations, oskar_mem_void_const(u),
            oskar_mem_void_const(v), oskar_mem_void_const(w), wavenumber,
            oskar_mem_void_const(source_filter), source_filter_min,
            source_filter_max, ignore_w_components,
            oskar_mem_void(K->data), oskar_amd_current_stream()));
}

/* ---- Telescope ------------------------------------------------------------ */
oskar_Telescope* oskar_telescope_create(int type, int location,
        int num_stations, int* status)
{
    int i = 0;
    oskar_Telescope* t = 0;
    if (*status) return 0;
    if (type != OSKAR_SINGLE && type != OSKAR_DOUBLE)
    {
        *status = OSKAR_ERR_BAD_DATA_TYPE;
        return 0;
    }
    t = (oskar_Telescope*) calloc(1, sizeof(oskar_Telescope));
    if (!t)
    {
        *status = OSKAR_ERR_MEMORY_ALLOC_FAILURE;
        return 0;
    }
    t->precision = type;
    t->mem_location = location;
    t->num_stations = num_stations;
    t->uv_filter_max = FLT_MAX;
    t->uv_filter_units = OSKAR_METRES;
    for (i = 0; i < 3; ++i)
        t->station_true_offset_ecef_metres[i] =
                oskar_mem_create(type, location, (size_t) num_stations, status);
    /* Every station of type 0 (oskar_telescope_create.c:56-58). */
    t->station_type_map = oskar_mem_create(OSKAR_INT, OSKAR_CPU,
            (size_t) num_stations, status);
    oskar_mem_clear_contents(t->station_type_map, status);
    return t;
}

/* oskar/telescope/src/oskar_telescope_accessors.c:137-140, 187-196,
 * 300-304, 609-621. */
int oskar_telescope_allow_station_beam_duplication(const oskar_Telescope* t)
{ return t->allow_station_beam_duplication; }

void oskar_telescope_set_allow_station_beam_duplication(oskar_Telescope* t,
        int value)
{ t->allow_station_beam_duplication = value; }

oskar_Mem* oskar_telescope_station_type_map(oskar_Telescope* t)
{ return t->station_type_map; }

const oskar_Mem* oskar_telescope_station_type_map_const(
        const oskar_Telescope* t)
{ return t->station_type_map; }

void oskar_telescope_set_unique_stations(oskar_Telescope* t, int value,
        int* status)
{
    int i = 0;
    if (*status) return;
    if (value)
    {
        int* map = (int*) oskar_mem_void(t->station_type_map);
        for (i = 0; i < t->num_stations; ++i) map[i] = i;
    }
    else oskar_mem_clear_contents(t->station_type_map, status);
}

void oskar_telescope_free(oskar_Telescope* t, int* status)
{
    int i = 0;
    if (!t) return;
    for (i = 0; i < 3; ++i)
        oskar_mem_free(t->station_true_offset_ecef_metres[i], status);
    oskar_mem_free(t->station_type_map, status);
    for (i = 0; i < t->num_station_models; ++i)
        oskar_station_free(t->station[i], status);
    free(t->station);
    free(t);
}

/* oskar/telescope/src/oskar_telescope_resize_station_array.c: new station
 * models are created empty, in the telescope's precision and location;
 * shrinking frees the surplus ones. */
void oskar_telescope_resize_station_array(oskar_Telescope* t, int size,
        int* status)
{
    int i = 0;
    if (*status) return;
    if (size < 0)
    {
        *status = OSKAR_ERR_INVALID_ARGUMENT;
        return;
    }
    if (size == t->num_station_models) return;
    for (i = size; i < t->num_station_models; ++i)
        oskar_station_free(t->station[i], status);
    if (size > 0)
    {
        oskar_Station** s = (oskar_Station**) realloc(t->station,
                (size_t) size * sizeof(oskar_Station*));
        if (!s)
        {
            *status = OSKAR_ERR_MEMORY_ALLOC_FAILURE;
            return;
        }
        t->station = s;
        for (i = t->num_station_models; i < size; ++i)
            t->station[i] = oskar_station_create(t->precision,
                    t->mem_location, 0, status);
    }
    else
    {
        free(t->station);
        t->station = 0;
    }
    t->num_station_models = size;
}

int oskar_telescope_num_station_models(const oskar_Telescope* t)
{ return t->num_station_models; }

oskar_Station* oskar_telescope_station(oskar_Telescope* t, int i)
{ return (i >= 0 && i < t->num_station_models) ? t->station[i] : 0; }

const oskar_Station* oskar_telescope_station_const(const oskar_Telescope* t,
        int i)
{ return (i >= 0 && i < t->num_station_models) ? t->station[i] : 0; }

int oskar_telescope_precision(const oskar_Telescope* m) { return m->precision; }
int oskar_telescope_mem_location(const oskar_Telescope* m)
{ return m->mem_location; }
int oskar_telescope_num_stations(const oskar_Telescope* m)
{ return m->num_stations; }
int oskar_telescope_num_baselines(const oskar_Telescope* m)
{ return (m->num_stations * (m->num_stations - 1)) / 2; }
int oskar_telescope_phase_centre_coord_type(const oskar_Telescope* m)
{ return m->phase_centre_coord_type; }
double oskar_telescope_phase_centre_longitude_rad(const oskar_Telescope* m)
{ return m->phase_centre_rad[0]; }
double oskar_telescope_phase_centre_latitude_rad(const oskar_Telescope* m)
{ return m->phase_centre_rad[1]; }
double oskar_telescope_channel_bandwidth_hz(const oskar_Telescope* m)
{ return m->channel_bandwidth_hz; }
double oskar_telescope_time_average_sec(const oskar_Telescope* m)
{ return m->time_average_sec; }
double oskar_telescope_uv_filter_min(const oskar_Telescope* m)
{ return m->uv_filter_min; }
double oskar_telescope_uv_filter_max(const oskar_Telescope* m)
{ return m->uv_filter_max; }
int oskar_telescope_uv_filter_units(const oskar_Telescope* m)
{ return m->uv_filter_units; }
oskar_Mem* oskar_telescope_station_true_offset_ecef_metres(
        oskar_Telescope* m, int dim)
{ return (dim >= 0 && dim < 3) ? m->station_true_offset_ecef_metres[dim] : 0; }
const oskar_Mem* oskar_telescope_station_true_offset_ecef_metres_const(
        const oskar_Telescope* m, int dim)
{ return (dim >= 0 && dim < 3) ? m->station_true_offset_ecef_metres[dim] : 0; }

void oskar_telescope_set_channel_bandwidth(oskar_Telescope* m, double bw)
{ m->channel_bandwidth_hz = bw; }
void oskar_telescope_set_time_average(oskar_Telescope* m, double t)
{ m->time_average_sec = t; }
void oskar_telescope_set_phase_centre(oskar_Telescope* m, int coord_type,
        double longitude_rad, double latitude_rad)
{
    m->phase_centre_coord_type = coord_type;
    m->phase_centre_rad[0] = longitude_rad;
    m->phase_centre_rad[1] = latitude_rad;
}

void oskar_telescope_set_uv_filter(oskar_Telescope* m, double uv_filter_min,
        double uv_filter_max, const char* units, int* status)
{
    if (*status) return;
    if (!units || !units[0])
    {
        *status = OSKAR_ERR_BAD_UNITS;
        return;
    }
    if (units[0] == 'M' || units[0] == 'm') m->uv_filter_units = OSKAR_METRES;
    else if (units[0] == 'W' || units[0] == 'w')
        m->uv_filter_units = OSKAR_WAVELENGTHS;
    else
    {
        *status = OSKAR_ERR_BAD_UNITS;
        return;
    }
    m->uv_filter_min = uv_filter_min;
    m->uv_filter_max = uv_filter_max;
}
