/*
 * Aperture-array station model, element model, work buffers and the
 * station-beam evaluation chain.
 *
 * Reference:
 *  - oskar/telescope/station/src/oskar_station_create.c:20-120 (defaults),
 *    oskar_station_set_element_coords.c, oskar_station_analyse.c:20-175,
 *    oskar_station_set_element_feed_angle.c:28-47;
 *  - oskar/telescope/station/src/oskar_evaluate_station_beam_aperture_array.c
 *    :70-240 (element patterns per type into a signal buffer of
 *    num_types x (num_points + 1), weights per feed, oskar_dftw over the true
 *    element positions with the element types as data indices, blanking);
 *  - oskar/telescope/station/element/src/oskar_element_evaluate.c:22-170
 *    (theta/phi with delta_phi = pi/2 - orientation, dipole or isotropic,
 *    optional normalisation by an extra point at the pole);
 *  - oskar/telescope/station/src/oskar_station_evaluate_element_weights.c:15-66;
 *  - oskar/telescope/station/src/oskar_station_beam_horizon_direction.c:17-90;
 *  - oskar/convert/define_convert_apparent_ra_dec_to_enu_directions.h.
 */
#include <float.h>
#include <limits.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "private.h"
#include "oskar_dftw.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif
#define C_0 299792458.0

/* ---- Element ------------------------------------------------------------ */
static oskar_Element* element_create(int precision, int location)
{
    oskar_Element* e = (oskar_Element*) calloc(1, sizeof(oskar_Element));
    if (!e) return 0;
    e->precision = precision;
    e->mem_location = location;
    e->element_type = OSKAR_ELEMENT_TYPE_DIPOLE;
    e->dipole_length = 0.5;
    e->dipole_length_units = OSKAR_WAVELENGTHS;
    return e;
}

void oskar_element_set_element_type(oskar_Element* data, const char* type,
        int* status)
{
    if (*status || !data) return;
    if (!type || !type[0]) { *status = OSKAR_ERR_INVALID_ARGUMENT; return; }
    if (type[0] == 'D' || type[0] == 'd')
        data->element_type = OSKAR_ELEMENT_TYPE_DIPOLE;
    else if (type[0] == 'I' || type[0] == 'i')
        data->element_type = OSKAR_ELEMENT_TYPE_ISOTROPIC;
    else *status = OSKAR_ERR_INVALID_ARGUMENT;
}

void oskar_element_set_dipole_length(oskar_Element* data, double value,
        const char* units, int* status)
{
    if (*status || !data) return;
    if (!units || !units[0]) { *status = OSKAR_ERR_BAD_UNITS; return; }
    if (units[0] == 'M' || units[0] == 'm')
        data->dipole_length_units = OSKAR_METRES;
    else if (units[0] == 'W' || units[0] == 'w')
        data->dipole_length_units = OSKAR_WAVELENGTHS;
    else { *status = OSKAR_ERR_BAD_UNITS; return; }
    data->dipole_length = value;
}

int oskar_element_type(const oskar_Element* d) { return d->element_type; }
double oskar_element_dipole_length(const oskar_Element* d)
{ return d->dipole_length; }
int oskar_element_dipole_length_units(const oskar_Element* d)
{ return d->dipole_length_units; }

/* ---- Station ------------------------------------------------------------ */
/* Per-feed element arrays: slot(s, feed, f) is the field's array for that
 * feed; feed_mem falls back to feed 0 where feed 1 has none
 * (oskar_station_accessors.c:259-275). */
enum { F_TRUE = 0, F_MEAS = 3, F_CABLE = 6, F_GAIN, F_GAIN_ERR, F_PHASE,
    F_PHASE_ERR, F_WEIGHT, F_COUNT };

static oskar_Mem** slot(oskar_Station* s, int feed, int f)
{
    if (f < F_MEAS) return feed ? &s->y_true_enu_metres[f] :
            &s->element_true_enu_metres[f];
    if (f < F_CABLE) return feed ? &s->y_measured_enu_metres[f - F_MEAS] :
            &s->element_measured_enu_metres[f - F_MEAS];
    switch (f)
    {
    case F_CABLE: return feed ? &s->y_cable_length_error :
            &s->element_cable_length_error;
    case F_GAIN: return feed ? &s->y_gain : &s->element_gain;
    case F_GAIN_ERR: return feed ? &s->y_gain_error : &s->element_gain_error;
    case F_PHASE: return feed ? &s->y_phase_offset_rad :
            &s->element_phase_offset_rad;
    case F_PHASE_ERR: return feed ? &s->y_phase_error_rad :
            &s->element_phase_error_rad;
    default: return feed ? &s->y_weight : &s->element_weight;
    }
}

static oskar_Mem* feed_mem(const oskar_Station* s, int feed, int f)
{
    oskar_Mem* m = *slot((oskar_Station*) s, feed ? 1 : 0, f);
    return m ? m : *slot((oskar_Station*) s, 0, f);
}

static int has_feed1_data(const oskar_Station* s)
{
    int f = 0;
    for (f = 0; f < F_COUNT; ++f)
        if (*slot((oskar_Station*) s, 1, f)) return 1;
    return 0;
}

oskar_Station* oskar_station_create(int type, int location, int num_elements,
        int* status)
{
    int i = 0, f = 0;
    oskar_Station* s = 0;
    if (*status) return 0;
    if (type != OSKAR_SINGLE && type != OSKAR_DOUBLE)
    {
        *status = OSKAR_ERR_BAD_DATA_TYPE;
        return 0;
    }
    if (location != OSKAR_CPU && location != OSKAR_GPU)
    {
        *status = OSKAR_ERR_BAD_LOCATION;
        return 0;
    }
    s = (oskar_Station*) calloc(1, sizeof(oskar_Station));
    if (!s) { *status = OSKAR_ERR_MEMORY_ALLOC_FAILURE; return 0; }
    s->precision = type;
    s->mem_location = location;
    s->station_type = OSKAR_STATION_TYPE_AA;
    s->normalise_final_beam = OSKAR_TRUE;
    s->beam_coord_type = OSKAR_COORDS_RADEC;
    s->num_elements = num_elements;
    s->enable_array_pattern = OSKAR_TRUE;
    s->common_element_orientation = OSKAR_TRUE;
    s->common_pol_beams = OSKAR_TRUE;
    s->seed_time_variable_errors = 1;   /* oskar_station_create.c:100 */
    for (i = 0; i < 3; ++i)
    {
        s->element_true_enu_metres[i] =
                oskar_mem_create(type, location, num_elements, status);
        s->element_measured_enu_metres[i] =
                oskar_mem_create(type, location, num_elements, status);
    }
    s->element_cable_length_error =
            oskar_mem_create(type, location, num_elements, status);
    for (f = 0; f < 2; ++f)
        for (i = 0; i < 3; ++i)
            s->element_euler_cpu[f][i] = oskar_mem_create(OSKAR_DOUBLE,
                    OSKAR_CPU, num_elements, status);
    s->element_types = oskar_mem_create(OSKAR_INT, location, num_elements,
            status);
    s->element_types_cpu = oskar_mem_create(OSKAR_INT, OSKAR_CPU,
            num_elements, status);
    oskar_mem_clear_contents(s->element_cable_length_error, status);
    oskar_mem_clear_contents(s->element_types, status);
    for (i = 0; i < 3; ++i)
    {
        oskar_mem_clear_contents(s->element_true_enu_metres[i], status);
        oskar_mem_clear_contents(s->element_measured_enu_metres[i], status);
    }
    oskar_station_resize_element_types(s, 1, status);
    return s;
}

void oskar_station_free(oskar_Station* s, int* status)
{
    int i = 0, f = 0;
    if (!s) return;
    for (i = 0; i < 3; ++i)
    {
        oskar_mem_free(s->element_true_enu_metres[i], status);
        oskar_mem_free(s->element_measured_enu_metres[i], status);
    }
    oskar_mem_free(s->element_cable_length_error, status);
    for (f = 0; f < 2; ++f)
        for (i = 0; i < 3; ++i) oskar_mem_free(s->element_euler_cpu[f][i],
                status);
    oskar_mem_free(s->element_types, status);
    oskar_mem_free(s->element_types_cpu, status);
    oskar_mem_free(s->permitted_beam_az_rad, status);
    oskar_mem_free(s->permitted_beam_el_rad, status);
    oskar_mem_free(s->element_gain, status);
    oskar_mem_free(s->element_gain_error, status);
    oskar_mem_free(s->element_phase_offset_rad, status);
    oskar_mem_free(s->element_phase_error_rad, status);
    oskar_mem_free(s->element_weight, status);
    for (f = 0; f < F_COUNT; ++f) oskar_mem_free(*slot(s, 1, f), status);
    for (i = 0; i < s->num_element_types; ++i) free(s->element[i]);
    free(s->element);
    if (s->child)
    {
        for (i = 0; i < s->num_elements; ++i)
            oskar_station_free(s->child[i], status);
        free(s->child);
    }
    free(s);
}

static void resize_zero(oskar_Mem* m, size_t n, int* status)
{
    const size_t old = oskar_mem_length(m);
    oskar_mem_realloc(m, n, status);
    if (*status || n <= old) return;
    /* Clear the new tail through a temporary host buffer. */
    {
        oskar_Mem* z = oskar_mem_create(oskar_mem_type(m), OSKAR_CPU,
                n - old, status);
        oskar_mem_copy_contents(m, z, old, 0, n - old, status);
        oskar_mem_free(z, status);
    }
}

void oskar_station_resize(oskar_Station* s, int num_elements, int* status)
{
    int i = 0, f = 0, feed = 0;
    const int old = s ? s->num_elements : 0;
    if (*status || !s) return;
    for (i = 0; i < 3; ++i)
    {
        resize_zero(s->element_true_enu_metres[i], num_elements, status);
        resize_zero(s->element_measured_enu_metres[i], num_elements, status);
    }
    resize_zero(s->element_cable_length_error, num_elements, status);
    for (f = 0; f < 2; ++f)
        for (i = 0; i < 3; ++i)
            resize_zero(s->element_euler_cpu[f][i], num_elements, status);
    resize_zero(s->element_types, num_elements, status);
    resize_zero(s->element_types_cpu, num_elements, status);
    /* Every per-feed array created so far (feed 0's gains, errors, phases
     * and weights are created on first use, feed 1's arrays all are) takes
     * the new length; new elements get gain 1 and weight 1 + 0i, the other
     * fields 0 (oskar_station_resize.c:48-66). */
    for (feed = 0; feed < 2; ++feed)
        for (f = feed ? 0 : F_GAIN; f < F_COUNT; ++f)
        {
            oskar_Mem* m = *slot(s, feed, f);
            if (!m) continue;
            resize_zero(m, (size_t) num_elements, status);
            if ((f == F_GAIN || f == F_WEIGHT) && num_elements > old)
                oskar_mem_set_value_real(m, 1.0, (size_t) old,
                        (size_t) (num_elements - old), status);
        }
    s->num_elements = num_elements;
}

oskar_Station* oskar_station_create_copy(const oskar_Station* src,
        int location, int* status)
{
    int i = 0, f = 0;
    oskar_Station* s = 0;
    if (*status || !src) return 0;
    s = oskar_station_create(src->precision, location, src->num_elements,
            status);
    if (*status) return s;
    s->station_type = src->station_type;
    s->normalise_final_beam = src->normalise_final_beam;
    memcpy(s->offset_ecef, src->offset_ecef, sizeof(src->offset_ecef));
    s->lon_rad = src->lon_rad;
    s->lat_rad = src->lat_rad;
    s->alt_metres = src->alt_metres;
    s->beam_lon_rad = src->beam_lon_rad;
    s->beam_lat_rad = src->beam_lat_rad;
    s->beam_coord_type = src->beam_coord_type;
    s->normalise_array_pattern = src->normalise_array_pattern;
    s->normalise_element_pattern = src->normalise_element_pattern;
    s->enable_array_pattern = src->enable_array_pattern;
    s->common_element_orientation = src->common_element_orientation;
    s->common_pol_beams = src->common_pol_beams;
    s->swap_xy = src->swap_xy;
    s->array_is_3d = src->array_is_3d;
    for (i = 0; i < 3; ++i)
    {
        oskar_mem_copy(s->element_true_enu_metres[i],
                src->element_true_enu_metres[i], status);
        oskar_mem_copy(s->element_measured_enu_metres[i],
                src->element_measured_enu_metres[i], status);
    }
    oskar_mem_copy(s->element_cable_length_error,
            src->element_cable_length_error, status);
    for (f = 0; f < 2; ++f)
        for (i = 0; i < 3; ++i)
            oskar_mem_copy(s->element_euler_cpu[f][i],
                    src->element_euler_cpu[f][i], status);
    oskar_mem_copy(s->element_types, src->element_types, status);
    oskar_mem_copy(s->element_types_cpu, src->element_types_cpu, status);
    oskar_station_resize_element_types(s, src->num_element_types, status);
    for (i = 0; i < src->num_element_types && !*status; ++i)
        *s->element[i] = *src->element[i];
    {
        oskar_Mem** d[5] = {&s->element_gain, &s->element_gain_error,
                &s->element_phase_offset_rad, &s->element_phase_error_rad,
                &s->element_weight};
        const oskar_Mem* m[5] = {src->element_gain, src->element_gain_error,
                src->element_phase_offset_rad, src->element_phase_error_rad,
                src->element_weight};
        for (i = 0; i < 5; ++i)
            if (m[i] && !*status)
                *d[i] = oskar_mem_create_copy(m[i], location, status);
        s->apply_element_errors = src->apply_element_errors;
        s->seed_time_variable_errors = src->seed_time_variable_errors;
        s->unique_id = src->unique_id;
        s->element_errors_random = src->element_errors_random;
        s->apply_element_weight = src->apply_element_weight;
    }
    if (src->num_permitted_beams > 0 && !*status)
    {
        s->permitted_beam_az_rad = oskar_mem_create_copy(
                src->permitted_beam_az_rad, OSKAR_CPU, status);
        s->permitted_beam_el_rad = oskar_mem_create_copy(
                src->permitted_beam_el_rad, OSKAR_CPU, status);
        s->num_permitted_beams = src->num_permitted_beams;
    }
    for (f = 0; f < F_COUNT && !*status; ++f)
        if (*slot((oskar_Station*) src, 1, f))
            *slot(s, 1, f) = oskar_mem_create_copy(
                    *slot((oskar_Station*) src, 1, f), location, status);
    s->identical_children = src->identical_children;
    if (src->child && !*status)
    {
        s->child = (oskar_Station**) calloc((size_t) src->num_elements,
                sizeof(oskar_Station*));
        if (!s->child) { *status = OSKAR_ERR_MEMORY_ALLOC_FAILURE; return s; }
        for (i = 0; i < src->num_elements; ++i)
            s->child[i] = oskar_station_create_copy(src->child[i], location,
                    status);
    }
    return s;
}

/* ---- Child stations (hierarchical stations) ----------------------------- */
/* oskar/telescope/station/src/oskar_station_create_child_stations.c:15-44:
 * one empty child station per element, of the parent's precision and
 * location; an existing child array is an allocation error. */
void oskar_station_create_child_stations(oskar_Station* s, int* status)
{
    int i = 0;
    if (*status || !s) return;
    if (s->child)
    {
        *status = OSKAR_ERR_MEMORY_ALLOC_FAILURE;
        return;
    }
    s->child = (oskar_Station**) calloc((size_t) (s->num_elements > 0 ?
            s->num_elements : 1), sizeof(oskar_Station*));
    if (!s->child) { *status = OSKAR_ERR_MEMORY_ALLOC_FAILURE; return; }
    for (i = 0; i < s->num_elements; ++i)
        s->child[i] = oskar_station_create(s->precision, s->mem_location, 0,
                status);
}

/* oskar_station_accessors.c:166-168, 422-436. */
int oskar_station_has_child(const oskar_Station* s)
{
    return s ? (s->child ? 1 : 0) : 0;
}

oskar_Station* oskar_station_child(oskar_Station* s, int i)
{
    return s ? s->child[i] : 0;
}

const oskar_Station* oskar_station_child_const(const oskar_Station* s, int i)
{
    return s ? s->child[i] : 0;
}

int oskar_station_identical_children(const oskar_Station* s)
{
    return s ? s->identical_children : 0;
}

/* Nonzero if n values of two arrays differ, or their type or length does
 * (oskar_mem_different for the arrays a station holds; either may be
 * NULL, which matches only NULL). */
static int mem_different(const oskar_Mem* a, const oskar_Mem* b, size_t n,
        int* status)
{
    int diff = 0;
    if (!a || !b) return a != b;
    if (oskar_mem_type(a) != oskar_mem_type(b)) return 1;
    if (n == 0)
    {
        if (oskar_mem_length(a) != oskar_mem_length(b)) return 1;
        n = oskar_mem_length(a);
    }
    if (oskar_mem_length(a) < n || oskar_mem_length(b) < n) return 1;
    {
        const size_t bytes = n * oskar_mem_element_size(oskar_mem_type(a));
        oskar_Mem* ha = oskar_mem_create_copy(a, OSKAR_CPU, status);
        oskar_Mem* hb = oskar_mem_create_copy(b, OSKAR_CPU, status);
        if (!*status)
            diff = memcmp(oskar_mem_void_const(ha), oskar_mem_void_const(hb),
                    bytes) != 0;
        oskar_mem_free(ha, status);
        oskar_mem_free(hb, status);
    }
    return diff;
}

/* oskar/telescope/station/src/oskar_station_different.c:15-200, over the
 * fields this library's stations hold: the meta-data, the element arrays
 * (coordinates, orientations, cable-length errors, gains, phase offsets,
 * apodisation weights, types), the element models, the permitted beams and,
 * recursively, the child stations.  The unique ID is not compared. */
int oskar_station_different(const oskar_Station* a, const oskar_Station* b,
        int* status)
{
    int i = 0, f = 0;
    size_t n = 0;
    if (!a || !b) return 0;
    if (*status) return 1;
    n = (size_t) a->num_elements;
    if (a->station_type != b->station_type ||
            a->normalise_final_beam != b->normalise_final_beam ||
            a->beam_coord_type != b->beam_coord_type ||
            a->beam_lon_rad != b->beam_lon_rad ||
            a->beam_lat_rad != b->beam_lat_rad ||
            a->identical_children != b->identical_children ||
            a->num_elements != b->num_elements ||
            a->num_element_types != b->num_element_types ||
            a->normalise_array_pattern != b->normalise_array_pattern ||
            a->normalise_element_pattern != b->normalise_element_pattern ||
            a->enable_array_pattern != b->enable_array_pattern ||
            a->common_element_orientation != b->common_element_orientation ||
            a->common_pol_beams != b->common_pol_beams ||
            a->array_is_3d != b->array_is_3d ||
            a->apply_element_errors != b->apply_element_errors ||
            a->apply_element_weight != b->apply_element_weight ||
            a->num_permitted_beams != b->num_permitted_beams ||
            (a->child != 0) != (b->child != 0))
        return 1;
    for (i = 0; i < 3; ++i)
        if (mem_different(a->element_measured_enu_metres[i],
                b->element_measured_enu_metres[i], n, status) ||
                mem_different(a->element_true_enu_metres[i],
                b->element_true_enu_metres[i], n, status))
            return 1;
    for (f = 0; f < 2; ++f)
        for (i = 0; i < 3; ++i)
            if (mem_different(a->element_euler_cpu[f][i],
                    b->element_euler_cpu[f][i], n, status))
                return 1;
    if (mem_different(a->element_gain, b->element_gain, n, status) ||
            mem_different(a->element_phase_offset_rad,
            b->element_phase_offset_rad, n, status) ||
            mem_different(a->element_weight, b->element_weight, n, status) ||
            mem_different(a->element_cable_length_error,
            b->element_cable_length_error, n, status) ||
            mem_different(a->element_types, b->element_types, n, status) ||
            mem_different(a->element_types_cpu, b->element_types_cpu, n,
            status) ||
            mem_different(a->permitted_beam_az_rad, b->permitted_beam_az_rad,
            0, status) ||
            mem_different(a->permitted_beam_el_rad, b->permitted_beam_el_rad,
            0, status))
        return 1;
    for (f = 0; f < F_COUNT; ++f)
        if (mem_different(*slot((oskar_Station*) a, 1, f),
                *slot((oskar_Station*) b, 1, f), n, status))
            return 1;
    for (i = 0; i < a->num_element_types; ++i)
        if (a->element[i]->element_type != b->element[i]->element_type ||
                a->element[i]->dipole_length != b->element[i]->dipole_length ||
                a->element[i]->dipole_length_units !=
                b->element[i]->dipole_length_units)
            return 1;
    if (a->child)
        for (i = 0; i < a->num_elements; ++i)
            if (oskar_station_different(a->child[i], b->child[i], status))
                return 1;
    return 0;
}

void oskar_station_resize_element_types(oskar_Station* s,
        int num_element_types, int* status)
{
    int i = 0;
    if (*status || !s) return;
    if (num_element_types < s->num_element_types)
    {
        for (i = num_element_types; i < s->num_element_types; ++i)
            free(s->element[i]);
    }
    {
        oskar_Element** e = (oskar_Element**) realloc(s->element,
                (num_element_types > 0 ? num_element_types : 1) *
                sizeof(oskar_Element*));
        if (!e) { *status = OSKAR_ERR_MEMORY_ALLOC_FAILURE; return; }
        s->element = e;
    }
    for (i = s->num_element_types; i < num_element_types; ++i)
    {
        s->element[i] = element_create(s->precision, s->mem_location);
        if (!s->element[i]) { *status = OSKAR_ERR_MEMORY_ALLOC_FAILURE; }
    }
    s->num_element_types = num_element_types;
}

static void set_real(oskar_Mem* m, int index, double v, int* status)
{
    oskar_mem_set_element_real(m, (size_t) index, v, status);
}

/* A per-element array of the station, created on first use with every
 * element at the default value (real part; imaginary parts zero). */
static oskar_Mem* element_array(oskar_Station* s, oskar_Mem** m, int type,
        double value, int* status)
{
    if (!*m && !*status)
    {
        *m = oskar_mem_create(type, s->mem_location, (size_t) s->num_elements,
                status);
        oskar_mem_set_value_real(*m, value, 0, (size_t) s->num_elements,
                status);
    }
    return *m;
}

/* The array of field f for one feed, created on first use.  Feed 1's
 * starts as a copy of feed 0's (or at the default where feed 0 has none),
 * so elements not set for feed 1 keep feed 0's values; the reference
 * leaves them uninitialised (oskar_station_set_element_coords.c:41-54). */
static oskar_Mem* field_array(oskar_Station* s, int feed, int f, int type,
        double value, int* status)
{
    oskar_Mem** m = slot(s, feed, f);
    if (feed && !*m && !*status && *slot(s, 0, f))
        *m = oskar_mem_create_copy(*slot(s, 0, f), s->mem_location, status);
    return element_array(s, m, type, value, status);
}

static int bad_feed_or_index(const oskar_Station* s, int feed, int index,
        int* status)
{
    if (*status || !s) return 1;
    if (feed < 0 || feed > 1 || index < 0 || index >= s->num_elements)
    {
        *status = OSKAR_ERR_OUT_OF_RANGE;
        return 1;
    }
    return 0;
}

void oskar_station_set_element_coords(oskar_Station* s, int feed, int index,
        const double measured_enu[3], const double true_enu[3], int* status)
{
    int i = 0;
    if (bad_feed_or_index(s, feed, index, status)) return;
    for (i = 0; i < 3; ++i)
    {
        set_real(field_array(s, feed, F_MEAS + i, s->precision, 0.0, status),
                index, measured_enu[i], status);
        set_real(field_array(s, feed, F_TRUE + i, s->precision, 0.0, status),
                index, true_enu[i], status);
    }
    if (measured_enu[2] != 0.0 || true_enu[2] != 0.0) s->array_is_3d = 1;
}

void oskar_station_set_element_cable_length_error(oskar_Station* s, int feed,
        int index, const double error_metres, int* status)
{
    if (bad_feed_or_index(s, feed, index, status)) return;
    set_real(field_array(s, feed, F_CABLE, s->precision, 0.0, status), index,
            error_metres, status);
}

void oskar_station_set_element_errors(oskar_Station* s, int feed, int index,
        double gain, double gain_error, double phase_offset_deg,
        double phase_error_deg, int* status)
{
    if (bad_feed_or_index(s, feed, index, status)) return;
    const int t = s->precision;
    set_real(field_array(s, feed, F_GAIN, t, 1.0, status), index, gain,
            status);
    set_real(field_array(s, feed, F_GAIN_ERR, t, 0.0, status), index,
            gain_error, status);
    set_real(field_array(s, feed, F_PHASE, t, 0.0, status), index,
            phase_offset_deg * M_PI / 180.0, status);
    set_real(field_array(s, feed, F_PHASE_ERR, t, 0.0, status), index,
            phase_error_deg * M_PI / 180.0, status);
}

void oskar_station_set_element_weight(oskar_Station* s, int feed, int index,
        double re, double im, int* status)
{
    if (bad_feed_or_index(s, feed, index, status)) return;
    oskar_Mem* w = field_array(s, feed, F_WEIGHT,
            s->precision | OSKAR_COMPLEX, 1.0, status);
    if (*status) return;
    {
        /* One complex value through a host alias. */
        double vd[2] = {re, im};
        float vf[2] = {(float) re, (float) im};
        oskar_Mem* src = oskar_mem_create_alias_from_raw(
                s->precision == OSKAR_DOUBLE ? (void*) vd : (void*) vf,
                s->precision | OSKAR_COMPLEX, OSKAR_CPU, 1, status);
        oskar_mem_copy_contents(w, src, (size_t) index, 0, 1, status);
        oskar_mem_free(src, status);
    }
}

unsigned int oskar_station_seed_time_variable_errors(const oskar_Station* s)
{
    return s ? s->seed_time_variable_errors : 0u;
}

void oskar_station_set_seed_time_variable_errors(oskar_Station* s,
        unsigned int value)
{
    if (s) s->seed_time_variable_errors = value;
}

int oskar_station_unique_id(const oskar_Station* s)
{
    return s ? s->unique_id : 0;
}

/* oskar_station_accessors.c:503-515: depth first, parent before its
 * children. */
void oskar_station_set_unique_ids(oskar_Station* s, int* counter)
{
    int i = 0;
    if (!s || !counter) return;
    s->unique_id = (*counter)++;
    if (s->child)
        for (i = 0; i < s->num_elements; ++i)
            oskar_station_set_unique_ids(s->child[i], counter);
}

int oskar_station_apply_element_errors(const oskar_Station* s)
{
    return s ? s->apply_element_errors : 0;
}

int oskar_station_apply_element_weight(const oskar_Station* s)
{
    return s ? s->apply_element_weight : 0;
}

/* The arrays a feed's beam uses (feed 1 falls back to feed 0's); NULL until
 * a value of that field has been set. */
const oskar_Mem* oskar_station_element_gain_const(const oskar_Station* s,
        int feed)
{
    return s ? feed_mem(s, feed, F_GAIN) : 0;
}

const oskar_Mem* oskar_station_element_gain_error_const(
        const oskar_Station* s, int feed)
{
    return s ? feed_mem(s, feed, F_GAIN_ERR) : 0;
}

const oskar_Mem* oskar_station_element_phase_offset_rad_const(
        const oskar_Station* s, int feed)
{
    return s ? feed_mem(s, feed, F_PHASE) : 0;
}

const oskar_Mem* oskar_station_element_phase_error_rad_const(
        const oskar_Station* s, int feed)
{
    return s ? feed_mem(s, feed, F_PHASE_ERR) : 0;
}

const oskar_Mem* oskar_station_element_weight_const(const oskar_Station* s,
        int feed)
{
    return s ? feed_mem(s, feed, F_WEIGHT) : 0;
}

void oskar_station_set_element_feed_angle(oskar_Station* s, int feed,
        int index, double alpha_deg, double beta_deg, double gamma_deg,
        int* status)
{
    const double deg2rad = M_PI / 180.0;
    if (*status || !s) return;
    if (index < 0 || index >= s->num_elements || feed < 0 || feed > 1)
    {
        *status = OSKAR_ERR_OUT_OF_RANGE;
        return;
    }
    ((double*) oskar_mem_void(s->element_euler_cpu[feed][0]))[index] =
            alpha_deg * deg2rad;
    ((double*) oskar_mem_void(s->element_euler_cpu[feed][1]))[index] =
            beta_deg * deg2rad;
    ((double*) oskar_mem_void(s->element_euler_cpu[feed][2]))[index] =
            gamma_deg * deg2rad;
}

void oskar_station_set_element_type(oskar_Station* s, int index,
        int element_type, int* status)
{
    if (*status || !s) return;
    if (index < 0 || index >= s->num_elements)
    {
        *status = OSKAR_ERR_OUT_OF_RANGE;
        return;
    }
    ((int*) oskar_mem_void(s->element_types_cpu))[index] = element_type;
    set_real(s->element_types, index, (double) element_type, status);
}

void oskar_station_analyse(oskar_Station* s,
        int* finished_identical_station_check, int* status)
{
    int f = 0, i = 0, d = 0;
    if (*status || !s) return;
    s->array_is_3d = 0;
    s->common_element_orientation = 1;
    s->common_pol_beams = 1;
    for (f = 0; f < 2; ++f)
    {
        const double* e[3];
        for (d = 0; d < 3; ++d)
            e[d] = (const double*) oskar_mem_void_const(
                    s->element_euler_cpu[f][d]);
        for (i = 1; i < s->num_elements; ++i)
            if (e[0][i] != e[0][0] || e[1][i] != e[1][0] || e[2][i] != e[2][0])
            {
                s->common_element_orientation = 0;
                break;
            }
    }
    /* Feed-1 data of any kind give separate polarisation beams, and both
     * feeds are then analysed (oskar_station_analyse.c:62-79). */
    if (has_feed1_data(s)) s->common_pol_beams = 0;
    const int num_feeds = s->common_pol_beams ? 1 : 2;
    /* Element errors and apodisation (oskar_station_analyse.c:80-133):
     * errors apply if any gain != 1 or phase offset != 0, or any error
     * scale != 0; apodisation if any weight != 1 + 0i. */
    s->apply_element_errors = 0;
    s->element_errors_random = 0;
    s->apply_element_weight = 0;
    for (f = 0; f < num_feeds && !*status; ++f)
    {
        const oskar_Mem* z[2] = { feed_mem(s, f, F_TRUE + 2),
                feed_mem(s, f, F_MEAS + 2) };
        for (d = 0; d < 2 && !*status; ++d)
        {
            oskar_Mem* h = oskar_mem_create_copy(z[d], OSKAR_CPU, status);
            for (i = 0; i < s->num_elements && !*status; ++i)
            {
                const double v = oskar_mem_get_element(h, (size_t) i, status);
                if (v != 0.0) { s->array_is_3d = 1; break; }
            }
            oskar_mem_free(h, status);
        }
        const oskar_Mem* m[4] = {feed_mem(s, f, F_GAIN),
                feed_mem(s, f, F_PHASE), feed_mem(s, f, F_GAIN_ERR),
                feed_mem(s, f, F_PHASE_ERR)};
        const double dflt[4] = {1.0, 0.0, 0.0, 0.0};
        for (d = 0; d < 4 && !*status; ++d)
        {
            if (!m[d]) continue;
            oskar_Mem* h = oskar_mem_create_copy(m[d], OSKAR_CPU, status);
            for (i = 0; i < s->num_elements && !*status; ++i)
                if (oskar_mem_get_element(h, (size_t) i, status) != dflt[d])
                {
                    s->apply_element_errors = 1;
                    /* Random errors make every station unique
                     * (oskar_station_analyse.c:119-125). */
                    if (d >= 2)
                    {
                        s->element_errors_random = 1;
                        *finished_identical_station_check = 1;
                    }
                    break;
                }
            oskar_mem_free(h, status);
        }
        const oskar_Mem* wm = feed_mem(s, f, F_WEIGHT);
        if (wm && !*status)
        {
            oskar_Mem* h = oskar_mem_create_copy(wm, OSKAR_CPU, status);
            const int dbl = s->precision == OSKAR_DOUBLE;
            const void* p = oskar_mem_void_const(h);
            for (i = 0; i < s->num_elements && !*status; ++i)
            {
                const double re = dbl ? ((const double*) p)[2 * i] :
                        ((const float*) p)[2 * i];
                const double im = dbl ? ((const double*) p)[2 * i + 1] :
                        ((const float*) p)[2 * i + 1];
                if (re != 1.0 || im != 0.0)
                {
                    s->apply_element_weight = 1;
                    break;
                }
            }
            oskar_mem_free(h, status);
        }
    }
    /* Keep the device copy of the element types in step with the host. */
    oskar_mem_copy(s->element_types, s->element_types_cpu, status);

    /* Children, then whether they are all alike
     * (oskar_station_analyse.c:193-220). */
    if (s->child)
    {
        for (i = 0; i < s->num_elements; ++i)
            oskar_station_analyse(s->child[i],
                    finished_identical_station_check, status);
        s->identical_children = 0;
        if (!*finished_identical_station_check)
        {
            s->identical_children = 1;
            for (i = 1; i < s->num_elements; ++i)
                if (oskar_station_different(s->child[0], s->child[i],
                        status))
                    s->identical_children = 0;
        }
    }
}

void oskar_station_set_station_type(oskar_Station* m, int type)
{ m->station_type = type; }
void oskar_station_set_normalise_final_beam(oskar_Station* m, int v)
{ m->normalise_final_beam = v; }
void oskar_station_set_position(oskar_Station* m, double lon, double lat,
        double alt, double x, double y, double z)
{
    m->lon_rad = lon; m->lat_rad = lat; m->alt_metres = alt;
    m->offset_ecef[0] = x; m->offset_ecef[1] = y; m->offset_ecef[2] = z;
}
void oskar_station_set_phase_centre(oskar_Station* m, int coord_type,
        double lon, double lat)
{
    m->beam_coord_type = coord_type;
    m->beam_lon_rad = lon;
    m->beam_lat_rad = lat;
}
void oskar_station_set_normalise_array_pattern(oskar_Station* m, int v)
{ m->normalise_array_pattern = v; }
void oskar_station_set_normalise_element_pattern(oskar_Station* m, int v)
{ m->normalise_element_pattern = v; }
void oskar_station_set_enable_array_pattern(oskar_Station* m, int v)
{ m->enable_array_pattern = v; }
void oskar_station_set_swap_xy(oskar_Station* m, int v) { m->swap_xy = v; }

int oskar_station_precision(const oskar_Station* m) { return m->precision; }
int oskar_station_mem_location(const oskar_Station* m)
{ return m->mem_location; }
int oskar_station_type(const oskar_Station* m) { return m->station_type; }
int oskar_station_num_elements(const oskar_Station* m)
{ return m->num_elements; }
int oskar_station_num_element_types(const oskar_Station* m)
{ return m->num_element_types; }
int oskar_station_normalise_final_beam(const oskar_Station* m)
{ return m->normalise_final_beam; }
int oskar_station_normalise_array_pattern(const oskar_Station* m)
{ return m->normalise_array_pattern; }
int oskar_station_normalise_element_pattern(const oskar_Station* m)
{ return m->normalise_element_pattern; }
int oskar_station_enable_array_pattern(const oskar_Station* m)
{ return m->enable_array_pattern; }
int oskar_station_swap_xy(const oskar_Station* m) { return m->swap_xy; }
int oskar_station_array_is_3d(const oskar_Station* m)
{ return m->array_is_3d; }
int oskar_station_common_element_orientation(const oskar_Station* m)
{ return m->common_element_orientation; }
int oskar_station_common_pol_beams(const oskar_Station* m)
{ return m->common_pol_beams; }
double oskar_station_lon_rad(const oskar_Station* m) { return m->lon_rad; }
double oskar_station_lat_rad(const oskar_Station* m) { return m->lat_rad; }
double oskar_station_beam_lon_rad(const oskar_Station* m)
{ return m->beam_lon_rad; }
double oskar_station_beam_lat_rad(const oskar_Station* m)
{ return m->beam_lat_rad; }
int oskar_station_beam_coord_type(const oskar_Station* m)
{ return m->beam_coord_type; }
oskar_Mem* oskar_station_element_true_enu_metres(oskar_Station* m, int feed,
        int dim)
{ return (feed >= 0 && feed <= 1 && dim >= 0 && dim < 3) ?
        feed_mem(m, feed, F_TRUE + dim) : 0; }
const oskar_Mem* oskar_station_element_true_enu_metres_const(
        const oskar_Station* m, int feed, int dim)
{ return (feed >= 0 && feed <= 1 && dim >= 0 && dim < 3) ?
        feed_mem(m, feed, F_TRUE + dim) : 0; }
oskar_Mem* oskar_station_element_measured_enu_metres(oskar_Station* m,
        int feed, int dim)
{ return (feed >= 0 && feed <= 1 && dim >= 0 && dim < 3) ?
        feed_mem(m, feed, F_MEAS + dim) : 0; }
const oskar_Mem* oskar_station_element_measured_enu_metres_const(
        const oskar_Station* m, int feed, int dim)
{ return (feed >= 0 && feed <= 1 && dim >= 0 && dim < 3) ?
        feed_mem(m, feed, F_MEAS + dim) : 0; }
oskar_Mem* oskar_station_element_cable_length_error_metres(oskar_Station* m,
        int feed)
{ return (feed >= 0 && feed <= 1) ? feed_mem(m, feed, F_CABLE) : 0; }
const oskar_Mem* oskar_station_element_cable_length_error_metres_const(
        const oskar_Station* m, int feed)
{ return (feed >= 0 && feed <= 1) ? feed_mem(m, feed, F_CABLE) : 0; }
oskar_Mem* oskar_station_element_types(oskar_Station* m)
{ return m->element_types; }
const oskar_Mem* oskar_station_element_types_const(const oskar_Station* m)
{ return m->element_types; }
double oskar_station_element_euler_index_rad(const oskar_Station* m,
        int feed, int dim, int index)
{
    if (feed < 0 || feed > 1 || dim < 0 || dim > 2 || index < 0 ||
            index >= m->num_elements) return 0.0;
    return ((const double*) oskar_mem_void_const(
            m->element_euler_cpu[feed][dim]))[index];
}
oskar_Element* oskar_station_element(oskar_Station* m, int i)
{ return (i >= 0 && i < m->num_element_types) ? m->element[i] : 0; }
const oskar_Element* oskar_station_element_const(const oskar_Station* m, int i)
{ return (i >= 0 && i < m->num_element_types) ? m->element[i] : 0; }

/* ---- Work buffers ------------------------------------------------------- */
oskar_StationWork* oskar_station_work_create(int type, int location,
        int* status)
{
    oskar_StationWork* w = 0;
    if (*status) return 0;
    if (type != OSKAR_SINGLE && type != OSKAR_DOUBLE)
    {
        *status = OSKAR_ERR_BAD_DATA_TYPE;
        return 0;
    }
    w = (oskar_StationWork*) calloc(1, sizeof(oskar_StationWork));
    if (!w) { *status = OSKAR_ERR_MEMORY_ALLOC_FAILURE; return 0; }
    w->precision = type;
    w->mem_location = location;
    w->weights = oskar_mem_create(type | OSKAR_COMPLEX, location, 0, status);
    w->theta = oskar_mem_create(type, location, 0, status);
    w->phi_x = oskar_mem_create(type, location, 0, status);
    w->phi_y = oskar_mem_create(type, location, 0, status);
    w->enu[0] = oskar_mem_create(type, location, 0, status);
    w->enu[1] = oskar_mem_create(type, location, 0, status);
    w->enu[2] = oskar_mem_create(type, location, 0, status);
    w->horizon_mask = oskar_mem_create(OSKAR_INT, location, 0, status);
    w->source_indices = oskar_mem_create(OSKAR_INT, location, 0, status);
    return w;
}

oskar_Mem* oskar_station_work_horizon_mask(oskar_StationWork* w)
{ return w->horizon_mask; }

oskar_Mem* oskar_station_work_source_indices(oskar_StationWork* w)
{ return w->source_indices; }

void oskar_station_work_free(oskar_StationWork* w, int* status)
{
    if (!w) return;
    oskar_mem_free(w->weights, status);
    oskar_mem_free(w->signal, status);
    oskar_mem_free(w->theta, status);
    oskar_mem_free(w->phi_x, status);
    oskar_mem_free(w->phi_y, status);
    oskar_mem_free(w->enu[0], status);
    oskar_mem_free(w->enu[1], status);
    oskar_mem_free(w->enu[2], status);
    oskar_mem_free(w->beam_out, status);
    oskar_mem_free(w->horizon_mask, status);
    oskar_mem_free(w->source_indices, status);
    oskar_mem_free(w->zero_idx, status);
    {
        int i = 0;
        for (i = 0; i < w->num_depths; ++i)
            oskar_mem_free(w->depth_beam[i], status);
        free(w->depth_beam);
    }
    free(w);
}

/* ---- Station-beam chain ------------------------------------------------- */
/* Great-circle distance by the haversine formula
 * (oskar/math/src/oskar_angular_distance.c:13-21). */
static double angular_distance(double lon1_rad, double lon2_rad,
        double lat1_rad, double lat2_rad)
{
    const double sin_dlat = sin(0.5 * (lat2_rad - lat1_rad));
    const double sin_dlon = sin(0.5 * (lon2_rad - lon1_rad));
    return 2.0 * asin(sqrt(sin_dlat * sin_dlat +
            cos(lat1_rad) * cos(lat2_rad) * sin_dlon * sin_dlon));
}

void oskar_station_beam_horizon_direction(const oskar_Station* s,
        const double gast_rad, double* x, double* y, double* z, int* status)
{
    if (*status) return;
    if (s->beam_coord_type == OSKAR_COORDS_RADEC)
    {
        const double lst = gast_rad + s->lon_rad;
        const double sin_lat = sin(s->lat_rad), cos_lat = cos(s->lat_rad);
        const double ha = lst - s->beam_lon_rad;
        const double sin_ha = sin(ha), cos_ha = cos(ha);
        const double sin_dec = sin(s->beam_lat_rad);
        const double cos_dec = cos(s->beam_lat_rad);
        *x = -cos_dec * sin_ha;
        *y = cos_lat * sin_dec - sin_lat * cos_dec * cos_ha;
        *z = sin_lat * sin_dec + cos_lat * cos_dec * cos_ha;
    }
    else if (s->beam_coord_type == OSKAR_COORDS_AZEL)
    {
        const double cos_el = cos(s->beam_lat_rad);
        *x = cos_el * sin(s->beam_lon_rad);
        *y = cos_el * cos(s->beam_lon_rad);
        *z = sin(s->beam_lat_rad);
    }
    else *status = OSKAR_ERR_INVALID_ARGUMENT;

    /* Snap to the nearest permitted (az, el), by great-circle distance, the
     * first of equal distances winning (:50-88). */
    if (s->num_permitted_beams > 0)
    {
        int i = 0, best = 0;
        double best_dist = DBL_MAX;
        const double az = atan2(*x, *y);
        const double el = atan2(*z, sqrt(*x * *x + *y * *y));
        const double* p_az = (const double*)
                oskar_mem_void_const(s->permitted_beam_az_rad);
        const double* p_el = (const double*)
                oskar_mem_void_const(s->permitted_beam_el_rad);
        for (i = 0; i < s->num_permitted_beams; ++i)
        {
            const double d = angular_distance(p_az[i], az, p_el[i], el);
            if (d < best_dist)
            {
                best_dist = d;
                best = i;
            }
        }
        {
            const double cos_el = cos(p_el[best]);
            *x = cos_el * sin(p_az[best]);
            *y = cos_el * cos(p_az[best]);
            *z = sin(p_el[best]);
        }
    }
}

int oskar_station_num_permitted_beams(const oskar_Station* model)
{
    return model ? model->num_permitted_beams : 0;
}

const oskar_Mem* oskar_station_permitted_beam_az_rad_const(
        const oskar_Station* model)
{
    return model ? model->permitted_beam_az_rad : 0;
}

const oskar_Mem* oskar_station_permitted_beam_el_rad_const(
        const oskar_Station* model)
{
    return model ? model->permitted_beam_el_rad : 0;
}

/* The first two numbers of a line (separators ',', ' ', '\t'; a token
 * starting with '#' ends the line), as oskar_string_to_array_d reads them
 * (oskar/utility/src/oskar_string_to_array.c). */
static int two_numbers(char* line, double* v)
{
    int n = 0;
    char* save = 0;
    char* tok = strtok_r(line, ", \t\r\n", &save);
    while (tok && n < 2)
    {
        char* end = 0;
        if (tok[0] == '#') break;
        const double x = strtod(tok, &end);
        if (end > tok) v[n++] = x;
        tok = strtok_r(0, ", \t\r\n", &save);
    }
    return n;
}

void oskar_station_load_permitted_beams(oskar_Station* station,
        const char* filename, int* status)
{
    int n = 0, cap = 0;
    double *az = 0, *el = 0;
    char line[4096];
    FILE* file = 0;
    if (*status || !station) return;
    if (station->precision != OSKAR_SINGLE &&
            station->precision != OSKAR_DOUBLE)
    {
        *status = OSKAR_ERR_BAD_DATA_TYPE;
        return;
    }
    file = fopen(filename, "r");
    if (!file)
    {
        *status = OSKAR_ERR_FILE_IO;
        return;
    }
    while (fgets(line, sizeof(line), file))
    {
        double par[2] = {0.0, 0.0};
        if (two_numbers(line, par) < 2) continue;
        if (n == cap)
        {
            cap += 100;
            double* a = (double*) realloc(az, (size_t) cap * sizeof(double));
            if (a) az = a;
            double* e = (double*) realloc(el, (size_t) cap * sizeof(double));
            if (e) el = e;
            if (!a || !e)
            {
                *status = OSKAR_ERR_MEMORY_ALLOC_FAILURE;
                break;
            }
        }
        az[n] = par[0] * M_PI / 180.0;
        el[n] = par[1] * M_PI / 180.0;
        ++n;
    }
    fclose(file);
    if (!*status)
    {
        oskar_mem_free(station->permitted_beam_az_rad, status);
        oskar_mem_free(station->permitted_beam_el_rad, status);
        station->permitted_beam_az_rad = oskar_mem_create(OSKAR_DOUBLE,
                OSKAR_CPU, (size_t) n, status);
        station->permitted_beam_el_rad = oskar_mem_create(OSKAR_DOUBLE,
                OSKAR_CPU, (size_t) n, status);
        if (!*status && n > 0)
        {
            memcpy(oskar_mem_void(station->permitted_beam_az_rad), az,
                    (size_t) n * sizeof(double));
            memcpy(oskar_mem_void(station->permitted_beam_el_rad), el,
                    (size_t) n * sizeof(double));
        }
        station->num_permitted_beams = *status ? 0 : n;
    }
    free(az);
    free(el);
}

static int check_gpu(const oskar_Mem* m, int* status)
{
    if (*status) return 0;
    if (oskar_mem_location(m) != OSKAR_GPU)
    {
        *status = OSKAR_ERR_BAD_LOCATION;
        return 0;
    }
    return 1;
}

void oskar_evaluate_element_weights_dft(int num_elements, const oskar_Mem* x,
        const oskar_Mem* y, const oskar_Mem* z,
        const oskar_Mem* cable_length_error, double wavenumber,
        double x_beam, double y_beam, double z_beam, oskar_Mem* weights,
        int* status)
{
    if (*status) return;
    const int type = oskar_mem_type(x);
    if (oskar_mem_type(y) != type || oskar_mem_type(z) != type ||
            oskar_mem_type(cable_length_error) != type ||
            oskar_mem_type(weights) != (type | OSKAR_COMPLEX))
    {
        *status = OSKAR_ERR_TYPE_MISMATCH;
        return;
    }
    if (oskar_mem_location(y) != oskar_mem_location(x) ||
            oskar_mem_location(z) != oskar_mem_location(x) ||
            oskar_mem_location(cable_length_error) != oskar_mem_location(x) ||
            oskar_mem_location(weights) != oskar_mem_location(x))
    {
        *status = OSKAR_ERR_LOCATION_MISMATCH;
        return;
    }
    if ((int) oskar_mem_length(weights) < num_elements)
    {
        *status = OSKAR_ERR_DIMENSION_MISMATCH;
        return;
    }
    if (!check_gpu(x, status)) return;
    *status = oskar_amd_hip_status(oskar_amd_element_weights_dft_launch(
            oskar_type_precision(type), num_elements, oskar_mem_void_const(x),
            oskar_mem_void_const(y), oskar_mem_void_const(z),
            oskar_mem_void_const(cable_length_error), wavenumber, x_beam,
            y_beam, z_beam, oskar_mem_void(weights),
            oskar_amd_current_stream()));
}

void oskar_convert_enu_directions_to_theta_phi(int offset_in, int num_points,
        const oskar_Mem* x, const oskar_Mem* y, const oskar_Mem* z,
        int extra_point_at_pole, double delta_phi1, double delta_phi2,
        oskar_Mem* theta, oskar_Mem* phi1, oskar_Mem* phi2, int* status)
{
    if (*status) return;
    const int type = oskar_mem_type(theta);
    if (oskar_mem_type(x) != type || oskar_mem_type(y) != type ||
            oskar_mem_type(z) != type || oskar_mem_type(phi1) != type ||
            oskar_mem_type(phi2) != type)
    {
        *status = OSKAR_ERR_TYPE_MISMATCH;
        return;
    }
    if (!check_gpu(x, status) || !check_gpu(theta, status)) return;
    {
        const size_t need = (size_t) num_points + (extra_point_at_pole ? 1 : 0);
        if (oskar_mem_length(theta) < need || oskar_mem_length(phi1) < need ||
                oskar_mem_length(phi2) < need ||
                oskar_mem_length(x) < (size_t) (offset_in + num_points))
        {
            *status = OSKAR_ERR_DIMENSION_MISMATCH;
            return;
        }
    }
    *status = oskar_amd_hip_status(oskar_amd_theta_phi_launch(type, offset_in,
            num_points, oskar_mem_void_const(x), oskar_mem_void_const(y),
            oskar_mem_void_const(z), extra_point_at_pole, delta_phi1,
            delta_phi2, oskar_mem_void(theta), oskar_mem_void(phi1),
            oskar_mem_void(phi2), oskar_amd_current_stream()));
}

void oskar_evaluate_dipole_pattern(int num_points, const oskar_Mem* theta,
        const oskar_Mem* phi_x, const oskar_Mem* phi_y, double freq_hz,
        double dipole_length_m, int swap_xy, int offset_out,
        oskar_Mem* pattern, int* status)
{
    if (*status) return;
    const int precision = oskar_mem_precision(pattern);
    const int location = oskar_mem_location(pattern);
    const double kL = dipole_length_m * (M_PI * freq_hz / C_0);
    const double cos_kL = cos(kL);
    if (oskar_mem_location(theta) != location ||
            oskar_mem_location(phi_x) != location ||
            oskar_mem_location(phi_y) != location)
    {
        *status = OSKAR_ERR_LOCATION_MISMATCH;
        return;
    }
    if (oskar_mem_type(theta) != precision ||
            oskar_mem_type(phi_x) != precision ||
            oskar_mem_type(phi_y) != precision)
    {
        *status = OSKAR_ERR_TYPE_MISMATCH;
        return;
    }
    if (!oskar_mem_is_complex(pattern))
    {
        *status = OSKAR_ERR_BAD_DATA_TYPE;
        return;
    }
    if (!check_gpu(pattern, status)) return;
    if (oskar_mem_length(pattern) < (size_t) (offset_out + num_points) ||
            (int) oskar_mem_length(theta) < num_points)
    {
        *status = OSKAR_ERR_DIMENSION_MISMATCH;
        return;
    }
    *status = oskar_amd_hip_status(oskar_amd_dipole_launch(precision,
            oskar_mem_is_matrix(pattern), num_points,
            oskar_mem_void_const(theta), oskar_mem_void_const(phi_x),
            oskar_mem_void_const(phi_y), kL, cos_kL, swap_xy, offset_out,
            oskar_mem_void(pattern), oskar_amd_current_stream()));
}

void oskar_blank_below_horizon(int offset_mask, int num_sources,
        const oskar_Mem* mask, int offset_out, oskar_Mem* jones, int* status)
{
    if (*status) return;
    if (oskar_mem_precision(mask) != oskar_mem_precision(jones))
    {
        *status = OSKAR_ERR_TYPE_MISMATCH;
        return;
    }
    if (oskar_mem_location(mask) != oskar_mem_location(jones))
    {
        *status = OSKAR_ERR_LOCATION_MISMATCH;
        return;
    }
    if (!oskar_mem_is_complex(jones))
    {
        *status = OSKAR_ERR_BAD_DATA_TYPE;
        return;
    }
    if (!check_gpu(jones, status)) return;
    if (oskar_mem_length(mask) < (size_t) (offset_mask + num_sources) ||
            oskar_mem_length(jones) < (size_t) (offset_out + num_sources))
    {
        *status = OSKAR_ERR_DIMENSION_MISMATCH;
        return;
    }
    *status = oskar_amd_hip_status(oskar_amd_blank_below_horizon_launch(
            oskar_mem_precision(jones), oskar_mem_is_matrix(jones),
            offset_mask, num_sources, oskar_mem_void_const(mask), offset_out,
            oskar_mem_void(jones), oskar_amd_current_stream()));
}

/* oskar_element_evaluate() for dipole and isotropic elements. */
static void element_evaluate(const oskar_Element* model, int normalise,
        int swap_xy, double orientation_x, double orientation_y,
        int offset_points, int num_points, const oskar_Mem* x, const oskar_Mem* y,
        const oskar_Mem* z, double frequency_hz, int offset_out,
        oskar_Mem* output, int* status)
{
    double dipole_length_m = 0.0;
    if (*status) return;
    const int num_points_norm = normalise ? 1 + num_points : num_points;
    oskar_mem_ensure(output, (size_t) (num_points_norm + offset_out), status);
    dipole_length_m = model->dipole_length;
    if (model->dipole_length_units == OSKAR_WAVELENGTHS)
        dipole_length_m *= (C_0 / frequency_hz);
    if (*status) return;
    if (model->element_type == OSKAR_ELEMENT_TYPE_ISOTROPIC)
        oskar_mem_set_value_real(output, 1.0, (size_t) offset_out,
                (size_t) num_points_norm, status);
    else if (model->element_type == OSKAR_ELEMENT_TYPE_DIPOLE)
    {
        /* oskar_convert_enu_directions_to_theta_phi followed by
         * oskar_evaluate_dipole_pattern, in one kernel (the reference's two
         * steps, oskar_element_evaluate.c:46-80, with theta and phi kept
         * in registers). */
        const int precision = oskar_mem_precision(output);
        const double kL = dipole_length_m * (M_PI * frequency_hz / C_0);
        if (oskar_mem_type(x) != precision || oskar_mem_type(y) != precision ||
                oskar_mem_type(z) != precision)
        {
            *status = OSKAR_ERR_TYPE_MISMATCH;
            return;
        }
        if (!check_gpu(x, status) || !check_gpu(output, status)) return;
        if (oskar_mem_length(x) < (size_t) (offset_points + num_points) ||
                oskar_mem_length(y) < (size_t) (offset_points + num_points) ||
                oskar_mem_length(z) < (size_t) (offset_points + num_points))
        {
            *status = OSKAR_ERR_DIMENSION_MISMATCH;
            return;
        }
        *status = oskar_amd_hip_status(oskar_amd_dipole_enu_launch(precision,
                oskar_mem_is_matrix(output), offset_points, num_points,
                oskar_mem_void_const(x), oskar_mem_void_const(y),
                oskar_mem_void_const(z), normalise,
                (M_PI / 2) - orientation_x, (M_PI / 2) - orientation_y, kL,
                cos(kL), swap_xy, offset_out, oskar_mem_void(output),
                oskar_amd_current_stream()));
    }
    else *status = OSKAR_ERR_FUNCTION_NOT_AVAILABLE;
    if (normalise)
        oskar_mem_normalise(output, (size_t) offset_out, (size_t) num_points,
                (size_t) (offset_out + num_points), status);
}

/* The work array for one recursion depth (oskar_station_work_beam,
 * oskar_station_work.c:256-274): depth 0 holds element patterns, depth d >= 1
 * the beams of the children of a station at depth d. */
static oskar_Mem* work_beam(oskar_StationWork* w, const oskar_Mem* beam,
        size_t length, int depth, int* status)
{
    oskar_Mem** m = 0;
    if (*status) return 0;
    if (depth == 0) m = &w->signal;
    else
    {
        if (depth >= w->num_depths)
        {
            int i = 0;
            oskar_Mem** b = (oskar_Mem**) realloc(w->depth_beam,
                    (size_t) (depth + 1) * sizeof(oskar_Mem*));
            if (!b) { *status = OSKAR_ERR_MEMORY_ALLOC_FAILURE; return 0; }
            for (i = w->num_depths; i <= depth; ++i) b[i] = 0;
            w->depth_beam = b;
            w->num_depths = depth + 1;
        }
        m = &w->depth_beam[depth];
    }
    if (*m && oskar_mem_type(*m) != oskar_mem_type(beam))
    {
        oskar_mem_free(*m, status);
        *m = 0;
    }
    if (!*m) *m = oskar_mem_create(oskar_mem_type(beam), OSKAR_GPU, 0, status);
    oskar_mem_ensure(*m, length, status);
    return *m;
}

/* At least n int zeros in the work struct (cleared when grown). */
static const oskar_Mem* work_zero_idx(oskar_StationWork* w, int n,
        int* status)
{
    if (*status) return 0;
    if (!w->zero_idx || (int) oskar_mem_length(w->zero_idx) < n)
    {
        oskar_mem_free(w->zero_idx, status);
        w->zero_idx = oskar_mem_create(OSKAR_INT, OSKAR_GPU, (size_t) n,
                status);
        oskar_mem_clear_contents(w->zero_idx, status);
    }
    return w->zero_idx;
}

/* Element weights of one station for the beam direction, with its element
 * errors and apodisation (oskar_station_evaluate_element_weights.c:15-66). */
static void station_weights(const oskar_Station* s, int feed,
        oskar_StationWork* work, double wavenumber, double beam_x,
        double beam_y, double beam_z, int time_index, int* status)
{
    if (*status) return;
    oskar_mem_ensure(work->weights, (size_t) s->num_elements, status);
    oskar_evaluate_element_weights_dft(s->num_elements,
            feed_mem(s, feed, F_MEAS), feed_mem(s, feed, F_MEAS + 1),
            feed_mem(s, feed, F_MEAS + 2), feed_mem(s, feed, F_CABLE),
            wavenumber, beam_x, beam_y, beam_z, work->weights, status);
    if ((s->apply_element_errors || s->apply_element_weight) && !*status)
    {
        const int err = s->apply_element_errors;
        const int rnd = err && s->element_errors_random;
        const int e = oskar_amd_weights_scale_launch(s->precision,
                s->num_elements, oskar_mem_void(work->weights),
                err ? oskar_mem_void_const(feed_mem(s, feed, F_GAIN)) : 0,
                rnd ? oskar_mem_void_const(feed_mem(s, feed, F_GAIN_ERR)) : 0,
                err ? oskar_mem_void_const(feed_mem(s, feed, F_PHASE)) : 0,
                rnd ? oskar_mem_void_const(feed_mem(s, feed, F_PHASE_ERR)) : 0,
                s->apply_element_weight ?
                oskar_mem_void_const(feed_mem(s, feed, F_WEIGHT)) : 0,
                s->seed_time_variable_errors, time_index, s->unique_id,
                oskar_amd_current_stream());
        if (e) *status = e;
    }
}

/* One station of a (possibly hierarchical) aperture array, for points
 * [offset_points, offset_points + num_points) of x, y, z, into beam from
 * offset_out (oskar_evaluate_station_beam_aperture_array.c:70-288, the
 * branches for element types and for child stations).  The points below
 * the horizon are blanked at the end unless `blank` is 0. */
static void aa_beam(oskar_Station* s, oskar_StationWork* work,
        int offset_points, int num_points, const oskar_Mem* x,
        const oskar_Mem* y, const oskar_Mem* z, int time_index,
        double gast_rad, double frequency_hz, int depth, int offset_out,
        oskar_Mem* beam, int blank, int* status)
{
    double beam_x = 0.0, beam_y = 0.0, beam_z = 0.0;
    oskar_Mem* signal = 0;
    int i = 0;
    if (*status) return;
    if (s->station_type != OSKAR_STATION_TYPE_AA)
    {
        /* Other station kinds are not on this library's path. */
        *status = OSKAR_ERR_FUNCTION_NOT_AVAILABLE;
        return;
    }
    if (s->mem_location != OSKAR_GPU)
    {
        *status = OSKAR_ERR_BAD_LOCATION;
        return;
    }
    if (s->precision != oskar_mem_precision(beam))
    {
        *status = OSKAR_ERR_TYPE_MISMATCH;
        return;
    }
    const double wavenumber = 2.0 * M_PI * frequency_hz / C_0;
    /* Separate beams for the X and Y feeds when the station holds feed-1
     * data and the beam is polarised (:89-90): each feed's DFT fills its
     * own row of the Jones matrix. */
    const int num_feeds = (s->common_pol_beams || !oskar_mem_is_matrix(beam)) ?
            1 : 2;
    int feed = 0;
    oskar_station_beam_horizon_direction(s, gast_rad, &beam_x, &beam_y,
            &beam_z, status);
    oskar_mem_ensure(beam, (size_t) (offset_out + num_points), status);

    if (s->child)
    {
        /* Each element is a child station: its beam is the element's
         * signal (:242-286).  Alike children are evaluated once. */
        if (!s->enable_array_pattern)
        {
            *status = OSKAR_ERR_SETTINGS_TELESCOPE;
            return;
        }
        const oskar_Mem* child_idx = 0;
        if (s->identical_children && s->num_elements > 0)
        {
            /* One evaluated child beam serves every element through an
             * all-zero data index (the reference copies it to each
             * element's row, :253-262; the DFT reads the same values). */
            signal = work_beam(work, beam, (size_t) num_points, depth,
                    status);
            aa_beam(s->child[0], work, offset_points, num_points, x, y, z,
                    time_index, gast_rad, frequency_hz, depth + 1, 0, signal,
                    1, status);
            child_idx = work_zero_idx(work, s->num_elements, status);
        }
        else
        {
            signal = work_beam(work, beam,
                    (size_t) s->num_elements * (size_t) num_points, depth,
                    status);
            for (i = 0; i < s->num_elements; ++i)
                aa_beam(s->child[i], work, offset_points, num_points, x, y, z,
                        time_index, gast_rad, frequency_hz, depth + 1,
                        i * num_points, signal, 1, status);
        }
        for (feed = 0; feed < num_feeds; ++feed)
        {
            station_weights(s, feed, work, wavenumber, beam_x, beam_y,
                    beam_z, time_index, status);
            oskar_dftw(s->normalise_array_pattern, s->num_elements,
                    wavenumber, work->weights, feed_mem(s, feed, F_TRUE),
                    feed_mem(s, feed, F_TRUE + 1),
                    feed_mem(s, feed, F_TRUE + 2), offset_points, num_points,
                    x, y, s->array_is_3d ? z : 0, child_idx, signal,
                    feed == 0 || num_feeds == 1, feed == 1 || num_feeds == 1,
                    offset_out, beam, status);
        }
        return;
    }

    const int num_types = s->num_element_types;
    const oskar_Mem* data_idx = 0;
    if (s->common_element_orientation)
    {
        /* Element patterns per element type: signal[type][num_points (+1)],
         * the DFT picking each element's row by its type (:157-177). */
        data_idx = s->element_types;
        signal = work_beam(work, beam,
                (size_t) num_types * (size_t) (num_points + 1), 0, status);
        for (i = 0; i < num_types; ++i)
        {
            element_evaluate(s->element[i], s->normalise_element_pattern,
                    s->swap_xy,
                    oskar_station_element_euler_index_rad(s, 0, 0, 0) +
                    M_PI / 2, oskar_station_element_euler_index_rad(s, 1, 0, 0),
                    offset_points, num_points, x, y, z, frequency_hz,
                    i * num_points,
                    signal, status);
        }
    }
    else
    {
        /* Elements with their own orientations: one pattern per element,
         * of its type, turned by its own Euler angles; row j of the signal
         * is element j's (:178-201). */
        const int* type = (const int*) oskar_mem_void_const(
                s->element_types_cpu);
        signal = work_beam(work, beam,
                (size_t) s->num_elements * (size_t) (num_points + 1), 0,
                status);
        for (i = 0; i < s->num_elements && !*status; ++i)
        {
            if (type[i] < 0 || type[i] >= num_types)
            {
                *status = OSKAR_ERR_OUT_OF_RANGE;
                break;
            }
            element_evaluate(s->element[type[i]], s->normalise_element_pattern,
                    s->swap_xy,
                    oskar_station_element_euler_index_rad(s, 0, 0, i) +
                    M_PI / 2, oskar_station_element_euler_index_rad(s, 1, 0, i),
                    offset_points, num_points, x, y, z, frequency_hz,
                    i * num_points,
                    signal, status);
        }
    }
    if (s->enable_array_pattern)
    {
        for (feed = 0; feed < num_feeds; ++feed)
        {
            station_weights(s, feed, work, wavenumber, beam_x, beam_y,
                    beam_z, time_index, status);
            oskar_dftw(s->normalise_array_pattern, s->num_elements,
                    wavenumber, work->weights, feed_mem(s, feed, F_TRUE),
                    feed_mem(s, feed, F_TRUE + 1),
                    feed_mem(s, feed, F_TRUE + 2), offset_points, num_points,
                    x, y, s->array_is_3d ? z : 0, data_idx, signal,
                    feed == 0 || num_feeds == 1, feed == 1 || num_feeds == 1,
                    offset_out, beam, status);
        }
    }
    else
    {
        /* Without the array pattern the beam is the first row's pattern
         * (:224-229). */
        if (num_types == 1)
            oskar_mem_copy_contents(beam, signal, (size_t) offset_out, 0,
                    (size_t) num_points, status);
        else
        {
            *status = OSKAR_ERR_SETTINGS_TELESCOPE;
            return;
        }
    }
    if (blank)
        oskar_blank_below_horizon(offset_points, num_points, z, offset_out,
                beam, status);
}

int oskar_amd_station_beam_aa_fused_covers(const oskar_Station* s,
        const oskar_Mem* beam)
{
    int i = 0;
    if (!s || s->station_type != OSKAR_STATION_TYPE_AA || s->child ||
            s->mem_location != OSKAR_GPU ||
            !s->common_element_orientation || s->num_element_types != 1 ||
            !s->enable_array_pattern || s->normalise_element_pattern ||
            s->num_elements <= 0 || !s->element || !s->element[0] ||
            s->element[0]->element_type != OSKAR_ELEMENT_TYPE_DIPOLE ||
            s->precision != oskar_mem_precision(beam) ||
            !oskar_mem_is_complex(beam))
        return 0;
    /* One feed beam for both rows of a matrix (:89-90). */
    if (!s->common_pol_beams && oskar_mem_is_matrix(beam)) return 0;
    /* Every element's data row is the one pattern. */
    if (!s->element_types_cpu ||
            (int) oskar_mem_length(s->element_types_cpu) < s->num_elements)
        return 0;
    {
        const int* type = (const int*) oskar_mem_void_const(
                s->element_types_cpu);
        for (i = 0; i < s->num_elements; ++i)
            if (type[i] != 0) return 0;
    }
    return 1;
}

void oskar_amd_station_beam_aa_fused(oskar_Station* s,
        oskar_StationWork* work, int num_points, const oskar_Mem* x,
        const oskar_Mem* y, const oskar_Mem* z, int normalise, double nx,
        double ny, double nz, int time_index, double gast_rad,
        double frequency_hz, int offset_out, oskar_Mem* beam, int* status)
{
    double beam_x = 0.0, beam_y = 0.0, beam_z = 0.0;
    int i = 0;
    if (*status) return;
    if (!oskar_amd_station_beam_aa_fused_covers(s, beam))
    {
        *status = OSKAR_ERR_FUNCTION_NOT_AVAILABLE;
        return;
    }
    for (i = 0; i < 3; ++i)
    {
        const oskar_Mem* d = i == 0 ? x : (i == 1 ? y : z);
        if (oskar_mem_type(d) != s->precision)
        {
            *status = OSKAR_ERR_TYPE_MISMATCH;
            return;
        }
        if (!check_gpu(d, status)) return;
        if (oskar_mem_length(d) < (size_t) num_points)
        {
            *status = OSKAR_ERR_DIMENSION_MISMATCH;
            return;
        }
    }
    const double wavenumber = 2.0 * M_PI * frequency_hz / C_0;
    oskar_station_beam_horizon_direction(s, gast_rad, &beam_x, &beam_y,
            &beam_z, status);
    oskar_mem_ensure(beam, (size_t) offset_out + (size_t) num_points, status);
    station_weights(s, 0, work, wavenumber, beam_x, beam_y, beam_z,
            time_index, status);
    if (*status || num_points == 0) return;
    {
        /* The element pattern's parameters as element_evaluate forms them
         * for oskar_convert_enu_directions_to_theta_phi and
         * oskar_evaluate_dipole_pattern. */
        const oskar_Element* model = s->element[0];
        double dipole_length_m = model->dipole_length;
        if (model->dipole_length_units == OSKAR_WAVELENGTHS)
            dipole_length_m *= (C_0 / frequency_hz);
        const double kL = dipole_length_m * (M_PI * frequency_hz / C_0);
        const double orientation_x =
                oskar_station_element_euler_index_rad(s, 0, 0, 0) + M_PI / 2;
        const double orientation_y =
                oskar_station_element_euler_index_rad(s, 1, 0, 0);
        const size_t es = oskar_mem_element_size(oskar_mem_type(beam));
        *status = oskar_amd_hip_status(oskar_amd_aa_fused_launch(
                s->precision, oskar_mem_is_matrix(beam), s->array_is_3d,
                normalise, s->num_elements, wavenumber,
                oskar_mem_void_const(work->weights),
                oskar_mem_void_const(feed_mem(s, 0, F_TRUE)),
                oskar_mem_void_const(feed_mem(s, 0, F_TRUE + 1)),
                oskar_mem_void_const(feed_mem(s, 0, F_TRUE + 2)),
                num_points, oskar_mem_void_const(x), oskar_mem_void_const(y),
                oskar_mem_void_const(z), nx, ny, nz, kL, cos(kL),
                (M_PI / 2) - orientation_x, (M_PI / 2) - orientation_y,
                s->swap_xy,
                s->normalise_array_pattern ? 1.0 / s->num_elements : 1.0,
                (char*) oskar_mem_void(beam) + (size_t) offset_out * es,
                oskar_amd_current_stream()));
    }
}

/* Points per pass through a hierarchical station: each depth holds one
 * beam per child for every point of the pass (the reference's
 * MAX_CHUNK_SIZE, oskar_evaluate_station_beam_aperture_array.c:21, is 49152;
 * a larger pass suits a GPU with HBM to spare and changes no result, since
 * every point is evaluated on its own).  The pass is at most
 * CHILD_PASS_POINTS, and smaller when the most children of any station
 * would make a depth's beams exceed CHILD_PASS_BYTES or a child offset
 * (child index x points, an int) overflow. */
#define CHILD_PASS_POINTS (1 << 20)
#define CHILD_PASS_BYTES ((size_t) 1 << 30)

/* Most children of any station in the tree (1 for a flat station). */
static int max_children(const oskar_Station* s)
{
    int i = 0, m = 1;
    if (!s->child) return 1;
    m = s->num_elements > 1 ? s->num_elements : 1;
    for (i = 0; i < s->num_elements; ++i)
    {
        const int c = max_children(s->child[i]);
        if (c > m) m = c;
        if (s->identical_children) break;
    }
    return m;
}

static void aa_entry(oskar_Station* s, oskar_StationWork* work,
        int num_points, const oskar_Mem* x, const oskar_Mem* y,
        const oskar_Mem* z, int time_index, double gast_rad,
        double frequency_hz, oskar_Mem* beam, int blank, int* status)
{
    int start = 0;
    if (*status) return;
    if (oskar_mem_location(beam) != OSKAR_GPU ||
            s->mem_location != OSKAR_GPU || work->mem_location != OSKAR_GPU)
    {
        *status = OSKAR_ERR_BAD_LOCATION;
        return;
    }
    if (!oskar_mem_is_complex(beam) ||
            oskar_mem_precision(beam) != s->precision)
    {
        *status = OSKAR_ERR_TYPE_MISMATCH;
        return;
    }
    if (!s->child)
    {
        aa_beam(s, work, 0, num_points, x, y, z, time_index, gast_rad,
                frequency_hz, 0, 0, beam, blank, status);
        return;
    }
    /* Depth 0 is the element level, so the recursion starts at 1 (:52-66). */
    const int children = max_children(s);
    const size_t by_bytes = CHILD_PASS_BYTES /
            ((size_t) children * oskar_mem_element_size(oskar_mem_type(beam)));
    int pass = CHILD_PASS_POINTS;
    if ((size_t) pass > by_bytes) pass = (int) by_bytes;
    if (pass > INT_MAX / children) pass = INT_MAX / children;
    if (pass < 1) pass = 1;
    for (start = 0; start < num_points; start += pass)
    {
        const int n = (num_points - start < pass) ? num_points - start : pass;
        aa_beam(s, work, start, n, x, y, z, time_index, gast_rad,
                frequency_hz, 1, start, beam, blank, status);
    }
}

void oskar_evaluate_station_beam_aperture_array(oskar_Station* s,
        oskar_StationWork* work, int num_points, const oskar_Mem* x,
        const oskar_Mem* y, const oskar_Mem* z, int time_index,
        double gast_rad, double frequency_hz, oskar_Mem* beam, int* status)
{
    aa_entry(s, work, num_points, x, y, z, time_index, gast_rad,
            frequency_hz, beam, 1, status);
}

void oskar_amd_station_beam_aa_unblanked(oskar_Station* s,
        oskar_StationWork* work, int num_points, const oskar_Mem* x,
        const oskar_Mem* y, const oskar_Mem* z, int time_index,
        double gast_rad, double frequency_hz, oskar_Mem* beam, int* status)
{
    aa_entry(s, work, num_points, x, y, z, time_index, gast_rad,
            frequency_hz, beam, 0, status);
}
