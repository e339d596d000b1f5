/*
 * oskar_Station (aperture-array subset), oskar_Element and oskar_StationWork.
 *
 * Reference: struct oskar/telescope/station/private_station.h:22-105 and
 * oskar/telescope/station/element/private_element.h; defaults from
 * oskar/telescope/station/src/oskar_station_create.c:20-120 (type AA,
 * normalise_final_beam TRUE, beam coordinates RADEC, enable_array_pattern
 * TRUE, element type dipole of 0.5 wavelengths).
 *
 * Supported by this library: one feed of element coordinates shared by both
 * polarisations (common_pol_beams), element types of kind dipole or
 * isotropic with a common orientation, optional cable-length errors, array
 * and element normalisation flags, swap_xy and 2D/3D arrays.  Child
 * (hierarchical) stations, HARP coefficients, spherical-wave element data,
 * tapers, time-variable element errors, gain tables and apodisation weights
 * are not on the hot path: oskar_evaluate_station_beam_aperture_array()
 * returns OSKAR_ERR_FUNCTION_NOT_AVAILABLE if a station needs them.
 */
#ifndef OSKAR_STATION_H_
#define OSKAR_STATION_H_

#include "oskar_mem.h"

#ifdef __cplusplus
extern "C" {
#endif

struct oskar_Station;
struct oskar_Element;
struct oskar_StationWork;
#ifndef OSKAR_STATION_TYPEDEF_
#define OSKAR_STATION_TYPEDEF_
typedef struct oskar_Station oskar_Station;
#endif
#ifndef OSKAR_ELEMENT_TYPEDEF_
#define OSKAR_ELEMENT_TYPEDEF_
typedef struct oskar_Element oskar_Element;
#endif
#ifndef OSKAR_STATION_WORK_TYPEDEF_
#define OSKAR_STATION_WORK_TYPEDEF_
typedef struct oskar_StationWork oskar_StationWork;
#endif

enum OSKAR_STATION_TYPE
{
    OSKAR_STATION_TYPE_AA,
    OSKAR_STATION_TYPE_ISOTROPIC,
    OSKAR_STATION_TYPE_GAUSSIAN_BEAM,
    OSKAR_STATION_TYPE_VLA_PBCOR
};

enum OSKAR_ELEMENT_TYPE
{
    OSKAR_ELEMENT_TYPE_DIPOLE,
    OSKAR_ELEMENT_TYPE_ISOTROPIC
};

/* ---- Station model ------------------------------------------------------ */
/* oskar/telescope/station/oskar_station_create.h */
OSKAR_EXPORT
oskar_Station* oskar_station_create(int type, int location, int num_elements,
        int* status);
/* oskar/telescope/station/oskar_station_create_copy.h */
OSKAR_EXPORT
oskar_Station* oskar_station_create_copy(const oskar_Station* src,
        int location, int* status);
OSKAR_EXPORT void oskar_station_free(oskar_Station* model, int* status);
OSKAR_EXPORT void oskar_station_resize(oskar_Station* station,
        int num_elements, int* status);
/* oskar/telescope/station/oskar_station_analyse.h: recomputes array_is_3d
 * and common_element_orientation from the element data. */
OSKAR_EXPORT void oskar_station_analyse(oskar_Station* station,
        int* finished_identical_station_check, int* status);

/* oskar/telescope/station/oskar_station_set_element_coords.h */
OSKAR_EXPORT
void oskar_station_set_element_coords(oskar_Station* station, int feed,
        int index, const double measured_enu[3], const double true_enu[3],
        int* status);
/* oskar/telescope/station/src/oskar_station_set_element_errors.c:16-80:
 * element gain, gain error scale, phase offset and phase error scale (the
 * phases in degrees).  Feed 0 only (OSKAR_ERR_FUNCTION_NOT_AVAILABLE for
 * feed 1, whose elements share feed 0's values here).  Arrays start at gain
 * 1 and zero phase and errors.  After oskar_station_analyse, the beam
 * multiplies each element's DFT weight by a (cos p, sin p) with
 * a = gain + g gain_error and p = phase_offset + h phase_error, where
 * (g, h) are the element's Gaussian draws of the reference's generator
 * (Philox4x32-10 keyed by the station's seed, counter (element / 2,
 * time index, station unique id, 0x12345678), Box-Muller), zero when no
 * error scale is set. */
OSKAR_EXPORT
void oskar_station_set_element_errors(oskar_Station* station, int feed,
        int index, double gain, double gain_error, double phase_offset_deg,
        double phase_error_deg, int* status);

/* oskar/telescope/station/src/oskar_station_set_element_weight.c:13-28:
 * complex apodisation weight of one element (feed 0; default 1 + 0i),
 * multiplied into its DFT weight after the errors. */
OSKAR_EXPORT
void oskar_station_set_element_weight(oskar_Station* station, int feed,
        int index, double re, double im, int* status);

/* oskar/telescope/station/src/oskar_station_accessors.c:21-24, 221-224,
 * 503-513, 615-620: the seed of the time-variable element errors (default
 * 1) and the station's unique id (the next value of *counter), both part of
 * the random draws' key and counter. */
OSKAR_EXPORT unsigned int oskar_station_seed_time_variable_errors(
        const oskar_Station* station);
OSKAR_EXPORT void oskar_station_set_seed_time_variable_errors(
        oskar_Station* station, unsigned int value);
OSKAR_EXPORT int oskar_station_unique_id(const oskar_Station* station);
OSKAR_EXPORT void oskar_station_set_unique_ids(oskar_Station* station,
        int* counter);

/* Hierarchical stations, whose elements are child stations:
 * oskar/telescope/station/src/oskar_station_create_child_stations.c:15-44
 * (one empty child per element; OSKAR_ERR_MEMORY_ALLOC_FAILURE if children
 * exist), oskar_station_accessors.c:166-168 and 422-436 (accessors), and
 * oskar_station_different.c:15-200 (nonzero if two stations differ in
 * anything but their unique id, children compared recursively).
 * oskar_station_analyse sets identical_children when every child equals the
 * first and no station below has random element errors; the aperture-array
 * beam then evaluates one child and copies it. */
OSKAR_EXPORT void oskar_station_create_child_stations(oskar_Station* station,
        int* status);
OSKAR_EXPORT int oskar_station_has_child(const oskar_Station* station);
OSKAR_EXPORT oskar_Station* oskar_station_child(oskar_Station* station,
        int i);
OSKAR_EXPORT const oskar_Station* oskar_station_child_const(
        const oskar_Station* station, int i);
OSKAR_EXPORT int oskar_station_identical_children(
        const oskar_Station* station);
OSKAR_EXPORT int oskar_station_different(const oskar_Station* a,
        const oskar_Station* b, int* status);

/* oskar/telescope/station/src/oskar_station_accessors.c:211-219: set by
 * oskar_station_analyse (any gain != 1 or phase offset != 0 or error scale
 * != 0; any apodisation weight != 1 + 0i). */
OSKAR_EXPORT int oskar_station_apply_element_errors(
        const oskar_Station* station);
OSKAR_EXPORT int oskar_station_apply_element_weight(
        const oskar_Station* station);
/* oskar/telescope/station/oskar_station_accessors.h:235-268: the element
 * gain, gain error, phase offset, phase error and apodisation weight arrays
 * of a feed (feed 1 shares feed 0's until set for feed 1; NULL until a
 * value of the field is set). */
OSKAR_EXPORT const oskar_Mem* oskar_station_element_gain_const(
        const oskar_Station* station, int feed);
OSKAR_EXPORT const oskar_Mem* oskar_station_element_gain_error_const(
        const oskar_Station* station, int feed);
OSKAR_EXPORT const oskar_Mem* oskar_station_element_phase_offset_rad_const(
        const oskar_Station* station, int feed);
OSKAR_EXPORT const oskar_Mem* oskar_station_element_phase_error_rad_const(
        const oskar_Station* station, int feed);
OSKAR_EXPORT const oskar_Mem* oskar_station_element_weight_const(
        const oskar_Station* station, int feed);

/* oskar/telescope/station/oskar_station_set_element_cable_length_error.h */
OSKAR_EXPORT
void oskar_station_set_element_cable_length_error(oskar_Station* station,
        int feed, int index, const double error_metres, int* status);
/* oskar/telescope/station/oskar_station_set_element_feed_angle.h */
OSKAR_EXPORT
void oskar_station_set_element_feed_angle(oskar_Station* station, int feed,
        int index, double alpha_deg, double beta_deg, double gamma_deg,
        int* status);
/* oskar/telescope/station/oskar_station_set_element_type.h */
OSKAR_EXPORT
void oskar_station_set_element_type(oskar_Station* dst, int index,
        int element_type, int* status);
/* oskar/telescope/station/oskar_station_resize_element_types.h */
OSKAR_EXPORT
void oskar_station_resize_element_types(oskar_Station* model,
        int num_element_types, int* status);

/* Setters, oskar/telescope/station/oskar_station_accessors.h */
OSKAR_EXPORT void oskar_station_set_station_type(oskar_Station* model,
        int type);
OSKAR_EXPORT void oskar_station_set_normalise_final_beam(oskar_Station* m,
        int value);
OSKAR_EXPORT void oskar_station_set_position(oskar_Station* model,
        double longitude_rad, double latitude_rad, double altitude_m,
        double offset_ecef_x, double offset_ecef_y, double offset_ecef_z);
OSKAR_EXPORT void oskar_station_set_phase_centre(oskar_Station* model,
        int coord_type, double longitude_rad, double latitude_rad);
OSKAR_EXPORT void oskar_station_set_normalise_array_pattern(oskar_Station* m,
        int value);
OSKAR_EXPORT void oskar_station_set_normalise_element_pattern(
        oskar_Station* model, int value);
OSKAR_EXPORT void oskar_station_set_enable_array_pattern(oskar_Station* m,
        int value);
OSKAR_EXPORT void oskar_station_set_swap_xy(oskar_Station* model, int value);

/* Accessors, oskar/telescope/station/oskar_station_accessors.h */
OSKAR_EXPORT int oskar_station_precision(const oskar_Station* model);
OSKAR_EXPORT int oskar_station_mem_location(const oskar_Station* model);
OSKAR_EXPORT int oskar_station_type(const oskar_Station* model);
OSKAR_EXPORT int oskar_station_num_elements(const oskar_Station* model);
OSKAR_EXPORT int oskar_station_num_element_types(const oskar_Station* m);
OSKAR_EXPORT int oskar_station_normalise_final_beam(const oskar_Station* m);
OSKAR_EXPORT int oskar_station_normalise_array_pattern(const oskar_Station* m);
OSKAR_EXPORT
int oskar_station_normalise_element_pattern(const oskar_Station* model);
OSKAR_EXPORT int oskar_station_enable_array_pattern(const oskar_Station* m);
OSKAR_EXPORT int oskar_station_swap_xy(const oskar_Station* model);
OSKAR_EXPORT int oskar_station_array_is_3d(const oskar_Station* model);
OSKAR_EXPORT
int oskar_station_common_element_orientation(const oskar_Station* model);
OSKAR_EXPORT int oskar_station_common_pol_beams(const oskar_Station* model);
OSKAR_EXPORT double oskar_station_lon_rad(const oskar_Station* model);
OSKAR_EXPORT double oskar_station_lat_rad(const oskar_Station* model);
OSKAR_EXPORT double oskar_station_beam_lon_rad(const oskar_Station* model);
OSKAR_EXPORT double oskar_station_beam_lat_rad(const oskar_Station* model);
OSKAR_EXPORT int oskar_station_beam_coord_type(const oskar_Station* model);
OSKAR_EXPORT
oskar_Mem* oskar_station_element_true_enu_metres(oskar_Station* model,
        int feed, int dim);
OSKAR_EXPORT
const oskar_Mem* oskar_station_element_true_enu_metres_const(
        const oskar_Station* model, int feed, int dim);
OSKAR_EXPORT
oskar_Mem* oskar_station_element_measured_enu_metres(oskar_Station* model,
        int feed, int dim);
OSKAR_EXPORT
const oskar_Mem* oskar_station_element_measured_enu_metres_const(
        const oskar_Station* model, int feed, int dim);
OSKAR_EXPORT
oskar_Mem* oskar_station_element_cable_length_error_metres(
        oskar_Station* model, int feed);
OSKAR_EXPORT
const oskar_Mem* oskar_station_element_cable_length_error_metres_const(
        const oskar_Station* model, int feed);
OSKAR_EXPORT oskar_Mem* oskar_station_element_types(oskar_Station* model);
OSKAR_EXPORT
const oskar_Mem* oskar_station_element_types_const(const oskar_Station* m);
OSKAR_EXPORT
double oskar_station_element_euler_index_rad(const oskar_Station* model,
        int feed, int dim, int index);
OSKAR_EXPORT oskar_Element* oskar_station_element(oskar_Station* model,
        int element_type_index);
OSKAR_EXPORT
const oskar_Element* oskar_station_element_const(const oskar_Station* model,
        int element_type_index);

/* ---- Element model (oskar/telescope/station/element/) ------------------ */
/* type: "Dipole" or "Isotropic" (first letter, case-insensitive). */
OSKAR_EXPORT void oskar_element_set_element_type(oskar_Element* data,
        const char* type, int* status);
/* units: "Metres" or "Wavelengths" (first letter, case-insensitive). */
OSKAR_EXPORT void oskar_element_set_dipole_length(oskar_Element* data,
        double value, const char* units, int* status);
OSKAR_EXPORT int oskar_element_type(const oskar_Element* data);
OSKAR_EXPORT double oskar_element_dipole_length(const oskar_Element* data);
OSKAR_EXPORT int oskar_element_dipole_length_units(const oskar_Element* d);

/* ---- Work buffers (oskar/telescope/station/oskar_station_work.h) ------- */
OSKAR_EXPORT
oskar_StationWork* oskar_station_work_create(int type, int location,
        int* status);
OSKAR_EXPORT void oskar_station_work_free(oskar_StationWork* work,
        int* status);
/* Integer work arrays of the horizon clip
 * (oskar/telescope/station/src/oskar_station_work.c:110-118):
 * the mask of sources above the horizon of any station model, and its
 * exclusive prefix sum with the total after the last source. */
OSKAR_EXPORT oskar_Mem* oskar_station_work_horizon_mask(
        oskar_StationWork* work);
OSKAR_EXPORT oskar_Mem* oskar_station_work_source_indices(
        oskar_StationWork* work);

/* ---- Permitted beam directions ----------------------------------------- */
/* oskar/telescope/station/src/oskar_station_accessors.c:455-470: the number
 * of permitted (azimuth, elevation) beam directions (0: any direction) and
 * their values in radians (double, host memory; NULL before a load). */
OSKAR_EXPORT int oskar_station_num_permitted_beams(const oskar_Station* model);
OSKAR_EXPORT const oskar_Mem* oskar_station_permitted_beam_az_rad_const(
        const oskar_Station* model);
OSKAR_EXPORT const oskar_Mem* oskar_station_permitted_beam_el_rad_const(
        const oskar_Station* model);

/* oskar/telescope/station/src/oskar_station_load_permitted_beams.c: one
 * direction per line, azimuth and elevation in degrees as the first two
 * numbers (separated by commas, spaces or tabs; '#' starts a comment; lines
 * with fewer than two numbers are skipped).  Replaces any loaded set.
 * OSKAR_ERR_FILE_IO if the file cannot be opened. */
OSKAR_EXPORT void oskar_station_load_permitted_beams(oskar_Station* station,
        const char* filename, int* status);

/* ---- Station-beam evaluation ------------------------------------------- */
/* oskar/telescope/station/oskar_station_beam_horizon_direction.h:
 * beam direction cosines (RADEC via the apparent-RA/Dec formula, or AZEL);
 * with permitted beams, the permitted (az, el) nearest on the sphere
 * (haversine distance, first of equals) replaces the direction. */
OSKAR_EXPORT
void oskar_station_beam_horizon_direction(const oskar_Station* station,
        const double gast_rad, double* x, double* y, double* z, int* status);

/* oskar/telescope/station/src/oskar_evaluate_station_beam_aperture_array.c
 * :31-42.  Element patterns per type -> DFT weights from measured element
 * positions -> oskar_dftw over true element positions -> zero below the
 * horizon.  The beam is written to beam[0 .. num_points). */
OSKAR_EXPORT
void oskar_evaluate_station_beam_aperture_array(oskar_Station* station,
        oskar_StationWork* work, int num_points, const oskar_Mem* x,
        const oskar_Mem* y, const oskar_Mem* z, int time_index,
        double gast_rad, double frequency_hz, oskar_Mem* beam, int* status);

/* oskar/telescope/station/oskar_evaluate_element_weights_dft.h:
 * w_i = exp(-i k (x_i x_b + y_i y_b + z_i z_b + cable_i)). */
OSKAR_EXPORT
void oskar_evaluate_element_weights_dft(int num_elements,
        const oskar_Mem* x, const oskar_Mem* y, const oskar_Mem* z,
        const oskar_Mem* cable_length_error, double wavenumber,
        double x_beam, double y_beam, double z_beam, oskar_Mem* weights,
        int* status);

/* oskar/telescope/station/element/oskar_evaluate_dipole_pattern.h */
OSKAR_EXPORT
void oskar_evaluate_dipole_pattern(int num_points, const oskar_Mem* theta,
        const oskar_Mem* phi_x, const oskar_Mem* phi_y, double freq_hz,
        double dipole_length_m, int swap_xy, int offset_out,
        oskar_Mem* pattern, int* status);

/* oskar/convert/oskar_convert_enu_directions_to_theta_phi.h */
OSKAR_EXPORT
void oskar_convert_enu_directions_to_theta_phi(int offset_in, int num_points,
        const oskar_Mem* x, const oskar_Mem* y, const oskar_Mem* z,
        int extra_point_at_pole, double delta_phi1, double delta_phi2,
        oskar_Mem* theta, oskar_Mem* phi1, oskar_Mem* phi2, int* status);

/* oskar/telescope/station/oskar_blank_below_horizon.h: beam = 0 where the
 * z direction cosine is negative. */
OSKAR_EXPORT
void oskar_blank_below_horizon(int offset_mask, int num_sources,
        const oskar_Mem* mask, int offset_out, oskar_Mem* jones, int* status);

/* oskar/telescope/station/oskar_station_beam.h
 * (src/oskar_station_beam.c:37-251): the beam of one station for sources in
 * any supported frame (OSKAR_COORDS_RADEC, _AZEL, _ENU_DIR, _REL_DIR about
 * (ref_lon_rad, ref_lat_rad)) at one time (UTC MJD; GAST from
 * oskar_convert_mjd_to_gast_fast) and frequency, written to
 * beam[offset_out ..].  With the station's normalise_final_beam flag (and a
 * non-isotropic station) every value is divided by the beam in the
 * normalisation direction (norm_coord_type: RADEC or AZEL).  If the beam
 * direction is below the horizon the output is zero.  Station types:
 * aperture array and isotropic; the Gaussian and VLA primary beams and the
 * ionospheric screen return OSKAR_ERR_FUNCTION_NOT_AVAILABLE. */
OSKAR_EXPORT
void oskar_station_beam(oskar_Station* station, oskar_StationWork* work,
        int source_coord_type, int num_points,
        const oskar_Mem* const source_coords[3], double ref_lon_rad,
        double ref_lat_rad, int norm_coord_type, double norm_lon_rad,
        double norm_lat_rad, int time_index, double time_start_mjd_utc,
        double time_mjd_utc, double frequency_hz, int offset_out,
        oskar_Mem* beam, int* status);

#ifdef __cplusplus
}
#endif

#endif /* OSKAR_STATION_H_ */
