/*
 * Private structures and internal launch interfaces of the library.
 * Host code is C99; kernels live in the .hip files and are reached through
 * the extern "C" launchers declared at the bottom of this file.
 */
#ifndef OSKAR_AMD_PRIVATE_H_
#define OSKAR_AMD_PRIVATE_H_

#include "oskar_global.h"
#include "oskar_mem.h"
#include "oskar_jones.h"
#include "oskar_telescope.h"
#include "oskar_station.h"
#include "oskar_sky.h"
#include "oskar_vis.h"
#include "oskar_timer.h"

#ifdef __cplusplus
extern "C" {
#endif

/* oskar/mem/private_mem.h:24-38 (the mutex is dropped: reference counting is
 * only touched from the owning host thread in this library). */
struct oskar_Mem
{
    int type;
    int location;
    size_t num_elements;
    int owner;
    void* data;
    int ref_count;
};

/* oskar/interferometer/private_jones.h:34-41 */
struct oskar_Jones
{
    int num_stations;
    int num_sources;
    int cap_stations;
    int cap_sources;
    oskar_Mem* data;
};

/* The fields of oskar/telescope/private_telescope.h:13-68 that the
 * correlator reads. */
struct oskar_Telescope
{
    int precision;
    int mem_location;
    int num_stations;
    int phase_centre_coord_type;
    double phase_centre_rad[2];
    double channel_bandwidth_hz;
    double time_average_sec;
    double uv_filter_min;
    double uv_filter_max;
    int uv_filter_units;
    oskar_Mem* station_true_offset_ecef_metres[3];
    int num_station_models;      /* private_telescope.h:54 */
    oskar_Station** station;     /* station models */
    oskar_Mem* station_type_map; /* CPU int, num_stations (private_telescope.h:61) */
    int allow_station_beam_duplication;
};

/* oskar/telescope/station/element/private_element.h (dipole/isotropic). */
struct oskar_Element
{
    int precision;
    int mem_location;
    int element_type;
    double dipole_length;
    int dipole_length_units;
};

/* Aperture-array subset of oskar/telescope/station/private_station.h. */
struct oskar_Station
{
    int precision;
    int mem_location;
    int station_type;
    int normalise_final_beam;
    double offset_ecef[3];
    double lon_rad, lat_rad, alt_metres;
    double beam_lon_rad, beam_lat_rad;
    int beam_coord_type;
    int num_elements;
    int num_element_types;
    int normalise_array_pattern;
    int normalise_element_pattern;
    int enable_array_pattern;
    int common_element_orientation;
    int common_pol_beams;
    int swap_xy;
    int array_is_3d;
    oskar_Mem* element_true_enu_metres[3];
    oskar_Mem* element_measured_enu_metres[3];
    oskar_Mem* element_cable_length_error;
    oskar_Mem* element_euler_cpu[2][3];
    oskar_Mem* element_types;     /* int, station location */
    oskar_Mem* element_types_cpu; /* int, host */
    oskar_Element** element;
    /* Per-element gain, phase offset, the scales of their time-variable
     * errors, and apodisation weights (private_station.h element_gain,
     * element_gain_error, element_phase_offset_rad, element_phase_error_rad,
     * element_weight; feed 0; NULL until first set).  The apply flags are
     * set by oskar_station_analyse. */
    oskar_Mem* element_gain;
    oskar_Mem* element_gain_error;
    oskar_Mem* element_phase_offset_rad;
    oskar_Mem* element_phase_error_rad;
    oskar_Mem* element_weight;        /* complex */
    int apply_element_errors;
    int element_errors_random;        /* a gain or phase error scale != 0 */
    unsigned int seed_time_variable_errors;   /* private_station.h; 1 */
    int unique_id;
    int apply_element_weight;
    int num_permitted_beams;          /* private_station.h:91-93 */
    oskar_Mem* permitted_beam_az_rad; /* double, host; NULL until loaded */
    oskar_Mem* permitted_beam_el_rad;
    /* Child stations, one per element, or NULL (private_station.h child,
     * identical_children): an element of a hierarchical station is itself
     * a station whose beam takes the place of an element pattern. */
    struct oskar_Station** child;
    int identical_children;
    /* Feed-1 (Y) element data (private_station.h arrays [1]), NULL until set
     * for feed 1: while all are NULL the two feeds share feed 0's data and
     * the station has common polarisation beams. */
    oskar_Mem* y_true_enu_metres[3];
    oskar_Mem* y_measured_enu_metres[3];
    oskar_Mem* y_cable_length_error;
    oskar_Mem* y_gain;
    oskar_Mem* y_gain_error;
    oskar_Mem* y_phase_offset_rad;
    oskar_Mem* y_phase_error_rad;
    oskar_Mem* y_weight;
};

/* oskar/telescope/station/private_station_work.h (subset). */
struct oskar_StationWork
{
    int precision;
    int mem_location;
    oskar_Mem* weights;
    oskar_Mem* signal;
    oskar_Mem* theta;
    oskar_Mem* phi_x;
    oskar_Mem* phi_y;
    oskar_Mem* enu[3];        /* source directions (+1 normalisation point) */
    oskar_Mem* beam_out;      /* beam of those directions */
    oskar_Mem* horizon_mask;  /* int, one per sky source */
    oskar_Mem* source_indices;/* int, one per sky source + 1 (the total) */
    /* Child-station beams per recursion depth >= 1 (private_station_work.h
     * beam[], num_depths); depth 0 is the element-pattern signal above. */
    oskar_Mem** depth_beam;
    int num_depths;
    /* int zeros, one per child of a station whose children are alike: the
     * data index that makes every child read the one evaluated beam. */
    oskar_Mem* zero_idx;
};

/* oskar/sky/private_sky.h: one table of num_columns x capacity values
 * (column c at offset c x capacity), per-row valid-column counts
 * (capacity x OSKAR_SKY_NUM_FIXED_COLUMN_TYPES ints), column aliases,
 * types and attributes. */
struct oskar_Sky
{
    double attr_double[OSKAR_SKY_NUM_ATTRIBUTES_DOUBLE];
    int attr_int[OSKAR_SKY_NUM_ATTRIBUTES_INT];
    oskar_Mem* table;
    oskar_Mem* num_valid_columns;
    oskar_Mem** columns;
    oskar_SkyColumn* column_type;
    int* column_attr;
};

/* oskar/vis/private_vis_block.h */
struct oskar_VisBlock
{
    int dim_start_size[6];   /* start time, start channel, times, channels,
                                baselines, stations */
    int has_cross_correlations;
    int has_auto_correlations;
    oskar_Mem* cross_correlations;
    oskar_Mem* auto_correlations;
    oskar_Mem* baseline_uvw_metres[3];
    oskar_Mem* station_uvw_metres[3];
};

/* The fields of oskar/vis/private_vis_header.h that the .vis format holds. */
struct oskar_VisHeader
{
    oskar_Mem* settings;
    oskar_Mem* telescope_path;
    int num_tags_header;
    int num_tags_per_block;
    int write_autocorr, write_crosscorr;
    int amp_type, coord_precision;
    int max_times_per_block, num_times_total;
    int max_channels_per_block, num_channels_total;
    int num_stations;
    int pol_type;
    int casa_phase_convention;
    int phase_centre_type;
    double phase_centre_deg[2];
    double freq_start_hz, freq_inc_hz, channel_bandwidth_hz;
    double time_start_mjd_utc, time_inc_sec, time_average_sec;
    double telescope_centre_lon_deg, telescope_centre_lat_deg;
    double telescope_centre_alt_m;
    oskar_Mem* station_diameter_m;
    oskar_Mem** station_name;
    oskar_Mem* station_offset_ecef_metres[3];
    oskar_Mem** element_enu_metres[3];
    oskar_Mem** element_euler_angle_rad[2][3];
};

struct oskar_Timer
{
    int type;
    int paused;
    double elapsed;   /* accumulated seconds */
    double start;     /* host wall clock at last resume */
    void* ev_start;   /* hipEvent_t */
    void* ev_end;     /* hipEvent_t */
};

/* ---- internal helpers (host.c) ---------------------------------------- */
void* oskar_amd_current_stream(void);
void oskar_amd_check_launch(int* status);
int oskar_amd_hip_status(int hip_error);
void oskar_amd_set_last_error(int error);

/* ---- station beam (station.c) -------------------------------------------
 * oskar_evaluate_station_beam_aperture_array without the final horizon
 * blank of the top-level station (child stations are still blanked), for
 * a caller that blanks while it copies the beam out. */
void oskar_amd_station_beam_aa_unblanked(oskar_Station* s,
        oskar_StationWork* work, int num_points, const oskar_Mem* x,
        const oskar_Mem* y, const oskar_Mem* z, int time_index,
        double gast_rad, double frequency_hz, oskar_Mem* beam, int* status);

/* The beam of a flat aperture-array station whose elements are all of one
 * dipole type with a common orientation and common polarisation beams, in
 * one kernel per call (dftw.hip aa_fused_kernel): array factor x dipole
 * pattern, blanked below the horizon and, with normalise, divided as
 * oskar_mem_normalise divides by the beam in the direction (nx, ny, nz).
 * covers() says whether a station and beam type are of that kind; the
 * beam goes to beam from offset_out. */
int oskar_amd_station_beam_aa_fused_covers(const oskar_Station* s,
        const oskar_Mem* beam);
void oskar_amd_station_beam_aa_fused(oskar_Station* s,
        oskar_StationWork* work, int num_points, const oskar_Mem* x,
        const oskar_Mem* y, const oskar_Mem* z, int normalise, double nx,
        double ny, double nz, int time_index, double gast_rad,
        double frequency_hz, int offset_out, oskar_Mem* beam, int* status);

/* ---- piecewise chunk I/O (binary.c) --------------------------------------
 * A chunk written in pieces: begin writes the tag (data_size must equal the
 * sum of the pieces), each piece extends the CRC, end writes it.  A chunk
 * read in pieces: begin seeks to the payload and returns its size, pieces
 * are read in order, end checks the CRC. */
void oskar_amd_binary_write_begin(oskar_Binary* h, unsigned char data_type,
        unsigned char id_group, unsigned char id_tag, int user_index,
        size_t data_size, unsigned int* crc, int* status);
void oskar_amd_binary_write_piece(oskar_Binary* h, const void* data,
        size_t n, unsigned int* crc, int* status);
void oskar_amd_binary_write_end(oskar_Binary* h, unsigned int crc,
        int* status);
/* Payload bytes only: the caller extends the CRC itself. */
void oskar_amd_binary_write_raw(oskar_Binary* h, const void* data,
        size_t n, int* status);
void oskar_amd_binary_read_raw(oskar_Binary* h, void* data, size_t n,
        int* status);
size_t oskar_amd_binary_read_begin(oskar_Binary* h, int chunk_index,
        unsigned int* crc, int* status);
void oskar_amd_binary_read_piece(oskar_Binary* h, void* data, size_t n,
        unsigned int* crc, int* status);
void oskar_amd_binary_read_end(oskar_Binary* h, int chunk_index,
        unsigned int crc, int* status);

/* ---- kernel launchers (extern "C" in the .hip files) ------------------ */
/* All return a hipError_t value (0 on success) and launch on `stream`. */

/* xcorr_direct.hip */
typedef struct oskar_XcorrParams
{
    int use_casa;
    int num_sources;
    int num_stations;
    int offset_out;
    int gaussian;
    int bandwidth_smearing;
    int time_smearing;
    double uv_min_lambda, uv_max_lambda;
    double inv_wavelength, frac_bandwidth;
    double time_int_sec, gha0_rad, dec0_rad;
} oskar_XcorrParams;

int oskar_amd_xcorr_launch(int precision, int matrix,
        const oskar_XcorrParams* p, const void* jones,
        const void* I, const void* Q, const void* U, const void* V,
        const void* l, const void* m, const void* n,
        const void* a, const void* b, const void* c,
        const void* st_u, const void* st_v, const void* st_w,
        const void* st_x, const void* st_y, void* vis, void* stream);
/* As oskar_amd_xcorr_launch, with the split partials in the device's fixed
 * scratch array (no allocation; the split count is capped by its size). */
int oskar_amd_xcorr_launch_no_alloc(int precision, int matrix,
        const oskar_XcorrParams* p, const void* jones,
        const void* I, const void* Q, const void* U, const void* V,
        const void* l, const void* m, const void* n,
        const void* a, const void* b, const void* c,
        const void* st_u, const void* st_v, const void* st_w,
        const void* st_x, const void* st_y, void* vis, void* stream);

/* acorr.hip */
int oskar_amd_acorr_launch(int precision, int matrix, int num_sources,
        int num_stations, int offset_out, const void* jones, const void* I,
        const void* Q, const void* U, const void* V, void* vis,
        void* stream);

/* sky_clip.hip */
/* Up to SKY_CLIP_MAX_MODELS station models per mask launch: (ll, mm, nn)
 * of each, in FP64 (rounded to the kernel precision on the device). */
#define SKY_CLIP_MAX_MODELS 16
typedef struct oskar_HorizonModels
{
    int num;
    double lmn[SKY_CLIP_MAX_MODELS][3];
} oskar_HorizonModels;
int oskar_amd_horizon_mask_launch(int precision, int num,
        const void* l, const void* m, const void* n,
        const oskar_HorizonModels* models, int* mask, void* stream);
int oskar_amd_prefix_sum_launch(int num, const int* in, int* out,
        void* stream);
int oskar_amd_sky_copy_launch(int precision, int num_in, int capacity_in,
        int capacity_out, const int* mask, const int* indices,
        int num_columns, const void* table_in, void* table_out,
        const int* valid_in, int* valid_out, void* stream);
/* Fused clip: mask over all models, block counts, block offsets, then the
 * block scan writes indices (total at [num_in]) and scatters the rows. */
int oskar_amd_horizon_clip_launch(int precision, int num_in,
        const void* l, const void* m, const void* n,
        const oskar_HorizonModels* models, int* mask, int* indices,
        int capacity_in, int capacity_out, int num_columns,
        const void* table_in, void* table_out, const int* valid_in,
        int* valid_out, void* stream);

/* dftw.hip */
int oskar_amd_dftw_launch(int precision, int matrix, int is_3d,
        int num_in, double wavenumber, const void* weights,
        const void* x_in, const void* y_in, const void* z_in,
        int offset_coord_out, int num_out, const void* x_out,
        const void* y_out, const void* z_out, const int* data_idx,
        const void* data, int num_rows, int eval_x, int eval_y,
        int offset_out, void* output, double norm_factor, void* stream);

/* beam_util.hip */
int oskar_amd_set_point3_launch(int precision, void* x, void* y, void* z,
        size_t index, double vx, double vy, double vz, void* stream);
/* out[0 .. n) = data[0 .. n) with elements whose mask value is negative
 * set to zero (mask may be null) and, when normalise is set, scaled as
 * oskar_mem_normalise scales by element idx (blanked alike; n <= idx). */
int oskar_amd_beam_finish_launch(int type, unsigned int n,
        const void* data, unsigned int idx, int normalise, const void* mask,
        void* out, void* stream);
int oskar_amd_copy3_point_launch(int precision, unsigned int n,
        const void* x, const void* y, const void* z, void* xo, void* yo,
        void* zo, int extra, double px, double py, double pz, void* stream);
int oskar_amd_element_weights_dft_launch(int precision, int num_elements,
        const void* x, const void* y, const void* z, const void* cable,
        double wavenumber, double x1, double y1, double z1, void* weights,
        void* stream);
int oskar_amd_theta_phi_launch(int precision, int offset_in, int num,
        const void* x, const void* y, const void* z, int extra_point_at_pole,
        double delta_phi1, double delta_phi2, void* theta, void* phi1,
        void* phi2, void* stream);
int oskar_amd_dipole_launch(int precision, int matrix, int num_points,
        const void* theta, const void* phi_x, const void* phi_y, double kL,
        double cos_kL, int swap_xy, int offset_out, void* pattern,
        void* stream);
/* dftw.hip: the fused aperture-array beam (see
 * oskar_amd_station_beam_aa_fused); output already offset. */
int oskar_amd_aa_fused_launch(int precision, int matrix, int is_3d,
        int normalise, int num_in, double wavenumber, const void* weights,
        const void* x_in, const void* y_in, const void* z_in, int num_out,
        const void* x_out, const void* y_out, const void* z_out, double nx,
        double ny, double nz, double kL, double cos_kL, double dphi_x,
        double dphi_y, int swap_xy, double norm_factor, void* output,
        void* stream);
/* The dipole pattern straight from ENU directions (theta_phi then dipole,
 * theta and phi never stored): directions [offset_in, offset_in + num),
 * plus the pole at index num with extra_point_at_pole. */
int oskar_amd_dipole_enu_launch(int precision, int matrix, int offset_in,
        int num, const void* x, const void* y, const void* z,
        int extra_point_at_pole, double delta_phi1, double delta_phi2,
        double kL, double cos_kL, int swap_xy, int offset_out, void* pattern,
        void* stream);
int oskar_amd_blank_below_horizon_launch(int precision, int matrix,
        int offset_mask, int num, const void* mask, int offset_out,
        void* jones, void* stream);
int oskar_amd_mem_normalise_launch(int type, unsigned int offset,
        unsigned int n, void* data, unsigned int idx, void* stream);
int oskar_amd_mem_set_value_real_launch(int type, size_t offset, size_t n,
        double value, void* data, void* stream);
int oskar_amd_jones_K_launch(int precision, int num_sources,
        const void* l, const void* m, const void* n, int num_stations,
        const void* u, const void* v, const void* w, double wavenumber,
        const void* filter, double fmin, double fmax, int ignore_w,
        void* jones, void* stream);
int oskar_amd_jones_join_launch(int precision, int a_matrix, int b_matrix,
        size_t n, void* out, const void* a, const void* b, void* stream);
int oskar_amd_station_to_baseline_launch(int precision, int casa,
        int num_times, int num_stations, int offset_in, const void* u,
        const void* v, const void* w, int offset_out, void* uu, void* vv,
        void* ww, void* stream);
/* weights[i] *= amplitude (cos, sin)(phase) if gain (amplitude = gain +
 * draw x gain_error, phase = phase + draw x phase_error, the Gaussian draws
 * of the reference's generator for (seed, time_index, station_id) when
 * gain_error is set, else none), then *= apod[i] if apod. */
int oskar_amd_weights_scale_launch(int precision, int num, void* weights,
        const void* gain, const void* gain_error, const void* phase,
        const void* phase_error, const void* apod, unsigned int seed,
        int time_index, int station_id, void* stream);
int oskar_amd_enu_convert_launch(int precision, int kind, int num,
        const void* a, const void* b, const void* c, const double* params,
        void* x, void* y, void* z, void* stream);

/* sky_params.hip */
int oskar_amd_relative_directions_launch(int precision, int num,
        const void* lon, const void* lat, double lon0, double cos_lat0,
        double sin_lat0, void* l, void* m, void* n, void* stream);
/* Columns of one sky table that the Gaussian parameter kernel reads and
 * writes, as column numbers (column k starts at k * capacity), with the
 * cosines and sines of the six ellipse angles from the host's libm. */
typedef struct oskar_GaussianColumns
{
    size_t ra, dec, maj, min, pa, a, b, c;
    double cos_t[6], sin_t[6];
} oskar_GaussianColumns;
/* Clears *num_failed_dev (device memory), then adds one per failed fit. */
int oskar_amd_gaussian_params_launch(int precision, int num, void* table,
        size_t capacity, const oskar_GaussianColumns* cols, double ra0,
        double cos_dec0, double sin_dec0, int* num_failed_dev, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* OSKAR_AMD_PRIVATE_H_ */
