/*
 * Visibility blocks and headers: the correlator's output container and the
 * .vis file writer/reader.
 *
 * Reference: oskar/vis/private_vis_block.h (dim_start_size[6] = start
 * time, start channel, times, channels, baselines, stations; the
 * amplitudes in [time][channel][baseline] order with the polarisation
 * fastest), oskar/vis/oskar_vis_block.h and oskar_vis_header.h (tag
 * numbers, kept identical: they are the file format),
 * src/oskar_vis_block_create.c, _resize.c, _clear.c, _write.c, _read.c,
 * _station_to_baseline_coords.c, src/oskar_vis_header_create.c, _write.c,
 * _read.c, oskar/mem/src/oskar_binary_write_mem.c, oskar_binary_read_mem.c,
 * oskar/convert/define_convert_station_uvw_to_baseline_uvw.h.
 *
 * A block's arrays live at the block's location (OSKAR_GPU: device memory,
 * so the correlator writes straight into them with offset_out =
 * num_baselines x (num_channels x t + c) for time t and channel c of the
 * block); writing copies them to the host.  Headers live on the host.
 */
#ifndef OSKAR_VIS_H_
#define OSKAR_VIS_H_

#include "oskar_mem.h"
#include "oskar_binary.h"

#ifdef __cplusplus
extern "C" {
#endif

struct oskar_VisBlock;
#ifndef OSKAR_VIS_BLOCK_TYPEDEF_
#define OSKAR_VIS_BLOCK_TYPEDEF_
typedef struct oskar_VisBlock oskar_VisBlock;
#endif
struct oskar_VisHeader;
#ifndef OSKAR_VIS_HEADER_TYPEDEF_
#define OSKAR_VIS_HEADER_TYPEDEF_
typedef struct oskar_VisHeader oskar_VisHeader;
#endif

enum OSKAR_VIS_BLOCK_TAGS
{
    OSKAR_VIS_BLOCK_TAG_DIM_START_AND_SIZE    = 1,
    OSKAR_VIS_BLOCK_TAG_AUTO_CORRELATIONS     = 2,
    OSKAR_VIS_BLOCK_TAG_CROSS_CORRELATIONS    = 3,
    OSKAR_VIS_BLOCK_TAG_BASELINE_UU           = 4,
    OSKAR_VIS_BLOCK_TAG_BASELINE_VV           = 5,
    OSKAR_VIS_BLOCK_TAG_BASELINE_WW           = 6,
    OSKAR_VIS_BLOCK_TAG_STATION_U             = 7,
    OSKAR_VIS_BLOCK_TAG_STATION_V             = 8,
    OSKAR_VIS_BLOCK_TAG_STATION_W             = 9
};

enum OSKAR_VIS_HEADER_TAGS
{
    OSKAR_VIS_HEADER_TAG_TELESCOPE_PATH           = 1,
    OSKAR_VIS_HEADER_TAG_NUM_TAGS_PER_BLOCK       = 2,
    OSKAR_VIS_HEADER_TAG_WRITE_AUTO_CORRELATIONS  = 3,
    OSKAR_VIS_HEADER_TAG_WRITE_CROSS_CORRELATIONS = 4,
    OSKAR_VIS_HEADER_TAG_AMP_TYPE                 = 5,
    OSKAR_VIS_HEADER_TAG_COORD_PRECISION          = 6,
    OSKAR_VIS_HEADER_TAG_MAX_TIMES_PER_BLOCK      = 7,
    OSKAR_VIS_HEADER_TAG_NUM_TIMES_TOTAL          = 8,
    OSKAR_VIS_HEADER_TAG_MAX_CHANNELS_PER_BLOCK   = 9,
    OSKAR_VIS_HEADER_TAG_NUM_CHANNELS_TOTAL       = 10,
    OSKAR_VIS_HEADER_TAG_NUM_STATIONS             = 11,
    OSKAR_VIS_HEADER_TAG_POL_TYPE                 = 12,
    OSKAR_VIS_HEADER_TAG_CASA_PHASE_CONVENTION    = 13,
    OSKAR_VIS_HEADER_TAG_PHASE_CENTRE_COORD_TYPE  = 21,
    OSKAR_VIS_HEADER_TAG_PHASE_CENTRE_DEG         = 22,
    OSKAR_VIS_HEADER_TAG_FREQ_START_HZ            = 23,
    OSKAR_VIS_HEADER_TAG_FREQ_INC_HZ              = 24,
    OSKAR_VIS_HEADER_TAG_CHANNEL_BANDWIDTH_HZ     = 25,
    OSKAR_VIS_HEADER_TAG_TIME_START_MJD_UTC       = 26,
    OSKAR_VIS_HEADER_TAG_TIME_INC_SEC             = 27,
    OSKAR_VIS_HEADER_TAG_TIME_AVERAGE_SEC         = 28,
    OSKAR_VIS_HEADER_TAG_TELESCOPE_REF_LON_DEG    = 29,
    OSKAR_VIS_HEADER_TAG_TELESCOPE_REF_LAT_DEG    = 30,
    OSKAR_VIS_HEADER_TAG_TELESCOPE_REF_ALT_M      = 31,
    OSKAR_VIS_HEADER_TAG_STATION_X_OFFSET_ECEF    = 32,
    OSKAR_VIS_HEADER_TAG_STATION_Y_OFFSET_ECEF    = 33,
    OSKAR_VIS_HEADER_TAG_STATION_Z_OFFSET_ECEF    = 34,
    OSKAR_VIS_HEADER_TAG_ELEMENT_X_ENU            = 35,
    OSKAR_VIS_HEADER_TAG_ELEMENT_Y_ENU            = 36,
    OSKAR_VIS_HEADER_TAG_ELEMENT_Z_ENU            = 37,
    OSKAR_VIS_HEADER_TAG_STATION_NAME             = 41,
    OSKAR_VIS_HEADER_TAG_STATION_DIAMETER         = 42,
    OSKAR_VIS_HEADER_TAG_ELEMENT_FEED_ANGLE_X_A   = 43,
    OSKAR_VIS_HEADER_TAG_ELEMENT_FEED_ANGLE_Y_A   = 44,
    OSKAR_VIS_HEADER_TAG_ELEMENT_FEED_ANGLE_X_B   = 45,
    OSKAR_VIS_HEADER_TAG_ELEMENT_FEED_ANGLE_Y_B   = 46,
    OSKAR_VIS_HEADER_TAG_ELEMENT_FEED_ANGLE_X_C   = 47,
    OSKAR_VIS_HEADER_TAG_ELEMENT_FEED_ANGLE_Y_C   = 48
};

/* oskar/vis/oskar_vis_header.h polarisation types. */
enum OSKAR_VIS_POL_TYPE
{
    OSKAR_VIS_POL_TYPE_STOKES_I_Q_U_V     =  0,
    OSKAR_VIS_POL_TYPE_STOKES_I           =  1,
    OSKAR_VIS_POL_TYPE_STOKES_Q           =  2,
    OSKAR_VIS_POL_TYPE_STOKES_U           =  3,
    OSKAR_VIS_POL_TYPE_STOKES_V           =  4,
    OSKAR_VIS_POL_TYPE_LINEAR_XX_XY_YX_YY = 10,
    OSKAR_VIS_POL_TYPE_LINEAR_XX          = 11,
    OSKAR_VIS_POL_TYPE_LINEAR_XY          = 12,
    OSKAR_VIS_POL_TYPE_LINEAR_YX          = 13,
    OSKAR_VIS_POL_TYPE_LINEAR_YY          = 14
};

/* ---- blocks -------------------------------------------------------------- */
/* amp_type: single or double complex (matrix for four polarisations). */
OSKAR_EXPORT
oskar_VisBlock* oskar_vis_block_create(int location, int amp_type,
        int num_times, int num_channels, int num_stations,
        int create_crosscorr, int create_autocorr, int* status);
OSKAR_EXPORT void oskar_vis_block_free(oskar_VisBlock* vis, int* status);
/* A block sized for the header's largest block. */
OSKAR_EXPORT
oskar_VisBlock* oskar_vis_block_create_from_header(int location,
        const oskar_VisHeader* hdr, int* status);
/* Sets the dimensions; the amplitude arrays are resized (contents kept up
 * to the smaller size), uvw arrays only if they already hold data. */
OSKAR_EXPORT
void oskar_vis_block_resize(oskar_VisBlock* vis, int num_times,
        int num_channels, int num_stations, int* status);
/* Zeroes every array (on the current HIP stream for device blocks). */
OSKAR_EXPORT void oskar_vis_block_clear(oskar_VisBlock* vis, int* status);

OSKAR_EXPORT int oskar_vis_block_num_times(const oskar_VisBlock* vis);
OSKAR_EXPORT int oskar_vis_block_num_channels(const oskar_VisBlock* vis);
OSKAR_EXPORT int oskar_vis_block_num_baselines(const oskar_VisBlock* vis);
OSKAR_EXPORT int oskar_vis_block_num_stations(const oskar_VisBlock* vis);
OSKAR_EXPORT int oskar_vis_block_num_pols(const oskar_VisBlock* vis);
OSKAR_EXPORT int oskar_vis_block_start_time_index(const oskar_VisBlock* vis);
OSKAR_EXPORT
int oskar_vis_block_start_channel_index(const oskar_VisBlock* vis);
OSKAR_EXPORT
void oskar_vis_block_set_start_time_index(oskar_VisBlock* vis, int index);
OSKAR_EXPORT
void oskar_vis_block_set_start_channel_index(oskar_VisBlock* vis, int index);
OSKAR_EXPORT
int oskar_vis_block_has_auto_correlations(const oskar_VisBlock* vis);
OSKAR_EXPORT
int oskar_vis_block_has_cross_correlations(const oskar_VisBlock* vis);
OSKAR_EXPORT
oskar_Mem* oskar_vis_block_cross_correlations(oskar_VisBlock* vis);
OSKAR_EXPORT
oskar_Mem* oskar_vis_block_auto_correlations(oskar_VisBlock* vis);
/* dim 0, 1, 2 = u, v, w. */
OSKAR_EXPORT
oskar_Mem* oskar_vis_block_baseline_uvw_metres(oskar_VisBlock* vis,
        int dim);
OSKAR_EXPORT
oskar_Mem* oskar_vis_block_station_uvw_metres(oskar_VisBlock* vis, int dim);
/* Baseline (u, v, w) of every time from the station (u, v, w):
 * uu[b] = sign (u[q] - u[p]) for station pairs p < q in baseline order,
 * sign -1 for the CASA convention (a HIP kernel for device blocks). */
OSKAR_EXPORT
void oskar_vis_block_station_to_baseline_coords(oskar_VisBlock* vis,
        const oskar_VisHeader* hdr, int* status);
/* Chunks: dims (6 ints); auto-correlations; cross-correlations; station
 * (u, v, w) if they hold data, else baseline (u, v, w) if they do. */
OSKAR_EXPORT
void oskar_vis_block_write(const oskar_VisBlock* vis, oskar_Binary* h,
        int block_index, int* status);
OSKAR_EXPORT
void oskar_vis_block_read(oskar_VisBlock* vis, const oskar_VisHeader* hdr,
        oskar_Binary* h, int block_index, int* status);

/* Station (u, v, w) of num stations from offset_in to baseline (u, v, w)
 * from offset_out (oskar/convert/oskar_convert_station_uvw_to_baseline_uvw.h).
 * All six arrays share one location and real type; lengths are checked
 * (OSKAR_ERR_DIMENSION_MISMATCH) before anything is written. */
OSKAR_EXPORT
void oskar_convert_station_uvw_to_baseline_uvw(int use_casa_phase_convention,
        int num_stations, int offset_in, const oskar_Mem* u,
        const oskar_Mem* v, const oskar_Mem* w, int offset_out,
        oskar_Mem* uu, oskar_Mem* vv, oskar_Mem* ww, int* status);

/* ---- headers -------------------------------------------------------------- */
/* max_*_per_block <= 0 means the totals; tags per block = 1 + 4 (cross) +
 * 1 (auto); pol type linear XX..YY for matrix amplitudes, else Stokes I;
 * station arrays sized for num_stations. */
OSKAR_EXPORT
oskar_VisHeader* oskar_vis_header_create(int amp_type, int coord_precision,
        int max_times_per_block, int num_times_total,
        int max_channels_per_block, int num_channels_total, int num_stations,
        int write_autocorr, int write_crosscorr, int* status);
OSKAR_EXPORT void oskar_vis_header_free(oskar_VisHeader* hdr, int* status);

OSKAR_EXPORT int oskar_vis_header_amp_type(const oskar_VisHeader* hdr);
OSKAR_EXPORT int oskar_vis_header_coord_precision(const oskar_VisHeader* hdr);
OSKAR_EXPORT int oskar_vis_header_num_stations(const oskar_VisHeader* hdr);
OSKAR_EXPORT int oskar_vis_header_num_tags_per_block(const oskar_VisHeader* h);
OSKAR_EXPORT int oskar_vis_header_num_tags_header(const oskar_VisHeader* h);
OSKAR_EXPORT
int oskar_vis_header_max_times_per_block(const oskar_VisHeader* hdr);
OSKAR_EXPORT
int oskar_vis_header_max_channels_per_block(const oskar_VisHeader* hdr);
OSKAR_EXPORT int oskar_vis_header_num_times_total(const oskar_VisHeader* hdr);
OSKAR_EXPORT
int oskar_vis_header_num_channels_total(const oskar_VisHeader* hdr);
OSKAR_EXPORT
int oskar_vis_header_write_auto_correlations(const oskar_VisHeader* hdr);
OSKAR_EXPORT
int oskar_vis_header_write_cross_correlations(const oskar_VisHeader* hdr);
OSKAR_EXPORT int oskar_vis_header_pol_type(const oskar_VisHeader* hdr);
OSKAR_EXPORT
int oskar_vis_header_casa_phase_convention(const oskar_VisHeader* hdr);
OSKAR_EXPORT double oskar_vis_header_freq_start_hz(const oskar_VisHeader* h);
OSKAR_EXPORT double oskar_vis_header_freq_inc_hz(const oskar_VisHeader* h);
OSKAR_EXPORT
double oskar_vis_header_time_start_mjd_utc(const oskar_VisHeader* hdr);
OSKAR_EXPORT double oskar_vis_header_time_inc_sec(const oskar_VisHeader* h);
OSKAR_EXPORT
double oskar_vis_header_phase_centre_ra_deg(const oskar_VisHeader* hdr);
OSKAR_EXPORT
double oskar_vis_header_phase_centre_dec_deg(const oskar_VisHeader* hdr);
/* Host arrays of num_stations values in coord_precision; names are char. */
OSKAR_EXPORT
oskar_Mem* oskar_vis_header_station_offset_ecef_metres(oskar_VisHeader* h,
        int dim);
OSKAR_EXPORT oskar_Mem* oskar_vis_header_telescope_path(oskar_VisHeader* h);
OSKAR_EXPORT oskar_Mem* oskar_vis_header_settings(oskar_VisHeader* h);
OSKAR_EXPORT
oskar_Mem* oskar_vis_header_station_diameter_m(oskar_VisHeader* h);
OSKAR_EXPORT
oskar_Mem* oskar_vis_header_station_name(oskar_VisHeader* h, int station);
OSKAR_EXPORT
oskar_Mem* oskar_vis_header_element_enu_metres(oskar_VisHeader* h, int dim,
        int station);
/* Feed angles: dim 0 = X, 1 = Y polarisation; angle 0, 1, 2 = Euler
 * angles a, b, c in radians (OSKAR_DOUBLE). */
OSKAR_EXPORT
oskar_Mem* oskar_vis_header_element_feed_angle_rad(oskar_VisHeader* h,
        int dim, int angle, int station);
OSKAR_EXPORT double oskar_vis_header_channel_bandwidth_hz(
        const oskar_VisHeader* hdr);
OSKAR_EXPORT double oskar_vis_header_time_average_sec(
        const oskar_VisHeader* hdr);
OSKAR_EXPORT int oskar_vis_header_phase_centre_type(const oskar_VisHeader* h);
OSKAR_EXPORT double oskar_vis_header_telescope_lon_deg(
        const oskar_VisHeader* hdr);
OSKAR_EXPORT double oskar_vis_header_telescope_lat_deg(
        const oskar_VisHeader* hdr);
OSKAR_EXPORT double oskar_vis_header_telescope_alt_metres(
        const oskar_VisHeader* hdr);

OSKAR_EXPORT
void oskar_vis_header_set_pol_type(oskar_VisHeader* hdr, int value,
        int* status);
OSKAR_EXPORT
void oskar_vis_header_set_casa_phase_convention(oskar_VisHeader* hdr,
        int value);
OSKAR_EXPORT
void oskar_vis_header_set_freq_start_hz(oskar_VisHeader* hdr, double value);
OSKAR_EXPORT
void oskar_vis_header_set_freq_inc_hz(oskar_VisHeader* hdr, double value);
OSKAR_EXPORT
void oskar_vis_header_set_channel_bandwidth_hz(oskar_VisHeader* hdr,
        double value);
OSKAR_EXPORT
void oskar_vis_header_set_time_start_mjd_utc(oskar_VisHeader* hdr,
        double value);
OSKAR_EXPORT
void oskar_vis_header_set_time_inc_sec(oskar_VisHeader* hdr, double value);
OSKAR_EXPORT
void oskar_vis_header_set_time_average_sec(oskar_VisHeader* hdr,
        double value);
OSKAR_EXPORT
void oskar_vis_header_set_phase_centre(oskar_VisHeader* hdr, int coord_type,
        double lon_deg, double lat_deg);
OSKAR_EXPORT
void oskar_vis_header_set_telescope_centre(oskar_VisHeader* hdr,
        double lon_deg, double lat_deg, double alt_metres);

/* Creates the file, writes the metadata and header chunks, returns the
 * open handle for the blocks (oskar_vis_block_write with index 0, 1, ...). */
OSKAR_EXPORT
oskar_Binary* oskar_vis_header_write(const oskar_VisHeader* hdr,
        const char* filename, int* status);
/* Reads a header from a file opened with oskar_binary_create(.., 'r'). */
OSKAR_EXPORT
oskar_VisHeader* oskar_vis_header_read(oskar_Binary* h, int* status);

/* ---- oskar_Mem chunks (oskar/mem/oskar_binary_write_mem.h,
 * oskar_binary_read_mem.h) --------------------------------------------- */
/* num_to_write 0 = all elements; device memory is staged through the host. */
OSKAR_EXPORT
void oskar_binary_write_mem(oskar_Binary* h, const oskar_Mem* mem,
        unsigned char id_group, unsigned char id_tag, int user_index,
        size_t num_to_write, int* status);
/* Resizes mem to the chunk's length in mem's type, then reads into it. */
OSKAR_EXPORT
void oskar_binary_read_mem(oskar_Binary* h, oskar_Mem* mem,
        unsigned char id_group, unsigned char id_tag, int user_index,
        int* status);

#ifdef __cplusplus
}
#endif

#endif /* OSKAR_VIS_H_ */
