/*
 * oskar_Sky: the column-table sky model, with the horizon clip that feeds
 * the correlator the sources above the horizon of a time sample.
 *
 * Reference: oskar/sky/private_sky.h (one table of num_columns x capacity
 * values, column c at offset c x capacity; per-row counts of valid columns
 * for each fixed column type; column aliases, types and attributes),
 * oskar/sky/oskar_sky_enum.h (column and attribute numbering, kept
 * identical so that column types mean the same in both libraries),
 * oskar_sky_create.c, oskar_sky_free.c, oskar_sky_resize.c,
 * oskar_sky_column.c, oskar_sky_accessors.c, oskar_sky_copy_source_data.c,
 * oskar_sky_horizon_clip.c, oskar_update_horizon_mask.c and
 * oskar/math/src/oskar_prefix_sum.c.
 *
 * Memory: the table, the valid-column counts and every work array live at
 * the sky's location.  The HIP kernels of this library run on the current
 * HIP stream for OSKAR_GPU; the data-moving functions (horizon mask,
 * prefix sum, copy, clip) return OSKAR_ERR_BAD_LOCATION for OSKAR_CPU.
 */
#ifndef OSKAR_SKY_H_
#define OSKAR_SKY_H_

#include "oskar_mem.h"
#include "oskar_telescope.h"
#include "oskar_station.h"

#ifdef __cplusplus
extern "C" {
#endif

struct oskar_Sky;
#ifndef OSKAR_SKY_TYPEDEF_
#define OSKAR_SKY_TYPEDEF_
typedef struct oskar_Sky oskar_Sky;
#endif

/* Column types (oskar/sky/oskar_sky_enum.h). */
enum oskar_SkyColumn
{
    OSKAR_SKY_CUSTOM = 0,
    OSKAR_SKY_RA_RAD = 1,
    OSKAR_SKY_DEC_RAD = 2,
    OSKAR_SKY_I_JY = 3,
    OSKAR_SKY_Q_JY = 4,
    OSKAR_SKY_U_JY = 5,
    OSKAR_SKY_V_JY = 6,
    OSKAR_SKY_REF_HZ = 7,
    OSKAR_SKY_SPEC_IDX = 8,
    OSKAR_SKY_RM_RAD = 9,
    OSKAR_SKY_MAJOR_RAD = 10,
    OSKAR_SKY_MINOR_RAD = 11,
    OSKAR_SKY_PA_RAD = 12,
    OSKAR_SKY_LIN_SI,
    OSKAR_SKY_POLA_RAD,
    OSKAR_SKY_POLF,
    OSKAR_SKY_REF_WAVE_M,
    OSKAR_SKY_SPEC_CURV,
    OSKAR_SKY_LINE_WIDTH_HZ,
    OSKAR_SKY_INC_HZ,
    OSKAR_SKY_NUM_FIXED_COLUMN_TYPES,
    OSKAR_SKY_RA_DEG,
    OSKAR_SKY_DEC_DEG,
    OSKAR_SKY_SEMI_MAJOR,
    OSKAR_SKY_SEMI_MINOR,
    OSKAR_SKY_SCRATCH_START = 0x100,
    OSKAR_SKY_SCRATCH_EXT_A = OSKAR_SKY_SCRATCH_START,
    OSKAR_SKY_SCRATCH_EXT_B,
    OSKAR_SKY_SCRATCH_EXT_C,
    OSKAR_SKY_SCRATCH_L,
    OSKAR_SKY_SCRATCH_M,
    OSKAR_SKY_SCRATCH_N,
    OSKAR_SKY_SCRATCH_I_JY,
    OSKAR_SKY_SCRATCH_Q_JY,
    OSKAR_SKY_SCRATCH_U_JY,
    OSKAR_SKY_SCRATCH_V_JY,
    OSKAR_SKY_SCRATCH_END
};
typedef enum oskar_SkyColumn oskar_SkyColumn;

enum oskar_SkyAttribInt
{
    OSKAR_SKY_PRECISION,
    OSKAR_SKY_MEM_LOCATION,
    OSKAR_SKY_CAPACITY,
    OSKAR_SKY_NUM_SOURCES,
    OSKAR_SKY_NUM_COLUMNS,
    OSKAR_SKY_USE_EXTENDED,
    OSKAR_SKY_NUM_ATTRIBUTES_INT
};
typedef enum oskar_SkyAttribInt oskar_SkyAttribInt;

enum oskar_SkyAttribDouble
{
    OSKAR_SKY_REF_RA_RAD,
    OSKAR_SKY_REF_DEC_RAD,
    OSKAR_SKY_NUM_ATTRIBUTES_DOUBLE
};
typedef enum oskar_SkyAttribDouble oskar_SkyAttribDouble;

/* oskar_sky_create.c: precision OSKAR_SINGLE or OSKAR_DOUBLE; capacity =
 * num_sources + 1 rounded up to a multiple of 256; no columns yet. */
OSKAR_EXPORT
oskar_Sky* oskar_sky_create(int type, int location, int num_sources,
        int* status);
OSKAR_EXPORT void oskar_sky_free(oskar_Sky* sky, int* status);

/* oskar_sky_accessors.c */
OSKAR_EXPORT double oskar_sky_double(const oskar_Sky* sky,
        oskar_SkyAttribDouble attribute);
OSKAR_EXPORT int oskar_sky_int(const oskar_Sky* sky,
        oskar_SkyAttribInt attribute);
OSKAR_EXPORT void oskar_sky_set_double(oskar_Sky* sky,
        oskar_SkyAttribDouble attribute, double value);
OSKAR_EXPORT void oskar_sky_set_int(oskar_Sky* sky,
        oskar_SkyAttribInt attribute, int value);
OSKAR_EXPORT int oskar_sky_column_attribute(const oskar_Sky* sky,
        int column_index);
OSKAR_EXPORT oskar_SkyColumn oskar_sky_column_type(const oskar_Sky* sky,
        int column_index);
/* Count of valid columns of a fixed type for one source (one element read
 * from the sky's location), and its setter. */
OSKAR_EXPORT int oskar_sky_num_valid_columns_of_type(const oskar_Sky* sky,
        oskar_SkyColumn column_type, int index);
OSKAR_EXPORT void oskar_sky_set_num_valid_columns_of_type(oskar_Sky* sky,
        oskar_SkyColumn column_type, int index, int value);

/* oskar_sky_column.c: the column of (type, attribute), created (zeroed, at
 * the end of the table) if it does not exist; the _const form returns NULL
 * instead.  Creating a column re-points the aliases of all columns. */
OSKAR_EXPORT
oskar_Mem* oskar_sky_column(oskar_Sky* sky, oskar_SkyColumn column_type,
        int column_attribute, int* status);
OSKAR_EXPORT
const oskar_Mem* oskar_sky_column_const(const oskar_Sky* sky,
        oskar_SkyColumn column_type, int column_attribute);

/* oskar_sky_resize.c: new capacity num_sources + 1 rounded up to 256;
 * existing rows kept, new rows zero. */
OSKAR_EXPORT void oskar_sky_resize(oskar_Sky* sky, int num_sources,
        int* status);

/* oskar_update_horizon_mask.c: mask[i] |= (l_i ll + m_i mm + n_i nn > 0),
 * with ll = cos(lat) sin(ha0), mm = sin(lat) cos(dec0) - cos(lat) cos(ha0)
 * sin(dec0), nn = sin(lat) sin(dec0) + cos(lat) cos(ha0) cos(dec0) (the sine
 * of the elevation of direction (l, m, n) about the reference point),
 * computed in FP64 and rounded to the precision of l, m, n.  The sum is
 * evaluated as ((l ll + m mm) + n nn) with every product and sum rounded,
 * as the reference's CPU body, so the mask is identical to it. */
OSKAR_EXPORT
void oskar_update_horizon_mask(int num_sources, const oskar_Mem* l,
        const oskar_Mem* m, const oskar_Mem* n, const double ha0_rad,
        const double dec0_rad, const double lat_rad, oskar_Mem* mask,
        int* status);

/* oskar/math/src/oskar_prefix_sum.c: out[i] = sum_{j<i} in[j] (OSKAR_INT),
 * out[num_elements] = the total; out must hold num_elements + 1 values. */
OSKAR_EXPORT
void oskar_prefix_sum(size_t num_elements, const oskar_Mem* in,
        oskar_Mem* out, int* status);

/* oskar_sky_copy_source_data.c: copies every column value and the valid-
 * column counts of the sources with mask != 0 to rows indices[i] of out
 * (indices = the exclusive prefix sum of the mask, total at num_in); the
 * output gets the input's columns in the same order (OSKAR_ERR_TYPE_MISMATCH
 * if it already holds others first), its precision, location, extended-
 * source flag and reference point, and NUM_SOURCES = the total.  Unlike the
 * reference, which also copies the input's capacity and column count into
 * the output's attributes, the output keeps its own (they describe its own
 * table). */
OSKAR_EXPORT
void oskar_sky_copy_source_data(const oskar_Sky* in,
        const oskar_Mem* horizon_mask, const oskar_Mem* indices,
        oskar_Sky* out, int* status);

/* oskar_sky_horizon_clip.c: out = the sources of in above the horizon of
 * any station model of the telescope (oskar_update_horizon_mask with
 * ha0 = (gast + longitude) - RA0, dec0 = the sky's reference Dec, per
 * model), in their input order; the work's horizon mask and source indices
 * hold the mask and its prefix sum afterwards.  out is resized if its
 * capacity is below the input's source count.  The clip runs as three
 * kernels (mask and per-block counts; block offsets; block scan, index
 * write and row scatter) and one 4-byte read of the total. */
OSKAR_EXPORT
void oskar_sky_horizon_clip(oskar_Sky* out, const oskar_Sky* in,
        const oskar_Telescope* telescope, double gast_rad,
        oskar_StationWork* work, int* status);

/* oskar/sky/src/oskar_sky_evaluate_relative_directions.c:20-37: the
 * scratch columns L, M, N (created if missing) get the direction cosines
 * of each source about (ra0, dec0):
 *   l = cos(dec) sin(ra - ra0),
 *   m = cos(dec0) sin(dec) - sin(dec0) cos(dec) cos(ra - ra0),
 *   n = sin(dec0) sin(dec) + cos(dec0) cos(dec) cos(ra - ra0),
 * and the sky's reference point is set to (ra0, dec0).  A sky in host
 * memory is staged through device memory; the arithmetic runs on the GPU
 * either way.  OSKAR_ERR_INVALID_ARGUMENT if the RA or Dec column is
 * missing. */
OSKAR_EXPORT
void oskar_sky_evaluate_relative_directions(oskar_Sky* sky, double ra0_rad,
        double dec0_rad, int* status);

/* oskar/sky/src/oskar_sky_evaluate_gaussian_source_parameters.c:32-262:
 * for each source with a non-zero major or minor FWHM, the scratch columns
 * EXT_A, EXT_B, EXT_C (created if missing) get the Gaussian parameters
 * a, b, c that the correlator's exp(-(a u^2 + 2 b u v + c v^2)) uses, from
 * an ellipse fitted to six points of the source's ellipse projected onto
 * the tangent plane at (ra0, dec0).  Sources whose fit fails keep their
 * a, b, c and add one to *num_failed (which is not cleared first).
 * zero_failed_sources changes nothing, as in the reference, whose
 * flux-clearing call returns at once because *status still holds the fit's
 * error code at that point (:152-160); failed sources keep their fluxes.
 * Returns without doing anything if the major, minor or position-angle
 * column is missing.
 * The reference accepts only a sky in host memory and fits serially on the
 * CPU; here each source is fitted by one GPU thread.  A sky in device
 * memory is processed in place; one in host memory is staged through device
 * memory.  The call waits for the kernel (it returns the failure count).
 * OSKAR_ERR_INVALID_ARGUMENT if the RA or Dec column is missing. */
OSKAR_EXPORT
void oskar_sky_evaluate_gaussian_source_parameters(oskar_Sky* sky,
        int zero_failed_sources, double ra0, double dec0, int* num_failed,
        int* status);

#ifdef __cplusplus
}
#endif

#endif /* OSKAR_SKY_H_ */
