/*
 * Per-source sky parameters on the GPU: relative direction cosines and
 * Gaussian source parameters.
 *
 * Reference: oskar/sky/src/oskar_sky_evaluate_relative_directions.c:20-37,
 * oskar/sky/src/oskar_sky_evaluate_gaussian_source_parameters.c:32-262.
 * Both are called on the host's sky chunks before the simulation starts
 * (oskar/interferometer/src/oskar_interferometer_check_init.c:66-73), so a
 * sky in host memory is accepted here too: its table is staged through
 * device memory and the results are copied back.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <hip/hip_runtime_api.h>

#include "private.h"
#include "oskar_sky.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* Column number of (type, attribute 0), or -1. */
static int column_index(const oskar_Sky* sky, oskar_SkyColumn type)
{
    int i = 0;
    for (i = 0; i < sky->attr_int[OSKAR_SKY_NUM_COLUMNS]; ++i)
        if (sky->column_type[i] == type && sky->column_attr[i] == 0) return i;
    return -1;
}

/* Device pointer to the sky table: the table itself for a device sky, or a
 * device copy of a host table (*staged = 1; free with unstage_table). */
static void* stage_table(oskar_Sky* sky, int* staged, int* status)
{
    *staged = 0;
    if (*status) return 0;
    if (sky->attr_int[OSKAR_SKY_MEM_LOCATION] == OSKAR_GPU)
        return oskar_mem_void(sky->table);
    if (sky->attr_int[OSKAR_SKY_MEM_LOCATION] != OSKAR_CPU)
    {
        *status = OSKAR_ERR_BAD_LOCATION;
        return 0;
    }
    const size_t bytes = oskar_mem_length(sky->table) *
            oskar_mem_element_size(oskar_mem_type(sky->table));
    void* d = 0;
    int e = (int) hipMalloc(&d, bytes > 0 ? bytes : 1);
    if (!e && bytes > 0)
        e = (int) hipMemcpy(d, oskar_mem_void(sky->table), bytes,
                hipMemcpyHostToDevice);
    if (e)
    {
        if (d) (void) hipFree(d);
        *status = e;
        return 0;
    }
    *staged = 1;
    return d;
}

/* Copies a staged table back (after the stream has finished) and frees it. */
static void unstage_table(oskar_Sky* sky, void* d, int staged, int* status)
{
    if (!staged) return;
    const size_t bytes = oskar_mem_length(sky->table) *
            oskar_mem_element_size(oskar_mem_type(sky->table));
    int e = (int) hipStreamSynchronize(
            (hipStream_t) oskar_amd_current_stream());
    if (!e && !*status && bytes > 0)
        e = (int) hipMemcpy(oskar_mem_void(sky->table), d, bytes,
                hipMemcpyDeviceToHost);
    (void) hipFree(d);
    if (e && !*status) *status = e;
}

void oskar_sky_evaluate_relative_directions(oskar_Sky* sky, double ra0_rad,
        double dec0_rad, int* status)
{
    int staged = 0;
    if (*status) return;
    (void) oskar_sky_column(sky, OSKAR_SKY_SCRATCH_L, 0, status);
    (void) oskar_sky_column(sky, OSKAR_SKY_SCRATCH_M, 0, status);
    (void) oskar_sky_column(sky, OSKAR_SKY_SCRATCH_N, 0, status);
    if (*status) return;
    const int c_ra = column_index(sky, OSKAR_SKY_RA_RAD);
    const int c_dec = column_index(sky, OSKAR_SKY_DEC_RAD);
    if (c_ra < 0 || c_dec < 0)
    {
        *status = OSKAR_ERR_INVALID_ARGUMENT;
        return;
    }
    const int c_l = column_index(sky, OSKAR_SKY_SCRATCH_L);
    const int c_m = column_index(sky, OSKAR_SKY_SCRATCH_M);
    const int c_n = column_index(sky, OSKAR_SKY_SCRATCH_N);
    const int type = sky->attr_int[OSKAR_SKY_PRECISION];
    const int num = sky->attr_int[OSKAR_SKY_NUM_SOURCES];
    const size_t cap = (size_t) sky->attr_int[OSKAR_SKY_CAPACITY];
    const size_t esz = type == OSKAR_DOUBLE ? 8 : 4;
    if (num > 0)
    {
        char* t = (char*) stage_table(sky, &staged, status);
        if (*status) return;
        const int e = oskar_amd_relative_directions_launch(type, num,
                t + (size_t) c_ra * cap * esz, t + (size_t) c_dec * cap * esz,
                ra0_rad, cos(dec0_rad), sin(dec0_rad),
                t + (size_t) c_l * cap * esz, t + (size_t) c_m * cap * esz,
                t + (size_t) c_n * cap * esz, oskar_amd_current_stream());
        if (e && !*status) *status = e;
        unstage_table(sky, t, staged, status);
    }
    sky->attr_double[OSKAR_SKY_REF_RA_RAD] = ra0_rad;
    sky->attr_double[OSKAR_SKY_REF_DEC_RAD] = dec0_rad;
}

/* Per-device failure counter (4 bytes, kept until the process ends). */
static int* failure_counter(int* status)
{
    static int* ptr[64];
    int dev = 0;
    int e = (int) hipGetDevice(&dev);
    if (!e && (dev < 0 || dev >= 64)) e = (int) hipErrorInvalidDevice;
    if (!e && !ptr[dev]) e = (int) hipMalloc((void**) &ptr[dev], sizeof(int));
    if (e)
    {
        *status = e;
        return 0;
    }
    return ptr[dev];
}

void oskar_sky_evaluate_gaussian_source_parameters(oskar_Sky* sky,
        int zero_failed_sources, double ra0, double dec0, int* num_failed,
        int* status)
{
    int i = 0, staged = 0, failed = 0;
    oskar_GaussianColumns cols;
    /* The reference's zero_failed_sources has no effect: its flux-clearing
     * call runs while *status still holds the fit's error code and returns
     * at once (oskar_sky_evaluate_gaussian_source_parameters.c:152-160,
     * oskar_sky_clear_source_flux.c:17).  Fluxes are left alone here too,
     * so that simulations agree with the reference's. */
    (void) zero_failed_sources;
    if (*status) return;
    if (!oskar_sky_column_const(sky, OSKAR_SKY_MAJOR_RAD, 0) ||
            !oskar_sky_column_const(sky, OSKAR_SKY_MINOR_RAD, 0) ||
            !oskar_sky_column_const(sky, OSKAR_SKY_PA_RAD, 0))
        return;
    const int type = sky->attr_int[OSKAR_SKY_PRECISION];
    const int num = sky->attr_int[OSKAR_SKY_NUM_SOURCES];
    const size_t cap = (size_t) sky->attr_int[OSKAR_SKY_CAPACITY];
    (void) oskar_sky_column(sky, OSKAR_SKY_SCRATCH_EXT_A, 0, status);
    (void) oskar_sky_column(sky, OSKAR_SKY_SCRATCH_EXT_B, 0, status);
    (void) oskar_sky_column(sky, OSKAR_SKY_SCRATCH_EXT_C, 0, status);
    if (*status) return;
    memset(&cols, 0, sizeof(cols));
    {
        const int c_ra = column_index(sky, OSKAR_SKY_RA_RAD);
        const int c_dec = column_index(sky, OSKAR_SKY_DEC_RAD);
        if (c_ra < 0 || c_dec < 0)
        {
            *status = OSKAR_ERR_INVALID_ARGUMENT;
            return;
        }
        cols.ra = (size_t) c_ra;
        cols.dec = (size_t) c_dec;
    }
    cols.maj = (size_t) column_index(sky, OSKAR_SKY_MAJOR_RAD);
    cols.min = (size_t) column_index(sky, OSKAR_SKY_MINOR_RAD);
    cols.pa = (size_t) column_index(sky, OSKAR_SKY_PA_RAD);
    cols.a = (size_t) column_index(sky, OSKAR_SKY_SCRATCH_EXT_A);
    cols.b = (size_t) column_index(sky, OSKAR_SKY_SCRATCH_EXT_B);
    cols.c = (size_t) column_index(sky, OSKAR_SKY_SCRATCH_EXT_C);

    /* The six ellipse angles, rounded to the sky's precision before the
     * host's cos and sin, as the reference's loop variable is. */
    for (i = 0; i < 6; ++i)
    {
        const double t = i * 60.0 * M_PI / 180.0;
        const double tt = (type == OSKAR_SINGLE) ? (double) (float) t : t;
        cols.cos_t[i] = cos(tt);
        cols.sin_t[i] = sin(tt);
    }
    if (num <= 0) return;

    int* counter = failure_counter(status);
    void* t = stage_table(sky, &staged, status);
    if (*status) return;
    {
        const int e = oskar_amd_gaussian_params_launch(type, num, t, cap,
                &cols, ra0, cos(dec0), sin(dec0), counter,
                oskar_amd_current_stream());
        if (e && !*status) *status = e;
    }
    if (!*status)
    {
        const int e = (int) hipMemcpyAsync(&failed, counter, sizeof(int),
                hipMemcpyDeviceToHost,
                (hipStream_t) oskar_amd_current_stream());
        const int e2 = (int) hipStreamSynchronize(
                (hipStream_t) oskar_amd_current_stream());
        if (e || e2) *status = e ? e : e2;
    }
    unstage_table(sky, t, staged, status);
    if (!*status) *num_failed += failed;
}
