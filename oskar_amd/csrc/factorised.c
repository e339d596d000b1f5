/*
 * oskar_cross_correlate_factorised(): checks and dispatch to the MFMA
 * kernels in xcorr_factorised.hip.  Derived scalars and checks follow
 * oskar/correlate/src/oskar_cross_correlate.c:42-110.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "private.h"
#include "oskar_cross_correlate_factorised.h"

int oskar_amd_fact_launch(int precision, int mode, int use_casa,
        int num_sources, int num_stations, const void* beam,
        const void* I, const void* Q, const void* U, const void* V,
        const void* l, const void* m, const void* n, const void* st_u,
        const void* st_v, const void* st_w, double frequency_hz,
        double uv_min_lambda, double uv_max_lambda, int offset_out,
        void* vis, int src_per_part, void* stream);

int oskar_amd_fact_stations_launch(int precision, int mode, int use_casa,
        int num_sources, int num_stations, const void* E, size_t ld_e,
        const void* I, const void* Q, const void* U, const void* V,
        const void* l, const void* m, const void* n, const void* st_u,
        const void* st_v, const void* st_w, double frequency_hz,
        double uv_min_lambda, double uv_max_lambda, int offset_out,
        void* vis, int src_per_part, void* stream);
void oskar_amd_fact_stations_set_profiling(int on);
int oskar_amd_fact_stations_last_times(double* ms);

void oskar_amd_fact_set_profiling(int on);
int oskar_amd_fact_set_tile_shape(int shape);
int oskar_amd_fact_last_times(double* ms);

void* oskar_amd_fact_batch_create(int precision, int mode, int num_stations,
        int max_units, int* err);
void oskar_amd_fact_batch_free(void* b);
void oskar_amd_fact_batch_set_unit(void* b, int unit, int num_sources,
        const void* beam, const void* I, const void* Q, const void* U,
        const void* V, const void* l, const void* m, const void* n,
        const void* st_u, const void* st_v, const void* st_w,
        double frequency_hz, int offset_out, void* vis);
int oskar_amd_fact_batch_unit_is_set(const void* b, int unit);
int oskar_amd_fact_batch_set_config(void* b, int shape, int src_per_part,
        int workgroups, int use_graph);
void oskar_amd_fact_batch_get_config(const void* b, int* out);
int oskar_amd_fact_batch_run(void* b, int num_units, int use_casa,
        const double* uv_min, const double* uv_max, void* stream);
int oskar_amd_fact_batch_last_ms(void* b, double* ms);
int oskar_amd_fact_batch_autotune(void* b, int num_units, int reps,
        void* stream, double* best_ms);

static int part_size = 8192;
static int part_size_jones = 2048;

void oskar_cross_correlate_factorised_set_profiling(int on)
{
    oskar_amd_fact_set_profiling(on);
    oskar_amd_fact_stations_set_profiling(on);
}

int oskar_cross_correlate_factorised_jones_last_times(double ms[3])
{
    return oskar_amd_fact_stations_last_times(ms);
}

int oskar_cross_correlate_factorised_last_times(double ms[3])
{
    return oskar_amd_fact_last_times(ms);
}

int oskar_cross_correlate_factorised_set_tile_shape(int shape)
{
    return oskar_amd_fact_set_tile_shape(shape);
}

int oskar_cross_correlate_factorised_set_part_size(int sources)
{
    const int old = part_size;
    if (sources > 0) part_size = ((sources + 7) / 8) * 8;
    return old;
}

int oskar_cross_correlate_factorised_jones_set_part_size(int sources)
{
    const int old = part_size_jones;
    if (sources > 0)
        part_size_jones = sources > 2048 ? 2048 : ((sources + 3) / 4) * 4;
    return old;
}

/* Checks of the arguments one unit shares with oskar_cross_correlate(),
 * in the reference's order: location, then types, then dimensions
 * (oskar/correlate/src/oskar_cross_correlate.c:73-110). */
static int check_unit(int location, int num_stations, int num_sources,
        const oskar_Mem* station_beam, const oskar_Mem* const src_flux[4],
        const oskar_Mem* const src_dir[3],
        const oskar_Mem* const station_uvw[3], int offset_out,
        const oskar_Mem* vis)
{
    int i = 0;
    const int prec = oskar_mem_precision(vis);
    const size_t num_baselines = (size_t) num_stations *
            (size_t) (num_stations - 1) / 2;
    if (oskar_mem_location(vis) != location) return OSKAR_ERR_LOCATION_MISMATCH;
    for (i = 0; i < 3; ++i)
    {
        if (!station_uvw[i] || !src_dir[i] ||
                oskar_mem_location(station_uvw[i]) != location ||
                oskar_mem_location(src_dir[i]) != location)
            return OSKAR_ERR_LOCATION_MISMATCH;
    }
    for (i = 0; i < 4; ++i)
    {
        if (!src_flux[i] || oskar_mem_location(src_flux[i]) != location)
            return OSKAR_ERR_LOCATION_MISMATCH;
    }
    if (station_beam && oskar_mem_location(station_beam) != location)
        return OSKAR_ERR_LOCATION_MISMATCH;
    if (!oskar_mem_is_matrix(vis) || !oskar_mem_is_complex(vis) ||
            (prec != OSKAR_SINGLE && prec != OSKAR_DOUBLE))
        return OSKAR_ERR_BAD_DATA_TYPE;
    for (i = 0; i < 3; ++i)
    {
        if (oskar_mem_type(station_uvw[i]) != prec ||
                oskar_mem_type(src_dir[i]) != prec)
            return OSKAR_ERR_TYPE_MISMATCH;
    }
    for (i = 0; i < 4; ++i)
    {
        if (oskar_mem_type(src_flux[i]) != prec)
            return OSKAR_ERR_TYPE_MISMATCH;
    }
    if (station_beam && oskar_mem_type(station_beam) != oskar_mem_type(vis))
        return OSKAR_ERR_TYPE_MISMATCH;
    if (num_sources < 0 || offset_out < 0)
        return OSKAR_ERR_INVALID_ARGUMENT;
    for (i = 0; i < 3; ++i)
    {
        if ((int) oskar_mem_length(station_uvw[i]) != num_stations ||
                (int) oskar_mem_length(src_dir[i]) < num_sources)
            return OSKAR_ERR_DIMENSION_MISMATCH;
    }
    for (i = 0; i < 4; ++i)
    {
        if ((int) oskar_mem_length(src_flux[i]) < num_sources)
            return OSKAR_ERR_DIMENSION_MISMATCH;
    }
    if ((station_beam && (int) oskar_mem_length(station_beam) < num_sources)
            || oskar_mem_length(vis) < (size_t) offset_out + num_baselines)
        return OSKAR_ERR_DIMENSION_MISMATCH;
    if (location != OSKAR_GPU) return OSKAR_ERR_BAD_LOCATION;
    return 0;
}

/* Smearing factors depend on the baseline and do not factorise. */
static int smearing_on(const oskar_Telescope* tel)
{
    return oskar_telescope_channel_bandwidth_hz(tel) != 0.0 ||
            (oskar_telescope_time_average_sec(tel) != 0.0 &&
             oskar_telescope_phase_centre_coord_type(tel) !=
             OSKAR_COORDS_AZEL);
}

/* The uv filter in wavelengths at one frequency
 * (oskar/correlate/src/oskar_cross_correlate.c:60-71). */
static void uv_filter(const oskar_Telescope* tel, double frequency_hz,
        double* uv_min, double* uv_max)
{
    const double inv_wavelength = fabs(frequency_hz) / 299792458.0;
    *uv_min = oskar_telescope_uv_filter_min(tel);
    *uv_max = oskar_telescope_uv_filter_max(tel);
    if (oskar_telescope_uv_filter_units(tel) == OSKAR_METRES)
    {
        *uv_min *= inv_wavelength;
        *uv_max *= inv_wavelength;
    }
    if (*uv_max < 0.0 || *uv_max > FLT_MAX) *uv_max = FLT_MAX;
}

void oskar_cross_correlate_factorised(int use_casa, int num_sources,
        const oskar_Mem* station_beam, const oskar_Mem* const src_flux[4],
        const oskar_Mem* const src_dir[3], const oskar_Telescope* tel,
        const oskar_Mem* const station_uvw[3], double frequency_hz,
        int mode, int offset_out, oskar_Mem* vis, int* status)
{
    double uv_min = 0.0, uv_max = 0.0;
    if (*status) return;
    const int num_stations = oskar_telescope_num_stations(tel);
    frequency_hz = fabs(frequency_hz);
    uv_filter(tel, frequency_hz, &uv_min, &uv_max);
    const int location = oskar_mem_location(vis);
    const int prec = oskar_mem_precision(vis);
    if (oskar_telescope_mem_location(tel) != location)
    {
        *status = OSKAR_ERR_LOCATION_MISMATCH;
        return;
    }
    *status = check_unit(location, num_stations, num_sources, station_beam,
            src_flux, src_dir, station_uvw, offset_out, vis);
    if (*status) return;
    if (mode != OSKAR_FACTORISED_FP16X3 && mode != OSKAR_FACTORISED_BF16X3)
    {
        *status = OSKAR_ERR_INVALID_ARGUMENT;
        return;
    }
    if (smearing_on(tel))
    {
        *status = OSKAR_ERR_FUNCTION_NOT_AVAILABLE;
        return;
    }
    {
        const int e = oskar_amd_fact_launch(prec, mode, use_casa,
                num_sources, num_stations,
                station_beam ? oskar_mem_void_const(station_beam) : 0,
                oskar_mem_void_const(src_flux[0]),
                oskar_mem_void_const(src_flux[1]),
                oskar_mem_void_const(src_flux[2]),
                oskar_mem_void_const(src_flux[3]),
                oskar_mem_void_const(src_dir[0]),
                oskar_mem_void_const(src_dir[1]),
                oskar_mem_void_const(src_dir[2]),
                oskar_mem_void_const(station_uvw[0]),
                oskar_mem_void_const(station_uvw[1]),
                oskar_mem_void_const(station_uvw[2]), frequency_hz,
                uv_min, uv_max, offset_out, oskar_mem_void(vis), part_size,
                oskar_amd_current_stream());
        if (e) *status = e;
    }
}

void oskar_cross_correlate_factorised_jones(int use_casa, int num_sources,
        const oskar_Jones* E, const oskar_Mem* const src_flux[4],
        const oskar_Mem* const src_dir[3], const oskar_Telescope* tel,
        const oskar_Mem* const station_uvw[3], double frequency_hz,
        int mode, int offset_out, oskar_Mem* vis, int* status)
{
    double uv_min = 0.0, uv_max = 0.0;
    const oskar_Mem* e_mem = 0;
    if (*status) return;
    if (!E || !vis || !tel)
    {
        *status = OSKAR_ERR_INVALID_ARGUMENT;
        return;
    }
    const int num_stations = oskar_telescope_num_stations(tel);
    frequency_hz = fabs(frequency_hz);
    uv_filter(tel, frequency_hz, &uv_min, &uv_max);
    const int location = oskar_mem_location(vis);
    const int prec = oskar_mem_precision(vis);
    e_mem = oskar_jones_mem_const(E);
    if (oskar_telescope_mem_location(tel) != location ||
            oskar_jones_mem_location(E) != location)
    {
        *status = OSKAR_ERR_LOCATION_MISMATCH;
        return;
    }
    *status = check_unit(location, num_stations, num_sources, e_mem,
            src_flux, src_dir, station_uvw, offset_out, vis);
    if (*status) return;
    if (oskar_jones_num_stations(E) != num_stations ||
            oskar_jones_num_sources(E) < num_sources)
    {
        *status = OSKAR_ERR_DIMENSION_MISMATCH;
        return;
    }
    if (mode != OSKAR_FACTORISED_FP16X3 && mode != OSKAR_FACTORISED_BF16X3)
    {
        *status = OSKAR_ERR_INVALID_ARGUMENT;
        return;
    }
    if (smearing_on(tel))
    {
        *status = OSKAR_ERR_FUNCTION_NOT_AVAILABLE;
        return;
    }
    {
        const int e = oskar_amd_fact_stations_launch(prec, mode, use_casa,
                num_sources, num_stations, oskar_mem_void_const(e_mem),
                (size_t) oskar_jones_num_sources(E),
                oskar_mem_void_const(src_flux[0]),
                oskar_mem_void_const(src_flux[1]),
                oskar_mem_void_const(src_flux[2]),
                oskar_mem_void_const(src_flux[3]),
                oskar_mem_void_const(src_dir[0]),
                oskar_mem_void_const(src_dir[1]),
                oskar_mem_void_const(src_dir[2]),
                oskar_mem_void_const(station_uvw[0]),
                oskar_mem_void_const(station_uvw[1]),
                oskar_mem_void_const(station_uvw[2]), frequency_hz,
                uv_min, uv_max, offset_out, oskar_mem_void(vis),
                part_size_jones, oskar_amd_current_stream());
        if (e) *status = e;
    }
}

/* ------------------------------------------------------------------ */
/* Batches.                                                            */

struct oskar_FactorisedBatch
{
    void* impl;
    int precision, mode, num_stations, max_units;
    double* frequency;           /* per unit (Hz) */
    int* num_sources;            /* per unit */
};

/* Tuned configurations for this process, keyed by (stations, precision,
 * mode, source-count class = floor(log2(mean sources per unit))). */
#define TUNE_CACHE_SIZE 64
static struct
{
    int key[4];
    int config[3];
} tune_cache[TUNE_CACHE_SIZE];
static int tune_cache_len = 0;

oskar_FactorisedBatch* oskar_factorised_batch_create(int precision, int mode,
        int num_stations, int max_units, int* status)
{
    oskar_FactorisedBatch* b = 0;
    int e = 0;
    if (*status) return 0;
    if (precision != OSKAR_SINGLE && precision != OSKAR_DOUBLE)
    {
        *status = OSKAR_ERR_BAD_DATA_TYPE;
        return 0;
    }
    if ((mode != OSKAR_FACTORISED_FP16X3 && mode != OSKAR_FACTORISED_BF16X3)
            || num_stations < 2 || max_units < 1)
    {
        *status = OSKAR_ERR_INVALID_ARGUMENT;
        return 0;
    }
    b = (oskar_FactorisedBatch*) calloc(1, sizeof(*b));
    if (!b)
    {
        *status = OSKAR_ERR_MEMORY_ALLOC_FAILURE;
        return 0;
    }
    b->precision = precision;
    b->mode = mode;
    b->num_stations = num_stations;
    b->max_units = max_units;
    b->frequency = (double*) calloc((size_t) max_units, sizeof(double));
    b->num_sources = (int*) calloc((size_t) max_units, sizeof(int));
    if (!b->frequency || !b->num_sources)
    {
        free(b->frequency);
        free(b->num_sources);
        free(b);
        *status = OSKAR_ERR_MEMORY_ALLOC_FAILURE;
        return 0;
    }
    b->impl = oskar_amd_fact_batch_create(precision, mode, num_stations,
            max_units, &e);
    if (e)
    {
        free(b->frequency);
        free(b->num_sources);
        free(b);
        *status = e;
        return 0;
    }
    return b;
}

void oskar_factorised_batch_set_unit(oskar_FactorisedBatch* batch, int unit,
        int num_sources, const oskar_Mem* station_beam,
        const oskar_Mem* const src_flux[4], const oskar_Mem* const src_dir[3],
        const oskar_Mem* const station_uvw[3], double frequency_hz,
        int offset_out, oskar_Mem* vis, int* status)
{
    if (*status) return;
    if (!batch || unit < 0 || unit >= batch->max_units)
    {
        *status = OSKAR_ERR_INVALID_ARGUMENT;
        return;
    }
    *status = check_unit(OSKAR_GPU, batch->num_stations, num_sources,
            station_beam, src_flux, src_dir, station_uvw, offset_out, vis);
    if (*status) return;
    if (oskar_mem_precision(vis) != batch->precision)
    {
        *status = OSKAR_ERR_TYPE_MISMATCH;
        return;
    }
    frequency_hz = fabs(frequency_hz);
    batch->frequency[unit] = frequency_hz;
    batch->num_sources[unit] = num_sources;
    oskar_amd_fact_batch_set_unit(batch->impl, unit, num_sources,
            station_beam ? oskar_mem_void_const(station_beam) : 0,
            oskar_mem_void_const(src_flux[0]),
            oskar_mem_void_const(src_flux[1]),
            oskar_mem_void_const(src_flux[2]),
            oskar_mem_void_const(src_flux[3]),
            oskar_mem_void_const(src_dir[0]),
            oskar_mem_void_const(src_dir[1]),
            oskar_mem_void_const(src_dir[2]),
            oskar_mem_void_const(station_uvw[0]),
            oskar_mem_void_const(station_uvw[1]),
            oskar_mem_void_const(station_uvw[2]), frequency_hz, offset_out,
            oskar_mem_void(vis));
}

static int units_ready(const oskar_FactorisedBatch* batch, int num_units)
{
    int u = 0;
    if (!batch || num_units < 0 || num_units > batch->max_units)
        return OSKAR_ERR_INVALID_ARGUMENT;
    for (u = 0; u < num_units; ++u)
        if (!oskar_amd_fact_batch_unit_is_set(batch->impl, u))
            return OSKAR_ERR_INVALID_ARGUMENT;
    return 0;
}

void oskar_factorised_batch_run(oskar_FactorisedBatch* batch, int num_units,
        int use_casa, const oskar_Telescope* tel, int* status)
{
    int u = 0;
    double *uv_min = 0, *uv_max = 0;
    if (*status) return;
    *status = units_ready(batch, num_units);
    if (*status) return;
    if (oskar_telescope_mem_location(tel) != OSKAR_GPU)
    {
        *status = OSKAR_ERR_LOCATION_MISMATCH;
        return;
    }
    if (oskar_telescope_num_stations(tel) != batch->num_stations)
    {
        *status = OSKAR_ERR_DIMENSION_MISMATCH;
        return;
    }
    if (smearing_on(tel))
    {
        *status = OSKAR_ERR_FUNCTION_NOT_AVAILABLE;
        return;
    }
    if (num_units == 0) return;
    uv_min = (double*) malloc((size_t) num_units * sizeof(double));
    uv_max = (double*) malloc((size_t) num_units * sizeof(double));
    if (!uv_min || !uv_max)
    {
        free(uv_min);
        free(uv_max);
        *status = OSKAR_ERR_MEMORY_ALLOC_FAILURE;
        return;
    }
    for (u = 0; u < num_units; ++u)
        uv_filter(tel, batch->frequency[u], &uv_min[u], &uv_max[u]);
    *status = oskar_amd_fact_batch_run(batch->impl, num_units, use_casa,
            uv_min, uv_max, oskar_amd_current_stream());
    free(uv_min);
    free(uv_max);
}

void oskar_factorised_batch_autotune(oskar_FactorisedBatch* batch,
        int num_units, int* status)
{
    int u = 0, i = 0, key[4];
    double mean = 0.0;
    if (*status) return;
    *status = units_ready(batch, num_units);
    if (*status) return;
    if (num_units == 0) return;
    for (u = 0; u < num_units; ++u) mean += batch->num_sources[u];
    mean /= num_units;
    key[0] = batch->num_stations;
    key[1] = batch->precision;
    key[2] = batch->mode;
    key[3] = mean >= 1.0 ? (int) floor(log2(mean)) : 0;
    for (i = 0; i < tune_cache_len; ++i)
    {
        if (!memcmp(tune_cache[i].key, key, sizeof(key)))
        {
            int cfg[4];
            oskar_amd_fact_batch_get_config(batch->impl, cfg);
            *status = oskar_amd_fact_batch_set_config(batch->impl,
                    tune_cache[i].config[0], tune_cache[i].config[1],
                    tune_cache[i].config[2], cfg[3]);
            return;
        }
    }
    *status = oskar_amd_fact_batch_autotune(batch->impl, num_units, 2,
            oskar_amd_current_stream(), 0);
    if (*status) return;
    {
        int cfg[4];
        const int slot = tune_cache_len < TUNE_CACHE_SIZE ?
                tune_cache_len++ : TUNE_CACHE_SIZE - 1;
        oskar_amd_fact_batch_get_config(batch->impl, cfg);
        memcpy(tune_cache[slot].key, key, sizeof(key));
        memcpy(tune_cache[slot].config, cfg, 3 * sizeof(int));
    }
}

void oskar_factorised_batch_set_config(oskar_FactorisedBatch* batch,
        int tile_shape, int src_per_part, int num_workgroups, int use_graph,
        int* status)
{
    if (*status) return;
    if (!batch)
    {
        *status = OSKAR_ERR_INVALID_ARGUMENT;
        return;
    }
    if (oskar_amd_fact_batch_set_config(batch->impl, tile_shape,
            src_per_part, num_workgroups, use_graph))
        *status = OSKAR_ERR_INVALID_ARGUMENT;
}

void oskar_factorised_batch_config(const oskar_FactorisedBatch* batch,
        int config[4])
{
    if (batch) oskar_amd_fact_batch_get_config(batch->impl, config);
}

double oskar_factorised_batch_last_run_ms(oskar_FactorisedBatch* batch)
{
    double ms = -1.0;
    if (!batch || oskar_amd_fact_batch_last_ms(batch->impl, &ms)) return -1.0;
    return ms;
}

void oskar_factorised_batch_free(oskar_FactorisedBatch* batch, int* status)
{
    (void) status;
    if (!batch) return;
    oskar_amd_fact_batch_free(batch->impl);
    free(batch->frequency);
    free(batch->num_sources);
    free(batch);
}
