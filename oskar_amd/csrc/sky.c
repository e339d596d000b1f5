/*
 * oskar_Sky column table, the horizon mask, the prefix sum, the masked row
 * copy and the horizon clip.
 *
 * Reference: oskar/sky/src/oskar_sky_create.c, oskar_sky_free.c,
 * oskar_sky_resize.c, oskar_sky_column.c, oskar_sky_accessors.c,
 * oskar_sky_copy_source_data.c, oskar_sky_horizon_clip.c,
 * oskar_update_horizon_mask.c; oskar/math/src/oskar_prefix_sum.c.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <hip/hip_runtime_api.h>

#include "private.h"
#include "oskar_sky.h"

#define ROUND_CAPACITY(n) (((n) + 1 + 255) / 256 * 256)

oskar_Sky* oskar_sky_create(int type, int location, int num_sources,
        int* status)
{
    oskar_Sky* sky = 0;
    if (*status) return 0;
    if (type != OSKAR_SINGLE && type != OSKAR_DOUBLE)
    {
        *status = OSKAR_ERR_BAD_DATA_TYPE;
        return 0;
    }
    if (num_sources < 0)
    {
        *status = OSKAR_ERR_INVALID_ARGUMENT;
        return 0;
    }
    sky = (oskar_Sky*) calloc(1, sizeof(oskar_Sky));
    if (!sky)
    {
        *status = OSKAR_ERR_MEMORY_ALLOC_FAILURE;
        return 0;
    }
    const int capacity = ROUND_CAPACITY(num_sources);
    sky->attr_int[OSKAR_SKY_PRECISION] = type;
    sky->attr_int[OSKAR_SKY_MEM_LOCATION] = location;
    sky->attr_int[OSKAR_SKY_CAPACITY] = capacity;
    sky->attr_int[OSKAR_SKY_NUM_SOURCES] = num_sources;
    sky->table = oskar_mem_create(type, location, 0, status);
    sky->num_valid_columns = oskar_mem_create(OSKAR_INT, location,
            (size_t) capacity * OSKAR_SKY_NUM_FIXED_COLUMN_TYPES, status);
    oskar_mem_clear_contents(sky->num_valid_columns, status);
    if (*status)
    {
        int s = 0;
        oskar_sky_free(sky, &s);
        return 0;
    }
    return sky;
}

void oskar_sky_free(oskar_Sky* sky, int* status)
{
    int i = 0;
    if (!sky) return;
    for (i = 0; i < sky->attr_int[OSKAR_SKY_NUM_COLUMNS]; ++i)
        oskar_mem_free(sky->columns[i], status);
    oskar_mem_free(sky->table, status);
    oskar_mem_free(sky->num_valid_columns, status);
    free(sky->columns);
    free(sky->column_attr);
    free(sky->column_type);
    free(sky);
}

double oskar_sky_double(const oskar_Sky* sky, oskar_SkyAttribDouble a)
{
    return ((int) a < 0 || a >= OSKAR_SKY_NUM_ATTRIBUTES_DOUBLE) ? 0.0 :
            sky->attr_double[a];
}

int oskar_sky_int(const oskar_Sky* sky, oskar_SkyAttribInt a)
{
    return ((int) a < 0 || a >= OSKAR_SKY_NUM_ATTRIBUTES_INT) ? 0 :
            sky->attr_int[a];
}

void oskar_sky_set_double(oskar_Sky* sky, oskar_SkyAttribDouble a,
        double value)
{
    if (a == OSKAR_SKY_REF_RA_RAD || a == OSKAR_SKY_REF_DEC_RAD)
        sky->attr_double[a] = value;
}

void oskar_sky_set_int(oskar_Sky* sky, oskar_SkyAttribInt a, int value)
{
    /* Precision, location, capacity and column count are read-only; the
     * source count must stay below the capacity. */
    if (a == OSKAR_SKY_NUM_SOURCES)
    {
        if (value >= 0 && value < sky->attr_int[OSKAR_SKY_CAPACITY])
            sky->attr_int[a] = value;
    }
    else if (a == OSKAR_SKY_USE_EXTENDED)
        sky->attr_int[a] = value;
}

int oskar_sky_column_attribute(const oskar_Sky* sky, int i)
{
    return (i < 0 || i >= sky->attr_int[OSKAR_SKY_NUM_COLUMNS]) ? -1 :
            sky->column_attr[i];
}

oskar_SkyColumn oskar_sky_column_type(const oskar_Sky* sky, int i)
{
    return (i < 0 || i >= sky->attr_int[OSKAR_SKY_NUM_COLUMNS]) ?
            OSKAR_SKY_CUSTOM : sky->column_type[i];
}

static size_t valid_index(oskar_SkyColumn type, int index)
{
    return (size_t) index * OSKAR_SKY_NUM_FIXED_COLUMN_TYPES + (size_t) type;
}

int oskar_sky_num_valid_columns_of_type(const oskar_Sky* sky,
        oskar_SkyColumn type, int index)
{
    int value = 0;
    if ((int) type < 0 || type >= OSKAR_SKY_NUM_FIXED_COLUMN_TYPES ||
            index < 0 || index >= sky->attr_int[OSKAR_SKY_CAPACITY])
        return 0;
    const int* p = (const int*) oskar_mem_void_const(sky->num_valid_columns) +
            valid_index(type, index);
    if (sky->attr_int[OSKAR_SKY_MEM_LOCATION] == OSKAR_CPU) return *p;
    if (hipMemcpy(&value, p, sizeof(int), hipMemcpyDeviceToHost) !=
            hipSuccess)
        return 0;
    return value;
}

void oskar_sky_set_num_valid_columns_of_type(oskar_Sky* sky,
        oskar_SkyColumn type, int index, int value)
{
    if ((int) type < 0 || type >= OSKAR_SKY_NUM_FIXED_COLUMN_TYPES ||
            index < 0 || index >= sky->attr_int[OSKAR_SKY_CAPACITY])
        return;
    int* p = (int*) oskar_mem_void(sky->num_valid_columns) +
            valid_index(type, index);
    if (sky->attr_int[OSKAR_SKY_MEM_LOCATION] == OSKAR_CPU) *p = value;
    else (void) hipMemcpy(p, &value, sizeof(int), hipMemcpyHostToDevice);
}

oskar_Mem* oskar_sky_column(oskar_Sky* sky, oskar_SkyColumn type, int attr,
        int* status)
{
    int i = 0;
    if (*status || (int) type < 0) return 0;
    const int num = sky->attr_int[OSKAR_SKY_NUM_COLUMNS];
    for (i = 0; i < num; ++i)
        if (sky->column_type[i] == type && sky->column_attr[i] == attr)
            return sky->columns[i];

    /* Not found: append a zeroed column.  Growing the table moves it, so
     * every alias is re-pointed. */
    const int n = num + 1;
    oskar_SkyColumn* t = (oskar_SkyColumn*) realloc(sky->column_type,
            (size_t) n * sizeof(oskar_SkyColumn));
    if (t) sky->column_type = t;
    int* a = (int*) realloc(sky->column_attr, (size_t) n * sizeof(int));
    if (a) sky->column_attr = a;
    oskar_Mem** c = (oskar_Mem**) realloc(sky->columns,
            (size_t) n * sizeof(oskar_Mem*));
    if (c) sky->columns = c;
    if (!t || !a || !c)
    {
        *status = OSKAR_ERR_MEMORY_ALLOC_FAILURE;
        return 0;
    }
    const size_t capacity = (size_t) sky->attr_int[OSKAR_SKY_CAPACITY];
    oskar_mem_realloc(sky->table, capacity * n, status);
    if (*status) return 0;
    sky->columns[num] = oskar_mem_create_alias(0, 0, 0, status);
    if (*status) return 0;
    sky->column_type[num] = type;
    sky->column_attr[num] = attr;
    sky->attr_int[OSKAR_SKY_NUM_COLUMNS] = n;
    for (i = 0; i < n; ++i)
        oskar_mem_set_alias(sky->columns[i], sky->table, i * capacity,
                capacity, status);
    oskar_mem_clear_contents(sky->columns[num], status);
    return sky->columns[num];
}

const oskar_Mem* oskar_sky_column_const(const oskar_Sky* sky,
        oskar_SkyColumn type, int attr)
{
    int i = 0;
    if ((int) type < 0) return 0;
    for (i = 0; i < sky->attr_int[OSKAR_SKY_NUM_COLUMNS]; ++i)
        if (sky->column_type[i] == type && sky->column_attr[i] == attr)
            return sky->columns[i];
    return 0;
}

void oskar_sky_resize(oskar_Sky* sky, int num_sources, int* status)
{
    size_t i = 0;
    if (*status || sky->attr_int[OSKAR_SKY_NUM_SOURCES] == num_sources)
        return;
    if (num_sources < 0)
    {
        *status = OSKAR_ERR_INVALID_ARGUMENT;
        return;
    }
    const size_t new_cap = (size_t) ROUND_CAPACITY(num_sources);
    const size_t old_cap = (size_t) sky->attr_int[OSKAR_SKY_CAPACITY];
    const size_t ncol = (size_t) sky->attr_int[OSKAR_SKY_NUM_COLUMNS];
    oskar_Mem* table = oskar_mem_create(sky->attr_int[OSKAR_SKY_PRECISION],
            sky->attr_int[OSKAR_SKY_MEM_LOCATION], new_cap * ncol, status);
    oskar_mem_clear_contents(table, status);
    if (*status)
    {
        int s = 0;
        oskar_mem_free(table, &s);
        return;
    }
    /* Rows past the old capacity start with no valid columns. */
    {
        const size_t nf = OSKAR_SKY_NUM_FIXED_COLUMN_TYPES;
        oskar_Mem* valid = oskar_mem_create(OSKAR_INT,
                sky->attr_int[OSKAR_SKY_MEM_LOCATION], new_cap * nf, status);
        oskar_mem_clear_contents(valid, status);
        oskar_mem_copy_contents(valid, sky->num_valid_columns, 0, 0,
                (old_cap < new_cap ? old_cap : new_cap) * nf, status);
        if (*status)
        {
            int s = 0;
            oskar_mem_free(valid, &s);
            oskar_mem_free(table, &s);
            return;
        }
        oskar_mem_free(sky->num_valid_columns, status);
        sky->num_valid_columns = valid;
    }
    const size_t keep = old_cap < new_cap ? old_cap : new_cap;
    for (i = 0; i < ncol; ++i)
    {
        oskar_mem_copy_contents(table, sky->table, i * new_cap, i * old_cap,
                keep, status);
        oskar_mem_set_alias(sky->columns[i], table, i * new_cap, new_cap,
                status);
    }
    oskar_mem_free(sky->table, status);
    sky->table = table;
    sky->attr_int[OSKAR_SKY_CAPACITY] = (int) new_cap;
    sky->attr_int[OSKAR_SKY_NUM_SOURCES] = num_sources;
}

/* (ll, mm, nn) of oskar_update_horizon_mask.c:37-44. */
static void horizon_factors(double ha0, double dec0, double lat, double* f)
{
    const double cos_ha0 = cos(ha0), sin_dec0 = sin(dec0),
            cos_dec0 = cos(dec0), sin_lat = sin(lat), cos_lat = cos(lat);
    f[0] = cos_lat * sin(ha0);
    f[1] = sin_lat * cos_dec0 - cos_lat * cos_ha0 * sin_dec0;
    f[2] = sin_lat * sin_dec0 + cos_lat * cos_ha0 * cos_dec0;
}

void oskar_update_horizon_mask(int num_sources, const oskar_Mem* l,
        const oskar_Mem* m, const oskar_Mem* n, const double ha0_rad,
        const double dec0_rad, const double lat_rad, oskar_Mem* mask,
        int* status)
{
    if (*status) return;
    const int type = oskar_mem_precision(l);
    const int location = oskar_mem_location(mask);
    if (type != OSKAR_SINGLE && type != OSKAR_DOUBLE)
    {
        *status = OSKAR_ERR_BAD_DATA_TYPE;
        return;
    }
    if (oskar_mem_location(l) != location ||
            oskar_mem_location(m) != location ||
            oskar_mem_location(n) != location)
    {
        *status = OSKAR_ERR_LOCATION_MISMATCH;
        return;
    }
    if (oskar_mem_type(m) != type || oskar_mem_type(n) != type ||
            oskar_mem_type(mask) != OSKAR_INT)
    {
        *status = OSKAR_ERR_TYPE_MISMATCH;
        return;
    }
    if (num_sources < 0 || (int) oskar_mem_length(l) < num_sources ||
            (int) oskar_mem_length(m) < num_sources ||
            (int) oskar_mem_length(n) < num_sources ||
            (int) oskar_mem_length(mask) < num_sources)
    {
        *status = OSKAR_ERR_DIMENSION_MISMATCH;
        return;
    }
    if (location != OSKAR_GPU)
    {
        *status = OSKAR_ERR_BAD_LOCATION;
        return;
    }
    oskar_HorizonModels models;
    memset(&models, 0, sizeof(models));
    models.num = 1;
    horizon_factors(ha0_rad, dec0_rad, lat_rad, models.lmn[0]);
    const int e = oskar_amd_horizon_mask_launch(type, num_sources,
            oskar_mem_void_const(l), oskar_mem_void_const(m),
            oskar_mem_void_const(n), &models, oskar_mem_int(mask, status),
            oskar_amd_current_stream());
    if (e && !*status) *status = e;
}

void oskar_prefix_sum(size_t num_elements, const oskar_Mem* in,
        oskar_Mem* out, int* status)
{
    if (*status) return;
    const int location = oskar_mem_location(in);
    if (oskar_mem_location(out) != location)
    {
        *status = OSKAR_ERR_LOCATION_MISMATCH;
        return;
    }
    if (oskar_mem_type(in) != oskar_mem_type(out))
    {
        *status = OSKAR_ERR_TYPE_MISMATCH;
        return;
    }
    if (oskar_mem_type(in) != OSKAR_INT)
    {
        *status = OSKAR_ERR_BAD_DATA_TYPE;
        return;
    }
    if (oskar_mem_length(out) < num_elements + 1 ||
            oskar_mem_length(in) < num_elements || num_elements > 0x7FFFFFFE)
    {
        *status = OSKAR_ERR_DIMENSION_MISMATCH;
        return;
    }
    if (location != OSKAR_GPU)
    {
        *status = OSKAR_ERR_BAD_LOCATION;
        return;
    }
    const int e = oskar_amd_prefix_sum_launch((int) num_elements,
            oskar_mem_int_const(in, status), oskar_mem_int(out, status),
            oskar_amd_current_stream());
    if (e && !*status) *status = e;
}

/* Touches every input column in the output, in the input's order, and
 * checks that the order agrees (oskar_sky_copy_source_data.c:30-50). */
static void match_columns(const oskar_Sky* in, oskar_Sky* out, int* status)
{
    int c = 0;
    for (c = 0; c < in->attr_int[OSKAR_SKY_NUM_COLUMNS]; ++c)
    {
        const oskar_SkyColumn t = oskar_sky_column_type(in, c);
        const int a = oskar_sky_column_attribute(in, c);
        (void) oskar_sky_column(out, t, a, status);
        if (*status) return;
        if (t != oskar_sky_column_type(out, c) ||
                a != oskar_sky_column_attribute(out, c))
        {
            *status = OSKAR_ERR_TYPE_MISMATCH;
            return;
        }
    }
}

/* Metadata of oskar_sky_copy_source_data.c:142-154, keeping the output's
 * own capacity and column count. */
static void copy_metadata(const oskar_Sky* in, oskar_Sky* out, int num_out)
{
    out->attr_int[OSKAR_SKY_PRECISION] = in->attr_int[OSKAR_SKY_PRECISION];
    out->attr_int[OSKAR_SKY_MEM_LOCATION] =
            in->attr_int[OSKAR_SKY_MEM_LOCATION];
    out->attr_int[OSKAR_SKY_USE_EXTENDED] =
            in->attr_int[OSKAR_SKY_USE_EXTENDED];
    out->attr_double[OSKAR_SKY_REF_RA_RAD] =
            in->attr_double[OSKAR_SKY_REF_RA_RAD];
    out->attr_double[OSKAR_SKY_REF_DEC_RAD] =
            in->attr_double[OSKAR_SKY_REF_DEC_RAD];
    out->attr_int[OSKAR_SKY_NUM_SOURCES] = num_out;
}

static int read_total(const int* indices, int num_in, int* status)
{
    int total = 0;
    hipStream_t s = (hipStream_t) oskar_amd_current_stream();
    hipError_t e = hipMemcpyAsync(&total, indices + num_in, sizeof(int),
            hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess && !*status) *status = (int) e;
    return total;
}

void oskar_sky_copy_source_data(const oskar_Sky* in,
        const oskar_Mem* horizon_mask, const oskar_Mem* indices,
        oskar_Sky* out, int* status)
{
    if (*status) return;
    const int location = in->attr_int[OSKAR_SKY_MEM_LOCATION];
    const int type = out->attr_int[OSKAR_SKY_PRECISION];
    const int num_in = in->attr_int[OSKAR_SKY_NUM_SOURCES];
    if (type != in->attr_int[OSKAR_SKY_PRECISION])
    {
        *status = OSKAR_ERR_TYPE_MISMATCH;
        return;
    }
    if (out->attr_int[OSKAR_SKY_MEM_LOCATION] != location ||
            oskar_mem_location(horizon_mask) != location ||
            oskar_mem_location(indices) != location)
    {
        *status = OSKAR_ERR_LOCATION_MISMATCH;
        return;
    }
    if (oskar_mem_type(horizon_mask) != OSKAR_INT ||
            oskar_mem_type(indices) != OSKAR_INT)
    {
        *status = OSKAR_ERR_TYPE_MISMATCH;
        return;
    }
    if ((int) oskar_mem_length(horizon_mask) < num_in ||
            (int) oskar_mem_length(indices) < num_in + 1)
    {
        *status = OSKAR_ERR_DIMENSION_MISMATCH;
        return;
    }
    if (location != OSKAR_GPU)
    {
        *status = OSKAR_ERR_BAD_LOCATION;
        return;
    }
    match_columns(in, out, status);
    if (*status) return;
    /* Every kept row must fit: the total cannot exceed num_in, so an output
     * of at least the input's capacity is always large enough. */
    const int total = read_total(oskar_mem_int_const(indices, status), num_in,
            status);
    if (*status) return;
    if (total < 0 || total > num_in ||
            total >= out->attr_int[OSKAR_SKY_CAPACITY])
    {
        *status = OSKAR_ERR_DIMENSION_MISMATCH;
        return;
    }
    const int e = oskar_amd_sky_copy_launch(type, num_in,
            in->attr_int[OSKAR_SKY_CAPACITY],
            out->attr_int[OSKAR_SKY_CAPACITY],
            oskar_mem_int_const(horizon_mask, status),
            oskar_mem_int_const(indices, status),
            in->attr_int[OSKAR_SKY_NUM_COLUMNS],
            oskar_mem_void_const(in->table), oskar_mem_void(out->table),
            oskar_mem_int_const(in->num_valid_columns, status),
            oskar_mem_int(out->num_valid_columns, status),
            oskar_amd_current_stream());
    if (e && !*status) *status = e;
    copy_metadata(in, out, total);
}

void oskar_sky_horizon_clip(oskar_Sky* out, const oskar_Sky* in,
        const oskar_Telescope* telescope, double gast_rad,
        oskar_StationWork* work, int* status)
{
    int i = 0;
    if (*status) return;
    oskar_Mem* mask = oskar_station_work_horizon_mask(work);
    oskar_Mem* indices = oskar_station_work_source_indices(work);
    if (in->attr_int[OSKAR_SKY_PRECISION] !=
            out->attr_int[OSKAR_SKY_PRECISION])
    {
        *status = OSKAR_ERR_TYPE_MISMATCH;
        return;
    }
    const int location = out->attr_int[OSKAR_SKY_MEM_LOCATION];
    if (in->attr_int[OSKAR_SKY_MEM_LOCATION] != location ||
            oskar_mem_location(mask) != location ||
            oskar_mem_location(indices) != location)
    {
        *status = OSKAR_ERR_LOCATION_MISMATCH;
        return;
    }
    if (location != OSKAR_GPU)
    {
        *status = OSKAR_ERR_BAD_LOCATION;
        return;
    }
    const int type = in->attr_int[OSKAR_SKY_PRECISION];
    const int num_in = in->attr_int[OSKAR_SKY_NUM_SOURCES];
    const double ra0 = in->attr_double[OSKAR_SKY_REF_RA_RAD];
    const double dec0 = in->attr_double[OSKAR_SKY_REF_DEC_RAD];
    const oskar_Mem* l = oskar_sky_column_const(in, OSKAR_SKY_SCRATCH_L, 0);
    const oskar_Mem* m = oskar_sky_column_const(in, OSKAR_SKY_SCRATCH_M, 0);
    const oskar_Mem* n = oskar_sky_column_const(in, OSKAR_SKY_SCRATCH_N, 0);
    if (!l || !m || !n)
    {
        *status = OSKAR_ERR_DIMENSION_MISMATCH;
        return;
    }

    if (out->attr_int[OSKAR_SKY_CAPACITY] < num_in)
        oskar_sky_resize(out, num_in, status);
    oskar_mem_ensure(mask, (size_t) num_in, status);
    oskar_mem_ensure(indices, (size_t) num_in + 1, status);
    match_columns(in, out, status);
    if (*status) return;
    /* The column match may have grown the output table; the scatter needs
     * room for every input row. */
    if (out->attr_int[OSKAR_SKY_CAPACITY] <= num_in)
    {
        *status = OSKAR_ERR_DIMENSION_MISMATCH;
        return;
    }

    /* Station models' (ll, mm, nn); more than one launch's worth are ORed
     * into the mask first and the fused clip then reads it. */
    const int num_models = oskar_telescope_num_station_models(telescope);
    oskar_HorizonModels fused;
    memset(&fused, 0, sizeof(fused));
    void* stream = oskar_amd_current_stream();
    int* d_mask = oskar_mem_int(mask, status);
    int* d_idx = oskar_mem_int(indices, status);
    if (num_models <= SKY_CLIP_MAX_MODELS && num_models > 0)
    {
        fused.num = num_models;
        for (i = 0; i < num_models; ++i)
        {
            const oskar_Station* s = oskar_telescope_station_const(
                    telescope, i);
            horizon_factors((gast_rad + oskar_station_lon_rad(s)) - ra0,
                    dec0, oskar_station_lat_rad(s), fused.lmn[i]);
        }
    }
    else
    {
        oskar_mem_clear_contents(mask, status);
        for (i = 0; i < num_models && !*status; i += SKY_CLIP_MAX_MODELS)
        {
            int k = 0;
            oskar_HorizonModels part;
            memset(&part, 0, sizeof(part));
            part.num = num_models - i < SKY_CLIP_MAX_MODELS ?
                    num_models - i : SKY_CLIP_MAX_MODELS;
            for (k = 0; k < part.num; ++k)
            {
                const oskar_Station* s = oskar_telescope_station_const(
                        telescope, i + k);
                horizon_factors((gast_rad + oskar_station_lon_rad(s)) - ra0,
                        dec0, oskar_station_lat_rad(s), part.lmn[k]);
            }
            const int e = oskar_amd_horizon_mask_launch(type, num_in,
                    oskar_mem_void_const(l), oskar_mem_void_const(m),
                    oskar_mem_void_const(n), &part, d_mask, stream);
            if (e && !*status) *status = e;
        }
    }
    if (*status) return;
    {
        const int e = oskar_amd_horizon_clip_launch(type, num_in,
                oskar_mem_void_const(l), oskar_mem_void_const(m),
                oskar_mem_void_const(n), &fused, d_mask, d_idx,
                in->attr_int[OSKAR_SKY_CAPACITY],
                out->attr_int[OSKAR_SKY_CAPACITY],
                in->attr_int[OSKAR_SKY_NUM_COLUMNS],
                oskar_mem_void_const(in->table), oskar_mem_void(out->table),
                oskar_mem_int_const(in->num_valid_columns, status),
                oskar_mem_int(out->num_valid_columns, status), stream);
        if (e && !*status) *status = e;
    }
    if (*status) return;
    copy_metadata(in, out, read_total(d_idx, num_in, status));
}
