/*
 * Visibility blocks, headers and their .vis chunks.
 *
 * Reference: oskar/vis/src/oskar_vis_block_create.c, _create_from_header.c,
 * _resize.c, _clear.c, _free.c, _accessors.c, _write.c, _read.c,
 * _station_to_baseline_coords.c; oskar/vis/src/oskar_vis_header_create.c,
 * _free.c, _accessors.c, _write.c, _read.c;
 * oskar/mem/src/oskar_binary_write_mem.c, oskar_binary_read_mem.c;
 * oskar/convert/src/oskar_convert_station_uvw_to_baseline_uvw.c.
 *
 * The chunk sequence of a header and of a block, their groups, tags, user
 * indices and data types follow the reference writer chunk for chunk, so a
 * file written here reads with the reference reader and the reverse.  The
 * station-to-baseline conversion runs as a HIP kernel; a host block is
 * staged through device memory for it.
 */
#include <hip/hip_runtime_api.h>
#include <pthread.h>

#include "private.h"

#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

#define OSKAR_AMD_VERSION_STR "oskar_amd 1.0"

/* ---- oskar_Mem chunks ----------------------------------------------------
 * Device memory streams through two pinned staging buffers: the copy of
 * piece k + 1 (device to host) runs while piece k is checksummed and
 * written, and on reading, piece k + 1 is read and checksummed while piece
 * k is copied to the device.  The buffers (2 x STAGE_BYTES of pinned host
 * memory) are allocated on first use and kept for the life of the process;
 * a mutex serialises the calls that use them. */
enum { STAGE_BYTES = 8 << 20 };
static pthread_mutex_t stage_lock = PTHREAD_MUTEX_INITIALIZER;
static void* stage_buf[2];

static int stage_acquire(hipEvent_t ev[2])
{
    int k;
    hipError_t e = hipSuccess;
    pthread_mutex_lock(&stage_lock);
    for (k = 0; k < 2 && e == hipSuccess; ++k)
        if (!stage_buf[k])
            e = hipHostMalloc(&stage_buf[k], STAGE_BYTES,
                    hipHostMallocPortable);
    ev[0] = ev[1] = 0;
    for (k = 0; k < 2 && e == hipSuccess; ++k)
        e = hipEventCreateWithFlags(&ev[k], hipEventDisableTiming);
    if (e != hipSuccess)
    {
        for (k = 0; k < 2; ++k) if (ev[k]) (void) hipEventDestroy(ev[k]);
        pthread_mutex_unlock(&stage_lock);
    }
    return (int) e;
}

static void stage_release(hipEvent_t ev[2])
{
    (void) hipEventDestroy(ev[0]);
    (void) hipEventDestroy(ev[1]);
    pthread_mutex_unlock(&stage_lock);
}

static size_t min_size(size_t a, size_t b) { return a < b ? a : b; }

/* The CRC of a piece runs on a helper thread while the main thread moves
 * the same (write) or the next (read) piece through the file. */
typedef struct
{
    pthread_t thread;
    int running;
    const void* data;
    size_t n;
    unsigned int crc;
} CrcJob;

static void* crc_job_run(void* arg)
{
    CrcJob* j = (CrcJob*) arg;
    j->crc = oskar_binary_crc32c(j->crc, j->data, j->n);
    return 0;
}

static void crc_job_start(CrcJob* j, const void* data, size_t n)
{
    j->data = data;
    j->n = n;
    j->running = pthread_create(&j->thread, 0, crc_job_run, j) == 0;
    if (!j->running) (void) crc_job_run(j);
}

static void crc_job_join(CrcJob* j)
{
    if (j->running) (void) pthread_join(j->thread, 0);
    j->running = 0;
}

static void write_device(oskar_Binary* h, unsigned char type,
        unsigned char id_group, unsigned char id_tag, int user_index,
        const char* src, size_t bytes, int* status)
{
    hipEvent_t ev[2];
    unsigned int crc = 0;
    size_t off = 0;
    int k = 0;
    CrcJob job;
    const hipStream_t s = (hipStream_t) oskar_amd_current_stream();
    const int e0 = stage_acquire(ev);
    memset(&job, 0, sizeof(job));
    if (e0)
    {
        *status = oskar_amd_hip_status(e0);
        return;
    }
    oskar_amd_binary_write_begin(h, type, id_group, id_tag, user_index,
            bytes, &crc, status);
    if (!*status && bytes)
    {
        hipError_t e = hipMemcpyAsync(stage_buf[0], src,
                min_size(STAGE_BYTES, bytes), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipEventRecord(ev[0], s);
        if (e != hipSuccess) *status = oskar_amd_hip_status((int) e);
    }
    while (!*status && off < bytes)
    {
        const size_t n = min_size(STAGE_BYTES, bytes - off);
        const size_t next = off + n;
        hipError_t e = hipSuccess;
        if (next < bytes)
        {
            e = hipMemcpyAsync(stage_buf[k ^ 1], src + next,
                    min_size(STAGE_BYTES, bytes - next),
                    hipMemcpyDeviceToHost, s);
            if (e == hipSuccess) e = hipEventRecord(ev[k ^ 1], s);
        }
        if (e == hipSuccess) e = hipEventSynchronize(ev[k]);
        if (e != hipSuccess)
        {
            *status = oskar_amd_hip_status((int) e);
            break;
        }
        job.crc = crc;
        crc_job_start(&job, stage_buf[k], n);
        oskar_amd_binary_write_raw(h, stage_buf[k], n, status);
        crc_job_join(&job);
        crc = job.crc;
        off = next;
        k ^= 1;
    }
    /* No copy may still target the staging buffers when they are handed
     * on. */
    (void) hipStreamSynchronize(s);
    oskar_amd_binary_write_end(h, crc, status);
    stage_release(ev);
}

static void read_device(oskar_Binary* h, int chunk, char* dst, size_t bytes,
        int* status)
{
    hipEvent_t ev[2];
    unsigned int crc = 0;
    size_t off = 0;
    int k = 0, used[2] = {0, 0};
    CrcJob job;
    const hipStream_t s = (hipStream_t) oskar_amd_current_stream();
    const int e0 = stage_acquire(ev);
    memset(&job, 0, sizeof(job));
    if (e0)
    {
        *status = oskar_amd_hip_status(e0);
        return;
    }
    if (oskar_amd_binary_read_begin(h, chunk, &crc, status) != bytes &&
            !*status)
        *status = OSKAR_ERR_BINARY_FORMAT_BAD;
    while (!*status && off < bytes)
    {
        const size_t n = min_size(STAGE_BYTES, bytes - off);
        hipError_t e = hipSuccess;
        if (used[k]) e = hipEventSynchronize(ev[k]);
        if (e != hipSuccess)
        {
            *status = oskar_amd_hip_status((int) e);
            break;
        }
        oskar_amd_binary_read_raw(h, stage_buf[k], n, status);
        /* The previous piece's CRC ran meanwhile; this piece's starts
         * from it. */
        crc_job_join(&job);
        if (*status) break;
        job.crc = off ? job.crc : crc;
        crc_job_start(&job, stage_buf[k], n);
        e = hipMemcpyAsync(dst + off, stage_buf[k], n, hipMemcpyHostToDevice,
                s);
        if (e == hipSuccess) e = hipEventRecord(ev[k], s);
        if (e != hipSuccess)
        {
            *status = oskar_amd_hip_status((int) e);
            break;
        }
        used[k] = 1;
        off += n;
        k ^= 1;
    }
    crc_job_join(&job);
    if (off) crc = job.crc;
    {
        const hipError_t e = hipStreamSynchronize(s);
        if (e != hipSuccess && !*status)
            *status = oskar_amd_hip_status((int) e);
    }
    oskar_amd_binary_read_end(h, chunk, crc, status);
    stage_release(ev);
}

void oskar_binary_write_mem(oskar_Binary* h, const oskar_Mem* mem,
        unsigned char id_group, unsigned char id_tag, int user_index,
        size_t num_to_write, int* status)
{
    if (*status) return;
    const int type = oskar_mem_type(mem);
    if (num_to_write == 0) num_to_write = oskar_mem_length(mem);
    if (num_to_write > oskar_mem_length(mem))
    {
        *status = OSKAR_ERR_OUT_OF_RANGE;
        return;
    }
    const size_t bytes = num_to_write * oskar_mem_element_size(type);
    if (oskar_mem_location(mem) == OSKAR_GPU)
        write_device(h, (unsigned char) type, id_group, id_tag, user_index,
                (const char*) oskar_mem_void_const(mem), bytes, status);
    else if (oskar_mem_location(mem) == OSKAR_CPU)
        oskar_binary_write(h, (unsigned char) type, id_group, id_tag,
                user_index, bytes, oskar_mem_void_const(mem), status);
    else *status = OSKAR_ERR_BAD_LOCATION;
}

void oskar_binary_read_mem(oskar_Binary* h, oskar_Mem* mem,
        unsigned char id_group, unsigned char id_tag, int user_index,
        int* status)
{
    size_t bytes = 0;
    if (*status) return;
    const int type = oskar_mem_type(mem);
    const size_t es = oskar_mem_element_size(type);
    const int chunk = oskar_binary_query(h, (unsigned char) type, id_group,
            id_tag, user_index, &bytes, status);
    if (*status) return;
    if (bytes % es)
    {
        *status = OSKAR_ERR_BINARY_FORMAT_BAD;
        return;
    }
    oskar_mem_realloc(mem, bytes / es, status);
    if (*status) return;
    if (oskar_mem_location(mem) == OSKAR_GPU)
        read_device(h, chunk, (char*) oskar_mem_void(mem), bytes, status);
    else if (oskar_mem_location(mem) == OSKAR_CPU)
        oskar_binary_read_block(h, chunk, bytes, oskar_mem_void(mem), status);
    else *status = OSKAR_ERR_BAD_LOCATION;
}

/* ---- station to baseline coordinates ------------------------------------ */
static int real_type(int t) { return t == OSKAR_SINGLE || t == OSKAR_DOUBLE; }

static void station_to_baseline(int casa, int num_times, int num_stations,
        int offset_in, const oskar_Mem* u, const oskar_Mem* v,
        const oskar_Mem* w, int offset_out, oskar_Mem* uu, oskar_Mem* vv,
        oskar_Mem* ww, int* status)
{
    int i;
    const oskar_Mem* in[3] = {u, v, w};
    oskar_Mem* out[3] = {uu, vv, ww};
    if (*status) return;
    const int type = oskar_mem_type(u);
    if (!real_type(type))
    {
        *status = OSKAR_ERR_BAD_DATA_TYPE;
        return;
    }
    if (oskar_mem_location(u) != OSKAR_GPU)
    {
        *status = OSKAR_ERR_BAD_LOCATION;
        return;
    }
    if (num_stations < 0 || num_times < 0 || offset_in < 0 || offset_out < 0)
    {
        *status = OSKAR_ERR_INVALID_ARGUMENT;
        return;
    }
    const size_t nb = (size_t) num_stations * (num_stations - 1) / 2;
    const size_t need_in = (size_t) offset_in + (size_t) num_times *
            num_stations;
    const size_t need_out = (size_t) offset_out + (size_t) num_times * nb;
    for (i = 0; i < 3; ++i)
    {
        if (oskar_mem_type(in[i]) != type || oskar_mem_type(out[i]) != type)
        {
            *status = OSKAR_ERR_TYPE_MISMATCH;
            return;
        }
        if (oskar_mem_location(in[i]) != OSKAR_GPU ||
                oskar_mem_location(out[i]) != OSKAR_GPU)
        {
            *status = OSKAR_ERR_LOCATION_MISMATCH;
            return;
        }
        if ((num_times > 0 && num_stations > 1) &&
                (oskar_mem_length(in[i]) < need_in ||
                 oskar_mem_length(out[i]) < need_out))
        {
            *status = OSKAR_ERR_DIMENSION_MISMATCH;
            return;
        }
    }
    if (need_in > 0x7FFFFFFF || need_out > 0x7FFFFFFF)
    {
        *status = OSKAR_ERR_OUT_OF_RANGE;
        return;
    }
    *status = oskar_amd_hip_status(oskar_amd_station_to_baseline_launch(type,
            casa, num_times, num_stations, offset_in,
            oskar_mem_void_const(u), oskar_mem_void_const(v),
            oskar_mem_void_const(w), offset_out, oskar_mem_void(uu),
            oskar_mem_void(vv), oskar_mem_void(ww),
            oskar_amd_current_stream()));
}

void oskar_convert_station_uvw_to_baseline_uvw(int use_casa_phase_convention,
        int num_stations, int offset_in, const oskar_Mem* u,
        const oskar_Mem* v, const oskar_Mem* w, int offset_out,
        oskar_Mem* uu, oskar_Mem* vv, oskar_Mem* ww, int* status)
{
    station_to_baseline(use_casa_phase_convention, 1, num_stations,
            offset_in, u, v, w, offset_out, uu, vv, ww, status);
}

/* ---- blocks --------------------------------------------------------------- */
oskar_VisBlock* oskar_vis_block_create(int location, int amp_type,
        int num_times, int num_channels, int num_stations,
        int create_crosscorr, int create_autocorr, int* status)
{
    int i;
    oskar_VisBlock* vis = 0;
    if (*status) return 0;
    const int prec = oskar_type_precision(amp_type);
    if (!oskar_type_is_complex(amp_type) ||
            (prec != OSKAR_SINGLE && prec != OSKAR_DOUBLE))
    {
        *status = OSKAR_ERR_BAD_DATA_TYPE;
        return 0;
    }
    if (location != OSKAR_CPU && location != OSKAR_GPU)
    {
        *status = OSKAR_ERR_BAD_LOCATION;
        return 0;
    }
    vis = (oskar_VisBlock*) calloc(1, sizeof(oskar_VisBlock));
    if (!vis)
    {
        *status = OSKAR_ERR_MEMORY_ALLOC_FAILURE;
        return 0;
    }
    vis->has_cross_correlations = create_crosscorr;
    vis->has_auto_correlations = create_autocorr;
    for (i = 0; i < 3; ++i)
    {
        vis->baseline_uvw_metres[i] = oskar_mem_create(prec, location, 0,
                status);
        vis->station_uvw_metres[i] = oskar_mem_create(prec, location, 0,
                status);
    }
    vis->cross_correlations = oskar_mem_create(amp_type, location, 0, status);
    vis->auto_correlations = oskar_mem_create(amp_type, location, 0, status);
    oskar_vis_block_resize(vis, num_times, num_channels, num_stations,
            status);
    return vis;
}

oskar_VisBlock* oskar_vis_block_create_from_header(int location,
        const oskar_VisHeader* hdr, int* status)
{
    if (*status) return 0;
    return oskar_vis_block_create(location, hdr->amp_type,
            hdr->max_times_per_block, hdr->max_channels_per_block,
            hdr->num_stations, hdr->write_crosscorr, hdr->write_autocorr,
            status);
}

void oskar_vis_block_free(oskar_VisBlock* vis, int* status)
{
    int i;
    if (!vis) return;
    for (i = 0; i < 3; ++i)
    {
        oskar_mem_free(vis->baseline_uvw_metres[i], status);
        oskar_mem_free(vis->station_uvw_metres[i], status);
    }
    oskar_mem_free(vis->cross_correlations, status);
    oskar_mem_free(vis->auto_correlations, status);
    free(vis);
}

void oskar_vis_block_resize(oskar_VisBlock* vis, int num_times,
        int num_channels, int num_stations, int* status)
{
    int i;
    if (*status) return;
    if (num_times < 0 || num_channels < 0 || num_stations < 0)
    {
        *status = OSKAR_ERR_INVALID_ARGUMENT;
        return;
    }
    const int num_baselines = vis->has_cross_correlations ?
            num_stations * (num_stations - 1) / 2 : 0;
    vis->dim_start_size[2] = num_times;
    vis->dim_start_size[3] = num_channels;
    vis->dim_start_size[4] = num_baselines;
    vis->dim_start_size[5] = num_stations;
    for (i = 0; i < 3; ++i)
    {
        if (oskar_mem_length(vis->baseline_uvw_metres[i]) > 0)
            oskar_mem_realloc(vis->baseline_uvw_metres[i],
                    (size_t) num_times * num_baselines, status);
        if (oskar_mem_length(vis->station_uvw_metres[i]) > 0)
            oskar_mem_realloc(vis->station_uvw_metres[i],
                    (size_t) num_times * num_stations, status);
    }
    oskar_mem_realloc(vis->cross_correlations,
            (size_t) num_times * num_channels * num_baselines, status);
    oskar_mem_realloc(vis->auto_correlations, vis->has_auto_correlations ?
            (size_t) num_times * num_channels * num_stations : 0, status);
}

void oskar_vis_block_clear(oskar_VisBlock* vis, int* status)
{
    int i;
    if (*status) return;
    oskar_mem_clear_contents(vis->cross_correlations, status);
    oskar_mem_clear_contents(vis->auto_correlations, status);
    for (i = 0; i < 3; ++i)
    {
        oskar_mem_clear_contents(vis->baseline_uvw_metres[i], status);
        oskar_mem_clear_contents(vis->station_uvw_metres[i], status);
    }
}

int oskar_vis_block_num_times(const oskar_VisBlock* v)
{
    return v->dim_start_size[2];
}
int oskar_vis_block_num_channels(const oskar_VisBlock* v)
{
    return v->dim_start_size[3];
}
int oskar_vis_block_num_baselines(const oskar_VisBlock* v)
{
    return v->dim_start_size[4];
}
int oskar_vis_block_num_stations(const oskar_VisBlock* v)
{
    return v->dim_start_size[5];
}
int oskar_vis_block_num_pols(const oskar_VisBlock* v)
{
    return oskar_mem_is_matrix(v->cross_correlations) ? 4 : 1;
}
int oskar_vis_block_start_time_index(const oskar_VisBlock* v)
{
    return v->dim_start_size[0];
}
int oskar_vis_block_start_channel_index(const oskar_VisBlock* v)
{
    return v->dim_start_size[1];
}
void oskar_vis_block_set_start_time_index(oskar_VisBlock* v, int index)
{
    v->dim_start_size[0] = index;
}
void oskar_vis_block_set_start_channel_index(oskar_VisBlock* v, int index)
{
    v->dim_start_size[1] = index;
}
int oskar_vis_block_has_auto_correlations(const oskar_VisBlock* v)
{
    return v->has_auto_correlations && oskar_mem_length(v->auto_correlations);
}
int oskar_vis_block_has_cross_correlations(const oskar_VisBlock* v)
{
    return v->has_cross_correlations &&
            oskar_mem_length(v->cross_correlations);
}
oskar_Mem* oskar_vis_block_cross_correlations(oskar_VisBlock* v)
{
    return v->cross_correlations;
}
oskar_Mem* oskar_vis_block_auto_correlations(oskar_VisBlock* v)
{
    return v->auto_correlations;
}
oskar_Mem* oskar_vis_block_baseline_uvw_metres(oskar_VisBlock* v, int dim)
{
    return (dim >= 0 && dim < 3) ? v->baseline_uvw_metres[dim] : 0;
}
oskar_Mem* oskar_vis_block_station_uvw_metres(oskar_VisBlock* v, int dim)
{
    return (dim >= 0 && dim < 3) ? v->station_uvw_metres[dim] : 0;
}

void oskar_vis_block_station_to_baseline_coords(oskar_VisBlock* vis,
        const oskar_VisHeader* hdr, int* status)
{
    int i;
    oskar_Mem *in[3], *out[3];
    if (*status) return;
    const int nt = vis->dim_start_size[2], ns = vis->dim_start_size[5];
    const size_t nb = (size_t) ns * (ns - 1) / 2;
    const int type = oskar_mem_type(vis->station_uvw_metres[0]);
    const int on_host = oskar_mem_location(vis->station_uvw_metres[0]) ==
            OSKAR_CPU;
    for (i = 0; i < 3; ++i)
    {
        oskar_mem_ensure(vis->baseline_uvw_metres[i], (size_t) nt * nb,
                status);
        in[i] = vis->station_uvw_metres[i];
        out[i] = vis->baseline_uvw_metres[i];
    }
    if (on_host)
    {
        /* Host blocks go through device copies: the conversion is a HIP
         * kernel only. */
        for (i = 0; i < 3; ++i)
        {
            in[i] = oskar_mem_create_copy(vis->station_uvw_metres[i],
                    OSKAR_GPU, status);
            out[i] = oskar_mem_create(type, OSKAR_GPU, (size_t) nt * nb,
                    status);
        }
    }
    station_to_baseline(hdr ? hdr->casa_phase_convention : 0, nt, ns, 0,
            in[0], in[1], in[2], 0, out[0], out[1], out[2], status);
    if (!on_host) return;
    for (i = 0; i < 3; ++i)
    {
        oskar_mem_copy_contents(vis->baseline_uvw_metres[i], out[i], 0, 0,
                (size_t) nt * nb, status);
        oskar_mem_free(in[i], status);
        oskar_mem_free(out[i], status);
    }
}

void oskar_vis_block_write(const oskar_VisBlock* vis, oskar_Binary* h,
        int block_index, int* status)
{
    const unsigned char grp = OSKAR_TAG_GROUP_VIS_BLOCK;
    if (*status) return;
    oskar_binary_write(h, OSKAR_INT, grp,
            OSKAR_VIS_BLOCK_TAG_DIM_START_AND_SIZE, block_index,
            6 * sizeof(int), vis->dim_start_size, status);
    if (oskar_vis_block_has_auto_correlations(vis))
        oskar_binary_write_mem(h, vis->auto_correlations, grp,
                OSKAR_VIS_BLOCK_TAG_AUTO_CORRELATIONS, block_index, 0,
                status);
    if (!oskar_vis_block_has_cross_correlations(vis)) return;
    oskar_binary_write_mem(h, vis->cross_correlations, grp,
            OSKAR_VIS_BLOCK_TAG_CROSS_CORRELATIONS, block_index, 0, status);
    if (oskar_mem_length(vis->station_uvw_metres[0]) > 0)
    {
        oskar_binary_write_mem(h, vis->station_uvw_metres[0], grp,
                OSKAR_VIS_BLOCK_TAG_STATION_U, block_index, 0, status);
        oskar_binary_write_mem(h, vis->station_uvw_metres[1], grp,
                OSKAR_VIS_BLOCK_TAG_STATION_V, block_index, 0, status);
        oskar_binary_write_mem(h, vis->station_uvw_metres[2], grp,
                OSKAR_VIS_BLOCK_TAG_STATION_W, block_index, 0, status);
    }
    else if (oskar_mem_length(vis->baseline_uvw_metres[0]) > 0)
    {
        oskar_binary_write_mem(h, vis->baseline_uvw_metres[0], grp,
                OSKAR_VIS_BLOCK_TAG_BASELINE_UU, block_index, 0, status);
        oskar_binary_write_mem(h, vis->baseline_uvw_metres[1], grp,
                OSKAR_VIS_BLOCK_TAG_BASELINE_VV, block_index, 0, status);
        oskar_binary_write_mem(h, vis->baseline_uvw_metres[2], grp,
                OSKAR_VIS_BLOCK_TAG_BASELINE_WW, block_index, 0, status);
    }
}

void oskar_vis_block_read(oskar_VisBlock* vis, const oskar_VisHeader* hdr,
        oskar_Binary* h, int block_index, int* status)
{
    int tag_error = 0;
    const unsigned char grp = OSKAR_TAG_GROUP_VIS_BLOCK;
    if (*status) return;
    /* Chunks of block i start no earlier than this (metadata chunks the
     * header does not count only move them later). */
    oskar_binary_set_query_search_start(h, hdr->num_tags_header +
            block_index * hdr->num_tags_per_block, status);
    oskar_binary_read(h, OSKAR_INT, grp,
            OSKAR_VIS_BLOCK_TAG_DIM_START_AND_SIZE, block_index,
            6 * sizeof(int), vis->dim_start_size, status);
    if (hdr->write_autocorr)
        oskar_binary_read_mem(h, vis->auto_correlations, grp,
                OSKAR_VIS_BLOCK_TAG_AUTO_CORRELATIONS, block_index, status);
    if (!hdr->write_crosscorr) return;
    oskar_binary_read_mem(h, vis->cross_correlations, grp,
            OSKAR_VIS_BLOCK_TAG_CROSS_CORRELATIONS, block_index, status);
    /* Station coordinates (baselines follow from them), else the baseline
     * coordinates of older files. */
    oskar_binary_read_mem(h, vis->station_uvw_metres[0], grp,
            OSKAR_VIS_BLOCK_TAG_STATION_U, block_index, &tag_error);
    if (!tag_error)
    {
        oskar_binary_read_mem(h, vis->station_uvw_metres[1], grp,
                OSKAR_VIS_BLOCK_TAG_STATION_V, block_index, status);
        oskar_binary_read_mem(h, vis->station_uvw_metres[2], grp,
                OSKAR_VIS_BLOCK_TAG_STATION_W, block_index, status);
        oskar_vis_block_station_to_baseline_coords(vis, hdr, status);
    }
    else
    {
        oskar_binary_read_mem(h, vis->baseline_uvw_metres[0], grp,
                OSKAR_VIS_BLOCK_TAG_BASELINE_UU, block_index, status);
        oskar_binary_read_mem(h, vis->baseline_uvw_metres[1], grp,
                OSKAR_VIS_BLOCK_TAG_BASELINE_VV, block_index, status);
        oskar_binary_read_mem(h, vis->baseline_uvw_metres[2], grp,
                OSKAR_VIS_BLOCK_TAG_BASELINE_WW, block_index, status);
    }
}

/* ---- headers -------------------------------------------------------------- */
static int pol_type_allowed(int amp_type, int value)
{
    if (oskar_type_is_matrix(amp_type))
        return value == OSKAR_VIS_POL_TYPE_STOKES_I_Q_U_V ||
                value == OSKAR_VIS_POL_TYPE_LINEAR_XX_XY_YX_YY;
    return (value >= OSKAR_VIS_POL_TYPE_STOKES_I &&
            value <= OSKAR_VIS_POL_TYPE_STOKES_V) ||
            (value >= OSKAR_VIS_POL_TYPE_LINEAR_XX &&
             value <= OSKAR_VIS_POL_TYPE_LINEAR_YY);
}

static oskar_Mem** mem_list(int type, int n, size_t len, int* status)
{
    int i;
    oskar_Mem** list = (oskar_Mem**) calloc(n > 0 ? n : 1,
            sizeof(oskar_Mem*));
    if (!list)
    {
        *status = OSKAR_ERR_MEMORY_ALLOC_FAILURE;
        return 0;
    }
    for (i = 0; i < n; ++i)
        list[i] = oskar_mem_create(type, OSKAR_CPU, len, status);
    return list;
}

static void free_list(oskar_Mem** list, int n, int* status)
{
    int i;
    if (!list) return;
    for (i = 0; i < n; ++i) oskar_mem_free(list[i], status);
    free(list);
}

oskar_VisHeader* oskar_vis_header_create(int amp_type, int coord_precision,
        int max_times_per_block, int num_times_total,
        int max_channels_per_block, int num_channels_total, int num_stations,
        int write_autocorr, int write_crosscorr, int* status)
{
    int i, j;
    oskar_VisHeader* hdr = 0;
    if (*status) return 0;
    if (!oskar_type_is_complex(amp_type) ||
            (coord_precision != OSKAR_SINGLE &&
             coord_precision != OSKAR_DOUBLE))
    {
        *status = OSKAR_ERR_BAD_DATA_TYPE;
        return 0;
    }
    if (num_stations < 0)
    {
        *status = OSKAR_ERR_INVALID_ARGUMENT;
        return 0;
    }
    hdr = (oskar_VisHeader*) calloc(1, sizeof(oskar_VisHeader));
    if (!hdr)
    {
        *status = OSKAR_ERR_MEMORY_ALLOC_FAILURE;
        return 0;
    }
    hdr->amp_type = amp_type;
    hdr->coord_precision = coord_precision;
    /* One dims chunk, then the cross amplitudes and three coordinate
     * chunks, then the auto amplitudes: the block writer's sequence. */
    hdr->num_tags_per_block = 1 + (write_crosscorr ? 4 : 0) +
            (write_autocorr ? 1 : 0);
    hdr->max_times_per_block = max_times_per_block > 0 ?
            max_times_per_block : num_times_total;
    hdr->max_channels_per_block = max_channels_per_block > 0 ?
            max_channels_per_block : num_channels_total;
    hdr->num_times_total = num_times_total;
    hdr->num_channels_total = num_channels_total;
    hdr->num_stations = num_stations;
    hdr->write_autocorr = write_autocorr;
    hdr->write_crosscorr = write_crosscorr;
    hdr->pol_type = oskar_type_is_matrix(amp_type) ?
            OSKAR_VIS_POL_TYPE_LINEAR_XX_XY_YX_YY :
            OSKAR_VIS_POL_TYPE_STOKES_I;
    hdr->casa_phase_convention = 0;
    hdr->telescope_path = oskar_mem_create(OSKAR_CHAR, OSKAR_CPU, 0, status);
    hdr->settings = oskar_mem_create(OSKAR_CHAR, OSKAR_CPU, 0, status);
    hdr->station_diameter_m = oskar_mem_create(OSKAR_DOUBLE, OSKAR_CPU,
            num_stations, status);
    hdr->station_name = mem_list(OSKAR_CHAR, num_stations, 1, status);
    for (i = 0; i < 3; ++i)
    {
        hdr->station_offset_ecef_metres[i] = oskar_mem_create(
                coord_precision, OSKAR_CPU, num_stations, status);
        hdr->element_enu_metres[i] = mem_list(coord_precision, num_stations,
                0, status);
    }
    for (i = 0; i < 2; ++i)
        for (j = 0; j < 3; ++j)
            hdr->element_euler_angle_rad[i][j] = mem_list(OSKAR_DOUBLE,
                    num_stations, 0, status);
    return hdr;
}

void oskar_vis_header_free(oskar_VisHeader* hdr, int* status)
{
    int i, j;
    if (!hdr) return;
    oskar_mem_free(hdr->telescope_path, status);
    oskar_mem_free(hdr->settings, status);
    oskar_mem_free(hdr->station_diameter_m, status);
    free_list(hdr->station_name, hdr->num_stations, status);
    for (i = 0; i < 3; ++i)
    {
        oskar_mem_free(hdr->station_offset_ecef_metres[i], status);
        free_list(hdr->element_enu_metres[i], hdr->num_stations, status);
    }
    for (i = 0; i < 2; ++i)
        for (j = 0; j < 3; ++j)
            free_list(hdr->element_euler_angle_rad[i][j], hdr->num_stations,
                    status);
    free(hdr);
}

int oskar_vis_header_amp_type(const oskar_VisHeader* h) { return h->amp_type; }
int oskar_vis_header_coord_precision(const oskar_VisHeader* h)
{
    return h->coord_precision;
}
int oskar_vis_header_num_stations(const oskar_VisHeader* h)
{
    return h->num_stations;
}
int oskar_vis_header_num_tags_per_block(const oskar_VisHeader* h)
{
    return h->num_tags_per_block;
}
int oskar_vis_header_num_tags_header(const oskar_VisHeader* h)
{
    return h->num_tags_header;
}
int oskar_vis_header_max_times_per_block(const oskar_VisHeader* h)
{
    return h->max_times_per_block;
}
int oskar_vis_header_max_channels_per_block(const oskar_VisHeader* h)
{
    return h->max_channels_per_block;
}
int oskar_vis_header_num_times_total(const oskar_VisHeader* h)
{
    return h->num_times_total;
}
int oskar_vis_header_num_channels_total(const oskar_VisHeader* h)
{
    return h->num_channels_total;
}
int oskar_vis_header_write_auto_correlations(const oskar_VisHeader* h)
{
    return h->write_autocorr;
}
int oskar_vis_header_write_cross_correlations(const oskar_VisHeader* h)
{
    return h->write_crosscorr;
}
int oskar_vis_header_pol_type(const oskar_VisHeader* h) { return h->pol_type; }
int oskar_vis_header_casa_phase_convention(const oskar_VisHeader* h)
{
    return h->casa_phase_convention;
}
double oskar_vis_header_freq_start_hz(const oskar_VisHeader* h)
{
    return h->freq_start_hz;
}
double oskar_vis_header_freq_inc_hz(const oskar_VisHeader* h)
{
    return h->freq_inc_hz;
}
double oskar_vis_header_channel_bandwidth_hz(const oskar_VisHeader* h)
{
    return h->channel_bandwidth_hz;
}
double oskar_vis_header_time_start_mjd_utc(const oskar_VisHeader* h)
{
    return h->time_start_mjd_utc;
}
double oskar_vis_header_time_inc_sec(const oskar_VisHeader* h)
{
    return h->time_inc_sec;
}
double oskar_vis_header_time_average_sec(const oskar_VisHeader* h)
{
    return h->time_average_sec;
}
int oskar_vis_header_phase_centre_type(const oskar_VisHeader* h)
{
    return h->phase_centre_type;
}
double oskar_vis_header_phase_centre_ra_deg(const oskar_VisHeader* h)
{
    return h->phase_centre_deg[0];
}
double oskar_vis_header_phase_centre_dec_deg(const oskar_VisHeader* h)
{
    return h->phase_centre_deg[1];
}
double oskar_vis_header_telescope_lon_deg(const oskar_VisHeader* h)
{
    return h->telescope_centre_lon_deg;
}
double oskar_vis_header_telescope_lat_deg(const oskar_VisHeader* h)
{
    return h->telescope_centre_lat_deg;
}
double oskar_vis_header_telescope_alt_metres(const oskar_VisHeader* h)
{
    return h->telescope_centre_alt_m;
}
oskar_Mem* oskar_vis_header_station_offset_ecef_metres(oskar_VisHeader* h,
        int dim)
{
    return (dim >= 0 && dim < 3) ? h->station_offset_ecef_metres[dim] : 0;
}
oskar_Mem* oskar_vis_header_telescope_path(oskar_VisHeader* h)
{
    return h->telescope_path;
}
oskar_Mem* oskar_vis_header_settings(oskar_VisHeader* h)
{
    return h->settings;
}
oskar_Mem* oskar_vis_header_station_diameter_m(oskar_VisHeader* h)
{
    return h->station_diameter_m;
}
oskar_Mem* oskar_vis_header_station_name(oskar_VisHeader* h, int station)
{
    return (station >= 0 && station < h->num_stations) ?
            h->station_name[station] : 0;
}
oskar_Mem* oskar_vis_header_element_enu_metres(oskar_VisHeader* h, int dim,
        int station)
{
    return (dim >= 0 && dim < 3 && station >= 0 &&
            station < h->num_stations) ?
            h->element_enu_metres[dim][station] : 0;
}
oskar_Mem* oskar_vis_header_element_feed_angle_rad(oskar_VisHeader* h,
        int dim, int angle, int station)
{
    return (dim >= 0 && dim < 2 && angle >= 0 && angle < 3 && station >= 0 &&
            station < h->num_stations) ?
            h->element_euler_angle_rad[dim][angle][station] : 0;
}

void oskar_vis_header_set_pol_type(oskar_VisHeader* h, int value,
        int* status)
{
    if (*status) return;
    if (!pol_type_allowed(h->amp_type, value))
    {
        *status = OSKAR_ERR_TYPE_MISMATCH;
        return;
    }
    h->pol_type = value;
}
void oskar_vis_header_set_casa_phase_convention(oskar_VisHeader* h, int v)
{
    h->casa_phase_convention = v;
}
void oskar_vis_header_set_freq_start_hz(oskar_VisHeader* h, double v)
{
    h->freq_start_hz = v;
}
void oskar_vis_header_set_freq_inc_hz(oskar_VisHeader* h, double v)
{
    h->freq_inc_hz = v;
}
void oskar_vis_header_set_channel_bandwidth_hz(oskar_VisHeader* h, double v)
{
    h->channel_bandwidth_hz = v;
}
void oskar_vis_header_set_time_start_mjd_utc(oskar_VisHeader* h, double v)
{
    h->time_start_mjd_utc = v;
}
void oskar_vis_header_set_time_inc_sec(oskar_VisHeader* h, double v)
{
    h->time_inc_sec = v;
}
void oskar_vis_header_set_time_average_sec(oskar_VisHeader* h, double v)
{
    h->time_average_sec = v;
}
void oskar_vis_header_set_phase_centre(oskar_VisHeader* h, int coord_type,
        double lon_deg, double lat_deg)
{
    h->phase_centre_type = coord_type;
    h->phase_centre_deg[0] = lon_deg;
    h->phase_centre_deg[1] = lat_deg;
}
void oskar_vis_header_set_telescope_centre(oskar_VisHeader* h,
        double lon_deg, double lat_deg, double alt_metres)
{
    h->telescope_centre_lon_deg = lon_deg;
    h->telescope_centre_lat_deg = lat_deg;
    h->telescope_centre_alt_m = alt_metres;
}

static void write_string(oskar_Binary* h, unsigned char tag, const char* s,
        int* status)
{
    oskar_binary_write(h, OSKAR_CHAR, OSKAR_TAG_GROUP_METADATA, tag, 0,
            strlen(s) + 1, s, status);
}

oskar_Binary* oskar_vis_header_write(const oskar_VisHeader* hdr,
        const char* filename, int* status)
{
    int i, j, k;
    char text[4096];
    const unsigned char grp = OSKAR_TAG_GROUP_VIS_HEADER;
    const char* user;
    if (*status) return 0;
    oskar_Binary* h = oskar_binary_create(filename, 'w', status);
    if (*status)
    {
        oskar_binary_free(h);
        return 0;
    }

    /* Metadata: local date and time, library version, working directory,
     * user name. */
    {
        const time_t now = time(0);
        const struct tm* tm = localtime(&now);
        if (tm && strftime(text, sizeof(text), "%Y-%m-%d, %H:%M:%S (%Z)",
                tm) > 0)
            write_string(h, OSKAR_TAG_METADATA_DATE_TIME_STRING, text,
                    status);
    }
    write_string(h, OSKAR_TAG_METADATA_OSKAR_VERSION_STRING,
            OSKAR_AMD_VERSION_STR, status);
    if (getcwd(text, sizeof(text)))
        write_string(h, OSKAR_TAG_METADATA_CWD, text, status);
    user = getenv("USERNAME");
    if (!user) user = getenv("USER");
    if (user && user[0])
        write_string(h, OSKAR_TAG_METADATA_USERNAME, user, status);

    /* Settings text, when there is any. */
    {
        const size_t n = oskar_mem_length(hdr->settings);
        const char* s = oskar_mem_char_const(hdr->settings);
        if (n > 0 && s && s[0])
            oskar_binary_write_mem(h, hdr->settings, OSKAR_TAG_GROUP_SETTINGS,
                    OSKAR_TAG_SETTINGS, 0, 0, status);
    }

    oskar_binary_write_mem(h, hdr->telescope_path, grp,
            OSKAR_VIS_HEADER_TAG_TELESCOPE_PATH, 0, 0, status);
    {
        const struct { unsigned char tag; int value; } ints[] = {
            {OSKAR_VIS_HEADER_TAG_NUM_TAGS_PER_BLOCK, hdr->num_tags_per_block},
            {OSKAR_VIS_HEADER_TAG_WRITE_AUTO_CORRELATIONS,
                    hdr->write_autocorr},
            {OSKAR_VIS_HEADER_TAG_WRITE_CROSS_CORRELATIONS,
                    hdr->write_crosscorr},
            {OSKAR_VIS_HEADER_TAG_AMP_TYPE, hdr->amp_type},
            {OSKAR_VIS_HEADER_TAG_COORD_PRECISION, hdr->coord_precision},
            {OSKAR_VIS_HEADER_TAG_MAX_TIMES_PER_BLOCK,
                    hdr->max_times_per_block},
            {OSKAR_VIS_HEADER_TAG_NUM_TIMES_TOTAL, hdr->num_times_total},
            {OSKAR_VIS_HEADER_TAG_MAX_CHANNELS_PER_BLOCK,
                    hdr->max_channels_per_block},
            {OSKAR_VIS_HEADER_TAG_NUM_CHANNELS_TOTAL,
                    hdr->num_channels_total},
            {OSKAR_VIS_HEADER_TAG_NUM_STATIONS, hdr->num_stations},
            {OSKAR_VIS_HEADER_TAG_POL_TYPE, hdr->pol_type},
            {OSKAR_VIS_HEADER_TAG_CASA_PHASE_CONVENTION,
                    hdr->casa_phase_convention},
            {OSKAR_VIS_HEADER_TAG_PHASE_CENTRE_COORD_TYPE,
                    hdr->phase_centre_type}};
        for (i = 0; i < (int) (sizeof(ints) / sizeof(ints[0])); ++i)
            oskar_binary_write_int(h, grp, ints[i].tag, 0, ints[i].value,
                    status);
    }
    oskar_binary_write(h, OSKAR_DOUBLE, grp,
            OSKAR_VIS_HEADER_TAG_PHASE_CENTRE_DEG, 0, 2 * sizeof(double),
            hdr->phase_centre_deg, status);
    {
        const struct { unsigned char tag; double value; } reals[] = {
            {OSKAR_VIS_HEADER_TAG_FREQ_START_HZ, hdr->freq_start_hz},
            {OSKAR_VIS_HEADER_TAG_FREQ_INC_HZ, hdr->freq_inc_hz},
            {OSKAR_VIS_HEADER_TAG_CHANNEL_BANDWIDTH_HZ,
                    hdr->channel_bandwidth_hz},
            {OSKAR_VIS_HEADER_TAG_TIME_START_MJD_UTC,
                    hdr->time_start_mjd_utc},
            {OSKAR_VIS_HEADER_TAG_TIME_INC_SEC, hdr->time_inc_sec},
            {OSKAR_VIS_HEADER_TAG_TIME_AVERAGE_SEC, hdr->time_average_sec},
            {OSKAR_VIS_HEADER_TAG_TELESCOPE_REF_LON_DEG,
                    hdr->telescope_centre_lon_deg},
            {OSKAR_VIS_HEADER_TAG_TELESCOPE_REF_LAT_DEG,
                    hdr->telescope_centre_lat_deg},
            {OSKAR_VIS_HEADER_TAG_TELESCOPE_REF_ALT_M,
                    hdr->telescope_centre_alt_m}};
        for (i = 0; i < (int) (sizeof(reals) / sizeof(reals[0])); ++i)
            oskar_binary_write_double(h, grp, reals[i].tag, 0,
                    reals[i].value, status);
    }
    for (i = 0; i < 3; ++i)
        oskar_binary_write_mem(h, hdr->station_offset_ecef_metres[i], grp,
                (unsigned char) (OSKAR_VIS_HEADER_TAG_STATION_X_OFFSET_ECEF +
                        i), 0, 0, status);
    for (k = 0; k < hdr->num_stations; ++k)
        for (i = 0; i < 3; ++i)
            oskar_binary_write_mem(h, hdr->element_enu_metres[i][k], grp,
                    (unsigned char) (OSKAR_VIS_HEADER_TAG_ELEMENT_X_ENU + i),
                    k, 0, status);
    for (k = 0; k < hdr->num_stations; ++k)
        oskar_binary_write_mem(h, hdr->station_name[k], grp,
                OSKAR_VIS_HEADER_TAG_STATION_NAME, k, 0, status);
    oskar_binary_write_mem(h, hdr->station_diameter_m, grp,
            OSKAR_VIS_HEADER_TAG_STATION_DIAMETER, 0, 0, status);
    /* Feed angles: X a, Y a, X b, Y b, X c, Y c per station. */
    for (k = 0; k < hdr->num_stations; ++k)
        for (j = 0; j < 3; ++j)
            for (i = 0; i < 2; ++i)
                oskar_binary_write_mem(h,
                        hdr->element_euler_angle_rad[i][j][k], grp,
                        (unsigned char) (
                                OSKAR_VIS_HEADER_TAG_ELEMENT_FEED_ANGLE_X_A +
                                2 * j + i), k, 0, status);
    return h;
}

static void read_int(oskar_Binary* h, unsigned char tag, int* v, int* status)
{
    oskar_binary_read_int(h, OSKAR_TAG_GROUP_VIS_HEADER, tag, 0, v, status);
}

static void read_double(oskar_Binary* h, unsigned char tag, double* v,
        int* status)
{
    oskar_binary_read_double(h, OSKAR_TAG_GROUP_VIS_HEADER, tag, 0, v, status);
}

oskar_VisHeader* oskar_vis_header_read(oskar_Binary* h, int* status)
{
    int i, j, k, err;
    int write_auto = 0, write_cross = 0, amp_type = 0, coord_precision = 0;
    int max_times = 0, num_times = 0, max_channels = 0, num_channels = 0;
    int num_stations = 0, num_tags = 0;
    const unsigned char grp = OSKAR_TAG_GROUP_VIS_HEADER;
    oskar_VisHeader* hdr = 0;
    if (*status) return 0;

    /* Dimensions and types, required. */
    read_int(h, OSKAR_VIS_HEADER_TAG_WRITE_AUTO_CORRELATIONS, &write_auto,
            status);
    read_int(h, OSKAR_VIS_HEADER_TAG_WRITE_CROSS_CORRELATIONS, &write_cross,
            status);
    read_int(h, OSKAR_VIS_HEADER_TAG_AMP_TYPE, &amp_type, status);
    read_int(h, OSKAR_VIS_HEADER_TAG_COORD_PRECISION, &coord_precision,
            status);
    read_int(h, OSKAR_VIS_HEADER_TAG_MAX_TIMES_PER_BLOCK, &max_times, status);
    read_int(h, OSKAR_VIS_HEADER_TAG_NUM_TIMES_TOTAL, &num_times, status);
    read_int(h, OSKAR_VIS_HEADER_TAG_MAX_CHANNELS_PER_BLOCK, &max_channels,
            status);
    read_int(h, OSKAR_VIS_HEADER_TAG_NUM_CHANNELS_TOTAL, &num_channels,
            status);
    read_int(h, OSKAR_VIS_HEADER_TAG_NUM_STATIONS, &num_stations, status);
    if (*status) return 0;
    num_tags = 9;
    hdr = oskar_vis_header_create(amp_type, coord_precision, max_times,
            num_times, max_channels, num_channels, num_stations, write_auto,
            write_cross, status);
    if (*status) return hdr;
    read_int(h, OSKAR_VIS_HEADER_TAG_NUM_TAGS_PER_BLOCK,
            &hdr->num_tags_per_block, status);
    num_tags += 1;

    /* Settings, optional. */
    err = 0;
    oskar_binary_read_mem(h, hdr->settings, OSKAR_TAG_GROUP_SETTINGS,
            OSKAR_TAG_SETTINGS, 0, &err);
    if (!err) num_tags += 1;
    oskar_binary_read_mem(h, hdr->telescope_path, grp,
            OSKAR_VIS_HEADER_TAG_TELESCOPE_PATH, 0, status);
    num_tags += 1;

    /* Phase convention: optional (older files lack it; then 0). */
    err = 0;
    read_int(h, OSKAR_VIS_HEADER_TAG_CASA_PHASE_CONVENTION,
            &hdr->casa_phase_convention, &err);
    if (!err) num_tags += 1;
    else hdr->casa_phase_convention = 0;

    /* Polarisation type: optional too (files written without it keep the
     * default of the amplitude type). */
    err = 0;
    {
        int pol = 0;
        read_int(h, OSKAR_VIS_HEADER_TAG_POL_TYPE, &pol, &err);
        if (!err)
        {
            hdr->pol_type = pol;
            num_tags += 1;
        }
    }
    read_int(h, OSKAR_VIS_HEADER_TAG_PHASE_CENTRE_COORD_TYPE,
            &hdr->phase_centre_type, status);
    oskar_binary_read(h, OSKAR_DOUBLE, grp,
            OSKAR_VIS_HEADER_TAG_PHASE_CENTRE_DEG, 0, 2 * sizeof(double),
            hdr->phase_centre_deg, status);
    read_double(h, OSKAR_VIS_HEADER_TAG_FREQ_START_HZ, &hdr->freq_start_hz,
            status);
    read_double(h, OSKAR_VIS_HEADER_TAG_FREQ_INC_HZ, &hdr->freq_inc_hz,
            status);
    read_double(h, OSKAR_VIS_HEADER_TAG_CHANNEL_BANDWIDTH_HZ,
            &hdr->channel_bandwidth_hz, status);
    read_double(h, OSKAR_VIS_HEADER_TAG_TIME_START_MJD_UTC,
            &hdr->time_start_mjd_utc, status);
    read_double(h, OSKAR_VIS_HEADER_TAG_TIME_INC_SEC, &hdr->time_inc_sec,
            status);
    read_double(h, OSKAR_VIS_HEADER_TAG_TIME_AVERAGE_SEC,
            &hdr->time_average_sec, status);
    read_double(h, OSKAR_VIS_HEADER_TAG_TELESCOPE_REF_LON_DEG,
            &hdr->telescope_centre_lon_deg, status);
    read_double(h, OSKAR_VIS_HEADER_TAG_TELESCOPE_REF_LAT_DEG,
            &hdr->telescope_centre_lat_deg, status);
    read_double(h, OSKAR_VIS_HEADER_TAG_TELESCOPE_REF_ALT_M,
            &hdr->telescope_centre_alt_m, status);
    num_tags += 11;
    for (i = 0; i < 3; ++i)
        oskar_binary_read_mem(h, hdr->station_offset_ecef_metres[i], grp,
                (unsigned char) (OSKAR_VIS_HEADER_TAG_STATION_X_OFFSET_ECEF +
                        i), 0, status);
    num_tags += 3;

    /* Optional per-station chunks: each group counts only if complete. */
    err = 0;
    for (k = 0; k < num_stations && !err; ++k)
        for (i = 0; i < 3; ++i)
            oskar_binary_read_mem(h, hdr->element_enu_metres[i][k], grp,
                    (unsigned char) (OSKAR_VIS_HEADER_TAG_ELEMENT_X_ENU + i),
                    k, &err);
    if (!err) num_tags += 3 * num_stations;
    err = 0;
    for (k = 0; k < num_stations && !err; ++k)
        oskar_binary_read_mem(h, hdr->station_name[k], grp,
                OSKAR_VIS_HEADER_TAG_STATION_NAME, k, &err);
    if (!err) num_tags += num_stations;
    err = 0;
    oskar_binary_read_mem(h, hdr->station_diameter_m, grp,
            OSKAR_VIS_HEADER_TAG_STATION_DIAMETER, 0, &err);
    if (!err) num_tags += 1;
    err = 0;
    for (k = 0; k < num_stations && !err; ++k)
        for (j = 0; j < 3; ++j)
            for (i = 0; i < 2; ++i)
                oskar_binary_read_mem(h,
                        hdr->element_euler_angle_rad[i][j][k], grp,
                        (unsigned char) (
                                OSKAR_VIS_HEADER_TAG_ELEMENT_FEED_ANGLE_X_A +
                                2 * j + i), k, &err);
    if (!err) num_tags += 6 * num_stations;
    hdr->num_tags_header = num_tags;
    return hdr;
}
