/*
 * OSKAR binary container: writer, index reader, chunk reader, CRC-32C.
 *
 * Reference: oskar/binary/src/oskar_binary_create.c (header, index built by
 * reading every tag), oskar_binary_write.c (tag, payload, CRC),
 * oskar_binary_read.c, oskar_binary_query.c, oskar_crc.c.  The byte layout
 * is the file format and is kept exactly; the code is written here.
 */
#define _FILE_OFFSET_BITS 64
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/types.h>

#include "private.h"

#define HEADER_BYTES 64
#define TAG_BYTES 20
#define FLAG_EXTENDED 0x80
#define FLAG_CRC 0x40

typedef struct
{
    int extended;
    int data_type;
    int id_group, id_tag;
    char* name_group;
    char* name_tag;
    int user_index;
    int64_t payload_offset;
    size_t payload_size;
    int has_crc;
    unsigned int crc;          /* stored CRC of the chunk */
    unsigned int crc_head;     /* CRC of the tag and names */
} Chunk;

struct oskar_Binary
{
    FILE* stream;
    char mode;
    int num_chunks, cap_chunks, search_start;
    Chunk* chunk;
};

/* ---- CRC-32C (Castagnoli, reflected 0x82F63B78) ------------------------
 * x86-64 has the polynomial in hardware (SSE4.2 crc32, 8 bytes per
 * instruction); elsewhere, or without SSE4.2, a byte-wise table. */
static unsigned int crc_table[256];
static int crc_mode;                 /* 0 unset, 1 table, 2 hardware */

static void crc_init(void)
{
    unsigned int i, j;
    if (crc_mode) return;
    for (i = 0; i < 256; ++i)
    {
        unsigned int c = i;
        for (j = 0; j < 8; ++j) c = (c & 1u) ? (c >> 1) ^ 0x82F63B78u : c >> 1;
        crc_table[i] = c;
    }
#if defined(__x86_64__)
    __builtin_cpu_init();
    crc_mode = __builtin_cpu_supports("sse4.2") ? 2 : 1;
#else
    crc_mode = 1;
#endif
}

#if defined(__x86_64__)
#include <nmmintrin.h>
__attribute__((target("sse4.2")))
static unsigned int crc_hw(unsigned int c, const unsigned char* p, size_t n)
{
    uint64_t c64 = c;
    while (n && ((uintptr_t) p & 7u))
    {
        c64 = _mm_crc32_u8((unsigned int) c64, *p++);
        --n;
    }
    while (n >= 8)
    {
        uint64_t v;
        memcpy(&v, p, 8);
        c64 = _mm_crc32_u64(c64, v);
        p += 8;
        n -= 8;
    }
    while (n--) c64 = _mm_crc32_u8((unsigned int) c64, *p++);
    return (unsigned int) c64;
}
#endif

unsigned int oskar_binary_crc32c(unsigned int crc, const void* data,
        size_t num_bytes)
{
    const unsigned char* p = (const unsigned char*) data;
    crc_init();
    crc = ~crc;
#if defined(__x86_64__)
    if (crc_mode == 2) return ~crc_hw(crc, p, num_bytes);
#endif
    while (num_bytes--) crc = crc_table[(crc ^ *p++) & 0xFFu] ^ (crc >> 8);
    return ~crc;
}

/* ---- tags -------------------------------------------------------------- */
static int element_size(unsigned char type)
{
    int s = 0;
    if (type & OSKAR_CHAR) s = 1;
    else if (type & OSKAR_INT) s = 4;
    else if (type & OSKAR_SINGLE) s = 4;
    else if (type & OSKAR_DOUBLE) s = 8;
    else return 0;
    if (type & OSKAR_COMPLEX) s *= 2;
    if (type & OSKAR_MATRIX) s *= 4;
    return s;
}

static void put_le(unsigned char* dst, uint64_t v, int bytes)
{
    int i;
    for (i = 0; i < bytes; ++i) dst[i] = (unsigned char) (v >> (8 * i));
}

static uint64_t get_le(const unsigned char* src, int bytes)
{
    uint64_t v = 0;
    int i;
    for (i = bytes - 1; i >= 0; --i) v = (v << 8) | src[i];
    return v;
}

static void write_begin(oskar_Binary* h, unsigned char data_type,
        unsigned char flags, unsigned char group_byte,
        unsigned char tag_byte, const char* name_group, const char* name_tag,
        int user_index, size_t data_size, unsigned int* crc, int* status)
{
    unsigned char tag[TAG_BYTES];
    size_t names = 0;
    if (*status) return;
    if (h->mode != 'w' && h->mode != 'a')
    {
        *status = OSKAR_ERR_BINARY_NOT_OPEN_FOR_WRITE;
        return;
    }
    const int es = element_size(data_type);
    if (!es || es > 255)
    {
        *status = OSKAR_ERR_BINARY_TYPE_UNKNOWN;
        return;
    }
    if (flags & FLAG_EXTENDED) names = (size_t) group_byte + tag_byte;
    tag[0] = 'T';
    tag[1] = (unsigned char) (0x40 + OSKAR_BINARY_FORMAT_VERSION);
    tag[2] = 'G';
    tag[3] = (unsigned char) es;
    tag[4] = flags;
    tag[5] = data_type;
    tag[6] = group_byte;
    tag[7] = tag_byte;
    put_le(tag + 8, (uint32_t) user_index, 4);
    put_le(tag + 12, (uint64_t) (data_size + names + 4), 8);
    *crc = oskar_binary_crc32c(0, tag, TAG_BYTES);
    if (names)
    {
        *crc = oskar_binary_crc32c(*crc, name_group, group_byte);
        *crc = oskar_binary_crc32c(*crc, name_tag, tag_byte);
    }
    if (fwrite(tag, TAG_BYTES, 1, h->stream) != 1 ||
            (names && (fwrite(name_group, group_byte, 1, h->stream) != 1 ||
                       fwrite(name_tag, tag_byte, 1, h->stream) != 1)))
        *status = OSKAR_ERR_BINARY_WRITE_FAIL;
}

void oskar_amd_binary_write_begin(oskar_Binary* h, unsigned char data_type,
        unsigned char id_group, unsigned char id_tag, int user_index,
        size_t data_size, unsigned int* crc, int* status)
{
    write_begin(h, data_type, FLAG_CRC, id_group, id_tag, 0, 0, user_index,
            data_size, crc, status);
}

void oskar_amd_binary_write_raw(oskar_Binary* h, const void* data,
        size_t n, int* status)
{
    if (*status || !n) return;
    if (fwrite(data, 1, n, h->stream) != n)
        *status = OSKAR_ERR_BINARY_WRITE_FAIL;
}

void oskar_amd_binary_write_piece(oskar_Binary* h, const void* data,
        size_t n, unsigned int* crc, int* status)
{
    if (*status || !n) return;
    *crc = oskar_binary_crc32c(*crc, data, n);
    oskar_amd_binary_write_raw(h, data, n, status);
}

void oskar_amd_binary_write_end(oskar_Binary* h, unsigned int crc,
        int* status)
{
    unsigned char crc_le[4];
    if (*status) return;
    put_le(crc_le, crc, 4);
    if (fwrite(crc_le, 4, 1, h->stream) != 1)
        *status = OSKAR_ERR_BINARY_WRITE_FAIL;
}

static void write_chunk(oskar_Binary* h, unsigned char data_type,
        unsigned char flags, unsigned char group_byte,
        unsigned char tag_byte, const char* name_group, const char* name_tag,
        int user_index, size_t data_size, const void* data, int* status)
{
    unsigned int crc = 0;
    if (*status) return;
    if (data_size && !data)
    {
        *status = OSKAR_ERR_BINARY_MEMORY_NOT_ALLOCATED;
        return;
    }
    write_begin(h, data_type, flags, group_byte, tag_byte, name_group,
            name_tag, user_index, data_size, &crc, status);
    oskar_amd_binary_write_piece(h, data, data_size, &crc, status);
    oskar_amd_binary_write_end(h, crc, status);
}

void oskar_binary_write(oskar_Binary* handle, unsigned char data_type,
        unsigned char id_group, unsigned char id_tag, int user_index,
        size_t data_size, const void* data, int* status)
{
    write_chunk(handle, data_type, FLAG_CRC, id_group, id_tag, 0, 0,
            user_index, data_size, data, status);
}

void oskar_binary_write_ext(oskar_Binary* handle, unsigned char data_type,
        const char* name_group, const char* name_tag, int user_index,
        size_t data_size, const void* data, int* status)
{
    if (*status) return;
    const size_t lg = strlen(name_group), lt = strlen(name_tag);
    if (lg > 254 || lt > 254)
    {
        *status = OSKAR_ERR_BINARY_TAG_TOO_LONG;
        return;
    }
    write_chunk(handle, data_type, FLAG_EXTENDED | FLAG_CRC,
            (unsigned char) (lg + 1), (unsigned char) (lt + 1), name_group,
            name_tag, user_index, data_size, data, status);
}

void oskar_binary_write_int(oskar_Binary* handle, unsigned char id_group,
        unsigned char id_tag, int user_index, int value, int* status)
{
    oskar_binary_write(handle, OSKAR_INT, id_group, id_tag, user_index,
            sizeof(int), &value, status);
}

void oskar_binary_write_double(oskar_Binary* handle, unsigned char id_group,
        unsigned char id_tag, int user_index, double value, int* status)
{
    oskar_binary_write(handle, OSKAR_DOUBLE, id_group, id_tag, user_index,
            sizeof(double), &value, status);
}

/* ---- create / index ---------------------------------------------------- */
static void write_header(FILE* f, int* status)
{
    unsigned char hdr[HEADER_BYTES];
    memset(hdr, 0, sizeof(hdr));
    memcpy(hdr, "OSKARBIN", 9);                  /* with its NUL */
    hdr[9] = OSKAR_BINARY_FORMAT_VERSION;
    if (fseeko(f, 0, SEEK_SET) != 0 || fwrite(hdr, sizeof(hdr), 1, f) != 1)
        *status = OSKAR_ERR_BINARY_WRITE_FAIL;
}

static int add_chunk(oskar_Binary* h)
{
    if (h->num_chunks == h->cap_chunks)
    {
        const int cap = h->cap_chunks ? 2 * h->cap_chunks : 16;
        Chunk* c = (Chunk*) realloc(h->chunk, (size_t) cap * sizeof(Chunk));
        if (!c) return -1;
        h->chunk = c;
        h->cap_chunks = cap;
    }
    memset(&h->chunk[h->num_chunks], 0, sizeof(Chunk));
    return h->num_chunks++;
}

static void read_index(oskar_Binary* h, int* status)
{
    unsigned char hdr[HEADER_BYTES];
    if (fseeko(h->stream, 0, SEEK_SET) != 0 ||
            fread(hdr, sizeof(hdr), 1, h->stream) != 1)
    {
        *status = OSKAR_ERR_BINARY_READ_FAIL;
        return;
    }
    if (memcmp(hdr, "OSKARBIN", 8) != 0)
    {
        *status = OSKAR_ERR_BINARY_FILE_INVALID;
        return;
    }
    if ((int) hdr[9] > OSKAR_BINARY_FORMAT_VERSION)
    {
        *status = OSKAR_ERR_BINARY_VERSION_UNKNOWN;
        return;
    }
    for (;;)
    {
        unsigned char tag[TAG_BYTES];
        if (fread(tag, TAG_BYTES, 1, h->stream) != 1) break;
        if (tag[0] != 'T' || tag[2] != 'G' || (tag[4] & 0x1F) != 0)
        {
            *status = OSKAR_ERR_BINARY_FILE_INVALID;
            return;
        }
        const int version = (int) tag[1] - 0x40;
        if (version < 1 || version > OSKAR_BINARY_FORMAT_VERSION)
        {
            *status = OSKAR_ERR_BINARY_VERSION_UNKNOWN;
            return;
        }
        if (version > 1 && element_size(tag[5]) != (int) tag[3])
        {
            *status = OSKAR_ERR_BINARY_FORMAT_BAD;
            return;
        }
        const int i = add_chunk(h);
        if (i < 0)
        {
            *status = OSKAR_ERR_BINARY_MEMORY_NOT_ALLOCATED;
            return;
        }
        Chunk* c = &h->chunk[i];
        c->data_type = tag[5];
        c->id_group = tag[6];
        c->id_tag = tag[7];
        c->user_index = (int) (int32_t) get_le(tag + 8, 4);
        size_t block = (size_t) get_le(tag + 12, 8);
        c->has_crc = (tag[4] & FLAG_CRC) != 0;
        if (c->has_crc)
        {
            if (block < 4) { *status = OSKAR_ERR_BINARY_FILE_INVALID; return; }
            block -= 4;
        }
        c->crc_head = oskar_binary_crc32c(0, tag, TAG_BYTES);
        if (tag[4] & FLAG_EXTENDED)
        {
            const size_t lg = tag[6], lt = tag[7];
            c->extended = 1;
            if (block < lg + lt) { *status = OSKAR_ERR_BINARY_FILE_INVALID; return; }
            block -= lg + lt;
            c->name_group = (char*) calloc(lg + 1, 1);
            c->name_tag = (char*) calloc(lt + 1, 1);
            if (!c->name_group || !c->name_tag ||
                    (lg && fread(c->name_group, lg, 1, h->stream) != 1) ||
                    (lt && fread(c->name_tag, lt, 1, h->stream) != 1))
            {
                *status = OSKAR_ERR_BINARY_FILE_INVALID;
                return;
            }
            c->crc_head = oskar_binary_crc32c(c->crc_head, c->name_group, lg);
            c->crc_head = oskar_binary_crc32c(c->crc_head, c->name_tag, lt);
        }
        c->payload_size = block;
        c->payload_offset = (int64_t) ftello(h->stream);
        if (c->payload_offset < 0 ||
                fseeko(h->stream, (off_t) block, SEEK_CUR) != 0)
        {
            *status = OSKAR_ERR_BINARY_SEEK_FAIL;
            return;
        }
        if (c->has_crc)
        {
            unsigned char b[4];
            if (fread(b, 4, 1, h->stream) != 1)
            {
                *status = OSKAR_ERR_BINARY_READ_FAIL;
                return;
            }
            c->crc = (unsigned int) get_le(b, 4);
        }
    }
}

oskar_Binary* oskar_binary_create(const char* filename, char mode,
        int* status)
{
    oskar_Binary* h = 0;
    FILE* f = 0;
    if (*status) return 0;
    if (mode == 'r') f = fopen(filename, "rb");
    else if (mode == 'w') f = fopen(filename, "wb");
    else if (mode == 'a') f = fopen(filename, "a+b");
    if (!f)
    {
        *status = OSKAR_ERR_BINARY_OPEN_FAIL;
        return 0;
    }
    h = (oskar_Binary*) calloc(1, sizeof(oskar_Binary));
    if (!h)
    {
        (void) fclose(f);
        *status = OSKAR_ERR_BINARY_MEMORY_NOT_ALLOCATED;
        return 0;
    }
    h->stream = f;
    h->mode = mode;
    if (mode == 'w') write_header(f, status);
    else if (mode == 'a')
    {
        (void) fseeko(f, 0, SEEK_END);
        if (ftello(f) == 0) write_header(f, status);
    }
    else read_index(h, status);
    if (*status)
    {
        oskar_binary_free(h);
        return 0;
    }
    return h;
}

void oskar_binary_free(oskar_Binary* h)
{
    int i;
    if (!h) return;
    for (i = 0; i < h->num_chunks; ++i)
    {
        free(h->chunk[i].name_group);
        free(h->chunk[i].name_tag);
    }
    free(h->chunk);
    if (h->stream) (void) fclose(h->stream);
    free(h);
}

/* ---- queries and reads ------------------------------------------------- */
int oskar_binary_num_tags(const oskar_Binary* h) { return h->num_chunks; }

void oskar_binary_set_query_search_start(oskar_Binary* h, int start,
        int* status)
{
    if (*status) return;
    if (start < 0 || start >= h->num_chunks)
    {
        *status = OSKAR_ERR_BINARY_TAG_OUT_OF_RANGE;
        return;
    }
    h->search_start = start;
}

int oskar_binary_query(const oskar_Binary* h, unsigned char data_type,
        unsigned char id_group, unsigned char id_tag, int user_index,
        size_t* payload_size, int* status)
{
    int i;
    if (*status) return -1;
    for (i = h->search_start; i < h->num_chunks; ++i)
    {
        const Chunk* c = &h->chunk[i];
        if (!c->extended && (!data_type || c->data_type == data_type) &&
                c->id_group == id_group && c->id_tag == id_tag &&
                c->user_index == user_index)
        {
            if (payload_size) *payload_size = c->payload_size;
            return i;
        }
    }
    *status = OSKAR_ERR_BINARY_TAG_NOT_FOUND;
    return -1;
}

int oskar_binary_query_ext(const oskar_Binary* h, unsigned char data_type,
        const char* name_group, const char* name_tag, int user_index,
        size_t* payload_size, int* status)
{
    int i;
    if (*status) return -1;
    if (strlen(name_group) > 254 || strlen(name_tag) > 254)
    {
        *status = OSKAR_ERR_BINARY_TAG_TOO_LONG;
        return -1;
    }
    for (i = h->search_start; i < h->num_chunks; ++i)
    {
        const Chunk* c = &h->chunk[i];
        if (c->extended && (!data_type || c->data_type == data_type) &&
                c->user_index == user_index &&
                strcmp(c->name_group, name_group) == 0 &&
                strcmp(c->name_tag, name_tag) == 0)
        {
            if (payload_size) *payload_size = c->payload_size;
            return i;
        }
    }
    *status = OSKAR_ERR_BINARY_TAG_NOT_FOUND;
    return -1;
}

size_t oskar_amd_binary_read_begin(oskar_Binary* h, int chunk_index,
        unsigned int* crc, int* status)
{
    if (*status) return 0;
    if (h->mode != 'r')
    {
        *status = OSKAR_ERR_BINARY_NOT_OPEN_FOR_READ;
        return 0;
    }
    if (chunk_index < 0 || chunk_index >= h->num_chunks)
    {
        *status = OSKAR_ERR_BINARY_TAG_OUT_OF_RANGE;
        return 0;
    }
    const Chunk* c = &h->chunk[chunk_index];
    if (fseeko(h->stream, (off_t) c->payload_offset, SEEK_SET) != 0)
    {
        *status = OSKAR_ERR_BINARY_SEEK_FAIL;
        return 0;
    }
    *crc = c->crc_head;
    return c->payload_size;
}

void oskar_amd_binary_read_raw(oskar_Binary* h, void* data, size_t n,
        int* status)
{
    if (*status || !n) return;
    if (fread(data, 1, n, h->stream) != n)
        *status = OSKAR_ERR_BINARY_READ_FAIL;
}

void oskar_amd_binary_read_piece(oskar_Binary* h, void* data, size_t n,
        unsigned int* crc, int* status)
{
    oskar_amd_binary_read_raw(h, data, n, status);
    if (!*status && n) *crc = oskar_binary_crc32c(*crc, data, n);
}

void oskar_amd_binary_read_end(oskar_Binary* h, int chunk_index,
        unsigned int crc, int* status)
{
    if (*status) return;
    const Chunk* c = &h->chunk[chunk_index];
    if (c->has_crc && crc != c->crc) *status = OSKAR_ERR_BINARY_CRC_FAIL;
}

void oskar_binary_read_block(oskar_Binary* h, int chunk_index,
        size_t data_size, void* data, int* status)
{
    unsigned int crc = 0;
    const size_t n = oskar_amd_binary_read_begin(h, chunk_index, &crc,
            status);
    if (*status) return;
    if (n && (!data || data_size < n))
    {
        *status = OSKAR_ERR_BINARY_MEMORY_NOT_ALLOCATED;
        return;
    }
    oskar_amd_binary_read_piece(h, data, n, &crc, status);
    oskar_amd_binary_read_end(h, chunk_index, crc, status);
}

void oskar_binary_read(oskar_Binary* h, unsigned char data_type,
        unsigned char id_group, unsigned char id_tag, int user_index,
        size_t data_size, void* data, int* status)
{
    if (*status) return;
    if (h->mode != 'r')
    {
        *status = OSKAR_ERR_BINARY_NOT_OPEN_FOR_READ;
        return;
    }
    const int i = oskar_binary_query(h, data_type, id_group, id_tag,
            user_index, 0, status);
    oskar_binary_read_block(h, i, data_size, data, status);
}

void oskar_binary_read_ext(oskar_Binary* h, unsigned char data_type,
        const char* name_group, const char* name_tag, int user_index,
        size_t data_size, void* data, int* status)
{
    if (*status) return;
    if (h->mode != 'r')
    {
        *status = OSKAR_ERR_BINARY_NOT_OPEN_FOR_READ;
        return;
    }
    const int i = oskar_binary_query_ext(h, data_type, name_group, name_tag,
            user_index, 0, status);
    oskar_binary_read_block(h, i, data_size, data, status);
}

void oskar_binary_read_int(oskar_Binary* h, unsigned char id_group,
        unsigned char id_tag, int user_index, int* value, int* status)
{
    oskar_binary_read(h, OSKAR_INT, id_group, id_tag, user_index,
            sizeof(int), value, status);
}

void oskar_binary_read_double(oskar_Binary* h, unsigned char id_group,
        unsigned char id_tag, int user_index, double* value, int* status)
{
    oskar_binary_read(h, OSKAR_DOUBLE, id_group, id_tag, user_index,
            sizeof(double), value, status);
}
