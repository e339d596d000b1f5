/*
 * OSKAR binary container files (the .vis format's container).
 *
 * Reference: oskar/binary/oskar_binary.h (tag groups, metadata tags, error
 * codes and format version 2, kept identical: they are the file format),
 * oskar/binary/private_binary.h (64-byte file header, 20-byte chunk tag),
 * oskar/binary/src/oskar_binary_create.c, oskar_binary_write.c,
 * oskar_binary_read.c, oskar_binary_query.c, oskar_crc.c.
 *
 * File: a 64-byte header ("OSKARBIN\0", format version 2, the rest zero),
 * then chunks.  Chunk: a 20-byte tag (magic 'T', 0x40 + version, 'G',
 * element size; flags (bit 7 extended names, bit 6 CRC present); data type;
 * group id or group-name length; tag id or tag-name length; user index as
 * 4-byte little-endian; block size as 8-byte little-endian = payload
 * [+ names] + 4), for extended tags the two NUL-terminated names, the
 * payload, then the CRC-32C (Castagnoli, reflected polynomial 0x82F63B78,
 * initial value and final XOR 0xFFFFFFFF) of tag, names and payload as
 * 4 little-endian bytes.  Host byte order must be little-endian.
 */
#ifndef OSKAR_BINARY_H_
#define OSKAR_BINARY_H_

#include <stddef.h>

#include "oskar_global.h"

#ifdef __cplusplus
extern "C" {
#endif

#define OSKAR_BINARY_FORMAT_VERSION 2

struct oskar_Binary;
#ifndef OSKAR_BINARY_TYPEDEF_
#define OSKAR_BINARY_TYPEDEF_
typedef struct oskar_Binary oskar_Binary;
#endif

enum OSKAR_TAG_GROUPS
{
    OSKAR_TAG_GROUP_METADATA         = 1,
    OSKAR_TAG_GROUP_SYSTEM_INFO      = 2,
    OSKAR_TAG_GROUP_SETTINGS         = 3,
    OSKAR_TAG_GROUP_RUN              = 4,
    OSKAR_TAG_GROUP_VISIBILITY       = 5,
    OSKAR_TAG_GROUP_IMAGE            = 6,
    OSKAR_TAG_GROUP_SKY_MODEL        = 7,
    OSKAR_TAG_GROUP_TIME_FREQ_DATA   = 8,
    OSKAR_TAG_GROUP_SPLINE_DATA      = 9,
    OSKAR_TAG_GROUP_ELEMENT_DATA     = 10,
    OSKAR_TAG_GROUP_VIS_HEADER       = 11,
    OSKAR_TAG_GROUP_VIS_BLOCK        = 12
};

enum OSKAR_TAG_METADATA
{
    OSKAR_TAG_METADATA_DATE_TIME_STRING     = 1,
    OSKAR_TAG_METADATA_OSKAR_VERSION_STRING = 2,
    OSKAR_TAG_METADATA_USERNAME             = 3,
    OSKAR_TAG_METADATA_CWD                  = 4
};

enum OSKAR_TAG_SETTINGS
{
    OSKAR_TAG_SETTINGS_PATH = 1,
    OSKAR_TAG_SETTINGS      = 2
};

enum OSKAR_BINARY_ERRORS
{
    OSKAR_ERR_BINARY_OPEN_FAIL             = -100,
    OSKAR_ERR_BINARY_SEEK_FAIL             = -101,
    OSKAR_ERR_BINARY_READ_FAIL             = -102,
    OSKAR_ERR_BINARY_WRITE_FAIL            = -103,
    OSKAR_ERR_BINARY_NOT_OPEN_FOR_READ     = -104,
    OSKAR_ERR_BINARY_NOT_OPEN_FOR_WRITE    = -105,
    OSKAR_ERR_BINARY_FILE_INVALID          = -106,
    OSKAR_ERR_BINARY_FORMAT_BAD            = -107,
    OSKAR_ERR_BINARY_ENDIAN_MISMATCH       = -108,
    OSKAR_ERR_BINARY_VERSION_UNKNOWN       = -109,
    OSKAR_ERR_BINARY_TYPE_UNKNOWN          = -110,
    OSKAR_ERR_BINARY_INT_UNKNOWN           = -111,
    OSKAR_ERR_BINARY_FLOAT_UNKNOWN         = -112,
    OSKAR_ERR_BINARY_DOUBLE_UNKNOWN        = -113,
    OSKAR_ERR_BINARY_MEMORY_NOT_ALLOCATED  = -114,
    OSKAR_ERR_BINARY_TAG_NOT_FOUND         = -115,
    OSKAR_ERR_BINARY_TAG_TOO_LONG          = -116,
    OSKAR_ERR_BINARY_TAG_OUT_OF_RANGE      = -117,
    OSKAR_ERR_BINARY_CRC_FAIL              = -118
};

/* mode 'w' (truncate and write the header), 'a' (append; header only if
 * the file is empty) or 'r' (read the header and index every chunk). */
OSKAR_EXPORT
oskar_Binary* oskar_binary_create(const char* filename, char mode,
        int* status);
OSKAR_EXPORT void oskar_binary_free(oskar_Binary* handle);

OSKAR_EXPORT
void oskar_binary_write(oskar_Binary* handle, unsigned char data_type,
        unsigned char id_group, unsigned char id_tag, int user_index,
        size_t data_size, const void* data, int* status);
OSKAR_EXPORT
void oskar_binary_write_ext(oskar_Binary* handle, unsigned char data_type,
        const char* name_group, const char* name_tag, int user_index,
        size_t data_size, const void* data, int* status);
OSKAR_EXPORT
void oskar_binary_write_int(oskar_Binary* handle, unsigned char id_group,
        unsigned char id_tag, int user_index, int value, int* status);
OSKAR_EXPORT
void oskar_binary_write_double(oskar_Binary* handle, unsigned char id_group,
        unsigned char id_tag, int user_index, double value, int* status);

/* Index queries (read mode): number of chunks; the chunk of a (type,
 * group, tag, user index), type 0 matching any (-1 and
 * OSKAR_ERR_BINARY_TAG_NOT_FOUND if absent), with its payload size. */
OSKAR_EXPORT int oskar_binary_num_tags(const oskar_Binary* handle);
OSKAR_EXPORT
int oskar_binary_query(const oskar_Binary* handle, unsigned char data_type,
        unsigned char id_group, unsigned char id_tag, int user_index,
        size_t* payload_size, int* status);
OSKAR_EXPORT
int oskar_binary_query_ext(const oskar_Binary* handle,
        unsigned char data_type, const char* name_group,
        const char* name_tag, int user_index, size_t* payload_size,
        int* status);

/* Queries search the index from chunk start onwards (0 after create);
 * start beyond the last chunk gives OSKAR_ERR_BINARY_TAG_OUT_OF_RANGE. */
OSKAR_EXPORT
void oskar_binary_set_query_search_start(oskar_Binary* handle, int start,
        int* status);

/* Reads a chunk's payload into data (data_size >= payload) and checks its
 * CRC (OSKAR_ERR_BINARY_CRC_FAIL). */
OSKAR_EXPORT
void oskar_binary_read_block(oskar_Binary* handle, int chunk_index,
        size_t data_size, void* data, int* status);
OSKAR_EXPORT
void oskar_binary_read(oskar_Binary* handle, unsigned char data_type,
        unsigned char id_group, unsigned char id_tag, int user_index,
        size_t data_size, void* data, int* status);
OSKAR_EXPORT
void oskar_binary_read_ext(oskar_Binary* handle, unsigned char data_type,
        const char* name_group, const char* name_tag, int user_index,
        size_t data_size, void* data, int* status);
OSKAR_EXPORT
void oskar_binary_read_int(oskar_Binary* handle, unsigned char id_group,
        unsigned char id_tag, int user_index, int* value, int* status);
OSKAR_EXPORT
void oskar_binary_read_double(oskar_Binary* handle, unsigned char id_group,
        unsigned char id_tag, int user_index, double* value, int* status);

/* CRC-32C of num_bytes bytes continuing from a previous value (0 to
 * start): crc32c(A || B) == oskar_binary_crc32c(crc32c(A), B, n). */
OSKAR_EXPORT
unsigned int oskar_binary_crc32c(unsigned int crc, const void* data,
        size_t num_bytes);

#ifdef __cplusplus
}
#endif

#endif /* OSKAR_BINARY_H_ */
