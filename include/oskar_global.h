/*
 * Global definitions of the MI355X-native OSKAR hot-path library.
 *
 * Error codes, booleans, length units and coordinate types keep the numeric
 * values of the reference so that callers compiled against OSKAR's own
 * headers see identical semantics (oskar/oskar_global.h:20-97).
 *
 * Convention (oskar/oskar_global.h:20-55): every function takes a trailing
 * "int* status"; it returns immediately if *status is non-zero on entry and
 * sets a negative OSKAR_ERR_* code on failure.  Device (HIP runtime) errors
 * are reported as positive hipError_t values.
 */
#ifndef OSKAR_GLOBAL_H_
#define OSKAR_GLOBAL_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum OSKAR_ERROR_CODES
{
    OSKAR_ERR_EOF                                      = -1,
    OSKAR_ERR_FILE_IO                                  = -2,
    OSKAR_ERR_INVALID_ARGUMENT                         = -3,
    OSKAR_ERR_FUNCTION_NOT_AVAILABLE                   = -4,
    OSKAR_ERR_OUT_OF_RANGE                             = -5,
    OSKAR_ERR_MEMORY_ALLOC_FAILURE                     = -6,
    OSKAR_ERR_MEMORY_COPY_FAILURE                      = -7,
    OSKAR_ERR_MEMORY_NOT_ALLOCATED                     = -8,
    OSKAR_ERR_TYPE_MISMATCH                            = -9,
    OSKAR_ERR_LOCATION_MISMATCH                        = -10,
    OSKAR_ERR_DIMENSION_MISMATCH                       = -11,
    OSKAR_ERR_VALUE_MISMATCH                           = -12,
    OSKAR_ERR_BAD_DATA_TYPE                            = -13,
    OSKAR_ERR_BAD_LOCATION                             = -14,
    OSKAR_ERR_BAD_UNITS                                = -15,
    OSKAR_ERR_CUDA_NOT_AVAILABLE                       = -16,
    OSKAR_ERR_OPENCL_NOT_AVAILABLE                     = -17,
    OSKAR_ERR_KERNEL_LAUNCH_FAILURE                    = -18,
    OSKAR_ERR_COMPUTE_DEVICES                          = -19,
    OSKAR_ERR_ELLIPSE_FIT_FAILED                       = -22,
    OSKAR_ERR_SETTINGS_TELESCOPE                       = -23,
    OSKAR_ERR_SETUP_FAIL                               = -24
};

enum OSKAR_BOOL
{
    OSKAR_FALSE = 0,
    OSKAR_TRUE  = 1
};

enum OSKAR_LENGTH
{
    OSKAR_METRES = 0,
    OSKAR_WAVELENGTHS = 1
};

enum OSKAR_COORD_TYPE
{
    OSKAR_COORDS_REL_DIR,
    OSKAR_COORDS_ENU_DIR,
    OSKAR_COORDS_RADEC,
    OSKAR_COORDS_HADEC,
    OSKAR_COORDS_AZEL,
    OSKAR_COORDS_GALACTIC
};

#if defined(__GNUC__) && __GNUC__ >= 4
#    define OSKAR_EXPORT __attribute__ ((visibility ("default")))
#else
#    define OSKAR_EXPORT
#endif

/* Returns a human-readable string for an OSKAR_ERR_* code or a positive
 * HIP runtime error (mirrors oskar/utility/oskar_get_error_string.h). */
OSKAR_EXPORT
const char* oskar_get_error_string(int error);

#ifdef __cplusplus
}
#endif

#endif /* OSKAR_GLOBAL_H_ */
