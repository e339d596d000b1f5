/*
 * Device selection, stream and synchronisation helpers.
 *
 * Reference: oskar/utility/oskar_device.h (oskar_device_set,
 * oskar_device_synchronize, oskar_device_check_error_cuda).  The reference
 * launches every kernel on the default stream 0 of the current CUDA device
 * (oskar/utility/src/oskar_device.cpp:259).  This library launches on a
 * per-host-thread "current stream" that defaults to the null stream; callers
 * that capture HIP graphs or overlap work set their own stream here.
 */
#ifndef OSKAR_DEVICE_H_
#define OSKAR_DEVICE_H_

#include "oskar_global.h"

#ifdef __cplusplus
extern "C" {
#endif

/* oskar_device_set: location OSKAR_GPU selects HIP device id. */
OSKAR_EXPORT void oskar_device_set(int location, int id, int* status);
/* Blocks until all work on the current device has finished. */
OSKAR_EXPORT void oskar_device_synchronize(int location);
/* Maps the HIP runtime's last error into *status (positive hipError_t). */
OSKAR_EXPORT void oskar_device_check_error_hip(int* status);
/* Number of HIP devices (0 if the runtime finds none). */
OSKAR_EXPORT int oskar_device_count_hip(void);

/* No counterpart in the reference (which has only stream 0): the HIP stream
 * (a hipStream_t passed as void*) used by this thread's launches; NULL
 * selects the null stream. */
OSKAR_EXPORT void oskar_hip_set_stream(void* stream);
OSKAR_EXPORT void* oskar_hip_stream(void);

#ifdef __cplusplus
}
#endif

#endif /* OSKAR_DEVICE_H_ */
