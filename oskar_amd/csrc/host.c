/*
 * Device, stream, error-string and timer support (host C).
 *
 * Reference behaviour: oskar/utility/src/oskar_device.cpp:73-79 (runtime
 * errors surface as positive codes), oskar/utility/src/oskar_timer.c:71-209
 * (start/pause/resume/elapsed semantics).
 */
#include <hip/hip_runtime_api.h>
#include <stdlib.h>
#include <time.h>

#include "private.h"
#include "oskar_device.h"

static _Thread_local void* current_stream = 0;
static _Thread_local int last_error = 0;

void oskar_amd_set_last_error(int error) { last_error = error; }

void oskar_hip_set_stream(void* stream) { current_stream = stream; }
void* oskar_hip_stream(void) { return current_stream; }
void* oskar_amd_current_stream(void) { return current_stream; }

int oskar_amd_hip_status(int hip_error)
{
    return hip_error == (int) hipSuccess ? 0 : hip_error;
}

void oskar_amd_check_launch(int* status)
{
    const hipError_t e = hipPeekAtLastError();
    if (e != hipSuccess && !*status) *status = (int) e;
}

void oskar_device_check_error_hip(int* status)
{
    const hipError_t e = hipGetLastError();
    if (last_error && !*status) *status = last_error;
    last_error = 0;
    if (e != hipSuccess && !*status) *status = (int) e;
}

void oskar_device_set(int location, int id, int* status)
{
    if (*status) return;
    if (location == OSKAR_GPU)
    {
        const hipError_t e = hipSetDevice(id);
        if (e != hipSuccess) *status = (int) e;
    }
    else if (location != OSKAR_CPU)
    {
        *status = OSKAR_ERR_BAD_LOCATION;
    }
}

void oskar_device_synchronize(int location)
{
    if (location == OSKAR_GPU) (void) hipDeviceSynchronize();
}

int oskar_device_count_hip(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

const char* oskar_get_error_string(int error)
{
    if (error > 0) return hipGetErrorString((hipError_t) error);
    switch (error)
    {
    case 0:                                return "no error";
    case OSKAR_ERR_EOF:                    return "end of file";
    case OSKAR_ERR_FILE_IO:                return "file I/O error";
    case OSKAR_ERR_INVALID_ARGUMENT:       return "invalid function argument";
    case OSKAR_ERR_FUNCTION_NOT_AVAILABLE: return "function not available";
    case OSKAR_ERR_OUT_OF_RANGE:           return "out of range";
    case OSKAR_ERR_MEMORY_ALLOC_FAILURE:   return "memory allocation failure";
    case OSKAR_ERR_MEMORY_COPY_FAILURE:    return "memory copy failure";
    case OSKAR_ERR_MEMORY_NOT_ALLOCATED:   return "memory not allocated";
    case OSKAR_ERR_TYPE_MISMATCH:          return "data type mismatch";
    case OSKAR_ERR_LOCATION_MISMATCH:      return "data location mismatch";
    case OSKAR_ERR_DIMENSION_MISMATCH:     return "data dimension mismatch";
    case OSKAR_ERR_VALUE_MISMATCH:         return "value mismatch";
    case OSKAR_ERR_BAD_DATA_TYPE:          return "bad data type";
    case OSKAR_ERR_BAD_LOCATION:           return "bad memory location";
    case OSKAR_ERR_BAD_UNITS:              return "bad units";
    case OSKAR_ERR_CUDA_NOT_AVAILABLE:     return "GPU not available";
    case OSKAR_ERR_KERNEL_LAUNCH_FAILURE:  return "kernel launch failure";
    case OSKAR_ERR_COMPUTE_DEVICES:        return "compute device error";
    case OSKAR_ERR_SETTINGS_TELESCOPE:     return "bad telescope settings";
    default: break;
    }
    return "unknown error";
}

/* ---- Timer -------------------------------------------------------------- */
static double wall_time(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double) ts.tv_sec + 1e-9 * (double) ts.tv_nsec;
}

oskar_Timer* oskar_timer_create(int type)
{
    oskar_Timer* t = (oskar_Timer*) calloc(1, sizeof(oskar_Timer));
    if (!t) return 0;
    t->type = type;
    t->paused = 1;
    if (type == OSKAR_TIMER_HIP)
    {
        hipEvent_t a = 0, b = 0;
        if (hipEventCreate(&a) != hipSuccess ||
                hipEventCreate(&b) != hipSuccess)
        {
            /* Fall back to wall-clock timing when no device is present. */
            t->type = OSKAR_TIMER_NATIVE;
        }
        t->ev_start = (void*) a;
        t->ev_end = (void*) b;
    }
    return t;
}

void oskar_timer_free(oskar_Timer* timer)
{
    if (!timer) return;
    if (timer->ev_start) (void) hipEventDestroy((hipEvent_t) timer->ev_start);
    if (timer->ev_end) (void) hipEventDestroy((hipEvent_t) timer->ev_end);
    free(timer);
}

static double timer_interval(oskar_Timer* t)
{
    if (t->type == OSKAR_TIMER_HIP)
    {
        float ms = 0.0f;
        (void) hipEventRecord((hipEvent_t) t->ev_end,
                (hipStream_t) current_stream);
        (void) hipEventSynchronize((hipEvent_t) t->ev_end);
        (void) hipEventElapsedTime(&ms, (hipEvent_t) t->ev_start,
                (hipEvent_t) t->ev_end);
        return 1e-3 * (double) ms;
    }
    return wall_time() - t->start;
}

double oskar_timer_elapsed(oskar_Timer* timer)
{
    if (!timer) return 0.0;
    if (timer->paused) return timer->elapsed;
    /* Fold the running interval in and restart it. */
    timer->elapsed += timer_interval(timer);
    oskar_timer_resume(timer);
    timer->paused = 0;
    return timer->elapsed;
}

void oskar_timer_pause(oskar_Timer* timer)
{
    if (!timer || timer->paused) return;
    timer->elapsed += timer_interval(timer);
    timer->paused = 1;
}

void oskar_timer_reset(oskar_Timer* timer)
{
    if (!timer) return;
    timer->paused = 1;
    timer->elapsed = 0.0;
}

void oskar_timer_resume(oskar_Timer* timer)
{
    if (!timer) return;
    if (timer->type == OSKAR_TIMER_HIP)
        (void) hipEventRecord((hipEvent_t) timer->ev_start,
                (hipStream_t) current_stream);
    else
        timer->start = wall_time();
    timer->paused = 0;
}

void oskar_timer_restart(oskar_Timer* timer)
{
    oskar_timer_reset(timer);
    oskar_timer_resume(timer);
}

void oskar_timer_start(oskar_Timer* timer)
{
    oskar_timer_restart(timer);
}
