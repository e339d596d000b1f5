/*
 * Timers.
 *
 * Reference: oskar/utility/oskar_timer.h and src/oskar_timer.c:71-209.  A
 * timer created with OSKAR_TIMER_NATIVE measures host wall-clock time; one
 * created with OSKAR_TIMER_HIP (same value as the reference's
 * OSKAR_TIMER_CUDA) records HIP events on the current HIP stream, so it
 * times the device work queued between start/resume and pause/elapsed.
 */
#ifndef OSKAR_TIMER_H_
#define OSKAR_TIMER_H_

#include "oskar_global.h"

#ifdef __cplusplus
extern "C" {
#endif

struct oskar_Timer;
#ifndef OSKAR_TIMER_TYPEDEF_
#define OSKAR_TIMER_TYPEDEF_
typedef struct oskar_Timer oskar_Timer;
#endif

enum OSKAR_TIMER_TYPE
{
    OSKAR_TIMER_NATIVE = 0,
    OSKAR_TIMER_HIP = 1
};

OSKAR_EXPORT oskar_Timer* oskar_timer_create(int type);
OSKAR_EXPORT void oskar_timer_free(oskar_Timer* timer);
/* Total elapsed time in seconds; synchronises on the stop event. */
OSKAR_EXPORT double oskar_timer_elapsed(oskar_Timer* timer);
OSKAR_EXPORT void oskar_timer_pause(oskar_Timer* timer);
OSKAR_EXPORT void oskar_timer_reset(oskar_Timer* timer);
OSKAR_EXPORT void oskar_timer_resume(oskar_Timer* timer);
OSKAR_EXPORT void oskar_timer_restart(oskar_Timer* timer);
OSKAR_EXPORT void oskar_timer_start(oskar_Timer* timer);

#ifdef __cplusplus
}
#endif

#endif /* OSKAR_TIMER_H_ */
