/*
 * oskar_Jones: per-station, per-source Jones terms.
 *
 * Reference: struct oskar/interferometer/private_jones.h:34-41 (stations are
 * the slowest-varying dimension, sources the fastest: element (a, s) is at
 * a * num_sources + s); API oskar/interferometer/oskar_jones*.h.
 */
#ifndef OSKAR_JONES_H_
#define OSKAR_JONES_H_

#include "oskar_mem.h"

#ifdef __cplusplus
extern "C" {
#endif

struct oskar_Jones;
#ifndef OSKAR_JONES_TYPEDEF_
#define OSKAR_JONES_TYPEDEF_
typedef struct oskar_Jones oskar_Jones;
#endif

/* oskar/interferometer/oskar_jones_create.h */
OSKAR_EXPORT
oskar_Jones* oskar_jones_create(int type, int location, int num_stations,
        int num_sources, int* status);
/* oskar/interferometer/oskar_jones_create_copy.h */
OSKAR_EXPORT
oskar_Jones* oskar_jones_create_copy(const oskar_Jones* src, int location,
        int* status);
/* oskar/interferometer/oskar_jones_free.h */
OSKAR_EXPORT void oskar_jones_free(oskar_Jones* jones, int* status);
/* oskar/interferometer/oskar_jones_set_size.h */
OSKAR_EXPORT void oskar_jones_set_size(oskar_Jones* jones, int num_stations,
        int num_sources, int* status);
/* oskar/interferometer/oskar_jones_accessors.h */
OSKAR_EXPORT int oskar_jones_num_sources(const oskar_Jones* jones);
OSKAR_EXPORT int oskar_jones_num_stations(const oskar_Jones* jones);
OSKAR_EXPORT int oskar_jones_type(const oskar_Jones* jones);
OSKAR_EXPORT int oskar_jones_mem_location(const oskar_Jones* jones);
OSKAR_EXPORT oskar_Mem* oskar_jones_mem(oskar_Jones* jones);
OSKAR_EXPORT const oskar_Mem* oskar_jones_mem_const(const oskar_Jones* jones);

/* oskar/interferometer/oskar_jones_join.h: J3 = J1 * J2 element-wise
 * (2x2 matrix product per station and source); j3 may alias j1. */
OSKAR_EXPORT
void oskar_jones_join(oskar_Jones* j3, oskar_Jones* j1, const oskar_Jones* j2,
        int* status);

#ifdef __cplusplus
}
#endif

#endif /* OSKAR_JONES_H_ */
