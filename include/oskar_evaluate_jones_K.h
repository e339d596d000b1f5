/*
 * Interferometer phase Jones K.
 *
 * Reference: oskar/interferometer/oskar_evaluate_jones_K.h and
 * src/oskar_evaluate_jones_K.c:19-157, kernel define_evaluate_jones_K.h:56-75:
 *   K[a][s] = exp(+i k (u_a l_s + v_a m_s + w_a (n_s - 1))),  k = 2 pi f / c,
 * for sources with source_filter_min < source_filter[s] <= source_filter_max,
 * else 0; the w term is dropped when ignore_w_components is set.  K must be
 * a scalar complex Jones (stations slow, sources fast).  OSKAR_GPU runs a
 * HIP kernel; OSKAR_CPU returns OSKAR_ERR_BAD_LOCATION.
 */
#ifndef OSKAR_EVALUATE_JONES_K_H_
#define OSKAR_EVALUATE_JONES_K_H_

#include "oskar_jones.h"

#ifdef __cplusplus
extern "C" {
#endif

OSKAR_EXPORT
void oskar_evaluate_jones_K(
        oskar_Jones* K,
        int num_sources,
        const oskar_Mem* l,
        const oskar_Mem* m,
        const oskar_Mem* n,
        const oskar_Mem* u,
        const oskar_Mem* v,
        const oskar_Mem* w,
        double frequency_hz,
        const oskar_Mem* source_filter,
        double source_filter_min,
        double source_filter_max,
        int ignore_w_components,
        int* status);

#ifdef __cplusplus
}
#endif

#endif /* OSKAR_EVALUATE_JONES_K_H_ */
