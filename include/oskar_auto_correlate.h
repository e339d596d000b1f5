/*
 * Auto-correlation: vis[offset_out + s] += sum_i J_s(i) B_i J_s(i)^H.
 *
 * Reference: oskar/correlate/oskar_auto_correlate.h and
 * oskar/correlate/src/oskar_auto_correlate.c:22-151 (checks in the order
 * location -> type -> dimension, then the data type), kernel bodies
 * oskar/correlate/define_auto_correlate.h:7-104.  Location OSKAR_GPU runs the
 * gfx950 HIP kernels of this library on the current HIP stream; OSKAR_CPU
 * returns OSKAR_ERR_BAD_LOCATION.
 *
 * Semantics kept from the reference:
 *  - jones: num_stations x (>= num_sources) values, read with a source
 *    stride of num_sources, as the reference's kernels do;
 *  - matrix Jones: B_i = [[I+Q, U+iV], [U-iV, I-Q]]; the imaginary parts of
 *    the diagonal terms (a, d) of vis are left as they were;
 *  - scalar Jones: vis[offset_out + s].re += sum_i |J_s(i)|^2 I_i, the
 *    imaginary part is left as it was;
 *  - the result is ADDED to vis; vis is never cleared.
 * Checks beyond the reference's (which reads these unchecked): the source
 * fluxes must be device arrays of the Jones precision with >= num_sources
 * values (only I for scalar Jones), and vis must hold >= offset_out +
 * num_stations values (OSKAR_ERR_DIMENSION_MISMATCH otherwise).
 * Sums: FP per thread over a few sources and across a wave, then FP64.
 */
#ifndef OSKAR_AUTO_CORRELATE_H_
#define OSKAR_AUTO_CORRELATE_H_

#include "oskar_jones.h"

#ifdef __cplusplus
extern "C" {
#endif

OSKAR_EXPORT
void oskar_auto_correlate(
        int num_sources,
        const oskar_Jones* jones,
        const oskar_Mem* const src_flux[4],
        int offset_out,
        oskar_Mem* vis,
        int* status);

#ifdef __cplusplus
}
#endif

#endif /* OSKAR_AUTO_CORRELATE_H_ */
