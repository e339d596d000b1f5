/*
 * Station-factorised cross-correlation on the MI355X matrix cores.
 *
 * No counterpart in the reference: this is the fast path for the case the
 * reference computes with oskar_cross_correlate() on Jones matrices
 * J_p(s) = K_p(s) E(s) (oskar/interferometer/src/oskar_interferometer_run_block.c
 * :251-333), when every station shares one station beam E(s), all sources
 * are points and neither bandwidth nor time smearing is on.  Then
 *
 *   V_pq = sum_s K_p(s) conj(K_q(s)) E(s) B_s E(s)^H,
 *   K_p(s) = exp(+i 2 pi f/c (u_p l + v_p m + w_p (n - 1)))
 *
 * (K as oskar/interferometer/define_evaluate_jones_K.h:56-75, B as
 * oskar/correlate/define_correlate_utils.h:36-39), evaluated as one complex
 * GEMM on split-16-bit MFMA with FP32 accumulation; phases, their range
 * reduction and direction cosines in FP64; partial sums added in FP64.
 * The result matches oskar_cross_correlate() on the materialised J to the
 * error each mode states (relative RMS of the visibilities vs FP64):
 *   OSKAR_FACTORISED_FP16X3  (default) FP16 hi+lo with a power-of-two scale,
 *                            about 1e-6;
 *   OSKAR_FACTORISED_BF16X3  BF16 hi+lo, about 1e-5.
 *
 * Arguments follow oskar_cross_correlate() (oskar/correlate/oskar_cross_correlate.h
 * :48-63): vis[offset_out + b] is ADDED to, baselines outside the uv filter
 * are untouched, the CASA convention conjugate-transposes each 2x2 result.
 * station_beam: complex matrix, >= num_sources values (NULL = identity).
 * src_flux: I, Q, U, V; src_dir: l, m, n.  Status OSKAR_ERR_FUNCTION_NOT_AVAILABLE
 * if the telescope has bandwidth or time smearing enabled (use
 * oskar_cross_correlate() for those).
 */
#ifndef OSKAR_CROSS_CORRELATE_FACTORISED_H_
#define OSKAR_CROSS_CORRELATE_FACTORISED_H_

#include "oskar_jones.h"
#include "oskar_telescope.h"

#ifdef __cplusplus
extern "C" {
#endif

enum OSKAR_FACTORISED_MODE
{
    OSKAR_FACTORISED_FP16X3 = 0,
    OSKAR_FACTORISED_BF16X3 = 1
};

OSKAR_EXPORT
void oskar_cross_correlate_factorised(
        int use_casa_phase_convention,
        int num_sources,
        const oskar_Mem* station_beam,
        const oskar_Mem* const src_flux[4],
        const oskar_Mem* const src_dir[3],
        const oskar_Telescope* tel,
        const oskar_Mem* const station_uvw[3],
        double frequency_hz,
        int mode,
        int offset_out,
        oskar_Mem* vis,
        int* status);

/* Sources accumulated in FP32 by one workgroup before the FP64 sum
 * (default 8192; multiple of 8).  Returns the previous value. */
OSKAR_EXPORT int oskar_cross_correlate_factorised_set_part_size(int sources);

/* Stage timing with HIP events on the launch stream (off by default).
 * last_times() waits for the last profiled call on the current device and
 * returns 0 with ms[0..2] = preparation, MFMA product and reduction times;
 * -1 if nothing has been profiled. */
OSKAR_EXPORT void oskar_cross_correlate_factorised_set_profiling(int on);
OSKAR_EXPORT int oskar_cross_correlate_factorised_last_times(double ms[3]);

#ifdef __cplusplus
}
#endif

#endif /* OSKAR_CROSS_CORRELATE_FACTORISED_H_ */
