/*
 * Cross-correlation: V_b += sum_s smear_s(b) * J_P(s) B_s J_Q(s)^H.
 *
 * Reference: oskar/correlate/oskar_cross_correlate.h:48-63 and
 * oskar/correlate/src/oskar_cross_correlate.c:20-739 (derived scalars,
 * checks in the order location -> types -> dimensions, dispatch).  Location
 * OSKAR_GPU runs the hand-written gfx950 HIP kernels of this library on the
 * current HIP stream (see oskar_device.h); OSKAR_CPU returns
 * OSKAR_ERR_BAD_LOCATION, because this library carries no CPU path.
 *
 * Semantics kept from the reference:
 *  - baselines are ordered q = 0..N-2, p = q+1..N-1 and indexed
 *    b = q(N-1) - (q-1)q/2 + p - q - 1 (correlate/define_correlate_utils.h:42);
 *  - non-CASA: V = J_p B J_q^H with p the higher station index; CASA swaps
 *    the roles (correlate/define_cross_correlate.h:53-58);
 *  - the result is ADDED to vis[offset_out + b]; vis is never cleared and
 *    baselines outside the uv filter are left untouched;
 *  - bandwidth smearing sinc(pi dnu/f (uu l + vv m + ww (n-1))), time
 *    smearing with omega_E = 7.272205217e-5 rad/s only for phase centres that
 *    are not AZEL, Gaussian factor exp(-(a uu^2 + b 2uu vv + c vv^2))
 *    (correlate/define_correlate_utils.h:5-26, define_cross_correlate.h:16-32);
 *  - source_type 1 selects the Gaussian kernel; src_ext is ignored otherwise.
 */
#ifndef OSKAR_CROSS_CORRELATE_H_
#define OSKAR_CROSS_CORRELATE_H_

#include "oskar_jones.h"
#include "oskar_telescope.h"

#ifdef __cplusplus
extern "C" {
#endif

OSKAR_EXPORT
void oskar_cross_correlate(
        int use_casa_phase_convention,
        int source_type,
        int num_sources,
        const oskar_Jones* jones,
        const oskar_Mem* const src_flux[4],
        const oskar_Mem* const src_dir[3],
        const oskar_Mem* const src_ext[3],
        const oskar_Telescope* tel,
        const oskar_Mem* const station_uvw[3],
        double gast,
        double frequency_hz,
        int offset_out,
        oskar_Mem* vis,
        int* status);

/*
 * oskar_cross_correlate() without any device allocation.
 *
 * No counterpart in the reference.  For callers that cannot allocate
 * device memory: code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc.  Arguments, checks, errors and results
 * as oskar_cross_correlate(); the split-source partial sums go to a fixed
 * 64 MiB array that every device holds until the process ends, instead of
 * a workspace the library allocates.  No input is rejected for its size:
 * when the partials of the usual split count do not fit, the sources are
 * split fewer times (down to one split, which adds to vis directly and
 * uses no scratch), so results then differ from oskar_cross_correlate()'s
 * by the FP64 summation order only.  Up to 512 stations the split count,
 * and so every result bit, is that of oskar_cross_correlate().  Every call
 * on one device shares that device's array: calls that overlap on one
 * device (two streams or two host threads) race and must be serialised by
 * the caller.
 */
OSKAR_EXPORT
void oskar_cross_correlate_no_alloc(
        int use_casa_phase_convention,
        int source_type,
        int num_sources,
        const oskar_Jones* jones,
        const oskar_Mem* const src_flux[4],
        const oskar_Mem* const src_dir[3],
        const oskar_Mem* const src_ext[3],
        const oskar_Telescope* tel,
        const oskar_Mem* const station_uvw[3],
        double gast,
        double frequency_hz,
        int offset_out,
        oskar_Mem* vis,
        int* status);

/*
 * Backend seam: raw device pointers, launched on the current HIP stream.
 * These replace oskar_cross_correlate_{point,gaussian}_cuda_{f,d}
 * (oskar/correlate/oskar_cross_correlate_cuda.h:59-244) and
 * oskar_cross_correlate_scalar_{point,gaussian}_cuda_{f,d}
 * (oskar/correlate/oskar_cross_correlate_scalar_cuda.h) with the same
 * argument lists.  Errors from the launch are returned through the HIP
 * runtime's last-error state, as in the reference.
 */
OSKAR_EXPORT
void oskar_cross_correlate_point_hip_f(
        int use_casa_phase_convention,
        int num_sources, int num_stations, int offset_out,
        const float4c* d_jones, const float* d_I, const float* d_Q,
        const float* d_U, const float* d_V,
        const float* d_l, const float* d_m, const float* d_n,
        const float* d_station_u, const float* d_station_v,
        const float* d_station_w,
        const float* d_station_x, const float* d_station_y,
        float uv_min_lambda, float uv_max_lambda, float inv_wavelength,
        float frac_bandwidth, float time_int_sec, float gha0_rad,
        float dec0_rad, float4c* d_vis);
OSKAR_EXPORT
void oskar_cross_correlate_point_hip_d(
        int use_casa_phase_convention,
        int num_sources, int num_stations, int offset_out,
        const double4c* d_jones, const double* d_I, const double* d_Q,
        const double* d_U, const double* d_V,
        const double* d_l, const double* d_m, const double* d_n,
        const double* d_station_u, const double* d_station_v,
        const double* d_station_w,
        const double* d_station_x, const double* d_station_y,
        double uv_min_lambda, double uv_max_lambda, double inv_wavelength,
        double frac_bandwidth, double time_int_sec, double gha0_rad,
        double dec0_rad, double4c* d_vis);
OSKAR_EXPORT
void oskar_cross_correlate_gaussian_hip_f(
        int use_casa_phase_convention,
        int num_sources, int num_stations, int offset_out,
        const float4c* d_jones, const float* d_I, const float* d_Q,
        const float* d_U, const float* d_V,
        const float* d_l, const float* d_m, const float* d_n,
        const float* d_a, const float* d_b, const float* d_c,
        const float* d_station_u, const float* d_station_v,
        const float* d_station_w,
        const float* d_station_x, const float* d_station_y,
        float uv_min_lambda, float uv_max_lambda, float inv_wavelength,
        float frac_bandwidth, float time_int_sec, float gha0_rad,
        float dec0_rad, float4c* d_vis);
OSKAR_EXPORT
void oskar_cross_correlate_gaussian_hip_d(
        int use_casa_phase_convention,
        int num_sources, int num_stations, int offset_out,
        const double4c* d_jones, const double* d_I, const double* d_Q,
        const double* d_U, const double* d_V,
        const double* d_l, const double* d_m, const double* d_n,
        const double* d_a, const double* d_b, const double* d_c,
        const double* d_station_u, const double* d_station_v,
        const double* d_station_w,
        const double* d_station_x, const double* d_station_y,
        double uv_min_lambda, double uv_max_lambda, double inv_wavelength,
        double frac_bandwidth, double time_int_sec, double gha0_rad,
        double dec0_rad, double4c* d_vis);
OSKAR_EXPORT
void oskar_cross_correlate_scalar_point_hip_f(
        int use_casa_phase_convention,
        int num_sources, int num_stations, int offset_out,
        const float2* d_jones, const float* d_I,
        const float* d_l, const float* d_m, const float* d_n,
        const float* d_station_u, const float* d_station_v,
        const float* d_station_w,
        const float* d_station_x, const float* d_station_y,
        float uv_min_lambda, float uv_max_lambda, float inv_wavelength,
        float frac_bandwidth, float time_int_sec, float gha0_rad,
        float dec0_rad, float2* d_vis);
OSKAR_EXPORT
void oskar_cross_correlate_scalar_point_hip_d(
        int use_casa_phase_convention,
        int num_sources, int num_stations, int offset_out,
        const double2* d_jones, const double* d_I,
        const double* d_l, const double* d_m, const double* d_n,
        const double* d_station_u, const double* d_station_v,
        const double* d_station_w,
        const double* d_station_x, const double* d_station_y,
        double uv_min_lambda, double uv_max_lambda, double inv_wavelength,
        double frac_bandwidth, double time_int_sec, double gha0_rad,
        double dec0_rad, double2* d_vis);
OSKAR_EXPORT
void oskar_cross_correlate_scalar_gaussian_hip_f(
        int use_casa_phase_convention,
        int num_sources, int num_stations, int offset_out,
        const float2* d_jones, const float* d_I,
        const float* d_l, const float* d_m, const float* d_n,
        const float* d_a, const float* d_b, const float* d_c,
        const float* d_station_u, const float* d_station_v,
        const float* d_station_w,
        const float* d_station_x, const float* d_station_y,
        float uv_min_lambda, float uv_max_lambda, float inv_wavelength,
        float frac_bandwidth, float time_int_sec, float gha0_rad,
        float dec0_rad, float2* d_vis);
OSKAR_EXPORT
void oskar_cross_correlate_scalar_gaussian_hip_d(
        int use_casa_phase_convention,
        int num_sources, int num_stations, int offset_out,
        const double2* d_jones, const double* d_I,
        const double* d_l, const double* d_m, const double* d_n,
        const double* d_a, const double* d_b, const double* d_c,
        const double* d_station_u, const double* d_station_v,
        const double* d_station_w,
        const double* d_station_x, const double* d_station_y,
        double uv_min_lambda, double uv_max_lambda, double inv_wavelength,
        double frac_bandwidth, double time_int_sec, double gha0_rad,
        double dec0_rad, double2* d_vis);

#ifdef __cplusplus
}
#endif

#endif /* OSKAR_CROSS_CORRELATE_H_ */
