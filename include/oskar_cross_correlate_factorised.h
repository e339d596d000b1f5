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

/*
 * One station beam per station (no counterpart in the reference; the case
 * oskar_evaluate_jones_E produces when station beams are not duplicated,
 * oskar/interferometer/src/oskar_evaluate_jones_E.c:126-140):
 *
 *   V_pq = sum_s K_p(s) conj(K_q(s)) E_p(s) B_s E_q(s)^H,
 *
 * with E an oskar_Jones of num_stations x (>= num_sources) complex 2x2
 * matrices in the precision of vis.  Each B_s is factorised as
 * L_s D_s L_s^H (D = identity for any source with |pol| <= I; +-1 signs
 * otherwise, so every B is allowed), A_p(s) = K_p(s) E_p(s) L_s, and
 * V = A D A^H is one Hermitian rank-2S product on split-16-bit MFMA with
 * FP32 accumulation per part of sources and FP64 sums over parts.  The
 * station phases and beams enter once per (station, source).  Other
 * arguments, checks, accumulation and errors as
 * oskar_cross_correlate_factorised(); the Jones must be on the device of
 * vis.  Relative RMS error vs FP64 about 1e-6 (FP16X3).
 */
OSKAR_EXPORT
void oskar_cross_correlate_factorised_jones(
        int use_casa_phase_convention,
        int num_sources,
        const oskar_Jones* E,
        const oskar_Mem* const src_flux[4],
        const oskar_Mem* const src_dir[3],
        const oskar_Telescope* tel,
        const oskar_Mem* const station_uvw[3],
        double frequency_hz,
        int mode,
        int offset_out,
        oskar_Mem* vis,
        int* status);

/* Sources per part of oskar_cross_correlate_factorised_jones() (default
 * and maximum 2048; multiple of 4).  Returns the previous value. */
OSKAR_EXPORT int oskar_cross_correlate_factorised_jones_set_part_size(
        int sources);

/* As oskar_cross_correlate_factorised_last_times() for the last profiled
 * oskar_cross_correlate_factorised_jones() call: ms[0..2] = operands
 * (sources, phases, A = K E L, split), MFMA product, reduction. */
OSKAR_EXPORT int oskar_cross_correlate_factorised_jones_last_times(
        double ms[3]);

/* Sources accumulated in FP32 by one workgroup before the FP64 sum
 * (default 8192; multiple of 8).  Returns the previous value. */
OSKAR_EXPORT int oskar_cross_correlate_factorised_set_part_size(int sources);

/* Tile shape of oskar_cross_correlate_factorised(), as for batches
 * (0 = 128 x 64, 1 = 64 x 64 (default), 2 = 128 x 32, 3 = 64 x 64 with
 * separate MFMA and builder waves, 4 = 128 x 64 with 512 threads).
 * Returns the previous shape; a negative shape only returns the current
 * one; a shape past the last returns -1 and changes nothing. */
OSKAR_EXPORT int oskar_cross_correlate_factorised_set_tile_shape(int shape);


/* Stage timing with HIP events on the launch stream (off by default).
 * last_times() waits for the last profiled call on the current device and
 * returns 0 with ms[0..2] = preparation, MFMA product and reduction times;
 * -1 if nothing has been profiled. */
OSKAR_EXPORT void oskar_cross_correlate_factorised_set_profiling(int on);
OSKAR_EXPORT int oskar_cross_correlate_factorised_last_times(double ms[3]);

/*
 * Batches of (channel, time) units in one persistent launch.
 *
 * No counterpart in the reference: OSKAR calls the correlator once per
 * (time, channel, source chunk) from sim_baselines
 * (oskar/interferometer/src/oskar_interferometer_run_block.c:88-145,
 * 240-333).  A batch holds up to max_units units, each with its own source
 * count, station beam, fluxes, directions, station (u, v, w), frequency and
 * output; oskar_factorised_batch_run() computes, for every unit u < num_units,
 * exactly what oskar_cross_correlate_factorised() computes for it, in four
 * kernel launches: stations, sources, a persistent MFMA product over all
 * (unit, source part, tile) work items, and the FP64 reduction.  These are
 * captured in a hipGraph on the first run and replayed by later runs with
 * any units (sizes and pointers live in device-side descriptors).
 *
 * The unit inputs are referenced, not copied: they must stay valid and
 * unchanged until the run has completed on the stream.  Units that share
 * inputs (the directions of one time sample across channels) may pass the
 * same oskar_Mem.  Runs on one batch are ordered by the stream; a batch
 * belongs to the device current at creation.
 */
typedef struct oskar_FactorisedBatch oskar_FactorisedBatch;

/* precision: OSKAR_SINGLE or OSKAR_DOUBLE (of every unit's inputs and vis). */
OSKAR_EXPORT
oskar_FactorisedBatch* oskar_factorised_batch_create(int precision, int mode,
        int num_stations, int max_units, int* status);

/* Arguments as oskar_cross_correlate_factorised(), for unit `unit`. */
OSKAR_EXPORT
void oskar_factorised_batch_set_unit(oskar_FactorisedBatch* batch, int unit,
        int num_sources, const oskar_Mem* station_beam,
        const oskar_Mem* const src_flux[4], const oskar_Mem* const src_dir[3],
        const oskar_Mem* const station_uvw[3], double frequency_hz,
        int offset_out, oskar_Mem* vis, int* status);

/* Adds every unit's visibilities to its output (the uv filter and
 * smearing checks of oskar_cross_correlate_factorised() apply, per unit). */
OSKAR_EXPORT
void oskar_factorised_batch_run(oskar_FactorisedBatch* batch, int num_units,
        int use_casa_phase_convention, const oskar_Telescope* tel,
        int* status);

/* Picks the tile shape, part size and persistent grid by timing every
 * candidate on units [0, num_units) (outputs go to scratch memory; the
 * units' visibilities are not touched).  The result is cached per
 * (stations, precision, mode, source-count class) for the process, so a
 * second batch of the same kind reuses it without timing. */
OSKAR_EXPORT
void oskar_factorised_batch_autotune(oskar_FactorisedBatch* batch,
        int num_units, int* status);

/* Tile shapes: 0 = 128 x 64 stations, 1 = 64 x 64, 2 = 128 x 32,
 * 3 = 64 x 64 with separate MFMA and builder waves (512 threads),
 * 4 = 128 x 64 with 512 threads (two waves per SIMD).
 * src_per_part: sources summed in FP32 per work item (multiple of 8).
 * num_workgroups: persistent grid (0 = compute units x occupancy).
 * use_graph: capture and replay a hipGraph (default 1). */
OSKAR_EXPORT
void oskar_factorised_batch_set_config(oskar_FactorisedBatch* batch,
        int tile_shape, int src_per_part, int num_workgroups, int use_graph,
        int* status);

/* config[0..3] = tile shape, part size, workgroups, graph on/off. */
OSKAR_EXPORT
void oskar_factorised_batch_config(const oskar_FactorisedBatch* batch,
        int config[4]);

/* Milliseconds from the start of the last run's first kernel to the end of
 * its last (waits for it); negative if there is none. */
OSKAR_EXPORT
double oskar_factorised_batch_last_run_ms(oskar_FactorisedBatch* batch);

OSKAR_EXPORT
void oskar_factorised_batch_free(oskar_FactorisedBatch* batch, int* status);

#ifdef __cplusplus
}
#endif

#endif /* OSKAR_CROSS_CORRELATE_FACTORISED_H_ */
