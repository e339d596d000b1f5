/*
 * Learned surrogate of an aperture-array station beam.
 *
 * No counterpart in the reference, which evaluates the station beam exactly
 * with oskar_evaluate_station_beam_aperture_array()
 * (oskar/telescope/station/src/oskar_evaluate_station_beam_aperture_array.c
 * :31-288: element pattern, array-factor DFT, normalisation).  This is a
 * multilayer perceptron over direction and frequency,
 *   E(l, m, f) = W4 a3 + b4,  a_i = silu(W_i a_{i-1} + b_i),  a_{-1} = x(l, m, f),
 * with four hidden layers of 256 and a 128-feature sinusoidal encoding x:
 * sin and cos of 56 spatial frequencies (a Vogel spiral filling the disc of
 * radius station_radius * f_max / c cycles) of the scaled offsets
 * u = (f / f_max)(l - l0), v = (f / f_max)(m - m0) from the beam direction
 * (l0, m0), plus 16 smooth terms of (l, m, n, f).  Output: the 2x2 complex
 * Jones matrix of each direction, laid out as OSKAR's complex matrices.
 *
 * It is trained on the GPU (BF16 matrix-core forward and backward passes,
 * FP32 accumulation, FP32 master weights, Adam) from samples of the exact
 * beam, and evaluated by one fused kernel.  The exact beam remains the
 * judge of its error and the fallback where the surrogate is least
 * trustworthy: near nulls, where |E| is below a fraction of the beam's
 * peak (oskar_beam_surrogate_evaluate_with_fallback).
 */
#ifndef OSKAR_BEAM_SURROGATE_H_
#define OSKAR_BEAM_SURROGATE_H_

#include "oskar_mem.h"
#include "oskar_station.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oskar_BeamSurrogate oskar_BeamSurrogate;

/* Network sizes (fixed): parameters W0, b0, ..., W4, b4 flattened in that
 * order, W_i row-major [outputs][inputs]; the output layer has 32 rows of
 * which the first 8 are used (the rest stay zero). */
OSKAR_EXPORT int oskar_beam_surrogate_num_params(void);
OSKAR_EXPORT int oskar_beam_surrogate_num_features(void);
OSKAR_EXPORT int oskar_beam_surrogate_num_fourier(void);

/* A new network on the current device, weights drawn from
 * U(-1/sqrt(inputs), 1/sqrt(inputs)) with the given seed, biases zero. */
OSKAR_EXPORT
oskar_BeamSurrogate* oskar_beam_surrogate_create(double freq_min_hz,
        double freq_max_hz, double beam_l0, double beam_m0,
        double station_radius_m, unsigned int seed, int* status);

OSKAR_EXPORT
void oskar_beam_surrogate_free(oskar_BeamSurrogate* s, int* status);

/* The encoding's spatial frequencies (num_fourier x 2, cycles per unit of
 * (u, v)) and the network parameters (num_params floats), host memory. */
OSKAR_EXPORT
void oskar_beam_surrogate_fourier(const oskar_BeamSurrogate* s,
        float* fourier);
OSKAR_EXPORT
void oskar_beam_surrogate_get_params(oskar_BeamSurrogate* s, float* params,
        int* status);
/* Also resets the Adam state. */
OSKAR_EXPORT
void oskar_beam_surrogate_set_params(oskar_BeamSurrogate* s,
        const float* params, int* status);

/* Peak |E_ij| of the exact beam (set by oskar_beam_surrogate_train() from
 * its samples; 1 before).  It is the reference level of the fallback
 * threshold and the network's output unit: training targets are E / peak,
 * evaluation returns peak x the network output.  The parameters of a
 * trained network therefore go with its peak. */
OSKAR_EXPORT double oskar_beam_surrogate_peak(const oskar_BeamSurrogate* s);
OSKAR_EXPORT void oskar_beam_surrogate_set_peak(oskar_BeamSurrogate* s,
        double peak);

/* Samples of the exact station beam for training: num_freqs frequencies
 * evenly spaced over [freq_min_hz, freq_max_hz], each with dirs_per_freq
 * directions drawn over the sky above the horizon (a quarter of them within
 * 0.2 of (l0, m0) in direction cosines).  Outputs (resized, GPU, single
 * precision): l, m, f and the exact beam (complex matrix) of every sample,
 * from oskar_evaluate_station_beam_aperture_array(). */
OSKAR_EXPORT
void oskar_beam_surrogate_training_set(oskar_Station* station,
        oskar_StationWork* work, int num_freqs, int dirs_per_freq,
        double freq_min_hz, double freq_max_hz, double l0, double m0,
        unsigned int seed, oskar_Mem* l, oskar_Mem* m, oskar_Mem* f,
        oskar_Mem* beam, int* status);

/* num_steps Adam steps on random batches of batch_size samples (a multiple
 * of 256) of the training set; learning rate cosine-decayed from lr_start
 * to lr_end; loss_history (host, num_steps values, may be NULL) receives
 * the mean-square error of each step's batch before its update. */
OSKAR_EXPORT
void oskar_beam_surrogate_train(oskar_BeamSurrogate* s, int num_samples,
        const oskar_Mem* l, const oskar_Mem* m, const oskar_Mem* f,
        const oskar_Mem* beam, int batch_size, int num_steps,
        double lr_start, double lr_end, unsigned int seed,
        double* loss_history, int* status);

/* Loss and gradient (FP32, num_params values, GPU) for num_samples (a
 * multiple of 256) given encoded inputs: features [num_samples][128] and
 * targets [num_samples][8] in network units, single precision, GPU.
 * Parameters unchanged. */
OSKAR_EXPORT
void oskar_beam_surrogate_loss_and_gradient(oskar_BeamSurrogate* s,
        int num_samples, const oskar_Mem* features, const oskar_Mem* targets,
        double* loss, oskar_Mem* gradient, int* status);

/* Surrogate beam of num_points directions (l, m) at one frequency, into
 * beam (complex matrix, same precision as l and m, GPU). */
OSKAR_EXPORT
void oskar_beam_surrogate_evaluate(oskar_BeamSurrogate* s, int num_points,
        const oskar_Mem* l, const oskar_Mem* m, double frequency_hz,
        oskar_Mem* beam, int* status);

/* As oskar_beam_surrogate_evaluate() for directions (x, y, z) = (l, m, n),
 * then the exact beam (oskar_evaluate_station_beam_aperture_array() of the
 * station) wherever the surrogate's largest |E_ij| is below
 * threshold * peak; *num_exact receives the number of such directions. */
OSKAR_EXPORT
void oskar_beam_surrogate_evaluate_with_fallback(oskar_BeamSurrogate* s,
        oskar_Station* station, oskar_StationWork* work, int num_points,
        const oskar_Mem* x, const oskar_Mem* y, const oskar_Mem* z,
        double frequency_hz, double threshold, oskar_Mem* beam,
        int* num_exact, int* status);

#ifdef __cplusplus
}
#endif

#endif /* OSKAR_BEAM_SURROGATE_H_ */
