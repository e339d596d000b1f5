/*
 * oskar_BeamSurrogate: checks, training-set generation from the exact beam,
 * and the exact-beam fallback, around the network kernels in mlp.hip.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "private.h"
#include "oskar_beam_surrogate.h"

#define NF 56
#define D_IN 128

int oskar_amd_surr_num_params(void);
int oskar_amd_surr_num_fourier(void);
void* oskar_amd_surr_create(const float* fourier, double l0, double m0,
        double f_min, double f_max, const float* params, int* err);
void oskar_amd_surr_free(void* h);
int oskar_amd_surr_get_params(void* h, float* params);
int oskar_amd_surr_set_params(void* h, const float* params);
int oskar_amd_surr_loss_grad(void* h, int B, const float* feat,
        const float* target, double* loss, float* grad_out, void* stream);
int oskar_amd_surr_train(void* h, int num_samples, const float* l,
        const float* m, const float* f, const float* E, int B, int num_steps,
        double lr0, double lr1, unsigned long long seed, double* loss_hist,
        void* stream);
int oskar_amd_surr_infer(void* h, int precision, int num_points,
        const void* l, const void* m, double freq_hz, double flag_level,
        void* beam, int* flags, void* stream);
int oskar_amd_surr_max_abs(const float* v, long long n, double* out);
void oskar_amd_surr_set_out_scale(void* h, double scale);
int oskar_amd_surr_compact(int precision, int n, const int* flags,
        const void* x, const void* y, const void* z, int* work, void* xc,
        void* yc, void* zc, int* count, void* stream);
int oskar_amd_surr_scatter(int precision, int count, const int* idx,
        const void* src, void* dst, void* stream);

struct oskar_BeamSurrogate
{
    void* impl;
    float fourier[NF][2];
    double peak;
};

/* splitmix64: the weights' and directions' random numbers (host). */
static uint64_t next_u64(uint64_t* s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static double next_uniform(uint64_t* s)
{
    return (double) (next_u64(s) >> 11) * (1.0 / 9007199254740992.0);
}

int oskar_beam_surrogate_num_params(void)
{
    return oskar_amd_surr_num_params();
}

int oskar_beam_surrogate_num_features(void) { return D_IN; }

int oskar_beam_surrogate_num_fourier(void) { return NF; }

oskar_BeamSurrogate* oskar_beam_surrogate_create(double freq_min_hz,
        double freq_max_hz, double beam_l0, double beam_m0,
        double station_radius_m, unsigned int seed, int* status)
{
    static const int n_out[5] = {256, 256, 256, 256, 32};
    static const int n_in[5] = {D_IN, 256, 256, 256, 256};
    oskar_BeamSurrogate* s = 0;
    float* params = 0;
    uint64_t rng = 0x5EEDull ^ ((uint64_t) seed << 17);
    int i = 0, o = 0, e = 0;
    if (*status) return 0;
    if (!(freq_min_hz > 0.0) || !(freq_max_hz > freq_min_hz) ||
            !(station_radius_m > 0.0))
    {
        *status = OSKAR_ERR_INVALID_ARGUMENT;
        return 0;
    }
    s = (oskar_BeamSurrogate*) calloc(1, sizeof(*s));
    params = (float*) calloc((size_t) oskar_amd_surr_num_params(),
            sizeof(float));
    if (!s || !params)
    {
        free(s);
        free(params);
        *status = OSKAR_ERR_MEMORY_ALLOC_FAILURE;
        return 0;
    }
    /* Spatial frequencies: a Vogel spiral filling the disc of radius
     * station_radius * f_max / c cycles per unit of (u, v). */
    {
        const double rho = station_radius_m * freq_max_hz / 299792458.0;
        const double golden = M_PI * (3.0 - sqrt(5.0));
        for (i = 0; i < NF; ++i)
        {
            const double r = rho * sqrt((i + 0.5) / NF), t = i * golden;
            s->fourier[i][0] = (float) (r * cos(t));
            s->fourier[i][1] = (float) (r * sin(t));
        }
    }
    for (i = 0; i < 5; ++i)
    {
        const double lim = 1.0 / sqrt((double) n_in[i]);
        const int rows = (i == 4) ? 8 : n_out[i];
        int k = 0;
        for (k = 0; k < n_out[i] * n_in[i]; ++k)
            params[o + k] = (k / n_in[i] < rows) ?
                    (float) (lim * (2.0 * next_uniform(&rng) - 1.0)) : 0.0f;
        o += n_out[i] * n_in[i] + n_out[i];   /* biases stay zero */
    }
    s->peak = 1.0;
    s->impl = oskar_amd_surr_create(&s->fourier[0][0], beam_l0, beam_m0,
            freq_min_hz, freq_max_hz, params, &e);
    free(params);
    if (e)
    {
        free(s);
        *status = e;
        return 0;
    }
    return s;
}

void oskar_beam_surrogate_free(oskar_BeamSurrogate* s, int* status)
{
    (void) status;
    if (!s) return;
    oskar_amd_surr_free(s->impl);
    free(s);
}

void oskar_beam_surrogate_fourier(const oskar_BeamSurrogate* s,
        float* fourier)
{
    memcpy(fourier, s->fourier, sizeof(s->fourier));
}

void oskar_beam_surrogate_get_params(oskar_BeamSurrogate* s, float* params,
        int* status)
{
    if (*status) return;
    *status = oskar_amd_surr_get_params(s->impl, params);
}

void oskar_beam_surrogate_set_params(oskar_BeamSurrogate* s,
        const float* params, int* status)
{
    if (*status) return;
    *status = oskar_amd_surr_set_params(s->impl, params);
}

double oskar_beam_surrogate_peak(const oskar_BeamSurrogate* s)
{
    return s->peak;
}

void oskar_beam_surrogate_set_peak(oskar_BeamSurrogate* s, double peak)
{
    s->peak = peak;
    oskar_amd_surr_set_out_scale(s->impl, peak > 0.0 ? peak : 1.0);
}

static int check_gpu_single(const oskar_Mem* m, int matrix, size_t n)
{
    if (oskar_mem_location(m) != OSKAR_GPU) return OSKAR_ERR_BAD_LOCATION;
    if (oskar_mem_precision(m) != OSKAR_SINGLE ||
            (matrix && !(oskar_mem_is_matrix(m) && oskar_mem_is_complex(m)))
            || (!matrix && !oskar_mem_is_real(m)))
        return OSKAR_ERR_BAD_DATA_TYPE;
    if (oskar_mem_length(m) < n) return OSKAR_ERR_DIMENSION_MISMATCH;
    return 0;
}

void oskar_beam_surrogate_training_set(oskar_Station* station,
        oskar_StationWork* work, int num_freqs, int dirs_per_freq,
        double freq_min_hz, double freq_max_hz, double l0, double m0,
        unsigned int seed, oskar_Mem* l, oskar_Mem* m, oskar_Mem* f,
        oskar_Mem* beam, int* status)
{
    const size_t total = (size_t) num_freqs * (size_t) dirs_per_freq;
    float *hl = 0, *hm = 0, *hn = 0, *hf = 0;
    oskar_Mem *dx = 0, *dy = 0, *dz = 0, *out = 0, *tmp = 0;
    uint64_t rng = 0xB3A7ull ^ ((uint64_t) seed << 21);
    int k = 0;
    size_t i = 0;
    if (*status) return;
    if (num_freqs < 1 || dirs_per_freq < 1)
    {
        *status = OSKAR_ERR_INVALID_ARGUMENT;
        return;
    }
    if (oskar_station_mem_location(station) != OSKAR_GPU ||
            oskar_mem_location(l) != OSKAR_GPU ||
            oskar_mem_location(m) != OSKAR_GPU ||
            oskar_mem_location(f) != OSKAR_GPU ||
            oskar_mem_location(beam) != OSKAR_GPU)
    {
        *status = OSKAR_ERR_BAD_LOCATION;
        return;
    }
    if (oskar_station_precision(station) != OSKAR_SINGLE)
    {
        *status = OSKAR_ERR_BAD_DATA_TYPE;
        return;
    }
    oskar_mem_realloc(l, total, status);
    oskar_mem_realloc(m, total, status);
    oskar_mem_realloc(f, total, status);
    oskar_mem_realloc(beam, total, status);
    *status = *status ? *status : check_gpu_single(l, 0, total);
    *status = *status ? *status : check_gpu_single(m, 0, total);
    *status = *status ? *status : check_gpu_single(f, 0, total);
    *status = *status ? *status : check_gpu_single(beam, 1, total);
    if (*status) return;
    hl = (float*) malloc(total * sizeof(float));
    hm = (float*) malloc(total * sizeof(float));
    hn = (float*) malloc(total * sizeof(float));
    hf = (float*) malloc(total * sizeof(float));
    if (!hl || !hm || !hn || !hf)
    {
        *status = OSKAR_ERR_MEMORY_ALLOC_FAILURE;
        goto done;
    }
    for (k = 0; k < num_freqs; ++k)
    {
        const double fk = num_freqs > 1 ? freq_min_hz + (freq_max_hz -
                freq_min_hz) * k / (num_freqs - 1) : freq_min_hz;
        int j = 0;
        for (j = 0; j < dirs_per_freq; ++j)
        {
            double x, y, z;
            i = (size_t) k * dirs_per_freq + j;
            if ((j & 3) == 3)
            {
                /* Near the beam direction (and above the horizon). */
                do
                {
                    const double r = 0.2 * sqrt(next_uniform(&rng));
                    const double t = 2.0 * M_PI * next_uniform(&rng);
                    x = l0 + r * cos(t);
                    y = m0 + r * sin(t);
                } while (x * x + y * y >= 1.0);
                z = sqrt(1.0 - x * x - y * y);
            }
            else
            {
                const double t = 2.0 * M_PI * next_uniform(&rng);
                double s;
                z = 1.0 - next_uniform(&rng);      /* (0, 1] */
                s = sqrt(1.0 - z * z);
                x = s * sin(t);
                y = s * cos(t);
            }
            hl[i] = (float) x;
            hm[i] = (float) y;
            hn[i] = (float) z;
            hf[i] = (float) fk;
        }
    }
    {
        oskar_Mem* a = oskar_mem_create_alias_from_raw(hl, OSKAR_SINGLE,
                OSKAR_CPU, total, status);
        oskar_mem_copy_contents(l, a, 0, 0, total, status);
        oskar_mem_free(a, status);
        a = oskar_mem_create_alias_from_raw(hm, OSKAR_SINGLE, OSKAR_CPU,
                total, status);
        oskar_mem_copy_contents(m, a, 0, 0, total, status);
        oskar_mem_free(a, status);
        a = oskar_mem_create_alias_from_raw(hf, OSKAR_SINGLE, OSKAR_CPU,
                total, status);
        oskar_mem_copy_contents(f, a, 0, 0, total, status);
        oskar_mem_free(a, status);
        a = oskar_mem_create_alias_from_raw(hn, OSKAR_SINGLE, OSKAR_CPU,
                total, status);
        dz = oskar_mem_create(OSKAR_SINGLE, OSKAR_GPU, total, status);
        oskar_mem_copy_contents(dz, a, 0, 0, total, status);
        oskar_mem_free(a, status);
    }
    if (*status) goto done;
    /* The exact beam, one frequency at a time. */
    out = oskar_mem_create(oskar_mem_type(beam), OSKAR_GPU,
            (size_t) dirs_per_freq, status);
    for (k = 0; k < num_freqs && !*status; ++k)
    {
        const size_t o = (size_t) k * dirs_per_freq;
        const double fk = hf[o];
        dx = oskar_mem_create_alias_from_raw(
                (float*) oskar_mem_void(l) + o, OSKAR_SINGLE, OSKAR_GPU,
                (size_t) dirs_per_freq, status);
        dy = oskar_mem_create_alias_from_raw(
                (float*) oskar_mem_void(m) + o, OSKAR_SINGLE, OSKAR_GPU,
                (size_t) dirs_per_freq, status);
        tmp = oskar_mem_create_alias_from_raw(
                (float*) oskar_mem_void(dz) + o, OSKAR_SINGLE, OSKAR_GPU,
                (size_t) dirs_per_freq, status);
        oskar_evaluate_station_beam_aperture_array(station, work,
                dirs_per_freq, dx, dy, tmp, 0, 0.0, fk, out, status);
        oskar_mem_copy_contents(beam, out, o, 0, (size_t) dirs_per_freq,
                status);
        oskar_mem_free(dx, status);
        oskar_mem_free(dy, status);
        oskar_mem_free(tmp, status);
    }
done:
    oskar_mem_free(out, status);
    oskar_mem_free(dz, status);
    free(hl);
    free(hm);
    free(hn);
    free(hf);
}

void oskar_beam_surrogate_train(oskar_BeamSurrogate* s, int num_samples,
        const oskar_Mem* l, const oskar_Mem* m, const oskar_Mem* f,
        const oskar_Mem* beam, int batch_size, int num_steps,
        double lr_start, double lr_end, unsigned int seed,
        double* loss_history, int* status)
{
    if (*status) return;
    if (!s || num_samples < 1 || num_steps < 0 || batch_size < 256 ||
            batch_size % 256)
    {
        *status = OSKAR_ERR_INVALID_ARGUMENT;
        return;
    }
    *status = check_gpu_single(l, 0, (size_t) num_samples);
    *status = *status ? *status : check_gpu_single(m, 0, (size_t) num_samples);
    *status = *status ? *status : check_gpu_single(f, 0, (size_t) num_samples);
    *status = *status ? *status :
            check_gpu_single(beam, 1, (size_t) num_samples);
    if (*status) return;
    {
        double peak = 0.0;
        *status = oskar_amd_surr_max_abs(
                (const float*) oskar_mem_void_const(beam),
                8LL * num_samples, &peak);
        if (*status) return;
        oskar_beam_surrogate_set_peak(s, peak);
    }
    *status = oskar_amd_surr_train(s->impl, num_samples,
            (const float*) oskar_mem_void_const(l),
            (const float*) oskar_mem_void_const(m),
            (const float*) oskar_mem_void_const(f),
            (const float*) oskar_mem_void_const(beam), batch_size, num_steps,
            lr_start, lr_end, (unsigned long long) seed, loss_history,
            oskar_amd_current_stream());
}

void oskar_beam_surrogate_loss_and_gradient(oskar_BeamSurrogate* s,
        int num_samples, const oskar_Mem* features, const oskar_Mem* targets,
        double* loss, oskar_Mem* gradient, int* status)
{
    if (*status) return;
    if (!s || num_samples < 256 || num_samples % 256)
    {
        *status = OSKAR_ERR_INVALID_ARGUMENT;
        return;
    }
    *status = check_gpu_single(features, 0, (size_t) num_samples * D_IN);
    *status = *status ? *status :
            check_gpu_single(targets, 0, (size_t) num_samples * 8);
    *status = *status ? *status : check_gpu_single(gradient, 0,
            (size_t) oskar_amd_surr_num_params());
    if (*status) return;
    *status = oskar_amd_surr_loss_grad(s->impl, num_samples,
            (const float*) oskar_mem_void_const(features),
            (const float*) oskar_mem_void_const(targets), loss,
            (float*) oskar_mem_void(gradient), oskar_amd_current_stream());
}

static int check_eval(int num_points, const oskar_Mem* l, const oskar_Mem* m,
        const oskar_Mem* beam)
{
    const int prec = oskar_mem_precision(beam);
    if (oskar_mem_location(l) != OSKAR_GPU ||
            oskar_mem_location(m) != OSKAR_GPU ||
            oskar_mem_location(beam) != OSKAR_GPU)
        return OSKAR_ERR_BAD_LOCATION;
    if (!oskar_mem_is_matrix(beam) || !oskar_mem_is_complex(beam) ||
            (prec != OSKAR_SINGLE && prec != OSKAR_DOUBLE))
        return OSKAR_ERR_BAD_DATA_TYPE;
    if (oskar_mem_type(l) != prec || oskar_mem_type(m) != prec)
        return OSKAR_ERR_TYPE_MISMATCH;
    if (num_points < 0 || (int) oskar_mem_length(l) < num_points ||
            (int) oskar_mem_length(m) < num_points ||
            (int) oskar_mem_length(beam) < num_points)
        return OSKAR_ERR_DIMENSION_MISMATCH;
    return 0;
}

void oskar_beam_surrogate_evaluate(oskar_BeamSurrogate* s, int num_points,
        const oskar_Mem* l, const oskar_Mem* m, double frequency_hz,
        oskar_Mem* beam, int* status)
{
    if (*status) return;
    *status = check_eval(num_points, l, m, beam);
    if (*status) return;
    *status = oskar_amd_surr_infer(s->impl, oskar_mem_precision(beam),
            num_points, oskar_mem_void_const(l), oskar_mem_void_const(m),
            frequency_hz, 0.0, oskar_mem_void(beam), 0,
            oskar_amd_current_stream());
}

void oskar_beam_surrogate_evaluate_with_fallback(oskar_BeamSurrogate* s,
        oskar_Station* station, oskar_StationWork* work, int num_points,
        const oskar_Mem* x, const oskar_Mem* y, const oskar_Mem* z,
        double frequency_hz, double threshold, oskar_Mem* beam,
        int* num_exact, int* status)
{
    oskar_Mem *flags = 0, *wk = 0, *xc = 0, *yc = 0, *zc = 0, *ex = 0;
    int prec = 0, count = 0;
    if (num_exact) *num_exact = 0;
    if (*status) return;
    *status = check_eval(num_points, x, y, beam);
    if (!*status && (oskar_mem_location(z) != OSKAR_GPU))
        *status = OSKAR_ERR_BAD_LOCATION;
    if (!*status && oskar_mem_type(z) != oskar_mem_type(x))
        *status = OSKAR_ERR_TYPE_MISMATCH;
    if (!*status && (int) oskar_mem_length(z) < num_points)
        *status = OSKAR_ERR_DIMENSION_MISMATCH;
    if (!*status && oskar_station_precision(station) !=
            oskar_mem_precision(beam))
        *status = OSKAR_ERR_TYPE_MISMATCH;
    if (*status || num_points == 0) return;
    prec = oskar_mem_precision(beam);
    flags = oskar_mem_create(OSKAR_INT, OSKAR_GPU, (size_t) num_points,
            status);
    wk = oskar_mem_create(OSKAR_INT, OSKAR_GPU, (size_t) num_points + 1,
            status);
    xc = oskar_mem_create(prec, OSKAR_GPU, (size_t) num_points, status);
    yc = oskar_mem_create(prec, OSKAR_GPU, (size_t) num_points, status);
    zc = oskar_mem_create(prec, OSKAR_GPU, (size_t) num_points, status);
    if (!*status)
        *status = oskar_amd_surr_infer(s->impl, prec, num_points,
                oskar_mem_void_const(x), oskar_mem_void_const(y),
                frequency_hz, threshold * s->peak, oskar_mem_void(beam),
                (int*) oskar_mem_void(flags), oskar_amd_current_stream());
    if (!*status)
        *status = oskar_amd_surr_compact(prec, num_points,
                (const int*) oskar_mem_void_const(flags),
                oskar_mem_void_const(x), oskar_mem_void_const(y),
                oskar_mem_void_const(z), (int*) oskar_mem_void(wk),
                oskar_mem_void(xc), oskar_mem_void(yc), oskar_mem_void(zc),
                &count, oskar_amd_current_stream());
    if (!*status && count > 0)
    {
        oskar_Mem *ax = 0, *ay = 0, *az = 0;
        ex = oskar_mem_create(oskar_mem_type(beam), OSKAR_GPU,
                (size_t) count, status);
        ax = oskar_mem_create_alias_from_raw(oskar_mem_void(xc), prec,
                OSKAR_GPU, (size_t) count, status);
        ay = oskar_mem_create_alias_from_raw(oskar_mem_void(yc), prec,
                OSKAR_GPU, (size_t) count, status);
        az = oskar_mem_create_alias_from_raw(oskar_mem_void(zc), prec,
                OSKAR_GPU, (size_t) count, status);
        oskar_evaluate_station_beam_aperture_array(station, work, count, ax,
                ay, az, 0, 0.0, frequency_hz, ex, status);
        if (!*status)
            *status = oskar_amd_surr_scatter(prec, count,
                    (const int*) oskar_mem_void_const(wk),
                    oskar_mem_void_const(ex), oskar_mem_void(beam),
                    oskar_amd_current_stream());
        oskar_mem_free(ax, status);
        oskar_mem_free(ay, status);
        oskar_mem_free(az, status);
    }
    if (num_exact) *num_exact = count;
    oskar_mem_free(ex, status);
    oskar_mem_free(flags, status);
    oskar_mem_free(wk, status);
    oskar_mem_free(xc, status);
    oskar_mem_free(yc, status);
    oskar_mem_free(zc, status);
}
