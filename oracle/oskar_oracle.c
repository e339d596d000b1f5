/*
 * CPU oracle: plain-C / OpenMP restatement of the reference's CPU path for
 * the visibility hot path.  TEST INFRASTRUCTURE ONLY (see oskar_oracle.h):
 * the checker for tests/, smoke() and the CPU-baseline leg of bench.py.
 *
 * Each function cites the reference source it restates.  Parity status of
 * each function is pinned by tests/test_oracle.py (golden vectors from the
 * compiled reference, closed forms and invariants).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>

#include "oskar_oracle.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif
#define OMEGA_EARTH 7.272205217e-5   /* define_correlate_utils.h:3 */
#define C_0 299792458.0

/* sinc(x) = sin(x)/x, 1 at 0 (oskar/utility/oskar_kernel_macros.h:281). */
static inline double sinc_d(double x) { return x == 0.0 ? 1.0 : sin(x) / x; }
static inline float sinc_f(float x) { return x == 0.0f ? 1.0f : sinf(x) / x; }

/*
 * Cross-correlation body.  One macro instance per precision, Jones kind and
 * smearing flags, so the inner loop carries no run-time branches, as the
 * reference's template does (oskar_cross_correlate_omp.cpp:27-156).
 *
 * Per baseline (SQ < SP):
 *   uu,vv,ww = (uvw_P - uvw_Q)/lambda; uu2, vv2, uv_len; uuvv = 2 uu vv;
 *   (uu,vv,ww) *= pi dnu/f                     (define_correlate_utils.h:5-14)
 *   du,dv,dw from ECEF x,y, gha0, dec0        (define_correlate_utils.h:16-26)
 *   smear = exp(-(a uu2 + b uuvv + c vv2)) sinc(uu l + vv m + ww (n-1))
 *           sinc(du l + dv m + dw (n-1))
 *   matrix: sum += smear * J_first B J_second^H with B = [[I+Q, U+iV],
 *           [U-iV, I-Q]];  scalar: sum += smear I J_first conj(J_second).
 * FP32 accumulates with Kahan compensation (oskar_kahan_sum.h:22-30).
 */
#define KAHAN(FP, SUM, VAL, GUARD) { const FP y__ = (VAL) - GUARD;            \
        const FP t__ = SUM + y__; GUARD = (t__ - SUM) - y__; SUM = t__; }

#define XCORR_BODY(NAME, FP, SINC, EXP, SQRT, SIN, COS, MATRIX, GAUSS, BS, TS,\
        KAHAN_ON)                                                             \
static void NAME(int use_casa, int num_sources, int num_stations,            \
        int offset_out, const FP* jones, const FP* I, const FP* Q,            \
        const FP* U, const FP* V, const FP* l, const FP* m, const FP* n,      \
        const FP* a, const FP* b, const FP* c, const FP* st_u,                \
        const FP* st_v, const FP* st_w, const FP* st_x, const FP* st_y,       \
        FP uv_min, FP uv_max, FP inv_wl, FP frac_bw, FP time_int, FP gha0,    \
        FP dec0, FP* vis)                                                     \
{                                                                             \
    const int NR = MATRIX ? 8 : 2;                                            \
    int SQ;                                                                   \
    _Pragma("omp parallel for schedule(dynamic, 1)")                          \
    for (SQ = 0; SQ < num_stations; ++SQ)                                     \
    {                                                                         \
        int SP;                                                               \
        for (SP = SQ + 1; SP < num_stations; ++SP)                            \
        {                                                                     \
            FP uu, vv, ww, uu2, vv2, uuvv, uv_len, du = 0, dv = 0, dw = 0;    \
            FP sum[8] = {0, 0, 0, 0, 0, 0, 0, 0};                             \
            FP guard[8] = {0, 0, 0, 0, 0, 0, 0, 0};                           \
            int i, r;                                                         \
            const FP* jp = jones + (size_t) NR * num_sources *                \
                    (size_t) (use_casa ? SQ : SP);                            \
            const FP* jq = jones + (size_t) NR * num_sources *                \
                    (size_t) (use_casa ? SP : SQ);                            \
            uu = (st_u[SP] - st_u[SQ]) * inv_wl;                              \
            vv = (st_v[SP] - st_v[SQ]) * inv_wl;                              \
            ww = (st_w[SP] - st_w[SQ]) * inv_wl;                              \
            uu2 = uu * uu; vv2 = vv * vv;                                     \
            uv_len = SQRT(uu2 + vv2);                                         \
            uuvv = 2 * uu * vv;                                               \
            { const FP f = ((FP) M_PI) * frac_bw;                             \
              uu *= f; vv *= f; ww *= f; }                                    \
            if (uv_len < uv_min || uv_len > uv_max) continue;                 \
            if (TS)                                                           \
            {                                                                 \
                FP xx, yy, rot, temp;                                         \
                const FP sin_ha = SIN(gha0), cos_ha = COS(gha0);              \
                const FP sin_dec = SIN(dec0), cos_dec = COS(dec0);            \
                temp = ((FP) M_PI) * inv_wl;                                  \
                xx = (st_x[SP] - st_x[SQ]) * temp;                            \
                yy = (st_y[SP] - st_y[SQ]) * temp;                            \
                rot = ((FP) OMEGA_EARTH) * time_int;                          \
                temp = (xx * sin_ha + yy * cos_ha) * rot;                     \
                du = (xx * cos_ha - yy * sin_ha) * rot;                       \
                dv = temp * sin_dec;                                          \
                dw = -temp * cos_dec;                                         \
            }                                                                 \
            for (i = 0; i < num_sources; ++i)                                 \
            {                                                                 \
                FP smear = (FP) 1, t[8];                                      \
                if (GAUSS)                                                    \
                    smear = EXP(-(a[i] * uu2 + b[i] * uuvv + c[i] * vv2));    \
                /* The scalar form applies I before the sinc factors         \
                 * (oskar_cross_correlate_scalar_omp.cpp:111). */            \
                if (!MATRIX) smear *= I[i];                                   \
                if (BS || TS)                                                 \
                {                                                             \
                    const FP ll = l[i], mm = m[i], nn = n[i] - (FP) 1;        \
                    if (BS) smear *= SINC(uu * ll + vv * mm + ww * nn);       \
                    if (TS) smear *= SINC(du * ll + dv * mm + dw * nn);       \
                }                                                             \
                if (MATRIX)                                                   \
                {                                                             \
                    const FP* A = jp + 8 * (size_t) i;                        \
                    const FP* Bq = jq + 8 * (size_t) i;                       \
                    const FP ba = I[i] + Q[i], bd = I[i] - Q[i];              \
                    const FP bu = U[i], bv = V[i];                            \
                    FP M[8];                                                  \
                    /* M = J_first * B, B Hermitian (define_multiply.h:69) */ \
                    M[0] = A[0] * ba + (A[2] * bu + A[3] * bv);               \
                    M[1] = A[1] * ba + (A[3] * bu - A[2] * bv);               \
                    M[2] = A[2] * bd + (A[0] * bu - A[1] * bv);               \
                    M[3] = A[3] * bd + (A[0] * bv + A[1] * bu);               \
                    M[4] = A[4] * ba + (A[6] * bu + A[7] * bv);               \
                    M[5] = A[5] * ba + (A[7] * bu - A[6] * bv);               \
                    M[6] = A[6] * bd + (A[4] * bu - A[5] * bv);               \
                    M[7] = A[7] * bd + (A[4] * bv + A[5] * bu);               \
                    /* t = M * J_second^H (define_multiply.h:54-67) */        \
                    t[0] = (M[0]*Bq[0] + M[1]*Bq[1]) + (M[2]*Bq[2] + M[3]*Bq[3]); \
                    t[1] = (M[1]*Bq[0] - M[0]*Bq[1]) + (M[3]*Bq[2] - M[2]*Bq[3]); \
                    t[2] = (M[0]*Bq[4] + M[1]*Bq[5]) + (M[2]*Bq[6] + M[3]*Bq[7]); \
                    t[3] = (M[1]*Bq[4] - M[0]*Bq[5]) + (M[3]*Bq[6] - M[2]*Bq[7]); \
                    t[4] = (M[4]*Bq[0] + M[5]*Bq[1]) + (M[6]*Bq[2] + M[7]*Bq[3]); \
                    t[5] = (M[5]*Bq[0] - M[4]*Bq[1]) + (M[7]*Bq[2] - M[6]*Bq[3]); \
                    t[6] = (M[4]*Bq[4] + M[5]*Bq[5]) + (M[6]*Bq[6] + M[7]*Bq[7]); \
                    t[7] = (M[5]*Bq[4] - M[4]*Bq[5]) + (M[7]*Bq[6] - M[6]*Bq[7]); \
                }                                                             \
                else                                                          \
                {                                                             \
                    const FP* A = jp + 2 * (size_t) i;                        \
                    const FP* Bq = jq + 2 * (size_t) i;                       \
                    t[0] = A[0] * Bq[0] + A[1] * Bq[1];                       \
                    t[1] = A[1] * Bq[0] - A[0] * Bq[1];                       \
                }                                                             \
                for (r = 0; r < NR; ++r)                                      \
                {                                                             \
                    if (KAHAN_ON) KAHAN(FP, sum[r], t[r] * smear, guard[r])   \
                    else sum[r] += t[r] * smear;                              \
                }                                                             \
            }                                                                 \
            {                                                                 \
                const int bl = SQ * (num_stations - 1) - (SQ - 1) * SQ / 2 +  \
                        SP - SQ - 1;                                          \
                FP* out = vis + (size_t) NR * (size_t) (offset_out + bl);     \
                for (r = 0; r < NR; ++r) out[r] += sum[r];                    \
            }                                                                 \
        }                                                                     \
    }                                                                         \
}

#define XCORR_FAMILY(P, FP, SINC, EXP, SQRT, SIN, COS, K)                     \
XCORR_BODY(xc_##P##_m000, FP, SINC, EXP, SQRT, SIN, COS, 1, 0, 0, 0, K)       \
XCORR_BODY(xc_##P##_m001, FP, SINC, EXP, SQRT, SIN, COS, 1, 0, 0, 1, K)       \
XCORR_BODY(xc_##P##_m010, FP, SINC, EXP, SQRT, SIN, COS, 1, 0, 1, 0, K)       \
XCORR_BODY(xc_##P##_m011, FP, SINC, EXP, SQRT, SIN, COS, 1, 0, 1, 1, K)       \
XCORR_BODY(xc_##P##_m100, FP, SINC, EXP, SQRT, SIN, COS, 1, 1, 0, 0, K)       \
XCORR_BODY(xc_##P##_m101, FP, SINC, EXP, SQRT, SIN, COS, 1, 1, 0, 1, K)       \
XCORR_BODY(xc_##P##_m110, FP, SINC, EXP, SQRT, SIN, COS, 1, 1, 1, 0, K)       \
XCORR_BODY(xc_##P##_m111, FP, SINC, EXP, SQRT, SIN, COS, 1, 1, 1, 1, K)       \
XCORR_BODY(xc_##P##_s000, FP, SINC, EXP, SQRT, SIN, COS, 0, 0, 0, 0, K)       \
XCORR_BODY(xc_##P##_s001, FP, SINC, EXP, SQRT, SIN, COS, 0, 0, 0, 1, K)       \
XCORR_BODY(xc_##P##_s010, FP, SINC, EXP, SQRT, SIN, COS, 0, 0, 1, 0, K)       \
XCORR_BODY(xc_##P##_s011, FP, SINC, EXP, SQRT, SIN, COS, 0, 0, 1, 1, K)       \
XCORR_BODY(xc_##P##_s100, FP, SINC, EXP, SQRT, SIN, COS, 0, 1, 0, 0, K)       \
XCORR_BODY(xc_##P##_s101, FP, SINC, EXP, SQRT, SIN, COS, 0, 1, 0, 1, K)       \
XCORR_BODY(xc_##P##_s110, FP, SINC, EXP, SQRT, SIN, COS, 0, 1, 1, 0, K)       \
XCORR_BODY(xc_##P##_s111, FP, SINC, EXP, SQRT, SIN, COS, 0, 1, 1, 1, K)

XCORR_FAMILY(d, double, sinc_d, exp, sqrt, sin, cos, 0)
XCORR_FAMILY(f, float, sinc_f, expf, sqrtf, sinf, cosf, 1)

typedef void (*xc_fn_d)(int, int, int, int, const double*, const double*,
        const double*, const double*, const double*, const double*,
        const double*, const double*, const double*, const double*,
        const double*, const double*, const double*, const double*,
        const double*, const double*, double, double, double, double, double,
        double, double, double*);
typedef void (*xc_fn_f)(int, int, int, int, const float*, const float*,
        const float*, const float*, const float*, const float*, const float*,
        const float*, const float*, const float*, const float*, const float*,
        const float*, const float*, const float*, const float*, float, float,
        float, float, float, float, float, float*);

static const xc_fn_d XC_D[16] = {
    xc_d_s000, xc_d_s001, xc_d_s010, xc_d_s011,
    xc_d_s100, xc_d_s101, xc_d_s110, xc_d_s111,
    xc_d_m000, xc_d_m001, xc_d_m010, xc_d_m011,
    xc_d_m100, xc_d_m101, xc_d_m110, xc_d_m111 };
static const xc_fn_f XC_F[16] = {
    xc_f_s000, xc_f_s001, xc_f_s010, xc_f_s011,
    xc_f_s100, xc_f_s101, xc_f_s110, xc_f_s111,
    xc_f_m000, xc_f_m001, xc_f_m010, xc_f_m011,
    xc_f_m100, xc_f_m101, xc_f_m110, xc_f_m111 };

/* Derived scalars (oskar_cross_correlate.c:42-71). */
static void derived(double channel_bw, double time_avg_in, int is_azel,
        double ra0, double dec0_in, double uv_min_in, double uv_max_in,
        int uv_wavelengths, double gast, double freq, double* inv_wl,
        double* frac_bw, double* time_avg, double* gha0, double* dec0,
        double* uv_min, double* uv_max)
{
    freq = fabs(freq);
    *inv_wl = freq / C_0;
    *frac_bw = channel_bw / freq;
    *time_avg = 0.0; *gha0 = 0.0; *dec0 = 0.0;
    if (!is_azel)
    {
        *time_avg = time_avg_in;
        *gha0 = gast - ra0;
        *dec0 = dec0_in;
    }
    *uv_min = uv_min_in;
    *uv_max = uv_max_in;
    if (!uv_wavelengths)
    {
        *uv_min *= *inv_wl;
        *uv_max *= *inv_wl;
    }
    if (*uv_max < 0.0 || *uv_max > FLT_MAX) *uv_max = FLT_MAX;
}

void oracle_cross_correlate_d(int use_casa, int source_type, int matrix,
        int num_sources, int num_stations, const double* jones,
        const double* I, const double* Q, const double* U, const double* V,
        const double* l, const double* m, const double* n,
        const double* a, const double* b, const double* c,
        const double* st_u, const double* st_v, const double* st_w,
        const double* st_x, const double* st_y,
        double channel_bandwidth_hz, double time_average_sec,
        int phase_centre_is_azel, double ra0, double dec0_in, double uv_min_in,
        double uv_max_in, int uv_units_wavelengths, double gast,
        double frequency_hz, int offset_out, double* vis)
{
    double inv_wl, frac_bw, time_avg, gha0, dec0, uv_min, uv_max;
    derived(channel_bandwidth_hz, time_average_sec, phase_centre_is_azel,
            ra0, dec0_in, uv_min_in, uv_max_in, uv_units_wavelengths, gast,
            frequency_hz, &inv_wl, &frac_bw, &time_avg, &gha0, &dec0, &uv_min,
            &uv_max);
    const int g = (source_type == 1), bs = (frac_bw != 0.0),
            ts = (time_avg != 0.0);
    XC_D[(matrix ? 8 : 0) + g * 4 + bs * 2 + ts](use_casa, num_sources,
            num_stations, offset_out, jones, I, Q, U, V, l, m, n, a, b, c,
            st_u, st_v, st_w, st_x, st_y, uv_min, uv_max, inv_wl, frac_bw,
            time_avg, gha0, dec0, vis);
}

void oracle_cross_correlate_f(int use_casa, int source_type, int matrix,
        int num_sources, int num_stations, const float* jones,
        const float* I, const float* Q, const float* U, const float* V,
        const float* l, const float* m, const float* n,
        const float* a, const float* b, const float* c,
        const float* st_u, const float* st_v, const float* st_w,
        const float* st_x, const float* st_y,
        double channel_bandwidth_hz, double time_average_sec,
        int phase_centre_is_azel, double ra0, double dec0_in, double uv_min_in,
        double uv_max_in, int uv_units_wavelengths, double gast,
        double frequency_hz, int offset_out, float* vis)
{
    double inv_wl, frac_bw, time_avg, gha0, dec0, uv_min, uv_max;
    derived(channel_bandwidth_hz, time_average_sec, phase_centre_is_azel,
            ra0, dec0_in, uv_min_in, uv_max_in, uv_units_wavelengths, gast,
            frequency_hz, &inv_wl, &frac_bw, &time_avg, &gha0, &dec0, &uv_min,
            &uv_max);
    const int g = (source_type == 1), bs = ((float) frac_bw != 0.0f),
            ts = ((float) time_avg != 0.0f);
    XC_F[(matrix ? 8 : 0) + g * 4 + bs * 2 + ts](use_casa, num_sources,
            num_stations, offset_out, jones, I, Q, U, V, l, m, n, a, b, c,
            st_u, st_v, st_w, st_x, st_y, (float) uv_min, (float) uv_max,
            (float) inv_wl, (float) frac_bw, (float) time_avg, (float) gha0,
            (float) dec0, vis);
}

/*
 * Weighted DFT (define_dftw_c2c.h:79-109, define_dftw_m2m.h:109-167): the
 * element loop in input order, one accumulator per output component.
 */
#define DFTW_BODY(NAME, FP, SIN, COS)                                         \
void NAME(int matrix, int normalise, int num_in, double wavenumber_d,        \
        const FP* weights, const FP* x_in, const FP* y_in, const FP* z_in,    \
        int offset_coord_out, int num_out, const FP* x_out, const FP* y_out,  \
        const FP* z_out, const int* data_idx, const FP* data, int eval_x,     \
        int eval_y, int offset_out, FP* output)                               \
{                                                                             \
    const FP wavenumber = (FP) wavenumber_d;                                  \
    const FP norm = (FP) (normalise ? 1.0 / num_in : 1.0);                    \
    const int is_3d = (z_in != 0 && z_out != 0);                              \
    int j;                                                                    \
    _Pragma("omp parallel for")                                               \
    for (j = 0; j < num_out; ++j)                                             \
    {                                                                         \
        FP acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};                                 \
        const FP xo = wavenumber * x_out[j + offset_coord_out];               \
        const FP yo = wavenumber * y_out[j + offset_coord_out];               \
        const FP zo = is_3d ? wavenumber * z_out[j + offset_coord_out] : 0;   \
        int i, k;                                                             \
        for (i = 0; i < num_in; ++i)                                          \
        {                                                                     \
            FP t = xo * x_in[i] + yo * y_in[i];                               \
            if (is_3d) t += zo * z_in[i];                                     \
            FP im = SIN(t), re = COS(t);                                      \
            const FP wr = weights[2 * i], wi = weights[2 * i + 1];            \
            const FP tr = re;                                                 \
            re = re * wr - wi * im;                                           \
            im = im * wr + wi * tr;                                           \
            const int row = data_idx ? data_idx[i] : i;                       \
            if (matrix)                                                       \
            {                                                                 \
                const FP* d = data + 8 * ((size_t) row * num_out + j);        \
                for (k = 0; k < 4; ++k)                                       \
                {                                                             \
                    if ((k < 2 && !eval_x) || (k >= 2 && !eval_y)) continue;  \
                    acc[2 * k] += d[2 * k] * re - d[2 * k + 1] * im;          \
                    acc[2 * k + 1] += d[2 * k + 1] * re + d[2 * k] * im;      \
                }                                                             \
            }                                                                 \
            else                                                              \
            {                                                                 \
                const FP* d = data + 2 * ((size_t) row * num_out + j);        \
                acc[0] += d[0] * re - d[1] * im;                              \
                acc[1] += d[1] * re + d[0] * im;                              \
            }                                                                 \
        }                                                                     \
        if (matrix)                                                           \
        {                                                                     \
            FP* o = output + 8 * (size_t) (j + offset_out);                   \
            for (k = 0; k < 4; ++k)                                           \
            {                                                                 \
                if ((k < 2 && !eval_x) || (k >= 2 && !eval_y)) continue;      \
                o[2 * k] = acc[2 * k] * norm;                                 \
                o[2 * k + 1] = acc[2 * k + 1] * norm;                         \
            }                                                                 \
        }                                                                     \
        else                                                                  \
        {                                                                     \
            FP* o = output + 2 * (size_t) (j + offset_out);                   \
            o[0] = acc[0] * norm;                                             \
            o[1] = acc[1] * norm;                                             \
        }                                                                     \
    }                                                                         \
}

DFTW_BODY(oracle_dftw_d, double, sin, cos)
DFTW_BODY(oracle_dftw_f, float, sinf, cosf)

/* define_evaluate_element_weights_dft.h:3-16 */
void oracle_element_weights_dft_d(int n, const double* x, const double* y,
        const double* z, const double* cable, double wavenumber, double x1,
        double y1, double z1, double* weights)
{
    int i;
    for (i = 0; i < n; ++i)
    {
        const double p = wavenumber * (x[i] * x1 + y[i] * y1 + z[i] * z1 +
                (cable ? cable[i] : 0.0));
        weights[2 * i] = cos(-p);
        weights[2 * i + 1] = sin(-p);
    }
}

void oracle_element_weights_dft_f(int n, const float* x, const float* y,
        const float* z, const float* cable, double wavenumber_d, double x1_d,
        double y1_d, double z1_d, float* weights)
{
    const float wavenumber = (float) wavenumber_d, x1 = (float) x1_d,
            y1 = (float) y1_d, z1 = (float) z1_d;
    int i;
    for (i = 0; i < n; ++i)
    {
        const float p = wavenumber * (x[i] * x1 + y[i] * y1 + z[i] * z1 +
                (cable ? cable[i] : 0.0f));
        weights[2 * i] = cosf(-p);
        weights[2 * i + 1] = sinf(-p);
    }
}

/* define_convert_enu_directions_to_theta_phi.h:3-31 */
void oracle_theta_phi_d(int off_in, int num, const double* x, const double* y,
        const double* z, int extra_point_at_pole, double delta_phi1,
        double delta_phi2, double* theta, double* phi1, double* phi2)
{
    const double twopi = 2.0 * M_PI;
    int i;
    for (i = 0; i < num; ++i)
    {
        const double xx = x[i + off_in], yy = y[i + off_in], zz = z[i + off_in];
        const double p0 = atan2(yy, xx);
        double p1 = fmod(p0 + delta_phi1, twopi);
        double p2 = fmod(p0 + delta_phi2, twopi);
        if (p1 < 0) p1 += twopi;
        if (p2 < 0) p2 += twopi;
        theta[i] = atan2(sqrt(xx * xx + yy * yy), zz);
        phi1[i] = p1;
        phi2[i] = p2;
    }
    if (extra_point_at_pole)
    {
        theta[num] = 0.0;
        phi1[num] = delta_phi1;
        phi2[num] = delta_phi2;
    }
}

/* define_evaluate_dipole_pattern.h:3-76 */
static void dipole_d(double phi, double sin_theta, double cos_theta,
        double kL, double cos_kL, double* e_theta, double* e_phi)
{
    const double sin_phi = sin(phi), cos_phi = cos(phi);
    const double denom = 1.0 + cos_phi * cos_phi * (cos_theta * cos_theta - 1.0);
    e_theta[1] = e_phi[1] = 0.0;
    if (denom == 0.0) { e_theta[0] = e_phi[0] = 0.0; return; }
    {
        const double t = (cos(kL * cos_phi * sin_theta) - cos_kL) / denom;
        e_theta[0] = -cos_phi * cos_theta * t;
        e_phi[0] = sin_phi * t;
    }
}

void oracle_dipole_pattern_d(int matrix, int n, const double* theta,
        const double* phi_x, const double* phi_y, double kL, double cos_kL,
        int swap_xy, int offset_out, double* pattern)
{
    int i;
    for (i = 0; i < n; ++i)
    {
        const double st = sin(theta[i]), ct = cos(theta[i]);
        double xt[2], xp[2], yt[2], yp[2];
        dipole_d(phi_x[i], st, ct, kL, cos_kL, xt, xp);
        dipole_d(phi_y[i], st, ct, kL, cos_kL, yt, yp);
        if (matrix)
        {
            double* o = pattern + 8 * (size_t) (i + offset_out);
            const double* q[4];
            int k;
            if (swap_xy) { q[0] = yt; q[1] = yp; q[2] = xt; q[3] = xp; }
            else { q[0] = xt; q[1] = xp; q[2] = yt; q[3] = yp; }
            for (k = 0; k < 4; ++k) { o[2 * k] = q[k][0]; o[2 * k + 1] = q[k][1]; }
        }
        else
        {
            double* o = pattern + 2 * (size_t) (i + offset_out);
            o[0] = sqrt((xt[0] * xt[0] + xp[0] * xp[0] + yt[0] * yt[0] +
                    yp[0] * yp[0]) / 2.0);
            o[1] = 0.0;
        }
    }
}

/* oskar_mem_normalise (define_mem_normalise.h): scale all but idx. */
static void normalise_d(int nr, double* a, int offset, int n, int idx)
{
    const double* v = a + (size_t) nr * idx;
    double amp = 0.0, scal;
    int i, r;
    for (r = 0; r < nr; ++r) amp += v[r] * v[r];
    if (nr == 8) amp /= 2.0;
    scal = 1.0 / sqrt(amp);
    for (i = offset; i < offset + n; ++i)
        if (i != idx) for (r = 0; r < nr; ++r) a[(size_t) nr * i + r] *= scal;
}

int oracle_station_beam_aa_d(int matrix, int num_elements,
        const double* x_true, const double* y_true, const double* z_true,
        const double* x_meas, const double* y_meas, const double* z_meas,
        const double* cable, const int* element_types, int num_types,
        const int* type_kind, const double* dipole_length_m,
        double euler_x0, double euler_y0, int swap_xy, int is_3d,
        int norm_array, int norm_element, int enable_array,
        double beam_x, double beam_y, double beam_z,
        int num_points, const double* x, const double* y, const double* z,
        double frequency_hz, const double* err, const double* apod,
        double* beam)
{
    const int nr = matrix ? 8 : 2;
    const double wavenumber = 2.0 * M_PI * frequency_hz / C_0;
    const size_t sig_len = (size_t) num_types * (size_t) (num_points + 1);
    double* signal = (double*) calloc(sig_len * nr, sizeof(double));
    double* theta = (double*) calloc((size_t) num_points + 1, sizeof(double));
    double* phx = (double*) calloc((size_t) num_points + 1, sizeof(double));
    double* phy = (double*) calloc((size_t) num_points + 1, sizeof(double));
    double* w = (double*) calloc(2 * (size_t) (num_elements > 0 ?
            num_elements : 1), sizeof(double));
    int t, i;
    if (!signal || !theta || !phx || !phy || !w) return -1;
    /* Element patterns per type (oskar_element_evaluate.c:22-170) with
     * orientation_x = euler_x0 + pi/2, orientation_y = euler_y0 and
     * delta_phi = pi/2 - orientation. */
    for (t = 0; t < num_types; ++t)
    {
        const int off = t * num_points;
        const int npn = norm_element ? num_points + 1 : num_points;
        oracle_theta_phi_d(0, num_points, x, y, z, norm_element,
                M_PI / 2 - (euler_x0 + M_PI / 2), M_PI / 2 - euler_y0,
                theta, phx, phy);
        if (type_kind[t] == 1)
        {
            for (i = 0; i < npn; ++i)
            {
                double* o = signal + (size_t) nr * (off + i);
                memset(o, 0, nr * sizeof(double));
                o[0] = 1.0;
                if (matrix) o[6] = 1.0;
            }
        }
        else
        {
            const double kL = dipole_length_m[t] * (M_PI * frequency_hz / C_0);
            oracle_dipole_pattern_d(matrix, npn, theta, phx, phy, kL, cos(kL),
                    swap_xy, off, signal);
        }
        if (norm_element)
            normalise_d(nr, signal, off, num_points, off + num_points);
    }
    if (enable_array)
    {
        oracle_element_weights_dft_d(num_elements, x_meas, y_meas, z_meas,
                cable, wavenumber, beam_x, beam_y, beam_z, w);
        /* Element error factors (oracle_element_errors_d), then
         * apodisation (oskar_station_evaluate_element_weights.c:34-64):
         * complex products in oskar_mem_multiply's term order. */
        for (i = 0; i < num_elements; ++i)
        {
            double a = w[2 * i], b = w[2 * i + 1];
            if (err)
            {
                const double ex = err[2 * i], ey = err[2 * i + 1];
                const double a2 = a * ex - b * ey, b2 = a * ey + b * ex;
                a = a2; b = b2;
            }
            if (apod)
            {
                const double ax = apod[2 * i], ay = apod[2 * i + 1];
                const double a2 = a * ax - b * ay, b2 = a * ay + b * ax;
                a = a2; b = b2;
            }
            w[2 * i] = a;
            w[2 * i + 1] = b;
        }
        oracle_dftw_d(matrix, norm_array, num_elements, wavenumber, w,
                x_true, y_true, is_3d ? z_true : 0, 0, num_points, x, y,
                is_3d ? z : 0, element_types, signal, 1, 1, 0, beam);
    }
    else
    {
        if (num_types != 1) { free(signal); free(theta); free(phx);
            free(phy); free(w); return -23; }
        memcpy(beam, signal, (size_t) nr * num_points * sizeof(double));
    }
    /* Blank below the horizon (define_blank_below_horizon.h). */
    for (i = 0; i < num_points; ++i)
        if (z[i] < 0.0) memset(beam + (size_t) nr * i, 0, nr * sizeof(double));
    free(signal); free(theta); free(phx); free(phy); free(w);
    return 0;
}

/*
 * Auto-correlation, oskar/correlate/define_auto_correlate.h:32-58 (matrix)
 * and :84-104 (scalar), CPU bodies: per station s,
 *   matrix: sum = sum_i J_s(i) B_i J_s(i)^H (M = J B with B Hermitian, then
 *           M J^H, as the cross-correlation body), plain sum in FP;
 *           the imaginary parts of sum.a and sum.d are set to zero, then
 *           vis[offset_out + s] += sum;
 *   scalar: sum = sum_i |J_s(i)|^2 I_i; vis[offset_out + s].re += sum.
 */
#define ACORR_BODY(NAME, FP)                                                  \
void NAME(int matrix, int num_sources, int num_stations, int offset_out,     \
        const FP* jones, const FP* I, const FP* Q, const FP* U, const FP* V,  \
        FP* vis)                                                              \
{                                                                             \
    int s;                                                                    \
    _Pragma("omp parallel for")                                               \
    for (s = 0; s < num_stations; ++s)                                        \
    {                                                                         \
        FP sum[8] = {0, 0, 0, 0, 0, 0, 0, 0};                                 \
        int i, r;                                                             \
        if (!matrix)                                                          \
        {                                                                     \
            const FP* J = jones + 2 * (size_t) num_sources * s;               \
            for (i = 0; i < num_sources; ++i)                                 \
                sum[0] += (J[2 * i] * J[2 * i] + J[2 * i + 1] *              \
                        J[2 * i + 1]) * I[i];                                 \
            vis[2 * ((size_t) offset_out + s)] += sum[0];                     \
            continue;                                                         \
        }                                                                     \
        for (i = 0; i < num_sources; ++i)                                     \
        {                                                                     \
            const FP* A = jones + 8 * ((size_t) num_sources * s + i);         \
            const FP ba = I[i] + Q[i], bd = I[i] - Q[i];                      \
            const FP bu = U[i], bv = V[i];                                    \
            FP M[8];                                                          \
            M[0] = A[0] * ba + (A[2] * bu + A[3] * bv);                       \
            M[1] = A[1] * ba + (A[3] * bu - A[2] * bv);                       \
            M[2] = A[2] * bd + (A[0] * bu - A[1] * bv);                       \
            M[3] = A[3] * bd + (A[0] * bv + A[1] * bu);                       \
            M[4] = A[4] * ba + (A[6] * bu + A[7] * bv);                       \
            M[5] = A[5] * ba + (A[7] * bu - A[6] * bv);                       \
            M[6] = A[6] * bd + (A[4] * bu - A[5] * bv);                       \
            M[7] = A[7] * bd + (A[4] * bv + A[5] * bu);                       \
            sum[0] += (M[0]*A[0] + M[1]*A[1]) + (M[2]*A[2] + M[3]*A[3]);      \
            sum[1] += (M[1]*A[0] - M[0]*A[1]) + (M[3]*A[2] - M[2]*A[3]);      \
            sum[2] += (M[0]*A[4] + M[1]*A[5]) + (M[2]*A[6] + M[3]*A[7]);      \
            sum[3] += (M[1]*A[4] - M[0]*A[5]) + (M[3]*A[6] - M[2]*A[7]);      \
            sum[4] += (M[4]*A[0] + M[5]*A[1]) + (M[6]*A[2] + M[7]*A[3]);      \
            sum[5] += (M[5]*A[0] - M[4]*A[1]) + (M[7]*A[2] - M[6]*A[3]);      \
            sum[6] += (M[4]*A[4] + M[5]*A[5]) + (M[6]*A[6] + M[7]*A[7]);      \
            sum[7] += (M[5]*A[4] - M[4]*A[5]) + (M[7]*A[6] - M[6]*A[7]);      \
        }                                                                     \
        sum[1] = 0; sum[7] = 0;                                               \
        for (r = 0; r < 8; ++r)                                               \
            vis[8 * ((size_t) offset_out + s) + r] += sum[r];                 \
    }                                                                         \
}

ACORR_BODY(oracle_auto_correlate_d, double)
ACORR_BODY(oracle_auto_correlate_f, float)

/*
 * Horizon mask (oskar/sky/src/oskar_update_horizon_mask.c:30-60 and
 * define_update_horizon_mask.h:3-12): (ll, mm, nn) in double, rounded to
 * float for the float body; mask[i] |= l ll + m mm + n nn > 0.
 */
static void horizon_factors_d(double ha0, double dec0, double lat, double* f)
{
    const double cos_ha0 = cos(ha0), sin_dec0 = sin(dec0),
            cos_dec0 = cos(dec0), sin_lat = sin(lat), cos_lat = cos(lat);
    f[0] = cos_lat * sin(ha0);
    f[1] = sin_lat * cos_dec0 - cos_lat * cos_ha0 * sin_dec0;
    f[2] = sin_lat * sin_dec0 + cos_lat * cos_ha0 * cos_dec0;
}

void oracle_update_horizon_mask_d(int num, const double* l, const double* m,
        const double* n, double ha0, double dec0, double lat, int* mask)
{
    double f[3];
    int i;
    horizon_factors_d(ha0, dec0, lat, f);
    for (i = 0; i < num; ++i)
        mask[i] |= ((l[i] * f[0] + m[i] * f[1] + n[i] * f[2]) > 0.0);
}

void oracle_update_horizon_mask_f(int num, const float* l, const float* m,
        const float* n, double ha0, double dec0, double lat, int* mask)
{
    double f[3];
    int i;
    horizon_factors_d(ha0, dec0, lat, f);
    {
        const float a = (float) f[0], b = (float) f[1], c = (float) f[2];
        for (i = 0; i < num; ++i)
            mask[i] |= ((l[i] * a + m[i] * b + n[i] * c) > 0.0f);
    }
}

/* Exclusive prefix sum with the total after the last value
 * (oskar/math/src/oskar_prefix_sum.c:45-57). */
void oracle_prefix_sum_int(int num, const int* in, int* out)
{
    int i, sum = 0;
    for (i = 0; i < num; ++i)
    {
        const int x = in[i];
        out[i] = sum;
        sum += x;
    }
    out[num] = sum;
}

/* Masked row copy of the sky table (the CPU branch of
 * oskar/sky/src/oskar_sky_copy_source_data.c:52-109): rows with mask != 0
 * go, in order, to the next output row; nfix valid-column counts per row.
 * Returns the number of rows copied. */
#define SKY_COPY_BODY(NAME, FP)                                               int NAME(int num, int cap_in, int cap_out, const int* mask, int num_col,              const FP* table_in, FP* table_out, int nfix, const int* valid_in,             int* valid_out)                                                       {                                                                                 int i, c, r = 0;                                                              for (i = 0; i < num; ++i)                                                     {                                                                                 if (!mask[i]) continue;                                                       for (c = 0; c < num_col; ++c)                                                     table_out[(size_t) c * cap_out + r] =                                                 table_in[(size_t) c * cap_in + i];                                memcpy(valid_out + (size_t) nfix * r, valid_in + (size_t) nfix * i,                  (size_t) nfix * sizeof(int));                                         r++;                                                                      }                                                                             return r;                                                                 }
SKY_COPY_BODY(oracle_sky_copy_source_data_d, double)
SKY_COPY_BODY(oracle_sky_copy_source_data_f, float)

/* define_evaluate_jones_K.h:56-75 (CPU body) */
void oracle_evaluate_jones_K_d(int num_sources, const double* l,
        const double* m, const double* n, int num_stations, const double* u,
        const double* v, const double* w, double wavenumber,
        const double* I, double source_filter_min, double source_filter_max,
        double* jones)
{
    int a;
    _Pragma("omp parallel for")
    for (a = 0; a < num_stations; ++a)
    {
        int s;
        for (s = 0; s < num_sources; ++s)
        {
            double re = 0.0, im = 0.0;
            if (I[s] > source_filter_min && I[s] <= source_filter_max)
            {
                double phase = u[a] * l[s] + v[a] * m[s];
                phase += w[a] * (n[s] - 1.0);
                phase *= wavenumber;
                re = cos(phase);
                im = sin(phase);
            }
            jones[2 * ((size_t) s + (size_t) num_sources * a)] = re;
            jones[2 * ((size_t) s + (size_t) num_sources * a) + 1] = im;
        }
    }
}

void oracle_factorised_subset_ld(int num_sources, const double* l,
        const double* m, const double* n, const double* I, const double* E,
        int num_stations, const double* u, const double* v, const double* w,
        double wavenumber, int num_baselines, const int* p, const int* q,
        double* vis_out)
{
    int k;
    (void) num_stations;
    _Pragma("omp parallel for schedule(dynamic, 1)")
    for (k = 0; k < num_baselines; ++k)
    {
        long double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        const int P = p[k], Q = q[k];
        const double du = u[P] - u[Q], dv = v[P] - v[Q], dw = w[P] - w[Q];
        int s, r;
        for (s = 0; s < num_sources; ++s)
        {
            /* K_p K_q^* = exp(i k ((u_p - u_q) l + (v_p - v_q) m
             *   + (w_p - w_q)(n - 1))); V += I K_p K_q^* E E^H. */
            const long double ph = (long double) wavenumber *
                    ((long double) du * l[s] + (long double) dv * m[s] +
                     (long double) dw * ((long double) n[s] - 1.0L));
            const long double kr = cosl(ph), ki = sinl(ph);
            const double* e = E + 8 * (size_t) s;
            long double eeh[8];
            /* E E^H: (E E^H)_{ij} = sum_k E_ik conj(E_jk) */
            int i2, j2;
            for (i2 = 0; i2 < 2; ++i2)
                for (j2 = 0; j2 < 2; ++j2)
                {
                    long double re = 0, im = 0;
                    int kk;
                    for (kk = 0; kk < 2; ++kk)
                    {
                        const long double ar = e[2 * (2 * i2 + kk)];
                        const long double ai = e[2 * (2 * i2 + kk) + 1];
                        const long double br = e[2 * (2 * j2 + kk)];
                        const long double bi = e[2 * (2 * j2 + kk) + 1];
                        re += ar * br + ai * bi;
                        im += ai * br - ar * bi;
                    }
                    eeh[2 * (2 * i2 + j2)] = re;
                    eeh[2 * (2 * i2 + j2) + 1] = im;
                }
            for (r = 0; r < 4; ++r)
            {
                const long double er = eeh[2 * r], ei = eeh[2 * r + 1];
                acc[2 * r] += I[s] * (kr * er - ki * ei);
                acc[2 * r + 1] += I[s] * (kr * ei + ki * er);
            }
        }
        for (r = 0; r < 8; ++r) vis_out[8 * (size_t) k + r] = (double) acc[r];
    }
}

/*
 * Relative direction cosines of sky positions (the 3D CPU body of
 * oskar/convert/define_convert_lon_lat_to_relative_directions.h:3-17, as
 * called by oskar/sky/src/oskar_sky_evaluate_relative_directions.c:20-37
 * with lon0 = ra0, cos_lat0 = cos(dec0), sin_lat0 = sin(dec0)):
 *   l = cos(lat) sin(lon - lon0)
 *   m = cos(lat0) sin(lat) - sin(lat0) cos(lat) cos(lon - lon0)
 *   n = sin(lat0) sin(lat) + cos(lat0) cos(lat) cos(lon - lon0)
 * The reference compiles these macro bodies as C99 with <tgmath.h>
 * (oskar/utility/oskar_kernel_macros.h:207-209), so its float instance calls
 * sinf and cosf: TSIN and TCOS name them here.
 */
#define REL_DIR_BODY(NAME, FP, TSIN, TCOS)                                    \
void NAME(int num, const FP* lon_rad, const FP* lat_rad, double lon0,         \
        double lat0, FP* l, FP* m, FP* n)                                     \
{                                                                             \
    int i;                                                                    \
    const FP lon0_ = (FP) lon0, cos_lat0 = (FP) cos(lat0),                    \
            sin_lat0 = (FP) sin(lat0);                                        \
    for (i = 0; i < num; ++i)                                                 \
    {                                                                         \
        FP sin_lon, cos_lon, sin_lat, cos_lat;                                \
        const FP lon = lon_rad[i] - lon0_, lat = lat_rad[i];                  \
        sin_lon = TSIN(lon); cos_lon = TCOS(lon);                             \
        sin_lat = TSIN(lat); cos_lat = TCOS(lat);                             \
        l[i] = cos_lat * sin_lon;                                             \
        m[i] = cos_lat0 * sin_lat - sin_lat0 * cos_lat * cos_lon;             \
        n[i] = sin_lat0 * sin_lat + cos_lat0 * cos_lat * cos_lon;             \
    }                                                                         \
}
REL_DIR_BODY(oracle_relative_directions_d, double, sin, cos)
REL_DIR_BODY(oracle_relative_directions_f, float, sinf, cosf)

/*
 * Gaussian source parameters (oskar/sky/src/
 * oskar_sky_evaluate_gaussian_source_parameters.c:32-262).  Per source with
 * major or minor FWHM non-zero:
 *  1. six points of the ellipse (FWHM/2 semi-axes, position angle PA) on
 *     the tangent plane at (0, 0), t_j = 60 j degrees (:99-140);
 *  2. tangent plane -> lon/lat at (0, 0) -> xyz, rotated on the sphere to
 *     the source's (RA, Dec), -> lon/lat, -> tangent plane at the phase
 *     centre (:141-149; oskar/math/src/oskar_rotate.c:31-48,
 *     oskar/convert/src/oskar_convert_lon_lat_to_xyz.c,
 *     oskar_convert_xyz_to_lon_lat.c, and the two 2D macro bodies);
 *  3. least-squares conic a x^2 + b xy + c y^2 + d x + e y = 1 through the
 *     mean-removed points by the normal equations (X^T X) p = X^T 1, solved
 *     by LU with partial pivoting (oskar/math/src/oskar_fit_ellipse.c:17-
 *     131; oskar_matrix_multiply.c with transpose_a; oskar_lapack_subset.c
 *     getrf2/getrs).  The recursive getrf2 applies the same rank-1 updates
 *     in the same column order as this right-looking loop (the reference's
 *     gemm adds (-u) l, the same rounding as subtracting u l), so the
 *     factors round identically.  A zero pivot, or a*c <= 0 after the
 *     rotation that removes b, marks the fit as failed;
 *  4. a, b, c of the Gaussian in the uv plane from the fitted FWHMs and PA
 *     (:170-178 and :249-257).
 * NSIN/NCOS/NATAN2/NSQRT/NASIN are the libm functions of the precision:
 * the reference's per-precision helpers call sinf etc. in their float
 * versions, and its two 2D conversion macro bodies are compiled with
 * <tgmath.h>, which picks the same float functions for float arguments.
 * The ellipse fit and the loop body of the sky function use the double
 * functions in both precisions, as the reference's plain C does.  A failed
 * source gets failed[i] = 1 and keeps its a, b, c; a source with both axes
 * zero is skipped (failed[i] = 0, a, b, c untouched).
 */
#define GAUSS_PARAMS_BODY(NAME, FP, FP_MIN, NSIN, NCOS, NATAN2, NSQRT,  \
        NASIN)                                                                \
static int NAME##_fit(FP* major, FP* minor, FP* pa_out, const FP* x,          \
        const FP* y)                                                          \
{                                                                             \
    int i, j, k, ok = 1;                                                      \
    FP mean_x = 0, mean_y = 0, X[6][5], XX[5][5], s[5];                       \
    FP a, b, c, d, e;                                                         \
    for (i = 0; i < 5; ++i) s[i] = 0;                                         \
    for (i = 0; i < 6; ++i) { mean_x += x[i]; mean_y += y[i]; }               \
    mean_x /= (FP) 6;                                                         \
    mean_y /= (FP) 6;                                                         \
    for (i = 0; i < 6; ++i)                                                   \
    {                                                                         \
        const FP x_ = x[i] - mean_x, y_ = y[i] - mean_y;                      \
        X[i][0] = x_ * x_; X[i][1] = x_ * y_; X[i][2] = y_ * y_;              \
        X[i][3] = x_; X[i][4] = y_;                                           \
        for (j = 0; j < 5; ++j) s[j] += X[i][j];                              \
    }                                                                         \
    /* XX[i][j] = sum_k X[k][j] X[k][i], skipping zero X[k][j]. */            \
    for (j = 0; j < 5; ++j)                                                   \
        for (i = 0; i < 5; ++i)                                               \
        {                                                                     \
            FP acc = 0;                                                       \
            for (k = 0; k < 6; ++k)                                           \
                if (X[k][j] != 0) acc += X[k][j] * X[k][i];                   \
            XX[i][j] = acc;                                                   \
        }                                                                     \
    /* LU with partial pivoting (first largest magnitude wins). */            \
    {                                                                         \
        int piv[5];                                                           \
        for (k = 0; k < 5; ++k)                                               \
        {                                                                     \
            int p = k;                                                        \
            FP mx = fabs(XX[k][k]);                                           \
            for (i = k + 1; i < 5; ++i)                                       \
                if (fabs(XX[i][k]) > mx) { mx = fabs(XX[i][k]); p = i; }      \
            piv[k] = p;                                                       \
            if (XX[p][k] != 0)                                                \
            {                                                                 \
                if (p != k)                                                   \
                    for (j = 0; j < 5; ++j)                                   \
                    { FP t = XX[k][j]; XX[k][j] = XX[p][j]; XX[p][j] = t; }   \
                if (fabs(XX[k][k]) >= FP_MIN)                                 \
                {                                                             \
                    const FP r = (FP) 1 / XX[k][k];                           \
                    for (i = k + 1; i < 5; ++i) XX[i][k] = r * XX[i][k];      \
                }                                                             \
                else                                                          \
                    for (i = k + 1; i < 5; ++i) XX[i][k] /= XX[k][k];         \
            }                                                                 \
            else ok = 0;                                                      \
            for (j = k + 1; j < 5; ++j)                                       \
                if (XX[k][j] != 0)                                            \
                    for (i = k + 1; i < 5; ++i)                               \
                        XX[i][j] -= XX[k][j] * XX[i][k];                      \
        }                                                                     \
        for (k = 0; k < 5; ++k)                                               \
            if (piv[k] != k)                                                  \
            { FP t = s[k]; s[k] = s[piv[k]]; s[piv[k]] = t; }                 \
        for (k = 0; k < 5; ++k)                                               \
            if (s[k] != 0)                                                    \
                for (i = k + 1; i < 5; ++i) s[i] -= s[k] * XX[i][k];          \
        for (k = 4; k >= 0; --k)                                              \
            if (s[k] != 0)                                                    \
            {                                                                 \
                s[k] /= XX[k][k];                                             \
                for (i = 0; i < k; ++i) s[i] -= s[k] * XX[i][k];              \
            }                                                                 \
    }                                                                         \
    a = s[0]; b = s[1]; c = s[2]; d = s[3]; e = s[4];                         \
    {                                                                         \
        const double r1 = fabs(b / a), r2 = fabs(b / c);                      \
        if ((r1 < r2 ? r1 : r2) > 1e-3)                                       \
        {                                                                     \
            FP a2, c2, d2, e2;                                                \
            *pa_out = 0.5 * atan2(b, (c - a)) - M_PI / 2.0;                   \
            const FP cos_phi = (FP) cos(*pa_out);                             \
            const FP sin_phi = (FP) sin(*pa_out);                             \
            a2 = a*cos_phi*cos_phi - b*cos_phi*sin_phi + c*sin_phi*sin_phi;   \
            c2 = a*sin_phi*sin_phi + b*cos_phi*sin_phi + c*cos_phi*cos_phi;   \
            d2 = d*cos_phi - e*sin_phi;                                       \
            e2 = d*sin_phi + e*cos_phi;                                       \
            a = a2; c = c2; d = d2; e = e2;                                   \
        }                                                                     \
        else *pa_out = (c - a < 0.0) ? 0.0 : M_PI / 2.0;                      \
    }                                                                         \
    if (a * c <= 0.0) return 0;                                               \
    if (a < 0.0) { a = -a; c = -c; d = -d; e = -e; }                          \
    {                                                                         \
        const FP F = 1.0 + (d * d) / (4.0 * a) + (e * e) / (4.0 * c);         \
        a = sqrt(F / a);                                                      \
        b = sqrt(F / c);                                                      \
        *major = 2.0 * (a > b ? a : b);                                       \
        *minor = 2.0 * (a < b ? a : b);                                       \
    }                                                                         \
    return ok;                                                                \
}                                                                             \
int NAME(int num, const FP* ra, const FP* dec, const FP* maj_, const FP* min_,\
        const FP* pa_, double ra0, double dec0, FP* a_, FP* b_, FP* c_,       \
        int* failed)                                                          \
{                                                                             \
    int i, j, num_failed = 0;                                                 \
    const double sin_dec0 = sin(dec0), cos_dec0 = cos(dec0);                  \
    const FP lon0 = (FP) ra0, cl0 = (FP) cos_dec0, sl0 = (FP) sin_dec0;       \
    for (i = 0; i < num; ++i)                                                 \
    {                                                                         \
        FP l[6], m[6], lon[6], lat[6], x[6], y[6], z[6];                      \
        FP maj, min, pa;                                                      \
        failed[i] = 0;                                                        \
        if (maj_[i] == 0.0 && min_[i] == 0.0) continue;                       \
        const FP ea = maj_[i] / 2.0, eb = min_[i] / 2.0;                      \
        const FP cos_pa = cos(pa_[i]), sin_pa = sin(pa_[i]);                  \
        for (j = 0; j < 6; ++j)                                               \
        {                                                                     \
            const FP t = j * 60.0 * M_PI / 180.0;                             \
            l[j] = ea * cos(t) * sin_pa + eb * sin(t) * cos_pa;               \
            m[j] = ea * cos(t) * cos_pa - eb * sin(t) * sin_pa;               \
        }                                                                     \
        /* Tangent plane at (0, 0) -> lon, lat (2D macro, lon0 = 0). */       \
        for (j = 0; j < 6; ++j)                                               \
        {                                                                     \
            const FP l0 = l[j], m0 = m[j];                                    \
            const FP n0 = NSQRT((FP) 1 - l0 * l0 - m0 * m0);                  \
            lat[j] = NASIN(n0 * (FP) 0 + m0 * (FP) 1);                        \
            lon[j] = (FP) 0 + NATAN2(l0, (FP) 1 * n0 - m0 * (FP) 0);          \
        }                                                                     \
        for (j = 0; j < 6; ++j)                                               \
        {                                                                     \
            const FP cos_lon = NCOS(lon[j]), sin_lon = NSIN(lon[j]);          \
            const FP cos_lat = NCOS(lat[j]), sin_lat = NSIN(lat[j]);          \
            x[j] = cos_lat * cos_lon;                                         \
            y[j] = cos_lat * sin_lon;                                         \
            z[j] = sin_lat;                                                   \
        }                                                                     \
        {                                                                     \
            const FP cos_lon = NCOS(ra[i]), sin_lon = NSIN(ra[i]);            \
            const FP cos_lat = NCOS(dec[i]), sin_lat = NSIN(dec[i]);          \
            for (j = 0; j < 6; ++j)                                           \
            {                                                                 \
                const FP x_ = x[j], y_ = y[j], z_ = z[j];                     \
                x[j] = x_ * cos_lon * cos_lat - y_ * sin_lon                  \
                        - z_ * cos_lon * sin_lat;                             \
                y[j] = x_ * cos_lat * sin_lon + y_ * cos_lon                  \
                        - z_ * sin_lon * sin_lat;                             \
                z[j] = x_ * sin_lat + z_ * cos_lat;                           \
            }                                                                 \
        }                                                                     \
        for (j = 0; j < 6; ++j)                                               \
        {                                                                     \
            const FP x_ = x[j], y_ = y[j], z_ = z[j];                         \
            lon[j] = NATAN2(y_, x_);                                          \
            lat[j] = NATAN2(z_, NSQRT(x_ * x_ + y_ * y_));                    \
        }                                                                     \
        for (j = 0; j < 6; ++j)                                               \
        {                                                                     \
            FP sin_lon, cos_lon, sin_lat, cos_lat;                            \
            const FP lo = lon[j] - lon0, la = lat[j];                         \
            sin_lon = NSIN(lo); cos_lon = NCOS(lo);                           \
            sin_lat = NSIN(la); cos_lat = NCOS(la);                           \
            l[j] = cos_lat * sin_lon;                                         \
            m[j] = cl0 * sin_lat - sl0 * cos_lat * cos_lon;                   \
        }                                                                     \
        if (!NAME##_fit(&maj, &min, &pa, l, m))                               \
        {                                                                     \
            failed[i] = 1;                                                    \
            ++num_failed;                                                     \
            continue;                                                         \
        }                                                                     \
        {                                                                     \
            const FP inv_std_maj_2 = 0.5 * (maj * maj) * M_PI_2_2_LN_2;       \
            const FP inv_std_min_2 = 0.5 * (min * min) * M_PI_2_2_LN_2;       \
            const FP cos_pa_2 = cos(pa) * cos(pa);                            \
            const FP sin_pa_2 = sin(pa) * sin(pa);                            \
            const FP sin_2pa = sin(2.0 * pa);                                 \
            a_[i] = cos_pa_2 * inv_std_min_2 + sin_pa_2 * inv_std_maj_2;      \
            b_[i] = -sin_2pa * inv_std_min_2 * 0.5                            \
                    + sin_2pa * inv_std_maj_2 * 0.5;                          \
            c_[i] = sin_pa_2 * inv_std_min_2 + cos_pa_2 * inv_std_maj_2;      \
        }                                                                     \
    }                                                                         \
    return num_failed;                                                        \
}
#define M_PI_2_2_LN_2 7.11941466249375271693034 /* pi^2 / (2 ln 2) */
GAUSS_PARAMS_BODY(oracle_gaussian_parameters_d, double, DBL_MIN, sin, cos,
        atan2, sqrt, asin)
GAUSS_PARAMS_BODY(oracle_gaussian_parameters_f, float, FLT_MIN, sinf, cosf,
        atan2f, sqrtf, asinf)

/* Great-circle distance (haversine; oskar/math/src/oskar_angular_distance.c
 * :13-21). */
double oracle_angular_distance(double lon1_rad, double lon2_rad,
        double lat1_rad, double lat2_rad)
{
    const double sin_delta_lat = sin(0.5 * (lat2_rad - lat1_rad));
    const double sin_delta_lon = sin(0.5 * (lon2_rad - lon1_rad));
    return 2.0 * asin(sqrt(sin_delta_lat * sin_delta_lat +
            cos(lat1_rad) * cos(lat2_rad) * sin_delta_lon * sin_delta_lon));
}

/* Snap of a beam direction to the nearest permitted (az, el)
 * (oskar/telescope/station/src/oskar_station_beam_horizon_direction.c
 * :50-88): direction cosines -> az, el; the permitted direction at the
 * smallest haversine distance (the first of equals); back to direction
 * cosines.  Returns its index. */
int oracle_snap_beam_direction(double* x, double* y, double* z, int n,
        const double* p_az, const double* p_el)
{
    int i, best = 0;
    double best_dist = DBL_MAX;
    const double az = atan2(*x, *y);
    const double el = atan2(*z, sqrt(*x * *x + *y * *y));
    for (i = 0; i < n; ++i)
    {
        const double d = oracle_angular_distance(p_az[i], az, p_el[i], el);
        if (d < best_dist)
        {
            best_dist = d;
            best = i;
        }
    }
    {
        const double cos_el = cos(p_el[best]);
        *x = cos_el * sin(p_az[best]);
        *y = cos_el * cos(p_az[best]);
        *z = sin(p_el[best]);
    }
    return best;
}

/*
 * Philox4x32-10 (Salmon, Moraes, Dror and Shaw, "Parallel random numbers:
 * as easy as 1, 2, 3", SC 2011), the counter-based generator the reference
 * takes from Random123 (oskar/math/private_random_generators.h): ten rounds
 * of two 32x32 -> 64-bit products with multipliers 0xD2511F53 and
 * 0xCD9E8D57, the key bumped by 0x9E3779B9 and 0xBB67AE85 between rounds.
 */
void oracle_philox4x32_10(const unsigned int* ctr, const unsigned int* key,
        unsigned int* out)
{
    unsigned int c[4], k0 = key[0], k1 = key[1];
    int r, i;
    for (i = 0; i < 4; ++i) c[i] = ctr[i];
    for (r = 0; r < 10; ++r)
    {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        {
            const unsigned long long p0 = 0xD2511F53ull * c[0];
            const unsigned long long p1 = 0xCD9E8D57ull * c[2];
            const unsigned int n0 = (unsigned int) (p1 >> 32) ^ c[1] ^ k0;
            const unsigned int n2 = (unsigned int) (p0 >> 32) ^ c[3] ^ k1;
            c[0] = n0;
            c[1] = (unsigned int) p1;
            c[2] = n2;
            c[3] = (unsigned int) p0;
        }
    }
    for (i = 0; i < 4; ++i) out[i] = c[i];
}

/* Box-Muller of two 32-bit integers, the reference's CPU form
 * (oskar/math/private_random_helpers.h): the angle pi x u0 / 2^31 + half
 * step (u0 read as a 64-bit long, so the angle spans [0, 2 pi)), the radius
 * sqrt(-2 ln((u1 + 1/2) / 2^32)). */
static void box_muller_d(unsigned int u0, unsigned int u1, double* f0,
        double* f1)
{
    const double fa = 1.0 / (1.0 + 0x7FFFFFFFuL);
    const double fb = 1.0 / (1.0 + 0xFFFFFFFFuL);
    double t = M_PI;
    t *= (((long) (unsigned long) u0) * fa) + 0.5 * fa;
    *f0 = sin(t);
    *f1 = cos(t);
    {
        const double r = sqrt(-2.0 * log((u1 * fb) + 0.5 * fb));
        *f0 *= r;
        *f1 *= r;
    }
}

/* Gaussian draws of oskar_mem_random_gaussian (oskar/mem/src/
 * oskar_mem_random_gaussian.c:90-129, double, std 1): value j comes from
 * Philox with key (seed, 0xCAFEF00D) and counter (j / 4, c1, c2, c3),
 * Box-Muller of its outputs 0-1 for j % 4 in {0, 1} and 2-3 for {2, 3}. */
void oracle_random_gaussian_d(unsigned int n, unsigned int seed,
        unsigned int c1, unsigned int c2, unsigned int c3, double* out)
{
    unsigned int i;
    for (i = 0; i < (n + 3) / 4; ++i)
    {
        const unsigned int ctr[4] = {i, c1, c2, c3};
        const unsigned int key[2] = {seed, 0xCAFEF00Du};
        unsigned int u[4];
        double t[4];
        unsigned int j;
        oracle_philox4x32_10(ctr, key, u);
        box_muller_d(u[0], u[1], &t[0], &t[1]);
        box_muller_d(u[2], u[3], &t[2], &t[3]);
        for (j = 0; j < 4 && 4 * i + j < n; ++j) out[4 * i + j] = t[j];
    }
}

/* Complex error factors of the station's elements
 * (oskar/telescope/station/src/oskar_evaluate_element_weights_errors.c,
 * define_evaluate_element_weights_errors.h:8-12): draws (r.x, r.y) per
 * element from oracle_random_gaussian_d(2 n, seed, time_index, station_id,
 * 0x12345678), amplitude r.x gain_error + gain, phase r.y phase_error +
 * phase, factor amplitude (cos, sin)(phase).  Output interleaved. */
void oracle_element_errors_d(int n, unsigned int seed, int time_index,
        int station_id, const double* gain, const double* gain_error,
        const double* phase, const double* phase_error, double* err)
{
    int i;
    oracle_random_gaussian_d((unsigned int) (2 * n), seed,
            (unsigned int) time_index, (unsigned int) station_id,
            0x12345678u, err);
    for (i = 0; i < n; ++i)
    {
        double rx = err[2 * i], ry = err[2 * i + 1];
        rx *= gain_error[i]; rx += gain[i];
        ry *= phase_error[i]; ry += phase[i];
        {
            const double im = sin(ry), re = cos(ry);
            err[2 * i] = re * rx;
            err[2 * i + 1] = im * rx;
        }
    }
}
