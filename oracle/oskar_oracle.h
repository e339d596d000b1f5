/*
 * CPU oracle for the visibility hot path: a plain-C (OpenMP) restatement of
 * the reference's CPU algorithms.  TEST INFRASTRUCTURE ONLY: it is used by
 * tests/, by __graft_entry__.smoke() and by bench.py's CPU-baseline leg as
 * the checker and the CPU timing baseline, never by the library itself.
 *
 * All arrays are host memory.  Complex values are interleaved (re, im);
 * a 2x2 complex matrix is 8 reals [a.re a.im b.re b.im c.re c.im d.re d.im]
 * (oskar/utility/oskar_vector_types.h:85,98).
 */
#ifndef OSKAR_ORACLE_H_
#define OSKAR_ORACLE_H_

#ifdef __cplusplus
extern "C" {
#endif

/* Cross-correlation, double precision.  Follows
 * oskar/correlate/src/oskar_cross_correlate_omp.cpp:33-156 (matrix) and
 * oskar_cross_correlate_scalar_omp.cpp (scalar); derived scalars as in
 * oskar/correlate/src/oskar_cross_correlate.c:42-71.  matrix: 1 for 2x2
 * Jones, 0 for scalar Jones.  Adds into vis[offset_out + b]. */
void oracle_cross_correlate_d(int use_casa, int source_type, int matrix,
        int num_sources, int num_stations, const double* jones,
        const double* I, const double* Q, const double* U, const double* V,
        const double* l, const double* m, const double* n,
        const double* a, const double* b, const double* c,
        const double* st_u, const double* st_v, const double* st_w,
        const double* st_x, const double* st_y,
        double channel_bandwidth_hz, double time_average_sec,
        int phase_centre_is_azel, double phase_centre_ra_rad,
        double phase_centre_dec_rad, double uv_min, double uv_max,
        int uv_units_wavelengths, double gast, double frequency_hz,
        int offset_out, double* vis);

/* Single precision with Kahan-compensated source sums, as the reference's
 * float instantiation (oskar_cross_correlate_omp.cpp:140-144). */
void oracle_cross_correlate_f(int use_casa, int source_type, int matrix,
        int num_sources, int num_stations, const float* jones,
        const float* I, const float* Q, const float* U, const float* V,
        const float* l, const float* m, const float* n,
        const float* a, const float* b, const float* c,
        const float* st_u, const float* st_v, const float* st_w,
        const float* st_x, const float* st_y,
        double channel_bandwidth_hz, double time_average_sec,
        int phase_centre_is_azel, double phase_centre_ra_rad,
        double phase_centre_dec_rad, double uv_min, double uv_max,
        int uv_units_wavelengths, double gast, double frequency_hz,
        int offset_out, float* vis);

/* Auto-correlation, oskar/correlate/define_auto_correlate.h:32-58 and
 * :84-104 (CPU bodies): matrix adds sum_i J_s(i) B_i J_s(i)^H to
 * vis[offset_out + s] with zero imaginary parts on the diagonal; scalar adds
 * sum_i |J_s(i)|^2 I_i to the real part only.  Plain sums in FP. */
void oracle_auto_correlate_d(int matrix, int num_sources, int num_stations,
        int offset_out, const double* jones, const double* I, const double* Q,
        const double* U, const double* V, double* vis);
void oracle_auto_correlate_f(int matrix, int num_sources, int num_stations,
        int offset_out, const float* jones, const float* I, const float* Q,
        const float* U, const float* V, float* vis);

/* Horizon mask: mask[i] |= (l ll + m mm + n nn > 0) with (ll, mm, nn) of
 * oskar/sky/src/oskar_update_horizon_mask.c:37-44 (float body: the factors
 * rounded to float, as there). */
void oracle_update_horizon_mask_d(int num, const double* l, const double* m,
        const double* n, double ha0, double dec0, double lat, int* mask);
void oracle_update_horizon_mask_f(int num, const float* l, const float* m,
        const float* n, double ha0, double dec0, double lat, int* mask);

/* Exclusive prefix sum, out[num] = total (oskar/math/src/oskar_prefix_sum.c). */
void oracle_prefix_sum_int(int num, const int* in, int* out);

/* Masked, order-keeping row copy of a column table (column c at
 * c x capacity) with nfix valid-column counts per row
 * (oskar/sky/src/oskar_sky_copy_source_data.c, CPU branch); returns the
 * number of rows copied. */
int oracle_sky_copy_source_data_d(int num, int cap_in, int cap_out,
        const int* mask, int num_col, const double* table_in,
        double* table_out, int nfix, const int* valid_in, int* valid_out);
int oracle_sky_copy_source_data_f(int num, int cap_in, int cap_out,
        const int* mask, int num_col, const float* table_in,
        float* table_out, int nfix, const int* valid_in, int* valid_out);

/* Weighted DFT, oskar/math/define_dftw_c2c.h:79-109 and
 * define_dftw_m2m.h:109-167 (CPU bodies).  z_in/z_out NULL -> 2D; data_idx
 * NULL -> identity. */
void oracle_dftw_d(int matrix, int normalise, int num_in, double wavenumber,
        const double* weights, const double* x_in, const double* y_in,
        const double* z_in, int offset_coord_out, int num_out,
        const double* x_out, const double* y_out, const double* z_out,
        const int* data_idx, const double* data, int eval_x, int eval_y,
        int offset_out, double* output);
void oracle_dftw_f(int matrix, int normalise, int num_in, double wavenumber,
        const float* weights, const float* x_in, const float* y_in,
        const float* z_in, int offset_coord_out, int num_out,
        const float* x_out, const float* y_out, const float* z_out,
        const int* data_idx, const float* data, int eval_x, int eval_y,
        int offset_out, float* output);

/* DFT weights, oskar/telescope/station/define_evaluate_element_weights_dft.h. */
void oracle_element_weights_dft_d(int n, const double* x, const double* y,
        const double* z, const double* cable, double wavenumber, double x1,
        double y1, double z1, double* weights);
void oracle_element_weights_dft_f(int n, const float* x, const float* y,
        const float* z, const float* cable, double wavenumber, double x1,
        double y1, double z1, float* weights);

/* ENU direction cosines -> theta, phi (oskar/convert/
 * define_convert_enu_directions_to_theta_phi.h). */
void oracle_theta_phi_d(int off_in, int num, const double* x, const double* y,
        const double* z, int extra_point_at_pole, double delta_phi1,
        double delta_phi2, double* theta, double* phi1, double* phi2);

/* Dipole pattern (oskar/telescope/station/element/
 * define_evaluate_dipole_pattern.h).  matrix: 2x2 output, else scalar. */
void oracle_dipole_pattern_d(int matrix, int n, const double* theta,
        const double* phi_x, const double* phi_y, double kL, double cos_kL,
        int swap_xy, int offset_out, double* pattern);

/* Aperture-array station beam for one station, double precision, following
 * oskar_evaluate_station_beam_aperture_array.c:70-240 for a station without
 * children: element_type[t] is 0 (dipole) or 1 (isotropic);
 * dipole_length_m[t] in metres; beam (matrix: 8 reals, else 2 per point)
 * holds num_points values. */
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
        double* beam);

/* Jones K: K[a][s] = exp(+i k (u_a l_s + v_a m_s + w_a (n_s - 1)))
 * (oskar/interferometer/define_evaluate_jones_K.h:22-75), where source s is
 * included only when source_filter_min < I_s <= source_filter_max. */
void oracle_evaluate_jones_K_d(int num_sources, const double* l,
        const double* m, const double* n, int num_stations, const double* u,
        const double* v, const double* w, double wavenumber,
        const double* I, double source_filter_min, double source_filter_max,
        double* jones);

/* Factorised-product reference: V_pq for the listed baselines, with
 * A_p(s) = exp(i k (u_p l + v_p m + w_p (n - 1))) E(s) sqrt(I_s) (one station
 * beam E shared by all stations, unpolarised point sources), summed in long
 * double; p > q, non-CASA, no smearing.  Output 8 reals per baseline. */
void oracle_factorised_subset_ld(int num_sources, const double* l,
        const double* m, const double* n, const double* I, const double* E,
        int num_stations, const double* u, const double* v, const double* w,
        double wavenumber, int num_baselines, const int* p, const int* q,
        double* vis_out);

/* Relative direction cosines of sky positions about (lon0, lat0)
 * (oskar/convert/define_convert_lon_lat_to_relative_directions.h, 3D body;
 * oskar/sky/src/oskar_sky_evaluate_relative_directions.c:20-37). */
void oracle_relative_directions_d(int num, const double* lon_rad,
        const double* lat_rad, double lon0, double lat0, double* l,
        double* m, double* n);
void oracle_relative_directions_f(int num, const float* lon_rad,
        const float* lat_rad, double lon0, double lat0, float* l, float* m,
        float* n);

/* Gaussian source parameters a, b, c from FWHMs and position angles
 * (oskar/sky/src/oskar_sky_evaluate_gaussian_source_parameters.c:32-262):
 * failed[i] = 1 where the ellipse fit fails (a, b, c left as they were);
 * sources with both axes zero are skipped.  Returns the number failed. */
int oracle_gaussian_parameters_d(int num, const double* ra,
        const double* dec, const double* maj, const double* min,
        const double* pa, double ra0, double dec0, double* a, double* b,
        double* c, int* failed);
int oracle_gaussian_parameters_f(int num, const float* ra, const float* dec,
        const float* maj, const float* min, const float* pa, double ra0,
        double dec0, float* a, float* b, float* c, int* failed);

/* Great-circle distance by the haversine formula
 * (oskar/math/src/oskar_angular_distance.c:13-21). */
double oracle_angular_distance(double lon1_rad, double lon2_rad,
        double lat1_rad, double lat2_rad);

/* Nearest permitted (az, el) beam direction, replacing (x, y, z)
 * (oskar_station_beam_horizon_direction.c:50-88); returns its index. */
int oracle_snap_beam_direction(double* x, double* y, double* z, int n,
        const double* p_az, const double* p_el);

/* Philox4x32-10 (the reference's Random123 generator). */
void oracle_philox4x32_10(const unsigned int* ctr, const unsigned int* key,
        unsigned int* out);
/* oskar_mem_random_gaussian's draws (double, std 1). */
void oracle_random_gaussian_d(unsigned int n, unsigned int seed,
        unsigned int c1, unsigned int c2, unsigned int c3, double* out);
/* Complex error factors of a station's elements for one time index. */
void oracle_element_errors_d(int n, unsigned int seed, int time_index,
        int station_id, const double* gain, const double* gain_error,
        const double* phase, const double* phase_error, double* err);

#ifdef __cplusplus
}
#endif

#endif /* OSKAR_ORACLE_H_ */
