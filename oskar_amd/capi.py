"""ctypes bindings of the C-ABI declared in include/*.h.

Every exported symbol that the headers declare is listed in ``SIGNATURES``;
tests check that the built library exports each of them.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboskar_amd.so")

P = C.c_void_p
I = C.c_int
D = C.c_double
F = C.c_float
FP = C.POINTER(C.c_float)
IP = C.POINTER(C.c_int)
DP = C.POINTER(C.c_double)
SZ = C.c_size_t
SZP = C.POINTER(C.c_size_t)
CP = C.c_char_p
UB = C.c_ubyte
MEMA = C.c_void_p * 4   # const oskar_Mem* const src_flux[4]
MEM3 = C.c_void_p * 3   # const oskar_Mem* const [3]

_XC_F = [I, I, I, I, P] + [P] * 4 + [P] * 3 + [P] * 5 + [F] * 7 + [P]
_XC_D = [I, I, I, I, P] + [P] * 4 + [P] * 3 + [P] * 5 + [D] * 7 + [P]
_XCG_F = [I, I, I, I, P] + [P] * 4 + [P] * 3 + [P] * 3 + [P] * 5 + [F] * 7 + [P]
_XCG_D = [I, I, I, I, P] + [P] * 4 + [P] * 3 + [P] * 3 + [P] * 5 + [D] * 7 + [P]
_XS_F = [I, I, I, I, P, P] + [P] * 3 + [P] * 5 + [F] * 7 + [P]
_XS_D = [I, I, I, I, P, P] + [P] * 3 + [P] * 5 + [D] * 7 + [P]
_XSG_F = [I, I, I, I, P, P] + [P] * 3 + [P] * 3 + [P] * 5 + [F] * 7 + [P]
_XSG_D = [I, I, I, I, P, P] + [P] * 3 + [P] * 3 + [P] * 5 + [D] * 7 + [P]

# name -> (restype, argtypes)
SIGNATURES = {
    # oskar_global.h
    "oskar_get_error_string": (CP, [I]),
    # oskar_mem.h
    "oskar_type_precision": (I, [I]),
    "oskar_type_is_single": (I, [I]),
    "oskar_type_is_double": (I, [I]),
    "oskar_type_is_complex": (I, [I]),
    "oskar_type_is_real": (I, [I]),
    "oskar_type_is_matrix": (I, [I]),
    "oskar_type_is_scalar": (I, [I]),
    "oskar_mem_element_size": (SZ, [I]),
    "oskar_mem_create": (P, [I, I, SZ, IP]),
    "oskar_mem_create_alias": (P, [P, SZ, SZ, IP]),
    "oskar_mem_set_alias": (None, [P, P, SZ, SZ, IP]),
    "oskar_mem_create_alias_from_raw": (P, [P, I, I, SZ, IP]),
    "oskar_mem_create_copy": (P, [P, I, IP]),
    "oskar_mem_free": (None, [P, IP]),
    "oskar_mem_ref_dec": (None, [P]),
    "oskar_mem_ref_inc": (P, [P]),
    "oskar_mem_copy": (None, [P, P, IP]),
    "oskar_mem_copy_contents": (None, [P, P, SZ, SZ, SZ, IP]),
    "oskar_mem_clear_contents": (None, [P, IP]),
    "oskar_mem_ensure": (None, [P, SZ, IP]),
    "oskar_mem_realloc": (None, [P, SZ, IP]),
    "oskar_mem_set_value_real": (None, [P, D, SZ, SZ, IP]),
    "oskar_mem_normalise": (None, [P, SZ, SZ, SZ, IP]),
    "oskar_mem_evaluate_relative_error": (None, [P, P, DP, DP, DP, DP, IP]),
    "oskar_mem_get_element": (D, [P, SZ, IP]),
    "oskar_mem_set_element_real": (None, [P, SZ, D, IP]),
    "oskar_mem_allocated": (I, [P]),
    "oskar_mem_length": (SZ, [P]),
    "oskar_mem_location": (I, [P]),
    "oskar_mem_type": (I, [P]),
    "oskar_mem_precision": (I, [P]),
    "oskar_mem_is_double": (I, [P]),
    "oskar_mem_is_single": (I, [P]),
    "oskar_mem_is_complex": (I, [P]),
    "oskar_mem_is_real": (I, [P]),
    "oskar_mem_is_matrix": (I, [P]),
    "oskar_mem_is_scalar": (I, [P]),
    "oskar_mem_buffer": (P, [P]),
    "oskar_mem_buffer_const": (P, [P]),
    "oskar_mem_void": (P, [P]),
    "oskar_mem_void_const": (P, [P]),
    "oskar_mem_char": (P, [P]),
    "oskar_mem_char_const": (P, [P]),
    "oskar_mem_int": (P, [P, IP]),
    "oskar_mem_int_const": (P, [P, IP]),
    "oskar_mem_float": (P, [P, IP]),
    "oskar_mem_float_const": (P, [P, IP]),
    "oskar_mem_float2": (P, [P, IP]),
    "oskar_mem_float2_const": (P, [P, IP]),
    "oskar_mem_float4c": (P, [P, IP]),
    "oskar_mem_float4c_const": (P, [P, IP]),
    "oskar_mem_double": (P, [P, IP]),
    "oskar_mem_double_const": (P, [P, IP]),
    "oskar_mem_double2": (P, [P, IP]),
    "oskar_mem_double2_const": (P, [P, IP]),
    "oskar_mem_double4c": (P, [P, IP]),
    "oskar_mem_double4c_const": (P, [P, IP]),
    # oskar_jones.h
    "oskar_jones_create": (P, [I, I, I, I, IP]),
    "oskar_jones_create_copy": (P, [P, I, IP]),
    "oskar_jones_free": (None, [P, IP]),
    "oskar_jones_set_size": (None, [P, I, I, IP]),
    "oskar_jones_num_sources": (I, [P]),
    "oskar_jones_num_stations": (I, [P]),
    "oskar_jones_type": (I, [P]),
    "oskar_jones_mem_location": (I, [P]),
    "oskar_jones_mem": (P, [P]),
    "oskar_jones_mem_const": (P, [P]),
    "oskar_jones_join": (None, [P, P, P, IP]),
    "oskar_evaluate_jones_K": (None, [P, I, P, P, P, P, P, P, D, P, D, D, I,
                                      IP]),
    # oskar_telescope.h
    "oskar_telescope_create": (P, [I, I, I, IP]),
    "oskar_telescope_resize_station_array": (None, [P, I, IP]),
    "oskar_telescope_num_station_models": (I, [P]),
    "oskar_telescope_station": (P, [P, I]),
    "oskar_telescope_station_const": (P, [P, I]),
    "oskar_telescope_free": (None, [P, IP]),
    "oskar_telescope_precision": (I, [P]),
    "oskar_telescope_mem_location": (I, [P]),
    "oskar_telescope_num_stations": (I, [P]),
    "oskar_telescope_num_baselines": (I, [P]),
    "oskar_telescope_phase_centre_coord_type": (I, [P]),
    "oskar_telescope_phase_centre_longitude_rad": (D, [P]),
    "oskar_telescope_phase_centre_latitude_rad": (D, [P]),
    "oskar_telescope_channel_bandwidth_hz": (D, [P]),
    "oskar_telescope_time_average_sec": (D, [P]),
    "oskar_telescope_uv_filter_min": (D, [P]),
    "oskar_telescope_uv_filter_max": (D, [P]),
    "oskar_telescope_uv_filter_units": (I, [P]),
    "oskar_telescope_station_true_offset_ecef_metres": (P, [P, I]),
    "oskar_telescope_station_true_offset_ecef_metres_const": (P, [P, I]),
    "oskar_telescope_set_channel_bandwidth": (None, [P, D]),
    "oskar_telescope_set_time_average": (None, [P, D]),
    "oskar_telescope_set_phase_centre": (None, [P, I, D, D]),
    "oskar_telescope_set_uv_filter": (None, [P, D, D, CP, IP]),
    # oskar_cross_correlate.h
    "oskar_cross_correlate": (None, [I, I, I, P, MEMA, MEM3, MEM3, P, MEM3,
                                     D, D, I, P, IP]),
    "oskar_cross_correlate_point_hip_f": (None, _XC_F),
    "oskar_cross_correlate_point_hip_d": (None, _XC_D),
    "oskar_cross_correlate_gaussian_hip_f": (None, _XCG_F),
    "oskar_cross_correlate_gaussian_hip_d": (None, _XCG_D),
    "oskar_cross_correlate_scalar_point_hip_f": (None, _XS_F),
    "oskar_cross_correlate_scalar_point_hip_d": (None, _XS_D),
    "oskar_cross_correlate_scalar_gaussian_hip_f": (None, _XSG_F),
    "oskar_cross_correlate_scalar_gaussian_hip_d": (None, _XSG_D),
    # oskar_auto_correlate.h
    "oskar_auto_correlate": (None, [I, P, MEMA, I, P, IP]),
    # oskar_cross_correlate_factorised.h
    "oskar_cross_correlate_factorised": (None, [I, I, P, MEMA, MEM3, P, MEM3,
                                                D, I, I, P, IP]),
    "oskar_cross_correlate_factorised_set_part_size": (I, [I]),
    "oskar_cross_correlate_factorised_set_tile_shape": (I, [I]),
    "oskar_cross_correlate_factorised_set_profiling": (None, [I]),
    "oskar_cross_correlate_factorised_last_times": (I, [DP]),
    "oskar_factorised_batch_create": (P, [I, I, I, I, IP]),
    "oskar_factorised_batch_set_unit": (None, [P, I, I, P, MEMA, MEM3, MEM3,
                                               D, I, P, IP]),
    "oskar_factorised_batch_run": (None, [P, I, I, P, IP]),
    "oskar_factorised_batch_autotune": (None, [P, I, IP]),
    "oskar_factorised_batch_set_config": (None, [P, I, I, I, I, IP]),
    "oskar_factorised_batch_config": (None, [P, IP]),
    "oskar_factorised_batch_last_run_ms": (D, [P]),
    "oskar_factorised_batch_free": (None, [P, IP]),
    # oskar_convert.h
    "oskar_convert_mjd_to_gmst": (D, [D]),
    "oskar_equation_of_equinoxes_fast": (D, [D]),
    "oskar_convert_mjd_to_gast_fast": (D, [D]),
    "oskar_convert_any_to_enu_directions": (None, [I, I, MEM3, D, D, D, D,
                                                   MEM3, IP]),
    "oskar_station_beam": (None, [P, P, I, I, MEM3, D, D, I, D, D, I, D, D,
                                  D, I, P, IP]),
    # oskar_beam_surrogate.h
    "oskar_beam_surrogate_num_params": (I, []),
    "oskar_beam_surrogate_num_features": (I, []),
    "oskar_beam_surrogate_num_fourier": (I, []),
    "oskar_beam_surrogate_create": (P, [D, D, D, D, D, C.c_uint, IP]),
    "oskar_beam_surrogate_free": (None, [P, IP]),
    "oskar_beam_surrogate_fourier": (None, [P, FP]),
    "oskar_beam_surrogate_get_params": (None, [P, FP, IP]),
    "oskar_beam_surrogate_set_params": (None, [P, FP, IP]),
    "oskar_beam_surrogate_peak": (D, [P]),
    "oskar_beam_surrogate_set_peak": (None, [P, D]),
    "oskar_beam_surrogate_training_set": (None, [P, P, I, I, D, D, D, D,
                                                 C.c_uint, P, P, P, P, IP]),
    "oskar_beam_surrogate_train": (None, [P, I, P, P, P, P, I, I, D, D,
                                          C.c_uint, DP, IP]),
    "oskar_beam_surrogate_loss_and_gradient": (None, [P, I, P, P, DP, P, IP]),
    "oskar_beam_surrogate_evaluate": (None, [P, I, P, P, D, P, IP]),
    "oskar_beam_surrogate_evaluate_with_fallback": (None, [P, P, P, I, P, P,
                                                           P, D, D, P, IP,
                                                           IP]),
    # oskar_dftw.h
    "oskar_dftw": (None, [I, I, D, P, P, P, P, I, I, P, P, P, P, P, I, I, I,
                          P, IP]),
    # oskar_device.h
    "oskar_device_set": (None, [I, I, IP]),
    "oskar_device_synchronize": (None, [I]),
    "oskar_device_check_error_hip": (None, [IP]),
    "oskar_device_count_hip": (I, []),
    "oskar_hip_set_stream": (None, [P]),
    "oskar_hip_stream": (P, []),
    # oskar_timer.h
    "oskar_timer_create": (P, [I]),
    "oskar_timer_free": (None, [P]),
    "oskar_timer_elapsed": (D, [P]),
    "oskar_timer_pause": (None, [P]),
    "oskar_timer_reset": (None, [P]),
    "oskar_timer_resume": (None, [P]),
    "oskar_timer_restart": (None, [P]),
    "oskar_timer_start": (None, [P]),
    # oskar_station.h
    "oskar_station_create": (P, [I, I, I, IP]),
    "oskar_station_create_copy": (P, [P, I, IP]),
    "oskar_station_free": (None, [P, IP]),
    "oskar_station_resize": (None, [P, I, IP]),
    "oskar_station_analyse": (None, [P, IP, IP]),
    "oskar_station_set_element_coords": (None, [P, I, I, DP, DP, IP]),
    "oskar_station_set_element_cable_length_error": (None, [P, I, I, D, IP]),
    "oskar_station_set_element_feed_angle": (None, [P, I, I, D, D, D, IP]),
    "oskar_station_set_element_type": (None, [P, I, I, IP]),
    "oskar_station_resize_element_types": (None, [P, I, IP]),
    "oskar_station_set_station_type": (None, [P, I]),
    "oskar_station_set_normalise_final_beam": (None, [P, I]),
    "oskar_station_set_position": (None, [P, D, D, D, D, D, D]),
    "oskar_station_set_phase_centre": (None, [P, I, D, D]),
    "oskar_station_set_normalise_array_pattern": (None, [P, I]),
    "oskar_station_set_normalise_element_pattern": (None, [P, I]),
    "oskar_station_set_enable_array_pattern": (None, [P, I]),
    "oskar_station_set_swap_xy": (None, [P, I]),
    "oskar_station_precision": (I, [P]),
    "oskar_station_mem_location": (I, [P]),
    "oskar_station_type": (I, [P]),
    "oskar_station_num_elements": (I, [P]),
    "oskar_station_num_element_types": (I, [P]),
    "oskar_station_normalise_final_beam": (I, [P]),
    "oskar_station_normalise_array_pattern": (I, [P]),
    "oskar_station_normalise_element_pattern": (I, [P]),
    "oskar_station_enable_array_pattern": (I, [P]),
    "oskar_station_swap_xy": (I, [P]),
    "oskar_station_array_is_3d": (I, [P]),
    "oskar_station_common_element_orientation": (I, [P]),
    "oskar_station_common_pol_beams": (I, [P]),
    "oskar_station_lon_rad": (D, [P]),
    "oskar_station_lat_rad": (D, [P]),
    "oskar_station_beam_lon_rad": (D, [P]),
    "oskar_station_beam_lat_rad": (D, [P]),
    "oskar_station_beam_coord_type": (I, [P]),
    "oskar_station_element_true_enu_metres": (P, [P, I, I]),
    "oskar_station_element_true_enu_metres_const": (P, [P, I, I]),
    "oskar_station_element_measured_enu_metres": (P, [P, I, I]),
    "oskar_station_element_measured_enu_metres_const": (P, [P, I, I]),
    "oskar_station_element_cable_length_error_metres": (P, [P, I]),
    "oskar_station_element_cable_length_error_metres_const": (P, [P, I]),
    "oskar_station_element_types": (P, [P]),
    "oskar_station_element_types_const": (P, [P]),
    "oskar_station_element_euler_index_rad": (D, [P, I, I, I]),
    "oskar_station_element": (P, [P, I]),
    "oskar_station_element_const": (P, [P, I]),
    "oskar_element_set_element_type": (None, [P, CP, IP]),
    "oskar_element_set_dipole_length": (None, [P, D, CP, IP]),
    "oskar_element_type": (I, [P]),
    "oskar_element_dipole_length": (D, [P]),
    "oskar_element_dipole_length_units": (I, [P]),
    "oskar_station_work_create": (P, [I, I, IP]),
    "oskar_station_work_horizon_mask": (P, [P]),
    "oskar_station_work_source_indices": (P, [P]),
    # oskar_sky.h
    "oskar_sky_create": (P, [I, I, I, IP]),
    "oskar_sky_free": (None, [P, IP]),
    "oskar_sky_double": (D, [P, I]),
    "oskar_sky_int": (I, [P, I]),
    "oskar_sky_set_double": (None, [P, I, D]),
    "oskar_sky_set_int": (None, [P, I, I]),
    "oskar_sky_column_attribute": (I, [P, I]),
    "oskar_sky_column_type": (I, [P, I]),
    "oskar_sky_num_valid_columns_of_type": (I, [P, I, I]),
    "oskar_sky_set_num_valid_columns_of_type": (None, [P, I, I, I]),
    "oskar_sky_column": (P, [P, I, I, IP]),
    "oskar_sky_column_const": (P, [P, I, I]),
    "oskar_sky_resize": (None, [P, I, IP]),
    "oskar_update_horizon_mask": (None, [I, P, P, P, D, D, D, P, IP]),
    "oskar_prefix_sum": (None, [SZ, P, P, IP]),
    "oskar_sky_copy_source_data": (None, [P, P, P, P, IP]),
    "oskar_sky_horizon_clip": (None, [P, P, P, D, P, IP]),
    "oskar_sky_evaluate_relative_directions": (None, [P, D, D, IP]),
    "oskar_sky_evaluate_gaussian_source_parameters": (
        None, [P, I, D, D, IP, IP]),
    "oskar_station_work_free": (None, [P, IP]),
    "oskar_station_beam_horizon_direction": (None, [P, D, DP, DP, DP, IP]),
    "oskar_evaluate_station_beam_aperture_array": (None, [P, P, I, P, P, P, I,
                                                          D, D, P, IP]),
    "oskar_evaluate_element_weights_dft": (None, [I, P, P, P, P, D, D, D, D,
                                                  P, IP]),
    "oskar_evaluate_dipole_pattern": (None, [I, P, P, P, D, D, I, I, P, IP]),
    "oskar_convert_enu_directions_to_theta_phi": (None, [I, I, P, P, P, I, D,
                                                         D, P, P, P, IP]),
    "oskar_blank_below_horizon": (None, [I, I, P, I, P, IP]),
    # oskar_binary.h
    "oskar_binary_create": (P, [CP, C.c_char, IP]),
    "oskar_binary_free": (None, [P]),
    "oskar_binary_write": (None, [P, UB, UB, UB, I, SZ, P, IP]),
    "oskar_binary_write_ext": (None, [P, UB, CP, CP, I, SZ, P, IP]),
    "oskar_binary_write_int": (None, [P, UB, UB, I, I, IP]),
    "oskar_binary_write_double": (None, [P, UB, UB, I, D, IP]),
    "oskar_binary_num_tags": (I, [P]),
    "oskar_binary_query": (I, [P, UB, UB, UB, I, SZP, IP]),
    "oskar_binary_query_ext": (I, [P, UB, CP, CP, I, SZP, IP]),
    "oskar_binary_set_query_search_start": (None, [P, I, IP]),
    "oskar_binary_read_block": (None, [P, I, SZ, P, IP]),
    "oskar_binary_read": (None, [P, UB, UB, UB, I, SZ, P, IP]),
    "oskar_binary_read_ext": (None, [P, UB, CP, CP, I, SZ, P, IP]),
    "oskar_binary_read_int": (None, [P, UB, UB, I, IP, IP]),
    "oskar_binary_read_double": (None, [P, UB, UB, I, DP, IP]),
    "oskar_binary_crc32c": (C.c_uint, [C.c_uint, P, SZ]),
    # oskar_vis.h
    "oskar_vis_block_create": (P, [I, I, I, I, I, I, I, IP]),
    "oskar_vis_block_create_from_header": (P, [I, P, IP]),
    "oskar_vis_block_free": (None, [P, IP]),
    "oskar_vis_block_resize": (None, [P, I, I, I, IP]),
    "oskar_vis_block_clear": (None, [P, IP]),
    "oskar_vis_block_num_times": (I, [P]),
    "oskar_vis_block_num_channels": (I, [P]),
    "oskar_vis_block_num_baselines": (I, [P]),
    "oskar_vis_block_num_stations": (I, [P]),
    "oskar_vis_block_num_pols": (I, [P]),
    "oskar_vis_block_start_time_index": (I, [P]),
    "oskar_vis_block_start_channel_index": (I, [P]),
    "oskar_vis_block_set_start_time_index": (None, [P, I]),
    "oskar_vis_block_set_start_channel_index": (None, [P, I]),
    "oskar_vis_block_has_auto_correlations": (I, [P]),
    "oskar_vis_block_has_cross_correlations": (I, [P]),
    "oskar_vis_block_cross_correlations": (P, [P]),
    "oskar_vis_block_auto_correlations": (P, [P]),
    "oskar_vis_block_baseline_uvw_metres": (P, [P, I]),
    "oskar_vis_block_station_uvw_metres": (P, [P, I]),
    "oskar_vis_block_station_to_baseline_coords": (None, [P, P, IP]),
    "oskar_vis_block_write": (None, [P, P, I, IP]),
    "oskar_vis_block_read": (None, [P, P, P, I, IP]),
    "oskar_convert_station_uvw_to_baseline_uvw": (None, [I, I, I, P, P, P, I,
                                                         P, P, P, IP]),
    "oskar_vis_header_create": (P, [I, I, I, I, I, I, I, I, I, IP]),
    "oskar_vis_header_free": (None, [P, IP]),
    "oskar_vis_header_amp_type": (I, [P]),
    "oskar_vis_header_coord_precision": (I, [P]),
    "oskar_vis_header_num_stations": (I, [P]),
    "oskar_vis_header_num_tags_per_block": (I, [P]),
    "oskar_vis_header_num_tags_header": (I, [P]),
    "oskar_vis_header_max_times_per_block": (I, [P]),
    "oskar_vis_header_max_channels_per_block": (I, [P]),
    "oskar_vis_header_num_times_total": (I, [P]),
    "oskar_vis_header_num_channels_total": (I, [P]),
    "oskar_vis_header_write_auto_correlations": (I, [P]),
    "oskar_vis_header_write_cross_correlations": (I, [P]),
    "oskar_vis_header_pol_type": (I, [P]),
    "oskar_vis_header_casa_phase_convention": (I, [P]),
    "oskar_vis_header_freq_start_hz": (D, [P]),
    "oskar_vis_header_freq_inc_hz": (D, [P]),
    "oskar_vis_header_channel_bandwidth_hz": (D, [P]),
    "oskar_vis_header_time_start_mjd_utc": (D, [P]),
    "oskar_vis_header_time_inc_sec": (D, [P]),
    "oskar_vis_header_time_average_sec": (D, [P]),
    "oskar_vis_header_phase_centre_type": (I, [P]),
    "oskar_vis_header_phase_centre_ra_deg": (D, [P]),
    "oskar_vis_header_phase_centre_dec_deg": (D, [P]),
    "oskar_vis_header_telescope_lon_deg": (D, [P]),
    "oskar_vis_header_telescope_lat_deg": (D, [P]),
    "oskar_vis_header_telescope_alt_metres": (D, [P]),
    "oskar_vis_header_station_offset_ecef_metres": (P, [P, I]),
    "oskar_vis_header_telescope_path": (P, [P]),
    "oskar_vis_header_settings": (P, [P]),
    "oskar_vis_header_station_diameter_m": (P, [P]),
    "oskar_vis_header_station_name": (P, [P, I]),
    "oskar_vis_header_element_enu_metres": (P, [P, I, I]),
    "oskar_vis_header_element_feed_angle_rad": (P, [P, I, I, I]),
    "oskar_vis_header_set_pol_type": (None, [P, I, IP]),
    "oskar_vis_header_set_casa_phase_convention": (None, [P, I]),
    "oskar_vis_header_set_freq_start_hz": (None, [P, D]),
    "oskar_vis_header_set_freq_inc_hz": (None, [P, D]),
    "oskar_vis_header_set_channel_bandwidth_hz": (None, [P, D]),
    "oskar_vis_header_set_time_start_mjd_utc": (None, [P, D]),
    "oskar_vis_header_set_time_inc_sec": (None, [P, D]),
    "oskar_vis_header_set_time_average_sec": (None, [P, D]),
    "oskar_vis_header_set_phase_centre": (None, [P, I, D, D]),
    "oskar_vis_header_set_telescope_centre": (None, [P, D, D, D]),
    "oskar_vis_header_write": (P, [P, CP, IP]),
    "oskar_vis_header_read": (P, [P, IP]),
    "oskar_binary_write_mem": (None, [P, P, UB, UB, I, SZ, IP]),
    "oskar_binary_read_mem": (None, [P, P, UB, UB, I, IP]),
}


class _Lib:
    """Holds the loaded CDLL with prototypes applied."""

    def __init__(self, path):
        if not os.path.exists(path):
            raise ImportError(
                "liboskar_amd.so not found at %s: build it first with "
                "`python -c 'import __graft_entry__ as g; g.build()'` "
                "(make -C oskar_amd/csrc)" % path)
        self.path = path
        self.dll = C.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(self.dll, name, None)
            if fn is None:
                continue  # reported by the export test
            fn.restype = res
            fn.argtypes = args

    def __getattr__(self, name):
        return getattr(self.dll, name)


def load(path=LIB_PATH):
    return _Lib(path)


lib = load()
