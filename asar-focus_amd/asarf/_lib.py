"""ctypes binding of libasarf.so (the C ABI declared in include/asarf.h).

The product path is the HIP library: if it is missing this module raises instead of falling back to anything.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
LIB_PATH = os.path.join(PKG, "lib", "libasarf.so")
HEADER_PATH = os.path.join(os.path.dirname(PKG), "include", "asarf.h")

OK, EINVAL, ESHAPE, EWORKSPACE, EHIP, ENOMEM, EUNSUPPORTED = range(7)
Q1_IQ_PADDED_WIDTH, Q2_COS_NO_SQRT, Q4_DC_EXTRA_BLOCK = 1, 2, 4
QUIRKS_REFERENCE = 7
CF32, F32 = 0, 1
RAW_CF32, RAW_U8, RAW_I8, RAW_ERS_LINES, RAW_ASAR_FBAQ4 = 0, 1, 2, 3, 4
FORWARD, INVERSE = 0, 1
FFT_ROWS, FFT_COLUMNS = 0, 1


class AsarfError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"asarf status {status}: {msg}")
        self.status = status


class Image(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("x_size", ctypes.c_int), ("y_size", ctypes.c_int),
                ("x_stride", ctypes.c_int), ("y_stride", ctypes.c_int), ("dtype", ctypes.c_int)]


class Workspace(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("bytes", ctypes.c_size_t)]


class ChirpInfo(ctypes.Structure):
    _fields_ = [("range_sampling_rate", ctypes.c_double), ("pulse_duration", ctypes.c_double),
                ("n_samples", ctypes.c_int), ("pulse_bandwidth", ctypes.c_double), ("kr", ctypes.c_double),
                ("apply_window", ctypes.c_int)]


class IqResults(ctypes.Structure):
    _fields_ = [("dc_i", ctypes.c_double), ("dc_q", ctypes.c_double), ("iq_gain", ctypes.c_double),
                ("phase_error", ctypes.c_double), ("stddev_i", ctypes.c_double), ("stddev_q", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class RdaParams(ctypes.Structure):
    _fields_ = [("wavelength", ctypes.c_double), ("range_spacing", ctypes.c_double),
                ("slant_range_first", ctypes.c_double), ("prf", ctypes.c_double),
                ("az_bandwidth_fraction", ctypes.c_double), ("dc_poly", ctypes.c_double * 3),
                ("vr_poly", ctypes.c_double * 3)]


class FocusParams(ctypes.Structure):
    _fields_ = [("na", ctypes.c_int), ("nr", ctypes.c_int), ("nc", ctypes.c_int), ("nrp", ctypes.c_int),
                ("nap", ctypes.c_int), ("raw_format", ctypes.c_int), ("prf", ctypes.c_double),
                ("wavelength", ctypes.c_double), ("range_spacing", ctypes.c_double),
                ("slant_range_first", ctypes.c_double), ("az_bandwidth_fraction", ctypes.c_double),
                ("vr_poly", ctypes.c_double * 3), ("n_total_samples", ctypes.c_size_t), ("quirks", ctypes.c_int),
                ("imp", ctypes.c_int), ("azimuth_spacing", ctypes.c_double), ("reference_range", ctypes.c_double),
                ("grsr_poly", ctypes.c_double * 5), ("gr_range_size", ctypes.c_int),
                ("ground_spacing", ctypes.c_double), ("lines_before", ctypes.c_int), ("lines_after", ctypes.c_int),
                ("samples_per_line", ctypes.c_int), ("line_bytes", ctypes.c_long), ("swst_mult", ctypes.c_int)]


class FocusResults(ctypes.Structure):
    _fields_ = [("iq", IqResults), ("dc_poly", ctypes.c_double * 3), ("n_blocks", ctypes.c_int),
                ("block_freq", ctypes.c_double * 64), ("block_x", ctypes.c_double * 64), ("nrv", ctypes.c_int),
                ("crop_front", ctypes.c_int), ("crop_back", ctypes.c_int), ("crop_near", ctypes.c_int),
                ("crop_far", ctypes.c_int), ("imp_rows", ctypes.c_int),
                ("azimuth_core_fused", ctypes.c_int), ("rcmc_shift_bound", ctypes.c_int)]


_vp, _ci, _cd, _cf, _sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_float, ctypes.c_size_t
_P = ctypes.POINTER
class Osv(ctypes.Structure):
    _fields_ = [("t", ctypes.c_double), ("pos", ctypes.c_double * 3), ("vel", ctypes.c_double * 3)]


class Mjd(ctypes.Structure):
    _fields_ = [("days", ctypes.c_int), ("seconds", ctypes.c_uint), ("micros", ctypes.c_uint)]

    def astuple(self):
        return (self.days, self.seconds, self.micros)


class DorisMeta(ctypes.Structure):
    """asarf_doris_meta: L1 metadata of a DORIS orbit file."""
    _fields_ = [("orbit_name", ctypes.c_char * 64), ("vector_source", ctypes.c_char * 3), ("phase", ctypes.c_char),
                ("cycle", ctypes.c_uint), ("delta_ut1", ctypes.c_double), ("abs_orbit", ctypes.c_uint),
                ("quality", ctypes.c_uint), ("state_vector_time", Mjd), ("pos", ctypes.c_double * 3),
                ("vel", ctypes.c_double * 3)]


class InsSwath(ctypes.Structure):
    """asarf_ins_swath: an image-mode swath's parameters from the instrument characterisation file."""
    _fields_ = [("radar_frequency", ctypes.c_float), ("radar_sampling_rate", ctypes.c_float),
                ("range_gate_bias", ctypes.c_float), ("chirp_amplitude", ctypes.c_float * 4),
                ("chirp_phase", ctypes.c_float * 4), ("chirp_duration", ctypes.c_float),
                ("swst_rank", ctypes.c_ushort)]


class Timing(ctypes.Structure):
    """asarf_timing: the radar timing of the level-0 flow."""
    _fields_ = [("prf", ctypes.c_double), ("line_time_interval", ctypes.c_double),
                ("two_way_slant_range_time", ctypes.c_double), ("slant_range_first", ctypes.c_double),
                ("wavelength", ctypes.c_double), ("range_spacing", ctypes.c_double), ("chirp_kr", ctypes.c_double),
                ("chirp_duration", ctypes.c_double), ("chirp_bandwidth", ctypes.c_double),
                ("chirp_samples", ctypes.c_int)]


class Patc(ctypes.Structure):
    _fields_ = [("cyclic_counter", ctypes.c_ushort), ("mission", ctypes.c_char * 2), ("orbit_number", ctypes.c_uint),
                ("epoch_1950", ctypes.c_int), ("milliseconds", ctypes.c_int), ("sbt_counter", ctypes.c_uint),
                ("sbt_period", ctypes.c_uint)]


class PatcReference(ctypes.Structure):
    _fields_ = [("time_us", ctypes.c_longlong), ("sbt_counter", ctypes.c_uint), ("sbt_period_ns", ctypes.c_uint)]


PRODUCT_SAR_IM0, PRODUCT_ASA_IM0, PRODUCT_SAR_IMS, PRODUCT_SAR_IMP, PRODUCT_ASA_IMS, PRODUCT_ASA_IMP, \
    PRODUCT_UNIDENTIFIED = range(7)
INSTRUMENT_SAR, INSTRUMENT_ASAR = 0, 1


class AsarEcho(ctypes.Structure):
    _uc, _us = ctypes.c_ubyte, ctypes.c_ushort
    _fields_ = [("isp_sensing_time", Mjd), ("fep_time", Mjd), ("isp_length", _us), ("packet_id", _us),
                ("sequence_count", _us), ("packet_length", _us), ("segmentation_flags", _uc), ("mode_id", _uc),
                ("antenna_beam_set_no", _uc), ("comp_ratio", _uc), ("echo_flag", _uc), ("noise_flag", _uc),
                ("cal_flag", _uc), ("cal_type", _uc), ("cycle_count", _us), ("mode_count", ctypes.c_uint),
                ("onboard_time", ctypes.c_ulonglong), ("pri_code", _us), ("swst_code", _us),
                ("echo_window_code", _us), ("upconverter_raw", _uc), ("downconverter_raw", _uc), ("tx_pol", _uc),
                ("rx_pol", _uc), ("cal_row_number", _uc), ("tx_pulse_code", _us), ("beam_adj_delta", _uc),
                ("chirp_pulse_bw_code", _uc), ("aux_tx_monitor_level", _uc), ("resampling_factor", _us),
                ("sample_bytes", ctypes.c_int), ("samples", ctypes.c_int), ("dsr_bytes", ctypes.c_int)]


class ErsEcho(ctypes.Structure):
    _uc = ctypes.c_ubyte
    _fields_ = [("isp_sensing_time", Mjd), ("isp_length", ctypes.c_ushort), ("data_record_number", ctypes.c_uint),
                ("packet_counter", _uc), ("subcommutation_counter", _uc), ("obrc_orbit", _uc), ("activity_task", _uc),
                ("sample_flags", _uc), ("sbt", ctypes.c_uint), ("image_format_counter", ctypes.c_uint),
                ("swst_code", ctypes.c_ushort), ("pri_code", ctypes.c_ushort)]


class Dsd(ctypes.Structure):
    _fields_ = [("ds_name", ctypes.c_char * 64), ("ds_type", ctypes.c_char * 16), ("filename", ctypes.c_char * 96),
                ("ds_offset", ctypes.c_ulonglong), ("ds_size", ctypes.c_ulonglong), ("num_dsr", ctypes.c_ulonglong),
                ("dsr_size", ctypes.c_ulonglong)]


class TiePoints(ctypes.Structure):
    _fields_ = [("samp_numbers", ctypes.c_uint * 11), ("slant_range_times", ctypes.c_float * 11),
                ("angles", ctypes.c_float * 11), ("lats", ctypes.c_int * 11), ("longs", ctypes.c_int * 11)]


class ResultsWindow(ctypes.Structure):
    _fields_ = [("lines_before_sensing_start", ctypes.c_long), ("lines_after_sensing_start", ctypes.c_long),
                ("lines_before_sensing_stop", ctypes.c_long), ("lines_after_sensing_stop", ctypes.c_long),
                ("near_range_pixels", ctypes.c_int), ("far_range_pixels", ctypes.c_int)]


class Geometry(ctypes.Structure):
    _fields_ = [("osv", ctypes.POINTER(Osv)), ("n_osv", ctypes.c_int), ("first_line_time", ctypes.c_double),
                ("line_time_interval", ctypes.c_double), ("azimuth_size", ctypes.c_int),
                ("range_size", ctypes.c_int), ("chirp_samples", ctypes.c_int),
                ("slant_range_first", ctypes.c_double), ("range_spacing", ctypes.c_double),
                ("carrier_frequency", ctypes.c_double), ("prf", ctypes.c_double),
                ("az_bandwidth_fraction", ctypes.c_double), ("vr_poly", ctypes.c_double * 3),
                ("dc_poly", ctypes.c_double * 3), ("center_point", ctypes.c_double * 3)]


class SqThresholds(ctypes.Structure):
    """asarf_sq_thresholds: summary-quality thresholds of the processor configuration file."""
    _fields_ = [(n, ctypes.c_float) for n in (
        "thresh_chirp_broadening", "thresh_chirp_sidelobe", "thresh_chirp_islr", "thresh_input_mean",
        "exp_input_mean", "thresh_input_std_dev", "exp_input_std_dev", "thresh_dop_cen", "thresh_dop_amb",
        "thresh_output_mean", "exp_output_mean", "thresh_output_std_dev", "exp_output_std_dev",
        "thresh_input_missing_lines", "thresh_input_gaps")] + [("lines_per_gap", ctypes.c_uint)]


class L1Inputs(ctypes.Structure):
    """asarf_l1_inputs: the metadata of a level-1 header."""
    _cs = ctypes.c_char_p
    _us = ctypes.c_ushort
    _fields_ = [("l0_product_name", _cs), ("detected", ctypes.c_int), ("acquisition_station", _cs),
                ("processing_station", _cs), ("software_version", _cs), ("processing_time", Mjd),
                ("sensing_start", Mjd), ("sensing_stop", Mjd), ("first_line_time", Mjd), ("last_line_time", Mjd),
                ("orbit", ctypes.POINTER(DorisMeta)), ("rel_orbit", ctypes.c_int), ("utc_sbt_time", Mjd),
                ("sat_binary_time", ctypes.c_uint), ("clock_step", ctypes.c_uint), ("leap_utc", Mjd),
                ("leap_sign", ctypes.c_int), ("leap_err", ctypes.c_char), ("product_err", ctypes.c_int),
                ("swath", _cs), ("ascending", ctypes.c_int), ("polarization", _cs), ("echo_comp_method", _cs),
                ("echo_comp_ratio", _cs), ("init_cal_method", _cs), ("init_cal_ratio", _cs), ("per_cal_method", _cs),
                ("per_cal_ratio", _cs), ("noise_method", _cs), ("noise_ratio", _cs), ("configuration_file", _cs),
                ("instrument_file", _cs), ("external_characterization_file", _cs),
                ("external_calibration_file", _cs), ("orbit_dataset_name", _cs),
                ("geometry", ctypes.POINTER(Geometry)), ("azimuth_spacing", ctypes.c_double),
                ("two_way_slant_range_time", ctypes.c_double), ("range_sampling_rate", ctypes.c_double),
                ("chirp_window", ctypes.c_int), ("chirp_alpha", ctypes.c_double), ("chirp_kr", ctypes.c_double),
                ("azimuth_window_alpha", ctypes.c_double), ("grsr5", ctypes.POINTER(ctypes.c_double)),
                ("first_sbt", ctypes.c_uint), ("first_mjd", Mjd), ("first_swst_code", _us), ("last_swst_code", _us),
                ("pri_code", _us), ("tx_pulse_len_code", _us), ("tx_bw_code", _us), ("echo_win_len_code", _us),
                ("up_code", _us), ("down_code", _us), ("resamp_code", _us), ("beam_adj_code", _us),
                ("beam_set_num_code", _us), ("tx_monitor_code", _us), ("swst_changes", ctypes.c_uint),
                ("swst_rank", ctypes.c_uint), ("tx_monitor_value", ctypes.c_float), ("iq", IqResults),
                ("statistics", ctypes.c_double * 4), ("sq", SqThresholds), ("mds_records", ctypes.c_int),
                ("mds_record_size", ctypes.c_int)]


SIGNATURES = {
    "asarf_ctx_create": (_ci, [_ci, _vp, _P(_vp)]),
    "asarf_ctx_destroy": (_ci, [_vp]),
    "asarf_sync": (_ci, [_vp]),
    "asarf_ctx_stream": (_vp, [_vp]),
    "asarf_last_error": (ctypes.c_char_p, []),
    "asarf_version": (ctypes.c_char_p, []),
    "asarf_image_alloc": (_ci, [_vp, _ci, _ci, _ci, _ci, _ci, _P(Image)]),
    "asarf_image_free": (_ci, [_vp, _P(Image)]),
    "asarf_image_zero": (_ci, [_vp, _P(Image)]),
    "asarf_image_bytes": (_sz, [_P(Image)]),
    "asarf_workspace_alloc": (_ci, [_vp, _sz, _P(Workspace)]),
    "asarf_workspace_free": (_ci, [_vp, _P(Workspace)]),
    "asarf_image_adopt": (_ci, [_P(Image), _ci, _ci, _ci, _ci, _ci, _P(Workspace)]),
    "asarf_image_release": (_ci, [_P(Image), _P(Workspace)]),
    "asarf_image_subset_copy": (_ci, [_vp, _P(Image), _ci, _ci, _ci, _ci, _vp, _sz]),
    "asarf_fft_padding_size": (_ci, [_ci]),
    "asarf_fft_static_column_length": (_ci, [_ci]),
    "asarf_fft_fast_range_size": (_ci, [_ci]),
    "asarf_generate_chirp": (_ci, [_P(ChirpInfo), _ci, _vp]),
    "asarf_polyfit": (_ci, [_vp, _vp, _ci, _ci, _vp]),
    "asarf_polyval": (_cd, [_vp, _ci, _cd]),
    "asarf_iq_correction": (_ci, [_vp, _P(Image), _sz, _ci, _P(IqResults)]),
    "asarf_doppler_centroid": (_ci, [_vp, _P(Image), _cd, _ci, _vp, _vp, _vp, _ci, _P(_ci)]),
    "asarf_doppler_centroid_noalloc": (_ci, [_vp, _P(Image), _cd, _ci, _vp, _vp, _vp, _ci, _P(_ci)]),
    "asarf_range_compression": (_ci, [_vp, _P(Image), _vp, _ci, _ci, _P(Workspace)]),
    "asarf_range_doppler": (_ci, [_vp, _P(RdaParams), _P(Image), _P(Image), _P(Workspace)]),
    "asarf_rcmc_window_size": (_ci, []),
    "asarf_az_multilook": (_ci, [_vp, _P(Image), _cd, _P(Workspace), _P(_ci), _P(_cd)]),
    "asarf_ingest_ers_im": (_ci, [_vp, _vp, ctypes.c_long, _ci, _ci, _vp, _ci, _P(Image)]),
    "asarf_ingest_asar_fbaq4": (_ci, [_vp, _vp, ctypes.c_long, _ci, _vp, _ci, _vp, _vp, _ci, _P(Image)]),
    "asarf_fbaq4_index": (_ci, [_ci, _ci]),
    "asarf_interpolate_orbit": (_ci, [_P(Osv), _ci, _cd, _P(Osv)]),
    "asarf_geolocate": (_ci, [_P(_cd), _P(_cd), _P(_cd), _cd, _P(_cd)]),
    "asarf_incidence_angle": (_cd, [_P(_cd), _P(_cd)]),
    "asarf_estimate_processing_velocity": (_ci, [_P(Geometry), _P(_cd), _P(_cd), _P(_cd), _P(_ci)]),
    "asarf_srgr_fit": (_ci, [_P(Geometry), _cd, _P(_cd), _P(_ci)]),
    "asarf_geo_to_xyz": (_ci, [_cd, _cd, _cd, _P(_cd)]),
    "asarf_xyz_to_geo": (_ci, [_P(_cd), _P(_cd)]),
    "asarf_scene_center": (_ci, [_P(Geometry), _P(_cd), _P(_cd), _P(_cd)]),
    "asarf_aperture_pixels": (_ci, [_P(Geometry), _P(_ci), _ci, _P(_ci)]),
    "asarf_geolocation_tie_points": (_ci, [_P(Geometry), _ci, ctypes.c_void_p, _P(TiePoints)]),
    "asarf_con_summary_quality": (_ci, [ctypes.c_char_p, ctypes.c_long, _ci, _P(SqThresholds)]),
    "asarf_l1_header_size": (ctypes.c_long, [_ci]),
    "asarf_l1_header": (_ci, [_P(L1Inputs), ctypes.c_void_p, ctypes.c_long, _P(ctypes.c_long)]),
    "asarf_doris_orbit": (_ci, [ctypes.c_char_p, ctypes.c_long, _P(Mjd), _P(Mjd), _P(Osv), _P(Mjd), _ci, _P(_ci),
                                _P(DorisMeta)]),
    "asarf_doris_filename_window": (_ci, [ctypes.c_char_p, _P(Mjd), _P(Mjd)]),
    "asarf_con_process_gains": (_ci, [ctypes.c_char_p, ctypes.c_long, _vp, _vp]),
    "asarf_xca_scaling_factors": (_ci, [ctypes.c_char_p, ctypes.c_long, _vp, _vp]),
    "asarf_calibration_constant": (_ci, [_ci, _ci, ctypes.c_float, ctypes.c_float, _P(ctypes.c_double)]),
    "asarf_ins_swath_params": (_ci, [ctypes.c_char_p, ctypes.c_long, _ci, _P(InsSwath)]),
    "asarf_radar_timing": (_ci, [_P(InsSwath), _ci, _ci, _ci, _ci, _P(Timing)]),
    "asarf_mjd_compare": (_ci, [_P(Mjd), _P(Mjd)]),
    "asarf_mjd_add_micros": (_ci, [_P(Mjd), ctypes.c_longlong, _P(Mjd)]),
    "asarf_counter_gap": (ctypes.c_ulonglong, [ctypes.c_ulonglong, ctypes.c_ulonglong, ctypes.c_ulonglong]),
    "asarf_parse_patc": (_ci, [ctypes.c_char_p, ctypes.c_long, _P(Patc)]),
    "asarf_patc_register": (_ci, [_P(Patc), _P(PatcReference)]),
    "asarf_adjust_isp_sensing_time": (_ci, [_P(PatcReference), _P(Mjd), ctypes.c_uint, _ci]),
    "asarf_product_type_from_name": (_ci, [ctypes.c_char_p]),
    "asarf_target_product": (_ci, [_ci, ctypes.c_char_p, _P(_ci)]),
    "asarf_instrument": (_ci, [_ci, _P(_ci)]),
    "asarf_parse_asar_im_packet": (_ci, [ctypes.c_char_p, ctypes.c_long, _P(AsarEcho), ctypes.c_void_p, ctypes.c_long]),
    "asarf_parse_ers_im_packet": (_ci, [ctypes.c_char_p, ctypes.c_long, _P(ErsEcho), ctypes.c_void_p]),
    "asarf_header_get": (_ci, [ctypes.c_char_p, ctypes.c_long, ctypes.c_char_p, ctypes.c_char_p, _ci]),
    "asarf_level0_mdsr": (_ci, [ctypes.c_char_p, ctypes.c_long, _P(Dsd)]),
    "asarf_ins_fbaq4_tables": (_ci, [ctypes.c_char_p, ctypes.c_long, ctypes.c_void_p, ctypes.c_void_p]),
    "asarf_ers_forecast": (_ci, [ctypes.c_char_p, ctypes.c_long, ctypes.c_long, _P(Mjd), _P(Mjd), _P(PatcReference),
                                 _P(ctypes.c_long), _P(ctypes.c_long), _P(ctypes.c_long), _P(Mjd), ctypes.c_long,
                                 _P(ctypes.c_long)]),
    "asarf_asar_forecast": (_ci, [ctypes.c_char_p, ctypes.c_long, ctypes.c_long, _P(Mjd), _P(Mjd),
                                  _P(ctypes.c_long), _P(ctypes.c_long), _P(ctypes.c_long), _P(Mjd), ctypes.c_long,
                                  _P(ctypes.c_long)]),
    "asarf_results_window": (_ci, [_cd, ctypes.c_long, ctypes.c_long, ctypes.c_long, _ci, _P(ResultsWindow)]),
    "asarf_condition_slc": (_ci, [_vp, _P(Image), ctypes.c_float, _ci, _vp, ctypes.c_size_t]),
    "asarf_condition_detected": (_ci, [_vp, _vp, _ci, _ci, ctypes.c_float, _ci, _vp, ctypes.c_size_t]),
    "asarf_image_statistics": (_ci, [_vp, _P(Image), _ci, _P(_cd)]),
    "asarf_range_spreading_loss": (_ci, [_vp, _vp, _ci, _ci, _cd, _cd, _cd]),
    "asarf_srgr": (_ci, [_vp, _vp, _ci, _ci, _vp, _cd, _cd, _ci, _vp]),
    "asarf_srgr_fit_spherical": (_ci, [_cd, _cd, _ci, _cd, _cd, _cd, _vp, _P(_ci)]),
    "asarf_fft_c2c": (_ci, [_vp, _vp, _ci, _ci, ctypes.c_long, _ci, _ci, _cf]),
    "asarf_plan_create": (_ci, [_vp, _P(FocusParams), _vp, _P(_vp)]),
    "asarf_plan_destroy": (_ci, [_vp]),
    "asarf_plan_device_bytes": (_sz, [_vp]),
    "asarf_plan_execute": (_ci, [_vp, _vp, _P(Image), _vp]),
    "asarf_plan_execute_lines": (_ci, [_vp, _vp, _vp, _ci, _P(Image), _vp]),
    "asarf_plan_execute_fbaq4": (_ci, [_vp, _vp, _vp, _ci, _vp, _vp, _ci, _P(Image), _vp]),
    "asarf_plan_imp_capacity": (_ci, [_vp, _P(_ci), _P(_ci)]),
    "asarf_plan_results": (_ci, [_vp, _P(FocusResults)]),
    "asarf_plan_set_timing": (_ci, [_vp, _ci]),
    "asarf_plan_kernel_times": (_ci, [_vp, _P(ctypes.c_char_p), _P(_cf), _ci, _P(_ci)]),
    "asarf_rcmc_net_create": (_ci, [_vp, ctypes.c_ulonglong, _P(_vp)]),
    "asarf_rcmc_net_destroy": (_ci, [_vp]),
    "asarf_rcmc_net_num_params": (_ci, []),
    "asarf_rcmc_net_get_params": (_ci, [_vp, _P(_cf), _ci]),
    "asarf_rcmc_net_set_params": (_ci, [_vp, _P(_cf), _ci]),
    "asarf_rcmc_net_grad": (_ci, [_vp, _vp, ctypes.c_long, _P(_cf), _P(_cd)]),
    "asarf_rcmc_net_train": (_ci, [_vp, _vp, ctypes.c_long, _ci, _ci, _cf, _P(_cf)]),
    "asarf_rcmc_net_build_table": (_ci, [_vp, _vp, ctypes.c_long, _ci, _ci, _ci, _cf, _P(_cf), _P(_cf), _P(_ci)]),
    "asarf_plan_rcmc_pairs": (_ci, [_vp, _vp, ctypes.c_long, ctypes.c_ulonglong, _vp]),
    "asarf_plan_set_rcmc": (_ci, [_vp, _ci, _vp]),
    "asarf_fft_c2c_split_bf16": (_ci, [_vp, _vp, _ci, _ci, ctypes.c_long, _ci, _ci, _cf]),
    "asarf_plan_set_fft": (_ci, [_vp, _ci]),
    "asarf_autotune_fft": (_ci, [_vp, _ci, _ci, _ci, _ci, _P(_ci), _P(_cf), _P(_ci)]),
    "asarf_tune_save": (_ci, [_vp, ctypes.c_char_p]),
    "asarf_tune_load": (_ci, [_vp, ctypes.c_char_p, _P(_ci)]),
    "asarf_batch_create": (_ci, [_vp, _ci, _P(_vp)]),
    "asarf_batch_add": (_ci, [_vp, _P(FocusParams), _vp, _vp, _vp, _P(_ci)]),
    "asarf_batch_execute": (_ci, [_vp, _ci]),
    "asarf_batch_sync": (_ci, [_vp]),
    "asarf_batch_size": (_ci, [_vp]),
    "asarf_batch_slc": (_ci, [_vp, _ci, _P(Image)]),
    "asarf_batch_results": (_ci, [_vp, _ci, _P(FocusResults)]),
    "asarf_batch_destroy": (_ci, [_vp]),
}

_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (the HIP library is required; run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def check(status: int):
    if status != OK:
        msg = lib().asarf_last_error()
        raise AsarfError(status, msg.decode() if msg else "")
    return status


def last_error() -> str:
    m = lib().asarf_last_error()
    return m.decode() if m else ""
