"""ctypes binding of VHSDecode.Hip.Native (include/vhsdecode_hip.h).

This is the binding a managed caller would write (the reference's P/Invoke
declarations, src/VHSDecode.Core/Dsp/Ipp/IppNativeMethods.cs:6-120 and
Dsp/CudaFast/CudaFastNativeRuntime.cs:131-800, are the model).  Loading fails
loudly when the library is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

_PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.environ.get("VHSDECODE_HIP_LIBRARY", os.path.join(_PKG_ROOT, "lib", "libvhsdecode_hip.so"))

ABI_VERSION = 1
NOALLOC_MAX_SAMPLES = 16 * 1024 * 1024
NOALLOC_MAX_EDGES = 448 * 1024

REQUIRED_EXPORTS = (
    "vhsdecode_hip_get_abi_version",
    "vhsdecode_hip_get_runtime_info",
    "vhsdecode_hip_get_last_error",
    "vhsdecode_hip_demod_create",
    "vhsdecode_hip_demod_destroy",
    "vhsdecode_hip_demod_span_device",
    "vhsdecode_hip_demod_span_host",
    "vhsdecode_hip_find_pulses",
    "vhsdecode_hip_refine_hsync",
    "vhsdecode_hip_set_sinc_table",
    "vhsdecode_hip_tbc_resample",
    "vhsdecode_hip_tbc_resample_device",
    "vhsdecode_hip_decode_fields_device",
    "vhsdecode_hip_serve_create",
    "vhsdecode_hip_serve_add_format",
    "vhsdecode_hip_serve_add_stream",
    "vhsdecode_hip_serve_submit",
    "vhsdecode_hip_serve_wait",
    "vhsdecode_hip_serve_destroy",
    "vhsdecode_hip_set_profiling",
    "vhsdecode_hip_get_profile",
    "vhsdecode_hip_debug_fft16k",
    "vhsdecode_hip_debug_replace_spikes",
    "vhsdecode_hip_set_ld_audio_phase2", "vhsdecode_hip_ld_audio_phase2_device",
    "vhsdecode_hip_batch_create",
    "vhsdecode_hip_batch_info",
    "vhsdecode_hip_batch_run_device",
    "vhsdecode_hip_batch_destroy",
    "vhsdecode_hip_hifi_create",
    "vhsdecode_hip_hifi_destroy",
    "vhsdecode_hip_hifi_demodulate_device",
    "vhsdecode_hip_hifi_demodulate_host",
    "vhsdecode_hip_run",
    "vhsdecode_hip_field_locations_device",
    "vhsdecode_hip_detect_dropouts_device",
    "vhsdecode_hip_find_pulses_noalloc",
    "vhsdecode_hip_decode_vhs_fields_device",
    "vhsdecode_hip_vhs_field_lines_device",
    "vhsdecode_hip_set_vhs_chroma",
    "vhsdecode_hip_decode_vhs_fields_chroma_device",
)

KERNEL_NAMES = ("block_demod", "spike_repair", "chroma", "pulses", "hsync", "tbc", "sub_deemphasis", "block_video",
                "locate", "dropouts", "chroma_field")

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_i16p = ctypes.POINTER(ctypes.c_int16)
_i32p = ctypes.POINTER(ctypes.c_int32)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_u16p = ctypes.POINTER(ctypes.c_uint16)
_u32p = ctypes.POINTER(ctypes.c_uint32)


class RuntimeInfoV1(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("device_count", ctypes.c_int32), ("device_id", ctypes.c_int32),
                ("compute_units", ctypes.c_int32), ("hbm_bytes", ctypes.c_uint64),
                ("lds_bytes_per_workgroup", ctypes.c_int32), ("device_name", ctypes.c_char * 128),
                ("gcn_arch", ctypes.c_char * 64)]


class SubDeemphV1(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("high_pass", _dp), ("amplitude_sos", _dp),
                ("deviation", ctypes.c_double), ("exponential_scaling", ctypes.c_double),
                ("scaling1", ctypes.c_double), ("scaling2", ctypes.c_double), ("logistic_mid", ctypes.c_double),
                ("logistic_rate", ctypes.c_double), ("static_factor", ctypes.c_double)]


class BatchSegmentV1(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("context", ctypes.c_void_p), ("rf_offset", ctypes.c_int64),
                ("length", ctypes.c_int64), ("out_offset", ctypes.c_int64)]


class BatchConfigV1(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("segments", ctypes.POINTER(BatchSegmentV1)),
                ("count", ctypes.c_int32), ("use_graph", ctypes.c_int32)]


class HifiConfigV1(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("device_id", ctypes.c_int32), ("block_length", ctypes.c_int64),
                ("sample_rate_hz", ctypes.c_int32), ("carrier_hz", ctypes.c_double), ("deviation_hz", ctypes.c_double)]


class LdAudioV1(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("bin_count", ctypes.c_int32), ("low_bin", ctypes.c_int32 * 2),
                ("stage1", _dp * 2), ("slice_rate_hz", ctypes.c_double), ("low_frequency_hz", ctypes.c_double * 2)]


class LaserDiscV1(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("rf_mtf", _dp), ("clip_low_hz", ctypes.c_double),
                ("clip_high_hz", ctypes.c_double), ("video_burst", _dp), ("video_burst_offset", ctypes.c_int32),
                ("video_pilot", _dp), ("analog_audio", ctypes.POINTER(LdAudioV1))]


class FiltersV1(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("block_length", ctypes.c_int32),
                ("sample_rate_hz", ctypes.c_double), ("rf_video", _dp), ("video", _dp),
                ("video_lowpass05", _dp), ("video_lowpass05_offset", ctypes.c_int32), ("envelope_sos", _dp),
                ("envelope_sections", ctypes.c_int32), ("chroma_sos", _dp), ("chroma_sections", ctypes.c_int32),
                ("chroma_offset", ctypes.c_int32), ("diff_demod_max", ctypes.c_double),
                ("plane_bias_hz", ctypes.c_double), ("sub_deemphasis", ctypes.POINTER(SubDeemphV1)),
                ("rf_boost_sos", _dp), ("rf_boost_multiplier", ctypes.c_double), ("video_notch", _dp),
                ("laserdisc", ctypes.POINTER(LaserDiscV1))]


class DemodConfigV1(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("device_id", ctypes.c_int32), ("block_cut", ctypes.c_int32),
                ("block_cut_end", ctypes.c_int32), ("max_blocks", ctypes.c_int32),
                ("filters", ctypes.POINTER(FiltersV1))]


class SpanV1(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("video", _fp), ("video_lowpass", _fp), ("envelope", _fp),
                ("chroma", _fp), ("block_flags", _u32p), ("pilot", _fp), ("audio_left", _fp),
                ("audio_right", _fp), ("chroma_stream", ctypes.c_void_p)]


class HsyncParamsV1(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("threshold", ctypes.c_double),
                ("sample_rate_mhz", ctypes.c_double), ("hsync_pulse_us", ctypes.c_double),
                ("ire0", ctypes.c_double), ("hz_ire", ctypes.c_double), ("is_pal", ctypes.c_int32),
                ("use_right_hsync", ctypes.c_int32)]


class TbcParamsV1(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("output_line_length", ctypes.c_int32),
                ("first_line", ctypes.c_int32), ("line_count", ctypes.c_int32),
                ("nominal_line_length", ctypes.c_double), ("wow_level_smoothing", ctypes.c_double),
                ("convert", ctypes.c_int32), ("ire0", ctypes.c_double), ("hz_ire", ctypes.c_double),
                ("output_zero", ctypes.c_int32), ("vsync_ire", ctypes.c_double), ("output_scale", ctypes.c_double)]


class FieldParamsV1(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("threshold", ctypes.c_double),
                ("sample_rate_mhz", ctypes.c_double), ("hsync_pulse_us", ctypes.c_double),
                ("ire0", ctypes.c_double), ("hz_ire", ctypes.c_double), ("is_pal", ctypes.c_int32),
                ("use_right_hsync", ctypes.c_int32), ("output_line_length", ctypes.c_int32),
                ("field_lines", ctypes.c_int32), ("nominal_line_length", ctypes.c_double),
                ("wow_level_smoothing", ctypes.c_double), ("output_zero", ctypes.c_int32),
                ("vsync_ire", ctypes.c_double), ("output_scale", ctypes.c_double)]


class LdAudioPhase2V1(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("block_length", ctypes.c_int32),
                ("center_hz", ctypes.c_double * 2), ("stage2", ctypes.c_void_p * 2)]


class ChromaParamsV1(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("is_ntsc", ctypes.c_int32), ("fsc_mhz", ctypes.c_double),
                ("color_under_hz", ctypes.c_double), ("output_rate_hz", ctypes.c_double),
                ("field_lines", ctypes.c_int32), ("output_line_length", ctypes.c_int32),
                ("burst_start", ctypes.c_int32), ("burst_end", ctypes.c_int32), ("burst_abs_ref", ctypes.c_double),
                ("rotation", ctypes.c_int32 * 2), ("sections", ctypes.c_int32), ("input_line_length", ctypes.c_int32),
                ("final_sos", (ctypes.c_double * 6) * 8)]


class VhsSyncV1(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("is_pal", ctypes.c_int32), ("line_period_us", ctypes.c_double),
                ("hsync_pulse_us", ctypes.c_double), ("eq_pulse_us", ctypes.c_double),
                ("vsync_pulse_us", ctypes.c_double), ("num_pulses", ctypes.c_int32),
                ("hsync_tolerance_us", ctypes.c_double), ("eq_tolerance_us", ctypes.c_double),
                ("frames_per_second", ctypes.c_double), ("frame_lines", ctypes.c_double),
                ("output_rate_mhz", ctypes.c_double), ("read_samples", ctypes.c_int64), ("block_length", ctypes.c_int32), ("block_cut", ctypes.c_int32),
                ("block_stride", ctypes.c_int32)]


class VhsFieldV1(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("status", ctypes.c_int32), ("begin", ctypes.c_int64),
                ("window_start", ctypes.c_int64), ("window_length", ctypes.c_int64),
                ("level_source", ctypes.c_int32), ("raw_pulses", ctypes.c_int32), ("refined_pulses", ctypes.c_int32),
                ("is_first_field", ctypes.c_int32), ("parity_confidence", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("sync_level", ctypes.c_double), ("blank_level", ctypes.c_double),
                ("dc_offset", ctypes.c_double), ("threshold", ctypes.c_double), ("hsync_median", ctypes.c_double),
                ("mean_line_length", ctypes.c_double), ("line0", ctypes.c_double), ("first_hsync", ctypes.c_double),
                ("first_hsync_line", ctypes.c_double), ("next_offset", ctypes.c_double),
                ("next_begin", ctypes.c_int64)]


class DropoutParamsV1(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("threshold_fraction", ctypes.c_double),
                ("absolute_threshold", ctypes.c_double), ("hysteresis", ctypes.c_double),
                ("merge_threshold", ctypes.c_int32), ("minimum_length", ctypes.c_int32),
                ("output_line_length", ctypes.c_int32), ("first_field_lines", ctypes.c_int32),
                ("second_field_lines", ctypes.c_int32), ("capacity", ctypes.c_int32)]


class ServeConfigV1(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("device_id", ctypes.c_int32), ("max_slots", ctypes.c_int32),
                ("hbm_budget_bytes", ctypes.c_uint64), ("max_window_samples", ctypes.c_int64),
                ("max_fields", ctypes.c_int32)]


class ServeFormatV1(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("filters", ctypes.POINTER(FiltersV1)),
                ("block_cut", ctypes.c_int32), ("block_cut_end", ctypes.c_int32),
                ("field", ctypes.POINTER(FieldParamsV1)), ("sync", ctypes.POINTER(VhsSyncV1)),
                ("sinc_table", _fp), ("sinc_count", ctypes.c_int64)]


READ_CALLBACK = ctypes.CFUNCTYPE(ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_size_t)
CANCEL_CALLBACK = ctypes.CFUNCTYPE(ctypes.c_int32, ctypes.c_void_p)
INPUT_FLOAT32, INPUT_INT16 = 0, 1


class RunConfigV1(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("device_id", ctypes.c_int32),
                ("format", ctypes.POINTER(ServeFormatV1)), ("dropouts", ctypes.POINTER(DropoutParamsV1)),
                ("total_samples", ctypes.c_uint64), ("input_sample_format", ctypes.c_uint32),
                ("maximum_output_fields", ctypes.c_uint32), ("span_samples", ctypes.c_int64),
                ("output_base_utf8", ctypes.c_char_p), ("overwrite", ctypes.c_int32),
                ("read_callback", READ_CALLBACK), ("cancel_callback", CANCEL_CALLBACK),
                ("user_data", ctypes.c_void_p), ("system_name", ctypes.c_char_p),
                ("colour_burst_start_us", ctypes.c_double), ("colour_burst_end_us", ctypes.c_double),
                ("chroma", ctypes.POINTER(ChromaParamsV1))]


class RunResultV1(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("fields_written", ctypes.c_uint32),
                ("output_line_length", ctypes.c_uint32), ("output_field_lines", ctypes.c_uint32),
                ("elapsed_seconds", ctypes.c_double), ("spans", ctypes.c_uint32), ("dropout_rows", ctypes.c_uint32)]


class WindowV1(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("stream_id", ctypes.c_int32), ("first_sample", ctypes.c_int64),
                ("sample_count", ctypes.c_int64), ("max_fields", ctypes.c_int32), ("fields_out", ctypes.c_void_p)]


class WindowResultV1(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("status", ctypes.c_int32), ("field_count", ctypes.c_int32),
                ("queued_ms", ctypes.c_double), ("latency_ms", ctypes.c_double)]


class ProfileV1(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("launches", ctypes.c_int64 * len(KERNEL_NAMES)),
                ("total_ms", ctypes.c_double * len(KERNEL_NAMES)), ("units", ctypes.c_int64 * len(KERNEL_NAMES))]


_lib = None


class HipError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"VHSDecode.Hip.Native status {status}: {message}")
        self.status = status


def load():
    """Load and bind the native library (raises if it is absent)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(
            f"VHSDecode.Hip.Native is not built ({LIB_PATH}); run `make -C vhs-decode-dotnet_amd` "
            "or __graft_entry__.build()")
    L = ctypes.CDLL(LIB_PATH)
    for name in REQUIRED_EXPORTS:
        getattr(L, name)
    L.vhsdecode_hip_get_abi_version.restype = ctypes.c_int32
    L.vhsdecode_hip_get_last_error.restype = ctypes.c_char_p
    L.vhsdecode_hip_get_runtime_info.argtypes = [ctypes.c_int32, ctypes.POINTER(RuntimeInfoV1)]
    L.vhsdecode_hip_demod_create.argtypes = [ctypes.POINTER(DemodConfigV1), ctypes.POINTER(ctypes.c_void_p)]
    L.vhsdecode_hip_demod_destroy.argtypes = [ctypes.c_void_p]
    L.vhsdecode_hip_demod_destroy.restype = None
    L.vhsdecode_hip_demod_span_device.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32,
                                                  ctypes.POINTER(SpanV1), ctypes.c_void_p]
    L.vhsdecode_hip_demod_span_host.argtypes = [ctypes.c_void_p, _i16p, ctypes.c_int64, ctypes.c_int64,
                                                ctypes.c_int32, ctypes.POINTER(SpanV1)]
    L.vhsdecode_hip_find_pulses.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_double,
                                            ctypes.c_int32, ctypes.c_int32, _i32p, _i32p, ctypes.c_int32, _i32p]
    L.vhsdecode_hip_find_pulses_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_double,
                                                    ctypes.c_double, ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p,
                                                    ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p,
                                                    ctypes.c_void_p]
    L.vhsdecode_hip_refine_hsync.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, _dp, _u8p,
                                             ctypes.c_int32, ctypes.POINTER(HsyncParamsV1), _dp, _u8p]
    L.vhsdecode_hip_set_sinc_table.argtypes = [ctypes.c_void_p, _fp, ctypes.c_int64]
    L.vhsdecode_hip_tbc_resample.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, _dp, ctypes.c_int32,
                                             ctypes.POINTER(TbcParamsV1), _u16p, _dp]
    L.vhsdecode_hip_tbc_resample_device.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                                    ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                                    ctypes.POINTER(TbcParamsV1), ctypes.c_void_p, ctypes.c_void_p,
                                                    ctypes.c_void_p]
    L.vhsdecode_hip_serve_create.argtypes = [ctypes.POINTER(ServeConfigV1), ctypes.POINTER(ctypes.c_void_p)]
    L.vhsdecode_hip_serve_add_format.argtypes = [ctypes.c_void_p, ctypes.POINTER(ServeFormatV1),
                                                 ctypes.POINTER(ctypes.c_int32)]
    L.vhsdecode_hip_serve_add_stream.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_int64,
                                                 ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]
    L.vhsdecode_hip_serve_submit.argtypes = [ctypes.c_void_p, ctypes.POINTER(WindowV1),
                                             ctypes.POINTER(ctypes.c_uint64)]
    L.vhsdecode_hip_serve_wait.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int32,
                                           ctypes.POINTER(WindowResultV1)]
    L.vhsdecode_hip_serve_destroy.argtypes = [ctypes.c_void_p]
    L.vhsdecode_hip_serve_destroy.restype = None
    L.vhsdecode_hip_field_locations_device.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p),
                                                       ctypes.POINTER(ctypes.c_int32)]
    L.vhsdecode_hip_detect_dropouts_device.argtypes = [
        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p,
        ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(DropoutParamsV1)] + [ctypes.c_void_p] * 8
    L.vhsdecode_hip_decode_fields_device.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                     ctypes.c_int64, _dp, _u8p, ctypes.POINTER(ctypes.c_int64),
                                                     ctypes.c_int32,
                                                     ctypes.POINTER(FieldParamsV1), ctypes.c_void_p, _dp,
                                                     ctypes.c_void_p]
    L.vhsdecode_hip_set_profiling.argtypes = [ctypes.c_void_p, ctypes.c_int32]
    L.vhsdecode_hip_get_profile.argtypes = [ctypes.c_void_p, ctypes.POINTER(ProfileV1)]
    L.vhsdecode_hip_debug_fft16k.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                                             ctypes.c_void_p]
    L.vhsdecode_hip_batch_create.argtypes = [ctypes.POINTER(BatchConfigV1), ctypes.POINTER(ctypes.c_void_p)]
    L.vhsdecode_hip_batch_info.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64),
                                           ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]
    L.vhsdecode_hip_batch_run_device.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                                 ctypes.POINTER(SpanV1), ctypes.c_void_p]
    L.vhsdecode_hip_batch_destroy.argtypes = [ctypes.c_void_p]
    L.vhsdecode_hip_batch_destroy.restype = None
    L.vhsdecode_hip_debug_replace_spikes.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                                                     ctypes.c_void_p, ctypes.c_void_p]
    L.vhsdecode_hip_set_ld_audio_phase2.argtypes = [ctypes.c_void_p, ctypes.POINTER(LdAudioPhase2V1)]
    L.vhsdecode_hip_ld_audio_phase2_device.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                       ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                                                       ctypes.c_void_p]
    L.vhsdecode_hip_hifi_create.argtypes = [ctypes.POINTER(HifiConfigV1), ctypes.POINTER(ctypes.c_void_p)]
    L.vhsdecode_hip_hifi_destroy.argtypes = [ctypes.c_void_p]
    L.vhsdecode_hip_hifi_destroy.restype = None
    L.vhsdecode_hip_hifi_demodulate_device.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                       ctypes.c_int32, ctypes.c_void_p]
    L.vhsdecode_hip_hifi_demodulate_host.argtypes = [ctypes.c_void_p, _fp, _fp, ctypes.c_int32]
    L.vhsdecode_hip_run.argtypes = [ctypes.POINTER(RunConfigV1), ctypes.POINTER(RunResultV1)]
    L.vhsdecode_hip_decode_vhs_fields_device.argtypes = [
        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int32,
        ctypes.c_int64, ctypes.POINTER(VhsSyncV1), ctypes.POINTER(FieldParamsV1), ctypes.c_int32, ctypes.c_int32,
        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    L.vhsdecode_hip_set_vhs_chroma.argtypes = [ctypes.c_void_p, ctypes.POINTER(ChromaParamsV1)]
    L.vhsdecode_hip_decode_vhs_fields_chroma_device.argtypes = [
        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64,
        ctypes.c_int32, ctypes.c_int64, ctypes.POINTER(VhsSyncV1), ctypes.POINTER(FieldParamsV1), ctypes.c_int32,
        ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p,
        ctypes.c_void_p, ctypes.c_void_p]
    L.vhsdecode_hip_vhs_field_lines_device.argtypes = [ctypes.c_void_p, ctypes.c_int32,
                                                       ctypes.POINTER(ctypes.c_void_p),
                                                       ctypes.POINTER(ctypes.c_int32)]
    for name in REQUIRED_EXPORTS:
        fn = getattr(L, name)
        if name not in ("vhsdecode_hip_get_last_error", "vhsdecode_hip_demod_destroy"):
            fn.restype = ctypes.c_int32
    if L.vhsdecode_hip_get_abi_version() != ABI_VERSION:
        raise HipError(-30006, "ABI version mismatch")
    _lib = L
    return L


def check(status: int):
    if status != 0:
        msg = load().vhsdecode_hip_get_last_error()
        raise HipError(status, msg.decode("utf-8", "replace") if msg else "")
