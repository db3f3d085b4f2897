"""ctypes view of the CPU oracle (oracle/oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, where it is the checker and the CPU baseline.
The product library never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_u16p = ctypes.POINTER(ctypes.c_uint16)


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.oracle_fft.argtypes = [_dp, ctypes.c_int, ctypes.c_int]
        L.oracle_vhs_atan2.argtypes = [ctypes.c_float, ctypes.c_float]
        L.oracle_vhs_atan2.restype = ctypes.c_float
        L.oracle_vhs_unwrap.argtypes = [_dp, ctypes.c_int, ctypes.c_double, _dp]
        L.oracle_sosfiltfilt_f32.argtypes = [_dp, ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int, _fp]
        L.oracle_demod_block.argtypes = [_dp, ctypes.c_int, ctypes.c_double, _dp, _dp, _dp, ctypes.c_int,
                                         _dp, ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                         _dp, _dp, _dp, _fp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.oracle_demod_block.restype = ctypes.c_int
        L.oracle_demod_block_ld.argtypes = [_dp, ctypes.c_int, ctypes.c_double, _dp, _dp, _dp, _dp, ctypes.c_int,
                                            _dp, ctypes.c_int, _dp, ctypes.c_double, ctypes.c_double,
                                            _dp, _dp, _dp, _dp, _dp, _dp]
        L.oracle_conj_unwrap.argtypes = [_dp, ctypes.c_int, ctypes.c_double, _dp]
        L.oracle_replace_spikes.argtypes = [_dp, _dp, ctypes.c_int, ctypes.c_double]
        L.oracle_ld_audio_channel.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp,
                                              ctypes.c_double, ctypes.c_double, _dp]
        L.oracle_hifi_analytic_phase.argtypes = [_fp, ctypes.c_int]
        L.oracle_hifi_demodulate_phase.argtypes = [_fp, ctypes.c_int, _fp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.oracle_pairwise_sum_f32.argtypes = [_fp, ctypes.c_int64]
        L.oracle_pairwise_sum_f32.restype = ctypes.c_float
        L.oracle_find_dropouts.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_int, ctypes.c_int, _ip, _ip, ctypes.c_int]
        L.oracle_find_dropouts.restype = ctypes.c_int
        L.oracle_map_tape_rf_to_tbc.argtypes = [_dp, _dp, ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                ctypes.c_int, ctypes.c_int, _ip, _ip, _ip, ctypes.c_int]
        L.oracle_map_tape_rf_to_tbc.restype = ctypes.c_int
        L.oracle_tape_dropouts.argtypes = [_fp, ctypes.c_int64, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int64,
                                           ctypes.c_int64, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_int, ctypes.c_int, ctypes.c_int, _ip, _ip, _ip, _ip, _ip, _ip,
                                           ctypes.c_int, _fp]
        L.oracle_tape_dropouts.restype = ctypes.c_int
        L.oracle_estimate_read_samples.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int64]
        L.oracle_estimate_read_samples.restype = ctypes.c_int64
        L.oracle_resolve_read_window.argtypes = [ctypes.c_int64] * 5 + [ctypes.POINTER(ctypes.c_int64)] * 2
        L.oracle_find_pulses.argtypes = [_dp, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                         _ip, _ip, ctypes.c_int]
        L.oracle_find_pulses.restype = ctypes.c_int
        L.oracle_zero_crossing.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                           ctypes.c_int, ctypes.c_int, _dp]
        L.oracle_zero_crossing.restype = ctypes.c_int
        L.oracle_refine_hsync.argtypes = [_dp, ctypes.c_int, _dp, _u8p, ctypes.c_int, ctypes.c_double,
                                          ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                          ctypes.c_int, ctypes.c_int, _dp, _u8p]
        L.oracle_tbc_resample.argtypes = [_dp, ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_double, ctypes.c_double, _fp, _dp,
                                          ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                          ctypes.c_double, ctypes.c_double, _u16p]
        L.oracle_tbc_resample.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


def fft(x: np.ndarray, inverse: bool = False) -> np.ndarray:
    buf = np.ascontiguousarray(x, dtype=np.complex128).copy()
    lib().oracle_fft(_p(buf.view(np.float64), _dp), buf.size, 1 if inverse else 0)
    return buf


def vhs_atan2(y: float, x: float) -> float:
    return lib().oracle_vhs_atan2(y, x)


def vhs_unwrap(analytic: np.ndarray, fs: float) -> np.ndarray:
    a = np.ascontiguousarray(analytic, dtype=np.complex128)
    out = np.empty(a.size)
    lib().oracle_vhs_unwrap(_p(a.view(np.float64), _dp), a.size, fs, _p(out, _dp))
    return out


def sosfiltfilt_f32(sos: np.ndarray, x: np.ndarray, pad: int = -1) -> np.ndarray:
    s = np.ascontiguousarray(sos, dtype=np.float64)
    xi = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty(xi.size, dtype=np.float32)
    lib().oracle_sosfiltfilt_f32(_p(s, _dp), s.shape[0], _p(xi, _dp), xi.size, pad, _p(out, _fp))
    return out


class DemodResult:
    def __init__(self, video, video_lowpass, envelope, chroma, demod, weak):
        self.video = video
        self.video_lowpass = video_lowpass
        self.envelope = envelope
        self.chroma = chroma
        self.demod = demod
        self.weak = weak


def demod_block(x: np.ndarray, filters, fs: Optional[float] = None) -> DemodResult:
    """One block through the exact VHS chain; `filters` is a TapeFilterSet."""
    xi = np.ascontiguousarray(x, dtype=np.float64)
    n = xi.size
    fs = float(filters.sample_rate_hz if fs is None else fs)
    rv = np.ascontiguousarray(filters.rf_video, dtype=np.complex128)
    vf = np.ascontiguousarray(filters.video, dtype=np.complex128)
    lf = np.ascontiguousarray(filters.video_lowpass05, dtype=np.complex128)
    env = np.ascontiguousarray(filters.envelope_sos, dtype=np.float64)
    chs = filters.chroma_sos
    ch = np.ascontiguousarray(chs if chs is not None else np.zeros((0, 6)), dtype=np.float64)
    video = np.empty(n)
    vlp = np.empty(n)
    envelope = np.empty(n)
    chroma = np.empty(n, dtype=np.float32)
    demod = np.empty(n)
    dmax = filters.diff_demod_max if filters.diff_demod_max is not None else -1.0
    sd = getattr(filters, "sub_deemphasis", None)
    if sd is not None:
        sd_hp = np.ascontiguousarray(sd.high_pass, dtype=np.complex128)
        sd_sos = np.ascontiguousarray(sd.amplitude_sos, dtype=np.float64)
        sd_par = sd.params()
        sd_args = (_p(sd_hp.view(np.float64), _dp), _p(sd_sos, _dp), _p(sd_par, _dp))
    else:
        sd_args = (None, None, None)
    bo = getattr(filters, "rf_high_boost", None)
    boost = None if bo is None else np.ascontiguousarray(np.concatenate([[bo.multiplier], np.ravel(bo.rf_top_sos)]),
                                                         dtype=np.float64)
    nt = getattr(filters, "fsc_notch", None)
    notch = None if nt is None else np.ascontiguousarray(np.concatenate([nt[0], nt[1]]), dtype=np.float64)
    sd_args = sd_args + (_p(boost, _dp) if boost is not None else None, _p(notch, _dp) if notch is not None else None)
    weak = lib().oracle_demod_block(
        _p(xi, _dp), n, fs, _p(rv.view(np.float64), _dp), _p(vf.view(np.float64), _dp),
        _p(lf.view(np.float64), _dp), int(filters.video_lowpass05_offset), _p(env, _dp), env.shape[0],
        _p(ch, _dp), ch.shape[0], int(filters.chroma_offset), float(dmax), _p(video, _dp), _p(vlp, _dp),
        _p(envelope, _dp), _p(chroma, _fp), _p(demod, _dp), *sd_args)
    return DemodResult(video, vlp, envelope, chroma if chs is not None else None, demod, bool(weak))


def find_pulses(x: np.ndarray, high: float, min_len: int = 0, max_len: int = 5000):
    xi = np.ascontiguousarray(x, dtype=np.float64)
    cap = max(16, xi.size // 2 + 1)
    starts = np.empty(cap, dtype=np.int32)
    lens = np.empty(cap, dtype=np.int32)
    count = lib().oracle_find_pulses(_p(xi, _dp), xi.size, high, min_len, max_len, _p(starts, _ip),
                                     _p(lens, _ip), cap)
    return starts[:count].copy(), lens[:count].copy()


def zero_crossing(data, start, target, edge=0, count=16, reverse=False):
    d = np.ascontiguousarray(data, dtype=np.float64)
    out = ctypes.c_double()
    ok = lib().oracle_zero_crossing(_p(d, _dp), d.size, start, target, edge, count, 1 if reverse else 0,
                                    ctypes.byref(out))
    return out.value if ok else None


def refine_hsync(sync_ref, locations, filled, threshold, fs_mhz, hsync_us, ire0, hz_ire, is_pal, use_right=True):
    s = np.ascontiguousarray(sync_ref, dtype=np.float64)
    loc = np.ascontiguousarray(locations, dtype=np.float64)
    fil = np.ascontiguousarray(filled, dtype=np.uint8)
    ro = np.empty_like(loc)
    fo = np.empty_like(fil)
    lib().oracle_refine_hsync(_p(s, _dp), s.size, _p(loc, _dp), _p(fil, _u8p), loc.size, threshold, fs_mhz,
                              hsync_us, ire0, hz_ire, 1 if is_pal else 0, 1 if use_right else 0,
                              _p(ro, _dp), _p(fo, _u8p))
    return ro, fo.astype(bool)


def tbc_resample(src, locations, first_line, line_count, out_len, lut, nominal=None, smoothing=0.0, conv=None):
    """conv = (ire0, hz_ire, output_zero, vsync_ire, output_scale) or None."""
    s = np.ascontiguousarray(src, dtype=np.float64)
    loc = np.ascontiguousarray(locations, dtype=np.float64)
    lt = np.ascontiguousarray(lut, dtype=np.float32).reshape(-1)
    out = np.empty(out_len * line_count)
    u16 = np.empty(out_len * line_count, dtype=np.uint16)
    c = conv if conv is not None else (0.0, 1.0, 0, 0.0, 1.0)
    rc = lib().oracle_tbc_resample(_p(s, _dp), s.size, _p(loc, _dp), loc.size, first_line, line_count, out_len,
                                   float(nominal) if nominal else 0.0, float(smoothing), _p(lt, _fp), _p(out, _dp),
                                   1 if conv is not None else 0, float(c[0]), float(c[1]), int(c[2]), float(c[3]),
                                   float(c[4]), _p(u16, _u16p))
    if rc != 0:
        raise ValueError("invalid line range for TBC resampling")
    return (out, u16) if conv is not None else out


class LdBlock:
    def __init__(self, n: int):
        self.video = np.zeros(n)
        self.demod = np.zeros(n)
        self.video_lowpass = np.zeros(n)
        self.envelope = np.zeros(n)
        self.video_burst = np.zeros(n)
        self.video_pilot = np.zeros(n)


def _cplx(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.complex128)


def demod_block_ld(block: np.ndarray, filters) -> LdBlock:
    """LaserDisc block demod (oracle_demod_block_ld) with a LaserDiscFilterSet."""
    x = np.ascontiguousarray(block, dtype=np.float64)
    n = x.size
    out = LdBlock(n)
    rv, mtf, vid, vlp = (_cplx(a) for a in (filters.rf_video, filters.rf_mtf, filters.video, filters.video_lowpass05))
    burst, pilot = _cplx(filters.video_burst), _cplx(filters.video_pilot)
    lib().oracle_demod_block_ld(
        _p(x, _dp), n, filters.sample_rate_hz, _p(rv.view(np.float64), _dp), _p(mtf.view(np.float64), _dp),
        _p(vid.view(np.float64), _dp), _p(vlp.view(np.float64), _dp), filters.video_lowpass05_offset,
        _p(burst.view(np.float64), _dp) if burst is not None else None, filters.video_burst_offset,
        _p(pilot.view(np.float64), _dp) if pilot is not None else None, filters.clip_low_hz, filters.clip_high_hz,
        _p(out.video, _dp), _p(out.demod, _dp), _p(out.video_lowpass, _dp), _p(out.envelope, _dp),
        _p(out.video_burst, _dp), _p(out.video_pilot, _dp))
    if burst is None:
        out.video_burst = None
    if pilot is None:
        out.video_pilot = None
    return out


def replace_spikes(demod: np.ndarray, diffed: np.ndarray, max_value: float) -> np.ndarray:
    """RfDemodulator.ReplaceSpikes on copies of the arrays; returns the repaired demod."""
    d = np.ascontiguousarray(demod, dtype=np.float64).copy()
    dd = np.ascontiguousarray(diffed, dtype=np.float64)
    lib().oracle_replace_spikes(_p(d, _dp), _p(dd, _dp), d.size, float(max_value))
    return d


def conj_unwrap(analytic: np.ndarray, fs: float) -> np.ndarray:
    a = np.ascontiguousarray(analytic, dtype=np.complex128)
    out = np.zeros(a.size)
    lib().oracle_conj_unwrap(_p(a.view(np.float64), _dp), a.size, fs, _p(out, _dp))
    return out


def ld_audio_block(block: np.ndarray, filters):
    """LaserDisc analog audio of one block: (left, right) at the slice rate."""
    spec = fft(np.asarray(block, dtype=np.float64).astype(np.complex128))
    res = []
    for ch in (filters.audio_left, filters.audio_right):
        st = _cplx(ch.stage1)
        out = np.zeros(ch.bin_count)
        lib().oracle_ld_audio_channel(_p(spec.view(np.float64), _dp), spec.size, ch.low_bin, ch.bin_count,
                                      _p(st.view(np.float64), _dp), ch.slice_rate_hz, ch.low_frequency_hz,
                                      _p(out, _dp))
        res.append(out)
    return res[0], res[1]


def hifi_analytic_phase(x: np.ndarray) -> np.ndarray:
    buf = np.ascontiguousarray(x, dtype=np.float32).copy()
    lib().oracle_hifi_analytic_phase(_p(buf, _fp), buf.size)
    return buf


def hifi_demodulate_phase(phase: np.ndarray, output: np.ndarray, fs: int, carrier: int, deviation: int) -> None:
    ph = np.ascontiguousarray(phase, dtype=np.float32)
    assert output.dtype == np.float32 and output.flags["C_CONTIGUOUS"] and output.size >= ph.size
    lib().oracle_hifi_demodulate_phase(_p(ph, _fp), ph.size, _p(output, _fp), fs, carrier, deviation)


def hifi_demodulate(x: np.ndarray, output: np.ndarray, fs: int, carrier: int, deviation: int) -> None:
    """HiFiHilbertDiscriminator.Demodulate (HiFi/HiFiHilbertDiscriminator.cs:42-46)."""
    hifi_demodulate_phase(hifi_analytic_phase(x), output, fs, carrier, deviation)


def pairwise_sum_f32(values: np.ndarray) -> np.float32:
    """NumpyReduction.PairwiseSumFloat32 of float32 values."""
    v = np.ascontiguousarray(values, dtype=np.float32)
    return np.float32(lib().oracle_pairwise_sum_f32(_p(v, _fp), v.size))


def find_dropouts(envelope, start, end, threshold, hysteresis, merge=30, min_len=10, cap=1 << 16):
    """RfDropoutDetector.FindDropouts: list of (start, end) sample ranges."""
    e = np.ascontiguousarray(envelope, dtype=np.float64)
    rs = np.zeros(cap, dtype=np.int32)
    re = np.zeros(cap, dtype=np.int32)
    n = lib().oracle_find_dropouts(_p(e, _dp), int(start), int(end), float(threshold), float(hysteresis), int(merge),
                                   int(min_len), _p(rs, _ip), _p(re, _ip), cap)
    n = min(n, cap)
    return [(int(rs[k]), int(re[k])) for k in range(n)]


def map_tape_rf_to_tbc(ranges, locations, out_len, start_line, end_line, line_offset, cap=1 << 16):
    """TbcDropoutMapper.MapTapeRfToTbc: (field_line, start_x, end_x) int arrays."""
    rs = np.ascontiguousarray([r[0] for r in ranges], dtype=np.float64)
    re = np.ascontiguousarray([r[1] for r in ranges], dtype=np.float64)
    loc = np.ascontiguousarray(locations, dtype=np.float64)
    fl, sx, ex = (np.zeros(cap, dtype=np.int32) for _ in range(3))
    n = lib().oracle_map_tape_rf_to_tbc(_p(rs, _dp), _p(re, _dp), rs.size, _p(loc, _dp), loc.size, int(out_len),
                                        int(start_line), int(end_line), int(line_offset), _p(fl, _ip), _p(sx, _ip),
                                        _p(ex, _ip), cap)
    if n < 0:
        raise ValueError(f"map_tape_rf_to_tbc failed ({n})")
    n = min(n, cap)
    return fl[:n].copy(), sx[:n].copy(), ex[:n].copy()


def tape_dropouts(envelope, locations, line_count, window, fraction=0.18, absolute=-1.0, hysteresis=1.25, merge=30,
                  min_len=10, out_len=910, cap=1 << 14):
    """DetectTapeDropouts for one field (span coordinates, output line k at
    locations[k]): dict with the threshold mean, ranges and the TBC map."""
    env = np.ascontiguousarray(envelope, dtype=np.float32)
    loc = np.ascontiguousarray(locations, dtype=np.float64)
    rs, re, fl, sx, ex = (np.zeros(cap, dtype=np.int32) for _ in range(5))
    nr = ctypes.c_int(0)
    mean = np.zeros(1, dtype=np.float32)
    n = lib().oracle_tape_dropouts(_p(env, _fp), env.size, _p(loc, _dp), loc.size, int(line_count), int(window[0]),
                                   int(window[1]), float(fraction), float(absolute), float(hysteresis), int(merge),
                                   int(min_len), int(out_len), _p(rs, _ip), _p(re, _ip), ctypes.byref(nr),
                                   _p(fl, _ip), _p(sx, _ip), _p(ex, _ip), cap, _p(mean, _fp))
    if n < 0:
        raise ValueError(f"tape_dropouts failed ({n})")
    k = min(nr.value, cap)
    n = min(n, cap)
    return dict(mean=mean[0], ranges=list(zip(rs[:k].tolist(), re[:k].tolist())), field_line=fl[:n].copy(),
                start_x=sx[:n].copy(), end_x=ex[:n].copy())


def estimate_read_samples(line_period_us: float, fs_hz: float, is_pal: bool, block_length: int = 32768) -> int:
    """DecodeReadWindowPlanner.EstimateReadSampleCount."""
    return int(lib().oracle_estimate_read_samples(line_period_us, fs_hz, 1 if is_pal else 0, block_length))


def resolve_read_window(intended: int, requested: int, block_length: int = 32768, block_cut: int = 1024,
                        stride: int = 30720):
    """DecodeReadWindowPlanner.Resolve: (start, length)."""
    a, b = ctypes.c_int64(), ctypes.c_int64()
    lib().oracle_resolve_read_window(int(intended), int(requested), block_length, block_cut, stride,
                                     ctypes.byref(a), ctypes.byref(b))
    return int(a.value), int(b.value)


def field_read_window(line0: float, lead_lines: float, line_len: float, requested: int, span_len: int,
                      block_length: int = 32768, block_cut: int = 1024, stride: int = 30720):
    """A located field's read window in span coordinates: Resolve of the intended start
    lead_lines before line 0 (truncated to a sample), clipped to the span."""
    intended = int(max(0.0, line0 - lead_lines * line_len))
    start, length = resolve_read_window(intended, requested, block_length, block_cut, stride)
    return min(start, span_len), min(start + length, span_len)


# ---------------------------------------------------------------- VHS field sync sequence (sync.cpp)
class SyncConfig(ctypes.Structure):
    _fields_ = [("sample_rate_hz", ctypes.c_double), ("is_pal", ctypes.c_int32),
                ("line_period_us", ctypes.c_double), ("hsync_us", ctypes.c_double), ("eq_us", ctypes.c_double),
                ("vsync_us", ctypes.c_double), ("num_pulses", ctypes.c_int32),
                ("hsync_tolerance_us", ctypes.c_double), ("eq_tolerance_us", ctypes.c_double),
                ("ire0", ctypes.c_double), ("hz_ire", ctypes.c_double), ("vsync_ire", ctypes.c_double),
                ("fps", ctypes.c_double), ("frame_lines", ctypes.c_double), ("output_lines", ctypes.c_int32),
                ("block_length", ctypes.c_int64), ("block_cut", ctypes.c_int64), ("block_stride", ctypes.c_int64),
                ("read_samples", ctypes.c_int64), ("level_divisor", ctypes.c_int32),
                ("serration_mode", ctypes.c_int32), ("tf_order", ctypes.c_int32 * 4),
                ("tf_b", (ctypes.c_double * 21) * 4), ("tf_a", (ctypes.c_double * 21) * 4),
                ("sos_sections", ctypes.c_int32 * 4), ("sos", ((ctypes.c_double * 6) * 10) * 4)]


class SyncField(ctypes.Structure):
    _fields_ = [("begin", ctypes.c_int64), ("window_start", ctypes.c_int64), ("window_length", ctypes.c_int64),
                ("status", ctypes.c_int32), ("level_source", ctypes.c_int32), ("sync_level", ctypes.c_double),
                ("blank_level", ctypes.c_double), ("dc_offset", ctypes.c_double), ("threshold", ctypes.c_double),
                ("serration_candidates", ctypes.c_int32), ("serration_measurements", ctypes.c_int32),
                ("raw_pulses", ctypes.c_int32), ("refined_pulses", ctypes.c_int32),
                ("hsync_median", ctypes.c_double), ("mean_line_length", ctypes.c_double),
                ("is_first_field", ctypes.c_int32), ("parity_confidence", ctypes.c_int32),
                ("line0", ctypes.c_double), ("first_hsync", ctypes.c_double), ("first_hsync_line", ctypes.c_double),
                ("processed_lines", ctypes.c_int32), ("next_offset", ctypes.c_double), ("next_begin", ctypes.c_int64)]


def _sync_lib():
    L = lib()
    if not getattr(L, "_sync_bound", False):
        L.oracle_sync_config_size.restype = ctypes.c_int
        L.oracle_sync_field_size.restype = ctypes.c_int
        assert L.oracle_sync_config_size() == ctypes.sizeof(SyncConfig)
        assert L.oracle_sync_field_size() == ctypes.sizeof(SyncField)
        L.oracle_vhs_field_sequence.argtypes = [_dp, ctypes.c_int64, ctypes.POINTER(SyncConfig), ctypes.c_int64,
                                                ctypes.c_int, ctypes.c_int, ctypes.POINTER(SyncField), _dp, _dp,
                                                ctypes.c_int64]
        L.oracle_vhs_field_sequence.restype = ctypes.c_int
        L.oracle_serration_levels.argtypes = [_dp, ctypes.c_int64, _dp, _dp]
        L.oracle_serration_levels.restype = ctypes.c_int
        for name in ("oracle_mean_f64", "oracle_mean_f64_fastmath", "oracle_median_f64"):
            getattr(L, name).argtypes = [_dp, ctypes.c_int64]
            getattr(L, name).restype = ctypes.c_double
        L.oracle_filtfilt_tf.argtypes = [_dp, _dp, ctypes.c_int, _dp, ctypes.c_int64]
        L.oracle_filtfilt_sos.argtypes = [_dp, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int64]
        L.oracle_refine_serration_levels.argtypes = [_dp, ctypes.c_int64, ctypes.POINTER(SyncConfig),
                                                     ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                                     ctypes.c_double, _dp, _dp, _ip, _ip]
        L.oracle_refine_serration_levels.restype = ctypes.c_int
        L.oracle_serration_analyze.argtypes = [_dp, ctypes.c_int64, ctypes.POINTER(SyncConfig), _ip, _ip, _ip,
                                               ctypes.c_int, _ip, _dp]
        L.oracle_serration_measure.argtypes = [_dp, ctypes.c_int64, ctypes.POINTER(SyncConfig), ctypes.c_int, _dp]
        L.oracle_serration_measure.restype = ctypes.c_int
        L.oracle_sync_timing.argtypes = [_dp, _ip, _ip, ctypes.c_int, ctypes.c_int, _dp]
        L.oracle_sync_classify.argtypes = [_dp, _ip, _ip, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp,
                                           ctypes.c_int64, ctypes.c_double, _ip, _ip, _ip, _ip, ctypes.c_int]
        L.oracle_sync_classify.restype = ctypes.c_int
        L.oracle_sync_mean_line_length.argtypes = [_dp, _ip, _ip, _ip, _ip, ctypes.c_int]
        L.oracle_sync_mean_line_length.restype = ctypes.c_double
        L.oracle_sync_upstream_lines.argtypes = [_ip, _ip, _ip, _ip, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                                 ctypes.c_double, ctypes.c_int, _dp]
        L.oracle_sync_resolve_cadence.argtypes = [_ip, _ip, _ip, _ip, ctypes.c_int, ctypes.c_double] + \
            [ctypes.c_int] * 9 + [_ip]
        L.oracle_sync_resolve_cadence.restype = ctypes.c_int
        L.oracle_sync_line0_transitions.argtypes = [_ip, _ip, _ip, _ip, ctypes.c_int, ctypes.c_double] + \
            [ctypes.c_int] * 5 + [_dp]
        L.oracle_sync_next_field.argtypes = [ctypes.c_int] * 4 + [ctypes.c_double] * 3 + [ctypes.c_int]
        L.oracle_sync_next_field.restype = ctypes.c_double
        L._sync_bound = True
    return L


def butterworth_serration_filter(fs: float, cutoff: float, transition: float, highpass: bool):
    """VsyncSerrationDetector.DesignFilter (Dsp/VsyncSerrationDetector.cs:582-603) with the
    order rule of IirFilterDesign.ButterworthLowPassOrder (Dsp/IirFilterDesign.cs:441-469):
    (order, b, a, sos).  The transfer function comes from scipy.signal.butter, whose
    zpk -> bilinear -> polynomial steps the reference ports."""
    import math

    import scipy.signal as ss
    nyq = fs / 2.0
    wp = cutoff / nyq
    ws = (cutoff + transition) / nyq
    warped_pass = math.tan(math.pi * wp / 2.0)
    warped_stop = math.tan(math.pi * ws / 2.0)
    ratio = warped_stop / warped_pass
    gpass = math.pow(10.0, 0.1 * 3.0) - 1.0
    gstop = math.pow(10.0, 0.1 * 30.0) - 1.0
    order = int(math.ceil(math.log10(gstop / gpass) / (2.0 * math.log10(ratio))))
    natural = math.pow(gpass, -1.0 / (2.0 * order)) * warped_pass
    wn = math.atan(natural) * 2.0 / math.pi
    order = min(order, 20)
    btype = "highpass" if highpass else "lowpass"
    b, a = ss.butter(order, wn, btype=btype, output="ba")
    sos = ss.butter(order, wn, btype=btype, output="sos")
    return order, np.asarray(b, dtype=np.float64), np.asarray(a, dtype=np.float64), np.asarray(sos, dtype=np.float64)


SERRATION_REFERENCE, SERRATION_SOS, SERRATION_NONE = 0, 1, 2


def sync_config(fmt, sample_rate_hz: float = 40e6, block_length: int = 32768, block_cut: int = 1024,
                stride: int = 30720, serration_mode: int = SERRATION_REFERENCE, level_divisor: int = 1) -> SyncConfig:
    """The vhs command's sync configuration of a tape format (SyncAnalyzer.FromParameters with the
    vhs tolerances 0.7/0.9 us, DecodeSession.cs:420-428; VsyncSerrationDetector of
    TbcFieldDecodePipeline.BuildVsyncSerrationDetector, :788-806; VideoOutputConverter levels)."""
    sp = fmt.sysparams
    is_pal = fmt.parent_system == "PAL"
    c = SyncConfig()
    c.sample_rate_hz = sample_rate_hz
    c.is_pal = 1 if is_pal else 0
    c.line_period_us = float(sp["line_period"])
    c.hsync_us, c.eq_us, c.vsync_us = float(sp["hsyncPulseUS"]), float(sp["eqPulseUS"]), float(sp["vsyncPulseUS"])
    c.num_pulses = int(sp["numPulses"])
    c.hsync_tolerance_us, c.eq_tolerance_us = 0.7, 0.9
    c.ire0 = float(fmt.rfparams["ire0"]) if "ire0" in fmt.rfparams else float(sp["ire0"])
    c.hz_ire, c.vsync_ire = float(sp["hz_ire"]), float(sp["vsync_ire"])
    c.fps, c.frame_lines = float(sp["FPS"]), float(sp["frame_lines"])
    c.output_lines = 313 if is_pal else 263
    c.block_length, c.block_cut, c.block_stride = block_length, block_cut, stride
    c.read_samples = estimate_read_samples(c.line_period_us, sample_rate_hz, is_pal, block_length)
    c.level_divisor = level_divisor
    c.serration_mode = serration_mode
    fsr = sample_rate_hz / level_divisor
    vf, hf = c.fps * 2.0, c.fps * c.frame_lines
    specs = [(vf * 5.0, 1000.0, False), (hf, hf, True), (hf, hf, False), (hf / 3.0, hf / 2.0, False)]
    for k, (cut, tw, hp) in enumerate(specs):
        order, b, a, sos = butterworth_serration_filter(fsr, cut, tw, hp)
        c.tf_order[k] = order
        for i in range(order + 1):
            c.tf_b[k][i] = b[i]
            c.tf_a[k][i] = a[i]
        c.sos_sections[k] = sos.shape[0]
        for s_ in range(sos.shape[0]):
            for j in range(6):
                c.sos[k][s_][j] = sos[s_, j]
    return c


def vhs_field_sequence(vlp: np.ndarray, cfg: SyncConfig, first_begin: int = 0, max_fields: int = 64,
                       want_shifted: bool = False):
    """The reference's VHS field sequence over a span of FP64 sync reference (global sample 0 on the
    stride grid): per field the read window, levels, threshold, parity, line 0 and the line
    locations (window-relative).  Returns (fields, locations[nfields][max_lines], shifted or None)."""
    L = _sync_lib()
    x = np.ascontiguousarray(vlp, dtype=np.float64)
    max_lines = int(cfg.output_lines) + 10
    fields = (SyncField * max_fields)()
    locs = np.zeros((max_fields, max_lines), dtype=np.float64)
    shifted = None
    stride_s = 0
    if want_shifted:
        stride_s = int(cfg.read_samples) + 2 * int(cfg.block_stride)
        shifted = np.zeros((max_fields, stride_s), dtype=np.float64)
    n = L.oracle_vhs_field_sequence(_p(x, _dp), x.size, ctypes.byref(cfg), int(first_begin), max_fields, max_lines,
                                    fields, _p(locs, _dp), _p(shifted, _dp) if shifted is not None else None,
                                    stride_s)
    return list(fields[:n]), locs[:n], (shifted[:n] if shifted is not None else None)


def serration_levels(serration: np.ndarray):
    L = _sync_lib()
    x = np.ascontiguousarray(serration, dtype=np.float64)
    s, b = ctypes.c_double(), ctypes.c_double()
    ok = L.oracle_serration_levels(_p(x, _dp), x.size, ctypes.byref(s), ctypes.byref(b))
    return (s.value, b.value) if ok else None


def refine_serration_levels(data: np.ndarray, cfg: SyncConfig, initial_sync: float, initial_blank: float,
                            reference_sync: float, hz_ire: float):
    L = _sync_lib()
    x = np.ascontiguousarray(data, dtype=np.float64)
    s, b = ctypes.c_double(), ctypes.c_double()
    vc, pc = ctypes.c_int(), ctypes.c_int()
    rc = L.oracle_refine_serration_levels(_p(x, _dp), x.size, ctypes.byref(cfg), initial_sync, initial_blank,
                                          reference_sync, hz_ire, ctypes.byref(s), ctypes.byref(b), ctypes.byref(vc),
                                          ctypes.byref(pc))
    return (rc, s.value, b.value, vc.value, pc.value)


# ---- SyncAnalyzer / FieldParityDetector / VBlankSyncResolver pieces (sync.cpp)
HSYNC, EQUALIZING, VSYNC, EQUALIZING_SECOND = 0, 1, 2, 3


def analyzer(sample_rate_hz, line_period_us, hsync_us, eq_us, vsync_us, num_pulses=6, hsync_tol=0.5, eq_tol=0.5):
    """The arguments of the SyncAnalyzer constructor (Dsp/SyncAnalysis.cs:35-66)."""
    return np.array([sample_rate_hz, line_period_us, hsync_us, eq_us, vsync_us, num_pulses, hsync_tol, eq_tol],
                    dtype=np.float64)


def _pulse_arrays(pulses):
    """pulses: [(kind, start, length, in_order)] -> four int32 arrays."""
    a = np.array([[p[0], p[1], p[2], 1 if p[3] else 0] for p in pulses], dtype=np.int32).reshape(-1, 4)
    return [np.ascontiguousarray(a[:, k]) for k in range(4)]


def sync_timing(an, pulses, preserve_units=False):
    """EstimateTiming over raw (start, length) pulses: dict of the SyncTiming fields."""
    L = _sync_lib()
    st = np.ascontiguousarray([p[0] for p in pulses], dtype=np.int32)
    ln = np.ascontiguousarray([p[1] for p in pulses], dtype=np.int32)
    out = np.zeros(8)
    L.oracle_sync_timing(_p(an, _dp), _p(st, _ip), _p(ln, _ip), st.size, 1 if preserve_units else 0, _p(out, _dp))
    keys = ("hsync_median", "hsync_lo", "hsync_hi", "eq_lo", "eq_hi", "vsync_lo", "vsync_hi", "nominal")
    return dict(zip(keys, out.tolist()))


def sync_classify(an, pulses, refine=False, preserve_units=False, reference=None, rescue_step=0.0):
    """ClassifyPulses / RefinePulses: [(kind, start, length, in_order)]."""
    L = _sync_lib()
    st = np.ascontiguousarray([p[0] for p in pulses], dtype=np.int32)
    ln = np.ascontiguousarray([p[1] for p in pulses], dtype=np.int32)
    cap = 4 * max(1, st.size) + 64
    k, s, l, o = (np.zeros(cap, dtype=np.int32) for _ in range(4))
    ref = None if reference is None else np.ascontiguousarray(reference, dtype=np.float64)
    n = L.oracle_sync_classify(_p(an, _dp), _p(st, _ip), _p(ln, _ip), st.size, 1 if preserve_units else 0,
                               1 if refine else 0, _p(ref, _dp) if ref is not None else None,
                               0 if ref is None else ref.size, rescue_step, _p(k, _ip), _p(s, _ip), _p(l, _ip),
                               _p(o, _ip), cap)
    return [(int(k[i]), int(s[i]), int(l[i]), bool(o[i])) for i in range(min(n, cap))]


def sync_mean_line_length(an, pulses):
    a = _pulse_arrays(pulses)
    return float(_sync_lib().oracle_sync_mean_line_length(_p(an, _dp), *[_p(x, _ip) for x in a], a[0].size))


def sync_upstream_lines(pulses, reference_pulse, reference_line, mean_line_length, lines):
    a = _pulse_arrays(pulses)
    out = np.zeros(lines)
    _sync_lib().oracle_sync_upstream_lines(*[_p(x, _ip) for x in a], a[0].size, reference_pulse, reference_line,
                                           mean_line_length, lines, _p(out, _dp))
    return out


def _tri(v):
    return -1 if v is None else (1 if v else 0)


def sync_resolve_cadence(pulses, mean_line_length, is_pal, field_lines, previous_first=None, minimum_confidence=100,
                         has_previous_hsync=False, has_fallback_line=False, fallback_first=None,
                         fallback_confidence=-1):
    a = _pulse_arrays(pulses)
    conf = ctypes.c_int()
    first = _sync_lib().oracle_sync_resolve_cadence(
        *[_p(x, _ip) for x in a], a[0].size, mean_line_length, 1 if is_pal else 0, field_lines[0], field_lines[1],
        _tri(previous_first), minimum_confidence, 1 if has_previous_hsync else 0, 1 if has_fallback_line else 0,
        _tri(fallback_first), fallback_confidence, ctypes.byref(conf))
    return bool(first), conf.value


def sync_line0_transitions(pulses, mean_line_length, is_pal, num_pulses, first_field, current_lines, first_lines):
    """EstimateLine0FromTransitions: (first, last, combined), each None or
    dict(line0, first_hsync, first_hsync_line, valid_count)."""
    a = _pulse_arrays(pulses)
    out = np.zeros(15)
    _sync_lib().oracle_sync_line0_transitions(*[_p(x, _ip) for x in a], a[0].size, mean_line_length,
                                              1 if is_pal else 0, num_pulses, 1 if first_field else 0, current_lines,
                                              first_lines, _p(out, _dp))
    res = []
    for k in range(3):
        r = out[5 * k:5 * k + 5]
        res.append(None if r[0] == 0 else dict(line0=r[1], first_hsync=r[2], first_hsync_line=r[3],
                                                valid_count=int(r[4])))
    return tuple(res)


def sync_next_field(is_pal, num_pulses, first_field, current_lines, mean_line_length, first_hsync, nominal=0.0,
                    cut=0):
    return float(_sync_lib().oracle_sync_next_field(1 if is_pal else 0, num_pulses, 1 if first_field else 0,
                                                    current_lines, mean_line_length, first_hsync, nominal, cut))


def serration_analyze(data: np.ndarray, cfg: SyncConfig, cap: int = 1 << 16):
    """VsyncSerrationDetector.Analyze on a fresh detector: dict of the minima lists,
    candidates, measurement count and first measured levels."""
    L = _sync_lib()
    x = np.ascontiguousarray(data, dtype=np.float64)
    e, h, c = (np.zeros(cap, dtype=np.int32) for _ in range(3))
    counts = np.zeros(4, dtype=np.int32)
    lv = np.zeros(2)
    L.oracle_serration_analyze(_p(x, _dp), x.size, ctypes.byref(cfg), _p(e, _ip), _p(h, _ip), _p(c, _ip), cap,
                               _p(counts, _ip), _p(lv, _dp))
    return dict(envelope_minima=e[:counts[0]].copy(), harmonic_minima=h[:counts[1]].copy(),
                candidates=c[:counts[2]].copy(), measurements=int(counts[3]), levels=tuple(lv))


def serration_measure(data: np.ndarray, cfg: SyncConfig, position: int):
    x = np.ascontiguousarray(data, dtype=np.float64)
    lv = np.zeros(2)
    ok = _sync_lib().oracle_serration_measure(_p(x, _dp), x.size, ctypes.byref(cfg), position, _p(lv, _dp))
    return tuple(lv) if ok else None


def serration_config(sample_rate_hz: float, fps: float, frame_lines: float, eq_us: float,
                     serration_mode: int = SERRATION_REFERENCE) -> SyncConfig:
    """A SyncConfig carrying only the VsyncSerrationDetector geometry and filters
    (new VsyncSerrationDetector(sampleRateHz, framesPerSecond, frameLines, equalizingPulseUs))."""
    c = SyncConfig()
    c.sample_rate_hz = sample_rate_hz
    c.fps, c.frame_lines, c.eq_us = fps, frame_lines, eq_us
    c.line_period_us = 1e6 / (fps * frame_lines)
    c.hsync_us, c.vsync_us, c.num_pulses = 4.7, 27.3, 5
    c.hsync_tolerance_us = c.eq_tolerance_us = 0.5
    c.hz_ire = 1.0
    c.output_lines = 313
    c.block_length, c.block_cut, c.block_stride, c.read_samples = 32768, 1024, 30720, 32768
    c.level_divisor = 1
    c.serration_mode = serration_mode
    vf, hf = fps * 2.0, fps * frame_lines
    specs = [(vf * 5.0, 1000.0, False), (hf, hf, True), (hf, hf, False), (hf / 3.0, hf / 2.0, False)]
    for k, (cut, tw, hp) in enumerate(specs):
        order, b, a, sos = butterworth_serration_filter(sample_rate_hz, cut, tw, hp)
        c.tf_order[k] = order
        for i in range(order + 1):
            c.tf_b[k][i] = b[i]
            c.tf_a[k][i] = a[i]
        c.sos_sections[k] = sos.shape[0]
        for s_ in range(sos.shape[0]):
            for j in range(6):
                c.sos[k][s_][j] = sos[s_, j]
    return c
