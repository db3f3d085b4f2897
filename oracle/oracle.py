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
