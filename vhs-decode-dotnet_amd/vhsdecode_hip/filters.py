"""Host-side FP64 design of the per-block filter spectra for colour-under tape formats.

This is the setup half of the decode chain: the managed caller of the native
backend builds these tables once per (format, block length, sample rate) and
hands them to the library, exactly as the reference's exact and ipp-fast
backends consume `DecodeFilterSet` (src/VHSDecode.Core/Dsp/DecodeFilterSet.cs:140-256,
`DecodeFilterSetBuilder.BuildBasic`).  The reference ports scipy.signal's
Butterworth/zpk2sos/freqz semantics (src/VHSDecode.Core/Dsp/IirFilterDesign.cs);
here scipy.signal itself provides those textbook routines and the custom
shelf/peaking/ramp responses are written from the formulas.

Spectra are sampled on the whole unit circle (N bins), like `FrequencyResponse`
with whole=true (IirFilterDesign.cs:962-1058).
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List, Optional

import numpy as np
from scipy import signal

from .formats import FormatParameters

VIDEO_LOWPASS05_OFFSET = 32  # DecodeFilterSet.DefaultVideoLowPass05Offset (DecodeFilterSet.cs:42)
VHS_ENVELOPE_CUTOFF_HZ = 700_000.0  # BuildVhsEnvelopeSos (DecodeFilterSet.cs:602-620)


@dataclass
class TapeFilterSet:
    """The subset of `DecodeFilterSet` the VHS real-FFT demod path reads
    (RfBlockDecodePipeline.cs:268-297 passes RfVideo, RfMtf, Video, VideoLowPass05,
    VideoLowPass05Offset, VhsEnvelopeSos, ChromaBurstSos, ChromaOffsetSamples)."""

    block_length: int
    sample_rate_hz: float
    rf_video: np.ndarray          # complex128[N]
    video: np.ndarray             # complex128[N]
    video_lowpass: np.ndarray     # complex128[N]
    video_lowpass05: np.ndarray   # complex128[N]
    envelope_sos: np.ndarray      # float64[sections, 6]
    chroma_sos: Optional[np.ndarray]
    chroma_offset: int
    video_lowpass05_offset: int
    diff_demod_max: Optional[float]
    plane_bias_hz: float = 0.0    # mid-level carrier; FP32 planes are stored as offsets from it
    sub_deemphasis: Optional["SubDeemphasis"] = None
    rf_high_boost: Optional["RfHighBoost"] = None
    fsc_notch: Optional[tuple] = None   # (b[3], a[3]) of the Betamax fsc notch on Video


@dataclass
class RfHighBoost:
    """RfHighBoostOptions (DecodeFilterSet.cs:90; DecodeSession.cs BuildRfHighBoostOptions)
    with its band-pass (BuildVhsRfTopSos, DecodeFilterSet.cs:622-641): the RF top band,
    scaled by 0.9 * mean(envelope) / envelope and `multiplier`, is added to the filtered
    RF before demodulation (RfDemodulator.cs:2374-2449)."""

    multiplier: float
    rf_top_sos: np.ndarray        # float64[1, 6]


@dataclass
class SubDeemphasis:
    """SubDeemphasisOptions (DecodeSession.cs:1250-1278) with its high-pass response
    (RfDemodulator.BuildNonlinearHighPassFilter, RfDemodulator.cs:3094-3142) and the
    amplitude low-pass (IirFilterDesign.ButterworthLowPass order 1).  Absent options
    carry their neutral values: scaling 1, logistic rate 0, static factor 0."""

    high_pass: np.ndarray         # complex128[N], whole circle
    amplitude_sos: np.ndarray     # float64[1, 6]
    deviation: float
    exponential_scaling: float
    scaling1: float = 1.0
    scaling2: float = 1.0
    logistic_mid: float = 0.0
    logistic_rate: float = 0.0
    static_factor: float = 0.0

    def params(self) -> np.ndarray:
        return np.array([self.deviation, self.exponential_scaling, self.scaling1, self.scaling2,
                         self.logistic_mid, self.logistic_rate, self.static_factor], dtype=np.float64)


# --------------------------------------------------------------------------- responses

def _freqz_ba(b, a, n: int) -> np.ndarray:
    """Whole-circle frequency response of b/a on n bins (IirFilterDesign.cs:962-1021)."""
    b = np.asarray(b, dtype=np.float64)
    a = np.asarray(a, dtype=np.float64)
    if a.size == 1 and b.size <= n and n >= 2 and (n & (n - 1)) == 0:
        padded = np.zeros(n)
        padded[: b.size] = b
        half = np.fft.rfft(padded) / a[0]
        out = np.empty(n, dtype=np.complex128)
        out[: half.size] = half
        out[half.size:] = np.conj(half[-2:0:-1])
        return out
    w = np.arange(n, dtype=np.float64) * (2.0 * math.pi / n)
    zm1 = np.exp(-1j * w)
    num = np.full(n, b[-1], dtype=np.complex128)
    for coef in b[-2::-1]:
        num = num * zm1 + coef
    den = np.full(n, a[-1], dtype=np.complex128)
    for coef in a[-2::-1]:
        den = den * zm1 + coef
    return num / den


def _freqz_sos(sos, n: int) -> np.ndarray:
    out = None
    for section in np.asarray(sos, dtype=np.float64):
        a0 = section[3]
        b = section[:3] / a0 if a0 != 1.0 else section[:3]
        a = section[3:] / a0 if a0 != 1.0 else section[3:]
        h = _freqz_ba(b, a, n)
        out = h if out is None else out * h
    return out


def _mirror_half(half: np.ndarray) -> np.ndarray:
    """FrequencyDomainFilter.MirrorHalfToFull (FrequencyDomainFilter.cs:46-61)."""
    n = (half.size - 1) * 2
    out = np.empty(n, dtype=np.float64)
    out[: half.size] = half
    out[half.size:] = half[-2:0:-1]
    return out


def _super_gaussian(x, frequency, order=1, center=0.0):
    """PortedMath.SuperGaussian (PortedMath.cs:474-478)."""
    scale = 2.0 * (x - center) * (math.log(2.0) / 2.0) ** (1.0 / (2.0 * order))
    return np.exp(-2.0 * (scale / frequency) ** (2.0 * order))


def _supergauss_half(corner, order, nyq, n, center=None, high=None):
    count = n // 2 + 1
    f = nyq * np.arange(count, dtype=np.float64) / (count - 1)
    if center is None:
        return _super_gaussian(f, corner, order)
    return _super_gaussian(f, high - corner, order, center)


def _peaking_constant_q(fnorm: float, gain_db: float, bandwidth_octaves: float):
    """IirFilterDesign.PeakingConstantQ (IirFilterDesign.cs:854-871) with the
    second-order bilinear map of BilinearSecondOrder (:1614-1636)."""
    q = 1.0 / (2.0 * math.sinh((math.log(2.0) / 2.0) * bandwidth_octaves))
    gain = 10.0 ** (gain_db / 20.0)
    zero_gain = gain if gain_db > 0 else 1.0
    pole_gain = 1.0 if gain_db > 0 else gain
    warped = 4.0 * math.tan(math.pi * fnorm / 2.0)
    an = [1.0, (zero_gain / q) * warped, warped * warped]
    ad = [1.0, (1.0 / (pole_gain * q)) * warped, warped * warped]
    return _bilinear2(an, ad, 2.0)


def _bilinear2(an, ad, fs):
    c = 2.0 * fs
    c2 = c * c
    n0 = an[0] * c2 + an[1] * c + an[2]
    n1 = -2.0 * an[0] * c2 + 2.0 * an[2]
    n2 = an[0] * c2 - an[1] * c + an[2]
    d0 = ad[0] * c2 + ad[1] * c + ad[2]
    d1 = -2.0 * ad[0] * c2 + 2.0 * ad[2]
    d2 = ad[0] * c2 - ad[1] * c + ad[2]
    return [n0 / d0, n1 / d0, n2 / d0], [1.0, d1 / d0, d2 / d0]


def _shelf(center_hz, gain_db, high: bool, fs, q):
    """IirFilterDesign.Shelf (IirFilterDesign.cs:873-922), RBJ cookbook shelf."""
    gain = 10.0 ** (gain_db / 40.0)
    omega = 2.0 * math.pi * (center_hz / fs)
    alpha = math.sin(omega) / (2.0 * q)
    cos = math.cos(omega)
    rg = math.sqrt(gain)
    if not high:
        b0 = gain * ((gain + 1.0) - ((gain - 1.0) * cos) + (2.0 * rg * alpha))
        b1 = 2.0 * gain * ((gain - 1.0) - ((gain + 1.0) * cos))
        b2 = gain * ((gain + 1.0) - ((gain - 1.0) * cos) - (2.0 * rg * alpha))
        a0 = (gain + 1.0) + ((gain - 1.0) * cos) + (2.0 * rg * alpha)
        a1 = -2.0 * ((gain - 1.0) + ((gain + 1.0) * cos))
        a2 = (gain + 1.0) + ((gain - 1.0) * cos) - (2.0 * rg * alpha)
    else:
        b0 = gain * ((gain + 1.0) + ((gain - 1.0) * cos) + (2.0 * rg * alpha))
        b1 = -2.0 * gain * ((gain - 1.0) + ((gain + 1.0) * cos))
        b2 = gain * ((gain + 1.0) + ((gain - 1.0) * cos) - (2.0 * rg * alpha))
        a0 = (gain + 1.0) - ((gain - 1.0) * cos) + (2.0 * rg * alpha)
        a1 = 2.0 * ((gain - 1.0) - ((gain + 1.0) * cos))
        a2 = (gain + 1.0) - ((gain - 1.0) * cos) - (2.0 * rg * alpha)
    return [b0, b1, b2], [a0, a1, a2]


def _emphasis_iir(t_zero, t_pole, fs):
    """IirFilterDesign.EmphasisIir (IirFilterDesign.cs:936-960) + BilinearFirstOrder (:1638-1663)."""
    fz = 2.0 * fs * math.tan((1.0 / t_zero) / (2.0 * fs))
    fp = 2.0 * fs * math.tan((1.0 / t_pole) / (2.0 * fs))
    gain = fp / fz
    an = [gain, gain * fz]
    ad = [1.0, fp]
    factor = math.sqrt(fs * 2.0)
    inv = 1.0 / factor
    nc = an[1] * inv + an[0] * -factor
    nl = an[1] * inv + an[0] * factor
    dc = ad[1] * inv + ad[0] * -factor
    dl = ad[1] * inv + ad[0] * factor
    return [nl / dl, nc / dl], [dl / dl, dc / dl]


def _ramp(start_hz, boost_start, boost_max, nyq, n):
    """FrequencyDomainFilter.RampFilter (FrequencyDomainFilter.cs:64-103)."""
    half = n // 2
    zero_count = int((start_hz / nyq) * half)
    zero_count = min(max(zero_count, 0), half)
    hs = np.zeros(half)
    ramp_count = half - zero_count
    if ramp_count == 1:
        hs[zero_count] = boost_start
    elif ramp_count > 1:
        end = boost_max * (nyq / 20e6)
        step = (end - boost_start) / (ramp_count - 1)
        hs[zero_count:] = boost_start + np.arange(ramp_count) * step
        hs[-1] = end
    out = np.empty(n)
    out[:half] = hs
    out[n - 1 - np.arange(half)] = hs
    return out


def _firwin_lowpass(numtaps: int, cutoff: float) -> np.ndarray:
    """DecodeFilterSetBuilder.FirWinLowPass (DecodeFilterSet.cs:1346-1375): Hamming-windowed
    sinc normalised to unit DC gain, i.e. scipy.signal.firwin(numtaps, cutoff)."""
    return signal.firwin(numtaps, cutoff, window="hamming", pass_zero=True, scale=True)


def _lowpass_first_order_sos(fnorm: float) -> np.ndarray:
    """IirFilterDesign.ButterworthLowPass(order 1) (IirFilterDesign.cs:1078-1109)."""
    k = math.tan(math.pi * fnorm / 2.0)
    norm = 1.0 / (1.0 + k)
    return np.array([[k * norm, k * norm, 0.0, 1.0, (k - 1.0) * norm, 0.0]])


def _butter_highpass_sos(order: int, fnorm: float) -> np.ndarray:
    """IirFilterDesign.ButterworthHighPass (IirFilterDesign.cs:108-111, 1078-1143):
    bilinear biquads with k = tan(pi f / 2), then a first-order section for odd orders."""
    k = math.tan(math.pi * fnorm / 2.0)
    sections = []
    for i in range(order // 2):
        q = 1.0 / (2.0 * math.sin(((2.0 * i) + 1.0) * math.pi / (2.0 * order)))
        k2 = k * k
        norm = 1.0 / (1.0 + (k / q) + k2)
        sections.append([norm, -2.0 * norm, norm, 1.0, 2.0 * (k2 - 1.0) * norm, (1.0 - (k / q) + k2) * norm])
    if order % 2:
        norm = 1.0 / (1.0 + k)
        sections.append([norm, -norm, 0.0, 1.0, (k - 1.0) * norm, 0.0])
    return np.array(sections, dtype=np.float64)


def _bandpass_sos(order: int, lo: float, hi: float) -> np.ndarray:
    """IirFilterDesign.ButterworthBandPassSos (IirFilterDesign.cs:594-668)."""
    if order == 1:
        fs = 2.0
        bil = 2.0 * fs
        wl = 2.0 * fs * math.tan(math.pi * lo / fs)
        wh = 2.0 * fs * math.tan(math.pi * hi / fs)
        bw = wh - wl
        center = math.sqrt(wl * wh)
        pr = (-1.0 * bw) / 2.0
        pi_ = math.sqrt(center * center - pr * pr)
        p1 = complex(pr, -pi_)
        p2 = complex(pr, pi_)
        d1 = (bil + p1) / (bil - p1)
        d2 = (bil + p2) / (bil - p2)
        gain = bw * (bil / ((bil - p1) * (bil - p2))).real
        return np.array([[gain, 0.0 * gain, -1.0 * gain, 1.0, -(d1.real + d2.real), (d1 * d2).real]])
    return signal.butter(order, [lo, hi], btype="bandpass", output="sos")


# --------------------------------------------------------------------------- builders

def _rf_video(fmt: FormatParameters, nyq: float, n: int) -> np.ndarray:
    """BuildRfVideo for tape (DecodeFilterSet.cs:307-406; no split skirts, magnitude only)."""
    rf = fmt.rfparams
    low = float(rf["video_bpf_low"])
    high = float(rf["video_bpf_high"])
    if rf.get("video_bpf_supergauss", False):
        half = _supergauss_half(low, int(rf.get("video_bpf_order") or 1), nyq, n, center=(low + high) / 2.0, high=high)
        out = _mirror_half(half).astype(np.complex128)
    else:
        order = rf.get("video_bpf_order")
        if order is not None and int(order) > 0:
            b, a = signal.butter(int(order), [low / nyq, high / nyq], btype="bandpass")
            out = _freqz_ba(b, a, n)
        else:
            out = np.ones(n, dtype=np.complex128)
    out = np.abs(out).astype(np.complex128)
    if rf.get("video_lpf_extra") is not None:
        sos = signal.butter(int(rf.get("video_lpf_extra_order") or 1), float(rf["video_lpf_extra"]) / nyq, output="sos")
        out = out * np.abs(_freqz_sos(sos, n))
    if rf.get("video_hpf_extra") is not None:
        sos = signal.butter(int(rf.get("video_hpf_extra_order") or 1), float(rf["video_hpf_extra"]) / nyq,
                            btype="highpass", output="sos")
        out = out * np.abs(_freqz_sos(sos, n))
    peak = rf.get("video_rf_peak_freq")
    if peak is not None and float(peak) > 0:
        gain = float(rf.get("video_rf_peak_gain", 3.0) if rf.get("video_rf_peak_gain") is not None else 3.0)
        bw = float(rf.get("video_rf_peak_bandwidth") or 2_500_000.0) / nyq
        b, a = _peaking_constant_q(float(peak) / nyq, gain, bw)
        out = out * np.abs(_freqz_ba(b, a, n))
    if rf.get("boost_rf_linear_0") is not None:
        ramp = _ramp(float(rf.get("start_rf_linear") or 0.0), float(rf["boost_rf_linear_0"]),
                     float(rf.get("boost_rf_linear_20", 1.0) if rf.get("boost_rf_linear_20") is not None else 1.0),
                     nyq, n)
        out = out * ramp
        if rf.get("boost_rf_linear_double", False):
            out = out * ramp
    return np.abs(out).astype(np.complex128)


def _video_lowpass(fmt: FormatParameters, nyq: float, n: int) -> np.ndarray:
    """BuildVideoLowPass with zeroPhaseMagnitude (DecodeFilterSet.cs:1019-1061)."""
    rf = fmt.rfparams
    corner = float(rf["video_lpf_freq"])
    order = int(rf.get("video_lpf_order") or 1)
    if rf.get("video_lpf_supergauss", False):
        return _mirror_half(_supergauss_half(corner, order, nyq, n)).astype(np.complex128)
    if order % 2 == 0:
        resp = _freqz_sos(signal.butter(order, corner / nyq, output="sos"), n)
    else:
        b, a = signal.butter(order, corner / nyq)
        resp = _freqz_ba(b, a, n)
    return np.abs(resp).astype(np.complex128)


def _deemphasis(fmt: FormatParameters, fs: float, n: int, configured_strength: bool) -> Optional[np.ndarray]:
    """ApplyVideoDeemphasis (DecodeFilterSet.cs:1104-1152)."""
    rf = fmt.rfparams
    if rf.get("deemph_gain") is not None and rf.get("deemph_mid") is not None:
        q = float(rf.get("deemph_q", 0.5) if rf.get("deemph_q") is not None else 0.5)
        b, a = _shelf(float(rf["deemph_mid"]), float(rf["deemph_gain"]), True, fs, q)
        # VideoDeEmphasisShelf swaps numerator and denominator of the pre-emphasis shelf.
        return _freqz_ba(a, b, n)
    tc = rf.get("video_deemp")
    if isinstance(tc, list) and len(tc) >= 2:
        b, a = _emphasis_iir(float(tc[0]), float(tc[1]), fs)
        resp = _freqz_ba(b, a, n)
        strength = float(rf.get("video_deemp_strength", 1.0) or 1.0) if configured_strength else 1.0
        return resp if strength == 1.0 else resp ** strength
    return None


_CUSTOM_FILTERS = None


def _custom_filter_file(name: str, fs: float, n: int) -> Optional[np.ndarray]:
    """DecodeFilterSetBuilder.TryLoadCustomFilter (DecodeFilterSet.cs:1658-1700): the
    response stored for `name` at this sample rate (the reference's
    Formats/custom/<name>-<rate>.txt, carried as data), a half spectrum mirrored to the
    whole circle; None when no file exists for the rate or its length does not fit."""
    global _CUSTOM_FILTERS
    if _CUSTOM_FILTERS is None:
        import os
        path = os.path.join(os.path.dirname(__file__), "data", "custom_luma_filters.npz")
        with np.load(path) as z:
            _CUSTOM_FILTERS = {k: z[k] for k in z.files}
    half = _CUSTOM_FILTERS.get(f"{name}-{int(fs)}")
    if half is None:
        return None
    if half.size == n:
        return half.copy()
    if half.size != n // 2 + 1:
        return None
    out = np.empty(n, dtype=np.complex128)
    out[:half.size] = half
    i = np.arange(1, half.size - 1)
    out[n - i] = np.conj(half[i])
    return out


def _apply_custom_video_filters(video: np.ndarray, fmt: FormatParameters, fs: float, n: int) -> None:
    """ApplyCustomVideoFilters (DecodeFilterSet.cs:1560-1605): file responses and
    high/low shelves (designed with half the sample rate as the rate argument, as the
    reference calls Shelf) multiply the Video filter in order."""
    for flt in fmt.rfparams.get("video_custom_luma_filters") or []:
        kind = flt.get("type")
        if kind == "file":
            resp = _custom_filter_file(flt.get("filename") or "", fs, n)
        elif kind in ("highshelf", "lowshelf"):
            b, a = _shelf(float(flt["midfreq"]), float(flt["gain"]), kind == "highshelf", fs / 2.0, float(flt["q"]))
            resp = _freqz_ba(b, a, n)
        else:
            resp = None
        if resp is not None and resp.size == video.size:
            video *= resp


def build_tape_filters(fmt: FormatParameters, sample_rate_hz: float = 40e6, block_length: int = 32768) -> TapeFilterSet:
    """`DecodeFilterSetBuilder.BuildBasic` for a colour-under tape format, reduced to the
    tables the VHS block demodulator consumes (DecodeFilterSet.cs:140-256)."""
    if block_length <= 0 or block_length % 2:
        raise ValueError("Block length must be a positive even value.")
    n = block_length
    fs = float(sample_rate_hz)
    nyq = fs / 2.0
    rf_video = _rf_video(fmt, nyq, n)
    vlp = _video_lowpass(fmt, nyq, n)
    deemph = _deemphasis(fmt, fs, n, configured_strength=True)
    video = vlp * deemph if deemph is not None else vlp.copy()
    _apply_custom_video_filters(video, fmt, fs, n)
    ref_deemph = _deemphasis(fmt, fs, n, configured_strength=False)
    reference = vlp * ref_deemph if ref_deemph is not None else vlp.copy()
    if 500_000.0 >= nyq:
        vlp05 = reference.copy()
    else:
        taps = _firwin_lowpass(65, 0.5 / (nyq / 1_000_000.0))
        vlp05 = reference * _freqz_ba(taps, [1.0], n)
    env = _lowpass_first_order_sos(VHS_ENVELOPE_CUTOFF_HZ / nyq)
    rf = fmt.rfparams
    chroma = None
    chroma_offset = 0
    if fmt.is_color_under and rf.get("color_under_carrier") is not None and rf.get("chroma_bpf_upper") is not None:
        lo = float(rf.get("chroma_bpf_lower") or 60_000.0)
        hi = float(rf["chroma_bpf_upper"])
        order = int(rf.get("chroma_bpf_order") or 4)
        chroma = _bandpass_sos(order, lo / nyq, hi / nyq)
        chroma_offset = int(float(rf.get("chroma_offset") or 5.0) * (fs / 40_000_000.0))
    sp = fmt.sysparams
    diff_max = (float(sp["ire0"]) + float(sp["hz_ire"]) * 100.0) * 2.0
    # the 50 IRE carrier, rounded to a whole Hz so it is exact in FP32
    bias = float(np.round(float(sp["ire0"]) + 50.0 * float(sp["hz_ire"])))
    boost = None
    if rf.get("boost_bpf_mult") is not None and rf.get("boost_bpf_low") is not None \
            and rf.get("boost_bpf_high") is not None and float(rf["boost_bpf_mult"]) != 0.0:
        lo, hi = float(rf["boost_bpf_low"]), float(rf["boost_bpf_high"])
        if lo <= 0.0 or hi <= lo or hi >= nyq:
            raise ValueError("VHS RF boost band-pass frequencies must be inside the Nyquist range.")
        boost = RfHighBoost(float(rf["boost_bpf_mult"]), _bandpass_sos(1, lo / nyq, hi / nyq))
    notch = None
    if fmt.tape_format in ("BETAMAX", "BETAMAX_HIFI"):
        notch = _notch(float(sp["fsc_mhz"]) * 1e6 / nyq, 2.0)
    return TapeFilterSet(n, fs, rf_video, video, vlp, vlp05, env, chroma, chroma_offset,
                         VIDEO_LOWPASS05_OFFSET, diff_max, bias, _sub_deemphasis(fmt, fs, n), boost, notch)


def _notch(fnorm: float, q: float):
    """IirFilterDesign.Notch (IirFilterDesign.cs:836-852)."""
    bandwidth = (fnorm / q) * math.pi
    omega = fnorm * math.pi
    beta = math.tan(bandwidth / 2.0)
    gain = 1.0 / (1.0 + beta)
    cos = math.cos(omega)
    return (np.array([gain, (-2.0 * cos) * gain, gain]), np.array([1.0, (-2.0 * gain) * cos, (2.0 * gain) - 1.0]))


def _sub_deemphasis(fmt: FormatParameters, fs: float, n: int) -> Optional[SubDeemphasis]:
    """BuildSubDeemphasisOptions (DecodeSession.cs:1250-1278), enabled by the format's
    use_sub_deemphasis."""
    rf, sp = fmt.rfparams, fmt.sysparams
    if not rf.get("use_sub_deemphasis"):
        return None
    nyq = fs / 2.0

    def opt(key, default=None):
        v = rf.get(key)
        return default if v is None else float(v)

    hpf = float(rf["nonlinear_highpass_freq"])
    upper = opt("nonlinear_bandpass_upper")
    order = int(rf.get("nonlinear_bandpass_order") or 1)
    if not 0.0 < hpf < nyq:
        raise ValueError("Nonlinear high-pass frequency must be between 0 Hz and Nyquist.")
    if upper is None:
        high_pass = _freqz_sos(_butter_highpass_sos(order, hpf / nyq), n)
    else:
        # IirFilterDesign.ButterworthBandPass transfer function (IirFilterDesign.cs:472-487)
        if not hpf < upper < nyq:
            raise ValueError("Nonlinear band-pass upper frequency must be above the high-pass frequency and below Nyquist.")
        b, a = signal.butter(order, [hpf / nyq, upper / nyq], btype="bandpass", output="ba")
        high_pass = _freqz_ba(b, a, n)
    amp_lpf = opt("nonlinear_amp_lpf_freq", 700_000.0)
    deviation = sp.get("nonlinear_deviation")
    deviation = float(deviation) if deviation is not None else float(sp["hz_ire"]) * (100.0 + -float(sp["vsync_ire"]))
    if deviation <= 0.0 or not 0.0 < amp_lpf < nyq:
        raise ValueError("invalid sub-deemphasis options")
    rate = opt("nonlinear_logistic_rate", 0.0)
    return SubDeemphasis(high_pass, _lowpass_first_order_sos(amp_lpf / nyq), deviation,
                         opt("nonlinear_exp_scaling", 0.25), opt("nonlinear_scaling_1", 1.0),
                         opt("nonlinear_scaling_2", 1.0), opt("nonlinear_logistic_mid", 0.0),
                         rate if rate > 0.0 else 0.0, opt("nonlinear_static_factor", 0.0))
