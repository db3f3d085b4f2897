"""Host-side FP64 design of the LaserDisc block filter spectra.

`DecodeFilterSetBuilder.BuildBasic` with TapeFormat "LD"
(src/VHSDecode.Core/Dsp/DecodeFilterSet.cs:140-256): the complex RF band-pass with
split skirts (:307-406), the MTF compensation (:408-444), the video low-pass kept as a
transfer-function response with its phase (:1019-1061), the de-emphasis and
group-delay equaliser (:1063-1091, :1104-1246), the 0.5 MHz sync reference
(:1248-1263), the colour-burst and pilot references (:1265-1345) and the analog-audio
slice filters (:527-562, :913-1017).

Unlike the tape filters these responses keep their phase, and the reference forms
their products with the fused multiply-add complex product
(NumpyVectorComplexMultiply, DecodeFilterSet.cs:2009-2020) and evaluates transfer
functions by an FMA Horner scheme (IirFilterDesign.FrequencyResponse,
IirFilterDesign.cs:962-1020).  numpy has no fused multiply-add, so `_fma` forms it
from an exact product (Veltkamp split) and an exact sum.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Optional

import numpy as np
from scipy import signal

from .formats import FormatParameters

VIDEO_LOWPASS05_OFFSET = 32   # DecodeFilterSet.DefaultVideoLowPass05Offset (DecodeFilterSet.cs:42)
VIDEO_BURST_OFFSET = 40       # DecodeFilterSet.DefaultLdVideoBurstOffset (DecodeFilterSet.cs:44)
CLIP_LOW_HZ = 1_500_000.0     # RfDemodulator.ClipDemodForVideo (RfDemodulator.cs:1891-1901)
CLIP_HIGH_FRACTION = 0.75


# --------------------------------------------------------------------------- exact arithmetic

def _two_prod(a, b):
    p = a * b
    f = 134217729.0  # 2**27 + 1
    ah = a * f
    ah = ah - (ah - a)
    al = a - ah
    bh = b * f
    bh = bh - (bh - b)
    bl = b - bh
    return p, ((ah * bh - p) + ah * bl + al * bh) + al * bl


def _fma(a, b, c):
    """a * b + c with one rounding (exact product, exact sum, one final add)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    c = np.asarray(c, dtype=np.float64)
    p, e = _two_prod(a, b)
    s = p + c
    bb = s - p
    t = (p - (s - bb)) + (c - bb)
    out = s + (t + e)
    # exact zeros and non-finite values take the plain expression
    plain = p + c
    return np.where(np.isfinite(out) & (out != 0.0), out, plain)


def cmul_fma(x, y):
    """NumpyVectorComplexMultiply (DecodeFilterSet.cs:2009-2020)."""
    x = np.asarray(x, dtype=np.complex128)
    y = np.asarray(y, dtype=np.complex128)
    re = _fma(x.real, y.real, -(x.imag * y.imag))
    im = _fma(x.real, y.imag, x.imag * y.real)
    return re + 1j * im


def _cdiv(x, y):
    """numpy's complex division (IirFilterDesign.NumpyComplexDivide, IirFilterDesign.cs:692-715)."""
    xr, xi, yr, yi = x.real, x.imag, y.real, y.imag
    big_re = np.abs(yr) >= np.abs(yi)
    with np.errstate(divide="ignore", invalid="ignore"):
        r1 = yi / yr
        s1 = 1.0 / (yr + yi * r1)
        a_re = (xr + xi * r1) * s1
        a_im = (xi - xr * r1) * s1
        r2 = yr / yi
        s2 = 1.0 / (yi + yr * r2)
        b_re = (xr * r2 + xi) * s2
        b_im = (xi * r2 - xr) * s2
    return np.where(big_re, a_re, b_re) + 1j * np.where(big_re, a_im, b_im)


def freqz_tf(b, a, n: int) -> np.ndarray:
    """IirFilterDesign.FrequencyResponse(TransferFunction, n, whole) (IirFilterDesign.cs:962-1020):
    an FIR over a power-of-two circle by a real FFT, otherwise FMA Horner evaluation of
    numerator and denominator at exp(-i w) and numpy complex division."""
    b = np.asarray(b, dtype=np.float64)
    a = np.asarray(a, dtype=np.float64)
    if a.size == 1 and n >= 2 and b.size <= n and (n & (n - 1)) == 0:
        padded = np.zeros(n)
        padded[: b.size] = b
        half = np.fft.rfft(padded) / a[0]
        out = np.empty(n, dtype=np.complex128)
        out[: half.size] = half
        out[half.size:] = np.conj(half[-2:0:-1])
        return out
    step = (2.0 * math.pi) / n
    w = np.arange(n, dtype=np.float64) * step
    z = np.array([math.cos(-v) for v in w]) + 1j * np.array([math.sin(-v) for v in w])

    def horner(coef):
        res = np.full(n, coef[-1], dtype=np.complex128)
        for c in coef[-2::-1]:
            prod = cmul_fma(res, z)
            res = (c + prod.real) + 1j * prod.imag
        return res

    return _cdiv(horner(b), horner(a))


def numpy_cpow(z: np.ndarray, p: float) -> np.ndarray:
    """NumpyComplexMath.Pow (Dsp/NumpyComplexMath.cs:9-56, 84-170): the ufunc square
    root for 1/2, binary powering with plain complex products for small integers, the
    C library's cpow otherwise (numpy's own power for the same cases)."""
    z = np.asarray(z, dtype=np.complex128)
    if p == 0.0:
        return np.ones_like(z)
    if p == 0.5:
        re, im = z.real, z.imag
        mag = np.hypot(re, im)
        with np.errstate(divide="ignore", invalid="ignore"):
            pos = np.sqrt((re + mag) * 0.5)
            neg = np.sqrt((-re + mag) * 0.5)
            out_re = np.where(re >= 0.0, pos, np.abs(im) / (2.0 * neg))
            out_im = np.where(re >= 0.0, im / (2.0 * pos), np.copysign(neg, im))
        zero = (re == 0.0) & (im == 0.0)
        out_re = np.where(zero, 0.0, out_re)
        out_im = np.where(zero, im, out_im)
        return out_re + 1j * out_im
    if -100.0 < p < 100.0 and float(p).is_integer():
        e = int(p)

        def mul(x, y):
            return (x.real * y.real - x.imag * y.imag) + 1j * (x.real * y.imag + x.imag * y.real)

        if e == 1:
            out = z.copy()
        elif e == 2:
            out = mul(z, z)
        elif e == 3:
            out = mul(z, mul(z, z))
        else:
            mag = abs(e)
            result = np.ones_like(z)
            power = z.copy()
            mask = 1
            while True:
                if mag & mask:
                    result = mul(result, power)
                mask <<= 1
                if mag < mask:
                    break
                power = mul(power, power)
            out = _cdiv(np.ones_like(result), result) if e < 0 else result
        zero = (z.real == 0.0) & (z.imag == 0.0)
        if np.any(zero):
            out = np.where(zero, 0.0 if p > 0 else complex(np.nan, np.nan), out)
        return out
    zero = (z.real == 0.0) & (z.imag == 0.0)
    with np.errstate(all="ignore"):
        out = np.power(z, p)
    if np.any(zero):
        out = np.where(zero, 0.0 if p > 0 else complex(np.nan, np.nan), out)
    return out


def _butter_tf(order: int, wn, btype: str):
    b, a = signal.butter(order, wn, btype=btype, output="ba")
    return np.asarray(b, dtype=np.float64), np.asarray(a, dtype=np.float64)


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


def _hamming(n: int) -> np.ndarray:
    """HammingWindowValue (DecodeFilterSet.cs:1431-1446)."""
    if n == 1:
        return np.ones(1)
    out = np.empty(n)
    delta = math.pi - (-math.pi)
    for i in range(n):
        phase = math.pi if i == n - 1 else (i * (delta / (n - 1.0))) - math.pi
        value = 0.54 * math.cos(0.0 * phase)
        value += (1.0 - 0.54) * math.cos(phase)
        out[i] = value
    return out


def _sinc(v: float) -> float:
    if abs(v) < 1e-12:
        return 1.0
    s = math.pi * v
    return math.sin(s) / s


def _numpy_pairwise_sum(values: np.ndarray) -> float:
    """numpy's pairwise summation of a float64 array (blocks of 128, eight accumulators)."""
    n = values.size
    if n < 8:
        total = -0.0
        for v in values:
            total += float(v)
        return total
    if n <= 128:
        r = [float(values[k]) for k in range(8)]
        i = 8
        while i < n - (n % 8):
            for k in range(8):
                r[k] += float(values[i + k])
            i += 8
        res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]))
        while i < n:
            res += float(values[i])
            i += 1
        return res
    half = n // 2
    half -= half % 8
    return _numpy_pairwise_sum(values[:half]) + _numpy_pairwise_sum(values[half:])


def firwin_lowpass(numtaps: int, cutoff: float) -> np.ndarray:
    return signal.firwin(numtaps, cutoff, window="hamming", pass_zero=True, scale=True)


def firwin_bandpass(numtaps: int, lo: float, hi: float) -> np.ndarray:
    """DecodeFilterSetBuilder.FirWinBandPass (DecodeFilterSet.cs:1377-1417): windowed
    difference of sincs, scaled to unit gain at the band centre."""
    if numtaps <= 0 or lo <= 0.0 or hi >= 1.0 or lo >= hi:
        raise ValueError("FIR band-pass cutoffs must satisfy 0 < low < high < 1")
    center = (numtaps - 1) / 2.0
    win = _hamming(numtaps)
    taps = np.empty(numtaps)
    for i in range(numtaps):
        s = i - center
        taps[i] = ((hi * _sinc(hi * s)) - (lo * _sinc(lo * s))) * win[i]
    mid = 0.5 * (lo + hi)
    prods = np.array([taps[i] * math.cos(math.pi * (i - center) * mid) for i in range(numtaps)])
    scale = _numpy_pairwise_sum(prods)
    if scale != 0.0:
        taps = taps / scale
    return taps


def unwrap_angles(phase: np.ndarray) -> np.ndarray:
    """PortedMath.UnwrapAnglesInPlace (Dsp/PortedMath.cs:404-434): each sample corrected
    against the already-unwrapped previous one."""
    out = np.array(phase, dtype=np.float64)
    tau = 2.0 * math.pi
    for i in range(1, out.size):
        diff = out[i] - out[i - 1]
        dm = math.fmod(diff + math.pi, tau)
        if dm < 0:
            dm += tau
        dm -= math.pi
        if dm == -math.pi and diff > 0.0:
            dm = math.pi
        corr = dm - diff
        if abs(diff) < math.pi:
            corr = 0.0
        out[i] += corr
    return out


def _interp(x, y, v):
    """InterpolatePiecewiseLinear (DecodeFilterSet.cs:1520-1556)."""
    out = np.empty(v.size)
    for i, value in enumerate(v):
        if value <= x[0]:
            out[i] = y[0]
            continue
        if value >= x[-1]:
            out[i] = y[-1]
            continue
        high = 1
        while high < len(x) - 1 and x[high] < value:
            high += 1
        low = high - 1
        frac = (value - x[low]) / (x[high] - x[low])
        out[i] = y[low] + ((y[high] - y[low]) * frac)
    return out


# --------------------------------------------------------------------------- filter set

@dataclass
class LdAudioChannel:
    """LaserDiscAnalogAudioChannelFilter (DecodeFilterSet.cs:913-948): a power-of-two
    slice of the block spectrum around the carrier, the FIR band-pass response on that
    slice times its Hilbert multiplier, and the field-level stage-2 response."""

    low_bin: int
    bin_count: int
    slice_rate_hz: float
    low_frequency_hz: float
    center_hz: float
    stage1: np.ndarray     # complex128[bin_count]
    stage2: np.ndarray     # complex128[block_length]


@dataclass
class LaserDiscFilterSet:
    block_length: int
    sample_rate_hz: float
    rf_video: np.ndarray
    rf_mtf: np.ndarray
    video: np.ndarray
    video_lowpass: np.ndarray
    video_reference: np.ndarray
    video_lowpass05: np.ndarray
    video_burst: Optional[np.ndarray]
    video_pilot: Optional[np.ndarray]
    video_lowpass05_offset: int
    video_burst_offset: int
    clip_low_hz: float
    clip_high_hz: float
    audio_left: Optional[LdAudioChannel]
    audio_right: Optional[LdAudioChannel]
    audio_decimation: int
    plane_bias_hz: float


def _rf_video(fmt: FormatParameters, nyq: float, n: int, ntsc_audio_notch: bool) -> np.ndarray:
    """BuildRfVideo with LaserDisc split skirts (DecodeFilterSet.cs:307-406)."""
    rf, sp = fmt.rfparams, fmt.sysparams
    low = float(rf["video_bpf_low"])
    high = float(rf["video_bpf_high"])
    if rf.get("video_bpf_supergauss", False):
        raise ValueError("super-Gaussian RF band-pass is not a LaserDisc design")
    low_order = int(rf.get("video_bpf_low_order") if rf.get("video_bpf_low_order") is not None else -1)
    high_order = int(rf.get("video_bpf_high_order") if rf.get("video_bpf_high_order") is not None else -1)
    order = rf.get("video_bpf_order")
    if low_order > 0 and high_order > 0:
        hp = freqz_tf(*_butter_tf(low_order, low / nyq, "highpass"), n)
        lp = freqz_tf(*_butter_tf(high_order, high / nyq, "lowpass"), n)
        out = cmul_fma(hp, lp)
    elif order is not None and int(order) > 0:
        out = freqz_tf(*_butter_tf(int(order), [low / nyq, high / nyq], "bandpass"), n)
    else:
        out = np.ones(n, dtype=np.complex128)
    for key, btype in (("video_lpf_extra", "lowpass"), ("video_hpf_extra", "highpass")):
        if rf.get(key) is not None:
            sos = signal.butter(int(rf.get(key + "_order") or 1), float(rf[key]) / nyq, btype=btype, output="sos")
            resp = None
            for sec in sos:
                r = freqz_tf(sec[:3] / sec[3], sec[3:] / sec[3], n)
                resp = r if resp is None else cmul_fma(resp, r)
            # MultiplyIfPresent scales by the magnitude (DecodeFilterSet.cs:2047-2063)
            out = out * np.abs(resp)
    if ntsc_audio_notch and sp.get("audio_lfreq") is not None and sp.get("audio_rfreq") is not None:
        # ApplyLdNtscAnalogAudioNotchIfPresent (DecodeFilterSet.cs:1930-1986)
        notch_order = int(rf.get("audio_notchorder") or 2)
        width = float(rf.get("audio_notchwidth") or 200_000.0)
        notches = []
        for carrier in (float(sp["audio_lfreq"]), float(sp["audio_rfreq"])):
            lo, hi = carrier - width, carrier + width
            if lo <= 0.0 or hi >= nyq or lo >= hi:
                raise ValueError("Band-stop notch frequencies must be between 0 Hz and Nyquist.")
            notches.append(freqz_tf(*_butter_tf(notch_order, [lo / nyq, hi / nyq], "bandstop"), n))
        out = out * (notches[0] * notches[1])
    return out


def _mtf(fmt: FormatParameters, level: Optional[float], offset: float, nyq: float, n: int) -> np.ndarray:
    """BuildLdMtf (DecodeFilterSet.cs:408-444)."""
    if level is None:
        return np.ones(n, dtype=np.complex128)
    rf = fmt.rfparams
    eff = (level + offset) * float(rf.get("MTF_basemult", 1.0) if rf.get("MTF_basemult") is not None else 1.0)
    if eff == 0.0:
        return np.ones(n, dtype=np.complex128)
    pole_dist = float(rf["MTF_poledist"])
    pole_freq = float(rf["MTF_freq"]) * 1_000_000.0
    if pole_dist <= 0.0 or pole_dist >= 1.0 or pole_freq <= 0.0 or pole_freq >= nyq:
        raise ValueError("LD MTF parameters must describe stable poles inside the Nyquist range.")
    th_lo = math.pi * (pole_freq / nyq)
    th_hi = math.pi * (2.0 - (pole_freq / nyq))
    lo = complex(pole_dist * math.cos(th_lo), pole_dist * math.sin(th_lo))
    hi = complex(pole_dist * math.cos(th_hi), pole_dist * math.sin(th_hi))
    a1 = -(lo + hi).real
    a2 = (lo * hi).real
    resp = freqz_tf([1.0], [1.0, a1, a2], n)
    return resp if eff == 1.0 else numpy_cpow(resp, eff)


def laserdisc_mtf(fmt: FormatParameters, level: float, target_mtf: float = 1.0,
                  sample_rate_hz: float = 40e6, block_length: int = 32768) -> np.ndarray:
    """DecodeFilterSetBuilder.BuildLaserDiscMtf (DecodeFilterSet.cs:258-289)."""
    if not math.isfinite(target_mtf) or sample_rate_hz <= 0.0 or block_length <= 0 or block_length % 2:
        raise ValueError("invalid MTF request")
    return _mtf(fmt, target_mtf * level, 0.0, sample_rate_hz / 2.0, block_length)


def _video_lowpass(fmt: FormatParameters, nyq: float, n: int) -> np.ndarray:
    """BuildVideoLowPass without zero-phase magnitude (DecodeFilterSet.cs:1019-1061)."""
    rf = fmt.rfparams
    corner = float(rf["video_lpf_freq"])
    order = int(rf.get("video_lpf_order") or 1)
    if rf.get("video_lpf_supergauss", False):
        raise ValueError("super-Gaussian video low-pass is not a LaserDisc design")
    return freqz_tf(*_butter_tf(order, corner / nyq, "lowpass"), n)


def _deemphasis(fmt: FormatParameters, fs: float, n: int, configured_strength: bool):
    """ApplyVideoDeemphasis, time-constant branch (DecodeFilterSet.cs:1104-1152)."""
    rf = fmt.rfparams
    tc = rf.get("video_deemp")
    if not (isinstance(tc, list) and len(tc) >= 2):
        return None
    resp = freqz_tf(*_emphasis_iir(float(tc[0]), float(tc[1]), fs), n)
    strength = 1.0
    if configured_strength and rf.get("video_deemp_strength") is not None:
        strength = float(rf["video_deemp_strength"])
    return resp if strength == 1.0 else numpy_cpow(resp, strength)


def _group_delay_equalizer(fmt: FormatParameters, vlp: np.ndarray, fs: float, n: int) -> np.ndarray:
    """BuildLdGroupDelayEqualizer (DecodeFilterSet.cs:1154-1246)."""
    if fmt.parent_system == "PAL":
        freqs = [0.0, 0.5e6, 2.0e6, 3.0e6, 4.0e6, 4.4336e6, 4.8e6, 5.5e6]
        delays = [0.0, 0.0, 10e-9, 35e-9, 85e-9, 135e-9, 200e-9, 200e-9]
    else:
        freqs = [0.0, 0.5e6, 2.0e6, 3.0e6, 3.58e6, 4.0e6, 4.2e6, 4.8e6]
        delays = [0.0, 0.0, 15e-9, 45e-9, 80e-9, 135e-9, 200e-9, 200e-9]
    i = np.arange(n)
    bins = np.where(i <= n // 2, i * fs / n, (n - i) * fs / n)
    target = _interp(freqs, delays, bins)
    phase = unwrap_angles(np.arctan2(vlp.imag, vlp.real))
    width = fs / n
    deriv = np.empty(n)
    deriv[0] = phase[1] - phase[0]
    deriv[-1] = phase[-1] - phase[-2]
    deriv[1:-1] = (phase[2:] - phase[:-2]) / 2.0
    gd = -deriv / ((2.0 * math.pi) * width)
    half = n // 2
    half_mhz = int(np.argmin(np.abs(bins - 500_000.0)))
    lpf = float(fmt.rfparams["video_lpf_freq"])
    taper_start = lpf + 300_000.0
    taper_end = lpf + 1_300_000.0
    eq = np.empty(n, dtype=np.complex128)
    cumulative = 0.0
    for k in range(half + 1):
        residual = target[k] - (gd[k] - gd[half_mhz])
        taper = min(max((taper_end - bins[k]) / (taper_end - taper_start), 0.0), 1.0)
        residual *= taper
        if bins[k] < 400_000.0:
            residual = 0.0
        cumulative += residual
        ph = -(2.0 * math.pi) * cumulative * width
        eq[k] = complex(math.cos(ph), math.sin(ph))
    for k in range(1, half):
        eq[n - k] = np.conj(eq[k])
    eq[0] = 1.0
    eq[half] = 1.0
    return eq


def _round_to_even(v: float) -> int:
    return int(np.rint(v))


def _next_pow2(v: int) -> int:
    p = 1
    while p < v:
        p <<= 1
    return p


def determine_fft_slice(center_hz: float, min_bandwidth_hz: float, fs: float, n: int):
    """DetermineFftSlice (DecodeFilterSet.cs:972-989)."""
    width = fs / n
    center = _round_to_even(center_hz / width)
    bw = max(1, _round_to_even(min_bandwidth_hz / width))
    count = 2 * _next_pow2(bw * 2)
    low = center - count // 4
    if low < 0 or low + count // 2 >= n:
        raise ValueError("LD analog audio FFT slice is outside the RF block spectrum.")
    return low, count, width * count


def slice_spectrum(spectrum: np.ndarray, low_bin: int, bin_count: int) -> np.ndarray:
    """SliceSpectrum (DecodeFilterSet.cs:991-1017): the bins [low, low + half) and the
    mirrored run [n - low - half, n - low)."""
    n = spectrum.size
    half = bin_count // 2
    if bin_count <= 0 or bin_count % 2 or low_bin < 0 or low_bin + half > n or n - low_bin < half:
        raise ValueError("invalid FFT slice")
    out = np.empty(bin_count, dtype=np.complex128)
    out[:half] = spectrum[low_bin:low_bin + half]
    out[half:] = spectrum[n - low_bin - half:n - low_bin]
    return out


def hilbert_multiplier(n: int) -> np.ndarray:
    """PortedMath.BuildHilbertMultiplier (Dsp/PortedMath.cs:456-472)."""
    out = np.zeros(n)
    out[0] = 1.0
    out[n // 2] = 1.0
    out[1:n // 2] = 2.0
    return out


def _stage2(slice_rate: float, n: int) -> np.ndarray:
    """BuildLdAnalogAudioStage2Filter (DecodeFilterSet.cs:950-970)."""
    nyq = slice_rate / 2.0
    order, wn = signal.buttord(20_000.0 / nyq, 24_000.0 / nyq, 1.0, 9.0)
    lp = freqz_tf(*_butter_tf(int(order), float(wn), "lowpass"), n)
    de = freqz_tf(*_emphasis_iir(5.3e-6, 75e-6, slice_rate), n)
    return cmul_fma(lp, de)


def _audio_channel(center, width, order, fs, nyq, n) -> LdAudioChannel:
    low, count, rate = determine_fft_slice(center, 200_000.0, fs, n)
    taps = firwin_bandpass(order, (center - width) / nyq, (center + width) / nyq)
    full = freqz_tf(taps, [1.0], n)
    stage1 = slice_spectrum(full, low, count) * hilbert_multiplier(count)
    return LdAudioChannel(low, count, rate, fs * low / n, center, stage1, _stage2(rate, n))


def build_laserdisc_filters(fmt: FormatParameters, sample_rate_hz: float = 40e6, block_length: int = 32768,
                            mtf_level: Optional[float] = 1.0, mtf_offset: float = 0.0,
                            analog_audio: bool = True) -> LaserDiscFilterSet:
    """BuildBasic for TapeFormat "LD" with the ld command's defaults (DecodeSession.cs:1081-1124:
    MTF level 1, NTSC analog-audio notch on NTSC, analog audio on, clip demod for video)."""
    if fmt.tape_format != "LD":
        raise ValueError("not a LaserDisc format")
    n = block_length
    if n <= 0 or n % 2:
        raise ValueError("Block length must be a positive even value.")
    fs = float(sample_rate_hz)
    nyq = fs / 2.0
    rf_video = _rf_video(fmt, nyq, n, ntsc_audio_notch=fmt.parent_system == "NTSC")
    mtf = _mtf(fmt, mtf_level, mtf_offset, nyq, n)
    vlp = _video_lowpass(fmt, nyq, n)
    video = vlp.copy()
    de = _deemphasis(fmt, fs, n, configured_strength=True)
    if de is not None:
        video = cmul_fma(video, de)
    video = cmul_fma(video, _group_delay_equalizer(fmt, vlp, fs, n))
    reference = vlp.copy()
    de1 = _deemphasis(fmt, fs, n, configured_strength=False)
    if de1 is not None:
        reference = cmul_fma(reference, de1)
    if 500_000.0 >= nyq:
        vlp05 = reference.copy()
    else:
        vlp05 = cmul_fma(reference, freqz_tf(firwin_lowpass(65, 0.5 / (nyq / 1_000_000.0)), [1.0], n))
    sp = fmt.sysparams
    nyq_mhz = nyq / 1_000_000.0
    fsc = float(sp["fsc_mhz"])
    if fsc - 0.2 <= 0.0 or fsc + 0.2 >= nyq_mhz:
        raise ValueError("Band-pass frequencies must be inside the Nyquist range.")
    burst = cmul_fma(reference, freqz_tf(firwin_bandpass(81, (fsc - 0.2) / nyq_mhz, (fsc + 0.2) / nyq_mhz), [1.0], n))
    pilot = None
    if fmt.parent_system == "PAL":
        pm = float(sp["pilot_mhz"])
        pilot = cmul_fma(reference, freqz_tf(*_butter_tf(1, [(pm - 0.1) / nyq_mhz, (pm + 0.1) / nyq_mhz], "bandpass"), n))
    left = right = None
    decimation = 0
    if analog_audio:
        rf = fmt.rfparams
        width = float(rf["audio_filterwidth"])
        order = int(rf.get("audio_filterorder") or 512)
        left = _audio_channel(float(sp["audio_lfreq"]), width, order, fs, nyq, n)
        right = _audio_channel(float(sp["audio_rfreq"]), width, order, fs, nyq, n)
        decimation = n // left.bin_count
    bias = float(np.round(float(sp["ire0"]) + 50.0 * float(sp["hz_ire"])))
    return LaserDiscFilterSet(n, fs, rf_video, mtf, video, vlp, reference, vlp05, burst, pilot,
                              VIDEO_LOWPASS05_OFFSET, VIDEO_BURST_OFFSET, CLIP_LOW_HZ, CLIP_HIGH_FRACTION * fs,
                              left, right, decimation, bias)
