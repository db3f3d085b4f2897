"""LaserDisc analog audio stage 2 on the CPU (numpy restatement).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the device's stage 2.
The product library never imports it.

Restates LaserDiscAnalogAudioPhase2.Apply (Decode/LaserDiscAnalogAudioPhase2.cs:12-220),
which RfBlockDecodePipeline.ApplyLaserDiscAnalogAudioPhase2 (:159-163) runs on the
assembled analog audio of each stream read (RfBlockStreamDecoder.cs:913-914): blocks of
the stage-2 filter's length overlapping by 512 samples; per block both channels as
float32 offsets from the float32 carrier, peaks of the left channel above 500 kHz
(compared against the block's last value, as the reference's loop does) zeroed within
8 samples in both channels, the spectrum times the stage-2 response (fused products),
the inverse transform's real part plus the carrier.  A block of the filter's length
takes the float32 real transform (PocketFftReal32.ForwardDucc: numpy's float32 pocketfft
here), a shorter one the FP64 complex transform of the zero-padded block.  Several blocks
store float32 values; one block keeps FP64.
"""
from __future__ import annotations

import ctypes
import ctypes.util
from typing import Tuple

import numpy as np

_LIBM = None


def _libm():
    """The C library's correctly rounded fma (Math.FusedMultiplyAdd)."""
    global _LIBM
    if _LIBM is None:
        _LIBM = ctypes.CDLL(ctypes.util.find_library("m") or "libm.so.6")
        _LIBM.fma.restype = ctypes.c_double
        _LIBM.fma.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double]
    return _LIBM


OVERLAP_SKIP = 512
PEAK_THRESHOLD_HZ = 500_000.0
PEAK_RADIUS = 8


def find_peaks(src: np.ndarray, low: float = PEAK_THRESHOLD_HZ) -> list:
    """FindPeaks (:131-150): index i - 1 for every i with current > last and next >
    current, where `last` is the block's final value (fixed for the whole loop) and
    values below `low` count as 0."""
    n = src.size
    if n < 2:
        return []
    s = src.astype(np.float64)
    z = np.where(s < low, 0.0, s)
    last = z[-1]
    cur, nxt = z[:-1], z[1:]
    idx = np.nonzero((cur > last) & (nxt > cur))[0]
    return [int(i) - 1 for i in idx]


def suppress(src: np.ndarray, peaks) -> None:
    """SuppressPeaks (:153-161), in place."""
    for p in peaks:
        a = max(0, p - PEAK_RADIUS)
        b = min(p + PEAK_RADIUS, src.size)
        if b > a:
            src[a:b] = 0.0


def _prepare(values: np.ndarray, start: int, count: int, center: float) -> np.ndarray:
    """PrepareChannel (:115-129): float32 value minus the float32 carrier."""
    return (values[start:start + count].astype(np.float32) - np.float32(center)).astype(np.float32)


def _filter(raw: np.ndarray, filt: np.ndarray, center: float) -> np.ndarray:
    """FilterChannel (:163-219)."""
    n = filt.size
    if raw.size == n:
        half = np.fft.rfft(raw.astype(np.float32))
        spec = np.empty(n, np.complex128)
        h = half.astype(np.complex128)
        spec[0] = complex(h[0].real, -0.0)
        spec[h.size - 1] = complex(h[-1].real, -0.0)
        spec[1:h.size - 1] = h[1:-1]
        spec[n - 1:n - h.size + 1:-1] = np.conj(h[1:-1])
    else:
        padded = np.zeros(n)
        padded[:raw.size] = raw
        spec = np.fft.fft(padded)
    # fused products: re = fma(a.re, b.re, -(a.im b.im)), im = fma(a.re, b.im, a.im b.re)
    fma = _libm().fma
    re = np.array([fma(x, y, -(u * v)) for x, y, u, v in
                   zip(spec.real.tolist(), filt.real.tolist(), spec.imag.tolist(), filt.imag.tolist())])
    im = np.array([fma(x, y, u * v) for x, y, u, v in
                   zip(spec.real.tolist(), filt.imag.tolist(), spec.imag.tolist(), filt.real.tolist())])
    out = np.fft.ifft(re + 1j * im)
    return out.real[:raw.size] + center


def _block(left, right, stage2_left, stage2_right, center_left, center_right, start, length):
    count = min(length, left.size - start)
    l = _prepare(left, start, count, center_left)
    peaks = find_peaks(l)
    suppress(l, peaks)
    r = _prepare(right, start, count, center_right)
    suppress(r, peaks)
    return _filter(l, stage2_left, center_left), _filter(r, stage2_right, center_right)


def phase2(left, right, stage2_left, stage2_right, center_left: float, center_right: float
           ) -> Tuple[np.ndarray, np.ndarray, bool]:
    """Apply (:12-84): returns (left, right, float32_storage)."""
    left = np.asarray(left, np.float64)
    right = np.asarray(right, np.float64)
    if left.size != right.size:
        raise ValueError("LD analog audio channel lengths did not match.")
    length = stage2_left.size
    if length <= OVERLAP_SKIP or stage2_right.size != length:
        raise ValueError("LD analog audio phase-2 filters must have matching lengths greater than 512.")
    n = left.size
    if n == 0:
        return left.copy(), right.copy(), False
    out_l = np.zeros(n)
    out_r = np.zeros(n)
    fl, fr = _block(left, right, stage2_left, stage2_right, center_left, center_right, 0, length)
    if fl.size >= n:
        return fl[:n].copy(), fr[:n].copy(), False
    f32 = lambda v: v.astype(np.float32).astype(np.float64)
    out_l[:fl.size] = f32(fl)
    out_r[:fr.size] = f32(fr)
    jump = length - OVERLAP_SKIP
    pos = fl.size
    sample = jump
    while sample < n - jump:
        bl, br = _block(left, right, stage2_left, stage2_right, center_left, center_right, sample, length)
        c = bl.size - OVERLAP_SKIP
        out_l[pos:pos + c] = f32(bl[OVERLAP_SKIP:])
        out_r[pos:pos + c] = f32(br[OVERLAP_SKIP:])
        pos += c
        sample += jump
    bl, br = _block(left, right, stage2_left, stage2_right, center_left, center_right, n - length - 1, length)
    c = bl.size - OVERLAP_SKIP
    out_l[n - c:] = f32(bl[OVERLAP_SKIP:])
    out_r[n - c:] = f32(br[OVERLAP_SKIP:])
    return out_l, out_r, True
