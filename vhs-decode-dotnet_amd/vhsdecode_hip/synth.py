"""Seeded synthetic FM tape RF (int16) with closed-form sync ground truth.

The composite baseband is built in IRE: sync tip at the format's vsync_ire,
blanking at 0, peak white at 100.  It is FM-modulated linearly,
f = ire0 + hz_ire * IRE, so the sync tip and white carriers are the format's
(e.g. PAL VHS 3.8 -> 4.8 MHz, NTSC VHS 3.4 -> 4.4 MHz), with the phase
accumulated in FP64, amplitude 8000 LSB and seeded Gaussian noise at the given
SNR relative to the carrier RMS.  A colour-under chroma carrier is added to
the RF so the chroma channel carries signal.

Field timing follows the 625/50 and 525/60 structure the reference's sync
analysis expects: broad (vsync) pulses, equalising pulses at half-line spacing
and normal hsync pulses of the format's hsyncPulseUS.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from .formats import FormatParameters


@dataclass
class SyntheticTape:
    rf: np.ndarray                 # int16 samples
    sample_rate_hz: float
    line_starts: np.ndarray        # FP64 sample position of every hsync leading edge (one per line)
    field_starts: np.ndarray       # index into line_starts of each field's first line
    lines_per_frame: int
    samples_per_line: float
    fmt: FormatParameters
    audio_carriers: tuple = ()     # (carrier Hz, deviation Hz, tone Hz, level) of each FM audio carrier

    def audio_frequency(self, index: int, sample_index: np.ndarray) -> np.ndarray:
        """Closed-form instantaneous frequency (Hz) of audio carrier `index` at the given samples."""
        fc, dev, tone, _ = self.audio_carriers[index]
        return fc + dev * np.sin(2.0 * np.pi * tone * np.asarray(sample_index, dtype=np.float64) / self.sample_rate_hz)


def _line_ire(kind: str, x_us: np.ndarray, line_us: float, p: dict, line_no: int, is_pal: bool) -> np.ndarray:
    """Composite level (IRE) of one line at offsets x_us (µs from the line start)."""
    sync = p["vsync_ire"]
    hs = p["hsyncPulseUS"]
    eq = p["eqPulseUS"]
    vs = p["vsyncPulseUS"]
    out = np.zeros_like(x_us)
    half = line_us / 2.0
    if kind == "broad":
        for base in (0.0, half):
            out[(x_us >= base) & (x_us < base + vs)] = sync
        return out
    if kind == "equal":
        for base in (0.0, half):
            out[(x_us >= base) & (x_us < base + eq)] = sync
        return out
    if kind == "eq_broad":
        out[(x_us >= 0.0) & (x_us < eq)] = sync
        out[(x_us >= half) & (x_us < half + vs)] = sync
        return out
    if kind == "broad_eq":
        out[(x_us >= 0.0) & (x_us < vs)] = sync
        out[(x_us >= half) & (x_us < half + eq)] = sync
        return out
    if kind == "eq_blank":
        out[(x_us >= 0.0) & (x_us < eq)] = sync
        return out
    if kind == "blank_eq":
        out[(x_us >= 0.0) & (x_us < hs)] = sync
        out[(x_us >= half) & (x_us < half + eq)] = sync
        return out
    # normal video line: hsync, burst on the back porch, active picture
    out[(x_us >= 0.0) & (x_us < hs)] = sync
    fsc = p["fsc_mhz"]
    b0, b1 = p["colorBurstUS"]
    burst = (x_us >= b0) & (x_us < b1)
    out[burst] += 20.0 * np.sin(2.0 * np.pi * fsc * x_us[burst])
    a0, a1 = p["activeVideoUS"]
    act = (x_us >= a0) & (x_us < a1)
    frac = (x_us[act] - a0) / (a1 - a0)
    if kind == "video_ramp":
        out[act] = 100.0 * frac
    else:
        # 75 % bars: luma steps white..black plus a subcarrier of the bar's chroma amplitude
        levels = np.array([75.0, 66.0, 52.0, 44.0, 30.0, 22.0, 8.0, 0.0]) if is_pal else \
            np.array([75.0, 69.0, 56.0, 48.0, 36.0, 28.0, 15.0, 7.5])
        chroma = np.array([0.0, 33.0, 47.0, 44.0, 44.0, 47.0, 33.0, 0.0])
        idx = np.minimum((frac * 8).astype(int), 7)
        out[act] = levels[idx] + 0.4 * chroma[idx] * np.sin(2.0 * np.pi * fsc * x_us[act] + 0.7 * idx)
    return out


def _frame_plan(is_pal: bool) -> List[str]:
    """Line kinds of one frame (two fields), starting at the first field's line 1."""
    kinds: List[str] = []
    if is_pal:
        # field 1: lines 1-2 broad, 3 broad/eq, 4-5 eq, 6.. video, 311-312 eq, 313 eq/broad
        kinds += ["broad", "broad", "broad_eq", "equal", "equal"]
        kinds += ["video_ramp" if (i % 40) == 0 else "video" for i in range(6, 311)]
        kinds += ["equal", "equal", "eq_broad"]
        # field 2: 314-315 broad, 316 eq, 317-318 eq, 319.. video, 623 blank/eq, 624-625 eq
        kinds += ["broad", "broad", "equal", "equal", "eq_blank"]
        kinds += ["video_ramp" if (i % 40) == 0 else "video" for i in range(319, 623)]
        kinds += ["blank_eq", "equal", "equal"]
        assert len(kinds) == 625, len(kinds)
    else:
        # NTSC field 1: lines 1-3 eq, 4-6 broad, 7-9 eq, 10.. video, 263 video/eq
        kinds += ["equal"] * 3 + ["broad"] * 3 + ["equal"] * 3
        kinds += ["video_ramp" if (i % 40) == 0 else "video" for i in range(10, 263)]
        kinds += ["blank_eq"]
        # field 2: 264-265 eq, 266 eq/broad, 267-268 broad, 269 broad/eq, 270-271 eq, 272 eq/blank
        kinds += ["equal", "equal", "eq_broad", "broad", "broad", "broad_eq", "equal", "equal", "eq_blank"]
        kinds += ["video_ramp" if (i % 40) == 0 else "video" for i in range(273, 526)]
        assert len(kinds) == 525, len(kinds)
    return kinds


def generate(fmt: FormatParameters, fields: int, sample_rate_hz: float = 40e6, snr_db: float = 30.0,
             seed: int = 1, amplitude: float = 8000.0, jitter_ns: float = 0.0, lead_in_lines: int = 20,
             lead_out_lines: int = 20, chroma_under_level: float = 0.12,
             audio_carriers: tuple = (), chroma_under: str = "carrier") -> SyntheticTape:
    """`audio_carriers`: (carrier Hz, deviation Hz, tone Hz, level relative to the video
    carrier) per FM audio carrier, e.g. the LaserDisc analog sound carriers near 2.3 and
    2.8 MHz, added with their phase accumulated in FP64.

    `chroma_under`: "carrier" adds a constant carrier at the format's colour-under
    frequency; "modulated" records the lines' chroma (burst and bar colours) under the
    luma FM as a colour-under recorder does: each video line's chroma moved to the
    colour-under carrier with the track's phase rotation (the format's chroma_rotation
    step per line, alternating tracks field by field)."""
    p = dict(fmt.sysparams)
    is_pal = fmt.parent_system == "PAL"
    line_us = float(p["line_period"])
    spl = line_us * sample_rate_hz / 1e6
    kinds = _frame_plan(is_pal)
    frame_lines = len(kinds)
    field_lines = [312.5, 312.5] if is_pal else [262.5, 262.5]
    total_lines = int(np.ceil(sum(field_lines[i % 2] for i in range(fields)))) + lead_in_lines + lead_out_lines + 2
    rng = np.random.default_rng(seed)
    # per-line start times (FP64 samples); optional 1/f-like jitter from integrated white noise
    starts = np.arange(total_lines + 1, dtype=np.float64) * spl
    if jitter_ns > 0:
        walk = np.cumsum(rng.standard_normal(total_lines + 1))
        walk -= np.linspace(walk[0], walk[-1], walk.size)
        walk *= (jitter_ns * 1e-9 * sample_rate_hz) / max(np.std(walk), 1e-12)
        starts = starts + walk
    n = int(np.floor(starts[-1]))
    ire = np.zeros(n, dtype=np.float64)
    sample_idx = np.arange(n, dtype=np.float64)
    first_kind = (frame_lines - lead_in_lines) % frame_lines
    cu = fmt.rfparams.get("color_under_carrier")
    modulated = chroma_under == "modulated" and cu and chroma_under_level > 0
    under = np.zeros(n, dtype=np.float64) if modulated else None
    rotation = [int(v) for v in (fmt.rfparams.get("chroma_rotation") or [0, 0])]
    for line in range(total_lines):
        a = int(np.ceil(starts[line]))
        b = min(int(np.ceil(starts[line + 1])), n)
        if a >= b:
            continue
        x_us = (sample_idx[a:b] - starts[line]) * 1e6 / sample_rate_hz
        frame_line = (first_kind + line) % frame_lines
        kind = kinds[frame_line]
        ire[a:b] = _line_ire(kind, x_us, line_us, p, line, is_pal)
        if modulated and kind.startswith("video"):
            # the line's chroma (amplitude in IRE, phase) on the colour-under carrier
            amp = np.zeros(b - a)
            ph = np.zeros(b - a)
            b0, b1 = p["colorBurstUS"]
            amp[(x_us >= b0) & (x_us < b1)] = 20.0
            if kind == "video":
                a0, a1 = p["activeVideoUS"]
                act = (x_us >= a0) & (x_us < a1)
                idx = np.minimum(((x_us[act] - a0) / (a1 - a0) * 8).astype(int), 7)
                bars = np.array([0.0, 33.0, 47.0, 44.0, 44.0, 47.0, 33.0, 0.0])
                amp[act] = 0.4 * bars[idx]
                ph[act] = 0.7 * idx
            track = rotation[0 if frame_line < frame_lines // 2 else 1]
            rot = (track * frame_line) % 4
            under[a:b] = (amp / 20.0) * np.sin(2.0 * np.pi * float(cu) * sample_idx[a:b] / sample_rate_hz
                                               + rot * (np.pi / 2.0) - ph)
    ire0 = float(fmt.rfparams.get("ire0", p["ire0"])) if "ire0" in fmt.rfparams else float(p["ire0"])
    hz_ire = float(p["hz_ire"])
    freq = ire0 + hz_ire * ire
    phase = 2.0 * np.pi * np.cumsum(freq) / sample_rate_hz
    rf = amplitude * np.cos(phase)
    if modulated:
        rf += chroma_under_level * amplitude * under
    elif cu and chroma_under_level > 0:
        rf += chroma_under_level * amplitude * np.sin(2.0 * np.pi * float(cu) * sample_idx / sample_rate_hz)
    for fc, dev, tone, level in audio_carriers:
        inst = fc + dev * np.sin(2.0 * np.pi * tone * sample_idx / sample_rate_hz)
        rf += level * amplitude * np.cos(2.0 * np.pi * np.cumsum(inst) / sample_rate_hz)
    sigma = (amplitude / np.sqrt(2.0)) / (10.0 ** (snr_db / 20.0))
    rf += rng.normal(0.0, sigma, n)
    rf = np.clip(np.rint(rf), -32768, 32767).astype(np.int16)
    # first line of each field (the line whose hsync opens field line 1)
    field_starts = []
    acc = float(lead_in_lines)
    for f in range(fields):
        field_starts.append(acc)
        acc += field_lines[f % 2]
    return SyntheticTape(rf, sample_rate_hz, starts, np.array(field_starts), frame_lines, spl, fmt,
                         tuple(tuple(float(v) for v in c) for c in audio_carriers))
