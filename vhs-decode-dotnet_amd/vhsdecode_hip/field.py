"""Field decode over the HIP backend: span demod -> sync pulses -> line locations ->
hsync refinement -> TBC resampling to 16-bit luma fields.

Mirrors the order of TbcFieldDecodePipeline.DecodePrepared
(src/VHSDecode.Core/Decode/TbcFieldDecodePipeline.cs:1071-1805): FindRawPulses
(:2375-2384), line locations, RefineLineLocationsFromHSync (:1344-1348) and the
TBC render (Tbc/TbcFieldRenderer.cs:289-340, TbcLineResampler.ResamplePreparedToUInt16).

Line locations: the reference derives them with SyncAnalyzer.EstimateTiming /
ClassifyPulses / RefinePulses, ResolveLine0Location and BuildUpstreamLineLocations
(Dsp/SyncAnalysis.cs:131-, TbcFieldDecodePipeline.cs:1176-1295), thousands of lines
of sequential heuristics for damaged tape.  `line_locations_from_pulses` is this
backend's own compact equivalent for well-formed sync: pulses are classified by
length, each vertical-sync group of broad pulses anchors a field, its parity comes
from the half-line phase of the broad group against the preceding hsync, and every
line start is snapped to the pulse that opens it (lines no pulse opens are
interpolated from their measured neighbours and marked `filled`).  Its agreement with the reference's sync
analysis on irregular signals is unpinned; on the synthetic tapes it is checked
against the generator's closed-form line starts.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from .formats import FormatParameters

SNAP_TOLERANCE = 0.12   # lines
LEAD_LINES = 8.0        # read-window lead before line 0


@dataclass
class FieldLines:
    line0: float                 # span position of line 0's sync leading edge (pulse start)
    locations: np.ndarray        # FP64 [lines + 1]
    filled: np.ndarray           # bool [lines + 1], True where no pulse opened the line
    first_field: bool


@dataclass
class SyncGeometry:
    fs_mhz: float
    line_len: float              # nominal input line length in samples (exact)
    nominal: float               # TbcLineResampler nominal input line length (rounded, DecodeSession.cs:408-410)
    hsync: float
    eq: float
    vsync: float
    field_lines: int             # output lines per field
    out_len: int                 # output samples per line
    threshold: float             # sync threshold (Hz): IreToHz(vsync_ire / 2)
    min_pulse: int
    max_pulse: int
    is_pal: bool
    smoothing: float             # WowLevelAdjustSmoothing (frame_lines / 2 for vhs, DecodeSession.cs:1029-1050)
    conv: tuple                  # (ire0, hz_ire, output_zero, vsync_ire, output_scale)


def sync_geometry(fmt: FormatParameters, sample_rate_hz: float = 40e6) -> SyncGeometry:
    sp = fmt.sysparams
    fs_mhz = sample_rate_hz / 1e6
    is_pal = fmt.parent_system == "PAL"
    line_len = float(sp["line_period"]) * fs_mhz
    vsync_ire = float(sp["vsync_ire"])
    ire0 = float(sp["ire0"])
    hz_ire = float(sp["hz_ire"])
    scale = ((0xD300 - 0x0100) if is_pal else (0xC800 - 0x0400)) / (100.0 - vsync_ire)
    # TbcFrameSpec (Tbc/TbcFrameSpec.cs:20-36): PAL 1135 x 313, NTSC 910 x 263 per field
    field_lines = 313 if is_pal else 263
    return SyncGeometry(
        fs_mhz=fs_mhz, line_len=line_len, nominal=float(np.round(line_len)),
        hsync=float(sp["hsyncPulseUS"]) * fs_mhz, eq=float(sp["eqPulseUS"]) * fs_mhz,
        vsync=float(sp["vsyncPulseUS"]) * fs_mhz, field_lines=field_lines, out_len=int(sp["outlinelen"]),
        threshold=ire0 + hz_ire * (vsync_ire / 2.0),
        min_pulse=max(0, int(np.ceil(float(sp["eqPulseUS"]) * fs_mhz / 8.0))),
        max_pulse=max(1, int(line_len * 5.0)), is_pal=is_pal, smoothing=float(sp["frame_lines"]) / 2.0,
        conv=(ire0, hz_ire, int(sp["outputZero"]), vsync_ire, scale))


def line_locations_from_pulses(starts: np.ndarray, lens: np.ndarray, g: SyncGeometry, span_len: int,
                               tol: float = SNAP_TOLERANCE) -> List[FieldLines]:
    starts = np.asarray(starts, dtype=np.float64)
    lens = np.asarray(lens, dtype=np.float64)
    out: List[FieldLines] = []
    if starts.size == 0:
        return out
    broad = lens > 0.5 * (g.hsync + g.vsync)
    hsync = (lens > 0.5 * (g.eq + g.hsync)) & ~broad
    hs_idx = np.flatnonzero(hsync)
    nl = g.field_lines
    n = np.arange(nl + 1, dtype=np.float64)
    for i in np.flatnonzero(broad & np.concatenate([[True], ~broad[:-1]])):
        k = np.searchsorted(hs_idx, i) - 1
        if k < 0:
            continue
        d = (starts[i] - starts[hs_idx[k]]) / g.line_len
        mid_line = abs(d - np.round(d)) > 0.25
        line0 = starts[i] - (0.5 * g.line_len if mid_line else 0.0)
        if line0 + (nl + 1) * g.line_len >= span_len:
            continue
        pred = line0 + n * g.line_len
        j = np.clip(np.searchsorted(starts, pred), 1, starts.size - 1)
        left, right = starts[j - 1], starts[j]
        best = np.where(np.abs(left - pred) <= np.abs(right - pred), left, right)
        hit = np.abs(best - pred) <= tol * g.line_len
        if not hit.any():
            continue
        # lines no pulse opened: interpolate between the measured neighbours
        locs = np.where(hit, best, np.interp(n, n[hit], best[hit] - n[hit] * g.line_len) + n * g.line_len)
        out.append(FieldLines(line0, locs, ~hit, first_field=not mid_line))
    return out


def field_origin(line0: float, g: SyncGeometry, block_length: int = 32768, block_cut: int = 1024,
                 stride: int = 30720, lead_lines: float = LEAD_LINES) -> int:
    """Start of a field's read window in span coordinates, placed the way
    DecodeReadWindowPlanner.Resolve places it (Decode/DecodeReadWindowPlanner.cs:37-63):
    the intended start (here `lead_lines` before line 0) less the block cut, aligned
    down to a block, expressed as a multiple of the block stride.  TBC source
    coordinates count from it, so their FP32 resolution matches the reference's."""
    intended = max(0.0, line0 - lead_lines * g.line_len)
    read = max(0, int(intended) - block_cut)
    aligned = (read // block_length) * block_length
    return (aligned // stride) * stride


def locator(g: SyncGeometry, block_length: int = 32768, block_cut: int = 1024, stride: int = 30720):
    """The device line locator's parameters (vhsdecode_hip_locator_v1) for geometry g;
    the same thresholds line_locations_from_pulses applies."""
    from . import _lib
    import ctypes
    return _lib.LocatorV1(ctypes.sizeof(_lib.LocatorV1), g.min_pulse, g.max_pulse, g.line_len,
                          0.5 * (g.hsync + g.vsync), 0.5 * (g.eq + g.hsync), SNAP_TOLERANCE, LEAD_LINES,
                          block_length, block_cut, stride)


def vhs_sync_params(fmt: FormatParameters, sample_rate_hz: float = 40e6, block_length: int = 32768,
                    block_cut: int = 1024, stride: int = 30720):
    """vhsdecode_hip_vhs_sync_v1 of a tape format: SyncAnalyzer.FromParameters with the vhs
    command's pulse tolerances (0.7 / 0.9 us, DecodeSession.cs:420-428) and the read-window
    geometry of DecodeReadWindowPlanner (Decode/DecodeReadWindowPlanner.cs:15-63)."""
    from . import _lib
    sp = fmt.sysparams
    return _lib.VhsSyncV1(ctypes.sizeof(_lib.VhsSyncV1), 1 if fmt.parent_system == "PAL" else 0,
                          float(sp["line_period"]), float(sp["hsyncPulseUS"]), float(sp["eqPulseUS"]),
                          float(sp["vsyncPulseUS"]), int(sp["numPulses"]), 0.7, 0.9, float(sp["FPS"]),
                          float(sp["frame_lines"]), float(sp["outfreq"]),
                          read_sample_count(fmt, sample_rate_hz, block_length),
                          block_length, block_cut, stride)


def read_sample_count(fmt: FormatParameters, sample_rate_hz: float = 40e6, block_length: int = 32768) -> int:
    """DecodeReadWindowPlanner.EstimateReadSampleCount (Decode/DecodeReadWindowPlanner.cs:15-35):
    350 lines (NTSC, aligned down to 16384) or 400 lines (PAL) of whole-sample line
    length, in whole blocks plus two."""
    line = int(np.round(float(fmt.sysparams["line_period"]) * (sample_rate_hz / 1e6)))
    if fmt.parent_system == "PAL":
        base = line * 400
    else:
        base = (line * 350 // 16384) * 16384
    return (base // block_length + 2) * block_length


def dropout_params(fmt: FormatParameters, sample_rate_hz: float = 40e6, fraction: float = 0.18,
                   absolute: float = -1.0, hysteresis: float = 1.25, capacity: int = 1024, merge: int = 30,
                   min_len: int = 10, block_length: int = 32768, block_cut: int = 1024, stride: int = 30720):
    """vhsdecode_hip_dropout_params_v1 for a tape format: TbcDropoutDetectionOptions.DefaultDdd
    (fraction 0.18, hysteresis 1.25; DefaultCxAdc uses 0.35), the field line counts of
    FieldLinesForParityDetection (TbcFieldDecodePipeline.cs:728-744) and the read window of
    DecodeReadWindowPlanner with this backend's line-0 lead."""
    from . import _lib
    g = sync_geometry(fmt, sample_rate_hz)
    lines = g.field_lines
    first, second = (lines - 1, lines) if g.is_pal else (lines, lines - 1)
    return _lib.DropoutParamsV1(ctypes.sizeof(_lib.DropoutParamsV1), fraction, absolute, hysteresis, merge, min_len,
                                g.out_len, first, second, capacity,
                                read_sample_count(fmt, sample_rate_hz, block_length), g.line_len, LEAD_LINES,
                                block_length, block_cut, stride)


class HipFieldDecoder:
    """Decodes every complete field of a device-resident demodulated span:
    pulses (device) -> line locations (host) -> batched hsync + TBC (device)."""

    def __init__(self, demod, fmt: FormatParameters, sample_rate_hz: float = 40e6):
        self.demod = demod
        self.fmt = fmt
        self.geom = sync_geometry(fmt, sample_rate_hz)

    def locate(self, span_vlp_ptr: int, span_len: int) -> List[FieldLines]:
        g = self.geom
        st, ln = self.demod.find_pulses(span_vlp_ptr, span_len, g.threshold, g.min_pulse, g.max_pulse)
        self.last_pulses = (st, ln)
        return line_locations_from_pulses(st, ln, g, span_len)

    def decode_async(self, span_video_ptr: int, span_vlp_ptr: int, span_len: int, out_ptr: int, max_fields: int,
                     count_ptr: int, line0_ptr: int = 0, first_ptr: int = 0, stream: int = 0):
        """Every step on the device with no host round trip; the field count lands in
        device int32 at count_ptr (and line 0 / first-field flags, when given)."""
        self.demod.decode_span_fields_device(span_vlp_ptr, span_video_ptr, span_len, self.geom, locator(self.geom),
                                             max_fields, out_ptr, count_ptr, line0_ptr, first_ptr, stream)

    def decode(self, span_video_ptr: int, span_vlp_ptr: int, span_len: int, out_ptr: int, max_fields: int,
               want_refined: bool = False):
        fields = self.locate(span_vlp_ptr, span_len)[:max_fields]
        if not fields:
            return fields, None
        locs = np.stack([f.locations for f in fields])
        filled = np.stack([f.filled for f in fields]).astype(np.uint8)
        origins = np.array([field_origin(f.line0, self.geom) for f in fields], dtype=np.int64)
        ref = self.demod.decode_fields_device(span_vlp_ptr, span_video_ptr, span_len, locs, filled, self.geom,
                                              out_ptr, origins=origins, want_refined=want_refined)
        return fields, ref
