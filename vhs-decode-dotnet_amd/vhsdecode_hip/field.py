"""Field decode over the HIP backend: the vhs field sequence of a demodulated span
(read windows, sync levels, pulses, parity, line 0, line locations, hsync refinement
and TBC resampling to 16-bit luma fields), all on the device.

Mirrors TbcFieldDecodePipeline.DecodePrepared
(src/VHSDecode.Core/Decode/TbcFieldDecodePipeline.cs:1071-1420) field after field as
TbcFieldSequenceDecodeEngine reads them (Decode/TbcFieldSequenceDecodeEngine.cs:161-171,
804-846): PrepareSyncSpan's level detection (:2776-3234), FindRawPulses,
SyncAnalyzer.EstimateTiming / RefinePulses / ComputeMeanLineLength
(Dsp/SyncAnalysis.cs:131-538), FieldParityDetector.ResolveCadence, ResolveLine0Location
(:5452-5651), BuildUpstreamLineLocations (SyncAnalysis.cs:585-657),
RefineLineLocationsFromHSync (:1344-1348) and the TBC render
(Tbc/TbcFieldRenderer.cs:289-340, TbcLineResampler.ResamplePreparedToUInt16).  The
serration candidate search of VsyncSerrationDetector is not run (its measurements are
taken as absent); the rest is the reference's arithmetic.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from .formats import FormatParameters


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


def input_line_length(fmt: FormatParameters, sample_rate_hz: float = 40e6) -> int:
    """max(1, round half even (SyncAnalyzer.NominalLineLength)): the chroma phase analysis's
    input line length (TbcFieldDecodePipeline.AnalyzeChromaPhase, :2536-2559), with the
    line period rounded to whole samples as SyncAnalyzer.FromParameters does
    (Dsp/SyncAnalysis.cs:88-114)."""
    fs_mhz = sample_rate_hz / 1e6
    lp = float(fmt.sysparams["line_period"])
    if fs_mhz > 0.0:
        lp = round(lp * fs_mhz) / fs_mhz
    return max(1, int(round(lp * fs_mhz)))


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
                   min_len: int = 10):
    """vhsdecode_hip_dropout_params_v1 for a tape format: TbcDropoutDetectionOptions.DefaultDdd
    (fraction 0.18, hysteresis 1.25; DefaultCxAdc uses 0.35) and the field line counts of
    FieldLinesForParityDetection (TbcFieldDecodePipeline.cs:728-744)."""
    from . import _lib
    g = sync_geometry(fmt, sample_rate_hz)
    lines = g.field_lines
    first, second = (lines - 1, lines) if g.is_pal else (lines, lines - 1)
    return _lib.DropoutParamsV1(ctypes.sizeof(_lib.DropoutParamsV1), fraction, absolute, hysteresis, merge, min_len,
                                g.out_len, first, second, capacity)


# record statuses (vhsdecode_hip_vhs_field_v1.status)
DECODED, NO_PULSES, NO_FIRST_HSYNC, INSUFFICIENT_DATA, MISPREDICTED, CAPACITY, NOT_READ = range(7)


@dataclass
class SpanDecode:
    records: list                # VhsFieldV1 of every record, in sequence order
    outputs: dict                # record index -> uint16 field (decoded records only)
    next_begin: int              # where the sequence continues (span coordinates)
    tables: Optional[list]       # per call: the four line tables [records][stride] (want_tables)
    chroma: Optional[dict] = None  # record index -> uint16 chroma field (chroma decoders)


class HipVhsFieldDecoder:
    """The vhs field sequence over device-resident demodulated spans.  `decode_async`
    enqueues one call (every step on the device, no host round trip); `decode_span`
    drives calls until the sequence's next read window leaves the span, carrying the
    sequence state between calls (and, with `state_shift`, between spans)."""

    def __init__(self, demod, fmt: FormatParameters, sample_rate_hz: float = 40e6, block_length: int = 32768,
                 block_cut: int = 1024, stride: int = 30720, chroma: bool = False):
        self.demod = demod
        self.fmt = fmt
        self.geom = sync_geometry(fmt, sample_rate_hz)
        self.sync = vhs_sync_params(fmt, sample_rate_hz, block_length, block_cut, stride)
        self.started = False
        self.chroma = None
        if chroma:
            from .chroma import chroma_params
            self.chroma = chroma_params(fmt)
            demod.set_vhs_chroma(self.chroma, input_line_length(fmt, sample_rate_hz))

    def decode_async(self, span_video_ptr: int, span_vlp_ptr: int, span_len: int, out_ptr: int, max_fields: int,
                     count_ptr: int, fields_ptr: int = 0, next_begin_ptr: int = 0, first_begin: int = 0,
                     reset: bool = True, state_shift: int = 0, rounds: int = 2, stream: int = 0,
                     span_chroma_ptr: int = 0, out_chroma_ptr: int = 0, field_number_base: int = 0):
        if self.chroma is not None:
            self.demod.decode_vhs_fields_chroma_device(span_vlp_ptr, span_video_ptr, span_chroma_ptr, span_len,
                                                       first_begin, self.sync, self.geom, max_fields, out_ptr,
                                                       out_chroma_ptr, count_ptr, fields_ptr, next_begin_ptr, reset,
                                                       state_shift, rounds, field_number_base, stream)
        else:
            self.demod.decode_vhs_fields_device(span_vlp_ptr, span_video_ptr, span_len, first_begin, self.sync,
                                                self.geom, max_fields, out_ptr, count_ptr, fields_ptr,
                                                next_begin_ptr, reset, state_shift, rounds, stream)
        self.started = True

    def decode_span(self, span_video_ptr: int, span_vlp_ptr: int, span_len: int, max_fields: int = 64,
                    first_begin: int = 0, reset: bool = True, state_shift: int = 0, rounds: int = 2,
                    want_tables: bool = False, span_chroma_ptr: int = 0, field_number_base: int = 0) -> SpanDecode:
        from . import _lib, hipmem
        g = self.geom
        per = g.field_lines * g.out_len
        size = ctypes.sizeof(_lib.VhsFieldV1)
        out = hipmem.DeviceBuffer(max_fields * per * 2)
        with_chroma = self.chroma is not None
        if with_chroma and not span_chroma_ptr:
            raise ValueError("a chroma decoder needs the span's chroma plane")
        out_c = hipmem.DeviceBuffer(max_fields * per * 2) if with_chroma else None
        chroma_outputs = {}
        recs = hipmem.DeviceBuffer(max_fields * size)
        meta = hipmem.DeviceBuffer(16)
        records, outputs, tables = [], {}, ([] if want_tables else None)
        begin = first_begin
        stall = 0
        while True:
            self.decode_async(span_video_ptr, span_vlp_ptr, span_len, out.ptr, max_fields, meta.ptr, recs.ptr,
                              meta.ptr + 8, begin, reset, state_shift, rounds, span_chroma_ptr=span_chroma_ptr,
                              out_chroma_ptr=out_c.ptr if with_chroma else 0, field_number_base=field_number_base)
            hipmem.synchronize()
            reset, state_shift = False, 0
            count = int(meta.download(np.int32, 1)[0])
            next_begin = int(meta.download(np.int64, 1, offset=8)[0])
            n = min(count + 1, max_fields)
            raw = recs.download(np.uint8, n * size)
            got = [_lib.VhsFieldV1.from_buffer_copy(raw[k * size:(k + 1) * size].tobytes()) for k in range(n)]
            fields = out.download(np.uint16, max(count, 1) * per).reshape(-1, per)
            cf = out_c.download(np.uint16, max(count, 1) * per).reshape(-1, per) if with_chroma else None
            for k in range(count):
                if got[k].status == DECODED:
                    outputs[len(records) + k] = fields[k].copy()
                    if with_chroma:
                        chroma_outputs[len(records) + k] = cf[k].copy()
            if want_tables:
                tabs = []
                for which in range(4):
                    ptr, stride = self.demod.vhs_field_lines_device(which)
                    rows = hipmem.download_ptr(ptr, np.float64, max(count, 1) * stride).reshape(-1, stride)
                    tabs.append(rows[:count])
                tables.append(tabs)
            records += got[:count]
            if count >= max_fields:
                begin = next_begin
                continue
            stop = got[count]
            if stop.status == CAPACITY:
                raise RuntimeError("a read window holds more sync edges or pulses than the device lists")
            if stop.status == NOT_READ and stop.window_start + stop.window_length > span_len:
                return SpanDecode(records, outputs, next_begin, tables, chroma_outputs if with_chroma else None)
            stall = stall + 1 if count == 0 and next_begin == begin else 0
            if stall > 2:
                raise RuntimeError("the field sequence made no progress")
            begin = next_begin
