"""Colour-under chroma parameters of a tape format (host side).

Restates the chroma field options the reference's `vhs` command builds by default
(TbcFieldDecodePipeline.BuildChromaFieldOptions, Decode/TbcFieldDecodePipeline.cs:2626-2760,
with the v0.4.0 behaviour profile): burst window, burst reference level, track rotation,
and the final chroma band-pass (DecodeFilterSetBuilder.BuildChromaFinalSosFilter,
Dsp/DecodeFilterSet.cs:710-751).  The device kernels take these as plain numbers and one
SOS table; the heterodyne and burst-carrier tables are computed by the library itself.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Tuple

import numpy as np

from .filters import _bandpass_sos
from .formats import FormatParameters

# TbcFieldDecodePipeline.VhsChromaPhaseAnalysisFirstLine (:227)
PHASE_ANALYSIS_FIRST_LINE = 1


@dataclass(frozen=True)
class ChromaParams:
    is_ntsc: bool
    fsc_mhz: float
    color_under_hz: float
    output_rate_mhz: float       # 4 fsc, the rate of the heterodyne and carrier tables
    frame_rate_hz: float         # TbcFrameSpec.OutputSampleRateHz (outfreq x 1e6)
    line_len: int                # output samples per line
    lines: int                   # output lines per field
    burst_start: int             # burst window in output samples (v0.4.0 padding)
    burst_end: int
    burst_abs_ref: float
    rotation: Tuple[int, ...]    # chroma_rotation (track phase steps per line)
    final_sos: np.ndarray        # (4, 6) final band-pass at 4 fsc
    sync_tip_len: int

    @property
    def field_samples(self) -> int:
        return self.lines * self.line_len

    @property
    def fsc_ratio(self) -> float:
        """RefineLineLocationsFromBurst's output samples per subcarrier cycle
        ((OutputSampleRateHz / 1e6) / FscMHz, TbcFieldDecodePipeline.cs:1381-1385)."""
        return (self.frame_rate_hz / 1_000_000.0) / self.fsc_mhz


def final_sos(fmt: FormatParameters) -> np.ndarray:
    """BuildChromaFinalSosFilter for a colour-under format: order-4 Butterworth band-pass
    from fsc - 0.9 x colour-under to fsc + 0.75 x colour-under at 4 fsc."""
    fsc = float(fmt.sysparams["fsc_mhz"])
    cu = float(fmt.rfparams["color_under_carrier"]) / 1e6
    nyq = (fsc * 4.0) / 2.0
    lo = fsc - (cu * 0.9)
    hi = fsc + (cu * 0.75)
    if lo <= 0.0 or hi <= lo or hi >= nyq:
        raise ValueError("final chroma band-pass outside the output Nyquist range")
    return np.ascontiguousarray(_bandpass_sos(4, lo / nyq, hi / nyq), dtype=np.float64)


def chroma_params(fmt: FormatParameters) -> ChromaParams:
    if not fmt.is_color_under:
        raise ValueError(f"{fmt.tape_format} is not a colour-under format")
    sp, rf = fmt.sysparams, fmt.rfparams
    out_per_us = float(sp["outfreq"])
    line_len = int(sp["outlinelen"])
    is_pal = fmt.parent_system == "PAL"
    lines = 313 if is_pal else 263
    cb = sp["colorBurstUS"]
    nominal_start = int(math.floor(float(cb[0]) * out_per_us))
    nominal_end = int(math.ceil(float(cb[1]) * out_per_us))
    start = min(max(nominal_start - 5, 0), line_len)
    end = min(max(nominal_end + 10, start), line_len)
    rot = tuple(int(v) for v in (rf.get("chroma_rotation") or []))
    return ChromaParams(
        is_ntsc=not is_pal, fsc_mhz=float(sp["fsc_mhz"]), color_under_hz=float(rf["color_under_carrier"]),
        output_rate_mhz=float(sp["fsc_mhz"]) * 4.0, frame_rate_hz=out_per_us * 1_000_000.0, line_len=line_len, lines=lines, burst_start=start,
        burst_end=end, burst_abs_ref=float(sp["burst_abs_ref"]), rotation=rot, final_sos=final_sos(fmt),
        sync_tip_len=int(math.floor(float(sp["hsyncPulseUS"]) * out_per_us)))


def to_struct(params: ChromaParams, input_line_length: int):
    """vhsdecode_hip_chroma_params_v1 of `params`."""
    import ctypes

    from . import _lib
    sos = np.asarray(params.final_sos, dtype=np.float64)
    if sos.shape[0] > 8 or len(params.rotation) != 2:
        raise ValueError("chroma needs at most 8 final band-pass sections and a two-entry rotation")
    cp = _lib.ChromaParamsV1()
    cp.struct_size = ctypes.sizeof(_lib.ChromaParamsV1)
    cp.is_ntsc = 1 if params.is_ntsc else 0
    cp.fsc_mhz = params.fsc_mhz
    cp.color_under_hz = params.color_under_hz
    cp.output_rate_hz = params.frame_rate_hz
    cp.field_lines = params.lines
    cp.output_line_length = params.line_len
    cp.burst_start = params.burst_start
    cp.burst_end = params.burst_end
    cp.burst_abs_ref = params.burst_abs_ref
    cp.rotation[0], cp.rotation[1] = params.rotation
    cp.sections = sos.shape[0]
    cp.input_line_length = int(input_line_length)
    for k in range(sos.shape[0]):
        for q in range(6):
            cp.final_sos[k][q] = float(sos[k, q])
    return cp
