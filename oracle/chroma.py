"""ctypes view of the chroma restatement (oracle/chroma.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/ and bench.py's parity leg as the checker of
the device chroma path.  The product library never imports it.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

from . import oracle as _o

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_u16p = ctypes.POINTER(ctypes.c_uint16)
_i = ctypes.c_int
_d = ctypes.c_double


def _lib():
    L = _o.lib()
    if not getattr(L, "_chroma_bound", False):
        L.oracle_chroma_heterodyne.argtypes = [_i, _d, _d, _d, _d, _dp]
        L.oracle_chroma_carrier.argtypes = [_i, _d, _d, _dp, _dp]
        L.oracle_numpy_trig_f32.argtypes = [ctypes.c_float, _i]
        L.oracle_numpy_trig_f32.restype = ctypes.c_float
        L.oracle_chroma_demod_burst.argtypes = [_dp, _i, _d, _i, _i, _dp, _dp, _i, _dp]
        L.oracle_chroma_probe.argtypes = [_dp, _i, _dp, _i, _i, _dp, _dp, _d, _i, _i, _i, _dp, _i, _i, _d, _dp]
        L.oracle_chroma_phase.argtypes = [_dp, _i, _dp, _i, _i, _i, _i, _i, _dp, _dp, _dp, _dp, _i, _i, _i, _ip, _i,
                                          _i, _i, _i, _i, _i, _i, _dp, _ip, _dp, _dp, _ip]
        L.oracle_chroma_phase.restype = _i
        L.oracle_chroma_refine_lines.argtypes = [_dp, _i, _i, _d, _i, _i, _dp, _dp, _i, _d, _d, _d, _i, _dp]
        L.oracle_chroma_burst_deemphasis.argtypes = [_dp, _i, _i, _i, _i, _i]
        L.oracle_chroma_upconvert.argtypes = [_dp, _i, _i, _i, _ip, _ip, _i, _dp, _i, _dp]
        L.oracle_chroma_upconvert_compensated.argtypes = [_dp, _i, _i, _i, _ip, _ip, _dp, _i, _d, _d, _d, _d, _dp]
        L.oracle_chroma_rms.argtypes = [_dp, _i, _i]
        L.oracle_chroma_rms.restype = _d
        L.oracle_chroma_acc.argtypes = [_dp, _i, _d, _i, _i, _i, _i, _i, _i, _dp]
        L.oracle_chroma_acc.restype = _d
        L.oracle_chroma_acc_comb_u16.argtypes = [_dp, _i, _d, _i, _i, _i, _i, _i, _i, _i, _i, _u16p]
        L.oracle_chroma_comb.argtypes = [_dp, _i, _i, _i, _i]
        L.oracle_chroma_ntsc_target.argtypes = [_i, _i, _ip]
        L.oracle_chroma_ntsc_target.restype = _d
        L.oracle_chroma_decode.argtypes = [_dp, _i, _i, _i, _ip, _ip, _dp, _i, _i, _i, _i, _i, _d, _d, _dp, _dp, _i,
                                           _i, _i, _d, _i, _u16p]
        L._chroma_bound = True
    return L


def _p(a, t):
    return a.ctypes.data_as(t)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def heterodyne(n: int, fsc_mhz: float, cuc_mhz: float, rate_mhz: float, drift: float = 0.0) -> np.ndarray:
    out = np.empty((4, n), np.float64)
    _lib().oracle_chroma_heterodyne(n, fsc_mhz, cuc_mhz, rate_mhz, drift, _p(out, _dp))
    return out


def carrier(n: int, carrier_mhz: float, rate_mhz: float):
    s = np.empty(n, np.float64)
    c = np.empty(n, np.float64)
    _lib().oracle_chroma_carrier(n, carrier_mhz, rate_mhz, _p(s, _dp), _p(c, _dp))
    return s, c


def numpy_trig_f32(x: float, cosine: bool) -> float:
    return float(_lib().oracle_numpy_trig_f32(x, 1 if cosine else 0))


def demod_burst(burst, line_scale, line_start, burst_start, bsin, bcos, f32=False) -> np.ndarray:
    """(phase, offset, magnitude, I, Q)"""
    b, s, c = _f64(burst), _f64(bsin), _f64(bcos)
    out = np.empty(5)
    _lib().oracle_chroma_demod_burst(_p(b, _dp), b.size, line_scale, line_start, burst_start, _p(s, _dp),
                                     _p(c, _dp), 1 if f32 else 0, _p(out, _dp))
    return out


def probe(chroma, het, burst_start, burst_end, bsin, bcos, line_scale, line_number, line_offset, line_len,
          sos: Optional[np.ndarray] = None, f32=False, filter_scale=0.0) -> np.ndarray:
    ch, h, s, c = _f64(chroma), _f64(het), _f64(bsin), _f64(bcos)
    so = None if sos is None else _f64(sos)
    out = np.empty(5)
    _lib().oracle_chroma_probe(_p(ch, _dp), ch.size, _p(h, _dp), burst_start, burst_end, _p(s, _dp), _p(c, _dp),
                               line_scale, line_number, line_offset, line_len,
                               None if so is None else _p(so, _dp), 0 if so is None else so.shape[0],
                               1 if f32 else 0, filter_scale, _p(out, _dp))
    return out


@dataclass
class PhaseResult:
    lines: np.ndarray      # line numbers
    rot: np.ndarray        # phase rotation per line
    burst: np.ndarray      # (lines, 5): phase, offset, magnitude, I, Q
    mag_avg: float
    phase_avg: float
    even_avg: float
    odd_avg: float
    next_index: int
    burst_line: int


def phase_sequence(chroma, locations, line_offset, lines_out, input_line_length, out_line_len, het4, bsin, bcos,
                   sos, burst_start, burst_end, rotation: Sequence[int], rotation_index: Optional[int] = None,
                   f32=True, color_killer=False, prev_burst_line=0, is_ntsc=True,
                   probe_table: Optional[np.ndarray] = None) -> PhaseResult:
    ch = _f64(chroma) if chroma is not None else np.zeros(1)
    loc = _f64(locations)
    h = _f64(het4) if het4 is not None else np.zeros(1)
    s = _f64(bsin) if bsin is not None else np.zeros(1)
    c = _f64(bcos) if bcos is not None else np.zeros(1)
    so = _f64(sos) if sos is not None else None
    rot = np.ascontiguousarray(rotation if rotation is not None else [], dtype=np.int32)
    pt = None if probe_table is None else _f64(probe_table)
    rot_out = np.zeros(lines_out, np.int32)
    bo = np.zeros((lines_out, 5))
    summ = np.zeros(4)
    ints = np.zeros(2, np.int32)
    n = ch.size if chroma is not None else 0
    rc = _lib().oracle_chroma_phase(
        _p(ch, _dp), n, _p(loc, _dp), loc.size, line_offset, lines_out, input_line_length, out_line_len,
        _p(h, _dp), _p(s, _dp), _p(c, _dp), None if so is None else _p(so, _dp),
        0 if so is None else so.shape[0], burst_start, burst_end, _p(rot, _ip), rot.size,
        0 if rotation_index is None else rotation_index, 0 if rotation_index is None else 1, 1 if f32 else 0,
        1 if color_killer else 0, prev_burst_line, 1 if is_ntsc else 0, None if pt is None else _p(pt, _dp),
        _p(rot_out, _ip), _p(bo, _dp), _p(summ, _dp), _p(ints, _ip))
    if rc != 0:
        raise ValueError(f"phase sequence failed ({rc})")
    return PhaseResult(np.arange(line_offset, line_offset + lines_out, dtype=np.int32), rot_out, bo,
                       float(summ[0]), float(summ[1]), float(summ[2]), float(summ[3]), int(ints[0]), int(ints[1]))


def refine_lines(locations, out_line_len, fsc_ratio, seq_first_line, phase, offset, burst_line, avg_all, avg_even,
                 avg_odd, is_ntsc) -> np.ndarray:
    loc = _f64(locations)
    ph, of = _f64(phase), _f64(offset)
    out = np.empty_like(loc)
    _lib().oracle_chroma_refine_lines(_p(loc, _dp), loc.size, out_line_len, fsc_ratio, seq_first_line, ph.size,
                                      _p(ph, _dp), _p(of, _dp), burst_line, avg_all, avg_even, avg_odd,
                                      1 if is_ntsc else 0, _p(out, _dp))
    return out


def refine_from_phase(locations, out_line_len, fsc_ratio, ph: PhaseResult, is_ntsc) -> np.ndarray:
    return refine_lines(locations, out_line_len, fsc_ratio, int(ph.lines[0]), ph.burst[:, 0], ph.burst[:, 1],
                        ph.burst_line, ph.phase_avg, ph.even_avg, ph.odd_avg, is_ntsc)


def burst_deemphasis(values, line_offset, lines_out, line_len, burst_end, after=5) -> np.ndarray:
    v = _f64(values).copy()
    _lib().oracle_chroma_burst_deemphasis(_p(v, _dp), line_offset, lines_out, line_len, burst_end, after)
    return v


def upconvert(chroma, line_offset, line_len, lines, rots, het_tables) -> np.ndarray:
    ch = _f64(chroma)
    h = _f64(np.asarray(het_tables).reshape(len(het_tables), -1))
    ln = np.ascontiguousarray(lines, np.int32)
    rt = np.ascontiguousarray(rots, np.int32)
    out = np.empty_like(ch)
    _lib().oracle_chroma_upconvert(_p(ch, _dp), ch.size, line_offset, line_len, _p(ln, _ip), _p(rt, _ip), ln.size,
                                   _p(h, _dp), h.shape[1], _p(out, _dp))
    return out


def upconvert_compensated(chroma, line_offset, line_len, lines, rots, burst_phase, cuc_hz, fsc_mhz, target_even,
                          target_odd) -> np.ndarray:
    ch = _f64(chroma)
    ln = np.ascontiguousarray(lines, np.int32)
    rt = np.ascontiguousarray(rots, np.int32)
    bp = _f64(burst_phase)
    out = np.empty_like(ch)
    _lib().oracle_chroma_upconvert_compensated(_p(ch, _dp), ch.size, line_offset, line_len, _p(ln, _ip),
                                               _p(rt, _ip), _p(bp, _dp), ln.size, cuc_hz, fsc_mhz, target_even,
                                               target_odd, _p(out, _dp))
    return out


def rms(values, f32=True) -> float:
    v = _f64(values)
    return float(_lib().oracle_chroma_rms(_p(v, _dp), v.size, 1 if f32 else 0))


def acc(chroma, burst_abs_ref, burst_start, burst_end, line_len, lines, burst_line, f32_rms=True):
    ch = _f64(chroma)
    out = np.empty_like(ch)
    m = _lib().oracle_chroma_acc(_p(ch, _dp), ch.size, burst_abs_ref, burst_start, burst_end, line_len, lines,
                                 burst_line, 1 if f32_rms else 0, _p(out, _dp))
    return out, float(m)


def acc_comb_u16(chroma, burst_abs_ref, burst_start, burst_end, line_len, lines, burst_line, distance,
                 retain_f32=True, f32_rms=True) -> np.ndarray:
    ch = _f64(chroma)
    out = np.empty(ch.size, np.uint16)
    _lib().oracle_chroma_acc_comb_u16(_p(ch, _dp), ch.size, burst_abs_ref, burst_start, burst_end, line_len, lines,
                                      burst_line, distance, 1 if retain_f32 else 0, 1 if f32_rms else 0,
                                      _p(out, _u16p))
    return out


def comb(values, line_len, distance, retain_f32=True) -> np.ndarray:
    v = _f64(values).copy()
    _lib().oracle_chroma_comb(_p(v, _dp), v.size, line_len, distance, 1 if retain_f32 else 0)
    return v


def ntsc_target(is_first_field: bool, field_number: int):
    pid = ctypes.c_int(0)
    t = _lib().oracle_chroma_ntsc_target(1 if is_first_field else 0, field_number, ctypes.byref(pid))
    return pid.value, float(t)


def decode(field, lines, line_len, line_offset, ph: PhaseResult, is_ntsc, is_first_field: Optional[bool],
           field_number, cuc_hz, fsc_mhz, het4, sos, burst_start, burst_end, burst_abs_ref) -> np.ndarray:
    """CompleteFieldWithPhaseCore (v0.4.0): uint16 chroma field."""
    f = _f64(field).copy()
    h = _f64(het4)
    so = _f64(sos)
    ln = np.ascontiguousarray(ph.lines, np.int32)
    rt = np.ascontiguousarray(ph.rot, np.int32)
    bp = _f64(ph.burst[:, 0])
    out = np.empty(lines * line_len, np.uint16)
    _lib().oracle_chroma_decode(_p(f, _dp), lines, line_len, line_offset, _p(ln, _ip), _p(rt, _ip), _p(bp, _dp),
                                ln.size, 1 if is_ntsc else 0, 0 if is_first_field is None else 1,
                                1 if is_first_field else 0, field_number, cuc_hz, fsc_mhz, _p(h, _dp), _p(so, _dp),
                                so.shape[0], burst_start, burst_end, burst_abs_ref, ph.burst_line, _p(out, _u16p))
    return out


_tables_cache = {}


def _tables(p):
    key = (p.lines * p.line_len, p.fsc_mhz, p.color_under_hz)
    if key not in _tables_cache:
        n = p.lines * p.line_len
        het = heterodyne(n, p.fsc_mhz, p.color_under_hz / 1e6, p.fsc_mhz * 4.0)
        s, c = carrier(n, p.fsc_mhz, p.fsc_mhz * 4.0)
        _tables_cache[key] = (het, s, c)
    return _tables_cache[key]


@dataclass
class FieldChroma:
    render: np.ndarray       # render line locations (burst-refined, less the NTSC fsc shift)
    phase: PhaseResult
    chroma: np.ndarray       # uint16 field


def field_chain(src, refined, p, is_first_field: bool, field_number: int, rotation_index: int, input_line_length,
                lut, nominal, smoothing, fsc_shift: float) -> FieldChroma:
    """One field's chroma on the oracle, as TbcFieldDecodePipeline.DecodePrepared runs it
    (Decode/TbcFieldDecodePipeline.cs:1366-1392, 1425-1470, 1869-1920): the analysis
    resample of the chroma window `src` from line 1 at the hsync-refined lines, the phase
    sequence from `rotation_index`, the burst-refined render locations, the render
    resample from the output first line and the chroma decode."""
    het, s, c = _tables(p)
    an = _o.tbc_resample(src, refined, 1, p.lines, p.line_len, lut, nominal=nominal, smoothing=smoothing)
    ph = phase_sequence(an, refined, 1, p.lines, input_line_length, p.line_len, het, s, c, p.final_sos,
                        p.burst_start, p.burst_end, p.rotation, rotation_index, f32=True, is_ntsc=p.is_ntsc)
    loc = refine_from_phase(refined, p.line_len, p.fsc_ratio, ph, p.is_ntsc)
    render = loc - fsc_shift if fsc_shift else loc
    first_line = 1 if p.is_ntsc else (3 if is_first_field else 4)
    fld = _o.tbc_resample(src, render, first_line, p.lines, p.line_len, lut, nominal=nominal, smoothing=smoothing)
    u16 = decode(fld, p.lines, p.line_len, first_line, ph, p.is_ntsc, is_first_field, field_number,
                 p.color_under_hz, p.fsc_mhz, het, p.final_sos, p.burst_start, p.burst_end, p.burst_abs_ref)
    return FieldChroma(render, ph, u16)
