"""Served decode windows over the HIP backend (vhsdecode_hip_serve_*): many streams,
seekable windows admitted into device slots as they free up, each window decoded end
to end (RF upload, block demod, field decode, download of the uint16 fields) on its
slot's own HIP stream.  The served counterpart of the reference's preview windows
(src/VHSDecode.Preview), made asynchronous."""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _lib, field
from .demod import FilterTables, _ptr
from .filters import TapeFilterSet
from .formats import FormatParameters


@dataclass
class WindowResult:
    status: int
    field_count: int
    queued_ms: float
    latency_ms: float


class HipServer:
    def __init__(self, max_slots: int = 16, hbm_budget_bytes: int = 64 << 30, max_window_samples: int = 80_000_000,
                 max_fields: int = 128, device: int = 0):
        self._L = _lib.load()
        cfg = _lib.ServeConfigV1(ctypes.sizeof(_lib.ServeConfigV1), device, max_slots, hbm_budget_bytes,
                                 max_window_samples, max_fields)
        h = ctypes.c_void_p()
        _lib.check(self._L.vhsdecode_hip_serve_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        self._keep = []
        self.geoms = []

    def add_format(self, fmt: FormatParameters, filters: TapeFilterSet, sinc_table: Optional[np.ndarray] = None,
                   block_cut: int = 1024, block_cut_end: int = 1024) -> int:
        tables = FilterTables(filters)
        g = field.sync_geometry(fmt, filters.sample_rate_hz)
        ire0, hz_ire, ozero, vsync, oscale = g.conv
        fp = _lib.FieldParamsV1(ctypes.sizeof(_lib.FieldParamsV1), g.threshold, g.fs_mhz, g.hsync / g.fs_mhz, ire0,
                                hz_ire, 1 if g.is_pal else 0, 1, g.out_len, g.field_lines, g.nominal, g.smoothing,
                                ozero, vsync, oscale)
        stride = filters.block_length - block_cut - block_cut_end
        sync = field.vhs_sync_params(fmt, filters.sample_rate_hz, filters.block_length, block_cut, stride)
        lut = None if sinc_table is None else np.ascontiguousarray(sinc_table, dtype=np.float32).reshape(-1)
        f = _lib.ServeFormatV1(ctypes.sizeof(_lib.ServeFormatV1), ctypes.pointer(tables.struct), block_cut,
                               block_cut_end, ctypes.pointer(fp), ctypes.pointer(sync),
                               _ptr(lut, _lib._fp) if lut is not None else None, 0 if lut is None else lut.size)
        fid = ctypes.c_int32()
        _lib.check(self._L.vhsdecode_hip_serve_add_format(self._h, ctypes.byref(f), ctypes.byref(fid)))
        self.geoms.append(g)
        return fid.value

    def add_stream(self, format_id: int, samples_ptr: int, length: int, loop: bool = True, keep=None) -> int:
        """samples_ptr: host int16 samples (pinned memory is uploaded by DMA), alive while the server is."""
        if keep is not None:
            self._keep.append(keep)
        sid = ctypes.c_int32()
        _lib.check(self._L.vhsdecode_hip_serve_add_stream(self._h, format_id, ctypes.c_void_p(samples_ptr), length,
                                                          1 if loop else 0, ctypes.byref(sid)))
        return sid.value

    def submit(self, stream_id: int, first_sample: int, sample_count: int, max_fields: int, out_ptr: int) -> int:
        w = _lib.WindowV1(ctypes.sizeof(_lib.WindowV1), stream_id, first_sample, sample_count, max_fields,
                          ctypes.c_void_p(out_ptr))
        t = ctypes.c_uint64()
        _lib.check(self._L.vhsdecode_hip_serve_submit(self._h, ctypes.byref(w), ctypes.byref(t)))
        return t.value

    def wait(self, ticket: int, timeout_ms: int = -1) -> WindowResult:
        r = _lib.WindowResultV1()
        r.struct_size = ctypes.sizeof(_lib.WindowResultV1)
        _lib.check(self._L.vhsdecode_hip_serve_wait(self._h, ticket, timeout_ms, ctypes.byref(r)))
        return WindowResult(r.status, r.field_count, r.queued_ms, r.latency_ms)

    def close(self):
        if getattr(self, "_h", None):
            self._L.vhsdecode_hip_serve_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
