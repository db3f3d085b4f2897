"""Whole-run decode over the HIP backend (vhsdecode_hip_run): RF read through a callback
in spans, every span decoded on the device, the fields written to <base>.tbc and
<base>.tbc.json.  The counterpart of the reference's vhsdecode_cuda_fast_run
(src/VHSDecode.CudaFast.Native/src/vhsdecode_cuda_fast.cpp:238-370) as
CudaFastDecodeRunner.TryDecodeAndWrite drives it."""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Callable, Optional

import numpy as np

from . import _lib, field
from .demod import FilterTables, _ptr
from .filters import TapeFilterSet
from .formats import FormatParameters


@dataclass
class RunResult:
    fields_written: int
    output_line_length: int
    output_field_lines: int
    elapsed_seconds: float
    spans: int
    dropout_rows: int


def _system_name(fmt: FormatParameters) -> bytes:
    """videoParameters.system as TBCWriter::write_json names it (tbc_writer.cpp:219-225)."""
    return b"PAL-M" if fmt.system in ("MPAL", "PAL-M") else (b"PAL" if fmt.parent_system == "PAL" else b"NTSC")


def _burst_us(fmt: FormatParameters) -> tuple:
    """The colour-burst window in microseconds (sysparams colorBurstUS)."""
    burst = fmt.sysparams.get("colorBurstUS", (0.0, 0.0))
    return float(burst[0]), float(burst[1])


def decode_run(rf: np.ndarray, fmt: FormatParameters, filters: TapeFilterSet, output_base: str,
               sinc_table: Optional[np.ndarray] = None, dropouts: bool = True, dropout_capacity: int = 1024,
               span_samples: int = 0, maximum_output_fields: int = 0, overwrite: bool = True,
               cancel: Optional[Callable[[], bool]] = None, block_cut: int = 1024, block_cut_end: int = 1024,
               device: int = 0, deliver: Optional[Callable[[int, int], int]] = None,
               chroma: bool = False, sqlite: bool = False) -> RunResult:
    """Decodes the whole of `rf` (int16, or float32 samples the library rounds to int16)
    read through the library's callback; returns the run's result record.  `deliver`
    (offset, count) -> samples, when given, bounds what each read returns (a short read
    ends the input, or reports a cancel).  `chroma`: also write the colour-under chroma to
    <output_base>_chroma.tbc (the luma then renders at the burst-refined locations).
    `sqlite`: also write <output_base>.tbc.db from the run's .tbc.json (tbc_db)."""
    L = _lib.load()
    src = np.ascontiguousarray(rf)
    if src.dtype == np.int16:
        fmt_code = _lib.INPUT_INT16
    elif src.dtype == np.float32:
        fmt_code = _lib.INPUT_FLOAT32
    else:
        raise TypeError("rf must be int16 or float32")
    item = src.itemsize
    base_addr = src.ctypes.data

    def read(_user, dst, offset, count):
        n = max(0, min(int(count), src.size - int(offset)))
        if deliver is not None:
            n = max(0, min(n, int(deliver(int(offset), int(count)))))
        if n:
            ctypes.memmove(dst, base_addr + int(offset) * item, n * item)
        return n

    def cancelled(_user):
        return 1 if cancel is not None and cancel() else 0

    read_cb = _lib.READ_CALLBACK(read)
    cancel_cb = _lib.CANCEL_CALLBACK(cancelled)
    tables = FilterTables(filters)
    g = field.sync_geometry(fmt, filters.sample_rate_hz)
    ire0, hz_ire, ozero, vsync, oscale = g.conv
    fp = _lib.FieldParamsV1(ctypes.sizeof(_lib.FieldParamsV1), g.threshold, g.fs_mhz, g.hsync / g.fs_mhz, ire0,
                            hz_ire, 1 if g.is_pal else 0, 1, g.out_len, g.field_lines, g.nominal, g.smoothing,
                            ozero, vsync, oscale)
    stride = filters.block_length - block_cut - block_cut_end
    sync = field.vhs_sync_params(fmt, filters.sample_rate_hz, filters.block_length, block_cut, stride)
    lut = None if sinc_table is None else np.ascontiguousarray(sinc_table, dtype=np.float32).reshape(-1)
    fm = _lib.ServeFormatV1(ctypes.sizeof(_lib.ServeFormatV1), ctypes.pointer(tables.struct), block_cut,
                            block_cut_end, ctypes.pointer(fp), ctypes.pointer(sync),
                            _ptr(lut, _lib._fp) if lut is not None else None, 0 if lut is None else lut.size)
    dp = field.dropout_params(fmt, filters.sample_rate_hz, capacity=dropout_capacity) if dropouts else None
    cfg = _lib.RunConfigV1(ctypes.sizeof(_lib.RunConfigV1), device, ctypes.pointer(fm),
                           ctypes.pointer(dp) if dp is not None else None, src.size, fmt_code,
                           maximum_output_fields, span_samples, output_base.encode("utf-8"), 1 if overwrite else 0,
                           read_cb, cancel_cb, None, _system_name(fmt), *_burst_us(fmt))
    cp = None
    if chroma:
        from .chroma import chroma_params, to_struct
        cp = to_struct(chroma_params(fmt), field.input_line_length(fmt, filters.sample_rate_hz))
        cfg.chroma = ctypes.pointer(cp)
    res = _lib.RunResultV1()
    res.struct_size = ctypes.sizeof(_lib.RunResultV1)
    _lib.check(L.vhsdecode_hip_run(ctypes.byref(cfg), ctypes.byref(res)))
    if sqlite:
        from .tbc_db import write_tbc_db
        write_tbc_db(output_base + ".tbc.json", output_base + ".tbc.db", dropouts_enabled=dropouts)
    return RunResult(res.fields_written, res.output_line_length, res.output_field_lines, res.elapsed_seconds,
                     res.spans, res.dropout_rows)
