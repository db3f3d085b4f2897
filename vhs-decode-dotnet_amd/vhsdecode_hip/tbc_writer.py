"""Raw `.tbc` output: little-endian uint16 fields, one after another, the layout
TbcOutputWriter writes (src/VHSDecode.Core/Tbc/TbcOutputWriter.cs:28-90) for the frame
shapes of TbcFrameSpec (Tbc/TbcFrameSpec.cs:20-36).  The device path already produces
fields in this layout (vhsdecode_hip_decode_vhs_fields_device, served windows); this
writes them."""
from __future__ import annotations

from typing import BinaryIO

import numpy as np

# TbcFrameSpec: output samples per line and lines per field
FRAME_SPECS = {"PAL": (1135, 313), "NTSC": (910, 263)}


def field_sample_count(system: str) -> int:
    width, lines = FRAME_SPECS[system]
    return width * lines


def to_little_endian_bytes(samples: np.ndarray) -> bytes:
    """TbcOutputWriter.ToLittleEndianBytes."""
    return np.ascontiguousarray(samples, dtype="<u2").tobytes()


def write_field(destination: BinaryIO, samples: np.ndarray, system: str) -> None:
    """TbcOutputWriter.WriteFrame: one field, rejected unless it has the frame shape's
    sample count."""
    samples = np.asarray(samples)
    if samples.size != field_sample_count(system):
        raise ValueError("TBC field sample count did not match the configured frame shape.")
    destination.write(to_little_endian_bytes(samples.reshape(-1)))


def write_fields(destination: BinaryIO, fields: np.ndarray, count: int, system: str) -> int:
    """Writes the first `count` fields of a [fields, lines * width] uint16 buffer; returns
    the bytes written."""
    per = field_sample_count(system)
    flat = np.asarray(fields).reshape(-1)
    if count < 0 or count * per > flat.size:
        raise ValueError("field count exceeds the buffer")
    for f in range(count):
        write_field(destination, flat[f * per:(f + 1) * per], system)
    return 2 * per * count


def dropouts_object(field_line, start_x, end_x, enabled: bool = True):
    """The field's `"dropOuts"` object of the `.tbc.json` metadata, as
    TbcOutputMetadataWriter.AddDropouts writes it (Decode/TbcOutputMetadataWriter.cs:1091-1107):
    three parallel integer arrays keyed fieldLine / startx / endx, or None (the key is left
    out) when dropout detection is off or the field has no rows."""
    rows = [np.asarray(a).reshape(-1) for a in (field_line, start_x, end_x)]
    if not (rows[0].size == rows[1].size == rows[2].size):
        raise ValueError("dropout row arrays differ in length")
    if not enabled or rows[0].size == 0:
        return None
    return {key: [int(v) for v in a] for key, a in zip(("fieldLine", "startx", "endx"), rows)}


def field_dropout_objects(row_count, field_line, start_x, end_x, capacity: int, fields: int,
                          enabled: bool = True) -> list:
    """Splits the per-field rows of vhsdecode_hip_detect_dropouts_device ([field][capacity]
    arrays, row_count[field] rows each, copied to the host) into one `"dropOuts"` object (or
    None) per field.  A field whose count exceeds `capacity` or is -1 (detection range over
    the device limit) has lost rows: that raises rather than writing a partial list."""
    counts = np.asarray(row_count).reshape(-1)
    if fields < 0 or fields > counts.size:
        raise ValueError("field count exceeds the row counts")
    planes = [np.asarray(a).reshape(-1) for a in (field_line, start_x, end_x)]
    if any(p.size < fields * capacity for p in planes):
        raise ValueError("row arrays are smaller than fields x capacity")
    out = []
    for f in range(fields):
        n = int(counts[f])
        if n < 0 or n > capacity:
            raise ValueError(f"field {f}: {n} dropout rows do not fit the capacity of {capacity}")
        base = f * capacity
        out.append(dropouts_object(*(p[base:base + n] for p in planes), enabled=enabled))
    return out
