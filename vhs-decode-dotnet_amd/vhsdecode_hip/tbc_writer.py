"""Raw `.tbc` output: little-endian uint16 fields, one after another, the layout
TbcOutputWriter writes (src/VHSDecode.Core/Tbc/TbcOutputWriter.cs:28-90) for the frame
shapes of TbcFrameSpec (Tbc/TbcFrameSpec.cs:20-36).  The device path already produces
fields in this layout (vhsdecode_hip_decode_span_fields_device, served windows); this
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
