"""The raw .tbc writer: little-endian uint16, field shape checked (TbcOutputWriter.cs:28-90)."""
import io

import numpy as np
import pytest

from vhsdecode_hip import tbc_writer


def test_fields_are_little_endian_uint16_in_order():
    per = tbc_writer.field_sample_count("NTSC")
    assert per == 910 * 263
    rng = np.random.default_rng(3)
    fields = rng.integers(0, 65536, size=(3, per), dtype=np.uint16)
    buf = io.BytesIO()
    n = tbc_writer.write_fields(buf, fields, 2, "NTSC")
    raw = buf.getvalue()
    assert n == len(raw) == 2 * 2 * per
    assert raw[0] == fields[0, 0] & 0xFF and raw[1] == fields[0, 0] >> 8
    back = np.frombuffer(raw, dtype="<u2").reshape(2, per)
    np.testing.assert_array_equal(back, fields[:2])


def test_field_shape_is_checked():
    with pytest.raises(ValueError):
        tbc_writer.write_field(io.BytesIO(), np.zeros(1135 * 312, np.uint16), "PAL")
    with pytest.raises(ValueError):
        tbc_writer.write_fields(io.BytesIO(), np.zeros((1, 1135 * 313), np.uint16), 2, "PAL")
