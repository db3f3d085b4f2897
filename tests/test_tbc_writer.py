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


def test_dropouts_object_reference_document():
    # the mapped rows and the document checks of CompatibilityTests.cs:16389-16423
    # (a PAL VHS field written with and without --noDOD)
    import json
    rows = ([0, 1, 2, 3, 0], [75, 0, 25, 0, 0], [100, 75, 100, 100, 25])
    doc = json.loads(json.dumps({"fields": [{"dropOuts": tbc_writer.dropouts_object(*rows)}]}))
    d = doc["fields"][0]["dropOuts"]
    assert list(d) == ["fieldLine", "startx", "endx"]
    assert len(d["fieldLine"]) == 5 and d["startx"][0] == 75 and d["endx"][3] == 100
    assert d["fieldLine"] == rows[0] and d["startx"] == rows[1] and d["endx"] == rows[2]
    assert tbc_writer.dropouts_object(*rows, enabled=False) is None
    assert tbc_writer.dropouts_object([], [], []) is None
    with pytest.raises(ValueError):
        tbc_writer.dropouts_object([1, 2], [3], [4])


def test_field_dropout_objects_split_device_layout():
    cap = 4
    counts = np.array([2, 0, 4], np.int32)
    fl = np.arange(3 * cap, dtype=np.int32)
    sx, ex = fl + 100, fl + 200
    objs = tbc_writer.field_dropout_objects(counts, fl, sx, ex, cap, 3)
    assert objs[0] == {"fieldLine": [0, 1], "startx": [100, 101], "endx": [200, 201]}
    assert objs[1] is None
    assert objs[2]["fieldLine"] == [8, 9, 10, 11]
    assert tbc_writer.field_dropout_objects(counts, fl, sx, ex, cap, 2, enabled=False) == [None, None]
    for bad in (5, -1):
        with pytest.raises(ValueError):
            tbc_writer.field_dropout_objects(np.array([bad], np.int32), fl, sx, ex, cap, 1)
    with pytest.raises(ValueError):
        tbc_writer.field_dropout_objects(counts, fl[:5], sx, ex, cap, 3)
