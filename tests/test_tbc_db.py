"""`.tbc.db` writer (vhsdecode_hip/tbc_db.py) against TbcSqliteMetadataWriter's contract
(src/VHSDecode.Core/Decode/TbcSqliteMetadataWriter.cs) and the reference's own checks of it
(tests/VHSDecode.Tests/CompatibilityTests.cs:13999-14011: with --noDOD the capture row is 1
and the field and VITS rows carry capture id 0)."""
import sqlite3
import struct

import pytest

from vhsdecode_hip import tbc_db


def _meta(n=3, with_dropouts=True):
    fields = []
    for k in range(n):
        f = {"isFirstField": k % 2 == 0, "seqNo": k + 1, "fileLoc": 1000 * k, "syncConf": 100,
             "fieldPhaseID": 0, "decodeFaults": 0 if k else 4}
        if with_dropouts:
            f["dropOuts"] = {"fieldLine": [7, 9], "startx": [10, 20], "endx": [30, 40]} if k == 1 else \
                {"fieldLine": [], "startx": [], "endx": []}
        fields.append(f)
    return {"videoParameters": {"system": "PAL", "decoder": "vhs-decode", "sampleRate": 17734375.0,
                                "activeVideoStart": 185, "activeVideoEnd": 1107, "fieldWidth": 1135,
                                "fieldHeight": 313, "numberOfSequentialFields": n, "colourBurstStart": 98,
                                "colourBurstEnd": 138, "white16bIre": 54016, "black16bIre": 16384,
                                "blanking16bIre": 16384, "isSubcarrierLocked": False},
            "pcmAudioParameters": {"bits": 16, "isLittleEndian": True, "isSigned": True, "sampleRate": 44100.0},
            "fields": fields}


def _q(path, sql):
    con = sqlite3.connect(path)
    try:
        return con.execute(sql).fetchall()
    finally:
        con.close()


def test_schema_and_header(tmp_path):
    db = str(tmp_path / "a.tbc.db")
    tbc_db.write_tbc_db(_meta(), db)
    with open(db, "rb") as f:
        head = f.read(100)
    assert struct.unpack(">i", head[96:100])[0] == 3_050_004
    assert _q(db, "PRAGMA user_version") == [(1,)]
    tables = [r[0] for r in _q(db, "SELECT name FROM sqlite_master WHERE type='table' ORDER BY rowid")]
    assert tables == ["capture", "pcm_audio_parameters", "field_record", "vits_metrics", "vbi", "drop_outs",
                      "vitc", "closed_caption"]
    sql = dict(_q(db, "SELECT name, sql FROM sqlite_master WHERE type='table'"))
    assert "PRIMARY KEY (capture_id, field_id, field_line, startx, endx)" in sql["drop_outs"]


def test_rows(tmp_path):
    db = str(tmp_path / "b.tbc.db")
    tbc_db.write_tbc_db(_meta(), db)
    cap = _q(db, "SELECT capture_id, system, decoder, field_width, field_height, number_of_sequential_fields, "
                 "colour_burst_start, colour_burst_end, is_mapped, is_subcarrier_locked, is_widescreen FROM capture")
    assert cap == [(1, "PAL", "vhs-decode", 1135, 313, 3, 98, 138, 0, 0, 0)]
    assert _q(db, "SELECT bits, is_little_endian, is_signed, sample_rate FROM pcm_audio_parameters") == \
        [(16, 1, 1, 44100.0)]
    rows = _q(db, "SELECT capture_id, field_id, decode_faults, file_loc, is_first_field, pad, sync_conf, "
                  "audio_samples, median_burst_ire FROM field_record ORDER BY field_id")
    assert rows == [(1, 0, 4, 0, 1, 0, 100, None, None), (1, 1, None, 1000, 0, 0, 100, None, None),
                    (1, 2, None, 2000, 1, 0, 100, None, None)]
    assert _q(db, "SELECT field_id, w_snr, b_psnr FROM vits_metrics ORDER BY field_id") == \
        [(0, 0.0, 0.0), (1, 0.0, 0.0), (2, 0.0, 0.0)]
    assert _q(db, "SELECT field_id, field_line, startx, endx FROM drop_outs ORDER BY field_line") == \
        [(1, 7, 10, 30), (1, 9, 20, 40)]


def test_no_dod_capture_ids(tmp_path):
    db = str(tmp_path / "c.tbc.db")
    tbc_db.write_tbc_db(_meta(1, with_dropouts=False), db, dropouts_enabled=False)
    assert _q(db, "SELECT capture_id FROM capture") == [(1,)]
    assert _q(db, "SELECT capture_id FROM field_record WHERE field_id = 0") == [(0,)]
    assert _q(db, "SELECT capture_id FROM vits_metrics WHERE field_id = 0") == [(0,)]
    assert _q(db, "SELECT COUNT(*) FROM drop_outs") == [(0,)]


def test_missing_seqno_is_an_error(tmp_path):
    m = _meta(1)
    del m["fields"][0]["seqNo"]
    with pytest.raises(ValueError):
        tbc_db.write_tbc_db(m, str(tmp_path / "d.tbc.db"))
