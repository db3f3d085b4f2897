"""The `.tbc.db` SQLite metadata beside a `.tbc` (ld-decode's metadata database).

Restates TbcSqliteMetadataWriter (src/VHSDecode.Core/Decode/TbcSqliteMetadataWriter.cs)
for the `vhs` command: the schema of :12-134 (user_version 1; tables capture,
pcm_audio_parameters, field_record, vits_metrics, vbi, drop_outs, vitc, closed_caption;
foreign keys not enforced), one capture row from the metadata's `videoParameters`
(:373-468: is_mapped 0, is_subcarrier_locked 1 for NTSC, is_widescreen 0), the PCM row
from `pcmAudioParameters` (:470-507), per field a field_record row (:509-552: field id
seqNo - 1; capture id 0 when dropout detection is off; decode faults 0 stored as NULL;
LaserDisc-only columns NULL; pad 0), a vits_metrics row (:554-581, missing metrics as 0.0),
a vbi row when the field carries vbiData (:583-606) and one drop_outs row per dropout
(:608-651, only with dropout detection on), and the SQLite version number the reference
stamps into the header (:351-359, 3050004 at byte 96).

The input is the `.tbc.json` document (as vhsdecode_hip_run writes it) or the same dict;
the database equals the reference's streaming writer's after its last field.
"""
from __future__ import annotations

import json
import os
import sqlite3
import struct
from typing import Optional, Union

UPSTREAM_SQLITE_VERSION = 3_050_004

SCHEMA = [
    "PRAGMA user_version = 1",
    """CREATE TABLE capture (
    capture_id INTEGER PRIMARY KEY,
    system TEXT NOT NULL CHECK (system IN ('NTSC','PAL','PAL_M')),
    decoder TEXT NOT NULL CHECK (decoder IN ('ld-decode','vhs-decode')),
    git_branch TEXT,
    git_commit TEXT,
    video_sample_rate REAL,
    active_video_start INTEGER,
    active_video_end INTEGER,
    field_width INTEGER,
    field_height INTEGER,
    number_of_sequential_fields INTEGER,
    colour_burst_start INTEGER,
    colour_burst_end INTEGER,
    is_mapped INTEGER CHECK (is_mapped IN (0,1)),
    is_subcarrier_locked INTEGER CHECK (is_subcarrier_locked IN (0,1)),
    is_widescreen INTEGER CHECK (is_widescreen IN (0,1)),
    white_16b_ire INTEGER,
    black_16b_ire INTEGER,
    blanking_16b_ire INTEGER,
    capture_notes TEXT
)""",
    """CREATE TABLE pcm_audio_parameters (
    capture_id INTEGER PRIMARY KEY REFERENCES capture(capture_id) ON DELETE CASCADE,
    bits INTEGER,
    is_signed INTEGER CHECK (is_signed IN (0,1)),
    is_little_endian INTEGER CHECK (is_little_endian IN (0,1)),
    sample_rate REAL
)""",
    """CREATE TABLE field_record (
    capture_id INTEGER NOT NULL REFERENCES capture(capture_id) ON DELETE CASCADE,
    field_id INTEGER NOT NULL,
    audio_samples INTEGER,
    decode_faults INTEGER,
    disk_loc REAL,
    efm_t_values INTEGER,
    field_phase_id INTEGER,
    file_loc INTEGER,
    is_first_field INTEGER CHECK (is_first_field IN (0,1)),
    median_burst_ire REAL,
    pad INTEGER CHECK (pad IN (0,1)),
    sync_conf INTEGER,
    ntsc_is_fm_code_data_valid INTEGER CHECK (ntsc_is_fm_code_data_valid IN (0,1)),
    ntsc_fm_code_data INTEGER,
    ntsc_field_flag INTEGER CHECK (ntsc_field_flag IN (0,1)),
    ntsc_is_video_id_data_valid INTEGER CHECK (ntsc_is_video_id_data_valid IN (0,1)),
    ntsc_video_id_data INTEGER,
    ntsc_white_flag INTEGER CHECK (ntsc_white_flag IN (0,1)),
    PRIMARY KEY (capture_id, field_id)
)""",
    """CREATE TABLE vits_metrics (
    capture_id INTEGER NOT NULL,
    field_id INTEGER NOT NULL,
    b_psnr REAL,
    w_snr REAL,
    FOREIGN KEY (capture_id, field_id)
        REFERENCES field_record(capture_id, field_id) ON DELETE CASCADE,
    PRIMARY KEY (capture_id, field_id)
)""",
    """CREATE TABLE vbi (
    capture_id INTEGER NOT NULL,
    field_id INTEGER NOT NULL,
    vbi0 INTEGER NOT NULL,
    vbi1 INTEGER NOT NULL,
    vbi2 INTEGER NOT NULL,
    FOREIGN KEY (capture_id, field_id)
        REFERENCES field_record(capture_id, field_id) ON DELETE CASCADE,
    PRIMARY KEY (capture_id, field_id)
)""",
    """CREATE TABLE drop_outs (
    capture_id INTEGER NOT NULL,
    field_id INTEGER NOT NULL,
    field_line INTEGER NOT NULL,
    startx INTEGER NOT NULL,
    endx INTEGER NOT NULL,
    FOREIGN KEY (capture_id, field_id)
        REFERENCES field_record(capture_id, field_id) ON DELETE CASCADE,
    PRIMARY KEY (capture_id, field_id, field_line, startx, endx)
)""",
    """CREATE TABLE vitc (
    capture_id INTEGER NOT NULL,
    field_id INTEGER NOT NULL,
    vitc0 INTEGER NOT NULL,
    vitc1 INTEGER NOT NULL,
    vitc2 INTEGER NOT NULL,
    vitc3 INTEGER NOT NULL,
    vitc4 INTEGER NOT NULL,
    vitc5 INTEGER NOT NULL,
    vitc6 INTEGER NOT NULL,
    vitc7 INTEGER NOT NULL,
    FOREIGN KEY (capture_id, field_id)
        REFERENCES field_record(capture_id, field_id) ON DELETE CASCADE,
    PRIMARY KEY (capture_id, field_id)
)""",
    """CREATE TABLE closed_caption (
    capture_id INTEGER NOT NULL,
    field_id INTEGER NOT NULL,
    data0 INTEGER,
    data1 INTEGER,
    FOREIGN KEY (capture_id, field_id)
        REFERENCES field_record(capture_id, field_id) ON DELETE CASCADE,
    PRIMARY KEY (capture_id, field_id)
)""",
]


def _value(obj: Optional[dict], name: str):
    """DbValue (:694-719): bool -> 0/1, finite numbers, strings; anything else NULL."""
    if not obj or name not in obj or obj[name] is None:
        return None
    v = obj[name]
    if isinstance(v, bool):
        return 1 if v else 0
    if isinstance(v, int):
        return v
    if isinstance(v, float):
        return v if v == v and abs(v) != float("inf") else None
    if isinstance(v, str):
        return v
    return None


def _number(obj, name):
    v = _value(obj, name)
    return float(v) if isinstance(v, (int, float)) else None


def _int(obj, name):
    v = _value(obj, name)
    return int(v) if isinstance(v, (int, float)) else None


def _bool_int(obj, name):
    v = _value(obj, name)
    return None if not isinstance(v, (int, float)) else (0 if v == 0 else 1)


def _string(obj, name):
    v = _value(obj, name)
    return v if isinstance(v, str) else None


def _array_int(arr, i) -> int:
    if i < len(arr):
        v = _int({"value": arr[i]}, "value")
        if v is not None:
            return v
    return 0


def write_tbc_db(metadata: Union[dict, str], db_path: str, dropouts_enabled: bool = True) -> None:
    """Writes the `vhs` command's `.tbc.db` for `metadata` (a `.tbc.json` path or dict)."""
    if isinstance(metadata, str):
        with open(metadata, "r", encoding="utf-8") as f:
            metadata = json.load(f)
    video = metadata.get("videoParameters")
    if video is None:
        raise ValueError("Metadata did not contain 'videoParameters'.")
    fields = metadata.get("fields")
    if fields is None:
        raise ValueError("Metadata did not contain 'fields'.")
    pcm = metadata.get("pcmAudioParameters") or {}
    if os.path.exists(db_path):
        os.remove(db_path)
    con = sqlite3.connect(db_path, isolation_level=None)
    try:
        con.execute("PRAGMA foreign_keys = OFF")
        for st in SCHEMA:
            con.execute(st)
        con.execute("BEGIN")
        system = _string(video, "system") or "NTSC"
        cur = con.execute(
            "INSERT INTO capture (system, decoder, git_branch, git_commit, video_sample_rate, active_video_start, "
            "active_video_end, field_width, field_height, number_of_sequential_fields, colour_burst_start, "
            "colour_burst_end, white_16b_ire, black_16b_ire, blanking_16b_ire, is_mapped, is_subcarrier_locked, "
            "is_widescreen) VALUES (?,?,?,?,?,?,?,?,?,?,?,?,?,?,?,?,?,?)",
            (system, _string(video, "decoder") or "ld-decode", _string(video, "gitBranch") or "",
             _string(video, "gitCommit") or "", _number(video, "sampleRate"), _value(video, "activeVideoStart"),
             _value(video, "activeVideoEnd"), _value(video, "fieldWidth"), _value(video, "fieldHeight"),
             _value(video, "numberOfSequentialFields"), _value(video, "colourBurstStart"),
             _value(video, "colourBurstEnd"), _value(video, "white16bIre"), _value(video, "black16bIre"),
             _value(video, "blanking16bIre"), 0, 1 if system.upper() == "NTSC" else 0, 0))
        capture_id = cur.lastrowid
        con.execute("INSERT INTO pcm_audio_parameters (capture_id, bits, is_little_endian, is_signed, sample_rate) "
                    "VALUES (?,?,?,?,?)", (capture_id, _value(pcm, "bits"), _bool_int(pcm, "isLittleEndian"),
                                           _bool_int(pcm, "isSigned"), _number(pcm, "sampleRate")))
        field_capture = capture_id if dropouts_enabled else 0
        for field in fields:
            seq = _int(field, "seqNo")
            if seq is None:
                raise ValueError("VHS/CVBS field metadata did not contain 'seqNo'.")
            fid = seq - 1
            faults = _int(field, "decodeFaults")
            con.execute(
                "INSERT INTO field_record (capture_id, field_id, audio_samples, decode_faults, disk_loc, "
                "efm_t_values, field_phase_id, file_loc, is_first_field, median_burst_ire, pad, sync_conf) "
                "VALUES (?,?,?,?,?,?,?,?,?,?,?,?)",
                (field_capture, fid, None, None if faults == 0 else faults, _number(field, "diskLoc"), None,
                 _value(field, "fieldPhaseID"), _value(field, "fileLoc"),
                 _bool_int(field, "isFirstField") or 0, None, 0, _value(field, "syncConf")))
            metrics = field.get("vitsMetrics")
            con.execute("INSERT INTO vits_metrics (capture_id, field_id, w_snr, b_psnr) VALUES (?,?,?,?)",
                        (field_capture, fid, _number(metrics, "wSNR") or 0.0, _number(metrics, "bPSNR") or 0.0))
            vbi = (field.get("vbi") or {}).get("vbiData")
            if vbi:
                con.execute("INSERT INTO vbi (capture_id, field_id, vbi0, vbi1, vbi2) VALUES (?,?,?,?,?)",
                            (field_capture, fid, _array_int(vbi, 0), _array_int(vbi, 1), _array_int(vbi, 2)))
            d = field.get("dropOuts")
            if dropouts_enabled and d:
                lines, starts, ends = d.get("fieldLine"), d.get("startx"), d.get("endx")
                if lines is not None and starts is not None and ends is not None:
                    for i in range(min(len(lines), len(starts), len(ends))):
                        con.execute("INSERT INTO drop_outs (capture_id, field_id, field_line, startx, endx) "
                                    "VALUES (?,?,?,?,?)", (field_capture, fid, _array_int(lines, i),
                                                           _array_int(starts, i), _array_int(ends, i)))
        con.execute("COMMIT")
    finally:
        con.close()
    with open(db_path, "r+b") as f:
        f.seek(96)
        f.write(struct.pack(">i", UPSTREAM_SQLITE_VERSION))
