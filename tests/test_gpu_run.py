"""The whole run (vhsdecode_hip_run) against one decode of the whole tape as a single span
(demod_span_device + decode_vhs_fields_device + detect_dropouts_device) on tapes with RF
dropouts: the run reads the tape through its callback in spans of five fields, carrying
the field sequence from span to span, and must write the same fields, in order, with the
same parity, read-window positions and dropout rows, and the .tbc.json that describes
them.  Runs on an MI355X only."""
import json
import os

import numpy as np
import pytest

from vhsdecode_hip import field, filters, formats, hipmem, synth
from vhsdecode_hip.demod import HipBlockDemodulator
from vhsdecode_hip.run import decode_run

pytestmark = pytest.mark.gpu

N = 32768
CUT = 1024
STRIDE = N - 2 * CUT


def _lut():
    import vhsdecode_hip
    return np.load(os.path.join(os.path.dirname(vhsdecode_hip.__file__), "data", "sinc_lut.npz"))["downscale_sinc_lut"]


@pytest.fixture(scope="module", params=[("NTSC", 9, 31), ("PAL", 7, 32)], ids=["ntsc_vhs", "pal_vhs"])
def tape(request):
    system, nfields, seed = request.param
    fmt = formats.get_format(system, "VHS", "sp")
    fl = filters.build_tape_filters(fmt, 40e6, N)
    t = synth.generate(fmt, fields=nfields, seed=seed, snr_db=30.0, jitter_ns=20.0)
    rf_host = t.rf.copy()
    rng = np.random.default_rng(seed)
    for _ in range(12):  # RF dropouts: stretches of low-level noise
        at = int(rng.integers(400_000, rf_host.size - 400_000))
        ln = int(rng.integers(40, 3000))
        rf_host[at:at + ln] = np.rint(rng.normal(0.0, 40.0, ln)).astype(np.int16)
    lut = _lut()
    dem = HipBlockDemodulator(fl, max_blocks=256)
    dem.set_sinc_table(lut)
    nb = dem.blocks_for(rf_host.size)
    n = nb * STRIDE
    rf = hipmem.DeviceBuffer.from_array(rf_host)
    planes = hipmem.DeviceBuffer(4 * n * 4)
    flags = hipmem.DeviceBuffer(nb * 4)
    video, vlp, env, chroma = (planes.ptr + k * n * 4 for k in range(4))
    dem.demod_span_device(rf.ptr, 0, nb, video, vlp, env, chroma, flags.ptr)
    dec = field.HipVhsFieldDecoder(dem, fmt)
    g = dec.geom
    mf = nfields + 4
    per = g.field_lines * g.out_len
    import ctypes
    from vhsdecode_hip import _lib
    size = ctypes.sizeof(_lib.VhsFieldV1)
    out = hipmem.DeviceBuffer(mf * per * 2)
    recs = hipmem.DeviceBuffer(mf * size)
    meta = hipmem.DeviceBuffer(16)
    dec.decode_async(video, vlp, n, out.ptr, mf, meta.ptr, recs.ptr, meta.ptr + 8, rounds=16)
    hipmem.synchronize()
    count = int(meta.download(np.int32, 1)[0])
    raw = recs.download(np.uint8, mf * size)
    records = [_lib.VhsFieldV1.from_buffer_copy(raw[k * size:(k + 1) * size].tobytes()) for k in range(mf)]
    # one call read the whole span: it stopped at a window past the span's end
    assert count < mf and records[count].status == field.NOT_READ
    assert records[count].window_start + records[count].window_length > n
    fields = out.download(np.uint16, mf * per).reshape(mf, per)
    # the whole span's dropout rows, the parameters decode_run uses
    cap = 1024
    dp = field.dropout_params(fmt, 40e6, capacity=cap)
    loc_ptr, nloc = dem.vhs_field_lines_device(3)
    drop = hipmem.DeviceBuffer(4 * mf * (1 + 3 * cap))
    rows = drop.ptr + 4 * mf
    dem.detect_dropouts_device(env, n, loc_ptr, nloc, recs.ptr, meta.ptr, mf, dp, drop.ptr, rows,
                               rows + 4 * mf * cap, rows + 8 * mf * cap)
    hipmem.synchronize()
    allr = drop.download(np.int32, mf * (1 + 3 * cap))
    nrow = allr[:mf]
    planes3 = allr[mf:].reshape(3, mf, cap)
    ok = [k for k in range(count) if records[k].status == field.DECODED]
    drops = [tuple(planes3[k, f, :nrow[f]].tolist() for k in range(3)) for f in ok]
    first = np.array([records[k].is_first_field for k in ok], dtype=np.int32)
    window = np.array([records[k].window_start for k in ok], dtype=np.int64)
    fields = fields[ok]
    count = len(ok)
    field_len = g.line_len * (g.field_lines - 0.5)
    print(f"\n[{system}] tape {t.rf.size} samples, whole-span fields {count}", flush=True)
    print(f"[{system}] whole-span dropout rows per field {[len(d[0]) for d in drops]}", flush=True)
    return dict(system=system, fmt=fmt, fl=fl, rf=rf_host, lut=lut, count=count, window=window, first=first,
                fields=fields, per=per, drops=drops, span=int(5 * field_len) + N)


def test_run_writes_the_whole_span_fields(tape, tmp_path):
    base = str(tmp_path / "capture")
    res = decode_run(tape["rf"], tape["fmt"], tape["fl"], base, sinc_table=tape["lut"], span_samples=tape["span"])
    data = np.fromfile(base + ".tbc", dtype="<u2")
    doc = json.load(open(base + ".tbc.json"))
    print(f"[{tape['system']}] run: {res.fields_written} fields in {res.spans} spans, {res.elapsed_seconds:.3f} s",
          flush=True)
    assert res.spans >= 2
    assert res.fields_written == tape["count"] == data.size // tape["per"]
    assert data.size == res.fields_written * tape["per"]
    vp = doc["videoParameters"]
    assert vp["numberOfSequentialFields"] == res.fields_written
    assert vp["fieldWidth"] * vp["fieldHeight"] == tape["per"]
    assert vp["system"] == tape["system"] and vp["isSourcePal"] == (tape["system"] == "PAL")
    # the burst window at the output rate, rounded as TBCWriter::write_json rounds it
    b0, b1 = tape["fmt"].sysparams["colorBurstUS"]
    assert (vp["colourBurstStart"], vp["colourBurstEnd"]) == (int(b0 * 1e-6 * vp["sampleRate"] + 0.5),
                                                              int(b1 * 1e-6 * vp["sampleRate"] + 0.5))
    assert list(vp).index("colourBurstStart") == list(vp).index("fieldHeight") + 1
    recs = doc["fields"]
    assert [r["seqNo"] for r in recs] == list(range(1, res.fields_written + 1))
    assert [r["isFirstField"] for r in recs] == [bool(x) for x in tape["first"]]
    # fileLoc: the RF sample where the field's read window starts (TbcDecodedField.StartSample)
    np.testing.assert_array_equal([r["fileLoc"] for r in recs], tape["window"])
    for r, (fl_, sx, ex) in zip(recs, tape["drops"]):
        d = r["dropOuts"]
        assert (d["fieldLine"], d["startx"], d["endx"]) == (fl_, sx, ex)
    assert res.dropout_rows == sum(len(x[0]) for x in tape["drops"]) > 0
    got = data.reshape(-1, tape["per"])
    diff = np.abs(got.astype(np.int32) - tape["fields"].astype(np.int32))
    print(f"[{tape['system']}] samples differing {np.count_nonzero(diff)} of {diff.size}, max {diff.max()}",
          flush=True)
    lines = diff.reshape(diff.shape[0], -1, vp["fieldWidth"]).max(axis=2)
    for f in range(diff.shape[0]):
        bad = np.nonzero(lines[f])[0]
        print(f"  field {f} window {tape['window'][f]}: {bad.size} lines differ"
              + (f" ({bad[:6].tolist()} .. {bad[-3:].tolist()}), max {lines[f].max()}" if bad.size else ""), flush=True)
    # the spans share the whole-span decode's block grid and read windows, so every sample is the same
    np.testing.assert_array_equal(got, tape["fields"])


def test_run_float32_input_and_field_limit(tape, tmp_path):
    a, b = str(tmp_path / "i16"), str(tmp_path / "f32")
    decode_run(tape["rf"], tape["fmt"], tape["fl"], a, sinc_table=tape["lut"], span_samples=tape["span"])
    res = decode_run(tape["rf"].astype(np.float32), tape["fmt"], tape["fl"], b, sinc_table=tape["lut"],
                     span_samples=tape["span"], maximum_output_fields=2)
    assert res.fields_written == 2
    ia = np.fromfile(a + ".tbc", dtype="<u2")
    ib = np.fromfile(b + ".tbc", dtype="<u2")
    np.testing.assert_array_equal(ib, ia[:ib.size])
    assert json.load(open(b + ".tbc.json"))["videoParameters"]["numberOfSequentialFields"] == 2


def test_run_cancel_through_a_short_read(tape, tmp_path):
    """A cancel reported by a short read of a later span (the read callback runs on the
    prefetch thread): the run stops with VHSDECODE_HIP_E_CANCELLED and writes no .tbc.json."""
    base = str(tmp_path / "cancelled")
    state = {"cancel": False}

    def deliver(offset, count):
        if offset > 0:
            state["cancel"] = True
            return 0
        return count

    from vhsdecode_hip import _lib
    with pytest.raises(_lib.HipError) as e:
        decode_run(tape["rf"], tape["fmt"], tape["fl"], base, sinc_table=tape["lut"], span_samples=tape["span"],
                   cancel=lambda: state["cancel"], deliver=deliver)
    assert e.value.status == -30009  # VHSDECODE_HIP_E_CANCELLED
    assert state["cancel"]
    assert not os.path.exists(base + ".tbc.json")
