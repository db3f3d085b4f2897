"""The device colour-under chroma (vhsdecode_hip_decode_vhs_fields_chroma_device) against the
oracle's restatement (oracle/chroma.py field_chain) on the same demodulated span: the
device's own chroma plane and hsync-refined line table are handed to the oracle, which
runs the phase analysis (track-phase start carried from field to field), the burst line
refinement, the render resample and the chroma decode.

Bars: the burst-refined render locations within 1e-9 samples (the device's FP64 atan2,
hypot and cos are the device library's, one ulp from the C library's at most); the
uint16 chroma within 1 LSB except at < 1e-4 of the samples (the wow smoothing's closed
form, as for the luma, and the FP64 transcendentals); the luma rendered at the
burst-refined locations within 1 LSB of the oracle's TBC at the same locations; the
plain sequence's records unchanged by the chroma (burst refinement only moves the render
locations).  Runs on an MI355X only."""
import os

import numpy as np
import pytest

from oracle import chroma as C
from oracle import oracle as O
from vhsdecode_hip import field, filters, formats, hipmem, synth
from vhsdecode_hip.demod import HipBlockDemodulator

pytestmark = pytest.mark.gpu

N, CUT = 32768, 1024
STRIDE = N - 2 * CUT


def _lut():
    import vhsdecode_hip
    return np.load(os.path.join(os.path.dirname(vhsdecode_hip.__file__), "data", "sinc_lut.npz"))[
        "downscale_sinc_lut"]


@pytest.fixture(scope="module", params=[("NTSC", 6, 81, 30.0), ("PAL", 5, 82, 30.0)], ids=["ntsc", "pal"])
def run(request):
    system, nf, seed, snr = request.param
    print(f"\n[chroma] {system} fields={nf} seed={seed} snr={snr}", flush=True)
    fmt = formats.get_format(system, "VHS", "sp")
    tape = synth.generate(fmt, fields=nf, seed=seed, snr_db=snr, jitter_ns=20.0, chroma_under="modulated")
    fl = filters.build_tape_filters(fmt, 40e6, N)
    dem = HipBlockDemodulator(fl, max_blocks=320)
    dem.set_sinc_table(_lut())
    nb = dem.blocks_for(tape.rf.size)
    n = nb * STRIDE
    rf = hipmem.DeviceBuffer.from_array(tape.rf)
    planes = hipmem.DeviceBuffer(4 * n * 4)
    flags = hipmem.DeviceBuffer(nb * 4)
    video, vlp, env, chroma = (planes.ptr + k * n * 4 for k in range(4))
    dem.demod_span_device(rf.ptr, 0, nb, video, vlp, env, chroma, flags.ptr)
    plain = field.HipVhsFieldDecoder(dem, fmt).decode_span(video, vlp, n, max_fields=12)
    dec = field.HipVhsFieldDecoder(dem, fmt, chroma=True)
    got = dec.decode_span(video, vlp, n, max_fields=12, want_tables=True, span_chroma_ptr=chroma)
    count = len(got.records)
    tables = [np.concatenate([t[w] for t in got.tables])[:count] for w in range(4)]
    planes_h = planes.download(np.float32, 4 * n).reshape(4, n)
    print(f"[chroma] {system} records {count}, decoded "
          f"{sum(r.status == field.DECODED for r in got.records)}, calls {len(got.tables)}", flush=True)
    return dict(system=system, fmt=fmt, dec=dec, got=got, plain=plain, tables=tables,
                video=planes_h[0].astype(np.float64) + dem.bias, chroma=planes_h[3].astype(np.float64),
                keep=(rf, planes, flags))


def _oracle_fields(run):
    fmt, dec, got = run["fmt"], run["dec"], run["got"]
    geom, p = dec.geom, dec.chroma
    fsc_shift = 0.0 if geom.is_pal else ((117.25 / 360.0) / ((float(fmt.sysparams["outfreq"]) * 1e6) / 4e6)) * 40.0
    refined = run["tables"][1]
    rot, number, out = 0, 0, {}
    for k, r in enumerate(got.records):
        if r.status != field.DECODED:
            continue
        src = run["chroma"][r.window_start:r.window_start + r.window_length]
        fc = C.field_chain(src, refined[k], p, bool(r.is_first_field), number, rot,
                           field.input_line_length(fmt), _lut(), geom.nominal, geom.smoothing, fsc_shift)
        out[k] = fc
        rot = fc.phase.next_index
        number += 1
    return out


@pytest.fixture(scope="module")
def oracle_fields(run):
    return _oracle_fields(run)


def test_chroma_does_not_change_the_sequence(run):
    a, b = run["plain"].records, run["got"].records
    assert len(a) == len(b) >= 3
    for x, y in zip(a, b):
        assert (x.status, x.window_start, x.window_length, x.line0, x.next_begin, x.is_first_field) == \
            (y.status, y.window_start, y.window_length, y.line0, y.next_begin, y.is_first_field)


def test_burst_refined_render_locations_match_oracle(run, oracle_fields):
    render = run["tables"][2]
    p = run["dec"].chroma
    worst = 0.0
    for k, fc in oracle_fields.items():
        nl = fc.render.size
        d = render[k, :nl]
        worst = max(worst, float(np.nanmax(np.abs(d - fc.render))))
    print(f"[chroma] {run['system']} render |device - oracle| max {worst:.3e} samples", flush=True)
    assert worst < 1e-9


def test_chroma_fields_match_oracle(run, oracle_fields):
    worst, over, total = 0, 0, 0
    for k, fc in oracle_fields.items():
        got = run["got"].chroma[k]
        diff = np.abs(got.astype(np.int32) - fc.chroma.astype(np.int32))
        worst = max(worst, int(diff.max()))
        over += int(np.count_nonzero(diff > 1))
        total += diff.size
        exact = np.count_nonzero(diff == 0) / diff.size
        print(f"[chroma] {run['system']} field {k}: max |lsb| {int(diff.max())}, exact {exact:.4f}", flush=True)
    assert total > 0
    assert over / total < 1e-4


def test_luma_at_burst_refined_locations_matches_oracle(run, oracle_fields):
    geom = run["dec"].geom
    worst = 0
    for k, fc in oracle_fields.items():
        r = run["got"].records[k]
        first_line = 1 if not geom.is_pal else (3 if r.is_first_field else 4)
        src = run["video"][r.window_start:r.window_start + r.window_length]
        _, want = O.tbc_resample(src, run["tables"][2][k, :fc.render.size], first_line, geom.field_lines,
                                 geom.out_len, _lut(), nominal=geom.nominal, smoothing=geom.smoothing, conv=geom.conv)
        diff = np.abs(run["got"].outputs[k].astype(np.int32) - want.astype(np.int32))
        worst = max(worst, int(diff.max()))
    assert worst <= 1


def test_decoded_burst_level_and_frequency(run):
    p = run["dec"].chroma
    for k, u16 in run["got"].chroma.items():
        u = u16.reshape(p.lines, p.line_len).astype(np.float64) - 32767.0
        burst = u[40:200, p.burst_start + 8:p.burst_end - 8]
        rms = np.sqrt(np.mean((burst - burst.mean(axis=1, keepdims=True)) ** 2))
        assert 0.5 * p.burst_abs_ref < rms < 1.5 * p.burst_abs_ref, rms
        spec = np.abs(np.fft.rfft(burst, axis=1)).mean(axis=0)
        assert abs((np.argmax(spec[1:]) + 1) / burst.shape[1] - 0.25) < 0.05


@pytest.mark.parametrize("system", ["NTSC", "PAL"])
def test_whole_run_writes_the_span_decode_chroma(system, tmp_path):
    """vhsdecode_hip_run with chroma: spans of about four fields, the field sequence and the
    chroma track-phase state carried from span to span; <base>.tbc and <base>_chroma.tbc
    equal one decode of the whole tape as a single span, field for field."""
    from vhsdecode_hip.run import decode_run
    fmt = formats.get_format(system, "VHS", "sp")
    tape = synth.generate(fmt, fields=9, seed=91, snr_db=30.0, jitter_ns=20.0, chroma_under="modulated")
    fl = filters.build_tape_filters(fmt, 40e6, N)
    lut = _lut()
    dem = HipBlockDemodulator(fl, max_blocks=320)
    dem.set_sinc_table(lut)
    nb = dem.blocks_for(tape.rf.size)
    n = nb * STRIDE
    rf = hipmem.DeviceBuffer.from_array(tape.rf)
    planes = hipmem.DeviceBuffer(4 * n * 4)
    flags = hipmem.DeviceBuffer(nb * 4)
    video, vlp, env, chroma = (planes.ptr + k * n * 4 for k in range(4))
    dem.demod_span_device(rf.ptr, 0, nb, video, vlp, env, chroma, flags.ptr)
    dec = field.HipVhsFieldDecoder(dem, fmt, chroma=True)
    one = dec.decode_span(video, vlp, n, max_fields=16, span_chroma_ptr=chroma)
    keys = sorted(one.outputs)
    base = str(tmp_path / "run")
    field_len = int(round(40e6 / (2 * float(fmt.sysparams["FPS"]))))
    res = decode_run(tape.rf, fmt, fl, base, sinc_table=lut, dropouts=False, span_samples=4 * field_len,
                     chroma=True, sqlite=True)
    import sqlite3
    con = sqlite3.connect(base + ".tbc.db")
    try:
        assert con.execute("SELECT COUNT(*), MIN(field_id), MAX(capture_id) FROM field_record").fetchone() == \
            (res.fields_written, 0, 0)
        assert con.execute("SELECT system, field_width FROM capture").fetchone()[0] == system
    finally:
        con.close()
    print(f"\n[chroma run] {system} fields {res.fields_written} spans {res.spans}; one span {len(keys)}", flush=True)
    assert res.spans >= 2 and res.fields_written == len(keys) >= 6
    g = dec.geom
    per = g.field_lines * g.out_len
    luma = np.fromfile(base + ".tbc", dtype="<u2").reshape(-1, per)
    chrom = np.fromfile(base + "_chroma.tbc", dtype="<u2").reshape(-1, per)
    assert luma.shape[0] == chrom.shape[0] == len(keys)
    for i, k in enumerate(keys):
        np.testing.assert_array_equal(luma[i], one.outputs[k])
        np.testing.assert_array_equal(chrom[i], one.chroma[k])
