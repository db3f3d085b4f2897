"""The device VHS field sequence (vhsdecode_hip_decode_vhs_fields_device) against the
oracle's restatement of the reference's sync path (oracle/sync.cpp, with no serration
search, the device's documented choice) on the same demodulated span: every field's
read window, sync levels, DC offset, threshold, pulse counts, parity, line 0, first
hsync and next read offset, its upstream line locations, its hsync-refined locations
and its TBC output.  The sync reference handed to the oracle is the device's own FP32
plane (plus its bias) promoted to FP64, so the comparison isolates the sync path.

Bars: integers and FP64 records bit-exact; line tables bit-exact; the uint16 TBC
output within 1 LSB (the wow level smoothing runs in closed form on the device, an
FP64 rounding difference from the sequential recurrence that can move a sample across
an integer boundary).  Runs on an MI355X only."""
import os

import numpy as np
import pytest

from oracle import oracle as O
from vhsdecode_hip import field, filters, formats, hipmem, synth
from vhsdecode_hip.demod import HipBlockDemodulator

pytestmark = pytest.mark.gpu

N, CUT = 32768, 1024
STRIDE = N - 2 * CUT
MAX_FIELDS = 12
RECORD_FIELDS = ("begin", "window_start", "window_length", "level_source", "raw_pulses", "refined_pulses",
                 "is_first_field", "parity_confidence", "sync_level", "blank_level", "dc_offset", "threshold",
                 "hsync_median", "mean_line_length", "line0", "first_hsync", "first_hsync_line", "next_offset",
                 "next_begin")


def _lut():
    import vhsdecode_hip
    return np.load(os.path.join(os.path.dirname(vhsdecode_hip.__file__), "data", "sinc_lut.npz"))[
        "downscale_sinc_lut"]


@pytest.fixture(scope="module", params=[("NTSC", 5, 31, 28.0), ("PAL", 4, 32, 30.0), ("NTSC", 4, 41, 34.0),
                                        ("PAL", 5, 42, 26.0)],
                ids=["ntsc_a", "pal_a", "ntsc_b", "pal_b"])
def decoded(request):
    system, nf, seed, snr = request.param
    print(f"\n[vhs sync] {system} fields={nf} seed={seed} snr={snr}", flush=True)
    fmt = formats.get_format(system, "VHS", "sp")
    tape = synth.generate(fmt, fields=nf, seed=seed, snr_db=snr, jitter_ns=20.0)
    fl = filters.build_tape_filters(fmt, 40e6, N)
    dem = HipBlockDemodulator(fl, max_blocks=160)
    dem.set_sinc_table(_lut())
    nb = dem.blocks_for(tape.rf.size)
    n = nb * STRIDE
    rf = hipmem.DeviceBuffer.from_array(tape.rf)
    planes = hipmem.DeviceBuffer(4 * n * 4)
    flags = hipmem.DeviceBuffer(nb * 4)
    video, vlp, env, chroma = (planes.ptr + k * n * 4 for k in range(4))
    dem.demod_span_device(rf.ptr, 0, nb, video, vlp, env, chroma, flags.ptr)
    geom = field.sync_geometry(fmt)
    sync = field.vhs_sync_params(fmt)
    per = geom.field_lines * geom.out_len
    out = hipmem.DeviceBuffer(MAX_FIELDS * per * 2)
    recs = hipmem.DeviceBuffer(MAX_FIELDS * 256)
    meta = hipmem.DeviceBuffer(64)
    import ctypes
    from vhsdecode_hip import _lib
    size = ctypes.sizeof(_lib.VhsFieldV1)
    # a call decodes fields until one's read window differs from the speculated one; the
    # sequence then continues from the next call's start with the carried state
    records, tables, u16 = [], [[], [], []], []
    begin, calls = 0, 0
    while len(records) < MAX_FIELDS and calls < MAX_FIELDS:
        out.fill(0)
        dem.decode_vhs_fields_device(vlp, video, n, begin, sync, geom, MAX_FIELDS - len(records), out.ptr, meta.ptr,
                                     fields_ptr=recs.ptr, next_begin_ptr=meta.ptr + 8, reset_state=calls == 0)
        hipmem.synchronize()
        calls += 1
        count = int(meta.download(np.int32, 1)[0])
        next_begin = int(meta.download(np.int64, 1, offset=8)[0])
        print(f"[vhs sync] call {calls}: begin={begin} fields={count} next={next_begin}", flush=True)
        if count == 0:
            break
        raw = recs.download(np.uint8, count * size)
        records += [_lib.VhsFieldV1.from_buffer_copy(raw[k * size:(k + 1) * size].tobytes()) for k in range(count)]
        for which in range(3):
            ptr, stride = dem.vhs_field_lines_device(which)
            tables[which].append(hipmem.download_ptr(ptr, np.float64, count * stride).reshape(count, stride))
        u16.append(out.download(np.uint16, count * per).reshape(count, per))
        begin = next_begin
    count = len(records)
    tables = [np.concatenate(t) if t else np.zeros((0, 1)) for t in tables]
    u16 = np.concatenate(u16) if u16 else np.zeros((0, per), dtype=np.uint16)
    planes_h = planes.download(np.float32, 4 * n).reshape(4, n)
    vlp64 = planes_h[1].astype(np.float64) + dem.bias
    cfg = O.sync_config(fmt, serration_mode=O.SERRATION_NONE)
    ofields, olocs, oshift = O.vhs_field_sequence(vlp64, cfg, 0, MAX_FIELDS, want_shifted=True)
    print(f"[vhs sync] {system} device fields={count} oracle fields={len(ofields)} "
          f"statuses={[f.status for f in ofields]} calls={calls}", flush=True)
    return dict(system=system, fmt=fmt, nf=nf, geom=geom, dem=dem, count=count, next_begin=begin, calls=calls,
                records=records, tables=tables, video=planes_h[0].astype(np.float64) + dem.bias, u16=u16,
                ofields=ofields, olocs=olocs, oshift=oshift, keep=(rf, planes, flags, out, recs, meta))


def _leading_ok(ofields):
    k = 0
    while k < len(ofields) and ofields[k].status == 0:
        k += 1
    return k


def test_field_records_match_oracle(decoded):
    count, records, ofields = decoded["count"], decoded["records"], decoded["ofields"]
    assert count == _leading_ok(ofields) >= decoded["nf"] - 2
    bad = []
    for k in range(count):
        d, o = records[k], ofields[k]
        assert d.status == 0
        for name in RECORD_FIELDS:
            if getattr(d, name) != getattr(o, name):
                bad.append((k, name, getattr(d, name), getattr(o, name)))
    for b in bad:
        print("record mismatch", b, flush=True)
    assert not bad
    # the sequence continues where the last decoded field says
    assert decoded["next_begin"] == records[count - 1].next_begin


def test_parity_alternates(decoded):
    firsts = [r.is_first_field for r in decoded["records"][:decoded["count"]]]
    assert all(a != b for a, b in zip(firsts, firsts[1:]))


def test_line_locations_match_oracle(decoded):
    count = decoded["count"]
    lines = decoded["tables"][0]
    nl = decoded["olocs"].shape[1]
    for k in range(count):
        d, o = lines[k, :nl], decoded["olocs"][k]
        if not np.array_equal(d, o):
            i = np.flatnonzero(d != o)
            print(f"field {k}: {i.size} line mismatches, first {i[:5]}, max |d| {np.abs(d - o).max():.3e}")
        np.testing.assert_array_equal(d, o)


def test_hsync_refined_locations_match_oracle(decoded):
    count, geom = decoded["count"], decoded["geom"]
    ire0, hz_ire = geom.conv[0], geom.conv[1]
    refined = decoded["tables"][1]
    nl = decoded["olocs"].shape[1]
    for k in range(count):
        o = decoded["ofields"][k]
        shifted = decoded["oshift"][k, :o.window_length]
        ref, _ = O.refine_hsync(shifted, decoded["olocs"][k], np.zeros(nl, dtype=np.uint8), o.threshold, geom.fs_mhz,
                                geom.hsync / geom.fs_mhz, ire0, hz_ire, geom.is_pal)
        d = refined[k, :nl]
        if not np.array_equal(d, ref):
            i = np.flatnonzero(d != ref)
            print(f"field {k}: {i.size} refined mismatches at {i[:8]}, max |d| {np.nanmax(np.abs(d - ref)):.3e}")
        np.testing.assert_array_equal(d, ref)


def test_tbc_output_within_one_lsb_of_oracle(decoded):
    count, geom, fmt = decoded["count"], decoded["geom"], decoded["fmt"]
    render = decoded["tables"][2]
    refined = decoded["tables"][1]
    nl = decoded["olocs"].shape[1]
    worst, over = 0, 0
    total = 0
    for k in range(count):
        o = decoded["ofields"][k]
        shift = 0.0
        if not geom.is_pal:
            # ApplyNtscFscPhaseShiftCore with fsc = TbcFrameSpec.OutputSampleRateHz / 4
            fsc = (float(fmt.sysparams["outfreq"]) * 1e6) / 4e6
            shift = ((117.25 / 360.0) / fsc) * geom.fs_mhz
        np.testing.assert_array_equal(render[k, :nl], refined[k, :nl] - shift if shift else refined[k, :nl])
        first_line = 1 if not geom.is_pal else (3 if o.is_first_field else 4)
        src = decoded["video"][o.window_start:o.window_start + o.window_length]
        _, want = O.tbc_resample(src, render[k, :nl], first_line, geom.field_lines, geom.out_len, _lut(),
                                 nominal=geom.nominal, smoothing=geom.smoothing, conv=geom.conv)
        got = decoded["u16"][k]
        diff = np.abs(got.astype(np.int32) - want.astype(np.int32))
        worst = max(worst, int(diff.max()))
        over += int(np.count_nonzero(diff > 0))
        total += diff.size
    print(f"[vhs sync] {decoded['system']} tbc max |lsb| {worst}, samples off by one {over}/{total}", flush=True)
    assert worst <= 1
