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


@pytest.fixture(scope="module", params=[("NTSC", 5, 31, 28.0, None), ("PAL", 4, 32, 30.0, None),
                                        ("NTSC", 4, 41, 34.0, None), ("PAL", 5, 42, 26.0, None),
                                        ("NTSC", 9, 43, 30.0, (0.42, 0.58)), ("PAL", 8, 44, 30.0, (0.40, 0.52))],
                ids=["ntsc_a", "pal_a", "ntsc_b", "pal_b", "ntsc_gap", "pal_gap"])
def decoded(request):
    """Synthetic tapes; the `gap` ones lose their RF over a stretch of a field or more
    (zero samples), so the sequence fails fields there and recovers with the reference's
    offsets."""
    system, nf, seed, snr, gap = request.param
    print(f"\n[vhs sync] {system} fields={nf} seed={seed} snr={snr} gap={gap}", flush=True)
    fmt = formats.get_format(system, "VHS", "sp")
    tape = synth.generate(fmt, fields=nf, seed=seed, snr_db=snr, jitter_ns=20.0)
    if gap is not None:
        tape.rf[int(gap[0] * tape.rf.size):int(gap[1] * tape.rf.size)] = 0
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
    dec = field.HipVhsFieldDecoder(dem, fmt)
    geom = dec.geom
    # calls of MAX_FIELDS records until the next read window leaves the span
    res = dec.decode_span(video, vlp, n, max_fields=MAX_FIELDS, want_tables=True)
    records = res.records
    count = len(records)
    tables = [np.concatenate([t[w][:len(t[w])] for t in res.tables])[:count] for w in range(4)]
    calls = len(res.tables)
    planes_h = planes.download(np.float32, 4 * n).reshape(4, n)
    vlp64 = planes_h[1].astype(np.float64) + dem.bias
    cfg = O.sync_config(fmt, serration_mode=O.SERRATION_NONE)
    ofields, olocs, oshift = O.vhs_field_sequence(vlp64, cfg, 0, 64, want_shifted=True)
    print(f"[vhs sync] {system} device fields={count} oracle fields={len(ofields)} "
          f"statuses={[f.status for f in ofields]} calls={calls}", flush=True)
    return dict(system=system, fmt=fmt, nf=nf, geom=geom, dem=dem, count=count, next_begin=res.next_begin,
                calls=calls, records=records, tables=tables, video=planes_h[0].astype(np.float64) + dem.bias,
                outputs=res.outputs, ofields=ofields, olocs=olocs, oshift=oshift, keep=(rf, planes, flags))


def test_field_records_match_oracle(decoded):
    count, records, ofields = decoded["count"], decoded["records"], decoded["ofields"]
    decoded_n = sum(r.status == field.DECODED for r in records)
    assert count == len(ofields) and decoded_n >= decoded["nf"] - 2
    bad = []
    for k in range(count):
        d, o = records[k], ofields[k]
        if d.status != o.status:
            bad.append((k, "status", d.status, o.status))
            continue
        names = RECORD_FIELDS if d.status == field.DECODED else ("begin", "window_start", "window_length",
                                                                   "next_offset", "next_begin")
        for name in names:
            if getattr(d, name) != getattr(o, name):
                bad.append((k, name, getattr(d, name), getattr(o, name)))
    for b in bad:
        print("record mismatch", b, flush=True)
    assert not bad
    # the sequence continues where the last record says
    assert decoded["next_begin"] == records[count - 1].next_begin


def test_parity_alternates(decoded):
    """Consecutive decoded fields alternate parity (a failed field between two breaks the
    cadence, as the field it lost)."""
    recs = decoded["records"]
    pairs = [(a, b) for a, b in zip(recs, recs[1:]) if a.status == b.status == field.DECODED]
    assert pairs and all(a.is_first_field != b.is_first_field for a, b in pairs)


def test_line_locations_match_oracle(decoded):
    count = decoded["count"]
    lines = decoded["tables"][0]
    nl = decoded["olocs"].shape[1]
    for k in range(count):
        if decoded["records"][k].status != field.DECODED:
            continue
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
        if o.status != 0:
            continue
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
        if o.status != 0:
            continue
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
        got = decoded["outputs"][k]
        # the dropout-detection table: render locations from the output first line, span coordinates
        span_tab = decoded["tables"][3][k]
        m = nl - first_line
        np.testing.assert_array_equal(span_tab[:m], render[k, first_line:nl] + o.window_start)
        diff = np.abs(got.astype(np.int32) - want.astype(np.int32))
        worst = max(worst, int(diff.max()))
        over += int(np.count_nonzero(diff > 0))
        total += diff.size
    print(f"[vhs sync] {decoded['system']} tbc max |lsb| {worst}, samples off by one {over}/{total}", flush=True)
    assert worst <= 1


# ---------------------------------------------------------------- full chain against the FP64 chain
def _oracle_planes(rf, fl, nb):
    """The FP64 oracle's span planes (video, sync reference) block by block."""
    from concurrent.futures import ThreadPoolExecutor

    def one(b):
        r = O.demod_block(rf[b * STRIDE:b * STRIDE + N].astype(np.float64), fl)
        return b, r.video[CUT:CUT + STRIDE], r.video_lowpass[CUT:CUT + STRIDE]

    video = np.empty(nb * STRIDE)
    vlp = np.empty(nb * STRIDE)
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        for b, v, s in ex.map(one, range(nb)):
            video[b * STRIDE:(b + 1) * STRIDE] = v
            vlp[b * STRIDE:(b + 1) * STRIDE] = s
    return video, vlp


@pytest.fixture(scope="module", params=[("NTSC", 4, 61, 28.0), ("PAL", 4, 62, 30.0)], ids=["ntsc", "pal"])
def chain(request):
    system, nf, seed, snr = request.param
    fmt = formats.get_format(system, "VHS", "sp")
    tape = synth.generate(fmt, fields=nf, seed=seed, snr_db=snr, jitter_ns=20.0)
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
    dec = field.HipVhsFieldDecoder(dem, fmt)
    got = dec.decode_span(video, vlp, n, max_fields=16, want_tables=True)
    print(f"\n[chain] {system} device records {len(got.records)}; FP64 oracle demod of {nb} blocks", flush=True)
    ovideo, ovlp = _oracle_planes(tape.rf, fl, nb)
    ofields, olocs, oshift = O.vhs_field_sequence(ovlp, O.sync_config(fmt, serration_mode=O.SERRATION_NONE), 0, 16,
                                                  want_shifted=True)
    return dict(system=system, fmt=fmt, dec=dec, got=got, ovideo=ovideo, ofields=ofields, olocs=olocs,
                oshift=oshift, keep=(rf, planes, flags))


def test_full_chain_fields_against_fp64_chain(chain):
    """The FP32 device chain (demod planes, then the field sequence) against the FP64 oracle
    chain on the same RF: the same read windows, parity and pulse counts, and TBC luma
    within +-1 LSB except where a sub-1e-3-sample difference in a line position moves the
    FP32 source coordinate SampleSinc reads (TbcLineResampler.SampleSinc) onto the next
    FP32 value on a steep edge (bound: fewer than 2e-4 of the samples)."""
    got, ofields, geom, fmt = chain["got"], chain["ofields"], chain["dec"].geom, chain["fmt"]
    recs = [r for r in got.records]
    assert len(recs) == len(ofields) >= 3
    worst, over, total, dl0 = 0, 0, 0, 0.0
    for k, (d, o) in enumerate(zip(recs, ofields)):
        assert (d.status, d.window_start, d.window_length, d.is_first_field) == \
            (o.status, o.window_start, o.window_length, o.is_first_field)
        assert (d.raw_pulses, d.refined_pulses) == (o.raw_pulses, o.refined_pulses)
        dl0 = max(dl0, abs(d.line0 - o.line0))
        if d.status != field.DECODED:
            continue
        nl = chain["olocs"].shape[1]
        ref, _ = O.refine_hsync(chain["oshift"][k, :o.window_length], chain["olocs"][k], np.zeros(nl, np.uint8),
                                o.threshold, geom.fs_mhz, geom.hsync / geom.fs_mhz, geom.conv[0], geom.conv[1],
                                geom.is_pal)
        if not geom.is_pal:
            ref = ref - ((117.25 / 360.0) / ((float(fmt.sysparams["outfreq"]) * 1e6) / 4e6)) * geom.fs_mhz
        first_line = 1 if not geom.is_pal else (3 if o.is_first_field else 4)
        src = chain["ovideo"][o.window_start:o.window_start + o.window_length]
        _, want = O.tbc_resample(src, ref, first_line, geom.field_lines, geom.out_len, _lut(), nominal=geom.nominal,
                                 smoothing=geom.smoothing, conv=geom.conv)
        diff = np.abs(got.outputs[k].astype(np.int32) - want.astype(np.int32))
        worst = max(worst, int(diff.max()))
        over += int(np.count_nonzero(diff > 1))
        total += diff.size
    print(f"[chain] {chain['system']}: line0 max |d| {dl0:.2e} samples; TBC max {worst} LSB, "
          f"beyond 1 LSB {over}/{total} ({over / max(total, 1):.2e})", flush=True)
    assert dl0 < 0.05
    assert over / total < 2e-4
