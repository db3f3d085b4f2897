"""The device field sequence on whole demodulated spans, run the ways a caller can run it:
one call against calls of a few records each (the sequence state carried between calls
on one span), against the same span entered at a later block-grid offset (state carried
across spans with `state_shift`), and pipelines on separate HIP streams; every record
and every uint16 output sample identical.  Also the chroma plane computed on its own
stream.  Runs on an MI355X only."""
import ctypes

import numpy as np
import pytest

from vhsdecode_hip import _lib, field, filters, formats, hipmem, synth
from vhsdecode_hip.demod import HipBlockDemodulator

pytestmark = pytest.mark.gpu

N = 32768
STRIDE = N - 2 * 1024
GRID = 491520  # lcm(32768, 30720): spans entered here keep every read window
SIZE = ctypes.sizeof(_lib.VhsFieldV1)
KEYS = ("begin", "window_start", "window_length", "status", "level_source", "is_first_field", "sync_level",
        "blank_level", "threshold", "line0", "first_hsync", "next_begin")


@pytest.fixture(scope="module", params=[("NTSC", 8, 11), ("PAL", 6, 12)], ids=["ntsc_vhs", "pal_vhs"])
def span(request):
    system, nfields, seed = request.param
    fmt = formats.get_format(system, "VHS", "sp")
    fl = filters.build_tape_filters(fmt, 40e6, N)
    tape = synth.generate(fmt, fields=nfields, seed=seed, snr_db=30.0, jitter_ns=20.0)
    dem = HipBlockDemodulator(fl, max_blocks=256)
    dem.set_sinc_table(np.load(_lut_path())["downscale_sinc_lut"])
    nb = dem.blocks_for(tape.rf.size)
    n = nb * STRIDE
    rf = hipmem.DeviceBuffer.from_array(tape.rf)
    planes = hipmem.DeviceBuffer(4 * n * 4)
    flags = hipmem.DeviceBuffer(nb * 4)
    video, vlp, env, chroma = (planes.ptr + k * n * 4 for k in range(4))
    dem.demod_span_device(rf.ptr, 0, nb, video, vlp, env, chroma, flags.ptr)
    hipmem.synchronize()
    dec = field.HipVhsFieldDecoder(dem, fmt)
    whole = dec.decode_span(video, vlp, n, max_fields=32)
    print(f"\n[{system}] blocks={nb} span={n} records={len(whole.records)} "
          f"statuses={[r.status for r in whole.records]}", flush=True)
    return dict(system=system, fmt=fmt, dem=dem, dec=dec, video=video, vlp=vlp, n=n, whole=whole,
                keep=(rf, planes, flags))


def _lut_path():
    import os
    import vhsdecode_hip
    return os.path.join(os.path.dirname(vhsdecode_hip.__file__), "data", "sinc_lut.npz")


def _same(a, b, shift=0):
    """Records equal, positions compared in the coordinates of `a` (b's span starts `shift`
    samples later)."""
    assert len(a.records) == len(b.records)
    for ra, rb in zip(a.records, b.records):
        for key in KEYS:
            va, vb = getattr(ra, key), getattr(rb, key)
            if key in ("begin", "window_start", "next_begin"):
                vb += shift
            assert va == vb, (key, va, vb)
    assert sorted(a.outputs) == sorted(b.outputs)
    for k in a.outputs:
        np.testing.assert_array_equal(a.outputs[k], b.outputs[k])


def test_whole_span_decodes_every_field(span):
    recs = span["whole"].records
    assert sum(r.status == field.DECODED for r in recs) >= (6 if span["system"] == "NTSC" else 4)
    firsts = [r.is_first_field for r in recs if r.status == field.DECODED]
    assert all(x != y for x, y in zip(firsts, firsts[1:]))


@pytest.mark.parametrize("per_call", [1, 2, 3])
def test_calls_of_a_few_records_equal_one_call(span, per_call):
    got = span["dec"].decode_span(span["video"], span["vlp"], span["n"], max_fields=per_call)
    _same(span["whole"], got)


@pytest.mark.parametrize("rounds", [1, 2])
def test_fewer_rounds_per_call_give_the_same_sequence(span, rounds):
    got = span["dec"].decode_span(span["video"], span["vlp"], span["n"], max_fields=32, rounds=rounds)
    _same(span["whole"], got)


def _one_call(dec, video, vlp, n, max_fields, first_begin=0, reset=True, shift=0, stream=0):
    """One decode call: (SpanDecode of its records, next begin)."""
    per = dec.geom.field_lines * dec.geom.out_len
    out = hipmem.DeviceBuffer(max_fields * per * 2)
    recs = hipmem.DeviceBuffer(max_fields * SIZE)
    meta = hipmem.DeviceBuffer(16)
    dec.decode_async(video, vlp, n, out.ptr, max_fields, meta.ptr, recs.ptr, meta.ptr + 8, first_begin, reset,
                     shift, stream=stream)
    hipmem.synchronize()
    count = int(meta.download(np.int32, 1)[0])
    nxt = int(meta.download(np.int64, 1, offset=8)[0])
    raw = recs.download(np.uint8, max(count, 1) * SIZE)
    records = [_lib.VhsFieldV1.from_buffer_copy(raw[k * SIZE:(k + 1) * SIZE].tobytes()) for k in range(count)]
    outs = out.download(np.uint16, max(count, 1) * per).reshape(-1, per)
    outputs = {k: outs[k].copy() for k in range(count) if records[k].status == field.DECODED}
    return field.SpanDecode(records, outputs, nxt, None), nxt


def test_replayed_launch_graph_matches_launches(span):
    """A call that repeats the previous call's launch arguments replays them as a captured
    HIP graph.  The launches made one by one (profiling on) give the reference; then the
    same call several times on fixed buffers (captured, then replayed), each followed by
    eager work on the same context and a synchronisation, gives the same records and
    fields every time."""
    dem, dec = span["dem"], span["dec"]
    per = dec.geom.field_lines * dec.geom.out_len
    mf = 32
    out = hipmem.DeviceBuffer(mf * per * 2)
    recs = hipmem.DeviceBuffer(mf * SIZE)
    meta = hipmem.DeviceBuffer(16)

    def call():
        out.zero()
        recs.zero()
        meta.zero()
        hipmem.synchronize()
        dec.decode_async(span["video"], span["vlp"], span["n"], out.ptr, mf, meta.ptr, recs.ptr, meta.ptr + 8)
        hipmem.synchronize()
        count = int(meta.download(np.int32, 1)[0])
        return (count, int(meta.download(np.int64, 1, offset=8)[0]), recs.download(np.uint8, (count + 1) * SIZE),
                out.download(np.uint16, max(count, 1) * per))

    dem.set_profiling(True)
    want = call()
    dem.take_profile()
    dem.set_profiling(False)
    assert want[0] >= 4
    for rep in range(4):
        got = call()
        assert got[:2] == want[:2], rep
        np.testing.assert_array_equal(got[2], want[2])
        np.testing.assert_array_equal(got[3], want[3])
        # eager work on the context between replays: a profiled call with other buffers
        dem.set_profiling(True)
        _one_call(dec, span["video"], span["vlp"], span["n"], 2)
        dem.take_profile()
        dem.set_profiling(False)
    print(f"[{span['system']}] graph replays equal the launches: {want[0]} records", flush=True)


def test_sequence_carried_into_a_later_span(span):
    """Two records decoded on the span, then the rest on the same planes entered at a
    block-grid multiple before the next read window (a later span): the level, parity and
    sync state carried with state_shift give the same sequence."""
    dec = span["dec"]
    part, nxt = _one_call(dec, span["video"], span["vlp"], span["n"], 2)
    assert len(part.records) == 2
    next_window = (nxt - 1024) // 32768 * 32768 // 30720 * 30720
    shift = next_window // GRID * GRID
    assert shift > 0
    rest = dec.decode_span(span["video"] + 4 * shift, span["vlp"] + 4 * shift, span["n"] - shift, max_fields=32,
                           first_begin=nxt - shift, reset=False, state_shift=shift)
    for r in rest.records:  # back to the first span's coordinates
        r.begin += shift
        r.window_start += shift
        r.next_begin += shift
    joined = field.SpanDecode(part.records + rest.records,
                              {**part.outputs, **{k + 2: v for k, v in rest.outputs.items()}}, 0, None)
    _same(span["whole"], joined)


def test_chroma_on_its_own_stream_is_identical(span):
    """span_v1.chroma_stream: the chroma plane computed on another stream (overlapping
    the rest of the chain) equals the one computed in order; the context's next call
    waits for it before reusing its descriptors."""
    dem, n = span["dem"], span["n"]
    rf, planes, flags = span["keep"]
    nb = n // STRIDE
    base = planes.download(np.float32, 4 * n).reshape(4, n)
    side = hipmem.Stream()
    other = hipmem.DeviceBuffer(4 * n * 4)
    other.fill(0xFF)
    p = [other.ptr + k * n * 4 for k in range(4)]
    for _ in range(2):
        dem.demod_span_device(rf.ptr, 0, nb, p[0], p[1], p[2], p[3], flags.ptr, chroma_stream=side.handle)
    hipmem.synchronize()
    got = other.download(np.float32, 4 * n).reshape(4, n)
    np.testing.assert_array_equal(got, base)


def test_pipelines_on_separate_streams_match(span):
    """Spans in flight: two contexts decoding on their own streams, their kernels
    interleaving on the device, give the same records and fields as one decode in order."""
    fmt, n = span["fmt"], span["n"]
    rf, _, _ = span["keep"]
    nb = n // STRIDE
    pipes = []
    for _ in range(2):
        d = HipBlockDemodulator(span["dem"].filters, max_blocks=nb)
        d.set_sinc_table(np.load(_lut_path())["downscale_sinc_lut"])
        planes = hipmem.DeviceBuffer(4 * n * 4)
        pipes.append(dict(dem=d, dec=field.HipVhsFieldDecoder(d, fmt), planes=planes,
                          p=[planes.ptr + k * n * 4 for k in range(4)], flags=hipmem.DeviceBuffer(nb * 4),
                          st=hipmem.Stream()))
    per = pipes[0]["dec"].geom.field_lines * pipes[0]["dec"].geom.out_len
    mf = 32
    for q in pipes:
        q.update(out=hipmem.DeviceBuffer(mf * per * 2), recs=hipmem.DeviceBuffer(mf * SIZE),
                 meta=hipmem.DeviceBuffer(16))
    for rep in range(3):
        for q in pipes:
            q["dem"].demod_span_device(rf.ptr, 0, nb, *q["p"], q["flags"].ptr, stream=q["st"].handle)
            q["dec"].decode_async(q["p"][0], q["p"][1], n, q["out"].ptr, mf, q["meta"].ptr, q["recs"].ptr,
                                  rounds=16, stream=q["st"].handle)
    hipmem.synchronize()
    whole = span["whole"]
    count = len(whole.records)
    for q in pipes:
        assert int(q["meta"].download(np.int32, 1)[0]) == count
        raw = q["recs"].download(np.uint8, count * SIZE)
        outs = q["out"].download(np.uint16, count * per).reshape(count, per)
        got = field.SpanDecode(
            [_lib.VhsFieldV1.from_buffer_copy(raw[k * SIZE:(k + 1) * SIZE].tobytes()) for k in range(count)],
            {k: outs[k] for k in whole.outputs}, 0, None)
        _same(whole, got)
