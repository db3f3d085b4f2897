"""The whole-span device field decode (vhsdecode_hip_decode_span_fields_device: pulses,
device line locator, hsync, TBC with no host round trip) against the host-located
path (find_pulses -> field.line_locations_from_pulses -> decode_fields_device) on the
same demodulated span: field count, line-0 positions, parity and every uint16 output
sample identical.  Runs on an MI355X only."""
import numpy as np
import pytest

from vhsdecode_hip import field, filters, formats, hipmem, synth
from vhsdecode_hip.demod import HipBlockDemodulator

pytestmark = pytest.mark.gpu

N = 32768
STRIDE = N - 2 * 1024


@pytest.fixture(scope="module", params=[("NTSC", 3, 11), ("PAL", 2, 12)], ids=["ntsc_vhs", "pal_vhs"])
def span(request):
    system, nfields, seed = request.param
    fmt = formats.get_format(system, "VHS", "sp")
    fl = filters.build_tape_filters(fmt, 40e6, N)
    tape = synth.generate(fmt, fields=nfields, seed=seed, snr_db=30.0, jitter_ns=20.0)
    dem = HipBlockDemodulator(fl, max_blocks=128)
    dem.set_sinc_table(np.load(_lut_path())["downscale_sinc_lut"])
    nb = dem.blocks_for(tape.rf.size)
    n = nb * STRIDE
    rf = hipmem.DeviceBuffer.from_array(tape.rf)
    planes = hipmem.DeviceBuffer(4 * n * 4)
    flags = hipmem.DeviceBuffer(nb * 4)
    video, vlp, env, chroma = (planes.ptr + k * n * 4 for k in range(4))
    dem.demod_span_device(rf.ptr, 0, nb, video, vlp, env, chroma, flags.ptr)
    hipmem.synchronize()
    print(f"\n[{system}] blocks={nb} span={n}", flush=True)
    return dict(system=system, fmt=fmt, dem=dem, video=video, vlp=vlp, n=n, keep=(rf, planes, flags))


def _lut_path():
    import os
    import vhsdecode_hip
    return os.path.join(os.path.dirname(vhsdecode_hip.__file__), "data", "sinc_lut.npz")


def _run_both(span, max_fields):
    dec = field.HipFieldDecoder(span["dem"], span["fmt"])
    g = dec.geom
    per = g.field_lines * g.out_len
    out_a = hipmem.DeviceBuffer(max_fields * per * 2)
    out_b = hipmem.DeviceBuffer(max_fields * per * 2)
    meta = hipmem.DeviceBuffer(4 + 8 * max_fields + 4 * max_fields + 64)
    count_ptr, line0_ptr = meta.ptr, meta.ptr + 8
    first_ptr = line0_ptr + 8 * max_fields
    fields, _ = dec.decode(span["video"], span["vlp"], span["n"], out_a.ptr, max_fields)
    dec.decode_async(span["video"], span["vlp"], span["n"], out_b.ptr, max_fields, count_ptr, line0_ptr, first_ptr)
    hipmem.synchronize()
    count = int(meta.download(np.int32, 1)[0])
    line0 = meta.download(np.float64, max_fields, offset=8)[:count]
    first = meta.download(np.int32, max_fields, offset=8 + 8 * max_fields)[:count]
    a = out_a.download(np.uint16, max_fields * per).reshape(max_fields, per)
    b = out_b.download(np.uint16, max_fields * per).reshape(max_fields, per)
    return fields, count, line0, first, a, b


def test_device_locator_matches_host_path(span):
    fields, count, line0, first, a, b = _run_both(span, 8)
    print(f"[{span['system']}] fields host={len(fields)} device={count}")
    assert count == len(fields) >= 1
    np.testing.assert_array_equal(line0, np.array([f.line0 for f in fields]))
    np.testing.assert_array_equal(first, np.array([1 if f.first_field else 0 for f in fields], dtype=np.int32))
    np.testing.assert_array_equal(a[:count], b[:count])


def test_device_locator_respects_max_fields(span):
    fields, count, line0, first, a, b = _run_both(span, 1)
    assert count == 1 == len(fields)
    np.testing.assert_array_equal(a[:1], b[:1])


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
    interleaving on the device, give the same fields as one decode in order."""
    fmt, n = span["fmt"], span["n"]
    rf, _, _ = span["keep"]
    nb = n // STRIDE
    ref_dec = field.HipFieldDecoder(span["dem"], fmt)
    g = ref_dec.geom
    per = g.field_lines * g.out_len
    pipes = []
    for _ in range(2):
        d = HipBlockDemodulator(span["dem"].filters, max_blocks=nb)
        d.set_sinc_table(np.load(_lut_path())["downscale_sinc_lut"])
        planes = hipmem.DeviceBuffer(4 * n * 4)
        pipes.append(dict(dem=d, dec=field.HipFieldDecoder(d, fmt), planes=planes,
                          p=[planes.ptr + k * n * 4 for k in range(4)], flags=hipmem.DeviceBuffer(nb * 4),
                          out=hipmem.DeviceBuffer(4 * per * 2), meta=hipmem.DeviceBuffer(64), st=hipmem.Stream()))
    for rep in range(3):
        for q in pipes:
            q["dem"].demod_span_device(rf.ptr, 0, nb, *q["p"], q["flags"].ptr, stream=q["st"].handle)
            q["dec"].decode_async(q["p"][0], q["p"][1], n, q["out"].ptr, 4, q["meta"].ptr, stream=q["st"].handle)
    hipmem.synchronize()
    outs = [q["out"].download(np.uint16, 4 * per) for q in pipes]
    counts = [int(q["meta"].download(np.int32, 1)[0]) for q in pipes]
    assert counts[0] == counts[1] >= 1
    np.testing.assert_array_equal(outs[0][:counts[0] * per], outs[1][:counts[0] * per])
    # and equal to the in-order decode of the fixture's planes
    ref = hipmem.DeviceBuffer(4 * per * 2)
    meta = hipmem.DeviceBuffer(64)
    ref_dec.decode_async(span["video"], span["vlp"], n, ref.ptr, 4, meta.ptr)
    hipmem.synchronize()
    np.testing.assert_array_equal(ref.download(np.uint16, 4 * per)[:counts[0] * per], outs[0][:counts[0] * per])
