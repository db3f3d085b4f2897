"""Served windows (vhsdecode_hip_serve_*): many windows from several looped streams in
flight at once, each result identical to the same window decoded alone, and to the
plain device path on the same RF.  Runs on an MI355X only."""
import numpy as np
import pytest

from vhsdecode_hip import field, filters, formats, hipmem, synth
from vhsdecode_hip.demod import HipBlockDemodulator
from vhsdecode_hip.serve import HipServer

pytestmark = pytest.mark.gpu

N = 32768
STRIDE = N - 2 * 1024
WINDOW = N + 100 * STRIDE  # 101 blocks, about 3.1 M samples (3-4 fields)
MAX_FIELDS = 4


def _lut():
    import os
    import vhsdecode_hip
    return np.load(os.path.join(os.path.dirname(vhsdecode_hip.__file__), "data", "sinc_lut.npz"))["downscale_sinc_lut"]


@pytest.fixture(scope="module")
def setup():
    lut = _lut()
    specs = [("NTSC", "VHS", "sp", 21), ("PAL", "VHS", "lp", 22)]
    bases = []
    for system, tf, speed, seed in specs:
        fmt = formats.get_format(system, tf, speed)
        fl = filters.build_tape_filters(fmt, 40e6, N)
        tape = synth.generate(fmt, fields=4, seed=seed, snr_db=30.0, jitter_ns=20.0)
        pin = hipmem.PinnedArray(tape.rf.size, np.int16)
        pin.array[:] = tape.rf
        bases.append((fmt, fl, pin))
    rng = np.random.default_rng(5)
    # (stream index, first sample) per request; three streams per format
    reqs = [(int(rng.integers(0, 6)), 2 * int(rng.integers(0, 2_000_000))) for _ in range(24)]
    return dict(lut=lut, bases=bases, reqs=reqs)


def _server(setup, slots):
    srv = HipServer(max_slots=slots, hbm_budget_bytes=32 << 30, max_window_samples=WINDOW, max_fields=MAX_FIELDS)
    fids = [srv.add_format(fmt, fl, setup["lut"]) for fmt, fl, _ in setup["bases"]]
    for s in range(6):
        fmt, fl, pin = setup["bases"][s // 3]
        srv.add_stream(fids[s // 3], pin.ptr, pin.array.size, loop=True, keep=pin)
    return srv


def _per(setup, s):
    g = field.sync_geometry(setup["bases"][s // 3][0])
    return g.field_lines * g.out_len


def test_windows_in_the_mix_equal_windows_alone(setup):
    outs = [hipmem.PinnedArray(MAX_FIELDS * _per(setup, s), np.uint16) for s, _ in setup["reqs"]]
    with _server(setup, 4) as srv:
        tickets = [srv.submit(s, first, WINDOW, MAX_FIELDS, outs[i].ptr) for i, (s, first) in enumerate(setup["reqs"])]
        mixed = [srv.wait(t) for t in tickets]
    assert all(r.status == 0 for r in mixed)
    assert all(r.field_count >= 1 for r in mixed)
    lat = np.array([r.latency_ms for r in mixed])
    print(f"\n24 windows, 4 slots: latency p50 {np.percentile(lat, 50):.1f} ms p99 {np.percentile(lat, 99):.1f} ms")
    with _server(setup, 1) as srv:
        for i in (0, 5, 11, 17, 23):
            s, first = setup["reqs"][i]
            alone = hipmem.PinnedArray(MAX_FIELDS * _per(setup, s), np.uint16)
            r = srv.wait(srv.submit(s, first, WINDOW, MAX_FIELDS, alone.ptr))
            assert r.status == 0 and r.field_count == mixed[i].field_count
            n = r.field_count * _per(setup, s)
            np.testing.assert_array_equal(alone.array[:n], outs[i].array[:n])


def test_served_window_equals_plain_device_path(setup):
    s, first = setup["reqs"][3]
    fmt, fl, pin = setup["bases"][s // 3]
    base = pin.array
    idx = (first + np.arange(WINDOW)) % base.size
    rf = np.ascontiguousarray(base[idx])
    per = _per(setup, s)
    out = hipmem.PinnedArray(MAX_FIELDS * per, np.uint16)
    with _server(setup, 2) as srv:
        r = srv.wait(srv.submit(s, first, WINDOW, MAX_FIELDS, out.ptr))
    dem = HipBlockDemodulator(fl, max_blocks=128, sinc_table=setup["lut"])
    nb = dem.blocks_for(rf.size)
    n = nb * STRIDE
    d_rf = hipmem.DeviceBuffer.from_array(rf)
    planes = hipmem.DeviceBuffer(4 * n * 4)
    flags = hipmem.DeviceBuffer(nb * 4)
    video, vlp, env, chroma = (planes.ptr + k * n * 4 for k in range(4))
    dout = hipmem.DeviceBuffer(MAX_FIELDS * per * 2)
    meta = hipmem.DeviceBuffer(64)
    dem.demod_span_device(d_rf.ptr, 0, nb, video, vlp, env, chroma, flags.ptr)
    field.HipFieldDecoder(dem, fmt).decode_async(video, vlp, n, dout.ptr, MAX_FIELDS, meta.ptr)
    hipmem.synchronize()
    count = int(meta.download(np.int32, 1)[0])
    assert r.status == 0 and r.field_count == count >= 1
    np.testing.assert_array_equal(out.array[:count * per], dout.download(np.uint16, count * per))


def test_pageable_output_matches_pinned(setup):
    s, first = setup["reqs"][7]
    per = _per(setup, s)
    pinned = hipmem.PinnedArray(MAX_FIELDS * per, np.uint16)
    plain = np.zeros(MAX_FIELDS * per, np.uint16)
    with _server(setup, 2) as srv:
        a = srv.wait(srv.submit(s, first, WINDOW, MAX_FIELDS, pinned.ptr))
        b = srv.wait(srv.submit(s, first, WINDOW, MAX_FIELDS, plain.ctypes.data))
    assert a.status == b.status == 0 and a.field_count == b.field_count
    n = a.field_count * per
    np.testing.assert_array_equal(pinned.array[:n], plain[:n])
