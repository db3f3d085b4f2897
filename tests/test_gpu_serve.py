"""Served windows (vhsdecode_hip_serve_*): many windows from several looped streams in
flight at once, each result identical to the same window decoded alone, and to the
plain device path on the same RF (a fresh vhs field sequence from 0.12 s before the
window, then PreviewFrameAssembler's frame rule, restated here).  Runs on an MI355X only."""
import numpy as np
import pytest

from vhsdecode_hip import field, filters, formats, hipmem, synth
from vhsdecode_hip.demod import HipBlockDemodulator
from vhsdecode_hip.serve import HipServer

pytestmark = pytest.mark.gpu

N = 32768
STRIDE = N - 2 * 1024
WINDOW = N + 100 * STRIDE  # 101 blocks, about 3.1 M samples (3-4 fields)
PREROLL = 4_800_000        # 0.12 s at 40 MSPS (DecodePreviewSegmentProvider.PrerollSeconds)
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
    srv = HipServer(max_slots=slots, hbm_budget_bytes=32 << 30, max_window_samples=WINDOW + PREROLL + 2,
                    max_fields=MAX_FIELDS)
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


def _kept(records, origin, target, half_frame, max_out):
    """PreviewFrameAssembler.Accept (src/VHSDecode.Preview/PreviewFrameAssembler.cs:55-95):
    record indices of the fields of the kept frames."""
    out, first, first_start = [], None, 0
    for k, r in enumerate(records):
        if len(out) + 2 > max_out:
            break
        if r.status != field.DECODED:
            continue
        start = origin + r.window_start
        if r.is_first_field:
            first, first_start = k, start
            continue
        if first is None:
            continue
        mid = max(first_start, start) + half_frame
        a, first = first, None
        if mid >= target:
            out += [a, k]
    return out


@pytest.mark.parametrize("req", [3, 8])
def test_served_window_equals_plain_device_path(setup, req):
    s, first = setup["reqs"][req]
    fmt, fl, pin = setup["bases"][s // 3]
    base = pin.array
    start = max(0, first - PREROLL) & ~1
    count = first - start + WINDOW
    idx = (start + np.arange(count)) % base.size
    rf = np.ascontiguousarray(base[idx])
    per = _per(setup, s)
    out = hipmem.PinnedArray(MAX_FIELDS * per, np.uint16)
    with _server(setup, 2) as srv:
        r = srv.wait(srv.submit(s, first, WINDOW, MAX_FIELDS, out.ptr))
    dem = HipBlockDemodulator(fl, max_blocks=300, sinc_table=setup["lut"])
    nb = dem.blocks_for(rf.size)
    n = nb * STRIDE
    d_rf = hipmem.DeviceBuffer.from_array(rf)
    planes = hipmem.DeviceBuffer(4 * n * 4)
    flags = hipmem.DeviceBuffer(nb * 4)
    video, vlp, env, chroma = (planes.ptr + k * n * 4 for k in range(4))
    dem.demod_span_device(d_rf.ptr, 0, nb, video, vlp, env, chroma, flags.ptr)
    dec = field.HipVhsFieldDecoder(dem, fmt)
    # the server's field budget: output fields, the preroll's fields and 8
    fps = float(fmt.sysparams["FPS"])
    budget = MAX_FIELDS + int(np.ceil(0.12 * fps * 2.0)) + 8
    res = dec.decode_span(video, vlp, n, max_fields=budget, rounds=4)
    half_frame = int(np.floor(40e6 / (fps * 2.0) + 0.5))
    kept = _kept(res.records[:budget], start + 1024, first, half_frame, MAX_FIELDS)
    print(f"[serve {req}] records {len(res.records)} kept {kept} served {r.field_count}", flush=True)
    assert r.status == 0 and r.field_count == len(kept) >= 2
    for j, k in enumerate(kept):
        np.testing.assert_array_equal(out.array[j * per:(j + 1) * per], res.outputs[k])


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
