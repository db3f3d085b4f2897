"""Served windows at the serving configuration's size: 16 streams (looped synthetic NTSC VHS
SP, PAL VHS SP and PAL VHS LP tapes in pinned host memory, in turn), full 2-second windows
(80 M samples, plus the 0.12 s preroll the server reads before each), four slots in
flight.  Every window decoded in the mix equals the same window decoded alone, bit for
bit, and one window of each system equals the plain device path on the same RF (the
block demod, a fresh vhs field sequence from the preroll start and the preview's frame
rule).  Runs on an MI355X only."""
import numpy as np
import pytest

from vhsdecode_hip import field, filters, formats, hipmem, synth
from vhsdecode_hip.demod import HipBlockDemodulator
from vhsdecode_hip.serve import HipServer

pytestmark = pytest.mark.gpu

N = 32768
STRIDE = N - 2 * 1024
FS = 40e6
WINDOW = int(2.0 * FS)           # 80 M samples
PREROLL = 4_800_000              # 0.12 s at 40 MSPS (DecodePreviewSegmentProvider.PrerollSeconds)
MAX_FIELDS = 122                 # two seconds of 60 fields/s, and two
STREAMS = 16


def _lut():
    import os
    import vhsdecode_hip
    return np.load(os.path.join(os.path.dirname(vhsdecode_hip.__file__), "data", "sinc_lut.npz"))["downscale_sinc_lut"]


@pytest.fixture(scope="module")
def setup():
    lut = _lut()
    specs = [("NTSC", "VHS", "sp"), ("PAL", "VHS", "sp"), ("PAL", "VHS", "lp")]
    bases = []
    for k, (system, tf, speed) in enumerate(specs):
        fmt = formats.get_format(system, tf, speed)
        fl = filters.build_tape_filters(fmt, FS, N)
        tape = synth.generate(fmt, fields=8, seed=400 + k, snr_db=30.0, jitter_ns=20.0)
        pin = hipmem.PinnedArray(tape.rf.size, np.int16)
        pin.array[:] = tape.rf
        bases.append((fmt, fl, pin))
    rng = np.random.default_rng(44)
    # one window per stream at a random even position past the preroll
    reqs = [(s, 2 * int(rng.integers(PREROLL // 2, 10**9))) for s in range(STREAMS)]
    return dict(lut=lut, bases=bases, reqs=reqs)


def _server(setup, slots):
    srv = HipServer(max_slots=slots, hbm_budget_bytes=64 << 30, max_window_samples=WINDOW + PREROLL + 2,
                    max_fields=MAX_FIELDS)
    fids = [srv.add_format(fmt, fl, setup["lut"]) for fmt, fl, _ in setup["bases"]]
    for s in range(STREAMS):
        fmt, fl, pin = setup["bases"][s % 3]
        srv.add_stream(fids[s % 3], pin.ptr, pin.array.size, loop=True, keep=pin)
    return srv


def _per(setup, s):
    g = field.sync_geometry(setup["bases"][s % 3][0])
    return g.field_lines * g.out_len


@pytest.fixture(scope="module")
def served(setup):
    outs = [hipmem.PinnedArray(MAX_FIELDS * _per(setup, s), np.uint16) for s, _ in setup["reqs"]]
    with _server(setup, 4) as srv:
        tickets = [srv.submit(s, first, WINDOW, MAX_FIELDS, outs[i].ptr) for i, (s, first) in enumerate(setup["reqs"])]
        mixed = [srv.wait(t) for t in tickets]
    lat = np.array([r.latency_ms for r in mixed])
    print(f"\n[config 4] {STREAMS} windows of {WINDOW} samples, 4 slots: fields {[r.field_count for r in mixed]}, "
          f"latency p50 {np.percentile(lat, 50):.1f} ms max {lat.max():.1f} ms", flush=True)
    return dict(outs=outs, mixed=mixed)


def test_every_window_in_the_mix_equals_the_window_alone(setup, served):
    mixed, outs = served["mixed"], served["outs"]
    assert all(r.status == 0 for r in mixed)
    assert all(r.field_count >= 96 for r in mixed)   # 2 s: 100 PAL or 120 NTSC fields, whole frames
    with _server(setup, 1) as srv:
        for i, (s, first) in enumerate(setup["reqs"]):
            alone = hipmem.PinnedArray(MAX_FIELDS * _per(setup, s), np.uint16)
            r = srv.wait(srv.submit(s, first, WINDOW, MAX_FIELDS, alone.ptr))
            assert r.status == 0 and r.field_count == mixed[i].field_count, i
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


@pytest.mark.parametrize("stream", [0, 1])
def test_window_equals_plain_device_path(setup, served, stream):
    s, first = setup["reqs"][stream]
    fmt, fl, pin = setup["bases"][s % 3]
    base = pin.array
    start = max(0, first - PREROLL) & ~1
    count = first - start + WINDOW
    idx = (start + np.arange(count)) % base.size
    rf = np.ascontiguousarray(base[idx])
    per = _per(setup, s)
    dem = HipBlockDemodulator(fl, max_blocks=(count - N) // STRIDE + 1, sinc_table=setup["lut"])
    nb = dem.blocks_for(rf.size)
    n = nb * STRIDE
    d_rf = hipmem.DeviceBuffer.from_array(rf)
    planes = hipmem.DeviceBuffer(4 * n * 4)
    flags = hipmem.DeviceBuffer(nb * 4)
    video, vlp, env, chroma = (planes.ptr + k * n * 4 for k in range(4))
    dem.demod_span_device(d_rf.ptr, 0, nb, video, vlp, env, chroma, flags.ptr)
    dec = field.HipVhsFieldDecoder(dem, fmt)
    fps = float(fmt.sysparams["FPS"])
    budget = MAX_FIELDS + int(np.ceil(0.12 * fps * 2.0)) + 8
    res = dec.decode_span(video, vlp, n, max_fields=budget, rounds=4)
    half_frame = int(np.floor(FS / (fps * 2.0) + 0.5))
    kept = _kept(res.records[:budget], start + 1024, first, half_frame, MAX_FIELDS)
    r = served["mixed"][stream]
    st = [int(x.status) for x in res.records]
    print(f"[config 4 stream {stream}] records {len(res.records)} (statuses {[(v, st.count(v)) for v in sorted(set(st))]}), "
          f"kept {len(kept)}, served {r.field_count}", flush=True)
    assert r.field_count == len(kept)
    out = served["outs"][stream].array
    bad = [(j, k, int(np.count_nonzero(out[j * per:(j + 1) * per] != res.outputs[k]))) for j, k in enumerate(kept)]
    bad = [b for b in bad if b[2]]
    assert not bad, f"fields differing (output slot, record, samples): {bad[:10]} of {len(kept)}"
