"""Tape RF-envelope dropout detection on the device (vhsdecode_hip_detect_dropouts_device:
float32 pairwise envelope mean over each field's read window, hysteresis ranges, TBC
rows) against the oracle's restatement of DetectTapeDropouts
(TbcFieldDecodePipeline.cs:6593-6642), whose parts tests/test_dropout_oracle.py pins to
the reference's vectors.  Bit-exact: the same ranges and rows, on the same envelope
plane, read windows and line locations (planted ones, and those of the device's vhs
field sequence).  Runs on an MI355X only."""
import numpy as np
import pytest

from oracle import oracle as O
from vhsdecode_hip import field, filters, formats, hipmem, synth
from vhsdecode_hip.demod import HipBlockDemodulator

pytestmark = pytest.mark.gpu

N = 32768
STRIDE = N - 2 * 1024


LEAD_LINES = 8.0  # the synthetic fields' read windows start this many lines before line 0


def _windows(line0, g, fmt, span_len):
    """Read windows (span coordinates, clipped) placed as DecodeReadWindowPlanner.Resolve
    places an intended start LEAD_LINES before each line 0."""
    req = field.read_sample_count(fmt)
    return [O.field_read_window(l0, LEAD_LINES, g.line_len, req, span_len) for l0 in line0]


def _records(windows, first):
    """Device records (vhsdecode_hip_vhs_field_v1) with the given windows and parities."""
    import ctypes
    from vhsdecode_hip import _lib
    size = ctypes.sizeof(_lib.VhsFieldV1)
    raw = bytearray()
    for (w0, w1), fst in zip(windows, first):
        r = _lib.VhsFieldV1()
        r.struct_size = size
        r.window_start, r.window_length = int(w0), int(w1 - w0)
        r.is_first_field = int(fst)
        raw += bytes(r)
    return hipmem.DeviceBuffer.from_array(np.frombuffer(bytes(raw), dtype=np.uint8).copy())


def _run(dem, env_ptr, span_len, loc_ptr, nloc, windows, first, count, max_fields, params, with_ranges=True):
    cap = params.capacity
    recs = _records(windows, first)
    meta = hipmem.DeviceBuffer.from_array(np.array([count, 0], dtype=np.int32))
    out = hipmem.DeviceBuffer(4 * max_fields * (2 + 5 * cap))
    out.fill(0xF9)  # -1,553,... sentinel in untouched entries
    rc_ptr = out.ptr
    nr_ptr = rc_ptr + 4 * max_fields
    fl_ptr, sx_ptr, ex_ptr, rs_ptr, re_ptr = (nr_ptr + 4 * max_fields + 4 * k * max_fields * cap for k in range(5))
    dem.detect_dropouts_device(env_ptr, span_len, loc_ptr, nloc, recs.ptr, meta.ptr, max_fields, params,
                               rc_ptr, fl_ptr, sx_ptr, ex_ptr, nr_ptr if with_ranges else 0,
                               rs_ptr if with_ranges else 0, re_ptr if with_ranges else 0)
    hipmem.synchronize()
    a = out.download(np.int32, max_fields * (2 + 5 * cap))
    rows = a[:max_fields]
    nr = a[max_fields:2 * max_fields]
    rest = a[2 * max_fields:].reshape(5, max_fields, cap)
    return rows, nr, rest


def _oracle(env, loc, first, window, params):
    lines = params.first_field_lines if first else params.second_field_lines
    return O.tape_dropouts(env, loc, lines, window, fraction=params.threshold_fraction,
                           absolute=params.absolute_threshold, hysteresis=params.hysteresis,
                           merge=params.merge_threshold, min_len=params.minimum_length,
                           out_len=params.output_line_length)


def _compare(got, want, f, cap):
    rows, nr, rest = got
    fl, sx, ex, rs, re = rest
    assert nr[f] == len(want["ranges"])
    k = min(int(nr[f]), cap)
    assert list(zip(rs[f, :k].tolist(), re[f, :k].tolist())) == want["ranges"][:k]
    assert rows[f] == want["field_line"].size
    m = min(int(rows[f]), cap)
    np.testing.assert_array_equal(fl[f, :m], want["field_line"][:m])
    np.testing.assert_array_equal(sx[f, :m], want["start_x"][:m])
    np.testing.assert_array_equal(ex[f, :m], want["end_x"][:m])


@pytest.fixture(scope="module")
def synthetic():
    """Four NTSC fields over a 3.2 M-sample envelope with planted dips: shorter than the
    minimum, inside the merge distance, across line boundaries, several lines long,
    before the first line, and open at the end of the detection range."""
    fmt = formats.get_format("NTSC", "VHS", "sp")
    g = field.sync_geometry(fmt)
    rng = np.random.default_rng(41)
    n = 3_200_000
    env = rng.uniform(800.0, 1200.0, n).astype(np.float32)
    nloc = g.field_lines + 1
    line0 = np.array([300_000.25, 980_000.5, 1_650_000.75, 2_330_000.0])
    first = np.array([1, 0, 1, 0], dtype=np.int32)
    locs = np.stack([l0 + g.line_len * np.arange(nloc) + rng.uniform(-0.4, 0.4, nloc) for l0 in line0])
    dips = []
    for f, l0 in enumerate(line0[:3]):
        base = int(l0)
        for off, ln in [(-50, 30), (5000, 8), (9000, 11), (9030, 40), (20000, 2600), (40000, 9000),
                        (60_000, 12), (60_020, 12), (300_000, 500), (int(g.line_len * 262.5), 4000)]:
            a = base + off + f * 7
            dips.append((a, a + ln))
    for a, b in dips:
        env[a:b] = rng.uniform(0.0, 100.0, b - a).astype(np.float32)
    dem = HipBlockDemodulator(filters.build_tape_filters(fmt, 40e6, N), max_blocks=1)
    env_d = hipmem.DeviceBuffer.from_array(env)
    loc_d = hipmem.DeviceBuffer.from_array(locs)
    print(f"\n[dropouts synthetic] fields=4 dips={len(dips)}", flush=True)
    return dict(fmt=fmt, env=env, locs=locs, line0=line0, first=first, n=n, nloc=nloc, dem=dem, env_d=env_d,
                loc_d=loc_d, windows=_windows(line0, g, fmt, n))


def test_synthetic_fields_match_oracle(synthetic):
    s = synthetic
    p = field.dropout_params(s["fmt"], capacity=256)
    got = _run(s["dem"], s["env_d"].ptr, s["n"], s["loc_d"].ptr, s["nloc"], s["windows"] + [(0, 0)],
               list(s["first"]) + [0], 4, 5, p)
    total = 0
    for f in range(4):
        want = _oracle(s["env"], s["locs"][f], s["first"][f], s["windows"][f], p)
        _compare(got, want, f, p.capacity)
        total += len(want["ranges"])
    # field 4 is beyond the device field count: untouched
    assert got[0][4] == np.frombuffer(b"\xf9" * 4, np.int32)[0]
    print(f"ranges {total}, rows {int(got[0][:4].sum())}")
    assert total >= 18


def test_absolute_threshold_boundaries(synthetic):
    """Samples exactly at the threshold open a range (<=) and exactly at threshold x
    hysteresis close it (>=)."""
    s = synthetic
    env = s["env"].copy()
    base = int(s["line0"][0]) + 100_000
    env[base:base + 40] = 100.0
    env[base + 40] = 125.0
    env[base + 41:base + 60] = 124.99
    env[base + 200:base + 260] = 100.0001
    env_d = hipmem.DeviceBuffer.from_array(env)
    p = field.dropout_params(s["fmt"], absolute=100.0, capacity=256)
    got = _run(s["dem"], env_d.ptr, s["n"], s["loc_d"].ptr, s["nloc"], s["windows"], s["first"], 4, 4, p)
    for f in range(4):
        want = _oracle(env, s["locs"][f], s["first"][f], s["windows"][f], p)
        _compare(got, want, f, p.capacity)
    want0 = _oracle(env, s["locs"][0], 1, s["windows"][0], p)
    assert (base, base + 40) in want0["ranges"]


def test_capacity_overflow_reports_full_count(synthetic):
    s = synthetic
    p = field.dropout_params(s["fmt"], capacity=3)
    got = _run(s["dem"], s["env_d"].ptr, s["n"], s["loc_d"].ptr, s["nloc"], s["windows"], s["first"], 4, 4, p)
    over = 0
    for f in range(4):
        want = _oracle(s["env"], s["locs"][f], s["first"][f], s["windows"][f], p)
        _compare(got, want, f, 3)
        over += int(got[0][f] > 3)
    assert over == 3  # the fourth field has no dips


@pytest.mark.parametrize("span_len", [5000, 12000, 70000, 400_000])
def test_short_spans_clip_the_window(synthetic, span_len):
    """Spans shorter than a read window: the window and the detection range are clipped
    to the span, and the pairwise tree of a short window has its leaves at or above the
    depth the mean kernel splits workgroups at."""
    s = synthetic
    env = s["env"][1_000_000:1_000_000 + span_len].copy()
    rng = np.random.default_rng(span_len)
    for _ in range(6):
        a = int(rng.integers(0, span_len - 100))
        env[a:a + int(rng.integers(5, 90))] = 10.0
    g = field.sync_geometry(s["fmt"])
    line0 = np.array([120.5, 50.25])
    locs = np.stack([l0 + g.line_len * np.arange(s["nloc"]) for l0 in line0])
    env_d = hipmem.DeviceBuffer.from_array(env)
    loc_d = hipmem.DeviceBuffer.from_array(locs)
    p = field.dropout_params(s["fmt"], capacity=256)
    windows = _windows(line0, g, s["fmt"], span_len)
    got = _run(s["dem"], env_d.ptr, span_len, loc_d.ptr, s["nloc"], windows, [1, 0], 2, 2, p)
    for f in range(2):
        want = _oracle(env, locs[f], f == 0, windows[f], p)
        _compare(got, want, f, p.capacity)
    assert got[1][0] >= 1


def test_rows_without_range_outputs(synthetic):
    s = synthetic
    p = field.dropout_params(s["fmt"], capacity=256)
    a = _run(s["dem"], s["env_d"].ptr, s["n"], s["loc_d"].ptr, s["nloc"], s["windows"], s["first"], 4, 4, p)
    b = _run(s["dem"], s["env_d"].ptr, s["n"], s["loc_d"].ptr, s["nloc"], s["windows"], s["first"], 4, 4, p,
             with_ranges=False)
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[2][:3], b[2][:3])


@pytest.fixture(scope="module", params=[("NTSC", 4, 51), ("PAL", 3, 52)], ids=["ntsc_vhs", "pal_vhs"])
def decoded(request):
    """A demodulated span with RF dropouts (stretches of low-level noise) and its device
    field decode: envelope plane, refined line locations, line 0 and parity."""
    system, nfields, seed = request.param
    fmt = formats.get_format(system, "VHS", "sp")
    fl = filters.build_tape_filters(fmt, 40e6, N)
    tape = synth.generate(fmt, fields=nfields, seed=seed, snr_db=30.0, jitter_ns=20.0)
    rf = tape.rf.copy()
    rng = np.random.default_rng(seed)
    drops = []
    for k in range(12):
        at = int(rng.integers(400_000, rf.size - 400_000))
        ln = int(rng.integers(40, 3000))
        rf[at:at + ln] = np.rint(rng.normal(0.0, 40.0, ln)).astype(np.int16)
        drops.append((at, ln))
    dem = HipBlockDemodulator(fl, max_blocks=160)
    dem.set_sinc_table(np.load(_lut_path())["downscale_sinc_lut"])
    nb = dem.blocks_for(rf.size)
    n = nb * STRIDE
    rf_d = hipmem.DeviceBuffer.from_array(rf)
    planes = hipmem.DeviceBuffer(4 * n * 4)
    flags = hipmem.DeviceBuffer(nb * 4)
    video, vlp, env, chroma = (planes.ptr + k * n * 4 for k in range(4))
    dem.demod_span_device(rf_d.ptr, 0, nb, video, vlp, env, chroma, flags.ptr)
    dec = field.HipVhsFieldDecoder(dem, fmt)
    max_fields = 8
    g = dec.geom
    out = hipmem.DeviceBuffer(max_fields * g.field_lines * g.out_len * 2)
    import ctypes
    from vhsdecode_hip import _lib
    recs = hipmem.DeviceBuffer(max_fields * ctypes.sizeof(_lib.VhsFieldV1))
    meta = hipmem.DeviceBuffer(16)
    dec.decode_async(video, vlp, n, out.ptr, max_fields, meta.ptr, recs.ptr)
    hipmem.synchronize()
    count = int(meta.download(np.int32, 1)[0])
    loc_ptr, nloc = dem.vhs_field_lines_device(3)
    raw = recs.download(np.uint8, count * ctypes.sizeof(_lib.VhsFieldV1))
    size = ctypes.sizeof(_lib.VhsFieldV1)
    records = [_lib.VhsFieldV1.from_buffer_copy(raw[k * size:(k + 1) * size].tobytes()) for k in range(count)]
    print(f"\n[dropouts {system}] blocks={nb} records={count} RF dropouts={len(drops)}", flush=True)
    return dict(fmt=fmt, dem=dem, env=env, n=n, count=count, count_ptr=meta.ptr, recs_ptr=recs.ptr, records=records,
                loc_ptr=loc_ptr, nloc=nloc, max_fields=max_fields, meta=meta, drops=drops,
                keep=(rf_d, planes, flags, out, recs))


def _lut_path():
    import os
    import vhsdecode_hip
    return os.path.join(os.path.dirname(vhsdecode_hip.__file__), "data", "sinc_lut.npz")


def test_decoded_span_dropouts_match_oracle(decoded):
    d = decoded
    p = field.dropout_params(d["fmt"], capacity=512)
    mf, cap = d["max_fields"], p.capacity
    out = hipmem.DeviceBuffer(4 * mf * (2 + 5 * cap))
    rc_ptr, nr_ptr = out.ptr, out.ptr + 4 * mf
    ptrs = [nr_ptr + 4 * mf + 4 * k * mf * cap for k in range(5)]
    d["dem"].detect_dropouts_device(d["env"], d["n"], d["loc_ptr"], d["nloc"], d["recs_ptr"], d["count_ptr"], mf, p,
                                    rc_ptr, ptrs[0], ptrs[1], ptrs[2], nr_ptr, ptrs[3], ptrs[4])
    hipmem.synchronize()
    a = out.download(np.int32, mf * (2 + 5 * cap))
    got = (a[:mf], a[mf:2 * mf], a[2 * mf:].reshape(5, mf, cap))
    env = hipmem.download_ptr(d["env"], np.float32, d["n"])
    locs = hipmem.download_ptr(d["loc_ptr"], np.float64, mf * d["nloc"]).reshape(mf, d["nloc"])
    found, spans = [], []
    for f in range(d["count"]):
        r = d["records"][f]
        if r.status != 0:
            assert got[0][f] == 0
            continue
        w = (min(r.window_start, d["n"]), min(r.window_start + r.window_length, d["n"]))
        want = _oracle(env, locs[f], r.is_first_field, w, p)
        _compare(got, want, f, cap)
        found += want["ranges"]
        end_line = min(d["nloc"] - 1, (p.first_field_lines if r.is_first_field else p.second_field_lines) + 1)
        spans.append((max(w[0], int(np.floor(locs[f][0]))), min(w[1], int(np.ceil(locs[f][end_line])))))
    print(f"fields {d['count']}: ranges {len(found)}, rows {int(got[0][:d['count']].sum())}")
    assert d["count"] >= 2
    # the RF dropouts inside a field's detection range are detected (span sample n is RF
    # sample n + 1024)
    inside = covered = 0
    for at, ln in d["drops"]:
        mid = at + ln // 2 - 1024
        if any(a <= mid < b for a, b in spans):
            inside += 1
            covered += any(s <= mid < e for s, e in found)
    print(f"RF dropouts inside detection ranges {inside}, detected {covered}")
    assert inside >= 2 and covered >= 0.75 * inside
