"""Oracle restatement of the tape RF-envelope dropout detection
(TbcFieldDecodePipeline.DetectTapeDropouts, Decode/TbcFieldDecodePipeline.cs:6593-6642)
pinned on the CPU: the reference's own vectors for RfDropoutDetector.FindDropouts and
TbcDropoutMapper.MapTapeRfToTbc (CompatibilityTests.cs:16320-16388), numpy.sum for the
float32 pairwise mean, an independent formulation of the hysteresis walk on random
envelopes, and the geometric meaning of the TBC rows."""
import numpy as np
import pytest

from oracle import oracle as O


def test_find_dropouts_reference_vector():
    """CompatibilityTests.cs:16323-16350: lows at 10..22, 25..35 and 60..65 of an 80-sample
    envelope of 10s; threshold 2, hysteresis 2, merge 5, minimum 10 -> one range 10..23
    (the second dip is inside the merge distance of nothing: it starts 2 samples after
    the first closed and is ignored; the third is too short)."""
    env = np.full(80, 10.0)
    env[10:23] = 1.0
    env[25:36] = 1.0
    env[60:66] = 1.0
    assert O.find_dropouts(env, 0, 80, 2.0, 2.0, merge=5, min_len=10) == [(10, 23)]


def test_map_tape_rf_to_tbc_reference_vector():
    """CompatibilityTests.cs:16378-16388: range 10..20 over lines at 0, 20, 40 with 100
    pixels per line -> rows (0, 50, 100) and (1, 0, 0)."""
    fl, sx, ex = O.map_tape_rf_to_tbc([(10, 20)], [0.0, 20.0, 40.0], 100, 0, 2, 0)
    assert fl.tolist() == [0, 1]
    assert sx.tolist() == [50, 0]
    assert ex.tolist() == [100, 0]


@pytest.mark.parametrize("n", list(range(0, 260)) + [511, 1000, 1023, 4097, 8000, 8191, 8192])
def test_pairwise_sum_matches_numpy(n):
    """numpy.sum of a float32 array is one pairwise call per 8192-value buffer; up to that
    length the restatement is numpy's summation bit for bit.  (At 32768 values it is pinned
    by the Betamax envelope mean inside the fixture hashes of tests/test_oracle.py.)"""
    x = (np.random.default_rng(n).standard_normal(n) * 1000.0).astype(np.float32)
    got = O.pairwise_sum_f32(x)
    want = np.sum(x, dtype=np.float32) if n else np.float32(-0.0)
    assert got.tobytes() == np.float32(want).tobytes()


def _walk(env, start, end, thr, hyst, merge, min_len):
    """The hysteresis ranges as a chain: a range opens at the first low sample at or
    after the merge point and closes at the first high sample after its start."""
    low = np.flatnonzero(env[start:end] <= thr) + start
    high = np.flatnonzero(env[start:end] >= thr * hyst) + start
    out, p = [], start
    while True:
        k = np.searchsorted(low, p)
        if k == low.size:
            break
        s = int(low[k])
        j = np.searchsorted(high, s + 1)
        e = int(high[j]) if j < high.size else end
        if e - s > min_len:
            out.append((s, e))
        if j == high.size:
            break
        p = e + merge + 1
    return out


@pytest.mark.parametrize("seed", range(12))
def test_find_dropouts_equals_chain_walk(seed):
    rng = np.random.default_rng(seed)
    n = 20000
    env = rng.uniform(0.5, 1.5, n)
    for _ in range(rng.integers(5, 60)):
        at = int(rng.integers(0, n))
        env[at:at + int(rng.integers(1, 400))] *= rng.uniform(0.0, 0.4)
    thr = float(rng.uniform(0.1, 0.3))
    start, end = int(rng.integers(0, 100)), int(n - rng.integers(0, 100))
    merge, min_len = int(rng.integers(0, 40)), int(rng.integers(0, 20))
    want = _walk(env, start, end, thr, 1.25, merge, min_len)
    assert O.find_dropouts(env, start, end, thr, 1.25, merge, min_len) == want


@pytest.mark.parametrize("seed", range(8))
def test_map_rows_cover_each_range(seed):
    """Each range maps to consecutive rows from the line holding its start to the line
    holding its end; mapped back to samples, the rows cover the range and overshoot it
    by less than one output pixel at either end."""
    rng = np.random.default_rng(100 + seed)
    W = 910
    lines = 40
    loc = np.cumsum(np.r_[rng.uniform(0, 50), rng.uniform(2500, 2600, lines)])
    pts = np.sort(rng.uniform(loc[0], loc[-2], 2 * int(rng.integers(1, 10))))
    ranges = [(float(np.floor(pts[2 * k])), float(np.ceil(pts[2 * k + 1])) + 1.0) for k in range(pts.size // 2)]
    ranges = [r for i, r in enumerate(ranges) if i == 0 or r[0] > ranges[i - 1][1]]
    fl, sx, ex = O.map_tape_rf_to_tbc(ranges, loc, W, 0, lines, -1)
    row = 0
    for s, e in ranges:
        first = int(np.searchsorted(loc, s, side="right")) - 1
        last = int(np.searchsorted(loc, e, side="right")) - 1
        for li in range(first, last + 1):
            assert fl[row] == li + 1
            ls, ln = loc[li], loc[li + 1] - loc[li]
            a = ls + sx[row] / W * ln
            b = ls + ex[row] / W * ln
            assert a <= max(s, ls) + 1e-6 and max(s, ls) - a < ln / W
            assert b >= min(e, loc[li + 1]) - 1e-6 and b - min(e, loc[li + 1]) < ln / W
            row += 1
    assert row == fl.size


def test_tape_dropouts_field_on_synthetic_envelope():
    """Planted envelope dips on a field of 263 lines: the ranges are the dips longer than
    the minimum length (threshold 0.18 x mean, nothing else comes near it) and the rows
    name the lines that hold them, one line per row."""
    rng = np.random.default_rng(3)
    L = 2542.2
    nloc = 264
    line0 = 40000.3
    loc = line0 + L * np.arange(nloc)
    n = int(loc[-1]) + 200000
    env = rng.uniform(900.0, 1100.0, n).astype(np.float32)
    dips = [(60000, 60120), (61000, 61008), (100000, 103000), (200000, 200011), (400000, 400012)]
    for a, b in dips:
        env[a:b] = 50.0
    r = O.tape_dropouts(env, loc, 263, (0, n), fraction=0.18, out_len=910)
    want = [(a, b) for a, b in dips if b - a > 10]
    assert r["ranges"] == want
    assert abs(float(r["mean"]) - float(np.mean(env, dtype=np.float64))) < 0.05
    lines = []
    for a, b in want:
        first = int(np.searchsorted(loc, a, side="right")) - 1
        last = int(np.searchsorted(loc, b, side="right")) - 1
        lines += list(range(first + 1, last + 2))
    assert r["field_line"].tolist() == lines


def test_read_window_reference_vectors():
    """CompatibilityTests.cs:2003-2024: request lengths and resolved windows for NTSC and
    PAL LaserDisc (stride 31712) and VHS (stride 30720)."""
    assert O.estimate_read_samples(63.5555555555, 40e6, False) == 950_272
    assert O.estimate_read_samples(64.0, 40e6, True) == 1_081_344
    assert O.resolve_read_window(0, 950_272, stride=31_712) == (0, 951_360)
    assert O.resolve_read_window(700_000, 950_272, stride=31_712) == (665_952, 983_072)
    assert O.resolve_read_window(0, 1_081_344, stride=31_712) == (0, 1_109_920)
    assert O.resolve_read_window(123, 950_272, stride=30_720) == (0, 952_320)


def test_host_read_sample_count_matches_oracle():
    from vhsdecode_hip import field, formats
    for system in ("NTSC", "PAL"):
        fmt = formats.get_format(system, "VHS", "sp")
        for fs in (20e6, 28.6e6, 40e6):
            want = O.estimate_read_samples(float(fmt.sysparams["line_period"]), fs, system == "PAL")
            assert field.read_sample_count(fmt, fs) == want
