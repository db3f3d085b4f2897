"""The chroma restatement (oracle/chroma.cpp) against the reference's own vectors.

Sources: tests/VHSDecode.Tests/CompatibilityTests.cs (VhsChromaDecoderUpconvertsLinePhases
:2830-2922, VhsChromaDecoderDetectsBurstPhaseSequence :2948-3160, the burst-magnitude
vector :3163-3176, the ACC/comb/de-emphasis vectors :2717-2825, the PAL v0.4.0 probe
and phase-sequence bit vectors :3179-3345).  Vectors the reference asserts only on
Windows (WindowsFrozenBitOracle: UCRT transcendentals) are checked where the Linux C
library reproduces them and otherwise reported, as the reference's own Linux run does.
"""
import hashlib
import math
import struct

import numpy as np
import pytest

from oracle import chroma as C
from vhsdecode_hip import chroma as HC
from vhsdecode_hip import formats


def _bits(x: float) -> int:
    return struct.unpack("<Q", struct.pack("<d", x))[0]


def _f32bits(x: float) -> int:
    return struct.unpack("<I", struct.pack("<f", x))[0]


def _sha(values) -> str:
    return hashlib.sha256(np.ascontiguousarray(values, dtype="<f8").tobytes()).hexdigest().upper()


def test_heterodyne_table_values():
    het = C.heterodyne(4, 3.0, 1.0, 16.0)
    assert het[0][1] == float(np.float32(-math.cos(math.tau * (4.0 / 16.0) * 1.0)))
    np.testing.assert_allclose(het[0], [-1.0, 0.0, 1.0, 0.0], atol=1e-12)
    np.testing.assert_allclose(het[1], [0.0, 1.0, 0.0, -1.0], atol=1e-12)


def test_upconvert_line_phases_and_numpy_slices():
    chroma = [10.0, 20.0, 30.0, 40.0, 50.0, 60.0]
    tables = [np.full(6, 2.0), np.full(6, -1.0)]
    up = C.upconvert(chroma, 1, 3, [1, 2], [0, 1], tables)
    np.testing.assert_allclose(up, [20.0, 40.0, 60.0, -40.0, -50.0, -60.0], atol=1e-12)
    sliced = C.upconvert(np.arange(1, 10) * 10.0, 3, 3, [1, 2, 3], [0, 1, 1], [np.full(9, 2.0), np.full(9, -1.0)])
    np.testing.assert_allclose(sliced, [-10.0, -20.0, -30.0, 80.0, 100.0, 120.0, 0.0, 0.0, 0.0], atol=1e-12)


def test_upconvert_phase_compensated_vectors():
    comp = C.upconvert_compensated([1.0, 1.0, 1.0, 1.0], 0, 2, [0, 1], [0, 1], [0.0, 90.0], 0.0, 1.0, 0.0, 0.0)
    np.testing.assert_allclose(comp, [-1.0, 0.0, -1.0, 0.0], atol=1e-12)
    numba = C.upconvert_compensated([0.5, -1.25, 2.0, -3.5, 4.25, -5.0, 6.5, -7.75, 8.0, -9.5], 1, 10, [1], [3],
                                    [55.0], 629_370.0, 315.0 / 88.0, -33.0, -33.0)
    expected = [3191852143, 1065141330, 1070476877, 3214743597, 3230108355,
                3212239443, 1085893508, 1084778306, 3229968243, 3239209412]
    assert [_f32bits(v) for v in numba] == expected
    neg = C.upconvert_compensated([1.0, 2.0, 3.0, 4.0], 1, 2, [0], [0], [0.0], 0.0, 1.0, 0.0, 0.0)
    np.testing.assert_allclose(neg, [0.0, 0.0, 3.0, 0.0], atol=1e-12)


def test_demod_burst_vectors():
    b = C.demod_burst([1.0, 2.0], 0.5, 0, 2, [0.0, 0.0, 0.0, 1.0], [0.0, 0.0, 1.0, 0.0])
    assert abs(b[0] - math.degrees(math.atan2(2.0, 1.0))) < 1e-12
    assert abs(b[1] - 90.0) < 1e-12 and abs(b[2] - math.sqrt(5.0)) < 1e-12
    assert abs(b[3] - 1.0) < 1e-12 and abs(b[4] - 2.0) < 1e-12
    burst = [(((i * 37) % 101 - 50) / 8.0) + math.ldexp(i + 1.0, -35) for i in range(70)]
    bsin = [float(np.float32(((i * 17) % 53 - 26) / 16.0)) for i in range(96)]
    bcos = [float(np.float32(((i * 29) % 67 - 33) / 16.0)) for i in range(96)]
    r = C.demod_burst(burst, 0.9989764459848242, 0, 7, bsin, bcos)
    assert [_bits(v) for v in (r[3], r[4], r[2], r[0], r[1])] == [
        0xC038BA000003D400, 0x400CDFFFFF95D000, 0x4038FD155992DCC5, 0x4065763E428A68F1, 0x3FE4A28575E4BBDA]
    r32 = C.demod_burst(burst, 0.9989764459848242, 0, 7, bsin, bcos, f32=True)
    assert [_bits(v) for v in (r32[3], r32[4], r32[2], r32[0], r32[1])] == [
        0xC038B9FFFFFFC000, 0x400CDFFFFFE90000, 0x4038FD1559905438, 0x4065763E42874AC4, 0x3FE4A28575E4BBDA]
    b4 = [-1937.88232421875, -110.7741928100586, 42.956520080566406, -197.7586212158203]
    m = C.demod_burst(b4, 1.0, 0, 0, [0.0, 0.0, 1.0, 1.0], [1.0, 1.0, 0.0, 0.0])
    # The reference's vector (0x40A00CFE60355380, CompatibilityTests.cs:3163-3176) is the
    # platform hypot it was frozen on; NumpyHypot calls libm's hypot on Linux, which here
    # returns the correctly rounded value one ulp below it (…537F).  The oracle follows
    # libm; the I and Q it is taken from are the vector's exact sums.
    i, q = b4[0] + b4[1], b4[2] + b4[3]
    assert m[3] == i and m[4] == q
    assert _bits(m[2]) in (0x40A00CFE60355380, 0x40A00CFE6035537F)
    assert abs(_bits(m[2]) - 0x40A00CFE60355380) <= 1


def test_probe_upconverted_burst_with_filter():
    chroma = np.arange(12, dtype=np.float64)
    s = np.zeros(12)
    c = np.zeros(12)
    c[8] = 1.0
    s[9] = 1.0
    r = C.probe(chroma, np.ones(12), 2, 4, s, c, 0.75, 1, 0, 6, filter_scale=2.0)
    assert abs(r[0] - math.degrees(math.atan2(18.0, 16.0))) < 1e-12
    assert abs(r[1] - 45.0) < 1e-12
    assert abs(r[2] - math.hypot(16.0, 18.0)) < 1e-12
    assert abs(r[3] - 16.0) < 1e-12 and abs(r[4] - 18.0) < 1e-12


def _table(fn):
    t = np.zeros((41, 4, 5))
    for line in range(41):
        for rot in range(4):
            t[line, rot] = fn(line, rot)
    return t


def test_phase_sequence_killer_rotation_and_next_index():
    loc = np.arange(41) * 100.0
    common = dict(locations=loc, line_offset=0, lines_out=40, input_line_length=100, out_line_len=100, het4=None,
                  bsin=None, bcos=None, sos=None, burst_start=0, burst_end=0, is_ntsc=True)
    seq = C.phase_sequence(None, rotation=[], rotation_index=0, color_killer=True, prev_burst_line=-1,
                           probe_table=_table(lambda l, r: (0.0, 0.0, 30_000.0, 30_000.0, 0.0)), **common)
    assert seq.next_index == 0 and seq.burst_line == 17 and abs(seq.mag_avg - 30_000.0) < 1e-12
    assert abs(seq.phase_avg) < 1e-12 and abs(seq.even_avg) < 1e-12 and abs(seq.odd_avg) < 1e-12
    low = C.phase_sequence(None, rotation=[], rotation_index=0, color_killer=True, prev_burst_line=0,
                           probe_table=_table(lambda l, r: (0.0, 0.0, 1_000.0, 1_000.0, 0.0)), **common)
    assert low.burst_line == -1 and abs(low.mag_avg - 1_000.0) < 1e-12

    def rotating(line, rot):
        ph = rot * math.pi / 2.0
        return (rot * 90.0, 0.0, 30_000.0, math.cos(ph) * 30_000.0, math.sin(ph) * 30_000.0)
    rotated = C.phase_sequence(None, rotation=[1, 3], rotation_index=0, probe_table=_table(rotating), **common)
    assert rotated.next_index == 1 and list(rotated.rot[:4]) == [1, 2, 3, 0]
    literal = C.phase_sequence(None, rotation=[1, 3, 2], rotation_index=1, probe_table=_table(rotating), **common)
    assert literal.next_index == 0


def test_refine_lines_from_burst():
    phase = np.zeros(11)
    offset = np.zeros(11)
    phase[12 - 3], offset[12 - 3] = 350.0, 0.0
    phase[13 - 3], offset[13 - 3] = 30.0, 5.0
    loc = np.arange(15) * 100.0
    ntsc = C.refine_lines(loc, 20, 4.0, 3, phase, offset, 0, 10.0, 40.0, 20.0, True)
    assert abs(ntsc[9] - 900.0) < 1e-12
    assert abs(ntsc[12] - (1200.0 + (20.0 / 360.0) * 4.0 * 5.0)) < 1e-12
    assert abs(ntsc[13] - (1300.0 + (-15.0 / 360.0) * 4.0 * 5.0)) < 1e-12
    pal = C.refine_lines(loc, 20, 4.0, 3, phase, offset, 0, 10.0, 40.0, 20.0, False)
    assert abs(pal[12] - (1200.0 + (50.0 / 360.0) * 4.0 * 5.0)) < 1e-12
    assert abs(pal[13] - (1300.0 + (-5.0 / 360.0) * 4.0 * 5.0)) < 1e-12
    killed = C.refine_lines(loc, 20, 4.0, 3, phase, offset, -1, 10.0, 40.0, 20.0, False)
    np.testing.assert_array_equal(killed, loc)


def test_burst_deemphasis_and_comb_vectors():
    d = C.burst_deemphasis(np.ones(16), 1, 2, 8, 2)
    np.testing.assert_allclose(d, [1, 1, 1, 1, 1, 1, 1, 2, 1, 1, 1, 1, 1, 1, 1, 2], atol=1e-12)
    comb_in = np.zeros(40)
    for line in range(20):
        comb_in[2 * line] = line * line
        comb_in[2 * line + 1] = line * line + 1
    ntsc = C.comb(comb_in, 2, 1)
    pal = C.comb(comb_in, 2, 2)
    assert abs(ntsc[32] + 0.5) < 1e-12 and abs(ntsc[33] + 0.5) < 1e-12
    assert abs(pal[32] + 2.0) < 1e-12 and abs(pal[33] + 2.0) < 1e-12
    assert ntsc[30] == comb_in[30] and ntsc[36] == comb_in[36]
    p = np.zeros(20)
    p[15], p[16], p[17] = -0.234567890123, 1.000000000123, 0.345678901234
    exact = ((p[16] * 2.0) - p[17] - p[15]) / 4.0
    f32 = C.comb(p, 1, 1)
    f64 = C.comb(p, 1, 1, retain_f32=False)
    assert f32[16] == float(np.float32(exact)) and f64[16] == exact and f32[16] != f64[16]


def test_acc_vectors():
    x = np.zeros(18 * 4)
    x[64:68] = [100.0, 102.0, 98.0, 100.0]
    x[68:72] = [200.0, 204.0, 196.0, 200.0]
    out, mean = C.acc(x, 10.0, 1, 3, 4, 18, 0)
    assert abs(mean - 3.0) < 1e-12
    np.testing.assert_allclose(out[64:72], [500, 510, 490, 500, 500, 510, 490, 500], atol=1e-12)
    killed, kmean = C.acc(x, 10.0, 1, 3, 4, 18, 17)
    assert abs(kmean - 2.0) < 1e-12 and not killed[64:68].any()
    L = 60
    probe = np.zeros(18 * L)
    for line in (16, 17):
        for i in range(L):
            probe[line * L + i] = float(np.float32(np.float32((((i * 37) + (line * 11)) % 101) - 50) / np.float32(10.0)
                                                   + np.float32(((i % 3) - 1) * 1000.0)))
    o32, _ = C.acc(probe, 4416.0, 4, 56, L, 18, 0)
    assert [_bits(o32[k]) for k in (960, 961, 977)] == [13886080547120341843, 13850057151597280334,
                                                        4662752321137288657]
    L2 = 52
    dp = np.zeros(17 * L2)
    for i in range(L2):
        dp[16 * L2 + i] = ((((i * 37) + 11) % 101) - 50) / 10.0 + (((i % 3) - 1) * 1000.0) + ((i % 7) * 1e-9)
    _, m64 = C.acc(dp, 4416.0, 0, L2, L2, 17, 0, f32_rms=False)
    assert _bits(m64) == 4650426021891215527


def test_acc_comb_u16_matches_combined_acc_conversion():
    # ApplyAutomaticChromaGainWithCombToU16 is ApplyAutomaticChromaGainWithComb followed by
    # the saturating conversion (VhsChromaU16SimdTests.cs:116-160): check the comb path
    # against the ACC vector restatement and the conversion rule on the same field
    rng = np.random.default_rng(3)
    L, lines = 40, 24
    x = rng.normal(0.0, 300.0, lines * L)
    u = C.acc_comb_u16(x, 4416.0, 4, 20, L, lines, 0, 1)
    assert (u[:16 * L] == 32767).all()
    combed = C.comb(x, L, 1)
    for line in range(16, lines - 2):
        seg = combed[line * L:(line + 1) * L]
        rms = C.rms(seg[4:20])
        want = np.clip(seg * (4416.0 / rms) + 32767.0, 0, 65535).astype(np.uint16)
        np.testing.assert_array_equal(u[line * L:(line + 1) * L], want)


def test_numpy_trig_float32_matches_numpy():
    xs = np.float32(np.linspace(-70000.0, 70000.0, 2001))
    got_s = np.array([C.numpy_trig_f32(float(v), False) for v in xs], np.float32)
    got_c = np.array([C.numpy_trig_f32(float(v), True) for v in xs], np.float32)
    # numpy's float32 sin/cos: the same reduction and polynomials (exact where numpy
    # uses them; allow one ulp where this numpy build takes another code path)
    np.testing.assert_allclose(got_s, np.sin(xs), atol=2e-7, rtol=0)
    np.testing.assert_allclose(got_c, np.cos(xs), atol=2e-7, rtol=0)


def _pal_options():
    fmt = formats.get_format("PAL", "VHS", "sp")
    return fmt, HC.chroma_params(fmt)


def test_pal_options_burst_window_and_final_filter():
    _, p = _pal_options()
    assert (p.burst_start, p.burst_end) == (94, 150)
    flat = p.final_sos.reshape(-1)
    h = _sha(flat)
    # Windows-frozen in the reference (E2A9F54E...): report a mismatch as a signal
    if h != "E2A9F54E967697937EBE3D44357B8805D0740FBC6C7651C8B596CC4A69992E86":
        pytest.xfail(f"final SOS filter hash differs from the Windows-frozen value ({h[:16]})")


def _pal_probe_inputs():
    fmt, p = _pal_options()
    n = 3 * p.line_len
    chroma = np.array([(((i * 37) % 101 - 50) / 8.0) + math.ldexp(i + 1.0, -35) for i in range(n)])
    het = C.heterodyne(n, p.fsc_mhz, p.color_under_hz / 1e6, p.fsc_mhz * 4.0)
    s, c = C.carrier(n, p.fsc_mhz, p.fsc_mhz * 4.0)
    return p, chroma, het, s, c


def test_pal_probe_release_bits():
    p, chroma, het, s, c = _pal_probe_inputs()
    r = C.probe(chroma, het[3], p.burst_start, p.burst_end, s, c, 0.9989764459848242, 2, 1, p.line_len,
                sos=p.final_sos)
    got = [_bits(v) for v in (r[0], r[1], r[2], r[3], r[4])]
    want = [0x4064C329592D3C5F, 0x4021518B70A91DA9, 0x402B4FBACBBD83FC, 0xC02A82F3B467500F, 0x400A3F02088DC1E6]
    if got != want:
        # Windows-frozen (UCRT cos/atan2/hypot); the offset does not depend on them
        assert got[1] == want[1]
        pytest.xfail("probe bits differ from the Windows-frozen vector")


def test_pal_phase_sequence_release_bits():
    fmt, p = _pal_options()
    n = p.field_samples
    chroma = np.array([(((i * 37) % 101 - 50) / 8.0) + math.ldexp(i + 1.0, -35) for i in range(n)])
    het = C.heterodyne(n, p.fsc_mhz, p.color_under_hz / 1e6, p.fsc_mhz * 4.0)
    s, c = C.carrier(n, p.fsc_mhz, p.fsc_mhz * 4.0)
    loc = np.arange(3 + p.lines + 2) * 2560.0
    r = C.phase_sequence(chroma, loc, 3, p.lines, 2560, p.line_len, het, s, c, p.final_sos, p.burst_start,
                         p.burst_end, p.rotation, f32=False, is_ntsc=False)
    assert r.next_index == 1 and len(r.rot) == 313 and r.burst_line == 0
    assert r.lines[0] == 3 and r.rot[0] == 0 and r.lines[-1] == 315 and r.rot[-1] == 0
    assert _bits(r.burst[0][1]) == 0 and _bits(r.burst[-1][1]) == 0
    flat = np.column_stack([r.lines.astype(np.float64), r.rot.astype(np.float64), r.burst]).reshape(-1)
    if _sha(flat) != "8BFBCFFD123B67693FA6EC24B9EFD91A5D90A9BF38BA1BA8926B41F24B9754FB":
        pytest.xfail("phase sequence differs from the Windows-frozen vector (UCRT transcendentals)")
    assert _bits(r.mag_avg) == 0x4037E9036918DF9A
