"""LaserDisc analog audio stage 2 on the oracle (oracle/ld_audio.py) against the
reference's own vectors.

Sources: tests/VHSDecode.Tests/CompatibilityTests.cs:6178-6236 (overlap, the float32
carrier, peak suppression, identity filter of 1024 points),
LaserDiscAudioNumericsCompatibilityTests.cs:14-57 (the float32 real FFT and the PAL v0.4.0
stage 2 over 34685 samples, both SHA-256 of float32 bits) and :59-91 (the output dtype).
The PAL stage-2 hashes are asserted by the reference on Windows only
(WindowsFrozenBitOracle); they are reported here where the Linux libraries differ.
"""
import hashlib

import numpy as np
import pytest

from oracle import ld_audio as LA
from vhsdecode_hip import formats, ld_filters


def _identity(n):
    return np.ones(n, np.complex128)


def test_short_block_passes_through_the_identity_filter():
    ident = _identity(1024)
    left = 1000.0 + np.arange(8.0)
    right = 2000.0 + 2.0 * np.arange(8.0)
    ol, orr, f32 = LA.phase2(left, right, ident, ident, 1000.0, 2000.0)
    assert not f32
    np.testing.assert_allclose(ol, left, atol=1e-8, rtol=0)
    np.testing.assert_allclose(orr, right, atol=1e-8, rtol=0)


def test_float32_carrier_subtraction():
    ident = _identity(1024)
    carrier = 2_301_136.3636363638
    src = np.array([2_313_564.0, 2_310_882.0, 2_308_120.0])
    ol, _, _ = LA.phase2(src, src, ident, ident, carrier, carrier)
    want = (src.astype(np.float32) - np.float32(carrier)).astype(np.float64) + carrier
    np.testing.assert_allclose(ol, want, atol=1e-8, rtol=0)


def test_overlap_places_the_final_block_one_sample_late():
    ident = _identity(1024)
    left = 1000.0 + np.arange(1500.0)
    right = 2000.0 + np.arange(1500.0)
    ol, orr, f32 = LA.phase2(left, right, ident, ident, 1000.0, 2000.0)
    assert f32
    assert abs(left[987] - ol[988]) < 1e-8
    assert abs(right[1498] - orr[1499]) < 1e-8


def test_peaks_of_the_left_channel_are_suppressed_in_both():
    ident = _identity(1024)
    ol, orr, _ = LA.phase2([1000.0, 601_001.0, 701_001.0, 1000.0], [2000.0, 2010.0, 2020.0, 2030.0],
                           ident, ident, 1000.0, 2000.0)
    assert np.all(np.abs(ol - 1000.0) < 1e-8)
    assert np.all(np.abs(orr - 2000.0) < 1e-8)


def test_output_dtype_follows_the_block_count():
    ident = _identity(1024)
    s = 1000.0 + np.arange(8.0)
    l = 1000.0 + np.arange(1500.0)
    assert LA.phase2(s, s, ident, ident, 1000.0, 1000.0)[2] is False
    assert LA.phase2(l, l, ident, ident, 1000.0, 1000.0)[2] is True


def test_peak_rule_against_brute_force():
    rng = np.random.default_rng(4)
    x = rng.uniform(0, 900_000.0, 300).astype(np.float32)
    z = [0.0 if float(v) < 500_000.0 else float(v) for v in x]
    want = [i - 1 for i in range(len(z) - 1) if z[i] > z[-1] and z[i + 1] > z[i]]
    assert LA.find_peaks(x) == want


def test_float32_real_fft_matches_the_reference_vector():
    idx = np.arange(32768)
    x = (((idx * 7919) % 65521) - 32760).astype(np.float32)
    spec = np.fft.rfft(x)
    assert spec.dtype == np.complex64  # computed in float32
    ref = np.fft.rfft(x.astype(np.float64))
    assert np.max(np.abs(spec - ref)) <= 2e-6 * np.max(np.abs(ref))
    h = hashlib.sha256(spec.tobytes()).hexdigest().upper()
    if h != "592EE58B19B8AAA305F6412284488191C5BD2B9BB97AB879006720DF51BD1C56":
        pytest.xfail(f"numpy's float32 pocketfft rounds differently from the reference's port ({h[:12]})")


def _pal_channel(center, multiplier, length=34685):
    i = np.arange(length, dtype=np.int64)
    off = ((((i * multiplier) % 200_003) - 100_001).astype(np.float32) * np.float32(0.25)).astype(np.float32)
    return (np.float32(center).astype(np.float64) + off.astype(np.float64))


def test_pal_stage2_v040_vector():
    fl = ld_filters.build_laserdisc_filters(formats.get_laserdisc_format("PAL"), 40e6, 32768)
    a, b = fl.audio_left, fl.audio_right
    left = _pal_channel(a.center_hz, 7919)
    right = _pal_channel(b.center_hz, 15427)
    ol, orr, f32 = LA.phase2(left, right, a.stage2, b.stage2, a.center_hz, b.center_hz)
    assert f32
    hl = hashlib.sha256(ol.astype(np.float32).tobytes()).hexdigest().upper()
    hr = hashlib.sha256(orr.astype(np.float32).tobytes()).hexdigest().upper()
    if (hl, hr) != ("68CB22D86AD3C6C85FFE3986CB2DC197C9DE6EA9ECF308E926E3480BC627C8DE",
                    "0A4B5F027D0072C3F50527544744A516336A1D9701438B32A953564064FE3091"):
        pytest.xfail(f"PAL stage-2 output differs from the Windows-frozen vector ({hl[:12]}, {hr[:12]})")
