"""Pins of the oracle's LaserDisc and HiFi restatements against the reference's own
golden vectors, and of the LaserDisc filter design against what the mathematics fixes.
CPU only."""
import hashlib
import math

import numpy as np
import pytest
from scipy import signal

from oracle import oracle as O
from vhsdecode_hip import formats, ld_filters, synth

N = 32768


def _sha(a, dtype):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=dtype).tobytes()).hexdigest().upper()


# ----------------------------------------------------------------- discriminator

def test_conjugate_product_unwrap_vector():
    """CompatibilityTests.cs:4748-4751 (PortedMath.UnwrapHilbertConjugateProduct)."""
    out = O.conj_unwrap(np.array([1 + 2j, -3 + 4j, -2 - 5j, 7 - 1j]), 8.0)
    np.testing.assert_array_equal(out, [0.0, 1.409665529398267, 2.6961931748400927, 2.3038068251599078])


def test_conjugate_product_unwrap_quarter_turns():
    """CompatibilityTests.cs:4722-4740: a quarter turn per sample at fs = 8 is 2 Hz, and
    a wrap from 7 pi / 4 to pi / 4 is a quarter turn forward."""
    a = np.exp(1j * np.array([0.0, math.pi / 2, math.pi, -math.pi / 2, 0.0]))
    np.testing.assert_allclose(O.conj_unwrap(a, 8.0), [0.0, 2.0, 2.0, 2.0, 2.0], rtol=0, atol=1e-12)
    w = np.exp(1j * np.array([7 * math.pi / 4, math.pi / 4]))
    np.testing.assert_allclose(O.conj_unwrap(w, 8.0), [0.0, 2.0], rtol=0, atol=1e-12)


# ----------------------------------------------------------------- LaserDisc block

def _lcg_u16_block():
    # LaserDiscVideoFilterCompatibilityTests.cs:96-100
    i = np.arange(N, dtype=np.uint64)
    return ((i * np.uint64(1_103_515_245) + np.uint64(12_345)) & np.uint64(0xFFFF)).astype(np.float64) - 32768.0


@pytest.fixture(scope="module")
def pal_custom_deemphasis_block():
    fmt = formats.get_laserdisc_format("PAL", deemp_low_mhz=0.25, deemp_high_mhz=1.5, deemp_strength=0.75)
    fl = ld_filters.build_laserdisc_filters(fmt, 40e6, N, analog_audio=False)
    return fl, O.demod_block_ld(_lcg_u16_block(), fl)


def test_ld_custom_deemphasis_time_constants():
    fmt = formats.get_laserdisc_format("PAL", deemp_low_mhz=0.25, deemp_high_mhz=1.5, deemp_strength=0.75)
    assert fmt.rfparams["video_deemp"] == [1.0 / 1_500_000.0, 1.0 / 250_000.0]
    assert fmt.rfparams["video_deemp_strength"] == 0.75


@pytest.mark.parametrize("channel,expected", [
    ("video", "5081BDE4C57623E1254A6FFB27E306215691829A1667BCF04FEC710A12B01C34"),
    ("demod", "E4AC0D4BB8769F27761AFEF067E3F935919E53707AA0C06D08988FD24E1C8CD2"),
    ("video_lowpass", "B24DE7F7A6DB3769580318592F21A421DCF5811FDC64AEB81382333EB52DED33"),
    ("video_pilot", "8AA6600E865D36EECE881C6C8FB8D7DB01778D66C674A3C9B26F3A159E3F576D"),
])
def test_ld_block_matches_reference_hashes(pal_custom_deemphasis_block, channel, expected):
    """LaserDiscVideoFilterCompatibilityTests.cs:94-134: float32 SHA-256 of the LCG block's
    Video, DemodRaw, VideoLowPass and VideoPilot (PAL, de-emphasis 0.25/1.5 MHz at 0.75).
    The VideoBurst hash (7692477D...) is not reproduced: the burst channel's parity with the
    reference is unpinned."""
    _, blk = pal_custom_deemphasis_block
    assert _sha(getattr(blk, channel), "<f4") == expected


def test_ld_burst_reference_is_the_fsc_band(pal_custom_deemphasis_block):
    """The burst plane is the demod through the 81-tap band-pass at fsc +- 0.2 MHz:
    its spectrum outside fsc +- 1 MHz is tiny next to the in-band part."""
    fl, blk = pal_custom_deemphasis_block
    spec = np.abs(np.fft.rfft(blk.video_burst))
    f = np.arange(spec.size) * 40e6 / N
    fsc = 4.43361875e6
    inband = spec[np.abs(f - fsc) < 0.2e6].mean()
    outband = spec[(np.abs(f - fsc) > 1.0e6) & (f > 0.2e6)].mean()
    assert outband < 1e-2 * inband


# ----------------------------------------------------------------- LaserDisc filter design

@pytest.mark.parametrize("system", ["PAL", "NTSC"])
def test_ld_mtf_closed_form(system):
    """BuildLdMtf: 1 / (1 + a1 z^-1 + a2 z^-2) with the two poles at MTF_poledist and
    +-MTF_freq, raised to level * MTF_basemult (numpy power)."""
    fmt = formats.get_laserdisc_format(system)
    rf = fmt.rfparams
    r, f0 = float(rf["MTF_poledist"]), float(rf["MTF_freq"]) * 1e6
    w = 2 * np.pi * np.arange(N) / N
    theta = np.pi * f0 / 20e6
    den = (1 - r * np.exp(1j * theta) * np.exp(-1j * w)) * (1 - r * np.exp(-1j * theta) * np.exp(-1j * w))
    for level in (1.0, 2.0, 0.5, 2.1):
        got = ld_filters.laserdisc_mtf(fmt, level)
        eff = level * float(rf["MTF_basemult"])
        want = (1.0 / den) ** eff
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)


def test_ld_filter_products_use_fused_multiply_add():
    """NumpyVectorComplexMultiply: real = fma(ar, br, -(ai bi)), imag = fma(ar, bi, ai br).
    With ar = 1 + 2^-30, br = 1 - 2^-30 the exact product 1 - 2^-60 rounds to 1 on its own,
    so only the fused form keeps the -2^-60 of the real part."""
    e = 2.0 ** -30
    a = np.array([(1.0 + e) + 1j * (1.0 - e)])
    b = np.array([(1.0 - e) + 1j * (1.0 + e)])
    got = ld_filters.cmul_fma(a, b)[0]
    assert got.real == -(2.0 ** -60)
    assert (1.0 + e) * (1.0 - e) - (1.0 - e) * (1.0 + e) == 0.0   # the separately rounded form
    assert got.imag == 2.0 - 2.0 ** -59


def test_ld_video_lowpass_is_the_butterworth_transfer_function():
    fmt = formats.get_laserdisc_format("NTSC")
    fl = ld_filters.build_laserdisc_filters(fmt, 40e6, N, analog_audio=False)
    b, a = signal.butter(int(fmt.rfparams["video_lpf_order"]), float(fmt.rfparams["video_lpf_freq"]) / 20e6)
    _, h = signal.freqz(b, a, worN=N, whole=True)
    np.testing.assert_allclose(fl.video_lowpass, h, rtol=1e-9, atol=1e-12)


def test_ld_group_delay_equalizer_is_unit_magnitude_and_hermitian():
    fmt = formats.get_laserdisc_format("NTSC")
    fl = ld_filters.build_laserdisc_filters(fmt, 40e6, N, analog_audio=False)
    eq = fl.video / ld_filters.cmul_fma(fl.video_lowpass, ld_filters._deemphasis(fmt, 40e6, N, True))
    np.testing.assert_allclose(np.abs(eq[np.abs(fl.video) > 1e-6]), 1.0, rtol=1e-9)
    k = np.arange(1, N // 2)
    np.testing.assert_allclose(fl.video[N - k], np.conj(fl.video[k]), rtol=1e-12, atol=1e-15)


def test_ld_fft_slice_geometry():
    """DetermineFftSlice for the NTSC carriers at 40 MHz: 1024-bin slices (decimation 32),
    the carrier a quarter of the slice above its low bin."""
    fmt = formats.get_laserdisc_format("NTSC")
    fl = ld_filters.build_laserdisc_filters(fmt, 40e6, N)
    for ch, fc in ((fl.audio_left, 2_301_136.3636363638), (fl.audio_right, 2_812_500.0)):
        assert ch.bin_count == 1024 and ch.slice_rate_hz == 40e6 / N * 1024
        assert ch.low_bin == int(np.rint(fc / (40e6 / N))) - 256
    assert fl.audio_decimation == 32
    spec = np.arange(N) + 1j * np.arange(N)
    sl = ld_filters.slice_spectrum(spec, 100, 8)
    np.testing.assert_array_equal(sl.real, [100, 101, 102, 103, N - 104, N - 103, N - 102, N - 101])


def test_ld_audio_recovers_fm_tones():
    """DecodeAnalogAudioChannel on a noiseless synthetic LaserDisc: each carrier's tone is
    recovered within 0.5 % of the deviation once the band-pass's linear-phase delay
    ((taps - 1) / 2 RF samples) and the discriminator's half-sample delay are taken out."""
    fmt = formats.get_laserdisc_format("NTSC")
    fl = ld_filters.build_laserdisc_filters(fmt, 40e6, N)
    car = ((2_301_136.3636363638, 100e3, 1000.0, 0.1), (2_812_500.0, 100e3, 3000.0, 0.1))
    tape = synth.generate(fmt, fields=1, seed=3, snr_db=80.0, audio_carriers=car)
    start = 40000
    left, right = O.ld_audio_block(tape.rf[start:start + N], fl)
    delay = (int(fmt.rfparams["audio_filterorder"]) - 1) / 2.0 + 16.0
    idx = start + np.arange(1024) * 32 - delay
    for ch, out in ((0, left), (1, right)):
        err = out[40:980] - tape.audio_frequency(ch, idx)[40:980]
        assert np.sqrt(np.mean(err ** 2)) < 0.005 * 100e3
        assert abs(np.mean(err)) < 0.002 * 100e3


# ----------------------------------------------------------------- HiFi Hilbert discriminator

def test_hifi_phase_kernel_vector():
    """HiFiCompatibilityTests.cs:576-597."""
    phase = np.array([0.0, 0.25, 3.0, -3.0, -2.5, 2.9, -3.1, 3.1, -3.14, 3.14, 0.0], dtype=np.float32)
    out = np.full(phase.size, 123.25, dtype=np.float32)
    O.hifi_demodulate_phase(phase, out, 8_388_608, 1_500_000, 100_000)
    assert _sha(out, "<f4") == "99FD09AB9488B7DFD95166DEA2D4753FFD9CA6D1DCFEF1FFCBC75BFE0C02FE98"
    assert out[-1] == np.float32(123.25)


_HIFI_RAMP = {64: "4CCCB11215C35B4CF6E3BBBA99AC2DCC0763C394D81D0A633B749144F48B0B01",
              128: "C82C86BA3413B0BC2CBDB53845D6CA478F4166DE5964A2C7C2E656E8210E8AE6",
              255: "81D40AF0ABE64834A0D153B6F43B8C2C4D38E3D5F058000C1531FD796975C1F8",
              256: "38C55E64B4E5E0AD9CD0160AB0DBCA59DE265043618C27CAA00A940DE0F89F45",
              257: "941DCA6294586C8B99583D35FB1CDBD1324C453DEC89573589A963B82F787CD1",
              512: "7AEA0D18364F6C89AF1F764816AB6416B24793756E339C20DF67A7470E789E6D",
              1000: "9489484FE45539857F30E34041E1A5DAED841230CB9BEC8FEAB10DFCB549EEDB",
              1024: "8F99555022C09B162EB9362AA13845B2B3C4250AFE3954393C328ABB9B6450B9"}
_HIFI_RANDOM = {64: "97E25C6FA9C1C1602A162A0364A10DBF5AD22AF068B08497F28682F96D59DA22",
                128: "51AA963F25C574E013CA3446036336635A1935F2872838CB015B59F8A3CE902B",
                255: "8F62CB790355C4328907DFC1D3547F1343B186790019CDEDA82BC82782ABD8EE",
                256: "9A512755DA76439087EE6A2B77B1D880B1EC15EFABD891A80A41F81B4713C7A5",
                257: "CFDB7727E2A94F9C2A6B2C4270EB1E7B59AE2C4790B9330A25FA0F164273BE28",
                512: "8F5D781234FC28EF7D8CBFE1A5FB802AC70380B3943FD393B8659D274D496230",
                1000: "A318DB9BA05C1FF8B244C2CFA7C2FFAA098A8FCB1552F1A1647631FEC74D6B7A",
                1024: "F1599DFA9591AC0A168F0BEE7F87E46B68B3CEB5CD9156F7D0FA83CE3C2C0395"}


def hifi_ramp_input(n):
    # HiFiCompatibilityTests.cs:612-614
    return np.array([np.float32(((i * 37) % 251) - 125) / np.float32(128.0) for i in range(n)], dtype=np.float32)


def hifi_random_input(n):
    # HiFiCompatibilityTests.cs:641-647
    out = np.empty(n, dtype=np.float32)
    state = 0x12345678
    for i in range(n):
        state = (state * 1_664_525 + 1_013_904_223) & 0xFFFFFFFF
        out[i] = np.float32((state >> 8) - 0x800000) / np.float32(0x800000)
    return out


@pytest.mark.parametrize("n", sorted(_HIFI_RAMP))
def test_hifi_fft_discriminator_vectors(n):
    """HiFiCompatibilityTests.cs:599-670: Demodulate at 2^23 Hz, 1.5 MHz carrier, 100 kHz
    deviation on a ramp and an LCG input; the last sample is left untouched."""
    for make, table in ((hifi_ramp_input, _HIFI_RAMP), (hifi_random_input, _HIFI_RANDOM)):
        out = np.full(n, 123.25, dtype=np.float32)
        O.hifi_demodulate(make(n), out, 8_388_608, 1_500_000, 100_000)
        assert _sha(out, "<f4") == table[n]
        assert out[-1] == np.float32(123.25)
