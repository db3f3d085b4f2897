"""Pins of the CPU oracle against the reference's own golden vectors and against
what the mathematics fixes.  CPU only."""
import hashlib
import json
import os
import struct

import numpy as np
import pytest
from scipy import signal

from oracle import oracle as O
from vhsdecode_hip import filters, formats

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
N = 32768


def _lcg_u8_block(n=N):
    # CompatibilityTests.cs:16657-16663
    out = np.empty(n, dtype=np.uint8)
    state = 0x12345678
    for i in range(n):
        state = (state * 1664525 + 1013904223) & 0xFFFFFFFF
        out[i] = state >> 24
    return out


def _sha_f32(a):
    return hashlib.sha256(np.asarray(a, dtype=np.float32).tobytes()).hexdigest().upper()


def _sha_f64(a):
    return hashlib.sha256(np.asarray(a, dtype=np.float64).tobytes()).hexdigest().upper()


# ----------------------------------------------------------------- TBC KATs
# TbcParallelResamplerTests.cs:76-183 (SHA-256 of the double output bits)

def test_tbc_sinc_interior_and_edges_kat(sinc_lut):
    out = O.tbc_resample(np.arange(64.0), [-4.25, 28.5, 67.75], 0, 2, 16, sinc_lut)
    assert _sha_f64(out) == "7E3B3F5DCF0E5C38FB7C46E6E5CF362D596F9DECD7416445247F0A4C4878F33C"


def test_tbc_negative_coordinates_kat(sinc_lut):
    out = O.tbc_resample(np.arange(64.0), [-4.25, 28.5], 0, 1, 16, sinc_lut)
    assert _sha_f64(out) == "426959AEB3440862EF1B3148CF09D4AF143D5F4F21944E04C6AAA62214F4C0BC"


def test_tbc_short_source_kat(sinc_lut):
    out = O.tbc_resample(np.arange(8.0), [0.0, 7.0], 0, 1, 16, sinc_lut)
    assert _sha_f64(out) == "5E5BED2929B57864FE4E2514C4D3754CF7168A6200A7C57FF45BDE2F97B1BD1F"


@pytest.mark.parametrize("case", ["nan_pos_neg", "nan_neg_pos", "inf", "signed_zero"])
def test_tbc_special_values_kat(sinc_lut, case):
    fin = np.array([np.sin(i * 0.0031) + np.cos(i * 0.0007) for i in range(4096)])
    pn = float(struct.unpack("<f", struct.pack("<I", 0x7FC12345))[0])
    nn = float(struct.unpack("<f", struct.pack("<I", 0xFFC54321))[0])
    s = fin.copy()
    expected = {
        "nan_pos_neg": ("6988C63B251226B8CCB1F26C93AC8B7D1026F954F65B6CF7B99339AD9FA6D573", 0xFFF8A86420000000),
        "nan_neg_pos": ("D849E74D2A52D9C9E9777B22FAE5AA4D239330505DD90128CDF1687FB8B784AD", 0x7FF82468A0000000),
        "inf": ("374D71C421F1738DDC02A924C9EA4D3DC80319D3008F0F77EB7E9F70EE45119F", 0xFFF0000000000000),
        "signed_zero": ("41267CF226CA5DD39BB245A02726A3862B35DABD1DF2F3C14CF150B46E808B05", 0),
    }[case]
    if case == "nan_pos_neg":
        s[993], s[994] = pn, nn
    elif case == "nan_neg_pos":
        s[993], s[994] = nn, pn
    elif case == "inf":
        s[993], s[994] = np.inf, -np.inf
    else:
        for t in range(16):
            s[993 + t] = 0.0 if t % 2 == 0 else -0.0
    out = O.tbc_resample(s, [1000.25, 2000.25], 0, 1, 16, sinc_lut, nominal=1000.0)
    assert _sha_f64(out) == expected[0]
    assert int(out[:1].view(np.uint64)[0]) == expected[1]


def test_tbc_linear_positions_closed_form(sinc_lut):
    # TbcParallelResamplerTests.cs:11-56: positions in closed form; a constant
    # source makes the resampled value equal the level adjust, and the level
    # adjust of a uniform line grid is exactly length/nominal.
    locs = 100.0 + 2000.125 * np.arange(6)
    src = np.full(20000, 3.0)
    out = O.tbc_resample(src, locs, 1, 3, 64, sinc_lut, nominal=2000.125)
    # the 16 taps of each phase sum to 1 within FP32 rounding
    assert np.allclose(out, 3.0, rtol=2e-6)


def test_tbc_non_monotonic_locations(sinc_lut):
    # TbcParallelResamplerTests.cs:58-74: positions 30, 28.75, 27.5, 26.25 for line 1.
    # A linear source (value == index) resampled at those positions returns them
    # up to the windowed-sinc interpolation error.
    out = O.tbc_resample(np.arange(128.0), [20.0, 30.0, 25.0, 40.0], 1, 1, 4, sinc_lut, nominal=10.0)
    lf = (25.0 - 30.0) / 10.0  # level adjust of line 1 is its length / nominal
    pos = np.array([30.0, 28.75, 27.5, 26.25])
    med = np.median([1.0, -0.5, 1.5])
    mad = np.median(np.abs(np.array([1.0, -0.5, 1.5]) - med))
    adj = med if abs(lf - med) > 15 * mad else lf
    np.testing.assert_allclose(out, pos * adj, rtol=1e-3)


# ----------------------------------------------------------------- block chain vs fixture hashes

@pytest.fixture(scope="module")
def family():
    with open(os.path.join(GOLDEN, "tape-family-v0.4.0.json")) as h:
        return json.load(h)


def _case(family, system, fmt, speed="sp"):
    for c in family["cases"]:
        if (c["system"], c["tape_format"], c["tape_speed"]) == (system, fmt, speed):
            return c
    raise KeyError


def test_pal_vhs_block_matches_upstream_channel_hashes(family):
    """tape-family-v0.4.0.json (CompatibilityTests.cs:16645-16719): SHA-256 of the
    float32-cast Video, VideoLowPass, Chroma and Envelope of one LCG u8 block."""
    x = _lcg_u8_block().astype(np.float64)
    fl = filters.build_tape_filters(formats.get_format("PAL", "VHS", "sp"), 40e6, N)
    r = O.demod_block(x, fl)
    c = _case(family, "PAL", "VHS")
    assert _sha_f32(r.video) == c["demod"]
    assert _sha_f32(r.video_lowpass) == c["demod_05"]
    assert _sha_f32(r.chroma) == c["demod_burst"]
    assert _sha_f32(r.envelope) == c["envelope"]


@pytest.mark.parametrize("system,fmt,speed", [("PAL", "VHS", "lp"), ("PAL", "VHS", "ep"), ("PAL", "VHS", "vp"),
                                              ("PAL", "VIDEO8", "sp"), ("PAL", "VIDEO8", "lp"),
                                              ("NTSC", "VIDEO8", "sp")])
def test_sub_deemphasis_matches_upstream_video_hash(family, system, fmt, speed):
    """Formats with use_sub_deemphasis (RfDemodulator.ApplySubDeemphasisRealCore,
    RfDemodulator.cs:2897-2999): the sub-deemphasised Video channel, the sync
    reference and the envelope match the fixture bit-exactly through the float32
    cast.  (Video8's order-10 chroma band-pass is designed differently; its
    chroma hash is not matched and not asserted.)"""
    x = _lcg_u8_block().astype(np.float64)
    fl = filters.build_tape_filters(formats.get_format(system, fmt, speed), 40e6, N)
    assert fl.sub_deemphasis is not None
    r = O.demod_block(x, fl)
    c = _case(family, system, fmt, speed)
    assert _sha_f32(r.video) == c["demod"]
    assert _sha_f32(r.video_lowpass) == c["demod_05"]
    assert _sha_f32(r.envelope) == c["envelope"]


@pytest.mark.parametrize("system", ["PAL", "NTSC"])
@pytest.mark.parametrize("speed", ["sp", "lp", "ep", "vp"])
def test_svhs_matches_upstream_channel_hashes(family, system, speed):
    """S-VHS: custom luma filters (the stored linear sub-deemphasis response and a high
    shelf, DecodeFilterSet.cs:1560-1700) on Video, sub-deemphasis, and all four
    channels bit-exact."""
    x = _lcg_u8_block().astype(np.float64)
    fl = filters.build_tape_filters(formats.get_format(system, "SVHS", speed), 40e6, N)
    r = O.demod_block(x, fl)
    c = _case(family, system, "SVHS", speed)
    assert _sha_f32(r.video) == c["demod"]
    assert _sha_f32(r.video_lowpass) == c["demod_05"]
    assert _sha_f32(r.chroma) == c["demod_burst"]
    assert _sha_f32(r.envelope) == c["envelope"]


@pytest.mark.parametrize("system", ["PAL", "NTSC"])
@pytest.mark.parametrize("speed", ["sp", "lp", "ep", "vp"])
def test_betamax_matches_upstream_hashes(family, system, speed):
    """Betamax: RF high boost on the complex path (RfDemodulator.cs:356-392, 2374-2449,
    with numpy's pairwise float32 envelope mean), band-pass sub-deemphasis and the fsc
    notch (RfDemodulator.cs:1018-1051): Video, VideoLowPass and Envelope bit-exact.  The
    PAL chroma band-pass pairs its sections differently (the reference ties them with
    the platform's csinh rounding), so only NTSC chroma is asserted."""
    x = _lcg_u8_block().astype(np.float64)
    fl = filters.build_tape_filters(formats.get_format(system, "BETAMAX", speed), 40e6, N)
    assert fl.rf_high_boost is not None and fl.fsc_notch is not None and fl.sub_deemphasis is not None
    r = O.demod_block(x, fl)
    c = _case(family, system, "BETAMAX", speed)
    assert _sha_f32(r.video) == c["demod"]
    assert _sha_f32(r.video_lowpass) == c["demod_05"]
    assert _sha_f32(r.envelope) == c["envelope"]
    if system == "NTSC":
        assert _sha_f32(r.chroma) == c["demod_burst"]


def test_rf_high_boost_changes_the_demod():
    """The boost moves the demod (so Video and VideoLowPass) and leaves the envelope."""
    x = _lcg_u8_block().astype(np.float64)
    fl = filters.build_tape_filters(formats.get_format("PAL", "BETAMAX", "sp"), 40e6, N)
    on = O.demod_block(x, fl)
    fl.rf_high_boost = None
    off = O.demod_block(x, fl)
    assert np.max(np.abs(on.video_lowpass - off.video_lowpass)) > 1.0
    np.testing.assert_array_equal(on.envelope, off.envelope)


def test_sub_deemphasis_changes_only_the_video_channel():
    """Switching sub-deemphasis off changes Video and leaves the other channels
    untouched (the correction is applied after the video inverse transform)."""
    x = _lcg_u8_block().astype(np.float64)
    fl = filters.build_tape_filters(formats.get_format("PAL", "VHS", "lp"), 40e6, N)
    on = O.demod_block(x, fl)
    fl.sub_deemphasis = None
    off = O.demod_block(x, fl)
    assert not np.array_equal(on.video, off.video)
    np.testing.assert_array_equal(on.video_lowpass, off.video_lowpass)
    np.testing.assert_array_equal(on.envelope, off.envelope)


def test_ntsc_vhs_chroma_matches_upstream_hash(family):
    """NTSC VHS: the chroma channel (input-only filter) matches bit-exactly.  The
    RF-side channels depend on the order-16 band-pass transfer-function design,
    whose FP64 rounding differs from the reference's numpy/OpenBLAS emulation."""
    x = _lcg_u8_block().astype(np.float64)
    fl = filters.build_tape_filters(formats.get_format("NTSC", "VHS", "sp"), 40e6, N)
    r = O.demod_block(x, fl)
    assert _sha_f32(r.chroma) == _case(family, "NTSC", "VHS")["demod_burst"]


# ----------------------------------------------------------------- FFT and filters

def test_fft_matches_numpy():
    rng = np.random.default_rng(3)
    for n in (8, 1024, 32768):
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        ref = np.fft.fft(x)
        got = O.fft(x)
        assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))
        back = O.fft(got, inverse=True)
        assert np.max(np.abs(back - x)) < 1e-13


def test_sosfiltfilt_f32_matches_scipy():
    rng = np.random.default_rng(5)
    x = rng.standard_normal(5000) * 100.0
    sos = signal.butter(4, [0.01, 0.1], btype="bandpass", output="sos")
    got = O.sosfiltfilt_f32(sos, x)
    ref = signal.sosfiltfilt(sos, x.astype(np.float32).astype(np.float64), padtype="odd", padlen=27)
    assert np.max(np.abs(got - ref)) < 2e-3 * np.max(np.abs(ref))
    # first-order low-pass used by the envelope: pad length 6
    sos1 = filters._lowpass_first_order_sos(700e3 / 20e6)
    got = O.sosfiltfilt_f32(sos1, np.abs(x))
    ref = signal.sosfiltfilt(sos1, np.abs(x).astype(np.float32).astype(np.float64), padtype="odd", padlen=6)
    assert np.max(np.abs(got - ref)) < 1e-4 * np.max(np.abs(ref))


def test_vhs_atan2_properties():
    # the polynomial is a 1e-5-accurate atan2 on all four quadrants and both branches
    rng = np.random.default_rng(7)
    for _ in range(2000):
        y, x = rng.standard_normal(2) * 1000
        a = O.vhs_atan2(np.float32(y), np.float32(x))
        assert abs(a - np.arctan2(y, x)) < 2e-5
    assert O.vhs_atan2(0.0, 0.0) == 0.0


def test_discriminator_constant_tone():
    # a complex tone of frequency f demodulates to f (out[0] = 0), within the
    # polynomial's error scaled by fs / tau
    fs = 40e6
    for f in (3.8e6, 4.8e6, 10e6):
        n = 4096
        a = np.exp(2j * np.pi * f * np.arange(n) / fs) * 8000.0
        d = O.vhs_unwrap(a, fs)
        assert d[0] == 0.0
        assert np.max(np.abs(d[1:] - f)) < 200.0


@pytest.mark.parametrize("fs", [20e6, 40e6])
@pytest.mark.parametrize("system,fmt_name", [("PAL", "VHS"), ("NTSC", "VHS"), ("PAL", "SVHS"), ("PAL", "BETAMAX"),
                                             ("PAL", "VIDEO8")])
def test_filter_sets_demodulate_a_constant_carrier(system, fmt_name, fs):
    """A constant FM carrier demodulates to its own frequency through the whole block chain
    at either sample rate: the RF band-pass passes it, the discriminator returns it and the
    video (low-pass and de-emphasis) and sync-reference filters keep DC.  This pins the 20
    MSPS filter sets, which the reference's fixtures (40 MSPS) do not cover: a design at the
    wrong rate, a band-pass that cuts the carrier or a filter that does not keep DC fails."""
    fmt = formats.get_format(system, fmt_name, "sp")
    sp = fmt.sysparams
    fl = filters.build_tape_filters(fmt, fs, N)
    mid = slice(2048, N - 2048)
    for ire in (-30.0, 50.0, 100.0):
        f = float(sp["ire0"]) + ire * float(sp["hz_ire"])
        x = np.round(8000.0 * np.cos(2.0 * np.pi * f * np.arange(N) / fs))
        r = O.demod_block(x, fl)
        assert np.max(np.abs(r.video[mid] - f)) < 40.0, (ire, np.max(np.abs(r.video[mid] - f)))
        assert np.max(np.abs(r.video_lowpass[mid] - f)) < 10.0, (ire, np.max(np.abs(r.video_lowpass[mid] - f)))


# ----------------------------------------------------------------- pulses

def _pulse_signal():
    # sync_pulses_test.cu:27-55: two fields of 128 samples
    sig = []
    f0 = np.ones(128)
    for s, l in ((0, 6), (12, 4), (30, 7), (50, 20), (80, 21), (110, 18)):
        f0[s:s + l] = -1.0
    f1 = np.ones(128)
    f1[9:14] = -1.0
    f1[10] = np.nan
    f1[40:43] = -1.0
    f1[58] = np.nan
    f1[59] = np.nan
    f1[60:68] = -1.0
    return f0, f1


def test_find_pulses_reference_vectors():
    f0, f1 = _pulse_signal()
    st, ln = O.find_pulses(f0, 0.0, 4, 20)
    assert st.tolist() == [12, 30, 50] and ln.tolist() == [4, 7, 20]
    st, ln = O.find_pulses(f1, 0.0, 4, 20)
    assert st.tolist() == [9, 60] and ln.tolist() == [5, 8]


def test_find_pulses_brute_force():
    rng = np.random.default_rng(11)
    for trial in range(20):
        x = rng.standard_normal(3000)
        x[rng.integers(0, 3000, 30)] = np.nan
        thr = 0.3
        st, ln = O.find_pulses(x, thr, 2, 9)
        # brute force: maximal runs of (<= thr) where NaN continues the current run state
        pulses = []
        state = x[0] <= thr
        start = 0
        for i in range(len(x)):
            v = x[i]
            if np.isnan(v):
                continue
            if state and v > thr:
                if 2 <= i - start <= 9 and start != 0:
                    pulses.append((start, i - start))
                state = False
            elif not state and v <= thr:
                start, state = i, True
        assert list(zip(st.tolist(), ln.tolist())) == pulses


# ----------------------------------------------------------------- hsync refine

def test_hsync_refine_recovers_synthetic_edges():
    """Ideal sync edges filtered like the sync reference: refined locations land on
    the known half-amplitude crossing (closed form for a linear ramp)."""
    fs_mhz = 40.0
    spl = 2560
    lines = 20
    n = spl * (lines + 2)
    ire0, hz_ire = 4.1e6, 7000.0
    sync = ire0 + hz_ire * -42.857
    blank = ire0
    x = np.full(n, blank)
    true = []
    for l in range(lines + 1):
        s = 1000.25 + l * spl
        true.append(s)
        t = np.arange(n) - s
        # 20-sample linear edges down to sync, 188 samples of sync
        ramp_dn = np.clip(t / 20.0, 0, 1)
        ramp_up = np.clip((t - 188) / 20.0, 0, 1)
        x -= (blank - sync) * (ramp_dn - ramp_up)
    locs = np.array(true) + 3.0  # a rough estimate 3 samples late
    filled = np.zeros(locs.size, dtype=np.uint8)
    thr = ire0 + hz_ire * (-42.857 / 2)
    ref, fil = O.refine_hsync(x, locs, filled, thr, fs_mhz, 4.7, ire0, hz_ire, is_pal=True, use_right=False)
    good = [l for l in range(locs.size) if not ((3 <= l <= 6) or (1 <= l <= 2))]
    # the midpoint crossing of a linear 20-sample ramp sits 10 samples after its start
    np.testing.assert_allclose(ref[good], np.array(true)[good] + 10.0, atol=1e-9)
