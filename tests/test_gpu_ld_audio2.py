"""LaserDisc analog audio stage 2 on the HIP device (vhsdecode_hip_ld_audio_phase2_device,
through the C ABI) against the oracle's restatement of LaserDiscAnalogAudioPhase2.Apply
(oracle/ld_audio.py), with the 40 MSPS stage-2 responses of NTSC and PAL.

Inputs: the reference's stage-2 test channel (LaserDiscAudioNumericsCompatibilityTests.cs:
203-217: float32 carrier plus a 0.25 Hz-step pseudo-random offset) at lengths that take one
block, a first and a last block, and middle blocks too, with planted left-channel peaks
above 500 kHz so the suppression of both channels is exercised.  The device transforms in
FP32 where the reference's short blocks use FP64, and both store float32 near 2.3-2.8 MHz
(a quarter-hertz step), so the bar is a float32 ulp or two at the carrier: max |device -
oracle| <= 1 Hz, RMS <= 0.25 Hz.  Without the suppression the planted peaks alone would
move the output by kilohertz.  Runs on an MI355X only."""
import numpy as np
import pytest

from oracle import ld_audio as LA
from vhsdecode_hip import formats, hipmem, ld_filters
from vhsdecode_hip.demod import HipBlockDemodulator

pytestmark = pytest.mark.gpu

MAX_HZ = 1.0
RMS_HZ = 0.25


def _channel(center, multiplier, length):
    i = np.arange(length, dtype=np.int64)
    off = ((((i * multiplier) % 200_003) - 100_001).astype(np.float32) * np.float32(0.25)).astype(np.float32)
    return (np.float32(center) + off).astype(np.float32)


@pytest.fixture(scope="module", params=["NTSC", "PAL"])
def ctx(request):
    fl = ld_filters.build_laserdisc_filters(formats.get_laserdisc_format(request.param), 40e6, 32768)
    with HipBlockDemodulator(fl, max_blocks=4) as dem:
        dem.set_ld_audio_phase2(fl)
        yield request.param, fl, dem


@pytest.mark.parametrize("count", [20_000, 34_685, 100_000])
def test_stage2_matches_oracle(ctx, count):
    system, fl, dem = ctx
    a, b = fl.audio_left, fl.audio_right
    left = _channel(a.center_hz, 7919, count)
    right = _channel(b.center_hz, 15427, count)
    rng = np.random.default_rng(count)
    for p in rng.choice(np.arange(40, count - 40), size=12, replace=False):
        left[p] = np.float32(a.center_hz + 650_000.0)
        left[p + 1] = np.float32(a.center_hz + 700_000.0)
    print(f"\n[ld audio 2] {system} count={count}", flush=True)
    want_l, want_r, _ = LA.phase2(left.astype(np.float64), right.astype(np.float64), a.stage2, b.stage2,
                                  a.center_hz, b.center_hz)
    dl, dr = hipmem.DeviceBuffer.from_array(left), hipmem.DeviceBuffer.from_array(right)
    ol, orr = hipmem.DeviceBuffer(4 * count), hipmem.DeviceBuffer(4 * count)
    dem.ld_audio_phase2_device(dl.ptr, dr.ptr, count, ol.ptr, orr.ptr)
    got_l = ol.download(np.float32, count).astype(np.float64)
    got_r = orr.download(np.float32, count).astype(np.float64)
    for name, g, w in (("left", got_l, want_l), ("right", got_r, want_r)):
        d = np.abs(g - w)
        rms = float(np.sqrt(np.mean(d ** 2)))
        print(f"[ld audio 2] {system} {count} {name}: max {d.max():.3f} Hz rms {rms:.4f} Hz", flush=True)
        assert d.max() <= MAX_HZ and rms <= RMS_HZ


def test_stage2_rejects_overlaid_output(ctx):
    _, _, dem = ctx
    buf = hipmem.DeviceBuffer(4 * 1000)
    other = hipmem.DeviceBuffer(4 * 1000)
    with pytest.raises(Exception):
        dem.ld_audio_phase2_device(buf.ptr, other.ptr, 1000, buf.ptr, other.ptr)
