"""LaserDisc complex path on the HIP device (vhs_block_demod_ld_kernel and the video
kernel's reference passes, through the C ABI) against the oracle's restatement of
RfDemodulator.DemodulateCore with the conjugate-product discriminator and of
RfBlockDecodePipeline.DecodeAnalogAudioChannel, which the reference's fixture hashes pin
(tests/test_laserdisc_oracle.py).  Runs on an MI355X only."""
import numpy as np
import pytest

from oracle import oracle as O
from vhsdecode_hip import formats, ld_filters, synth
from vhsdecode_hip.demod import HipBlockDemodulator

pytestmark = pytest.mark.gpu

N = 32768
CUT = 1024
STRIDE = N - 2 * CUT
# measured on MI355X: video / sync reference RMS 0.15-0.26 Hz (max 0.9 Hz), burst and
# pilot RMS <= 0.05 Hz, audio RMS <= 0.07 Hz, envelope max rel 6e-7
VIDEO_RMS_HZ = 1.0
VIDEO_MAX_HZ = 5.0
REF_RMS_HZ = 0.5
AUDIO_RMS_HZ = 1.0
ENV_REL = 1e-5


@pytest.fixture(scope="module", params=[("NTSC", 5), ("PAL", 6)], ids=["ntsc_ld", "pal_ld"])
def case(request):
    system, seed = request.param
    fmt = formats.get_laserdisc_format(system)
    fl = ld_filters.build_laserdisc_filters(fmt, 40e6, N)
    sp = fmt.sysparams
    car = ((float(sp["audio_lfreq"]), 100e3, 1000.0, 0.1), (float(sp["audio_rfreq"]), 100e3, 3000.0, 0.1))
    tape = synth.generate(fmt, fields=1, seed=seed, snr_db=30.0, audio_carriers=car)
    nb = 10
    rf = tape.rf[: N + (nb - 1) * STRIDE]
    with HipBlockDemodulator(fl, max_blocks=nb) as dem:
        gpu = dem.demod_span_host(rf, 0, nb)
    keys = ("video", "video_lowpass", "envelope", "video_burst", "video_pilot", "left", "right")
    ref = {k: [] for k in keys}
    dec = fl.audio_decimation
    for b in range(nb):
        blk = rf[b * STRIDE:b * STRIDE + N].astype(np.float64)
        r = O.demod_block_ld(blk, fl)
        left, right = O.ld_audio_block(blk, fl)
        for k in keys[:5]:
            v = getattr(r, k)
            if v is not None:
                ref[k].append(v[CUT:CUT + STRIDE])
        ref["left"].append(left[CUT // dec:(CUT + STRIDE) // dec])
        ref["right"].append(right[CUT // dec:(CUT + STRIDE) // dec])
    print(f"\n[LD {system}] blocks={nb}", flush=True)
    return dict(system=system, gpu=gpu, ref={k: np.concatenate(v) if v else None for k, v in ref.items()})


def test_video_plane(case):
    err = case["gpu"].video_hz - case["ref"]["video"]
    rms = float(np.sqrt(np.mean(err ** 2)))
    print(f"video RMS {rms:.3f} Hz, max {np.max(np.abs(err)):.2f} Hz")
    assert rms <= VIDEO_RMS_HZ
    assert np.max(np.abs(err)) <= VIDEO_MAX_HZ


def test_sync_reference_plane(case):
    err = case["gpu"].video_lowpass_hz - case["ref"]["video_lowpass"]
    rms = float(np.sqrt(np.mean(err ** 2)))
    print(f"sync reference RMS {rms:.3f} Hz, max {np.max(np.abs(err)):.2f} Hz")
    assert rms <= VIDEO_RMS_HZ
    assert np.max(np.abs(err)) <= VIDEO_MAX_HZ


def test_envelope_is_analytic_magnitude(case):
    ref = case["ref"]["envelope"]
    err = np.abs(case["gpu"].envelope.astype(np.float64) - ref)
    rel = float(err.max() / np.abs(ref).max())
    print(f"envelope max rel {rel:.3e}")
    assert rel <= ENV_REL


def test_burst_and_pilot_references(case):
    planes = [("burst", case["gpu"].chroma, case["ref"]["video_burst"])]
    if case["system"] == "PAL":
        planes.append(("pilot", case["gpu"].pilot, case["ref"]["video_pilot"]))
    else:
        assert case["gpu"].pilot is None
    for name, got, ref in planes:
        assert got is not None
        err = got.astype(np.float64) - ref
        rms = float(np.sqrt(np.mean(err ** 2)))
        print(f"{name} RMS {rms:.3f} Hz (signal RMS {np.sqrt(np.mean(ref ** 2)):.0f} Hz)")
        assert rms <= REF_RMS_HZ


def test_analog_audio(case):
    for ch in ("left", "right"):
        got = getattr(case["gpu"], "audio_" + ch).astype(np.float64)
        ref = case["ref"][ch]
        assert got.shape == ref.shape
        err = got - ref
        rms = float(np.sqrt(np.mean(err ** 2)))
        print(f"audio {ch} RMS {rms:.3f} Hz, max {np.max(np.abs(err)):.1f} Hz")
        assert rms <= AUDIO_RMS_HZ
