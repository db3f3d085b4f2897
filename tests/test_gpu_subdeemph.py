"""Sub-deemphasis on the HIP path (vhs_block_subdeemph_kernel, through the C ABI)
against the oracle's restatement of RfDemodulator.ApplySubDeemphasisRealCore
(Dsp/RfDemodulator.cs:2897-2999), which the upstream fixture hashes pin
(tests/test_oracle.py).  Runs on an MI355X only."""
import numpy as np
import pytest

from oracle import oracle as O
from vhsdecode_hip import filters, formats, synth
from vhsdecode_hip.demod import HipBlockDemodulator

pytestmark = pytest.mark.gpu

N = 32768
CUT = 1024
STRIDE = N - 2 * CUT
VIDEO_RMS_HZ = 2.0
VIDEO_MAX_HZ = 40.0


@pytest.fixture(scope="module", params=[("PAL", "VHS", "lp", 3), ("NTSC", "VHS", "ep", 4), ("PAL", "SVHS", "sp", 5),
                                        ("NTSC", "VIDEO8", "sp", 6), ("PAL", "BETAMAX", "sp", 7),
                                        ("NTSC", "BETAMAX", "lp", 8)],
                ids=["pal_vhs_lp", "ntsc_vhs_ep", "pal_svhs_sp", "ntsc_video8_sp", "pal_betamax_sp", "ntsc_betamax_lp"])
def case(request):
    system, tape_format, speed, seed = request.param
    fmt = formats.get_format(system, tape_format, speed)
    fl = filters.build_tape_filters(fmt, 40e6, N)
    assert fl.sub_deemphasis is not None
    tape = synth.generate(fmt, fields=1, seed=seed, snr_db=30.0)
    rf = tape.rf[: N + 11 * STRIDE]
    nb = 12
    with HipBlockDemodulator(fl, max_blocks=nb) as dem:
        gpu = dem.demod_span_host(rf, 0, nb)
    on, off, chroma, vlp = [], [], [], []
    for b in range(nb):
        blk = rf[b * STRIDE:b * STRIDE + N].astype(np.float64)
        r = O.demod_block(blk, fl)
        on.append(r.video[CUT:CUT + STRIDE])
        chroma.append(r.chroma[CUT:CUT + STRIDE])
        vlp.append(r.video_lowpass[CUT:CUT + STRIDE])
    sd = fl.sub_deemphasis
    fl.sub_deemphasis = None
    for b in range(nb):
        blk = rf[b * STRIDE:b * STRIDE + N].astype(np.float64)
        off.append(O.demod_block(blk, fl).video[CUT:CUT + STRIDE])
    fl.sub_deemphasis = sd
    print(f"\n[{system} {tape_format} {speed}] blocks={nb}", flush=True)
    return dict(gpu=gpu, on=np.concatenate(on), off=np.concatenate(off), chroma=np.concatenate(chroma),
                vlp=np.concatenate(vlp), sections=fl.chroma_sos.shape[0], boost=fl.rf_high_boost is not None)


def test_video_plane_with_sub_deemphasis(case):
    err = case["gpu"].video_hz - case["on"]
    rms = float(np.sqrt(np.mean(err ** 2)))
    print(f"video RMS {rms:.3f} Hz, max {np.max(np.abs(err)):.2f} Hz")
    assert rms <= VIDEO_RMS_HZ
    assert np.max(np.abs(err)) <= VIDEO_MAX_HZ


def test_sub_deemphasis_is_applied(case):
    """The correction is far larger than the FP32 error bar: the GPU plane is
    close to the sub-deemphasised oracle and far from the plain one."""
    corr = case["on"] - case["off"]
    err_off = case["gpu"].video_hz - case["off"]
    rms_corr = float(np.sqrt(np.mean(corr ** 2)))
    print(f"sub-deemphasis correction RMS {rms_corr:.1f} Hz")
    assert rms_corr > 10 * VIDEO_RMS_HZ
    assert float(np.sqrt(np.mean(err_off ** 2))) > 5 * VIDEO_RMS_HZ


def test_chroma_plane(case):
    """Chroma band-passes of 4 sections and of Video8's 10 sections (scan filter with
    20-state affine maps)."""
    ref = case["chroma"].astype(np.float64)
    err = np.abs(case["gpu"].chroma.astype(np.float64) - ref)
    rel = float(err.max() / np.abs(ref).max())
    print(f"chroma ({case['sections']} sections) max rel {rel:.3e}")
    assert rel <= 2e-3


def test_sync_reference_plane(case):
    """The sync reference follows the demod (RF-boosted for Betamax)."""
    err = np.abs(case["gpu"].video_lowpass_hz - case["vlp"])
    print(f"sync reference max {err.max():.3f} Hz (boost {case['boost']})")
    assert err.max() <= 20.0
