"""Diff-demod spike repair on the device (vhs_block_repair_kernel: spike list by a
workgroup-wide compaction, then one wave walking the windows in order) against the
oracle's restatement of RfDemodulator.ApplyDiffDemodRepairIfPresent / ReplaceSpikes
(RfDemodulator.cs:2527-2581, 966-1001).  Each block carries an RF dropout (a stretch of
low-level noise): the discriminator output there is spread over the whole band, far
above the repair threshold, so the sequential window logic runs on thousands of
spikes.  Runs on an MI355X only."""
import numpy as np
import pytest

from oracle import oracle as O
from vhsdecode_hip import filters, formats, hipmem, synth
from vhsdecode_hip.demod import HipBlockDemodulator

pytestmark = pytest.mark.gpu

N = 32768
CUT = 1024
STRIDE = N - 2 * CUT


@pytest.fixture(scope="module", params=[("NTSC", "VHS", 40e6, 31), ("PAL", "SVHS", 20e6, 32)],
                ids=["ntsc_vhs_40msps", "pal_svhs_20msps"])
def case(request):
    system, name, fs, seed = request.param
    fmt = formats.get_format(system, name, "sp")
    fl = filters.build_tape_filters(fmt, fs, N)
    tape = synth.generate(fmt, fields=1, sample_rate_hz=fs, seed=seed, snr_db=30.0)
    nb = 6
    rf = tape.rf[: N + (nb - 1) * STRIDE].copy()
    rng = np.random.default_rng(seed)
    drops = []
    for b in range(nb):
        at = b * STRIDE + 6000 + 4000 * b
        rf[at:at + 2500] = np.rint(rng.normal(0.0, 40.0, 2500)).astype(np.int16)
        drops.append(at)
    with HipBlockDemodulator(fl, max_blocks=nb) as dem:
        gpu = dem.demod_span_host(rf, 0, nb)
    video = []
    for b in range(nb):
        blk = rf[b * STRIDE:b * STRIDE + N].astype(np.float64)
        video.append(O.demod_block(blk, fl).video[CUT:CUT + STRIDE])
    print(f"\n[repair {system} {name} {fs / 1e6:.0f} MSPS] blocks={nb} flags={gpu.block_flags.tolist()}", flush=True)
    # kept-window sample n is RF sample n + CUT; mask each dropout with 4096 samples either side
    keep = np.ones(nb * STRIDE, dtype=bool)
    for at in drops:
        keep[max(0, at - CUT - 4096):max(0, at - CUT + 2500 + 4096)] = False
    return dict(gpu=gpu, video=np.concatenate(video), keep=keep, fl=fl)


def test_repair_runs(case):
    assert np.count_nonzero(case["gpu"].block_flags & 2) >= 1


def test_video_with_repair(case):
    """Inside a dropout the carrier is ~46 dB down and its FP32 and FP64 demods differ
    sample by sample; away from the dropouts the planes agree as elsewhere."""
    err = case["gpu"].video_hz - case["video"]
    k = case["keep"]
    rms_out = float(np.sqrt(np.mean(err[k] ** 2)))
    rms_all = float(np.sqrt(np.mean(err ** 2)))
    print(f"video RMS away from dropouts {rms_out:.3f} Hz (max {np.max(np.abs(err[k])):.2f}), overall {rms_all:.2f} Hz")
    assert rms_out <= 1.0
    assert rms_all <= 50.0


def test_replace_spikes_matches_oracle(case):
    """The kernel's ReplaceSpikes step (one wave walking the ordered spike list over the
    demod in LDS) on chosen inputs equals RfDemodulator.ReplaceSpikes bit for bit:
    windows overlap, spikes cluster, and both branches of the window test occur."""
    rng = np.random.default_rng(5)
    mx = 1.0e6
    dm = rng.uniform(0.0, 0.9e6, N).astype(np.float32)
    spikes = np.concatenate([rng.integers(0, N, 900), 5000 + np.arange(0, 400, 3), [0, 1, N - 2, N - 1]])
    dm[spikes] = rng.uniform(1.0e6, 4.0e6, spikes.size).astype(np.float32)
    dd = rng.uniform(0.0, 3.0e6, N).astype(np.float32)
    want = O.replace_spikes(dm.astype(np.float64), dd.astype(np.float64), mx).astype(np.float32)
    fl = case["fl"]
    with HipBlockDemodulator(fl, max_blocks=1) as dem:
        a = hipmem.DeviceBuffer.from_array(dm)
        b = hipmem.DeviceBuffer.from_array(dd)
        o = hipmem.DeviceBuffer(4 * N)
        from vhsdecode_hip import _lib
        import ctypes
        _lib.check(dem._L.vhsdecode_hip_debug_replace_spikes(dem._h, ctypes.c_void_p(a.ptr), ctypes.c_void_p(b.ptr),
                                                             mx, ctypes.c_void_p(o.ptr), None))
        hipmem.synchronize()
        got = o.download(np.float32, N)
    changed = int(np.count_nonzero(want != dm))
    print(f"spikes {int(np.count_nonzero(dm > mx))}, samples replaced {changed}")
    assert changed > 1000
    np.testing.assert_array_equal(got, want)
