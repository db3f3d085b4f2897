"""Parity of the HIP path (through the C ABI) with the CPU oracle on identical
seeded synthetic RF.  Runs on an MI355X only.

Bars:
 - integer outputs (pulse starts/lengths) identical;
 - FP64 stages fed identical input (hsync refinement, TBC resampling) bit-identical;
 - FP32 demod planes within the stated tolerances of the FP64 oracle;
 (the field sequence, and the full chain against the FP64 chain, are in
 tests/test_gpu_vhs_sync.py).
"""
import numpy as np
import pytest

from oracle import oracle as O
from vhsdecode_hip import filters, formats, synth
from vhsdecode_hip.demod import HipBlockDemodulator

pytestmark = pytest.mark.gpu

N = 32768
CUT = 1024
STRIDE = N - 2 * CUT

# FP32 tolerances (Hz unless stated); the TBC bar is +-1 LSB = 18.6 Hz (PAL VHS)
VIDEO_RMS_HZ = 2.0
VIDEO_MAX_HZ = 40.0
VLP_MAX_HZ = 20.0
ENV_REL = 2e-4
CHROMA_REL = 2e-3


def _dev(a):
    """Upload to HBM through the HIP runtime the library uses."""
    from vhsdecode_hip.hipmem import DeviceBuffer, device_count
    assert device_count() > 0, "GPU tests need a ROCm device"
    return DeviceBuffer.from_array(a)


def oracle_span(rf, fl, nblocks, first=0):
    planes = {k: [] for k in ("video", "vlp", "env", "chroma")}
    weak = []
    for b in range(nblocks):
        s = first + b * STRIDE
        r = O.demod_block(rf[s:s + N].astype(np.float64), fl)
        planes["video"].append(r.video[CUT:CUT + STRIDE])
        planes["vlp"].append(r.video_lowpass[CUT:CUT + STRIDE])
        planes["env"].append(r.envelope[CUT:CUT + STRIDE])
        planes["chroma"].append(r.chroma[CUT:CUT + STRIDE])
        weak.append(r.weak)
    return {k: np.concatenate(v) for k, v in planes.items()}, np.array(weak)


@pytest.fixture(scope="module", params=[("PAL", 1), ("NTSC", 2)], ids=["pal_vhs", "ntsc_vhs"])
def case(request):
    system, seed = request.param
    fmt = formats.get_format(system, "VHS", "sp")
    fl = filters.build_tape_filters(fmt, 40e6, N)
    tape = synth.generate(fmt, fields=1, seed=seed, snr_db=30.0 if system == "PAL" else 28.0,
                          jitter_ns=0.0 if system == "PAL" else 20.0)
    dem = HipBlockDemodulator(fl, max_blocks=64)
    nb = dem.blocks_for(tape.rf.size)
    gpu = dem.demod_span_host(tape.rf, 0, nb)
    ora, weak = oracle_span(tape.rf, fl, nb)
    print(f"\n[{system}] blocks={nb} span={gpu.video.size} flags={np.unique(gpu.block_flags)}", flush=True)
    return dict(system=system, fmt=fmt, fl=fl, tape=tape, dem=dem, gpu=gpu, ora=ora, weak=weak, nb=nb)


def test_video_plane_parity(case):
    g, o = case["gpu"], case["ora"]
    d = g.video_hz - o["video"]
    rms, mx = float(np.sqrt(np.mean(d * d))), float(np.max(np.abs(d)))
    print(f"[{case['system']}] video RMS {rms:.4f} Hz, max {mx:.3f} Hz", flush=True)
    assert rms < VIDEO_RMS_HZ and mx < VIDEO_MAX_HZ


def test_sync_reference_parity(case):
    g, o = case["gpu"], case["ora"]
    d = g.video_lowpass_hz - o["vlp"]
    print(f"[{case['system']}] vlp RMS {np.sqrt(np.mean(d * d)):.4f} Hz, max {np.max(np.abs(d)):.3f} Hz", flush=True)
    assert np.max(np.abs(d)) < VLP_MAX_HZ


def test_envelope_parity(case):
    g, o = case["gpu"], case["ora"]
    rel = np.max(np.abs(g.envelope - o["env"])) / np.max(np.abs(o["env"]))
    print(f"[{case['system']}] envelope max rel {rel:.3e}", flush=True)
    assert rel < ENV_REL
    assert bool((g.block_flags & 1).any()) == bool(case["weak"].any())


def test_chroma_parity(case):
    g, o = case["gpu"], case["ora"]
    rel = np.max(np.abs(g.chroma - o["chroma"])) / np.max(np.abs(o["chroma"]))
    print(f"[{case['system']}] chroma max rel {rel:.3e}", flush=True)
    assert rel < CHROMA_REL


def _thresholds(fmt):
    sp = fmt.sysparams
    thr = sp["ire0"] + sp["hz_ire"] * sp["vsync_ire"] / 2.0
    fs_mhz = 40.0
    min_len = max(0, int(np.ceil(sp["eqPulseUS"] * fs_mhz / 8.0)))
    max_len = max(1, int(sp["line_period"] * fs_mhz * 5.0))
    return thr, min_len, max_len


def test_pulses_identical(case):
    thr, mn, mx = _thresholds(case["fmt"])
    vlp = case["gpu"].video_lowpass
    dev = _dev(vlp)
    st, ln = case["dem"].find_pulses(dev.ptr, vlp.size, thr, mn, mx)
    # exactness of the detector on identical input
    ost, oln = O.find_pulses(case["gpu"].video_lowpass_hz, thr, mn, mx)
    assert st.tolist() == ost.tolist() and ln.tolist() == oln.tolist()
    # end to end: GPU chain vs FP64 oracle chain
    est, eln = O.find_pulses(case["ora"]["vlp"], thr, mn, mx)
    print(f"[{case['system']}] pulses gpu={len(st)} oracle={len(est)}", flush=True)
    assert st.tolist() == est.tolist() and ln.tolist() == eln.tolist()


def _line_locations(case):
    tape = case["tape"]
    span_len = case["gpu"].video.size
    locs = tape.line_starts - CUT
    locs = locs[(locs > 200) & (locs < span_len - 3000)]
    return locs


def test_hsync_refine_bit_identical(case):
    fmt = case["fmt"]
    sp = fmt.sysparams
    thr, _, _ = _thresholds(fmt)
    vlp = case["gpu"].video_lowpass
    locs = _line_locations(case) + 25.0
    filled = np.zeros(locs.size, dtype=np.uint8)
    dev = _dev(vlp)
    is_pal = case["system"] == "PAL"
    ref, fil = case["dem"].refine_hsync(dev.ptr, vlp.size, locs, filled, thr, 40.0, sp["hsyncPulseUS"],
                                        sp["ire0"], sp["hz_ire"], is_pal, True)
    oref, ofil = O.refine_hsync(case["gpu"].video_lowpass_hz, locs, filled, thr, 40.0, sp["hsyncPulseUS"], sp["ire0"],
                                sp["hz_ire"], is_pal, True)
    assert np.array_equal(ref.view(np.uint64), oref.view(np.uint64))
    assert np.array_equal(fil, ofil)
    # and the FP64-oracle chain's refinement agrees to a small fraction of a sample
    eref, _ = O.refine_hsync(case["ora"]["vlp"], locs, filled, thr, 40.0, sp["hsyncPulseUS"], sp["ire0"],
                             sp["hz_ire"], is_pal, True)
    print(f"[{case['system']}] hsync max |d| vs FP64 chain {np.max(np.abs(ref - eref)):.2e} samples", flush=True)
    assert np.max(np.abs(ref - eref)) < 0.01


def _conv(fmt):
    sp = fmt.sysparams
    is_pal = fmt.parent_system == "PAL"
    vsync = sp["vsync_ire"]
    scale = ((0xD300 - 0x0100) if is_pal else (0xC800 - 0x0400)) / (100.0 - vsync)
    return (sp["ire0"], sp["hz_ire"], int(sp["outputZero"]), vsync, scale)


def test_tbc_bit_identical_and_within_one_lsb(case, sinc_lut):
    fmt = case["fmt"]
    dem = case["dem"]
    dem.set_sinc_table(sinc_lut)
    video = case["gpu"].video
    locs = _line_locations(case)
    out_len = int(fmt.sysparams["outlinelen"])
    lines = locs.size - 1
    conv = _conv(fmt)
    dev = _dev(video)
    u16, f64 = dem.tbc_resample(dev.ptr, video.size, locs, 0, lines, out_len, conv=conv)
    ou16_same = O.tbc_resample(case["gpu"].video_hz, locs, 0, lines, out_len, sinc_lut, conv=conv)
    assert np.array_equal(u16, ou16_same[1])
    assert np.array_equal(f64.view(np.uint64), ou16_same[0].view(np.uint64))
    ou16 = O.tbc_resample(case["ora"]["video"], locs, 0, lines, out_len, sinc_lut, conv=conv)[1]
    dl = np.abs(u16.astype(np.int64) - ou16.astype(np.int64))
    print(f"[{case['system']}] TBC luma max |d| {dl.max()} LSB, frac>0 {np.mean(dl > 0):.2e}", flush=True)
    assert dl.max() <= 1


def test_fft16k_matches_numpy(case):
    from vhsdecode_hip.hipmem import synchronize
    rng = np.random.default_rng(0)
    x = (rng.standard_normal((3, 16384)) + 1j * rng.standard_normal((3, 16384))).astype(np.complex64)
    dev = _dev(x.view(np.float32).copy())
    case["dem"].debug_fft16k(dev.ptr, 3, inverse=False)
    synchronize()
    got = dev.download(np.float32, 2 * 3 * 16384).view(np.complex64).reshape(3, 16384)
    ref = np.fft.fft(x.astype(np.complex128), axis=1)
    err = np.sqrt(np.mean(np.abs(got - ref) ** 2)) / np.sqrt(np.mean(np.abs(ref) ** 2))
    print(f"fft16k relative RMS error {err:.2e}", flush=True)
    assert err < 1e-6
