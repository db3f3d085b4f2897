"""Batched decode of uneven segments of mixed formats and sample rates
(vhsdecode_hip_batch_*): every segment's planes equal, bit for bit, what the
per-segment span path (vhsdecode_hip_demod_span_device, whose parity with the oracle
the other GPU tests pin) writes for the same samples; graph replays and the
ungraphed batch give the same bits.  Runs on an MI355X only."""
import numpy as np
import pytest

from vhsdecode_hip import filters, formats, hipmem, ld_filters, synth
from vhsdecode_hip.batch import HipBatch, pack_segments
from vhsdecode_hip.demod import HipBlockDemodulator

pytestmark = pytest.mark.gpu

N = 32768
MIX = [("PAL", "VHS", "sp", 40e6), ("NTSC", "VHS", "sp", 20e6), ("PAL", "SVHS", "sp", 40e6),
       ("PAL", "BETAMAX", "sp", 20e6), ("PAL", "VIDEO8", "sp", 40e6), ("NTSC", "VHS", "sp", 40e6),
       ("NTSC", "LD", "", 40e6)]


@pytest.fixture(scope="module")
def batch_case():
    rng = np.random.default_rng(7)
    lengths = rng.integers(16384, 262145, size=20)
    dems, tapes = [], []
    for i, (system, fmt_name, speed, fs) in enumerate(MIX):
        if fmt_name == "LD":
            fmt = formats.get_laserdisc_format(system)
            fl = ld_filters.build_laserdisc_filters(fmt, fs, N, analog_audio=False)
        else:
            fmt = formats.get_format(system, fmt_name, speed)
            fl = filters.build_tape_filters(fmt, fs, N)
        dems.append(HipBlockDemodulator(fl, max_blocks=16))
        tapes.append(synth.generate(fmt, fields=1, sample_rate_hz=fs, seed=20 + i, snr_db=30.0).rf)
    pick = [dems[i % len(MIX)] for i in range(lengths.size)]
    segs = pack_segments(lengths, pick)
    rf = np.zeros(segs[-1].rf_offset + lengths[-1] + 2, dtype=np.int16)
    for i, s in enumerate(segs):
        tape = tapes[i % len(MIX)]
        start = 2 * int(rng.integers(0, (tape.size - s.length) // 2))
        rf[s.rf_offset:s.rf_offset + s.length] = tape[start:start + s.length]
    print(f"\n[batch] segments={len(segs)} blocks={sum(s.blocks() for s in segs)} "
          f"short={sum(1 for s in segs if s.blocks() == 0)}", flush=True)
    return dict(segs=segs, rf=rf, dems=dems)


def _run_batch(case, use_graph, runs=1):
    rf = hipmem.DeviceBuffer.from_array(case["rf"])
    with HipBatch(case["segs"], use_graph=use_graph) as b:
        n = b.output_samples
        planes = hipmem.DeviceBuffer(4 * n * 4)
        outs = []
        for _ in range(runs):
            planes.fill(0xFF)
            b.run_device(rf.ptr, case["rf"].size, planes.ptr, planes.ptr + 4 * n, planes.ptr + 8 * n,
                         planes.ptr + 12 * n)
            hipmem.synchronize()
            outs.append(planes.download(np.float32, 4 * n).reshape(4, n))
        return b.blocks, outs


def test_batch_matches_per_segment_span_path(batch_case):
    blocks, (got,) = _run_batch(batch_case, use_graph=True)
    rf = hipmem.DeviceBuffer.from_array(batch_case["rf"])
    checked = 0
    for s in batch_case["segs"]:
        nb = s.blocks()
        if nb == 0:
            continue
        d = s.demodulator
        n = nb * d.stride
        planes = hipmem.DeviceBuffer(4 * n * 4)
        flags = hipmem.DeviceBuffer(nb * 4)
        d.demod_span_device(rf.ptr, s.rf_offset, nb, planes.ptr, planes.ptr + 4 * n, planes.ptr + 8 * n,
                            planes.ptr + 12 * n, flags.ptr)
        hipmem.synchronize()
        ref = planes.download(np.float32, 4 * n).reshape(4, n)
        has_chroma = (d.filters.video_burst is not None) if hasattr(d.filters, "rf_mtf") else \
            (d.filters.chroma_sos is not None)
        for k in range(4 if has_chroma else 3):
            np.testing.assert_array_equal(got[k, s.out_offset:s.out_offset + n], ref[k])
        checked += nb
    assert checked == blocks > 0
    print(f"blocks checked {checked}")


def test_graph_replays_and_ungraphed_batch_agree(batch_case):
    _, graph_runs = _run_batch(batch_case, use_graph=True, runs=3)
    _, (plain,) = _run_batch(batch_case, use_graph=False)
    for r in graph_runs[1:]:
        np.testing.assert_array_equal(r, graph_runs[0])
    np.testing.assert_array_equal(plain, graph_runs[0])


def test_overlapping_outputs_are_rejected(batch_case):
    segs = batch_case["segs"]
    a, b = segs[0], segs[1]
    bad = [a, type(b)(b.demodulator, b.rf_offset, b.length, a.out_offset)]
    if a.blocks() == 0 or b.blocks() == 0:
        pytest.skip("needs two segments with blocks")
    with pytest.raises(Exception):
        HipBatch(bad)
