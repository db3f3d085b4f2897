"""The uneven batched launch of 64 segments (seeded lengths 16384..262144) of PAL and NTSC
VHS SP, PAL S-VHS, PAL Betamax and PAL Video8 at 20 and 40 MSPS (ten demod contexts, one
graph-captured batch) against the FP64 oracle: every block of every segment, plane by
plane, within the FP32 tolerances of the block chain (video RMS 2 Hz and max 40 Hz, sync
reference max 20 Hz, envelope and chroma relative 2e-4 and 2e-3 of the plane's range).
Segments shorter than one block decode nothing, as the stream decoder reads them.

A block whose unrepaired discriminator output exceeds the diff-demod repair limit
(RfDemodulator.ApplyDiffDemodRepairIfPresent) is repaired by windows placed on the samples
above the limit.  Here that happens where a segment's random start puts a discontinuity at
a block edge: the analytic signal rings there, and which edge samples cross the limit
depends on FP32 against FP64 rounding; the replaced values then spread through the
circular video filter over the whole block.  Such blocks (counted, and required to be
rare) are held to the envelope and chroma bars only; the repair itself is bit-exact to
the oracle on identical input (tests/test_gpu_repair.py).  Runs on an MI355X only."""
import copy
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from oracle import oracle as O
from vhsdecode_hip import filters, formats, hipmem, synth
from vhsdecode_hip.batch import HipBatch, pack_segments
from vhsdecode_hip.demod import HipBlockDemodulator

pytestmark = pytest.mark.gpu

N, CUT = 32768, 1024
STRIDE = N - 2 * CUT
MIX = [("PAL", "VHS", "sp"), ("NTSC", "VHS", "sp"), ("PAL", "SVHS", "sp"), ("PAL", "BETAMAX", "sp"),
       ("PAL", "VIDEO8", "sp")]
VIDEO_RMS_HZ, VIDEO_MAX_HZ, VLP_MAX_HZ, ENV_REL, CHROMA_REL = 2.0, 40.0, 20.0, 2e-4, 2e-3


def _raw_spikes(x, fl, lim):
    """Positions where the oracle's demod before the diff-demod repair (fl: the filter set
    with the repair off) exceeds the limit."""
    raw = O.demod_block(x, fl).demod
    return np.flatnonzero(raw[20:N - 20] > lim) + 20


@pytest.fixture(scope="module")
def config3():
    rng = np.random.default_rng(20240501)
    lengths = rng.integers(16384, 262145, size=64)
    kinds = [(m, fs) for m in MIX for fs in (20e6, 40e6)]
    dems, tapes, fls = [], [], []
    for i, ((system, name, speed), fs) in enumerate(kinds):
        fmt = formats.get_format(system, name, speed)
        fl = filters.build_tape_filters(fmt, fs, N)
        fls.append(fl)
        dems.append(HipBlockDemodulator(fl, max_blocks=16))
        tapes.append(synth.generate(fmt, fields=1, sample_rate_hz=fs, seed=20240501 + i, snr_db=30.0).rf)
    choice = rng.integers(0, len(kinds), size=lengths.size)
    segs = pack_segments(lengths, [dems[c] for c in choice])
    rf_host = np.zeros(segs[-1].rf_offset + int(lengths[-1]) + 2, dtype=np.int16)
    for s, c in zip(segs, choice):
        t = tapes[c]
        start = 2 * int(rng.integers(0, (t.size - s.length) // 2))
        rf_host[s.rf_offset:s.rf_offset + s.length] = t[start:start + s.length]
    rf = hipmem.DeviceBuffer.from_array(rf_host)
    with HipBatch(segs, use_graph=True) as b:
        n = b.output_samples
        planes = hipmem.DeviceBuffer(4 * n * 4)
        planes.fill(0xFF)
        b.run_device(rf.ptr, rf_host.size, planes.ptr, planes.ptr + 4 * n, planes.ptr + 8 * n, planes.ptr + 12 * n)
        hipmem.synchronize()
        got = planes.download(np.float32, 4 * n).reshape(4, n)
        blocks = b.blocks
    live = [(s, c) for s, c in zip(segs, choice) if s.blocks() > 0]
    # the same filter sets with the repair off (copies: the oracle runs on several threads)
    unrepaired = []
    for fl in fls:
        u = copy.copy(fl)
        u.diff_demod_max = None
        unrepaired.append(u)
    jobs = [(k, s, c, b) for k, (s, c) in enumerate(live) for b in range(s.blocks())]

    def one(job):
        k, s, c, b = job
        start = s.rf_offset + b * STRIDE
        x = rf_host[start:start + N].astype(np.float64)
        r = O.demod_block(x, fls[c])
        # the unrepaired demod: where it exceeds the repair limit
        lim = fls[c].diff_demod_max
        spikes = np.zeros(0, dtype=np.int64)
        if lim is not None:
            spikes = _raw_spikes(x, unrepaired[c], lim)
        return k, b, r, spikes

    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        ora = list(ex.map(one, jobs))
    print(f"\n[config 3] segments={len(segs)} live={len(live)} blocks={blocks} contexts={len(kinds)}", flush=True)
    return dict(kinds=kinds, live=live, got=got, ora=ora, dems=dems, keep=(rf, planes))


def test_every_block_of_every_segment_matches_the_oracle(config3):
    got, kinds = config3["got"], config3["kinds"]
    worst = {}
    spike_blocks = 0
    for k, b, r, spikes in config3["ora"]:
        s, c = config3["live"][k]
        o = s.out_offset + b * STRIDE
        bias = s.demodulator.bias
        name = f"{kinds[c][0][0]} {kinds[c][0][1]} {kinds[c][1] / 1e6:.0f}"
        dv = got[0, o:o + STRIDE].astype(np.float64) + bias - r.video[CUT:CUT + STRIDE]
        dl = got[1, o:o + STRIDE].astype(np.float64) + bias - r.video_lowpass[CUT:CUT + STRIDE]
        repaired = spikes.size > 0
        if repaired:
            spike_blocks += 1
            print(f"  {name} segment {k} block {b}: {spikes.size} samples over the repair limit "
                  f"(first {spikes[:3].tolist()}, last {spikes[-3:].tolist()}); video RMS "
                  f"{np.sqrt(np.mean(dv * dv)):.1f} Hz, not held to the plane bars", flush=True)
            dv = dl = np.zeros(1)
        oe = r.envelope[CUT:CUT + STRIDE]
        de = np.max(np.abs(got[2, o:o + STRIDE] - oe)) / np.max(np.abs(oe))
        oc = r.chroma[CUT:CUT + STRIDE]
        dc = np.max(np.abs(got[3, o:o + STRIDE] - oc)) / np.max(np.abs(oc))
        rms, mx, lmx = float(np.sqrt(np.mean(dv * dv))), float(np.max(np.abs(dv))), float(np.max(np.abs(dl)))
        w = worst.setdefault(name, [0.0, 0.0, 0.0, 0.0, 0.0, 0])
        for i, v in enumerate((rms, mx, lmx, de, dc)):
            w[i] = max(w[i], v)
        w[5] += 1
        assert rms < VIDEO_RMS_HZ and mx < VIDEO_MAX_HZ, (name, k, b, rms, mx)
        assert lmx < VLP_MAX_HZ, (name, k, b, lmx)
        assert de < ENV_REL, (name, k, b, de)
        assert dc < CHROMA_REL, (name, k, b, dc)
    print(f"  repaired blocks: {spike_blocks} of {len(config3['ora'])}", flush=True)
    assert spike_blocks <= 0.02 * len(config3["ora"])
    for name, w in sorted(worst.items()):
        print(f"  {name} MSPS: {w[5]} blocks, video RMS {w[0]:.3f} max {w[1]:.2f} Hz, sync ref max {w[2]:.2f} Hz, "
              f"envelope rel {w[3]:.1e}, chroma rel {w[4]:.1e}", flush=True)
    assert len(worst) == len(kinds)
