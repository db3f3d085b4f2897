#!/usr/bin/env python3
"""RF MSPS decoded (demod + TBC) on one MI355X.

Workload: synthetic NTSC VHS at 40 MSPS int16, 10 fields (525-line / 63.556 us,
SMPTE-like luma steps, 3.579545 MHz burst, FM 3.4 -> 4.4 MHz, 28 dB SNR, 1/f line
jitter sigma 20 ns), seeded.  One step decodes the whole tape resident in HBM:
  block demod of every 32768-sample overlap-save block (one launch) + chroma, the
  vhs field sequence (read windows, sync levels, pulses, parity, line 0, line
  locations), batched FP64 hsync refinement, wow level adjusts and TBC resampling of
  every field to uint16, and the RF-envelope dropout detection, all on the device with
  no host round trip inside the step.  The steps rotate over copies of the tape at
  different HBM addresses, so RF reads are not served by the last-level cache.
value = RF input samples per step * steps / wall time (MSPS).

The roofline object describes the dominant per-block kernel, timed with HIP events on
its launch stream with one span in flight (block_chain adds up the per-block kernels);
cpu_baseline is the FP64 oracle (oracle/) decoding the same tape on the host cores, and
`parity` compares the step's fields with that FP64 chain.
`serving` reports the served-window path (vhsdecode_hip_serve_*): p50/p99 latency of
2-second windows from 16 streams at 8 requests/s, and the capacity of a burst.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "vhs-decode-dotnet_amd"))
sys.path.insert(0, ROOT)

N = 32768
CUT = 1024
STRIDE = N - 2 * CUT
HBM_PEAK_GBS = 8000.0
DROPOUT_ROWS = 1024      # .tbc.json dropout rows kept per field
SEED = 20240501


def build_workload(fields: int, chroma: bool = True):
    from vhsdecode_hip import filters, formats, synth
    fmt = formats.get_format("NTSC", "VHS", "sp")
    # the colour-under chroma recorded with the track-phase rotation when the step decodes it
    tape = synth.generate(fmt, fields=fields, seed=SEED, snr_db=28.0, jitter_ns=20.0,
                          chroma_under="modulated" if chroma else "carrier")
    fl = filters.build_tape_filters(fmt, 40e6, N)
    nb = (tape.rf.size - N) // STRIDE + 1
    return fmt, tape, fl, nb


def cpu_baseline(fmt, tape, fl, nb, lut, seconds: float, threads: int, chroma: bool = True):
    """FP64 oracle on the same tape: block demod (threads over blocks), the vhs field
    sequence (read windows, levels, pulses, parity, line 0, line locations; serial, as the
    reference's sequence is), then per field hsync refinement, with chroma the burst-phase
    analysis and burst-refined lines (the track-phase start carried from field to field,
    so this part runs in field order), TBC, the chroma field and the envelope dropout
    detection (threads over fields); repeated for ~`seconds`, then one pass on one thread.
    Also returns the last pass's records and uint16 luma and chroma fields (the FP64 chain
    the device's output is compared with)."""
    from oracle import chroma as OC
    from oracle import oracle as O
    from vhsdecode_hip import chroma as HC
    from vhsdecode_hip import field
    g = field.sync_geometry(fmt)
    dp = field.dropout_params(fmt)
    cfg = O.sync_config(fmt, serration_mode=O.SERRATION_NONE)
    rf = tape.rf
    fsc_shift = 0.0 if g.is_pal else ((117.25 / 360.0) / ((float(fmt.sysparams["outfreq"]) * 1e6) / 4e6)) * g.fs_mhz

    cp = HC.chroma_params(fmt) if chroma else None
    ill = field.input_line_length(fmt)

    def one_block(b):
        r = O.demod_block(rf[b * STRIDE:b * STRIDE + N].astype(np.float64), fl)
        return (b, r.video[CUT:CUT + STRIDE], r.video_lowpass[CUT:CUT + STRIDE], r.envelope[CUT:CUT + STRIDE],
                r.chroma[CUT:CUT + STRIDE])

    def one_pass(pool):
        video = np.empty(nb * STRIDE)
        vlp = np.empty(nb * STRIDE)
        env = np.empty(nb * STRIDE, dtype=np.float32)
        chro = np.empty(nb * STRIDE)
        for b, v, l, e, c in pool.map(one_block, range(nb)):
            video[b * STRIDE:(b + 1) * STRIDE] = v
            vlp[b * STRIDE:(b + 1) * STRIDE] = l
            env[b * STRIDE:(b + 1) * STRIDE] = e
            chro[b * STRIDE:(b + 1) * STRIDE] = c
        recs, locs, shifted = O.vhs_field_sequence(vlp, cfg, 0, 64, want_shifted=True)
        nl = locs.shape[1] if len(recs) else 0

        def refined(k):
            o = recs[k]
            if o.status != 0:
                return None
            ref, _ = O.refine_hsync(shifted[k, :o.window_length], locs[k], np.zeros(nl, np.uint8), o.threshold,
                                    g.fs_mhz, g.hsync / g.fs_mhz, g.conv[0], g.conv[1], g.is_pal, True)
            return ref

        refs = list(pool.map(refined, range(len(recs))))
        # the chroma phase analysis in field order (the track-phase start is carried)
        chroma_fields, renders = {}, {}
        rot, number = 0, 0
        for k, o in enumerate(recs):
            if refs[k] is None:
                continue
            if cp is None:
                renders[k] = refs[k] - fsc_shift
                continue
            w0, w1 = o.window_start, o.window_start + o.window_length
            fc = OC.field_chain(chro[w0:w1], refs[k], cp, bool(o.is_first_field), number, rot, ill, lut,
                                g.nominal, g.smoothing, fsc_shift)
            rot = fc.phase.next_index
            number += 1
            renders[k] = fc.render
            chroma_fields[k] = fc.chroma

        def one_field(k):
            o = recs[k]
            if o.status != 0:
                return None
            ref = renders[k]
            first = 1 if not g.is_pal else (3 if o.is_first_field else 4)
            w0, w1 = o.window_start, o.window_start + o.window_length
            out = O.tbc_resample(video[w0:w1], ref, first, g.field_lines, g.out_len, lut, nominal=g.nominal,
                                 smoothing=g.smoothing, conv=g.conv)[1]
            lines = dp.first_field_lines if o.is_first_field else dp.second_field_lines
            O.tape_dropouts(env, ref[first:] + w0, lines, (w0, min(w1, env.size)), out_len=g.out_len)
            return out

        return recs, list(pool.map(one_field, range(len(recs)))), chroma_fields

    reps = 0
    with ThreadPoolExecutor(max_workers=threads) as pool:
        t0 = time.perf_counter()
        while True:
            recs, outs, couts = one_pass(pool)
            reps += 1
            el = time.perf_counter() - t0
            if el >= seconds:
                break
    samples = ((nb - 1) * STRIDE + N) * reps
    with ThreadPoolExecutor(max_workers=1) as pool:
        t1 = time.perf_counter()
        one_pass(pool)
        el1 = time.perf_counter() - t1
    return samples / el / 1e6, reps, el, ((nb - 1) * STRIDE + N) / el1 / 1e6, recs, outs, couts


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "hbm_traffic_block_kernels_ntsc_vhs_10fields.md")


def measured_traffic(kernel: str):
    """HBM bytes per launch of `kernel` on this workload, from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE passes (scaled by a plain copy of known size)."""
    key = f"traffic_bytes_per_launch[{kernel}] = "
    try:
        with open(TRAFFIC_PROFILE) as h:
            for line in h:
                if line.startswith(key):
                    return int(line[len(key):])
    except OSError:
        pass
    return None


def served_run(requests: int, rate: float, streams: int, window_s: float, seed: int, lut, sweep, sweep_requests: int):
    """Served windows, open loop: `streams` looped synthetic tapes (NTSC VHS SP, PAL VHS SP,
    PAL VHS LP in turn) in pinned host memory; `requests` seek requests for `window_s`-second
    windows at random positions arrive as a Poisson process of `rate` per second, each
    submitted at its scheduled time by one thread while another collects the results.  A
    request's latency runs from its scheduled arrival to its fields in host memory (the
    server's submit -> done time plus any lag of the submission behind the schedule).
    Then the same at each rate of `sweep` (`sweep_requests` requests each)."""
    import threading
    from vhsdecode_hip import field, filters, formats, hipmem, synth
    from vhsdecode_hip.serve import HipServer
    fs = 40e6
    window = int(window_s * fs) // 2 * 2
    preroll = int(round(0.12 * fs))  # samples the server decodes before each window
    specs = [("NTSC", "VHS", "sp"), ("PAL", "VHS", "sp"), ("PAL", "VHS", "lp")]
    bases, fmts, fls = [], [], []
    for k, (system, tf, speed) in enumerate(specs):
        fmt = formats.get_format(system, tf, speed)
        fl = filters.build_tape_filters(fmt, fs, N)
        tape = synth.generate(fmt, fields=8, seed=seed + 100 + k, snr_db=30.0, jitter_ns=20.0)
        pin = hipmem.PinnedArray(tape.rf.size, np.int16)
        pin.array[:] = tape.rf
        bases.append(pin)
        fmts.append(fmt)
        fls.append(fl)
    max_fields = int(window_s * 60) + 2
    srv = HipServer(max_slots=16, hbm_budget_bytes=64 << 30, max_window_samples=window + preroll + 2,
                    max_fields=max_fields)
    fids = [srv.add_format(fmts[k], fls[k], lut) for k in range(len(specs))]
    sids = [srv.add_stream(fids[i % 3], bases[i % 3].ptr, bases[i % 3].array.size, loop=True) for i in range(streams)]
    per = max(field.sync_geometry(f).field_lines * field.sync_geometry(f).out_len for f in fmts)
    pool = [hipmem.PinnedArray(max_fields * per, np.uint16) for _ in range(24)]
    rng = np.random.default_rng(seed)

    def run_point(n_req, r_rate):
        free = list(range(len(pool)))
        cv = threading.Condition()
        tickets = []  # (ticket, buffer, lag seconds)
        done_t = threading.Event()

        def submitter():
            arrivals = np.cumsum(rng.exponential(1.0 / r_rate, size=n_req))
            sids_r = rng.integers(0, streams, size=n_req)
            firsts = 2 * rng.integers(preroll // 2, 10**9, size=n_req)
            t0 = time.perf_counter()
            for i in range(n_req):
                now = time.perf_counter() - t0
                if arrivals[i] > now:
                    time.sleep(arrivals[i] - now)
                with cv:
                    while not free:
                        cv.wait()
                    b = free.pop()
                lag = (time.perf_counter() - t0) - arrivals[i]
                t = srv.submit(sids[int(sids_r[i])], int(firsts[i]), window, max_fields, pool[b].ptr)
                with cv:
                    tickets.append((t, b, lag))
                    cv.notify_all()
            done_t.set()

        th = threading.Thread(target=submitter)
        t_start = time.perf_counter()
        th.start()
        lat, fields, ok, k = [], 0, True, 0
        while k < n_req:
            with cv:
                while k >= len(tickets):
                    cv.wait()
                t, b, lag = tickets[k]
            r = srv.wait(t)
            ok &= r.status == 0
            fields += r.field_count
            lat.append(r.latency_ms + lag * 1e3)
            with cv:
                free.append(b)
                cv.notify_all()
            k += 1
        th.join()
        wall = time.perf_counter() - t_start
        lat = np.array(lat)
        return {"rate_per_s": r_rate, "requests": n_req, "all_ok": bool(ok),
                "p50_ms": round(float(np.percentile(lat, 50)), 2), "p99_ms": round(float(np.percentile(lat, 99)), 2),
                "max_ms": round(float(lat.max()), 2), "fields": int(fields),
                "offered_msps": round(r_rate * window / 1e6, 1),
                "aggregate_msps": round(n_req * window / wall / 1e6, 1), "seconds": round(wall, 2)}

    # warm every slot and format once (slot buffers, contexts and tables are created on first use)
    warm = [srv.submit(sids[k % len(sids)], preroll, window, max_fields, pool[k].ptr) for k in range(16)]
    for t in warm:
        srv.wait(t)
    head = run_point(requests, rate)
    points = [run_point(sweep_requests, r) for r in sweep]
    srv.close()
    return {
        "streams": streams, "window_s": window_s, "window_samples": window, "preroll_samples": preroll,
        "max_fields": max_fields, **head, "sweep": points,
        "formats": ["NTSC VHS SP", "PAL VHS SP", "PAL VHS LP"], "hbm_budget_gb": 64,
        "latency": "scheduled arrival -> uint16 fields in host memory, open-loop Poisson arrivals",
        "source": "pinned host RAM, looped 8-field synthetic tapes (30 dB, 20 ns jitter)",
    }


BATCH_MIX = [("PAL", "VHS", "sp"), ("NTSC", "VHS", "sp"), ("PAL", "SVHS", "sp"), ("PAL", "BETAMAX", "sp"),
             ("PAL", "VIDEO8", "sp")]


def batched_run(seed: int, segments: int, steps: int):
    """Uneven segments (seeded lengths 16384..262144) of PAL/NTSC VHS, PAL S-VHS, PAL
    Betamax and PAL Video8 at 20 and 40 MSPS, 30 dB: the graph-captured batch
    (vhsdecode_hip_batch_*) against the plain path, one demod_span_device call per
    segment on one stream.  Both decode exactly the segments' whole blocks."""
    from vhsdecode_hip import filters, formats, hipmem, synth
    from vhsdecode_hip.batch import HipBatch, pack_segments
    from vhsdecode_hip.demod import HipBlockDemodulator
    rng = np.random.default_rng(seed)
    lengths = rng.integers(16384, 262145, size=segments)
    kinds = [(m, fs) for m in BATCH_MIX for fs in (20e6, 40e6)]
    dems, tapes = [], []
    for i, ((system, name, speed), fs) in enumerate(kinds):
        fmt = formats.get_format(system, name, speed)
        dems.append(HipBlockDemodulator(filters.build_tape_filters(fmt, fs, N), max_blocks=16))
        tapes.append(synth.generate(fmt, fields=1, sample_rate_hz=fs, seed=seed + i, snr_db=30.0).rf)
    choice = rng.integers(0, len(kinds), size=segments)
    segs = pack_segments(lengths, [dems[c] for c in choice])
    rf_host = np.zeros(segs[-1].rf_offset + int(lengths[-1]) + 2, dtype=np.int16)
    for s, c in zip(segs, choice):
        t = tapes[c]
        start = 2 * int(rng.integers(0, (t.size - s.length) // 2))
        rf_host[s.rf_offset:s.rf_offset + s.length] = t[start:start + s.length]
    rf = hipmem.DeviceBuffer.from_array(rf_host)
    batch = HipBatch(segs, use_graph=True)
    n = batch.output_samples
    planes = hipmem.DeviceBuffer(4 * n * 4)
    flags = hipmem.DeviceBuffer(16 * 4)
    p = [planes.ptr + k * n * 4 for k in range(4)]
    live = [s for s in segs if s.blocks() > 0]

    def run_batch():
        batch.run_device(rf.ptr, rf_host.size, *p)

    def run_plain():
        for s in live:
            o = s.out_offset * 4
            s.demodulator.demod_span_device(rf.ptr, s.rf_offset, s.blocks(), p[0] + o, p[1] + o, p[2] + o,
                                            p[3] + o, flags.ptr)

    out = {}
    for name, fn in (("batched", run_batch), ("plain", run_plain)):
        for _ in range(2):
            fn()
        hipmem.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        hipmem.synchronize()
        out[name] = (time.perf_counter() - t0) / steps
    samples = sum(s.blocks() * s.demodulator.stride for s in segs)
    batch.close()
    for d in dems:
        d.close()
    return {"segments": segments, "contexts": len(kinds), "blocks": int(batch.blocks),
            "short_segments": sum(1 for s in segs if s.blocks() == 0), "decoded_samples": int(samples),
            "segment_samples": int(lengths.sum()),
            "batched_ms": round(out["batched"] * 1e3, 4), "plain_ms": round(out["plain"] * 1e3, 4),
            "batched_msps": round(samples / out["batched"] / 1e6, 1), "plain_msps": round(samples / out["plain"] / 1e6, 1),
            "speedup": round(out["plain"] / out["batched"], 2),
            "formats": [f"{m[0]} {m[1]} {m[2].upper()} {fs / 1e6:.0f} MSPS" for m, fs in kinds],
            "scope": "block demod chain (demod, repair, video, sub-deemphasis, chroma); no TBC (segments are not fields)"}


def whole_run(fmt, tape, fl, lut, copies: int, dropouts: bool, chroma: bool = True):
    """The whole-run entry (vhsdecode_hip_run) over `copies` back-to-back copies of the
    tape held in host RAM: callback reads, PCIe upload, device decode of each span, .tbc and
    .tbc.json written to a temporary directory.  End to end, so never the headline value."""
    import shutil
    import tempfile
    from vhsdecode_hip.run import decode_run
    rf = np.tile(tape.rf, copies)
    tmp = tempfile.mkdtemp(prefix="vhsd_run_")
    try:
        base = os.path.join(tmp, "capture")
        decode_run(rf[: 4 * (rf.size // copies)], fmt, fl, base, sinc_table=lut, dropouts=dropouts,
                   chroma=chroma)  # warm-up
        t0 = time.perf_counter()
        res = decode_run(rf, fmt, fl, base, sinc_table=lut, dropouts=dropouts, chroma=chroma)
        wall = time.perf_counter() - t0
        tbc_bytes = os.path.getsize(base + ".tbc")
        chroma_bytes = os.path.getsize(base + "_chroma.tbc") if chroma else 0
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    return {"samples": int(rf.size), "fields_written": res.fields_written, "spans": res.spans,
            "dropout_rows": res.dropout_rows, "seconds": round(wall, 4),
            "msps": round(rf.size / wall / 1e6, 1),
            "loop_seconds": round(res.elapsed_seconds, 4),
            "loop_msps": round(rf.size / res.elapsed_seconds / 1e6, 1), "tbc_bytes": tbc_bytes,
            "chroma_tbc_bytes": chroma_bytes,
            "path": "read callback from host RAM -> pinned staging -> PCIe -> device decode per span -> "
                    ".tbc/_chroma.tbc/.tbc.json on local disk; one span at a time"}


def laserdisc_run(seed: int, fields: int, steps: int, lut):
    """Synthetic NTSC LaserDisc at 40 MSPS int16 (7.6 -> 9.3 MHz video FM, analog sound
    carriers at 2.30/2.81 MHz with +-100 kHz deviation and 1 kHz / 3 kHz tones, 30 dB):
    LaserDisc block demod with both audio channels in the same launch (the LaserDisc
    field logic is not on the device, so no TBC).  Plus a HiFi stream: two channels (1.3
    and 1.7 MHz carriers, 440 / 880 Hz tones, +-150 kHz) of one second at the 2^23 Hz
    Hilbert IF rate through the HiFi discriminator."""
    from vhsdecode_hip import field, formats, hipmem, ld_filters, synth
    from vhsdecode_hip.demod import HipBlockDemodulator
    from vhsdecode_hip.hifi import HipHiFiHilbertDiscriminator
    fmt = formats.get_laserdisc_format("NTSC")
    fl = ld_filters.build_laserdisc_filters(fmt, 40e6, N)
    sp = fmt.sysparams
    car = ((float(sp["audio_lfreq"]), 100e3, 1000.0, 0.1), (float(sp["audio_rfreq"]), 100e3, 3000.0, 0.1))
    tape = synth.generate(fmt, fields=fields, seed=seed, snr_db=30.0, audio_carriers=car)
    nb = (tape.rf.size - N) // STRIDE + 1
    dem = HipBlockDemodulator(fl, max_blocks=nb)
    dem.set_sinc_table(lut)
    span = nb * STRIDE
    rf = hipmem.DeviceBuffer.from_array(tape.rf)
    planes = hipmem.DeviceBuffer(4 * span * 4)
    aud = hipmem.DeviceBuffer(2 * (span // fl.audio_decimation) * 4)
    flags = hipmem.DeviceBuffer(nb * 4)
    video, vlp, env, burst = (planes.ptr + k * span * 4 for k in range(4))

    def step():
        dem.demod_span_device(rf.ptr, 0, nb, video, vlp, env, burst, flags.ptr,
                              audio_left_ptr=aud.ptr, audio_right_ptr=aud.ptr + (span // fl.audio_decimation) * 4)

    step()
    hipmem.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    hipmem.synchronize()
    ld_s = (time.perf_counter() - t0) / steps
    rf_samples = (nb - 1) * STRIDE + N
    dem.close()
    if_rate, block = 1 << 23, 1 << 22
    t = np.arange(2 * block, dtype=np.float64) / if_rate
    hifi = []
    for carrier, tone in ((1_300_000.0, 440.0), (1_700_000.0, 880.0)):
        x = np.cos(2 * np.pi * np.cumsum(carrier + 150e3 * np.sin(2 * np.pi * tone * t)) / if_rate).astype(np.float32)
        hifi.append((HipHiFiHilbertDiscriminator(if_rate, carrier, 150e3, block),
                     hipmem.DeviceBuffer.from_array(x), hipmem.DeviceBuffer(x.nbytes)))

    def hifi_step():
        for d, xi, xo in hifi:
            d.demodulate_device(xi.ptr, xo.ptr, 2)

    hifi_step()
    hipmem.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        hifi_step()
    hipmem.synchronize()
    hifi_s = (time.perf_counter() - t0) / steps
    for d, _, _ in hifi:
        d.close()
    return {"workload": f"NTSC LaserDisc 40 MSPS int16, {fields} fields, 30 dB, analog audio 2.30/2.81 MHz",
            "rf_samples": int(rf_samples), "blocks": int(nb),
            "ms": round(ld_s * 1e3, 4), "msps": round(rf_samples / ld_s / 1e6, 1),
            "audio_samples_per_channel": int(span // fl.audio_decimation),
            "hifi": {"channels": 2, "if_rate_hz": if_rate, "block": block, "seconds": 1.0,
                     "ms": round(hifi_s * 1e3, 4),
                     "if_msps": round(2 * 2 * block / hifi_s / 1e6, 1)},
            "scope": "LaserDisc block demod and analog audio; no field decode"}


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--fields", type=int, default=10)
    ap.add_argument("--rf-copies", type=int, default=24,
                    help="copies of the tape in HBM the steps rotate over (24 x 13.3 MB > the 256 MB cache)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: min(16, cores)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=1000, help="served-window requests (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=8.0)
    ap.add_argument("--serve-streams", type=int, default=16)
    ap.add_argument("--serve-window", type=float, default=2.0, help="window length (s of RF)")
    ap.add_argument("--serve-sweep", type=str, default="16,32,64", help="further arrival rates (req/s)")
    ap.add_argument("--serve-sweep-requests", type=int, default=200)
    ap.add_argument("--no-dropouts", action="store_true", help="skip the per-field dropout detection")
    ap.add_argument("--no-chroma", action="store_true",
                    help="skip the colour-under chroma (the reference's --skip-chroma; the luma then renders at "
                         "the hsync-refined locations)")
    ap.add_argument("--pipeline", type=int, default=3,
                    help="spans in flight: consecutive steps alternate between this many pipelines")
    ap.add_argument("--batch-segments", type=int, default=64, help="batched-segment run (0: skip)")
    ap.add_argument("--ld-fields", type=int, default=20, help="LaserDisc + HiFi run (0: skip)")
    ap.add_argument("--run-copies", type=int, default=18,
                    help="whole-run entry over this many copies of the tape from host RAM (0: skip)")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="first run one device-to-device copy of this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported by this build "
              "(one process drives one MI355X).", file=sys.stderr)
        sys.exit(2)

    from vhsdecode_hip import field, hipmem
    from vhsdecode_hip.demod import HipBlockDemodulator

    with_chroma = not args.no_chroma
    fmt, tape, fl, nb = build_workload(args.fields, with_chroma)
    lut = np.load(os.path.join(ROOT, "vhs-decode-dotnet_amd", "vhsdecode_hip", "data",
                               "sinc_lut.npz"))["downscale_sinc_lut"]
    span = nb * STRIDE
    # the tape at RF_COPIES addresses in HBM, the steps rotating over them, so a step's RF
    # reads come from HBM and not from the 256 MB last-level cache a previous step filled
    rf_copies = [hipmem.DeviceBuffer.from_array(tape.rf) for _ in range(args.rf_copies)]
    max_fields = args.fields + 6

    # RF-envelope dropout detection of every field, on by default as in the reference's
    # VHS decode (TbcDropoutDetectionOptions.DefaultDdd)
    dropouts = None if args.no_dropouts else field.dropout_params(fmt, capacity=DROPOUT_ROWS)
    import ctypes
    from vhsdecode_hip import _lib
    rec_size = ctypes.sizeof(_lib.VhsFieldV1)

    class Pipe:
        """One decode pipeline: its own demod context (stash, descriptors, field scratch and
        sequence state), span planes, field outputs, records and stream."""

        def __init__(self, own_stream: bool):
            self.dem = HipBlockDemodulator(fl, max_blocks=nb)
            self.dem.set_sinc_table(lut)
            self.dec = field.HipVhsFieldDecoder(self.dem, fmt, chroma=with_chroma)
            g = self.dec.geom
            self.planes = hipmem.DeviceBuffer(4 * span * 4)
            self.flags = hipmem.DeviceBuffer(nb * 4)
            self.out = hipmem.DeviceBuffer(max_fields * g.field_lines * g.out_len * 2)
            self.out_c = hipmem.DeviceBuffer(max_fields * g.field_lines * g.out_len * 2) if with_chroma else None
            self.recs = hipmem.DeviceBuffer(max_fields * rec_size)
            self.meta = hipmem.DeviceBuffer(16)  # record count, next begin
            self.drop = hipmem.DeviceBuffer(4 * max_fields * (1 + 3 * DROPOUT_ROWS))
            self.stream = hipmem.Stream() if own_stream else None
            self.p = [self.planes.ptr + k * span * 4 for k in range(4)]

        def step(self, rf):
            # every stage on the device, no host round trip inside the step
            st = self.stream.handle if self.stream else 0
            self.dem.demod_span_device(rf.ptr, 0, nb, *self.p, self.flags.ptr, stream=st)
            self.dec.decode_async(self.p[0], self.p[1], span, self.out.ptr, max_fields, self.meta.ptr,
                                  self.recs.ptr, self.meta.ptr + 8, rounds=2, stream=st, span_chroma_ptr=self.p[3],
                                  out_chroma_ptr=self.out_c.ptr if with_chroma else 0)
            if dropouts is not None:
                loc_ptr, nloc = self.dem.vhs_field_lines_device(3)
                d = self.drop.ptr
                rows = d + 4 * max_fields
                self.dem.detect_dropouts_device(self.p[2], span, loc_ptr, nloc, self.recs.ptr, self.meta.ptr,
                                                max_fields, dropouts, d, rows, rows + 4 * max_fields * DROPOUT_ROWS,
                                                rows + 8 * max_fields * DROPOUT_ROWS, stream=st)

        def records(self):
            count = int(self.meta.download(np.int32, 1)[0])
            raw = self.recs.download(np.uint8, max(count, 1) * rec_size)
            return [_lib.VhsFieldV1.from_buffer_copy(raw[k * rec_size:(k + 1) * rec_size].tobytes())
                    for k in range(count)]

    # Consecutive spans alternate between pipelines on their own streams, so one span's
    # field decode (small latency-bound kernels) overlaps the next span's block kernels.
    pipes = [Pipe(args.pipeline > 1) for _ in range(max(1, args.pipeline))]
    dem = pipes[0].dem

    if args.copy_mib > 0:
        a = hipmem.DeviceBuffer(args.copy_mib << 20)
        b = hipmem.DeviceBuffer(args.copy_mib << 20)
        hipmem.copy_device(b.ptr, a.ptr, args.copy_mib << 20)
        del a, b

    for i in range(args.warmup):
        pipes[i % len(pipes)].step(rf_copies[i % len(rf_copies)])
    hipmem.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        pipes[i % len(pipes)].step(rf_copies[i % len(rf_copies)])
    t_enq = time.perf_counter()
    hipmem.synchronize()
    t1 = time.perf_counter()
    records = pipes[0].records()
    nfields = sum(1 for r in records if r.status == 0)
    # per-kernel times inside the pipelined loop, from a second pass with HIP events on
    # pipeline 0 (profiling makes its field sequence launch kernel by kernel instead of
    # replaying the captured graph, so this pass is not the timed one)
    dem.set_profiling(True)
    dem.take_profile()
    for i in range(min(args.steps, 12)):
        pipes[i % len(pipes)].step(rf_copies[i % len(rf_copies)])
    hipmem.synchronize()
    prof = dem.take_profile()
    # the same kernels with one span in flight: their own durations, no sharing of CUs
    # with another span's work (the roofline figure)
    for i in range(3):
        pipes[0].step(rf_copies[(args.steps + i) % len(rf_copies)])
        hipmem.synchronize()
    prof_iso = dem.take_profile()
    dem.set_profiling(False)
    dev_fields = pipes[0].out.download(np.uint16, len(records) * pipes[0].dec.geom.field_lines *
                                       pipes[0].dec.geom.out_len).reshape(len(records), -1)
    dev_chroma = pipes[0].out_c.download(np.uint16, len(records) * pipes[0].dec.geom.field_lines *
                                         pipes[0].dec.geom.out_len).reshape(len(records), -1) if with_chroma else None

    rf_samples = (nb - 1) * STRIDE + N
    elapsed = t1 - t0
    value = rf_samples * args.steps / elapsed / 1e6
    # algorithmic bytes per launch of each per-block kernel (int16 RF in, FP32 planes of
    # the kept stride out; spectra, stash and scratch are not algorithmic)
    alg_bytes = {
        "block_demod": ("vhs_block_demod_kernel", nb * (2 * N + 4 * STRIDE)),   # RF in, envelope out
        "block_video": ("vhs_block_video_kernel", nb * (2 * 4 * STRIDE)),        # video + sync reference out
        "chroma": ("vhs_block_chroma_kernel", nb * (2 * N + 4 * STRIDE)),       # RF in, chroma out
    }
    # one span in flight: each kernel's own duration (HIP events on its launch stream, no
    # other span's work sharing the CUs); the rocprofv3 kernel trace of `bench.py
    # --pipeline 1` reads the same durations
    avg_ms = {k: prof_iso[k]["ms"] / prof_iso[k]["launches"] for k in alg_bytes if prof_iso[k]["launches"]}
    # in the timed region: pipeline 0's launches, other spans' kernels overlapping them
    pipe_ms = {k: prof[k]["ms"] / prof[k]["launches"] for k in alg_bytes if prof[k]["launches"]}
    # the block demod kernel has the largest share of the step's kernel time (rocprofv3
    # summary)
    dominant = "block_demod"
    kernel_name, launch_bytes = alg_bytes[dominant]
    demod_ms = avg_ms[dominant]
    achieved = launch_bytes / (demod_ms * 1e-3) / 1e9
    chain_ms = sum(avg_ms.values())
    chain_bytes = nb * (2 * N + 16 * STRIDE)
    for name, p in prof.items():
        if p["launches"]:
            print(f"[bench] {name:13s} launches={p['launches']:4d} avg={p['ms'] / p['launches'] * 1e3:9.1f} us",
                  file=sys.stderr)
    print(f"[bench] blocks={nb} fields/step={nfields} step={elapsed / args.steps * 1e3:.3f} ms "
          f"(host enqueue {(t_enq - t0) / args.steps * 1e3:.3f} ms per step)", file=sys.stderr)
    print("[bench] one span in flight: " + ", ".join(
        f"{name} {p['ms'] / p['launches'] * 1e3:.1f} us" for name, p in prof_iso.items() if p["launches"]),
        file=sys.stderr)

    in_pipe_ms = pipe_ms[dominant]
    in_pipe_achieved = launch_bytes / (in_pipe_ms * 1e-3) / 1e9
    traffic = measured_traffic(kernel_name)
    result = {
        "metric": "RF MSps decoded (demod+TBC) on 1 GPU vs exact; p50/p99 2-s window latency",
        "value": round(value, 3),
        "unit": "MSPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"NTSC VHS 40 MSPS int16, {args.fields} fields, 28 dB SNR, 20 ns 1/f line jitter",
                   "rf_samples_per_step": rf_samples, "blocks_per_step": nb, "fields_per_step": nfields,
                   "records_per_step": len(records), "block_length": N, "parallelism": "single GPU",
                   "spans_in_flight": len(pipes), "dropout_detection": dropouts is not None,
                   "chroma": with_chroma,
                   "rf_copies_in_hbm": len(rf_copies),
                   "sync": "vhs field sequence (level detection, pulses, parity, line 0, line locations) on "
                           "the device, the serration search not run"},
        "roofline": {"bound": "hbm", "kernel": kernel_name, "achieved": round(achieved, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "bytes_per_launch": launch_bytes, "avg_launch_ms": round(demod_ms, 5),
                     "block_chain": {"kernels": sorted(avg_ms), "ms": round(chain_ms, 5),
                                     "achieved": round(chain_bytes / (chain_ms * 1e-3) / 1e9, 2)},
                     "timing": "HIP events on the launch stream, one span in flight",
                     "in_pipeline": {"avg_launch_ms": round(in_pipe_ms, 5), "achieved": round(in_pipe_achieved, 2),
                                     "frac": round(in_pipe_achieved / HBM_PEAK_GBS, 4)}},
    }
    if args.batch_segments > 0:
        bt = batched_run(SEED, args.batch_segments, steps=10)
        result["batched"] = bt
        print(f"[bench] batched: {bt['segments']} segments / {bt['blocks']} blocks in {bt['batched_ms']} ms "
              f"({bt['batched_msps']} MSPS) vs plain {bt['plain_ms']} ms ({bt['plain_msps']} MSPS)", file=sys.stderr)
    if args.run_copies > 0:
        wr = whole_run(fmt, tape, fl, lut, args.run_copies, dropouts is not None, with_chroma)
        result["whole_run"] = wr
        print(f"[bench] whole run: {wr['fields_written']} fields in {wr['spans']} spans, {wr['seconds']} s "
              f"({wr['msps']} MSPS end to end with set-up; span loop {wr['loop_msps']} MSPS)", file=sys.stderr)
    if args.ld_fields > 0:
        ldr = laserdisc_run(SEED, args.ld_fields, steps=5, lut=lut)
        result["laserdisc"] = ldr
        print(f"[bench] laserdisc: {ldr['msps']} MSPS; hifi {ldr['hifi']['if_msps']} IF MSPS",
              file=sys.stderr)
    if args.serve_requests > 0:
        for p in pipes:
            p.dem.close()
        sweep = [float(x) for x in args.serve_sweep.split(",") if x.strip()]
        sv = served_run(args.serve_requests, args.serve_rate, args.serve_streams, args.serve_window, SEED, lut,
                        sweep, args.serve_sweep_requests)
        result["serving"] = sv
        print(f"[bench] served: {sv['requests']} windows at {sv['rate_per_s']}/s p50 {sv['p50_ms']} ms "
              f"p99 {sv['p99_ms']} ms; sweep " + ", ".join(
                  f"{p['rate_per_s']}/s p99 {p['p99_ms']} ms {p['aggregate_msps']} MSPS" for p in sv["sweep"]),
              file=sys.stderr)
    if not args.no_cpu_baseline:
        nproc = os.cpu_count() or 1
        usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else nproc
        # every core this process may run on, up to the 16-core host share of one GPU
        threads = args.cpu_threads or max(1, min(16, usable))
        msps, reps, el, msps1, orecs, oouts, ochroma = cpu_baseline(fmt, tape, fl, nb, lut, args.cpu_seconds,
                                                                     threads, with_chroma)
        result["cpu_baseline"] = {"value": round(msps, 3), "unit": "MSPS", "cores": threads, "kind": "port",
                                  "sample": f"the full {args.fields}-field tape x {reps} passes "
                                            f"({el:.1f} s): FP64 oracle block demod and per-field hsync refine, "
                                            f"{'chroma burst analysis and decode, ' if with_chroma else ''}TBC "
                                            f"and dropout detection on {threads} threads; the field sequence"
                                            f"{' and the chroma track-phase chain' if with_chroma else ''} serial",
                                  "nproc": nproc, "affinity_cores": usable,
                                  "single_thread_value": round(msps1, 3)}
        # the device step's fields against this FP64 chain (TBC +-1 LSB bar)
        same = len(orecs) == len(records) and all(
            (a.status, a.window_start, a.is_first_field) == (b.status, b.window_start, b.is_first_field)
            for a, b in zip(records, orecs))
        worst, over, total = 0, 0, 0
        if same:
            for k, o in enumerate(oouts):
                if o is None:
                    continue
                d = np.abs(dev_fields[k].astype(np.int32) - o.astype(np.int32))
                worst = max(worst, int(d.max()))
                over += int(np.count_nonzero(d > 1))
                total += d.size
        result["parity"] = {"records_match": bool(same), "tbc_max_lsb": worst,
                            "tbc_frac_gt1": (over / total) if total else None, "tbc_samples": total,
                            "reference": "FP64 oracle chain (demod, field sequence, hsync, "
                                         f"{'chroma burst analysis, ' if with_chroma else ''}TBC) on the same RF"}
        if with_chroma and same:
            cw, cover, ctotal = 0, 0, 0
            for k, o in ochroma.items():
                d = np.abs(dev_chroma[k].astype(np.int32) - o.astype(np.int32))
                cw = max(cw, int(d.max()))
                cover += int(np.count_nonzero(d > 1))
                ctotal += d.size
            result["parity"].update({"chroma_max_lsb": cw, "chroma_frac_gt1": (cover / ctotal) if ctotal else None,
                                     "chroma_samples": ctotal})
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
