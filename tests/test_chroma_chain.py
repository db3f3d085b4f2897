"""The chroma restatement end to end on the CPU: a synthetic tape whose colour-under
chroma carries the burst with the format's track-phase rotation, demodulated and
synced by the oracle, then each field's chroma by oracle/chroma.py.  What the
mathematics fixes: the phase analysis follows the recorded rotation (one start for
every field, alternating as the tracks do), the burst-refined lines stay within a
fraction of a subcarrier cycle of the hsync-refined ones, and the decoded burst comes
out at the subcarrier with the reference burst level (the automatic gain's target).
Uses the block demod of a few blocks only (the chroma plane of one field)."""
import os

import numpy as np
import pytest

from oracle import chroma as C
from oracle import oracle as O
from vhsdecode_hip import chroma as HC
from vhsdecode_hip import field, filters, formats, synth

N, CUT = 32768, 1024
STRIDE = N - 2 * CUT


def _lut():
    import vhsdecode_hip
    return np.load(os.path.join(os.path.dirname(vhsdecode_hip.__file__), "data", "sinc_lut.npz"))[
        "downscale_sinc_lut"]


def _planes(rf, fl, nb):
    from concurrent.futures import ThreadPoolExecutor

    def one(b):
        r = O.demod_block(rf[b * STRIDE:b * STRIDE + N].astype(np.float64), fl)
        return b, r.video[CUT:CUT + STRIDE], r.video_lowpass[CUT:CUT + STRIDE], r.chroma[CUT:CUT + STRIDE]

    out = np.empty((3, nb * STRIDE))
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        for b, v, s, c in ex.map(one, range(nb)):
            out[:, b * STRIDE:(b + 1) * STRIDE] = (v, s, c)
    return out


@pytest.mark.parametrize("system", ["NTSC", "PAL"])
def test_chroma_chain_on_a_recorded_burst(system):
    fmt = formats.get_format(system, "VHS", "sp")
    p = HC.chroma_params(fmt)
    tape = synth.generate(fmt, fields=3, seed=71, snr_db=34.0, jitter_ns=10.0, chroma_under="modulated")
    fl = filters.build_tape_filters(fmt, 40e6, N)
    nb = (tape.rf.size - N) // STRIDE + 1
    video, vlp, chroma = _planes(tape.rf, fl, nb)
    geom = field.sync_geometry(fmt)
    ofields, olocs, oshift = O.vhs_field_sequence(vlp, O.sync_config(fmt, serration_mode=O.SERRATION_NONE), 0, 8,
                                                  want_shifted=True)
    done = [k for k, f in enumerate(ofields) if f.status == 0]
    assert len(done) >= 2
    fsc_shift = 0.0 if geom.is_pal else ((117.25 / 360.0) / ((float(fmt.sysparams["outfreq"]) * 1e6) / 4e6)) * 40.0
    rot_index, starts = 0, []
    nl = olocs.shape[1]
    for number, k in enumerate(done):
        o = ofields[k]
        ref, _ = O.refine_hsync(oshift[k, :o.window_length], olocs[k], np.zeros(nl, np.uint8), o.threshold,
                                geom.fs_mhz, geom.hsync / geom.fs_mhz, geom.conv[0], geom.conv[1], geom.is_pal)
        src = chroma[o.window_start:o.window_start + o.window_length]
        fc = C.field_chain(src, ref, p, bool(o.is_first_field), number, rot_index,
                           field.input_line_length(fmt), _lut(), geom.nominal, geom.smoothing, fsc_shift)
        starts.append(rot_index)
        rot_index = fc.phase.next_index
        # burst-refined lines move by less than half a subcarrier cycle (in input samples)
        moved = (fc.render + fsc_shift) - ref
        cycle = 40.0 / p.fsc_mhz
        assert np.nanmax(np.abs(moved[:p.lines])) < 0.5 * cycle
        # the decoded burst: a subcarrier at the reference level on the combed lines
        u = fc.chroma.reshape(p.lines, p.line_len).astype(np.float64) - 32767.0
        burst = u[40:200, p.burst_start + 8:p.burst_end - 8]
        rms = np.sqrt(np.mean((burst - burst.mean(axis=1, keepdims=True)) ** 2))
        assert 0.5 * p.burst_abs_ref < rms < 1.5 * p.burst_abs_ref, rms
        spec = np.abs(np.fft.rfft(burst, axis=1)).mean(axis=0)
        peak = np.argmax(spec[1:]) + 1
        assert abs(peak / burst.shape[1] - 0.25) < 0.05   # fsc at 4 fsc sampling
    # every field after the first takes the start the previous one handed on
    assert len(set(starts[1:]) | {starts[0]}) <= 2
