"""vhsdecode_hip_find_pulses_noalloc (the pulse search with no allocation: scratch in the
library's file-scope device array) against vhsdecode_hip_find_pulses on the same
demodulated span planes, through the C ABI: identical pulse starts, lengths and counts on
two calls in a row, the second on a smaller input, on the library's own stream and on a
caller stream.  Runs on an MI355X only."""
import ctypes

import numpy as np
import pytest

from vhsdecode_hip import _lib, field, filters, formats, hipmem, synth
from vhsdecode_hip.demod import HipBlockDemodulator

pytestmark = pytest.mark.gpu

N = 32768
STRIDE = N - 2 * 1024
CAPACITY = 8192


def _decoded_span(system, nfields, seed):
    fmt = formats.get_format(system, "VHS", "sp")
    fl = filters.build_tape_filters(fmt, 40e6, N)
    tape = synth.generate(fmt, fields=nfields, seed=seed, snr_db=30.0, jitter_ns=20.0)
    dem = HipBlockDemodulator(fl, max_blocks=128)
    nb = dem.blocks_for(tape.rf.size)
    n = nb * STRIDE
    rf = hipmem.DeviceBuffer.from_array(tape.rf)
    planes = hipmem.DeviceBuffer(4 * n * 4)
    flags = hipmem.DeviceBuffer(nb * 4)
    video, vlp, env, chroma = (planes.ptr + k * n * 4 for k in range(4))
    dem.demod_span_device(rf.ptr, 0, nb, video, vlp, env, chroma, flags.ptr)
    hipmem.synchronize()
    return dict(fmt=fmt, dem=dem, vlp=vlp, n=n, keep=(rf, planes, flags))


def _noalloc(ptr, length, bias, g, stream):
    L = _lib.load()
    out = hipmem.DeviceBuffer(2 * CAPACITY * 4 + 64)
    out.fill(0xFF)  # stale values must not survive: every stored word is overwritten
    starts, lens, counts = out.ptr, out.ptr + CAPACITY * 4, out.ptr + 2 * CAPACITY * 4
    _lib.check(L.vhsdecode_hip_find_pulses_noalloc(ctypes.c_void_p(ptr), length, bias, g.threshold, g.min_pulse,
                                                   g.max_pulse, ctypes.c_void_p(starts), ctypes.c_void_p(lens),
                                                   CAPACITY, ctypes.c_void_p(counts), ctypes.c_void_p(stream)))
    if stream:
        hipmem.hip().hipStreamSynchronize(ctypes.c_void_p(stream))
    hipmem.synchronize()
    c = out.download(np.int32, 2, offset=2 * CAPACITY * 4)
    n = int(c[1])
    assert 0 <= n <= CAPACITY and 0 <= int(c[0]) <= _lib.NOALLOC_MAX_EDGES
    return out.download(np.int32, n), out.download(np.int32, n, offset=CAPACITY * 4), c


@pytest.mark.parametrize("use_stream", [False, True], ids=["null_stream", "caller_stream"])
def test_noalloc_pulses_match_find_pulses_on_two_calls(use_stream):
    ntsc = _decoded_span("NTSC", 3, 21)
    pal = _decoded_span("PAL", 1, 22)
    stream = hipmem.Stream() if use_stream else None
    sh = stream.handle if stream else 0
    # first call: a 3-field NTSC span; second call: a smaller PAL span, then a cut of the NTSC span
    cases = [(ntsc, ntsc["n"]), (pal, pal["n"]), (ntsc, ntsc["n"] // 3 + 123)]
    assert cases[1][1] < cases[0][1] and cases[2][1] < cases[0][1]
    for k, (sp, length) in enumerate(cases):
        g = field.sync_geometry(sp["fmt"])
        print(f"[noalloc] call {k}: length={length} stream={'caller' if sh else 'null'}", flush=True)
        ref_st, ref_ln = sp["dem"].find_pulses(sp["vlp"], length, g.threshold, g.min_pulse, g.max_pulse, CAPACITY)
        st, ln, counts = _noalloc(sp["vlp"], length, sp["dem"].bias, g, sh)
        assert ref_st.size > 100
        np.testing.assert_array_equal(st, ref_st)
        np.testing.assert_array_equal(ln, ref_ln)
        assert int(counts[1]) == ref_st.size


def test_noalloc_rejects_oversize_and_handles_empty():
    L = _lib.load()
    out = hipmem.DeviceBuffer(64)
    rc = L.vhsdecode_hip_find_pulses_noalloc(ctypes.c_void_p(out.ptr), _lib.NOALLOC_MAX_SAMPLES + 1, 0.0, 0.0, 1, 10,
                                             None, None, 0, ctypes.c_void_p(out.ptr), None)
    assert rc == -30007
    assert b"exceeds the maximum" in L.vhsdecode_hip_get_last_error()
    out.fill(0x7F)
    _lib.check(L.vhsdecode_hip_find_pulses_noalloc(ctypes.c_void_p(out.ptr), 0, 0.0, 0.0, 1, 10, None, None, 0,
                                                   ctypes.c_void_p(out.ptr), None))
    hipmem.synchronize()
    np.testing.assert_array_equal(out.download(np.int32, 2), [0, 0])
