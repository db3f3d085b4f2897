"""Host-side checks of the whole-run entry (vhsdecode_hip_run): struct layouts against the
header, and the refusals that come before any device work (struct sizes, missing
callback, an existing output with overwrite off, a cancel before start-up)."""
import ctypes
import os

import numpy as np
import pytest

from vhsdecode_hip import _lib, filters, formats
from vhsdecode_hip.run import decode_run

E_INVALID, E_STRUCT, E_CANCELLED, E_IO = -30001, -30002, -30009, -30010


@pytest.fixture(scope="module")
def ntsc():
    fmt = formats.get_format("NTSC", "VHS", "sp")
    return fmt, filters.build_tape_filters(fmt, 40e6, 32768)


def test_run_struct_sizes_and_null_arguments():
    L = _lib.load()
    assert L.vhsdecode_hip_run(None, None) == E_INVALID
    cfg = _lib.RunConfigV1()
    res = _lib.RunResultV1()
    cfg.struct_size = ctypes.sizeof(cfg) - 8
    res.struct_size = ctypes.sizeof(res)
    assert L.vhsdecode_hip_run(ctypes.byref(cfg), ctypes.byref(res)) == E_STRUCT
    cfg.struct_size = ctypes.sizeof(cfg)
    assert L.vhsdecode_hip_run(ctypes.byref(cfg), ctypes.byref(res)) == E_INVALID  # no format


def test_run_refuses_existing_output(ntsc, tmp_path):
    fmt, fl = ntsc
    base = str(tmp_path / "capture")
    open(base + ".tbc", "wb").close()
    with pytest.raises(_lib.HipError) as e:
        decode_run(np.zeros(200_000, np.int16), fmt, fl, base, overwrite=False)
    assert e.value.status == E_IO


def test_run_cancel_before_start(ntsc, tmp_path):
    fmt, fl = ntsc
    with pytest.raises(_lib.HipError) as e:
        decode_run(np.zeros(200_000, np.int16), fmt, fl, str(tmp_path / "c"), cancel=lambda: True)
    assert e.value.status == E_CANCELLED
    assert not os.path.exists(str(tmp_path / "c.tbc"))


def test_run_rejects_bad_input_type(ntsc, tmp_path):
    fmt, fl = ntsc
    with pytest.raises(TypeError):
        decode_run(np.zeros(10, np.float64), fmt, fl, str(tmp_path / "d"))
