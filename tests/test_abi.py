"""The C-ABI library builds, loads and exports every symbol include/vhsdecode_hip.h
declares; struct layouts agree with the header.  CPU only (no compute calls)."""
import ctypes
import os
import re
import subprocess

import pytest

from vhsdecode_hip import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "vhsdecode_hip.h")


def _declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"VHSDECODE_HIP_API\s+[\w\s\*]+?\b(vhsdecode_hip_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    assert os.path.exists(_lib.LIB_PATH), "build the library first"
    names = _declared_functions()
    assert len(names) >= 12
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True, check=True)
    exported = set(line.split()[-1] for line in out.stdout.splitlines() if line.strip())
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    assert set(_lib.REQUIRED_EXPORTS) == set(names)


def test_library_loads_and_reports_abi():
    L = _lib.load()
    assert L.vhsdecode_hip_get_abi_version() == 1


def _c_sizeof(struct):
    src = f'#include "vhsdecode_hip.h"\n#include <stdio.h>\nint main(){{printf("%zu", sizeof({struct}));return 0;}}\n'
    exe = "/tmp/_vhsd_sizeof"
    subprocess.run(["gcc", "-x", "c", "-I", os.path.join(ROOT, "include"), "-o", exe, "-"], input=src, text=True,
                   check=True)
    return int(subprocess.run([exe], capture_output=True, text=True, check=True).stdout)


@pytest.mark.parametrize("name,cls", [
    ("vhsdecode_hip_runtime_info_v1", _lib.RuntimeInfoV1),
    ("vhsdecode_hip_filters_v1", _lib.FiltersV1),
    ("vhsdecode_hip_demod_config_v1", _lib.DemodConfigV1),
    ("vhsdecode_hip_span_v1", _lib.SpanV1),
    ("vhsdecode_hip_hsync_params_v1", _lib.HsyncParamsV1),
    ("vhsdecode_hip_tbc_params_v1", _lib.TbcParamsV1),
    ("vhsdecode_hip_profile_v1", _lib.ProfileV1),
    ("vhsdecode_hip_field_params_v1", _lib.FieldParamsV1),
    ("vhsdecode_hip_subdeemph_v1", _lib.SubDeemphV1),
    ("vhsdecode_hip_serve_config_v1", _lib.ServeConfigV1),
    ("vhsdecode_hip_serve_format_v1", _lib.ServeFormatV1),
    ("vhsdecode_hip_window_v1", _lib.WindowV1),
    ("vhsdecode_hip_window_result_v1", _lib.WindowResultV1),
    ("vhsdecode_hip_laserdisc_v1", _lib.LaserDiscV1),
    ("vhsdecode_hip_ld_audio_v1", _lib.LdAudioV1),
    ("vhsdecode_hip_hifi_config_v1", _lib.HifiConfigV1),
    ("vhsdecode_hip_batch_segment_v1", _lib.BatchSegmentV1),
    ("vhsdecode_hip_batch_config_v1", _lib.BatchConfigV1),
    ("vhsdecode_hip_dropout_params_v1", _lib.DropoutParamsV1),
    ("vhsdecode_hip_run_config_v1", _lib.RunConfigV1),
    ("vhsdecode_hip_run_result_v1", _lib.RunResultV1),
    ("vhsdecode_hip_vhs_sync_v1", _lib.VhsSyncV1),
    ("vhsdecode_hip_vhs_field_v1", _lib.VhsFieldV1),
])
def test_struct_layouts_match_header(name, cls):
    assert _c_sizeof(name) == ctypes.sizeof(cls)


def test_struct_size_is_validated():
    L = _lib.load()
    info = _lib.RuntimeInfoV1()
    info.struct_size = 3
    rc = L.vhsdecode_hip_get_runtime_info(0, ctypes.byref(info))
    assert rc == -30002
    assert b"struct_size" in L.vhsdecode_hip_get_last_error()


def test_create_rejects_bad_config_without_touching_a_device():
    L = _lib.load()
    cfg = _lib.DemodConfigV1()
    cfg.struct_size = 1
    handle = ctypes.c_void_p()
    assert L.vhsdecode_hip_demod_create(ctypes.byref(cfg), ctypes.byref(handle)) == -30002
    assert not handle.value
    L.vhsdecode_hip_demod_destroy(None)  # destroying NULL is a no-op
