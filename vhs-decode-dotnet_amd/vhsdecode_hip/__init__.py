"""VHSDecode.Hip.Native host package: MI355X per-block RF decode chain.

The compute path is the C-ABI library lib/libvhsdecode_hip.so (hand-written gfx950
HIP kernels); this package holds its ctypes binding, the host-side filter-table
builder, the format catalogue subset and the synthetic RF generator.
"""
from . import formats, filters  # noqa: F401

__all__ = ["formats", "filters", "synth", "demod", "_lib"]
