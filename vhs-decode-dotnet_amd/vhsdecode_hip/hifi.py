"""HiFi Hilbert FM discriminator on the device: the host mirror of
HiFiHilbertDiscriminator (src/VHSDecode.Core/HiFi/HiFiHilbertDiscriminator.cs:7-156)
over vhsdecode_hip_hifi_* (include/vhsdecode_hip.h).

Same constructor arguments and checks (sample rate, carrier rounded half to even,
deviation truncated) and the same Demodulate(input, output) contract: one block per
call by default, the output's last sample left as it was."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib


class HipHiFiHilbertDiscriminator:
    def __init__(self, sample_rate_hz: int, carrier_center_hz: float, deviation_hz: float, block_length: int,
                 device: int = 0):
        L = _lib.load()
        cfg = _lib.HifiConfigV1()
        cfg.struct_size = ctypes.sizeof(_lib.HifiConfigV1)
        cfg.device_id = device
        cfg.block_length = int(block_length)
        cfg.sample_rate_hz = int(sample_rate_hz)
        cfg.carrier_hz = float(carrier_center_hz)
        cfg.deviation_hz = float(deviation_hz)
        h = ctypes.c_void_p()
        _lib.check(L.vhsdecode_hip_hifi_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        self._L = L
        self.block_length = int(block_length)
        self.sample_rate_hz = int(sample_rate_hz)
        self.carrier_hz = int(np.rint(carrier_center_hz))
        self.deviation_hz = int(np.trunc(deviation_hz))

    def close(self):
        if getattr(self, "_h", None):
            self._L.vhsdecode_hip_hifi_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def demodulate(self, input: np.ndarray, output: np.ndarray) -> None:
        """Demodulate every whole block of `input` into `output` (float32, same length)."""
        x = np.ascontiguousarray(input, dtype=np.float32)
        if output.dtype != np.float32 or not output.flags["C_CONTIGUOUS"] or output.size < x.size:
            raise ValueError("output must be a contiguous float32 array at least as long as the input")
        if x.size % self.block_length:
            raise ValueError("input length must be a whole number of blocks")
        nblocks = x.size // self.block_length
        _lib.check(self._L.vhsdecode_hip_hifi_demodulate_host(self._h, x.ctypes.data_as(_lib._fp),
                                                              output.ctypes.data_as(_lib._fp), nblocks))

    def demodulate_device(self, input_ptr: int, output_ptr: int, nblocks: int, stream: int = 0) -> None:
        _lib.check(self._L.vhsdecode_hip_hifi_demodulate_device(self._h, ctypes.c_void_p(input_ptr),
                                                                ctypes.c_void_p(output_ptr), nblocks,
                                                                ctypes.c_void_p(stream)))
