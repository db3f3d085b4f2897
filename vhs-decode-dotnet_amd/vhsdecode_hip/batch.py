"""Batched decode of many RF segments of mixed formats and sample rates in one call
(vhsdecode_hip_batch_*, include/vhsdecode_hip.h): the segments' whole blocks are
grouped by demod context, each group is one launch of the block chain, the groups run
concurrently, and the batch replays as one captured HIP graph.

Each segment is read as the reference's stream decoder reads a capture
(RfBlockStreamDecoder.ReadCore, src/VHSDecode.Core/Decode/RfBlockStreamDecoder.cs:419-700):
whole 32768-sample blocks at the 30720 stride, kept windows stitched."""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import List, Sequence

from . import _lib
from .demod import HipBlockDemodulator


@dataclass
class Segment:
    demodulator: HipBlockDemodulator   # the format and sample rate
    rf_offset: int                     # first sample in the packed RF buffer (even)
    length: int
    out_offset: int                    # first kept sample in the span planes

    def blocks(self) -> int:
        d = self.demodulator
        return (self.length - d.block_length) // d.stride + 1 if self.length >= d.block_length else 0


def pack_segments(lengths: Sequence[int], demodulators: Sequence[HipBlockDemodulator]) -> List[Segment]:
    """Consecutive packing: RF offsets rounded up to even, output ranges back to back."""
    segs, rf, out = [], 0, 0
    for length, d in zip(lengths, demodulators):
        s = Segment(d, rf, int(length), out)
        segs.append(s)
        rf += (int(length) + 1) & ~1
        out += s.blocks() * d.stride
    return segs


class HipBatch:
    def __init__(self, segments: Sequence[Segment], use_graph: bool = True):
        L = _lib.load()
        arr = (_lib.BatchSegmentV1 * max(len(segments), 1))()
        for i, s in enumerate(segments):
            arr[i].struct_size = ctypes.sizeof(_lib.BatchSegmentV1)
            arr[i].context = s.demodulator._h
            arr[i].rf_offset = int(s.rf_offset)
            arr[i].length = int(s.length)
            arr[i].out_offset = int(s.out_offset)
        cfg = _lib.BatchConfigV1(ctypes.sizeof(_lib.BatchConfigV1), arr, len(segments), 1 if use_graph else 0)
        h = ctypes.c_void_p()
        _lib.check(L.vhsdecode_hip_batch_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h, self._L = h, L
        self.segments = list(segments)
        b, o, r = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        _lib.check(L.vhsdecode_hip_batch_info(self._h, ctypes.byref(b), ctypes.byref(o), ctypes.byref(r)))
        self.blocks, self.output_samples, self.rf_samples = b.value, o.value, r.value

    def run_device(self, rf_ptr: int, rf_count: int, video_ptr: int, vlp_ptr: int, env_ptr: int,
                   chroma_ptr: int = 0, stream: int = 0) -> None:
        span = _lib.SpanV1()
        span.struct_size = ctypes.sizeof(_lib.SpanV1)
        span.video = ctypes.cast(ctypes.c_void_p(video_ptr), _lib._fp)
        span.video_lowpass = ctypes.cast(ctypes.c_void_p(vlp_ptr), _lib._fp)
        span.envelope = ctypes.cast(ctypes.c_void_p(env_ptr), _lib._fp)
        span.chroma = ctypes.cast(ctypes.c_void_p(chroma_ptr), _lib._fp) if chroma_ptr else None
        _lib.check(self._L.vhsdecode_hip_batch_run_device(self._h, ctypes.c_void_p(rf_ptr), rf_count,
                                                          ctypes.byref(span), ctypes.c_void_p(stream)))

    def close(self):
        if getattr(self, "_h", None):
            self._L.vhsdecode_hip_batch_destroy(self._h)
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
