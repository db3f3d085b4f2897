"""Host-side mirror of the reference's block decode interface over the HIP backend.

`HipBlockDemodulator` plays the role of RfBlockDecodePipeline + RfBlockStreamDecoder
for the VHS real-FFT path (src/VHSDecode.Core/Decode/RfBlockDecodePipeline.cs:246-383,
Decode/RfBlockStreamDecoder.cs:419-700): it owns a native demod context built from a
`TapeFilterSet` (the DecodeFilterSet subset) and returns stitched field-span planes
(`RfDecodedSpan`, RfBlockStreamDecoder.cs:7-24).  Pulse detection, hsync refinement
and TBC resampling run on the same context.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Optional, Tuple

import numpy as np

from . import _lib
from .filters import TapeFilterSet

BLOCK_CUT = 1024       # DecodeSession.cs:259-260
BLOCK_CUT_END = 1024


def _ptr(a: np.ndarray, t):
    return a.ctypes.data_as(t)


@dataclass
class DecodedSpan:
    """Stitched demodulated planes of consecutive blocks (RfDecodedSpan)."""

    video: np.ndarray             # FP32 offsets from `bias` (Hz)
    video_lowpass: np.ndarray     # FP32 offsets from `bias` (Hz)
    envelope: np.ndarray
    chroma: Optional[np.ndarray]  # LaserDisc: the colour-burst reference (absolute Hz)
    block_flags: np.ndarray
    bias: float = 0.0
    pilot: Optional[np.ndarray] = None        # LaserDisc PAL pilot reference (absolute Hz)
    audio_left: Optional[np.ndarray] = None   # LaserDisc analog audio (Hz at the slice rate)
    audio_right: Optional[np.ndarray] = None

    @property
    def video_hz(self) -> np.ndarray:
        return self.video.astype(np.float64) + self.bias

    @property
    def video_lowpass_hz(self) -> np.ndarray:
        return self.video_lowpass.astype(np.float64) + self.bias

    @property
    def weak_rf(self) -> bool:
        return bool((self.block_flags & 1).any())


class FilterTables:
    """A TapeFilterSet as the C ABI's vhsdecode_hip_filters_v1 (plus its sub-deemphasis
    block), with the arrays it points into kept alive alongside."""

    def __init__(self, filters):
        if hasattr(filters, "rf_mtf"):
            self._init_laserdisc(filters)
            return
        self._rfv = np.ascontiguousarray(filters.rf_video, dtype=np.complex128).view(np.float64)
        self._vid = np.ascontiguousarray(filters.video, dtype=np.complex128).view(np.float64)
        self._vlp = np.ascontiguousarray(filters.video_lowpass05, dtype=np.complex128).view(np.float64)
        self._env = np.ascontiguousarray(filters.envelope_sos, dtype=np.float64)
        self._chr = None if filters.chroma_sos is None else np.ascontiguousarray(filters.chroma_sos, dtype=np.float64)
        f = _lib.FiltersV1()
        f.struct_size = ctypes.sizeof(_lib.FiltersV1)
        f.block_length = int(filters.block_length)
        f.sample_rate_hz = float(filters.sample_rate_hz)
        f.rf_video = _ptr(self._rfv, _lib._dp)
        f.video = _ptr(self._vid, _lib._dp)
        f.video_lowpass05 = _ptr(self._vlp, _lib._dp)
        f.video_lowpass05_offset = int(filters.video_lowpass05_offset)
        f.envelope_sos = _ptr(self._env, _lib._dp)
        f.envelope_sections = self._env.shape[0]
        f.chroma_sos = _ptr(self._chr, _lib._dp) if self._chr is not None else None
        f.chroma_sections = 0 if self._chr is None else self._chr.shape[0]
        f.chroma_offset = int(filters.chroma_offset)
        f.diff_demod_max = float(filters.diff_demod_max) if filters.diff_demod_max is not None else -1.0
        f.plane_bias_hz = float(filters.plane_bias_hz)
        sd = getattr(filters, "sub_deemphasis", None)
        if sd is not None:
            self._sd_hp = np.ascontiguousarray(sd.high_pass, dtype=np.complex128).view(np.float64)
            self._sd_sos = np.ascontiguousarray(sd.amplitude_sos, dtype=np.float64)
            d = _lib.SubDeemphV1()
            d.struct_size = ctypes.sizeof(_lib.SubDeemphV1)
            d.high_pass = _ptr(self._sd_hp, _lib._dp)
            d.amplitude_sos = _ptr(self._sd_sos, _lib._dp)
            (d.deviation, d.exponential_scaling, d.scaling1, d.scaling2, d.logistic_mid, d.logistic_rate,
             d.static_factor) = (float(v) for v in sd.params())
            self._sd = d
            f.sub_deemphasis = ctypes.pointer(d)
        bo = getattr(filters, "rf_high_boost", None)
        if bo is not None:
            self._boost = np.ascontiguousarray(bo.rf_top_sos, dtype=np.float64).reshape(-1)
            f.rf_boost_sos = _ptr(self._boost, _lib._dp)
            f.rf_boost_multiplier = float(bo.multiplier)
        nt = getattr(filters, "fsc_notch", None)
        if nt is not None:
            self._notch = np.ascontiguousarray(np.concatenate([nt[0], nt[1]]), dtype=np.float64)
            f.video_notch = _ptr(self._notch, _lib._dp)
        self.struct = f

    def _init_laserdisc(self, filters) -> None:
        """A LaserDiscFilterSet (ld_filters.build_laserdisc_filters) as filters_v1 with its
        laserdisc_v1 and ld_audio_v1 blocks."""
        def c(a):
            return None if a is None else np.ascontiguousarray(a, dtype=np.complex128).view(np.float64)

        self._rfv, self._vid, self._vlp = c(filters.rf_video), c(filters.video), c(filters.video_lowpass05)
        self._mtf, self._burst, self._pilot = c(filters.rf_mtf), c(filters.video_burst), c(filters.video_pilot)
        f = _lib.FiltersV1()
        f.struct_size = ctypes.sizeof(_lib.FiltersV1)
        f.block_length = int(filters.block_length)
        f.sample_rate_hz = float(filters.sample_rate_hz)
        f.rf_video = _ptr(self._rfv, _lib._dp)
        f.video = _ptr(self._vid, _lib._dp)
        f.video_lowpass05 = _ptr(self._vlp, _lib._dp)
        f.video_lowpass05_offset = int(filters.video_lowpass05_offset)
        f.diff_demod_max = -1.0
        f.plane_bias_hz = float(filters.plane_bias_hz)
        ld = _lib.LaserDiscV1()
        ld.struct_size = ctypes.sizeof(_lib.LaserDiscV1)
        ld.rf_mtf = _ptr(self._mtf, _lib._dp)
        ld.clip_low_hz = float(filters.clip_low_hz)
        ld.clip_high_hz = float(filters.clip_high_hz)
        ld.video_burst = _ptr(self._burst, _lib._dp) if self._burst is not None else None
        ld.video_burst_offset = int(filters.video_burst_offset)
        ld.video_pilot = _ptr(self._pilot, _lib._dp) if self._pilot is not None else None
        if filters.audio_left is not None:
            self._st1 = [c(filters.audio_left.stage1), c(filters.audio_right.stage1)]
            au = _lib.LdAudioV1()
            au.struct_size = ctypes.sizeof(_lib.LdAudioV1)
            au.bin_count = int(filters.audio_left.bin_count)
            if filters.audio_right.bin_count != au.bin_count:
                raise ValueError("LD analog audio channel lengths did not match.")
            au.low_bin[0] = int(filters.audio_left.low_bin)
            au.low_bin[1] = int(filters.audio_right.low_bin)
            au.stage1[0] = _ptr(self._st1[0], _lib._dp)
            au.stage1[1] = _ptr(self._st1[1], _lib._dp)
            au.slice_rate_hz = float(filters.audio_left.slice_rate_hz)
            au.low_frequency_hz[0] = float(filters.audio_left.low_frequency_hz)
            au.low_frequency_hz[1] = float(filters.audio_right.low_frequency_hz)
            self._audio = au
            ld.analog_audio = ctypes.pointer(au)
        self._ld = ld
        f.laserdisc = ctypes.pointer(ld)
        self.struct = f


class HipBlockDemodulator:
    def __init__(self, filters: TapeFilterSet, block_cut: int = BLOCK_CUT, block_cut_end: int = BLOCK_CUT_END,
                 max_blocks: int = 64, device: int = 0, sinc_table: Optional[np.ndarray] = None):
        L = _lib.load()
        self.filters = filters
        self.block_length = filters.block_length
        self.block_cut = block_cut
        self.block_cut_end = block_cut_end
        self.stride = self.block_length - block_cut - block_cut_end
        self.max_blocks = max_blocks
        self.device = device
        # keep FP64 tables alive for the duration of the create call
        self._tables = FilterTables(filters)
        f = self._tables.struct
        self.bias = float(filters.plane_bias_hz)
        cfg = _lib.DemodConfigV1()
        cfg.struct_size = ctypes.sizeof(_lib.DemodConfigV1)
        cfg.device_id = device
        cfg.block_cut = block_cut
        cfg.block_cut_end = block_cut_end
        cfg.max_blocks = max_blocks
        cfg.filters = ctypes.pointer(f)
        handle = ctypes.c_void_p()
        _lib.check(L.vhsdecode_hip_demod_create(ctypes.byref(cfg), ctypes.byref(handle)))
        self._h = handle
        self._L = L
        if sinc_table is not None:
            self.set_sinc_table(sinc_table)

    def close(self):
        if getattr(self, "_h", None):
            self._L.vhsdecode_hip_demod_destroy(self._h)
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

    # ------------------------------------------------------------------ blocks
    def blocks_for(self, rf_len: int, first_sample: int = 0) -> int:
        return max(0, (rf_len - first_sample - self.block_length) // self.stride + 1)

    def demod_span_host(self, rf: np.ndarray, first_sample: int = 0, nblocks: Optional[int] = None,
                        chroma: bool = True) -> DecodedSpan:
        rf = np.ascontiguousarray(rf, dtype=np.int16)
        if nblocks is None:
            nblocks = self.blocks_for(rf.size, first_sample)
        plane = nblocks * self.stride
        out = {k: np.empty(plane, dtype=np.float32) for k in ("video", "vlp", "env")}
        ld = hasattr(self.filters, "rf_mtf")
        has_chroma = (self.filters.video_burst is not None) if ld else (self.filters.chroma_sos is not None)
        ch = np.empty(plane, dtype=np.float32) if (chroma and has_chroma) else None
        pilot = np.empty(plane, dtype=np.float32) if ld and self.filters.video_pilot is not None else None
        aud = None
        if ld and self.filters.audio_left is not None:
            aud = [np.empty(plane // self.filters.audio_decimation, dtype=np.float32) for _ in range(2)]
        flags = np.zeros(nblocks, dtype=np.uint32)
        span = _lib.SpanV1()
        span.struct_size = ctypes.sizeof(_lib.SpanV1)
        span.video = _ptr(out["video"], _lib._fp)
        span.video_lowpass = _ptr(out["vlp"], _lib._fp)
        span.envelope = _ptr(out["env"], _lib._fp)
        span.chroma = _ptr(ch, _lib._fp) if ch is not None else None
        span.block_flags = _ptr(flags, _lib._u32p)
        span.pilot = _ptr(pilot, _lib._fp) if pilot is not None else None
        if aud is not None:
            span.audio_left = _ptr(aud[0], _lib._fp)
            span.audio_right = _ptr(aud[1], _lib._fp)
        _lib.check(self._L.vhsdecode_hip_demod_span_host(self._h, _ptr(rf, _lib._i16p), rf.size, first_sample,
                                                         nblocks, ctypes.byref(span)))
        return DecodedSpan(out["video"], out["vlp"], out["env"], ch, flags, self.bias, pilot,
                           aud[0] if aud else None, aud[1] if aud else None)

    def demod_span_device(self, rf_ptr: int, first_sample: int, nblocks: int, video_ptr: int, vlp_ptr: int,
                          env_ptr: int, chroma_ptr: int = 0, flags_ptr: int = 0, stream: int = 0,
                          pilot_ptr: int = 0, audio_left_ptr: int = 0, audio_right_ptr: int = 0,
                          chroma_stream: int = 0):
        """Device-pointer entry (HBM addresses, e.g. hipmem.DeviceBuffer.ptr); asynchronous on `stream`.
        LaserDisc: chroma_ptr receives the burst reference, pilot_ptr the PAL pilot and the audio
        pointers nblocks * stride / decimation samples each."""
        span = _lib.SpanV1()
        span.struct_size = ctypes.sizeof(_lib.SpanV1)
        span.video = ctypes.cast(ctypes.c_void_p(video_ptr), _lib._fp)
        span.video_lowpass = ctypes.cast(ctypes.c_void_p(vlp_ptr), _lib._fp)
        span.envelope = ctypes.cast(ctypes.c_void_p(env_ptr), _lib._fp)
        span.chroma = ctypes.cast(ctypes.c_void_p(chroma_ptr), _lib._fp) if chroma_ptr else None
        span.block_flags = ctypes.cast(ctypes.c_void_p(flags_ptr), _lib._u32p) if flags_ptr else None
        span.pilot = ctypes.cast(ctypes.c_void_p(pilot_ptr), _lib._fp) if pilot_ptr else None
        if audio_left_ptr and audio_right_ptr:
            span.audio_left = ctypes.cast(ctypes.c_void_p(audio_left_ptr), _lib._fp)
            span.audio_right = ctypes.cast(ctypes.c_void_p(audio_right_ptr), _lib._fp)
        span.chroma_stream = chroma_stream or None
        _lib.check(self._L.vhsdecode_hip_demod_span_device(self._h, ctypes.c_void_p(rf_ptr), first_sample, nblocks,
                                                           ctypes.byref(span), ctypes.c_void_p(stream)))

    # ------------------------------------------------------------------ sync
    def find_pulses(self, sync_ref_ptr: int, length: int, threshold: float, min_length: int, max_length: int,
                    capacity: int = 65536) -> Tuple[np.ndarray, np.ndarray]:
        starts = np.empty(capacity, dtype=np.int32)
        lens = np.empty(capacity, dtype=np.int32)
        count = ctypes.c_int32()
        _lib.check(self._L.vhsdecode_hip_find_pulses(self._h, ctypes.c_void_p(sync_ref_ptr), length, threshold,
                                                     min_length, max_length, _ptr(starts, _lib._i32p),
                                                     _ptr(lens, _lib._i32p), capacity, ctypes.byref(count)))
        n = count.value
        if n > capacity:
            raise _lib.HipError(-30007, f"{n} pulses exceed the capacity {capacity}")
        return starts[:n].copy(), lens[:n].copy()

    def refine_hsync(self, sync_ref_ptr: int, length: int, locations: np.ndarray, filled: np.ndarray,
                     threshold: float, sample_rate_mhz: float, hsync_pulse_us: float, ire0: float, hz_ire: float,
                     is_pal: bool, use_right_hsync: bool = True):
        loc = np.ascontiguousarray(locations, dtype=np.float64)
        fil = np.ascontiguousarray(filled, dtype=np.uint8)
        ro = np.empty_like(loc)
        fo = np.empty_like(fil)
        p = _lib.HsyncParamsV1(ctypes.sizeof(_lib.HsyncParamsV1), threshold, sample_rate_mhz, hsync_pulse_us, ire0,
                               hz_ire, 1 if is_pal else 0, 1 if use_right_hsync else 0)
        _lib.check(self._L.vhsdecode_hip_refine_hsync(self._h, ctypes.c_void_p(sync_ref_ptr), length,
                                                      _ptr(loc, _lib._dp), _ptr(fil, _lib._u8p), loc.size,
                                                      ctypes.byref(p), _ptr(ro, _lib._dp), _ptr(fo, _lib._u8p)))
        return ro, fo.astype(bool)

    # ------------------------------------------------------------------ TBC
    def set_sinc_table(self, table: np.ndarray):
        t = np.ascontiguousarray(table, dtype=np.float32).reshape(-1)
        _lib.check(self._L.vhsdecode_hip_set_sinc_table(self._h, _ptr(t, _lib._fp), t.size))

    def tbc_resample(self, source_ptr: int, source_length: int, locations: np.ndarray, first_line: int,
                     line_count: int, output_line_length: int, nominal: float = 0.0, smoothing: float = 0.0,
                     conv: Optional[tuple] = None, want_f64: bool = True):
        """conv = (ire0, hz_ire, output_zero, vsync_ire, output_scale) or None."""
        loc = np.ascontiguousarray(locations, dtype=np.float64)
        dest = output_line_length * line_count
        u16 = np.empty(dest, dtype=np.uint16) if conv is not None else None
        f64 = np.empty(dest, dtype=np.float64) if want_f64 else None
        c = conv if conv is not None else (0.0, 1.0, 0, 0.0, 1.0)
        p = _lib.TbcParamsV1(ctypes.sizeof(_lib.TbcParamsV1), output_line_length, first_line, line_count,
                             float(nominal), float(smoothing), 1 if conv is not None else 0, float(c[0]), float(c[1]),
                             int(c[2]), float(c[3]), float(c[4]))
        _lib.check(self._L.vhsdecode_hip_tbc_resample(self._h, ctypes.c_void_p(source_ptr), source_length,
                                                      _ptr(loc, _lib._dp), loc.size, ctypes.byref(p),
                                                      _ptr(u16, _lib._u16p) if u16 is not None else None,
                                                      _ptr(f64, _lib._dp) if f64 is not None else None))
        return u16, f64

    def field_params(self, geom):
        ire0, hz_ire, ozero, vsync, oscale = geom.conv
        return _lib.FieldParamsV1(ctypes.sizeof(_lib.FieldParamsV1), geom.threshold, geom.fs_mhz,
                                  geom.hsync / geom.fs_mhz, ire0, hz_ire, 1 if geom.is_pal else 0, 1, geom.out_len,
                                  geom.field_lines, geom.nominal, geom.smoothing, ozero, vsync, oscale)

    def decode_fields_device(self, sync_ref_ptr: int, video_ptr: int, span_length: int, locations: np.ndarray,
                             filled: np.ndarray, geom, out_ptr: int, origins: Optional[np.ndarray] = None,
                             want_refined: bool = False, stream: int = 0):
        """Batched hsync refinement + level adjusts + TBC for fields [nfields, lines + 1] of line locations;
        uint16 luma to device memory at out_ptr ([nfields, lines * out_len])."""
        loc = np.ascontiguousarray(locations, dtype=np.float64)
        fil = np.ascontiguousarray(filled, dtype=np.uint8)
        nfields = loc.shape[0]
        p = self.field_params(geom)
        ref = np.empty_like(loc) if want_refined else None
        org = None if origins is None else np.ascontiguousarray(origins, dtype=np.int64)
        _lib.check(self._L.vhsdecode_hip_decode_fields_device(
            self._h, ctypes.c_void_p(sync_ref_ptr), ctypes.c_void_p(video_ptr), span_length, _ptr(loc, _lib._dp),
            _ptr(fil, _lib._u8p), _ptr(org, ctypes.POINTER(ctypes.c_int64)) if org is not None else None, nfields,
            ctypes.byref(p), ctypes.c_void_p(out_ptr),
            _ptr(ref, _lib._dp) if ref is not None else None, ctypes.c_void_p(stream)))
        return ref

    def decode_vhs_fields_device(self, sync_ref_ptr: int, video_ptr: int, span_length: int, first_begin: int,
                                 sync, geom, max_fields: int, out_ptr: int, count_ptr: int, fields_ptr: int = 0,
                                 next_begin_ptr: int = 0, reset_state: bool = True, state_shift: int = 0,
                                 rounds: int = 4, stream: int = 0):
        """The vhs field sequence from `first_begin` on: read windows, sync levels, pulses,
        parity, line 0, line locations, hsync refinement and TBC, all on the device and
        asynchronous.  `sync` is a VhsSyncV1; records (VhsFieldV1) land at fields_ptr and
        their count at count_ptr."""
        p = self.field_params(geom)
        v = ctypes.c_void_p
        _lib.check(self._L.vhsdecode_hip_decode_vhs_fields_device(
            self._h, v(sync_ref_ptr), v(video_ptr), span_length, first_begin, 1 if reset_state else 0, state_shift,
            ctypes.byref(sync), ctypes.byref(p), max_fields, rounds, v(out_ptr), v(fields_ptr or None), v(count_ptr),
            v(next_begin_ptr or None), v(stream)))

    def set_vhs_chroma(self, params, input_line_length: int):
        """Colour-under chroma for the vhs field sequence (`params` a chroma.ChromaParams, or
        None to clear): heterodyne and burst-carrier tables built on the device."""
        from .chroma import to_struct
        if params is None:
            _lib.check(self._L.vhsdecode_hip_set_vhs_chroma(self._h, None))
            return
        cp = to_struct(params, input_line_length)
        _lib.check(self._L.vhsdecode_hip_set_vhs_chroma(self._h, ctypes.byref(cp)))

    def decode_vhs_fields_chroma_device(self, sync_ref_ptr: int, video_ptr: int, chroma_ptr: int, span_length: int,
                                        first_begin: int, sync, geom, max_fields: int, out_ptr: int,
                                        out_chroma_ptr: int, count_ptr: int, fields_ptr: int = 0,
                                        next_begin_ptr: int = 0, reset_state: bool = True, state_shift: int = 0,
                                        rounds: int = 4, field_number_base: int = 0, stream: int = 0):
        """decode_vhs_fields_device with the colour-under chroma of every decoded field
        (uint16 at out_chroma_ptr, slot by slot as the luma); the luma renders at the
        burst-refined line locations."""
        p = self.field_params(geom)
        v = ctypes.c_void_p
        _lib.check(self._L.vhsdecode_hip_decode_vhs_fields_chroma_device(
            self._h, v(sync_ref_ptr), v(video_ptr), v(chroma_ptr), span_length, first_begin,
            1 if reset_state else 0, state_shift, ctypes.byref(sync), ctypes.byref(p), max_fields, rounds,
            v(out_ptr), v(out_chroma_ptr), field_number_base, v(fields_ptr or None), v(count_ptr),
            v(next_begin_ptr or None), v(stream)))

    def vhs_field_lines_device(self, which: int):
        """(device pointer, stride) of a line table of the last vhs field decode, by record:
        0 line locations, 1 hsync-refined, 2 render locations (window coordinates), 3 render
        locations from the output first line (span coordinates)."""
        ptr = ctypes.c_void_p()
        stride = ctypes.c_int32()
        _lib.check(self._L.vhsdecode_hip_vhs_field_lines_device(self._h, which, ctypes.byref(ptr),
                                                                ctypes.byref(stride)))
        return int(ptr.value), int(stride.value)

    def field_locations_device(self):
        """(device pointer, locations per field) of the refined line locations of the last
        field decode on this context; valid until the next field decode."""
        ptr = ctypes.c_void_p()
        cnt = ctypes.c_int32()
        _lib.check(self._L.vhsdecode_hip_field_locations_device(self._h, ctypes.byref(ptr), ctypes.byref(cnt)))
        return int(ptr.value), int(cnt.value)

    def detect_dropouts_device(self, envelope_ptr: int, span_length: int, locations_ptr: int, location_count: int,
                               fields_ptr: int, count_ptr: int, max_fields: int, params, row_count_ptr: int,
                               field_line_ptr: int, start_x_ptr: int, end_x_ptr: int, range_count_ptr: int = 0,
                               range_start_ptr: int = 0, range_end_ptr: int = 0, stream: int = 0):
        """Tape RF-envelope dropout detection of the records (VhsFieldV1) of a vhs field
        decode (every pointer device memory; asynchronous on `stream`): .tbc.json rows per
        record at [record][capacity]."""
        v = ctypes.c_void_p
        _lib.check(self._L.vhsdecode_hip_detect_dropouts_device(
            self._h, v(envelope_ptr), span_length, v(locations_ptr), location_count, v(fields_ptr), v(count_ptr),
            max_fields, ctypes.byref(params), v(row_count_ptr), v(field_line_ptr), v(start_x_ptr), v(end_x_ptr),
            v(range_count_ptr or None), v(range_start_ptr or None), v(range_end_ptr or None), v(stream)))

    def set_profiling(self, enable: bool = True):
        _lib.check(self._L.vhsdecode_hip_set_profiling(self._h, 1 if enable else 0))

    def take_profile(self) -> dict:
        """Per-kernel launches, device ms and units since the last call (HIP events)."""
        p = _lib.ProfileV1()
        p.struct_size = ctypes.sizeof(_lib.ProfileV1)
        _lib.check(self._L.vhsdecode_hip_get_profile(self._h, ctypes.byref(p)))
        return {name: {"launches": p.launches[i], "ms": p.total_ms[i], "units": p.units[i]}
                for i, name in enumerate(_lib.KERNEL_NAMES)}

    def set_ld_audio_phase2(self, filters):
        """LaserDisc analog audio stage 2 (LaserDiscAnalogAudioPhase2.Apply) for this context:
        `filters` a LaserDiscFilterSet with analog audio (its channels' stage-2 responses and
        carriers), or None to clear."""
        if filters is None:
            _lib.check(self._L.vhsdecode_hip_set_ld_audio_phase2(self._h, None))
            return
        chans = (filters.audio_left, filters.audio_right)
        st = [np.ascontiguousarray(np.asarray(c.stage2, np.complex128)).view(np.float64) for c in chans]
        p = _lib.LdAudioPhase2V1()
        p.struct_size = ctypes.sizeof(_lib.LdAudioPhase2V1)
        p.block_length = int(chans[0].stage2.size)
        for k, c in enumerate(chans):
            p.center_hz[k] = float(c.center_hz)
            p.stage2[k] = st[k].ctypes.data
        _lib.check(self._L.vhsdecode_hip_set_ld_audio_phase2(self._h, ctypes.byref(p)))

    def ld_audio_phase2_device(self, left_ptr: int, right_ptr: int, count: int, out_left_ptr: int,
                               out_right_ptr: int, stream: int = 0):
        """Stage 2 over `count` samples per channel of device FP32 audio (Hz at the slice rate)."""
        v = ctypes.c_void_p
        _lib.check(self._L.vhsdecode_hip_ld_audio_phase2_device(self._h, v(left_ptr), v(right_ptr), count,
                                                                v(out_left_ptr), v(out_right_ptr), v(stream)))

    def debug_fft16k(self, data_ptr: int, count: int, inverse: bool = False, stream: int = 0):
        _lib.check(self._L.vhsdecode_hip_debug_fft16k(self._h, ctypes.c_void_p(data_ptr), count,
                                                      1 if inverse else 0, ctypes.c_void_p(stream)))


def runtime_info(device: int = 0):
    L = _lib.load()
    info = _lib.RuntimeInfoV1()
    info.struct_size = ctypes.sizeof(_lib.RuntimeInfoV1)
    _lib.check(L.vhsdecode_hip_get_runtime_info(device, ctypes.byref(info)))
    return {"device_count": info.device_count, "compute_units": info.compute_units, "hbm_bytes": info.hbm_bytes,
            "lds_bytes_per_workgroup": info.lds_bytes_per_workgroup,
            "device_name": info.device_name.decode(), "gcn_arch": info.gcn_arch.decode()}
