/*
 * VHSDecode.Hip.Native — C ABI of the MI355X (gfx950) per-block RF decode backend.
 *
 * Conventions follow the reference's native bridges
 * (src/VHSDecode.CudaFast.Native/include/vhsdecode_cuda_fast.h:1-180,
 *  src/VHSDecode.Ipp.Native/include/vhsdecode_ipp.h:1-366):
 *  - extern "C", cdecl, default visibility; no C++ or torch types cross the ABI;
 *  - every struct is versioned (_v1) and starts with uint32_t struct_size, which
 *    the library validates;
 *  - int32_t status: 0 = OK, negative codes in the -30000 range; exceptions never
 *    cross the ABI and vhsdecode_hip_get_last_error() returns a thread-local
 *    message for the last failure on the calling thread;
 *  - opaque context handles with create/destroy; destroying NULL is a no-op;
 *  - a context serialises its own calls; different contexts may run concurrently.
 *
 * Pointer arguments documented as "device" are HBM addresses on the context's
 * device (hipMalloc allocations, or any HBM allocation of the same HIP runtime) and the call is asynchronous on the
 * given stream (NULL = the context's own stream); "host" arguments are ordinary
 * process memory and the call is synchronous.
 */
#ifndef VHSDECODE_HIP_H
#define VHSDECODE_HIP_H

#include <stddef.h>
#include <stdint.h>

#if defined(_WIN32)
#define VHSDECODE_HIP_API __declspec(dllexport)
#else
#define VHSDECODE_HIP_API __attribute__((visibility("default")))
#endif

#ifdef __cplusplus
extern "C" {
#endif

#define VHSDECODE_HIP_ABI_VERSION 1

enum {
  VHSDECODE_HIP_OK = 0,
  VHSDECODE_HIP_E_INVALID_ARGUMENT = -30001,
  VHSDECODE_HIP_E_STRUCT_SIZE = -30002,
  VHSDECODE_HIP_E_NO_DEVICE = -30003,
  VHSDECODE_HIP_E_HIP = -30004,
  VHSDECODE_HIP_E_OUT_OF_MEMORY = -30005,
  VHSDECODE_HIP_E_UNSUPPORTED = -30006,
  VHSDECODE_HIP_E_CAPACITY = -30007,
  VHSDECODE_HIP_E_TIMEOUT = -30008,
  VHSDECODE_HIP_E_CANCELLED = -30009,      /* the run's cancel callback asked to stop */
  VHSDECODE_HIP_E_IO = -30010,             /* an output file could not be created or written */
  VHSDECODE_HIP_E_INTERNAL = -30099
};

typedef struct vhsdecode_hip_demod_context vhsdecode_hip_demod_context;

/* ---------------------------------------------------------------- runtime
 * Replaces vhsdecode_cuda_fast_get_abi_version / _get_runtime_info / _get_last_error
 * (vhsdecode_cuda_fast.h:145-152, vhsdecode_ipp.h:98-116); the managed loader checks
 * the ABI version together with the exported symbol list
 * (src/VHSDecode.Core/Dsp/Ipp/IppNativeMethods.cs:8-47, IppRuntime.cs:48-122). */
typedef struct vhsdecode_hip_runtime_info_v1 {
  uint32_t struct_size;
  int32_t device_count;
  int32_t device_id;
  int32_t compute_units;
  uint64_t hbm_bytes;
  int32_t lds_bytes_per_workgroup;
  char device_name[128];
  char gcn_arch[64];
} vhsdecode_hip_runtime_info_v1;

VHSDECODE_HIP_API int32_t vhsdecode_hip_get_abi_version(void);
VHSDECODE_HIP_API int32_t vhsdecode_hip_get_runtime_info(int32_t device_id, vhsdecode_hip_runtime_info_v1* info);
VHSDECODE_HIP_API const char* vhsdecode_hip_get_last_error(void);

/* ---------------------------------------------------------------- block demodulator
 * The exact/ipp-fast block seam: one call demodulates a run of overlap-save blocks
 * the way RfDemodulator.DemodulateCore does per block
 * (src/VHSDecode.Core/Dsp/RfDemodulator.cs:288-918, VHS real-FFT path) plus
 * RfBlockDecodePipeline.DecodeChromaBurstFloat32 (Decode/RfBlockDecodePipeline.cs:539-573),
 * and writes the trimmed windows stitched as RfBlockStreamDecoder.CopyTrimmedWindow
 * does (Decode/RfBlockStreamDecoder.cs:497-549).
 *
 * The filter tables are the caller's DecodeFilterSet (Dsp/DecodeFilterSet.cs:9-82):
 * full-circle complex128 spectra of block_length bins, interleaved (re, im); SOS
 * sections as [sections][6] = b0 b1 b2 a0 a1 a2. */
/* Sub-deemphasis of the video plane (RfDemodulator.ApplySubDeemphasisRealCore,
 * src/VHSDecode.Core/Dsp/RfDemodulator.cs:2897-2999; options of
 * Decode/DecodeSession.cs:1250-1278), enabled by use_sub_deemphasis for VHS LP/EP,
 * Video8 and others.  Absent options carry neutral values. */
typedef struct vhsdecode_hip_subdeemph_v1 {
  uint32_t struct_size;
  const double* high_pass;         /* BuildNonlinearHighPassFilter: whole-circle response, N complex (re, im) */
  const double* amplitude_sos;     /* [1][6]: ButterworthLowPass(1, amplitude low-pass / Nyquist) */
  double deviation;                /* SubDeemphasisOptions.Deviation (> 0) */
  double exponential_scaling;
  double scaling1;                 /* 1 when absent */
  double scaling2;                 /* 1 when absent */
  double logistic_mid;             /* 0 when absent */
  double logistic_rate;            /* <= 0: no logistic term */
  double static_factor;            /* 0 when absent */
} vhsdecode_hip_subdeemph_v1;

/* LaserDisc analog audio (DecodeFilterSetBuilder.BuildLdAnalogAudio,
 * src/VHSDecode.Core/Dsp/DecodeFilterSet.cs:527-562, 913-1017; decoded per block by
 * RfBlockDecodePipeline.DecodeAnalogAudioChannel, Decode/RfBlockDecodePipeline.cs:732-752,
 * and stitched by RfBlockStreamDecoder.CopyAudioTrimmedWindow, Decode/RfBlockStreamDecoder.cs:1690-1709). */
typedef struct vhsdecode_hip_ld_audio_v1 {
  uint32_t struct_size;
  int32_t bin_count;               /* slice length: a power of two <= 2048 dividing block_length, block_cut
                                      and the stride by the same decimation block_length / bin_count */
  int32_t low_bin[2];              /* left, right: first positive bin of each slice (DetermineFftSlice) */
  const double* stage1[2];         /* bin_count complex (re, im) each: the slice of the FIR band-pass
                                      response times the slice's Hilbert multiplier */
  double slice_rate_hz;            /* sample_rate_hz * bin_count / block_length */
  double low_frequency_hz[2];      /* sample_rate_hz * low_bin / block_length */
} vhsdecode_hip_ld_audio_v1;

/* The LaserDisc complex path (RfDemodulator.DemodulateCore with the conjugate-product
 * discriminator the ld command selects, src/VHSDecode.Core/Decode/DecodeSession.cs:1121-1123;
 * RfDemodulator.cs:564-605, 722-743, 769-775, 824-907). */
typedef struct vhsdecode_hip_laserdisc_v1 {
  uint32_t struct_size;
  const double* rf_mtf;            /* DecodeFilterSet.RfMtf (BuildLdMtf), NULL = all ones */
  double clip_low_hz;              /* ClipDemodForVideo bounds (1.5 MHz and 0.75 fs); clip_high_hz <= clip_low_hz: none */
  double clip_high_hz;
  const double* video_burst;       /* DecodeFilterSet.LdVideoBurst, NULL: none; written to span.chroma (absolute Hz) */
  int32_t video_burst_offset;      /* DecodeFilterSet.LdVideoBurstOffset (40): even, within the tail cut */
  const double* video_pilot;       /* DecodeFilterSet.LdVideoPilot (PAL), NULL: none; written to span.pilot */
  const vhsdecode_hip_ld_audio_v1* analog_audio;  /* NULL: no analog audio */
} vhsdecode_hip_laserdisc_v1;

typedef struct vhsdecode_hip_filters_v1 {
  uint32_t struct_size;
  int32_t block_length;            /* N; this build supports 32768 */
  double sample_rate_hz;           /* RF sample rate */
  const double* rf_video;          /* DecodeFilterSet.RfVideo (RfMtf is all-ones for tape) */
  const double* video;             /* DecodeFilterSet.Video */
  const double* video_lowpass05;   /* DecodeFilterSet.VideoLowPass05 */
  int32_t video_lowpass05_offset;  /* DecodeFilterSet.VideoLowPass05Offset (32) */
  const double* envelope_sos;      /* DecodeFilterSet.VhsEnvelopeSos */
  int32_t envelope_sections;
  const double* chroma_sos;        /* DecodeFilterSet.ChromaBurstSos, may be NULL */
  int32_t chroma_sections;
  int32_t chroma_offset;           /* DecodeFilterSet.ChromaOffsetSamples */
  double diff_demod_max;           /* DiffDemodRepairOptions.MaxValue; <= 0 disables the repair */
  double plane_bias_hz;            /* video and video_lowpass planes hold FP32 (value - plane_bias_hz); an
                                      FP32 plane at the carrier (~4 MHz) would keep only 1/4-1/2 Hz, offsets
                                      near a mid-level carrier keep 1/16 Hz.  0 writes absolute Hz. */
  const vhsdecode_hip_subdeemph_v1* sub_deemphasis;  /* NULL: the format does not use it */
  const double* rf_boost_sos;      /* RF high boost band-pass [1][6] (BuildVhsRfTopSos,
                                      Dsp/DecodeFilterSet.cs:622-641); NULL: no boost */
  double rf_boost_multiplier;      /* RfHighBoostOptions.Multiplier */
  const double* video_notch;       /* Betamax fsc notch transfer function b0 b1 b2 a0 a1 a2
                                      (RfDemodulator.cs:1018-1051); NULL: none.  Needs
                                      sub_deemphasis, which Betamax formats enable */
  const vhsdecode_hip_laserdisc_v1* laserdisc;  /* non-NULL: LaserDisc complex path.  envelope_sos, chroma_sos,
                                      sub_deemphasis, rf_boost_sos and video_notch are then ignored, the
                                      envelope plane holds |analytic| and diff_demod_max must be <= 0 */
} vhsdecode_hip_filters_v1;

typedef struct vhsdecode_hip_demod_config_v1 {
  uint32_t struct_size;
  int32_t device_id;
  int32_t block_cut;               /* head cut (1024) */
  int32_t block_cut_end;           /* tail cut (1024) */
  int32_t max_blocks;              /* largest run of blocks one call may submit */
  const vhsdecode_hip_filters_v1* filters;
} vhsdecode_hip_demod_config_v1;

/* Output planes of a stitched span: nblocks * stride samples each
 * (stride = block_length - block_cut - block_cut_end).  chroma may be NULL.
 * video and video_lowpass are offsets from the context's plane_bias_hz; the sync,
 * hsync and TBC entries below take planes in that form and add the bias in FP64. */
typedef struct vhsdecode_hip_span_v1 {
  uint32_t struct_size;
  float* video;
  float* video_lowpass;            /* sync reference, rolled by -video_lowpass05_offset */
  float* envelope;
  float* chroma;
  uint32_t* block_flags;           /* [nblocks]: bit0 weak RF (an envelope sample == 0), bit1 spike repaired */
  float* pilot;                    /* LaserDisc PAL pilot reference (absolute Hz), NULL: not written */
  float* audio_left;               /* LaserDisc analog audio, nblocks * stride / decimation samples each (Hz), */
  float* audio_right;              /* NULL: not written */
  void* chroma_stream;             /* device entry only: non-NULL computes the chroma plane on this stream,
                                      after the work already enqueued on the call's stream, so it overlaps the
                                      rest of the chain and later work; the caller orders its own uses of the
                                      plane.  The context's next span call waits for it. NULL: the call's stream */
} vhsdecode_hip_span_v1;

VHSDECODE_HIP_API int32_t vhsdecode_hip_demod_create(const vhsdecode_hip_demod_config_v1* config,
                                                     vhsdecode_hip_demod_context** context);
VHSDECODE_HIP_API void vhsdecode_hip_demod_destroy(vhsdecode_hip_demod_context* context);

/* Block b (0 <= b < nblocks) reads RF samples [first_sample + b*stride, + block_length)
 * of `rf` and writes stitched outputs [b*stride, (b+1)*stride).
 * Device variant: rf and every span pointer are device memory; asynchronous on `stream`. */
VHSDECODE_HIP_API int32_t vhsdecode_hip_demod_span_device(vhsdecode_hip_demod_context* context, const int16_t* rf,
                                                          int64_t first_sample, int32_t nblocks,
                                                          const vhsdecode_hip_span_v1* out, void* stream);
/* Host variant: rf holds rf_count samples in host memory; outputs are host memory. */
VHSDECODE_HIP_API int32_t vhsdecode_hip_demod_span_host(vhsdecode_hip_demod_context* context, const int16_t* rf,
                                                        int64_t rf_count, int64_t first_sample, int32_t nblocks,
                                                        const vhsdecode_hip_span_v1* out);

/* ---------------------------------------------------------------- HiFi Hilbert discriminator
 * HiFiHilbertDiscriminator (src/VHSDecode.Core/HiFi/HiFiHilbertDiscriminator.cs:7-156): the
 * phase of each block's analytic signal by a circular FFT of the whole block
 * (ComputeAnalyticPhase :48-86), then one output per phase step (DemodulatePhase :88-156).
 * Blocks are demodulated independently, as HiFiBlockDecoder calls Demodulate per block
 * (HiFi/HiFiBlockDecoder.cs:186-214); the reference's Hilbert IF block is 2^22 samples
 * (2^23 Hz, HiFiConstants.cs:55-57). */
typedef struct vhsdecode_hip_hifi vhsdecode_hip_hifi;

typedef struct vhsdecode_hip_hifi_config_v1 {
  uint32_t struct_size;
  int32_t device_id;
  int64_t block_length;            /* 16384 * 2^k, k = 0 .. 8 (2^14 .. 2^22 samples) */
  int32_t sample_rate_hz;          /* IF rate (> 0) */
  double carrier_hz;               /* rounded half-to-even to whole Hz, as the reference's constructor */
  double deviation_hz;             /* truncated to whole Hz */
} vhsdecode_hip_hifi_config_v1;

VHSDECODE_HIP_API int32_t vhsdecode_hip_hifi_create(const vhsdecode_hip_hifi_config_v1* config,
                                                    vhsdecode_hip_hifi** hifi);
VHSDECODE_HIP_API void vhsdecode_hip_hifi_destroy(vhsdecode_hip_hifi* hifi);
/* input and output: nblocks * block_length floats of device memory; asynchronous on
 * `stream` (NULL: the context's stream).  The last output sample of each block is not
 * written, as in the reference. */
VHSDECODE_HIP_API int32_t vhsdecode_hip_hifi_demodulate_device(vhsdecode_hip_hifi* hifi, const float* input,
                                                               float* output, int32_t nblocks, void* stream);
/* Host variant (synchronous). */
VHSDECODE_HIP_API int32_t vhsdecode_hip_hifi_demodulate_host(vhsdecode_hip_hifi* hifi, const float* input,
                                                             float* output, int32_t nblocks);

/* ---------------------------------------------------------------- batched segments
 * Many RF segments of mixed formats and sample rates (one demod context per format
 * and rate) decoded by one call.  Each segment is read as RfBlockStreamDecoder reads a
 * capture (Decode/RfBlockStreamDecoder.cs:419-700): whole blocks at the context's
 * stride from rf_offset while a whole block fits in `length` (a segment shorter than a
 * block yields nothing), each block's kept window stitched at out_offset + b * stride
 * of the shared span planes.  The blocks of a context form one launch of the block
 * chain; the launches of different contexts run concurrently, and with use_graph the
 * whole batch is captured once into a HIP graph and replayed while the rf and span
 * pointers stay the same.  The batch keeps its own descriptors and per-block scratch;
 * it must not outlive its contexts. */
typedef struct vhsdecode_hip_batch vhsdecode_hip_batch;

typedef struct vhsdecode_hip_batch_segment_v1 {
  uint32_t struct_size;
  vhsdecode_hip_demod_context* context;  /* decode format and sample rate */
  int64_t rf_offset;               /* first sample of the segment in the packed RF buffer (even) */
  int64_t length;                  /* samples in the segment */
  int64_t out_offset;              /* first kept sample of the segment in the span planes; output ranges must
                                      not overlap (LaserDisc audio: a multiple of the decimation) */
} vhsdecode_hip_batch_segment_v1;

typedef struct vhsdecode_hip_batch_config_v1 {
  uint32_t struct_size;
  const vhsdecode_hip_batch_segment_v1* segments;
  int32_t count;
  int32_t use_graph;               /* 1: capture the batch into a graph on the first run and replay it */
} vhsdecode_hip_batch_config_v1;

VHSDECODE_HIP_API int32_t vhsdecode_hip_batch_create(const vhsdecode_hip_batch_config_v1* config,
                                                     vhsdecode_hip_batch** batch);
/* Blocks decoded per run, the span plane length the segments need and the RF samples they read. */
VHSDECODE_HIP_API int32_t vhsdecode_hip_batch_info(const vhsdecode_hip_batch* batch, int64_t* blocks,
                                                   int64_t* output_samples, int64_t* rf_samples);
/* rf and every span plane are device memory; asynchronous on `stream` (NULL: the batch's stream). */
VHSDECODE_HIP_API int32_t vhsdecode_hip_batch_run_device(vhsdecode_hip_batch* batch, const int16_t* rf,
                                                         int64_t rf_count, const vhsdecode_hip_span_v1* out,
                                                         void* stream);
VHSDECODE_HIP_API void vhsdecode_hip_batch_destroy(vhsdecode_hip_batch* batch);

/* ---------------------------------------------------------------- sync pulses
 * PulseDetection.FindPulses (src/VHSDecode.Core/Dsp/PulseDetection.cs:45-222) with the
 * VHS bounds of TbcFieldDecodePipeline.FindRawPulses (Decode/TbcFieldDecodePipeline.cs:2375-2384).
 * sync_ref is device memory; starts/lengths/count are host memory (synchronous).
 * *count receives the number of pulses found; only min(count, capacity) are stored. */
VHSDECODE_HIP_API int32_t vhsdecode_hip_find_pulses(vhsdecode_hip_demod_context* context, const float* sync_ref,
                                                    int64_t length, double threshold, int32_t min_length,
                                                    int32_t max_length, int32_t* starts, int32_t* lengths,
                                                    int32_t capacity, int32_t* count);

/* The same pulse search with no allocation and no context.  This entry point has no
 * counterpart in the reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops).  The kernels keep their tile
 * counts and edge list in a file-scope device array of the library (one per device,
 * ~4 MiB, held until the process ends); sync_ref (a span plane holding offsets from
 * plane_bias_hz), starts, lengths and counts are device memory; asynchronous on `stream`.
 * counts[0] receives the number of pulse edges, counts[1] the number of pulses (only
 * min(counts[1], capacity) are stored).  If counts[0] exceeds
 * VHSDECODE_HIP_NOALLOC_MAX_EDGES the edge list was cut there and the pulse list is
 * incomplete.  Lengths above VHSDECODE_HIP_NOALLOC_MAX_SAMPLES are rejected with
 * VHSDECODE_HIP_E_CAPACITY.  Every call on a device shares that device's array: calls
 * that overlap on one device (different streams or host threads) race. */
#define VHSDECODE_HIP_NOALLOC_MAX_SAMPLES (16LL * 1024 * 1024)
#define VHSDECODE_HIP_NOALLOC_MAX_EDGES (448 * 1024)
VHSDECODE_HIP_API int32_t vhsdecode_hip_find_pulses_noalloc(const float* sync_ref, int64_t length,
                                                            double plane_bias_hz, double threshold,
                                                            int32_t min_length, int32_t max_length, int32_t* starts,
                                                            int32_t* lengths, int32_t capacity, int32_t* counts,
                                                            void* stream);

/* ---------------------------------------------------------------- hsync refinement
 * TbcFieldDecodePipeline.RefineLineLocationsFromHSync, vhs decode type
 * (Decode/TbcFieldDecodePipeline.cs:4570-4684 with :4807-5052), FP64 on the device. */
typedef struct vhsdecode_hip_hsync_params_v1 {
  uint32_t struct_size;
  double threshold;                /* detected sync threshold (Hz) */
  double sample_rate_mhz;          /* SyncAnalyzer.SampleRateMHz */
  double hsync_pulse_us;           /* SyncAnalyzer.HSyncPulseUs */
  double ire0;                     /* VideoOutputConverter.Ire0 */
  double hz_ire;                   /* VideoOutputConverter.HzIre */
  int32_t is_pal;                  /* PAL lines 1-2 are vsync lines too */
  int32_t use_right_hsync;         /* HSyncRefineOptions.UseRightHSync */
} vhsdecode_hip_hsync_params_v1;

/* sync_ref: device; locations/filled (in) and refined/filled_out (out): host. */
VHSDECODE_HIP_API int32_t vhsdecode_hip_refine_hsync(vhsdecode_hip_demod_context* context, const float* sync_ref,
                                                     int64_t length, const double* locations, const uint8_t* filled,
                                                     int32_t lines, const vhsdecode_hip_hsync_params_v1* params,
                                                     double* refined, uint8_t* filled_out);

/* ---------------------------------------------------------------- TBC line resampling
 * TbcLineResampler (Tbc/TbcLineResampler.cs:154-162, 732-878, 997-1011, 1167-1234,
 * 1495-1525) with the linear line-location interpolator and
 * VideoOutputConverter.FastMathConversion (Tbc/VideoOutputConverter.cs:74-90). */
typedef struct vhsdecode_hip_tbc_params_v1 {
  uint32_t struct_size;
  int32_t output_line_length;      /* TbcFrameSpec output line length (PAL 1135, NTSC 910) */
  int32_t first_line;
  int32_t line_count;
  double nominal_line_length;      /* <= 0: median of the line lengths, as the reference */
  double wow_level_smoothing;      /* WowLevelAdjustSmoothing (0 disables) */
  int32_t convert;                 /* write uint16 output through the converter */
  double ire0;
  double hz_ire;
  int32_t output_zero;
  double vsync_ire;
  double output_scale;
} vhsdecode_hip_tbc_params_v1;

/* Install the 65537 x 16 FP32 windowed-sinc table (the reference's sinc_lut.npz
 * payload).  Without it the library builds the Kaiser(beta=5) fallback of
 * TbcLineResampler.BuildKaiserSincLookup (Tbc/TbcLineResampler.cs:1297-1330). */
VHSDECODE_HIP_API int32_t vhsdecode_hip_set_sinc_table(vhsdecode_hip_demod_context* context, const float* table,
                                                       int64_t count);

/* source: device (span plane); locations: host; out_u16 / out_f64: host, either may be NULL. */
VHSDECODE_HIP_API int32_t vhsdecode_hip_tbc_resample(vhsdecode_hip_demod_context* context, const float* source,
                                                     int64_t source_length, const double* locations,
                                                     int32_t location_count, const vhsdecode_hip_tbc_params_v1* params,
                                                     uint16_t* out_u16, double* out_f64);

/* Device variant: every pointer is device memory, asynchronous on `stream`.
 * level_adjust holds one FP64 factor per output line (or per output sample when
 * per_sample != 0), indexed from line 0. */
VHSDECODE_HIP_API int32_t vhsdecode_hip_tbc_resample_device(vhsdecode_hip_demod_context* context, const float* source,
                                                            int64_t source_length, const double* locations,
                                                            const double* level_adjust, int32_t per_sample,
                                                            const vhsdecode_hip_tbc_params_v1* params,
                                                            uint16_t* out_u16, double* out_f64, void* stream);

/* ---------------------------------------------------------------- batched field decode
 * The per-field tail of TbcFieldDecodePipeline.DecodePrepared for a run of fields
 * (Decode/TbcFieldDecodePipeline.cs:1344-1348 hsync refinement, then the TBC render of
 * Tbc/TbcFieldRenderer.cs:289-340): one launch refines every field's line locations
 * (FP64), one computes the per-field wow level adjusts (median/MAD clamp and the
 * exponential smoothing of TbcLineResampler.cs:818-878, 997-1011) and one resamples
 * every output pixel to uint16.  Output line 0 of each field starts at its location 0. */
typedef struct vhsdecode_hip_field_params_v1 {
  uint32_t struct_size;
  double threshold;                /* sync threshold (Hz) */
  double sample_rate_mhz;
  double hsync_pulse_us;
  double ire0;
  double hz_ire;
  int32_t is_pal;
  int32_t use_right_hsync;
  int32_t output_line_length;      /* TbcFrameSpec output line length */
  int32_t field_lines;             /* output lines per field; each field supplies field_lines + 1 locations */
  double nominal_line_length;      /* round(line_period * MHz), DecodeSession.cs:408-410 */
  double wow_level_smoothing;      /* frame_lines / 2 for vhs (DecodeSession.cs:1029-1050); 0 disables */
  int32_t output_zero;
  double vsync_ire;
  double output_scale;
} vhsdecode_hip_field_params_v1;

/* sync_ref, video: device span planes; line_locations / filled: host
 * [nfields][field_lines + 1]; origins: host [nfields] or NULL — each field's TBC source
 * coordinates count from span position origins[f] (the start of the field's read window,
 * as TbcLineResampler sees it; DecodeReadWindowPlanner.Resolve, Decode/DecodeReadWindowPlanner.cs:37-63);
 * out: device [nfields][field_lines * output_line_length];
 * refined_out: optional host [nfields][field_lines + 1] (makes the call synchronous). */
VHSDECODE_HIP_API int32_t vhsdecode_hip_decode_fields_device(vhsdecode_hip_demod_context* context,
                                                             const float* sync_ref, const float* video,
                                                             int64_t span_length, const double* line_locations,
                                                             const uint8_t* filled, const int64_t* origins,
                                                             int32_t nfields,
                                                             const vhsdecode_hip_field_params_v1* params,
                                                             uint16_t* out, double* refined_out, void* stream);

/* ---------------------------------------------------------------- VHS field sequence
 * The field decode of the vhs command with its default (v0.4.0) sync path, for the fields
 * of a demodulated span read in sequence: TbcFieldDecodePipeline.DecodePrepared
 * (Decode/TbcFieldDecodePipeline.cs:1071-1420) from PrepareSyncSpan's level detection
 * (:2776-3234: the fallback serration-level search and refinement, the moving averages of
 * VsyncSerrationDetector and VhsFieldLevelState, the DC shift of the sync reference) to
 * FindRawPulses, SyncAnalyzer.EstimateTiming/RefinePulses/ComputeMeanLineLength
 * (Dsp/SyncAnalysis.cs:131-538), FieldParityDetector.ResolveCadence, ResolveLine0Location
 * (:5452-5651) with VBlankSyncResolver.EstimateLine0FromTransitions, BuildUpstreamLineLocations
 * (SyncAnalysis.cs:585-657), RefineLineLocationsFromHSync on the shifted reference at the
 * field's threshold, the NTSC fsc phase shift (:4022-4038) and the TBC render from the
 * output's first line (OutputFirstLine, :2475-2483).  Each field is read from the window
 * DecodeReadWindowPlanner.Resolve places at the previous field's next-field offset
 * (TbcFieldSequenceDecodeEngine.cs:161-171); the first from `first_begin`.
 * VsyncSerrationDetector's candidate search is not run (no serration measurement). */
typedef struct vhsdecode_hip_vhs_sync_v1 {
  uint32_t struct_size;
  int32_t is_pal;
  double line_period_us;           /* sysparams line_period */
  double hsync_pulse_us;           /* hsyncPulseUS */
  double eq_pulse_us;              /* eqPulseUS */
  double vsync_pulse_us;           /* vsyncPulseUS */
  int32_t num_pulses;              /* numPulses */
  double hsync_tolerance_us;       /* 0.7 for vhs (DecodeSession.cs:420-428) */
  double eq_tolerance_us;          /* 0.9 for vhs */
  double frames_per_second;        /* FPS */
  double frame_lines;              /* frame_lines */
  double output_rate_mhz;          /* outfreq (TbcFrameSpec.OutputSampleRateHz / 1e6) */
  int64_t read_samples;            /* DecodeReadWindowPlanner.EstimateReadSampleCount */
  int32_t block_length;            /* read-window alignment */
  int32_t block_cut;
  int32_t block_stride;
} vhsdecode_hip_vhs_sync_v1;

/* One field of the sequence (device output record; struct_size is written). */
typedef struct vhsdecode_hip_vhs_field_v1 {
  uint32_t struct_size;
  int32_t status;                  /* 0 decoded; 1 no sync pulses, 2 no first hsync, 3 insufficient
                                      data (the reference's recovery kinds: next_begin is the intended
                                      start plus the recovery offset); 4 window not the speculated one
                                      and 6 window not read in this call (next_begin = begin: call again
                                      from there, on this span if the window fits in it); 5 a field
                                      window holds more sync edges or pulses than the device lists */
  int64_t begin;                   /* intended read start (span coordinates) */
  int64_t window_start;            /* read window (span coordinates) */
  int64_t window_length;
  int32_t level_source;            /* 0 default threshold, 1 serration levels, 2 refined levels,
                                      3 saved levels, 4 reversed defaults */
  int32_t raw_pulses;
  int32_t refined_pulses;
  int32_t is_first_field;
  int32_t parity_confidence;
  int32_t reserved;
  double sync_level, blank_level;  /* Hz */
  double dc_offset;                /* added to the sync reference */
  double threshold;                /* sync threshold on the shifted reference */
  double hsync_median;
  double mean_line_length;
  double line0;                    /* window coordinates */
  double first_hsync;
  double first_hsync_line;
  double next_offset;              /* ComputeNextFieldOffsetSamples (window coordinates) */
  int64_t next_begin;              /* span coordinates */
} vhsdecode_hip_vhs_field_v1;

/* sync_ref, video: device span planes whose sample 0 lies on the block-stride grid;
 * first_begin: the first field's intended read start (span coordinates); reset_state: 1
 * starts a new sequence (no level history, parity or sync history), 0 continues the one
 * this context's last call left (its span coordinates shifted by `state_shift` samples,
 * the distance from that call's span start to this one's).  Up to `rounds` rounds run
 * on the device, each speculating the read windows of the fields ahead from the span's
 * vertical-sync pulses, decoding them in parallel and verifying them in sequence order;
 * a round ends at a mispredicted window or a failed field and the next one carries on
 * from there.  out: device [max_fields][field_lines * output_line_length], field k's
 * output at slot k; fields: optional device records [max_fields]; field_count: device
 * int32, the records written (0 .. count - 1, in sequence order: status 0 decoded, its
 * field in out; status 1-3 failed, with the reference's recovery applied, its out slot
 * untouched); record `count`, when count < max_fields, says why the call stopped (its
 * status) and where the sequence continues.  next_begin: optional device int64 receiving
 * where the sequence continues (span coordinates).  params->threshold is the nominal
 * sync threshold used only to speculate the read windows.  Asynchronous on `stream`. */
VHSDECODE_HIP_API int32_t vhsdecode_hip_decode_vhs_fields_device(
    vhsdecode_hip_demod_context* context, const float* sync_ref, const float* video, int64_t span_length,
    int64_t first_begin, int32_t reset_state, int64_t state_shift, const vhsdecode_hip_vhs_sync_v1* sync,
    const vhsdecode_hip_field_params_v1* params, int32_t max_fields, int32_t rounds, uint16_t* out,
    vhsdecode_hip_vhs_field_v1* fields, int32_t* field_count, int64_t* next_begin, void* stream);

/* ---------------------------------------------------------------- VHS colour-under chroma
 * The chroma of the vhs command for a colour-under format (VHS, S-VHS), the reference's
 * default with the v0.4.0 behaviour profile: TbcFieldDecodePipeline.DecodePrepared
 * (Decode/TbcFieldDecodePipeline.cs:1366-1392, 1425-1470, 1869-1920) with
 * VhsChromaDecoder (Decode/VhsChromaDecoder.cs): the burst phase of every line probed from
 * the chroma plane resampled at the hsync-refined locations (AnalyzeFieldPhaseWithWorkspace
 * :574-702, ProbeUpconvertedBurstValue :1660-1713, DemodBurstFloat32 :1521-1626, the
 * track-phase sequence :1906-1965, 3927-4293, carried from field to field), the line
 * locations moved by the burst phase (RefineLineLocationsFromBurst :2342-2428) for the
 * luma render as well, and the chroma field: render resample, NTSC burst de-emphasis,
 * heterodyne upconversion (NTSC phase compensated, :2009-2045, 2167-2249), the final
 * band-pass forward and backward in float32, comb and automatic gain to uint16
 * (CompleteFieldWithPhaseCore :953-1251, ApplyAutomaticChromaGainWithCombToU16
 * :3504-3656).  Not covered: the colour killer, track-phase detection, chroma AFC
 * (Betamax PAL), the Video8 chroma de-emphasis and the "current" behaviour profile. */
typedef struct vhsdecode_hip_chroma_params_v1 {
  uint32_t struct_size;
  int32_t is_ntsc;                 /* 1 NTSC (phase-compensated upconversion, 1H comb), 0 PAL (2H comb) */
  double fsc_mhz;                  /* sysparams fsc_mhz */
  double color_under_hz;           /* rfparams color_under_carrier */
  double output_rate_hz;           /* TbcFrameSpec.OutputSampleRateHz (outfreq x 1e6) */
  int32_t field_lines;             /* output lines per field (frame_lines / 2 + 1) */
  int32_t output_line_length;      /* outlinelen */
  int32_t burst_start, burst_end;  /* burst window in output samples (BuildChromaFieldOptions,
                                      TbcFieldDecodePipeline.cs:2626-2690) */
  double burst_abs_ref;            /* sysparams burst_abs_ref */
  int32_t rotation[2];             /* rfparams chroma_rotation */
  int32_t sections;                /* final band-pass sections, 1..8 */
  int32_t input_line_length;       /* max(1, round half even (nominal line length, samples)) */
  double final_sos[8][6];          /* BuildChromaFinalSosFilter (Dsp/DecodeFilterSet.cs:710-751):
                                      b0 b1 b2 a0 a1 a2 per section */
} vhsdecode_hip_chroma_params_v1;

/* Sets (params non-NULL) or clears the context's colour-under chroma format; builds the
 * heterodyne and burst-carrier tables of one field (BuildHeterodyneTable :1967-2007,
 * BuildCarrierTables :3793-3893) on the device.  No counterpart entry in the reference:
 * there the chroma options are part of the field decoder's construction. */
VHSDECODE_HIP_API int32_t vhsdecode_hip_set_vhs_chroma(vhsdecode_hip_demod_context* context,
                                                       const vhsdecode_hip_chroma_params_v1* params);

/* vhsdecode_hip_decode_vhs_fields_device with the chroma of every decoded field: `chroma`
 * is the span's chroma plane (the block chroma output, same span geometry as sync_ref);
 * out_chroma: device [max_fields][field_lines * output_line_length], field k's chroma at
 * slot k as its luma in `out`.  The luma render uses the burst-refined line locations, as
 * the reference's does whenever it writes chroma (vhsdecode_hip_decode_vhs_fields_device
 * matches its decode with chroma off).  field_number_base: the number of fields decoded
 * before this sequence (NtscFieldPhaseTarget's field number), used when reset_state is 1;
 * a continued sequence counts on.  Requires vhsdecode_hip_set_vhs_chroma. */
VHSDECODE_HIP_API int32_t vhsdecode_hip_decode_vhs_fields_chroma_device(
    vhsdecode_hip_demod_context* context, const float* sync_ref, const float* video, const float* chroma,
    int64_t span_length, int64_t first_begin, int32_t reset_state, int64_t state_shift,
    const vhsdecode_hip_vhs_sync_v1* sync, const vhsdecode_hip_field_params_v1* params, int32_t max_fields,
    int32_t rounds, uint16_t* out, uint16_t* out_chroma, int32_t field_number_base,
    vhsdecode_hip_vhs_field_v1* fields, int32_t* field_count, int64_t* next_begin, void* stream);

/* The line tables of the last vhsdecode_hip_decode_vhs_fields_device call on this context,
 * [records][stride] (rows of decoded records only): which 0, BuildUpstreamLineLocations
 * (TbcFieldDecodePipeline.cs:1244-1295); 1, after RefineLineLocationsFromHSync
 * (:1344-1348); 2, the TBC render locations (less the NTSC fsc phase shift, :1394-1400),
 * all three in window coordinates (the record's window_start is the origin); 3, the
 * render locations from the field's output first line (:2475-2483) in span coordinates,
 * the locations vhsdecode_hip_detect_dropouts_device takes.  Device memory owned by the
 * context, valid until the next vhs field decode on it. */
VHSDECODE_HIP_API int32_t vhsdecode_hip_vhs_field_lines_device(vhsdecode_hip_demod_context* context, int32_t which,
                                                               const double** table, int32_t* stride);

/* The refined line locations of the last vhsdecode_hip_decode_fields_device call on this context:
 * device memory owned by the context, [fields][location_count] in span coordinates,
 * valid until the next field decode call on it. */
VHSDECODE_HIP_API int32_t vhsdecode_hip_field_locations_device(vhsdecode_hip_demod_context* context,
                                                               const double** locations, int32_t* location_count);

/* ---------------------------------------------------------------- dropout detection
 * TbcFieldDecodePipeline.DetectTapeDropouts (Decode/TbcFieldDecodePipeline.cs:6593-6642)
 * for a run of fields, on the device: the float32 pairwise mean of the RF envelope
 * over each field's read window (NumpyReduction.MeanFloat32; the window of the field's
 * sequence record) gives the threshold, RfDropoutDetector.FindDropouts
 * (Decode/TbcDropoutMapping.cs:52-107) finds the hysteresis ranges between the field's
 * first and last detection line, and TbcDropoutMapper.MapTapeRfToTbc (:288-366) turns
 * them into .tbc.json rows (fieldLine, startx, endx).  Output line k of a field starts
 * at its location k, which is the reference's first rendered location, so a row's field
 * line is k + 1. */
typedef struct vhsdecode_hip_dropout_params_v1 {
  uint32_t struct_size;
  double threshold_fraction;       /* TbcDropoutDetectionOptions.ThresholdFraction: 0.18 (DDD), 0.35 (CxADC) */
  double absolute_threshold;       /* AbsoluteThreshold when >= 0; negative: fraction x envelope mean */
  double hysteresis;               /* 1.25 */
  int32_t merge_threshold;         /* RfDropoutDetector.DefaultMergeThreshold (30) */
  int32_t minimum_length;          /* RfDropoutDetector.DefaultMinimumLength (10) */
  int32_t output_line_length;      /* TbcFrameSpec output line length */
  int32_t first_field_lines;       /* CurrentFieldLineCount of a first field (NTSC 263, PAL 312) */
  int32_t second_field_lines;      /* ... of a second field (NTSC 262, PAL 313) */
  int32_t capacity;                /* rows (and ranges) stored per field */
} vhsdecode_hip_dropout_params_v1;

/* envelope: device span plane; locations: device [max_fields][location_count], each
 * field's locations from its output first line in span coordinates (table 3 of
 * vhsdecode_hip_vhs_field_lines_device); fields (read window, parity, status) and
 * field_count (int32, records to process, <= max_fields): device, as
 * vhsdecode_hip_decode_vhs_fields_device writes them.  Records whose status is not 0 get
 * no rows.  Outputs, device: row_count [max_fields]; field_line, start_x, end_x
 * [max_fields][capacity]; range_count, range_start, range_end (sample positions in the
 * span) likewise, or all three NULL.  A count above capacity means only the first
 * `capacity` entries were stored; -1 means the field's read window exceeded 3.6 M
 * samples or its detection range 8388608.  Asynchronous on `stream`. */
VHSDECODE_HIP_API int32_t vhsdecode_hip_detect_dropouts_device(
    vhsdecode_hip_demod_context* context, const float* envelope, int64_t span_length, const double* locations,
    int32_t location_count, const vhsdecode_hip_vhs_field_v1* fields, const int32_t* field_count,
    int32_t max_fields, const vhsdecode_hip_dropout_params_v1* params, int32_t* row_count, int32_t* field_line,
    int32_t* start_x, int32_t* end_x, int32_t* range_count, int32_t* range_start, int32_t* range_end, void* stream);

/* ---------------------------------------------------------------- served windows
 * A device-resident request loop for seekable decode windows from many concurrent
 * streams, the served counterpart of the reference's preview windows
 * (src/VHSDecode.Preview: preview_create / decode_window / destroy,
 * vhsdecode_cuda_fast.h:55-180) made asynchronous.  Requests queue in arrival order
 * and are admitted as device slots free up, within an HBM budget; each admitted
 * window runs the whole chain (RF upload, block demod, field decode, download of
 * the uint16 fields) on its slot's own HIP stream, so windows overlap on the
 * device.  Filter tables and the sinc table stay resident per slot and format.
 * A worker thread owned by the server issues all device work; submit and wait may be
 * called from any thread.  A window's result does not depend on what else is in
 * flight. */
typedef struct vhsdecode_hip_serve vhsdecode_hip_serve;

typedef struct vhsdecode_hip_serve_config_v1 {
  uint32_t struct_size;
  int32_t device_id;
  int32_t max_slots;               /* windows in flight at once */
  uint64_t hbm_budget_bytes;       /* admission budget for the slots' device memory */
  int64_t max_window_samples;      /* largest window a request may ask for */
  int32_t max_fields;              /* largest field count a request may ask for */
} vhsdecode_hip_serve_config_v1;

/* One decode format: the block filters and cuts, the field geometry and the vhs sync
 * parameters.  Everything is copied; the caller's arrays may be freed after the call. */
typedef struct vhsdecode_hip_serve_format_v1 {
  uint32_t struct_size;
  const vhsdecode_hip_filters_v1* filters;
  int32_t block_cut;
  int32_t block_cut_end;
  const vhsdecode_hip_field_params_v1* field;
  const vhsdecode_hip_vhs_sync_v1* sync;
  const float* sinc_table;         /* 65537 x 16 FP32, or NULL for the Kaiser fallback */
  int64_t sinc_count;
} vhsdecode_hip_serve_format_v1;

typedef struct vhsdecode_hip_window_v1 {
  uint32_t struct_size;
  int32_t stream_id;
  int64_t first_sample;            /* even */
  int64_t sample_count;            /* >= block_length, <= max_window_samples */
  int32_t max_fields;              /* <= the server's max_fields */
  uint16_t* fields_out;            /* host [max_fields][field_lines * output_line_length]; pinned
                                      memory (hipHostMalloc) is written by DMA, other memory
                                      through a blocking copy */
} vhsdecode_hip_window_v1;

typedef struct vhsdecode_hip_window_result_v1 {
  uint32_t struct_size;
  int32_t status;                  /* 0, or the error the window failed with */
  int32_t field_count;             /* fields written to fields_out */
  double queued_ms;                /* submit -> admitted to a slot */
  double latency_ms;               /* submit -> fields in host memory */
} vhsdecode_hip_window_result_v1;

VHSDECODE_HIP_API int32_t vhsdecode_hip_serve_create(const vhsdecode_hip_serve_config_v1* config,
                                                     vhsdecode_hip_serve** server);
VHSDECODE_HIP_API int32_t vhsdecode_hip_serve_add_format(vhsdecode_hip_serve* server,
                                                         const vhsdecode_hip_serve_format_v1* format,
                                                         int32_t* format_id);
/* A stream reads `samples` (host, int16, caller-owned for the server's lifetime);
 * with loop != 0 positions wrap modulo `length` (an endless recording). */
VHSDECODE_HIP_API int32_t vhsdecode_hip_serve_add_stream(vhsdecode_hip_serve* server, int32_t format_id,
                                                         const int16_t* samples, int64_t length, int32_t loop,
                                                         int32_t* stream_id);
VHSDECODE_HIP_API int32_t vhsdecode_hip_serve_submit(vhsdecode_hip_serve* server,
                                                     const vhsdecode_hip_window_v1* window, uint64_t* ticket);
/* Waits up to timeout_ms (< 0: forever) for a ticket; VHSDECODE_HIP_E_TIMEOUT if it is
 * still in flight.  A ticket's result is returned once. */
VHSDECODE_HIP_API int32_t vhsdecode_hip_serve_wait(vhsdecode_hip_serve* server, uint64_t ticket, int32_t timeout_ms,
                                                   vhsdecode_hip_window_result_v1* result);
/* Finishes the windows in flight, drops queued ones, frees every device resource. */
VHSDECODE_HIP_API void vhsdecode_hip_serve_destroy(vhsdecode_hip_serve* server);

/* ---------------------------------------------------------------- whole run
 * Replaces vhsdecode_cuda_fast_run (src/VHSDecode.CudaFast.Native/include/vhsdecode_cuda_fast.h:55-67,
 * 83-106, 153-156; src/VHSDecode.CudaFast.Native/src/vhsdecode_cuda_fast.cpp:238-370): RF read
 * through the caller's callback in spans, each span decoded on the device (block demod,
 * pulses, line locations, hsync, TBC and, with `dropouts`, RF-envelope dropout detection),
 * and the fields written in order to <output_base>.tbc with <output_base>.tbc.json
 * (videoParameters, then per field isFirstField, seqNo, fileLoc and dropOuts, the layout of
 * the reference's TBCWriter::write_json, src/VHSDecode.CudaFast.Native/src/tbc_writer.cpp:209-312).
 * Colour-under chroma is not decoded: no _chroma.tbc is written.
 * The fields are the vhs field sequence (vhsdecode_hip_decode_vhs_fields_device) from the
 * start of the input, carried from span to span.  Spans sit at fixed places on multiples
 * of lcm(block length, block stride) (491520 samples for 1024/1024 cuts) and overlap by
 * the longest read window and a grid step, so the blocks two spans share give the same
 * planes and every field's read window, hence every output sample, is what one span from
 * the start of the input gives.  A prefetch thread reads the next span through the
 * callback while the device decodes the current one: the callbacks must be callable from
 * a thread other than the caller's.  A cancel (the cancel callback, polled between spans
 * and after a short read) ends the run with VHSDECODE_HIP_E_CANCELLED and no .tbc.json. */
enum { VHSDECODE_HIP_INPUT_FLOAT32 = 0, VHSDECODE_HIP_INPUT_INT16 = 1 };

/* Reads sample_count samples from sample_offset into destination (int16 or float, as the
 * run's input_sample_format says); a short count means the end of the input (or a cancel). */
typedef size_t (*vhsdecode_hip_read_callback)(void* user_data, void* destination, uint64_t sample_offset,
                                              size_t sample_count);
/* Nonzero asks the run to stop (checked before every span). */
typedef int32_t (*vhsdecode_hip_cancel_callback)(void* user_data);

typedef struct vhsdecode_hip_run_config_v1 {
  uint32_t struct_size;
  int32_t device_id;
  const vhsdecode_hip_serve_format_v1* format;  /* filters, block cuts, field and sync params, sinc table */
  const vhsdecode_hip_dropout_params_v1* dropouts; /* NULL: no dropout detection (the reference's --noDOD) */
  uint64_t total_samples;
  uint32_t input_sample_format;    /* VHSDECODE_HIP_INPUT_*; FP32 samples are rounded to int16 */
  uint32_t maximum_output_fields;  /* 0: every field */
  int64_t span_samples;            /* RF samples per decode span (at least the overlap and one grid step);
                                      0: twenty-four fields' worth */
  const char* output_base_utf8;
  int32_t overwrite;               /* 0: fail when an output file exists */
  vhsdecode_hip_read_callback read_callback;
  vhsdecode_hip_cancel_callback cancel_callback; /* may be NULL */
  void* user_data;
  /* .tbc.json videoParameters (TBCWriter::write_json, tbc_writer.cpp:219-255):
   * system: NULL writes "PAL" or "NTSC" by the field params, else this name ("PAL-M");
   * colourBurstStart/End = (int)(us * 1e-6 * output rate + 0.5) (sysparams colorBurstUS). */
  const char* system_name;
  double colour_burst_start_us;
  double colour_burst_end_us;
  /* Colour-under chroma (vhsdecode_hip_set_vhs_chroma's parameters): the fields' chroma to
   * <base>_chroma.tbc beside the luma, which then renders at the burst-refined line
   * locations (the reference's vhs default); NULL writes no chroma file (its --skip-chroma). */
  const vhsdecode_hip_chroma_params_v1* chroma;
} vhsdecode_hip_run_config_v1;

typedef struct vhsdecode_hip_run_result_v1 {
  uint32_t struct_size;
  uint32_t fields_written;
  uint32_t output_line_length;
  uint32_t output_field_lines;
  double elapsed_seconds;
  uint32_t spans;                  /* decode spans the run took */
  uint32_t dropout_rows;           /* .tbc.json dropout rows written over all fields */
} vhsdecode_hip_run_result_v1;

VHSDECODE_HIP_API int32_t vhsdecode_hip_run(const vhsdecode_hip_run_config_v1* config,
                                            vhsdecode_hip_run_result_v1* result);

/* ---------------------------------------------------------------- observability
 * Per-kernel device time, the counterpart of vhsdecode_cuda_fast_result_v1.elapsed_seconds
 * (vhsdecode_cuda_fast.h:100-106) at kernel granularity.  When enabled, the context
 * brackets each launch with HIP events on the launch stream; reading the profile
 * synchronises the recorded events and clears them. */
enum {
  VHSDECODE_HIP_KERNEL_BLOCK_DEMOD = 0,
  VHSDECODE_HIP_KERNEL_SPIKE_REPAIR = 1,
  VHSDECODE_HIP_KERNEL_CHROMA = 2,
  VHSDECODE_HIP_KERNEL_PULSES = 3,
  VHSDECODE_HIP_KERNEL_HSYNC = 4,
  VHSDECODE_HIP_KERNEL_TBC = 5,
  VHSDECODE_HIP_KERNEL_SUB_DEEMPHASIS = 6,
  VHSDECODE_HIP_KERNEL_BLOCK_VIDEO = 7,
  VHSDECODE_HIP_KERNEL_LOCATE = 8,
  VHSDECODE_HIP_KERNEL_DROPOUTS = 9,
  VHSDECODE_HIP_KERNEL_CHROMA_FIELD = 10,  /* colour-under burst analysis and chroma field decode */
  VHSDECODE_HIP_KERNEL_COUNT = 11
};

typedef struct vhsdecode_hip_profile_v1 {
  uint32_t struct_size;
  int64_t launches[VHSDECODE_HIP_KERNEL_COUNT];
  double total_ms[VHSDECODE_HIP_KERNEL_COUNT];
  int64_t units[VHSDECODE_HIP_KERNEL_COUNT];  /* blocks / samples / fields / pixels processed */
} vhsdecode_hip_profile_v1;

VHSDECODE_HIP_API int32_t vhsdecode_hip_set_profiling(vhsdecode_hip_demod_context* context, int32_t enable);
VHSDECODE_HIP_API int32_t vhsdecode_hip_get_profile(vhsdecode_hip_demod_context* context,
                                                    vhsdecode_hip_profile_v1* profile);

/* Diagnostic: forward 16384-point complex FFTs of `count` transforms (device,
 * interleaved FP32) with the library's workgroup FFT.  No reference counterpart;
 * exposed so the transform can be checked in isolation. */
VHSDECODE_HIP_API int32_t vhsdecode_hip_debug_fft16k(vhsdecode_hip_demod_context* context, float* data,
                                                     int32_t count, int32_t inverse, void* stream);

/* ---------------------------------------------------------------- LaserDisc audio stage 2
 * LaserDiscAnalogAudioPhase2.Apply (src/VHSDecode.Core/Decode/LaserDiscAnalogAudioPhase2.cs:12-220),
 * which RfBlockDecodePipeline.ApplyLaserDiscAnalogAudioPhase2 (:159-163) runs on the
 * assembled analog audio of each stream read (RfBlockStreamDecoder.cs:913-914): per block of
 * the stage-2 length overlapping by 512, left-channel peak suppression in both channels, the
 * stage-2 response (BuildLdAnalogAudioStage2Filter, Dsp/DecodeFilterSet.cs:950-970), float32
 * output.  stage2[ch]: block_length complex values (re, im interleaved, FP64); block_length
 * must be 32768 (the 40 MSPS filter set's). */
typedef struct vhsdecode_hip_ld_audio_phase2_v1 {
  uint32_t struct_size;
  int32_t block_length;
  double center_hz[2];             /* LaserDiscAnalogAudioChannelFilter.CenterFrequencyHz, left / right */
  const double* stage2[2];         /* Stage2Filter, left / right */
} vhsdecode_hip_ld_audio_phase2_v1;

/* Sets (params non-NULL) or clears the context's stage-2 filters and carriers. */
VHSDECODE_HIP_API int32_t vhsdecode_hip_set_ld_audio_phase2(vhsdecode_hip_demod_context* context,
                                                            const vhsdecode_hip_ld_audio_phase2_v1* params);

/* Stage 2 of `count` samples per channel (device FP32 Hz at the slice rate, e.g. the span
 * planes audio_left / audio_right) into out_left / out_right (device FP32, `count` each;
 * must not overlay the inputs).  stream NULL: the context's stream, returning when done. */
VHSDECODE_HIP_API int32_t vhsdecode_hip_ld_audio_phase2_device(vhsdecode_hip_demod_context* context,
                                                               const float* left, const float* right, int64_t count,
                                                               float* out_left, float* out_right, void* stream);

/* Diagnostic: the block repair kernel's ReplaceSpikes step (RfDemodulator.ReplaceSpikes,
 * src/VHSDecode.Core/Dsp/RfDemodulator.cs:966-1001) on caller-supplied 32768-sample
 * demod and diffed-demod arrays (device FP32); writes the repaired demod to `out`.
 * Exposed so the sequential window logic can be checked against the reference's
 * semantics on chosen inputs. */
VHSDECODE_HIP_API int32_t vhsdecode_hip_debug_replace_spikes(vhsdecode_hip_demod_context* context, const float* demod,
                                                             const float* diffed, double max_value, float* out,
                                                             void* stream);

#ifdef __cplusplus
}
#endif

#endif /* VHSDECODE_HIP_H */
