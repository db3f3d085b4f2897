// Kernel argument blocks shared by the host orchestration and the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sos_scan.hpp"
#include "vhsdecode_hip.h"

namespace vhsd {

// Per-block scratch of the demod kernels: two spectra of 16384 complex FP32 values,
// then the demod spectrum's Nyquist bin at index 2 * 16384.
constexpr long long kStashEntries = 2 * 16384 + 64;
// Per-block auxiliary scratch of the RF-boost and notch paths: one half spectrum or one
// block of sample pairs.
constexpr long long kAuxEntries = 16384 + 64;

struct DemodArgs {
  const int16_t* rf;             // RF samples (device)
  const long long* block_start;  // input sample offset of each block
  const long long* out_offset;   // stitched output offset of each block's first kept sample
  int out_begin;                 // first kept sample inside the block (block cut)
  int out_count;                 // kept samples per block (stride)
  const float2* rf_video;        // [M+1] half spectra
  const float2* video;
  const float2* vlp05;
  const float2* oct;             // twiddle octant [kOctantEntries]
  int vlp_offset;                // sync reference roll (samples)
  float fs;                      // sample rate as FP32 (the discriminator's frequency)
  double diff_max;               // spike-repair threshold; <= 0 disables
  float bias;                    // plane bias (Hz), exactly representable in FP32
  float video_corr;              // bias * (Re Video[0] - 1): restores the DC the biased FFT removed
  float vlp_corr;                // bias * (Re VideoLowPass05[0] - 1)
  SosScanArgs env;               // envelope filter
  float* out_video;
  float* out_vlp;
  float* out_env;
  unsigned* flags;               // per block: bit0 weak RF, bit1 spike repair needed
  float2* stash;                 // [nblocks][kStashEntries] scratch, private to each block's workgroup
  // RF high boost (Betamax): the RF top band, scaled by 0.9 mean(envelope) / envelope and
  // boost_mult, is added to the filtered RF before the discriminator
  SosScanArgs boost;             // RF top band-pass (one section)
  float boost_mult;
  float2* aux;                   // [nblocks][kAuxEntries]
  // The video kernel reads the stash as clamp(x, clip_lo, clip_hi) - stash_sub: the tape
  // demod kernels stash absolute Hz (stash_sub = bias, no clip); the LaserDisc kernel
  // stashes offsets from the bias (stash_sub = 0) and clips for video (clip_hi > clip_lo).
  float stash_sub;
  float clip_lo, clip_hi;
  // reference planes after video and sync reference (LaserDisc burst and pilot)
  int nref;
  const float2* ref_filt[2];     // [M+1] half spectra
  float* ref_out[2];             // absolute Hz planes
  int ref_shift[2];              // roll (samples)
  float ref_corr[2];             // bias * Re H[0]
  // LaserDisc conjugate-product discriminator and analog audio
  double fs_d;                   // sample rate in FP64
  int audio;                     // 1: two audio slices per block
  int audio_low[2];              // first positive bin of each slice
  int audio_bins;                // slice length (power of two, <= 2048)
  int audio_log2;
  const float2* audio_stage1[2]; // [audio_bins] stage-1 responses
  double audio_scale;            // slice rate / 2 pi
  double audio_low_freq[2];      // Hz added to each demodulated slice
  int audio_cut;                 // kept window of each block's slice: [cut, cut + stride)
  int audio_stride;
  int audio_decimation;          // block_length / audio_bins
  float* out_audio[2];
};

struct ChromaArgs {
  const int16_t* rf;
  const long long* block_start;
  const long long* out_offset;
  int out_begin;
  int out_count;
  int chroma_offset;  // roll (samples)
  SosScanArgs sos;
  float* out_chroma;
};

// Sub-deemphasis of the kept video window of each block, from the demod spectrum the
// block demod (or repair) kernel leaves in the block's stash: D at [0, M), D[M] at [2M].
struct SubDeemphArgs {
  const long long* out_offset;
  int out_begin;
  int out_count;
  const float2* vh;        // [M+1] Video x HighPass half spectrum
  const float2* oct;
  float2* stash;
  SosScanArgs amp;         // amplitude low-pass (one section)
  float half_deviation;    // deviation / 2
  float exp_scaling;
  float scaling1;
  float scaling2;
  float logistic_mid;
  float logistic_rate;     // <= 0: off
  float static_factor;
  float* out_video;
  // fsc notch (Betamax): applied to the whole sub-deemphasised block before the kept
  // window is written; the block's video is rebuilt from the stashed demod spectrum
  int notch;
  SosScanArgs notch_sos;         // the notch transfer function as one section
  const float2* video;           // [M+1] Video half spectrum
  float video_corr;
  float2* aux;                   // [nblocks][kAuxEntries]
};

// Batched block kernels: one launch decodes blocks of several formats; a workgroup
// takes a global block and that block's record recs[block_rec[block]] (the per-block
// arrays of the records are indexed by the global block).
// A batch's first stage in one launch: workgroups [0, n_demod) run the demod stage of
// block demod_ids[i] (plain, RF boost or LaserDisc by its record's kind), the rest the
// chroma filter of block chroma_ids[i - n_demod] (by its record's section count).
struct StageTable {
  const int* demod_ids;
  int n_demod;
  const int* chroma_ids;
  const int* block_rec;     // block -> record
  const int* rec_kind;      // record -> kTableDemod / kTableDemodBoost / kTableDemodLd
  const int* rec_sections;  // record -> chroma sections
};

// A batch's second stage in one launch, per block: the spike repair (when its record
// has it and the block is flagged), the video stage, then the sub-deemphasis (when
// its record has it).
struct StageTable2 {
  const int* block_rec;       // block -> record
  const int* rec_repair;      // record -> 1 when the repair may run
  const int* rec_subdeemph;   // record -> 1 when the sub-deemphasis runs
};

// Demod-stage kinds of a batch record (StageTable::rec_kind).
enum BlockDemodKind {
  kTableDemod = 0,
  kTableDemodBoost = 1,
  kTableDemodLd = 2,
};

struct RepairArgs {
  DemodArgs d;
  float2* scratch_analytic;  // [nblocks][32768] complex
  float* scratch_demod;      // [nblocks][65536]: demod, diffed demod
};

struct HsyncParams {
  double threshold;
  double fs_mhz;       // sample rate in MHz (SyncAnalyzer.SampleRateMHz)
  double hsync_us;     // HSyncPulseUs
  double ire0;
  double hz_ire;
  int is_pal;
  int use_right;
};

struct TbcParams {
  int out_len;         // output line length
  int first_line;
  int line_count;
  double nominal;      // nominal input line length
  int has_conv;
  double ire0;
  double out_scale;
  double inv_hz_ire;
  double conv_offset;  // (outputZero + 0.5) - vsync_ire * outputScale
};

// RF-envelope dropout detection of a run of fields (dropout_kernels.hip).
constexpr int kDropMaxLocations = 512;
constexpr int kDropSummaryWords = 4096;  // summary bits: 131072 mask words per field

struct DropoutField {               // per-field state the mask pass hands to the walk
  double threshold, up;
  long long a0, b0;                 // detection range [a0, b0) in span samples
  int nwords, end_line, status;     // status: 1 walk, 0 empty range, -1 range over capacity
};

struct DropoutArgs {
  const float* envelope;           // span plane
  long long span_len;
  const double* locations;         // [fields][nloc], span coordinates, from the output first line
  int nloc;
  const vhsdecode_hip_vhs_field_v1* fields;  // [fields]: read window and parity
  const int* field_count;          // device count of fields to process
  double fraction;
  double absolute_threshold;       // >= 0: used instead of fraction x mean
  double hysteresis;
  int merge;
  int min_len;
  int out_len;
  int first_lines, second_lines;
  int capacity;                    // rows and ranges stored per field
  // scratch
  float* partials;                 // [fields][128]
  float* group_min;                // [fields][128] smallest envelope value under each group
  DropoutField* info;              // [fields]
  unsigned long long* masks;       // [fields][2][words_cap]
  unsigned* summaries;             // [fields][2][kDropSummaryWords]
  int words_cap;
  // outputs
  int* row_count;
  int* field_line;
  int* start_x;
  int* end_x;
  int* range_count;                // optional (NULL)
  int* range_start;
  int* range_end;
};

}  // namespace vhsd

namespace vhsd {
// LaserDisc analog audio stage 2 over one read's assembled audio (both channels).
struct LdAudio2Args {
  const float* in[2];
  float* out[2];
  long long count;         // samples per channel
  int nmid;                // middle blocks (between the first and the last)
  double center[2];        // carrier of each channel (Hz)
  const float2* filt[2];   // stage-2 half spectra, 16385 entries each
  const float2* oct;
};
}  // namespace vhsd

extern "C" {
size_t vhsd_block_lds_bytes();
hipError_t vhsd_launch_block_demod(const vhsd::DemodArgs* a, int nblocks, hipStream_t stream);
hipError_t vhsd_launch_block_chroma(const vhsd::ChromaArgs* a, int nblocks, hipStream_t stream);
hipError_t vhsd_launch_block_repair(const vhsd::RepairArgs* a, int nblocks, hipStream_t stream);
hipError_t vhsd_launch_dropouts(const vhsd::DropoutArgs* a, int max_fields, hipStream_t stream);
hipError_t vhsd_launch_block_demod_boost(const vhsd::DemodArgs* a, int nblocks, hipStream_t stream);
hipError_t vhsd_launch_block_demod_ld(const vhsd::DemodArgs* a, int nblocks, hipStream_t stream);
hipError_t vhsd_launch_stage1_table(const vhsd::DemodArgs* drecs, const vhsd::ChromaArgs* crecs,
                                    vhsd::StageTable t, int n_chroma, hipStream_t stream);
hipError_t vhsd_launch_stage2_table(const vhsd::RepairArgs* rrecs, const vhsd::DemodArgs* drecs,
                                    const vhsd::SubDeemphArgs* srecs, vhsd::StageTable2 t, int blocks,
                                    hipStream_t stream);
hipError_t vhsd_launch_fill_u32(unsigned* p, long long n, unsigned value, hipStream_t stream);
hipError_t vhsd_launch_ld_audio_phase2(const vhsd::LdAudio2Args* a, int nblocks, hipStream_t stream);
hipError_t vhsd_launch_block_video(const vhsd::DemodArgs* a, int nblocks, hipStream_t stream);
hipError_t vhsd_launch_block_subdeemph(const vhsd::SubDeemphArgs* a, int nblocks, hipStream_t stream);
hipError_t vhsd_launch_debug_fft(float* data, int count, int inverse, const float2* oct, hipStream_t stream);
hipError_t vhsd_launch_debug_replace_spikes(const float* dm, const float* dd, int* list, double mx, float* out,
                                           hipStream_t stream);
hipError_t vhsd_launch_pulses(const float* s, long long n, double bias, double high, int min_len, int max_len,
                              int* tile_counts,
                              int n_tiles_cap, long long* edge_pos, unsigned char* edge_kind, int edge_cap,
                              int* counters, int* starts, int* lengths, int capacity, hipStream_t stream);
hipError_t vhsd_launch_hsync(const float* const* spans, const long long* span_len, const double* const* locs,
                             const unsigned char* const* filled, const int* line_counts, const vhsd::HsyncParams* p,
                             double* const* refined_out, unsigned char* const* filled_out, int fields, int max_lines,
                             double bias, void* work, const double* field_threshold, const double* field_dc,
                             hipStream_t stream);
size_t vhsd_hsync_work_bytes(int fields, int max_lines);
hipError_t vhsd_launch_tbc_fields(const float* src, long long src_len, double bias, const long long* origins,
                                  const double* refined, int nloc, int lines, const int* first_lines,
                                  int max_first_line, int relative,
                                  int fields, double nominal, double line_decay, const double* pow_table,
                                  double* line_x, double* line_y0, int out_len, int smooth, double ire0,
                                  double out_scale, double inv_hz_ire, double conv_offset, const float* lut,
                                  unsigned short* out, const int* out_index, const int* field_count,
                                  hipStream_t stream);
hipError_t vhsd_launch_tbc(const float* src, long long src_len, double bias, const double* locs, const double* level_adjust,
                           int level_per_line, const vhsd::TbcParams* p, const float* lut, unsigned short* out_u16,
                           double* out_f64, hipStream_t stream);
}
