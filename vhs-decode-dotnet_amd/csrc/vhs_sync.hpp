// Device structures of the VHS field sync path (vhs_sync.hip) shared with the host.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vhsdecode_hip.h"

namespace vhsd {
namespace vsync {

constexpr int kPulseLds = 4096;   // pulse lists of a field staged in LDS up to this length
constexpr int kHistBins = 512;    // hsync-length histogram (3.75 us of lengths)
constexpr int kLogicAux = 17;     // bytes per pulse of a field's classification lists in global memory
constexpr int kAvgCap = 16;       // moving-average values kept
constexpr int kMaxGroups = 256;   // vertical-sync groups per span
constexpr int kStatusMispredicted = 4;
constexpr int kStatusCapacity = 5;
constexpr int kStatusNotRead = 6;

struct SyncCfg {
  double fs_mhz;
  double hsync_us, eq_us, vsync_us;
  double hsync_tol, eq_tol;          // microseconds
  int num_pulses;
  int is_pal;
  double ire0, hz_ire, vsync_ire;
  double nominal;                    // SyncAnalyzer.NominalLineLength
  int field_lines0, field_lines1;    // FieldLinesForParityDetection
  int output_lines;                  // TbcFrameSpec.OutputLineCount
  long long block_length, block_cut, stride, read_samples;
  int ser_eq_len, ser_line_len;      // VsyncSerrationDetector original-rate lengths
  int pulse_min, pulse_max;          // FindRawPulses bounds for vhs
  int fs_window;                     // VhsFieldLevelState window
  double bias;                       // plane bias (Hz)
  int pulse_cap;                     // pulses kept per field window (the most a window can hold)
  int line_stride;                   // line locations per field in the output table
  int max_records;                   // field records the caller holds
};

struct LevelStateDev {
  double ser_sync[kAvgCap], ser_blank[kAvgCap];
  int ser_n;
  double fs_sync[kAvgCap], fs_blank[kAvgCap];
  int fs_sync_n, fs_blank_n;
  int has_last;
  double last_sync, last_blank;
};

struct SyncHistoryDev {
  int prev_first;                    // -1: none
  int has_prev_first_hsync;
  double prev_first_hsync;
  long long prev_first_hsync_read;   // window start of that field (span coordinates)
  double prev_hsync_difference;
  int pulse_search_initialized;
};

struct LevelSearch {
  int first_ok;
  double first_sync, first_blank;
  int first_fail, first_has_measured;
  double first_measured;
  int ref_ok;
  double ref_sync, ref_blank;
  int ref_fail, ref_has_measured;
  double ref_measured;
};

struct FieldLevels {
  int source;
  double sync, blank, dc, threshold;
};

struct Estimate {
  int ok;
  double line0, first_hsync, unaligned;
  int valid_count;
};

struct Transitions {
  Estimate first, last, combined;
};

struct FieldLogic {
  int raw, refined;
  double hsync_median, mll;
  int consensus, detected, first_conf, second_conf;
  Transitions tr[2];  // first field, second field
};

typedef vhsdecode_hip_vhs_field_v1 FieldRecord;

constexpr int kSegLen = 4096;   // samples per segment of the segmented pulse searches
constexpr int kSpecR = 8;       // level-search thresholds evaluated on segment workgroups
constexpr int kSpecHalf = 4;    // of which the first pass counts these (most searches decide in them)

// A segment's pulse-search summary (vhs_sync.hip, seg_scan).
struct SegSummary {
  int first_pos;  // first transition, -1 for none
  int first_end;  // 1 when it ends a pulse
  int last_pos;
  int n_in;       // pulses in range starting and ending in the segment
  int n_vs;       // of those, longer than vs_min
};

// Scratch of the segmented pulse searches ([field] unless stated).
struct SegBuffers {
  int segs;             // segments per window (the longest)
  double* seg_min;      // [field][segs]
  int* seg_nan;         // [field][segs]
  double* thr;          // [field][kSpecR]: the level search's threshold sequence
  int* choice;          // its list threshold (index), -1: the whole search on one workgroup
  int* fallback;
  int* active;
  int* pending;         // the search needs the second half of the threshold sequence
  int* count;           // pulses of the search's list
  double* thr2;         // the second refinement's threshold
  int* active2;
  int* count2;
  double* thr3;         // the field threshold and DC offset
  double* dc3;
  int* count3;
  SegSummary* sum;      // [field][kSpecR][segs]
  int* seg_carry;       // [field][segs]
  int* seg_off;         // [field][segs]
};

// The sequence position across the rounds of one call: each round speculates windows from
// `begin`, decodes them into records [base, base + decoded) and moves `begin` and `records`
// on; `halted` ends the call (the next window runs past the span, a list overflowed, or
// the records are full).
struct SeqCtl {
  long long begin;
  int records;
  int base;
  int decoded;
  int halted;
};

struct SyncWork {
  const float* sync_ref;
  long long span_len;
  SeqCtl* ctl;
  long long state_shift;   // span coordinates of the incoming history less this span's
  int* spec_count;
  long long* win_start;
  long long* win_len;
  int* pulse_start;
  int* pulse_len;
  int* overflow;
  LevelSearch* search;
  FieldLevels* levels;
  LevelStateDev* level_before;
  const LevelStateDev* level_state_in;
  LevelStateDev* level_state_out;
  LevelStateDev* level_state_after_all;
  int* raw_count;
  FieldLogic* logic;
  int* ref_kind;
  int* ref_start;
  unsigned char* logic_aux;  // [max_fields][kLogicAux * pulse_cap]: a field's classification lists past kPulseLds pulses
  FieldRecord* records;      // [max_records], the whole call
  const SyncHistoryDev* history_in;
  SyncHistoryDev* history_out;
  double* lines;             // [max_fields][line_stride], this round
  SegBuffers seg;
};

}  // namespace vsync
}  // namespace vhsd

namespace vhsd {
namespace vsync {
// Tables for the hsync refinement and the TBC render of the decoded fields.
struct FieldTail {
  const float** spans;
  long long* span_len;
  const double** locs;
  const unsigned char** filled;
  double** refined;
  unsigned char** filled_out;
  int* line_counts;
  long long* origins;
  int* first_lines;
  double* threshold;
  double* dc;
  const unsigned char* zeros;     // line_stride zero bytes (no line was filled by interpolation)
  double* refined_base;           // [max_fields][line_stride]
  unsigned char* filled_out_base;
  double* render;                 // [max_fields][line_stride]
  int* out_index;                 // [max_fields]: the field's record (and output) index
  // the call's tables, by record: line locations, refined, render (window coordinates)
  // and the render locations from the output first line in span coordinates
  double* acc_lines;
  double* acc_refined;
  double* acc_render;
  double* acc_span;
  double fsc_shift;               // NTSC fsc phase shift (samples), 0 for PAL
  int out_first_line[2];          // OutputFirstLine: first field, second field
};
}  // namespace vsync
}  // namespace vhsd

extern "C" hipError_t vhsd_launch_vhs_results(const vhsd::vsync::SeqCtl* ctl, int* record_count, int64_t* next_begin,
                                              hipStream_t stream);
extern "C" hipError_t vhsd_launch_vhs_ctl_init(vhsd::vsync::SeqCtl* ctl, long long first_begin, int first_record,
                                               hipStream_t stream);
extern "C" hipError_t vhsd_launch_vhs_export(const vhsd::vsync::SyncCfg* c, const vhsd::vsync::SyncWork* w,
                                             const vhsd::vsync::FieldTail* t, int max_fields, hipStream_t stream);
extern "C" hipError_t vhsd_launch_vhs_sync_reset(vhsd::vsync::LevelStateDev* level, vhsd::vsync::SyncHistoryDev* hist,
                                                 hipStream_t stream);
extern "C" hipError_t vhsd_launch_vhs_tail_tables(const vhsd::vsync::SyncCfg* c, const vhsd::vsync::SyncWork* w,
                                                  const vhsd::vsync::FieldTail* t, int max_fields,
                                                  hipStream_t stream);
extern "C" hipError_t vhsd_launch_vhs_render(const vhsd::vsync::SyncCfg* c, const vhsd::vsync::SyncWork* w,
                                             const vhsd::vsync::FieldTail* t, int max_fields, hipStream_t stream);
extern "C" hipError_t vhsd_launch_vhs_sync(const vhsd::vsync::SyncCfg* c, const vhsd::vsync::SyncWork* w,
                                           const int* starts, const int* lens, const int* pulse_count, int capacity,
                                           long long span_len, int max_fields, hipStream_t stream);
