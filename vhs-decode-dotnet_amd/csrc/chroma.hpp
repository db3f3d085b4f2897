// Colour-under chroma of the vhs field sequence on the device: shared definitions.
//
// The reference's default `vhs` decode of a colour-under format (v0.4.0 profile,
// Decode/TbcFieldDecodePipeline.cs:1366-1392, 1869-1920; Decode/VhsChromaDecoder.cs):
//   1. the chroma plane resampled at the hsync-refined line locations from line 1, only
//      where the burst probes read it (ResampleChromaBurst, :2504-2534);
//   2. per line, the burst heterodyned by the line's track phase, band-passed forward and
//      backward in float32 and demodulated against the subcarrier (ProbeUpconvertedBurstValue,
//      DemodBurstFloat32), for both track-phase starts; the flip test and the phase
//      averages per start; the start carried from field to field (GetPhaseRotationSequenceCore);
//   3. the line locations moved by the burst phase (RefineLineLocationsFromBurst), which
//      the luma render uses too;
//   4. the chroma plane resampled at the render locations, NTSC burst de-emphasis, the
//      heterodyne upconversion (NTSC phase compensated), the final band-pass forward and
//      backward in float32 over the whole field, comb and automatic gain to uint16
//      (CompleteFieldWithPhaseCore).
#pragma once
#include <hip/hip_runtime.h>

#include "vhs_sync.hpp"

namespace vhsd {
namespace chroma {

constexpr int kMaxSections = 8;
constexpr int kProbeThreads = 64;
constexpr int kIirChunk = 64;  // whole-field band-pass: samples per chunk

// Per context and format, built on the host (vhsdecode_hip_set_vhs_chroma).
struct ChromaCfg {
  int is_ntsc;
  int lines;               // output lines per field
  int line_len;            // output samples per line
  int burst_start, burst_end;
  int win_max;             // longest probe window: burst_end + (burst_end - burst_start)
  int probe_ext;           // win_max + 2 * pad: a probe's filter buffer
  int pad;                 // the final band-pass's odd-extension length (DefaultPadLength)
  int iir_chunk;           // whole-field band-pass: samples per chunk
  int iir_warm;            // and the warm-up run before each chunk
  int sections;
  float sos[kMaxSections][6];  // b0 b1 b2 a0 a1 a2, FP32
  float zi[kMaxSections][2];   // steady-state initial conditions (FP32, scaled as the reference)
  int rot[2];              // chroma_rotation
  int input_line_length;   // max(1, round(nominal line length)) for ComputeLineScale
  int disable_phase_correction;
  double fsc_ratio;        // (OutputSampleRateHz / 1e6) / fsc MHz
  double cuc_hz;           // colour-under carrier
  double fsc_mhz;
  double burst_abs_ref;
  const float* het;        // [4][lines * line_len] BuildHeterodyneTable (float values)
  const float* bsin;       // [lines * line_len] BuildCarrierTables
  const float* bcos;
};

// Carried from call to call (reset with the field sequence).
struct ChromaState {
  int rot_index;           // the next field's track-phase start (NextChromaRotationIndex)
  int field_number;        // the next decoded field's number (NtscFieldPhaseTarget)
};

// Per call.
struct ChromaWork {
  const float* plane;      // chroma span plane (the block chroma output)
  long long plane_len;
  float* win;              // [F][lines][win_max] resampled burst windows
  double* probe;           // [F][2][lines][5]: phase, offset, magnitude, I, Q per start
  int* flip;               // [F][2]
  double* summ;            // [F][2][4]: mean magnitude, phase average, even, odd
  int* chosen;             // [F]: the start whose sequence the field uses
  int* field_no;           // [F]
  int* ones;               // [F] ones: the analysis resample's first line
  double* an_x;            // [F][lvl_stride] analysis level adjusts (lines 0 .. lines)
  double* an_y0;
  int lvl_stride;          // level-adjust entries per field (output lines + the largest first line)
  float* field;            // [F][lines * line_len]
  double* theta;           // [F][lines * line_len] NTSC upconversion phases
  float* fa;               // [F][iir_chunks * iir_chunk] the band-pass's sequence buffers (chunk-major)
  float* fb;
  float* iir_state;        // [F][iir_chunks][2][2 kMaxSections] chunk entry and end states
  int iir_chunks;
  ChromaState* state;
  unsigned short* out;     // [records][lines * line_len]
};

}  // namespace chroma
}  // namespace vhsd

extern "C" {
hipError_t vhsd_launch_chroma_state(vhsd::chroma::ChromaState* st, int reset, int field_number, int* ones,
                                    int max_fields, hipStream_t stream);
hipError_t vhsd_launch_chroma_analysis(const vhsd::chroma::ChromaCfg* cc, const vhsd::chroma::ChromaWork* cw,
                                       const double* refined, double* render, int line_stride, int nloc,
                                       const long long* origins, const double* pow_table, double nominal, int smooth,
                                       const float* lut, double fsc_shift, int max_fields, const int* decoded,
                                       hipStream_t stream);
hipError_t vhsd_launch_chroma_fields(const vhsd::chroma::ChromaCfg* cc, const vhsd::chroma::ChromaWork* cw,
                                     const double* render, int line_stride, const long long* origins,
                                     const int* first_lines, const double* line_x, const double* line_y0,
                                     const double* pow_table, double nominal, int smooth, const float* lut,
                                     const vhsd::vsync::FieldRecord* records, const vhsd::vsync::SeqCtl* ctl,
                                     const int* out_index, int max_fields, const int* decoded, hipStream_t stream);
hipError_t vhsd_launch_tbc_levels(const double* locs, int nloc, int lines, const int* first_lines, int line_stride,
                                  double nominal, double line_decay, double* line_x, double* line_y0,
                                  const int* field_count, int fields, hipStream_t stream);
}
