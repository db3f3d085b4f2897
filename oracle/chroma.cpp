// Oracle: CPU restatement of the reference's VHS colour-under chroma path in its default
// (v0.4.0) profile: burst-phase analysis of the TBC-resampled chroma field, burst-based
// line-location refinement, heterodyne upconversion, final band-pass, comb filter and
// automatic chroma gain to uint16.  TEST INFRASTRUCTURE ONLY: linked by tests/ and by
// bench.py's CPU baseline as the checker; the shipped library never links or calls it.
//
// Every function names the reference lines it restates (Decode/VhsChromaDecoder.cs
// unless noted).  Compiled with -ffp-contract=off; std::fma appears exactly where the
// reference calls Math.FusedMultiplyAdd / MathF.FusedMultiplyAdd.  Transcendentals are
// the C library's (cos, atan2, hypot, sinf, cosf), as the reference's Math.* calls are
// on Linux.
//
// Pinned in tests/test_chroma_oracle.py by the reference's vectors:
//  - heterodyne table values (CompatibilityTests.cs:2830-2842);
//  - DemodBurst in both the FP64 FMA form and the float32 form, bit for bit
//    (CompatibilityTests.cs:2949-2996, 3166-3176);
//  - the probe of an upconverted burst with a caller filter (:3005-3022);
//  - the phase sequence's rotation, colour killer, next index and averages (:3024-3117);
//  - UpconvertChroma / UpconvertChromaPhaseCompensated including the numba bit vector and
//    numpy negative-index slices (:2845-2921);
//  - RefineLineLocationsFromBurst for NTSC and PAL (:3119-3160);
//  - burst de-emphasis, comb and ACC vectors (:2717-2825);
//  - the PAL v0.4.0 phase-sequence bit vector (:3271-3345), which the reference freezes
//    on Windows (UCRT transcendentals): reproduced or reported as a signal.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

constexpr int kStartingLine = 16;          // StartingLine (:398)
constexpr double kBurstMagnitudeThreshold = 2.5e4;  // (:399)
constexpr int kBurstCheckSkipLines = 16;   // (:400)
constexpr double kS16AbsMax = 32767.0;     // (:404)

float bits_f(uint32_t u) {
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

// PositiveDegrees (:4430-4434): C#'s % on doubles is the truncated remainder.
double positive_degrees(double d) {
  const double r = std::fmod(d, 360.0);
  return r < 0.0 ? r + 360.0 : r;
}

int positive_modulo(int v, int m) {
  const int r = v % m;
  return r < 0 ? r + m : r;
}

// NumpyTrigFloat32 (:3834-3893): numpy's float32 sin/cos (Cody-Waite reduction and
// minimax polynomials in FMA form); arguments beyond the reduction's range take the C
// library's float sin/cos.
float numpy_trig_f32(float input, bool cosine) {
  if (std::isnan(input)) return NAN;
  const float max_cw = cosine ? 71476.0625f : 117435.992f;
  if (std::fabs(input) > max_cw) return cosine ? std::cos(input) : std::sin(input);
  const float two_over_pi = bits_f(0x3F22F983u);
  float quadrant = input * two_over_pi;
  const float round_magic = 12582912.0f;
  quadrant = (quadrant + round_magic) - round_magic;
  float reduced = std::fma(quadrant, bits_f(0xBFC90FD8u), input);
  reduced = std::fma(quadrant, bits_f(0xB4A8885Au), reduced);
  reduced = std::fma(quadrant, bits_f(0xA7C234C4u), reduced);
  const float squared = reduced * reduced;
  float c = std::fma(bits_f(0x37CC730Bu), squared, bits_f(0xBAB6036Eu));
  c = std::fma(c, squared, bits_f(0x3D2AAA9Eu));
  c = std::fma(c, squared, -0.5f);
  c = std::fma(c, squared, 1.0f);
  float s = std::fma(bits_f(0x363E9DDEu), squared, bits_f(0xB95035DDu));
  s = std::fma(s, squared, bits_f(0x3C0888CDu));
  s = std::fma(s, squared, -1.0f / 6.0f);
  s = std::fma(s, squared, 0.0f);
  s = std::fma(s, reduced, reduced);
  int iq = (int)std::nearbyint(quadrant);
  if (cosine) iq++;
  const float r = (iq & 1) == 0 ? s : c;
  return (iq & 2) != 0 ? -r : r;
}

struct Burst {
  double phase, offset, mag, i, q;
};

// DemodBurstValue FP64 form (:1438-1519) and DemodBurstFloat32 (:1521-1626).
Burst demod_burst(const double* burst, int len, double line_scale, int line_start, int burst_start,
                  const double* bsin, const double* bcos, bool f32) {
  double ic = 0.0, qc = 0.0;
  if (f32) {
    double ia[4][4] = {}, qa[4][4] = {};
    auto acc = [&](int index, double& ia1, double& qa1) {
      const float sample = (float)burst[index];
      const float ip = sample * (float)bcos[burst_start + index];
      const float qp = sample * (float)bsin[burst_start + index];
      ia1 += ip;
      qa1 += qp;
    };
    const int main_end = len & ~15;
    for (int index = 0; index < main_end; index += 16)
      for (int lane = 0; lane < 4; lane++)
        for (int g = 0; g < 4; g++) acc(index + lane + 4 * g, ia[g][lane], qa[g][lane]);
    auto hsum = [](double a[4][4]) {
      double l[4];
      for (int k = 0; k < 4; k++) l[k] = ((a[1][k] + a[0][k]) + a[2][k]) + a[3][k];
      return (l[0] + l[2]) + (l[1] + l[3]);
    };
    ic = hsum(ia);
    qc = hsum(qa);
    const int epi_end = len & ~3;
    if (epi_end > main_end) {
      double ei[4] = {ic, 0.0, 0.0, 0.0}, eq[4] = {qc, 0.0, 0.0, 0.0};
      for (int index = main_end; index < epi_end; index += 4)
        for (int lane = 0; lane < 4; lane++) acc(index + lane, ei[lane], eq[lane]);
      ic = (ei[0] + ei[2]) + (ei[1] + ei[3]);
      qc = (eq[0] + eq[2]) + (eq[1] + eq[3]);
    }
    for (int index = epi_end; index < len; index++) acc(index, ic, qc);
  } else {
    double i1[4] = {}, i2[4] = {}, q1[4] = {}, q2[4] = {};
    const int vec = len & ~7;
    for (int index = 0; index < vec; index += 8)
      for (int lane = 0; lane < 4; lane++) {
        const int a = index + lane, b = a + 4;
        i1[lane] = std::fma(burst[a], (double)(float)bcos[burst_start + a], i1[lane]);
        i2[lane] = std::fma(burst[b], (double)(float)bcos[burst_start + b], i2[lane]);
        q1[lane] = std::fma(burst[a], (double)(float)bsin[burst_start + a], q1[lane]);
        q2[lane] = std::fma(burst[b], (double)(float)bsin[burst_start + b], q2[lane]);
      }
    double i[4], q[4];
    for (int k = 0; k < 4; k++) {
      i[k] = i2[k] + i1[k];
      q[k] = q2[k] + q1[k];
    }
    ic = (i[0] + i[2]) + (i[1] + i[3]);
    qc = (q[0] + q[2]) + (q[1] + q[3]);
    for (int index = vec; index < len; index++) {
      ic = std::fma(burst[index], (double)(float)bcos[burst_start + index], ic);
      qc = std::fma(burst[index], (double)(float)bsin[burst_start + index], qc);
    }
  }
  Burst r;
  r.phase = positive_degrees(std::atan2(qc, ic) * (180.0 / M_PI));
  r.offset = positive_degrees((double)(burst_start - line_start) * std::fma(-line_scale, 90.0, 90.0));
  r.mag = std::hypot(ic, qc);
  r.i = ic;
  r.q = qc;
  return r;
}

}  // namespace

// Shared with oracle.cpp (same library).
namespace oracle_shared {
void sosfiltfilt_f32(const double* sos, int sections, const double* input, int n, int pad, float* output);
void sosfiltfilt_f64(const double* sos, int sections, const double* input, int n, int pad, double* output);
int default_pad_length(const double* sos, int sections);
}  // namespace oracle_shared

namespace {

// The burst filter of the phase analysis: the final SOS band-pass applied forward and
// backward in float32 (useFloat32Samples, AnalyzeFieldPhaseWithWorkspace :597-611) or in
// FP64 (SosFilter.ApplyForwardBackwardTo); in place on the padded window.
void burst_filter(const double* sos, int sections, bool f32, std::vector<double>& v) {
  const int n = (int)v.size();
  const int pad = oracle_shared::default_pad_length(sos, sections);
  if (n <= pad) return;  // the reference's filter requires a longer input; windows here are longer
  if (f32) {
    std::vector<float> out(n);
    oracle_shared::sosfiltfilt_f32(sos, sections, v.data(), n, pad, out.data());
    for (int i = 0; i < n; i++) v[i] = out[i];
  } else {
    std::vector<double> out(n);
    oracle_shared::sosfiltfilt_f64(sos, sections, v.data(), n, pad, out.data());
    v = out;
  }
}

// ProbeUpconvertedBurstValue (:1660-1713).  `het` is the heterodyne table of the
// line's phase rotation; sos == nullptr means no burst filter; filter_scale != 0 stands
// in for a caller filter that multiplies the window by a constant (tests).
Burst probe_burst(const double* chroma, int n, const double* het, int burst_start, int burst_end,
                  const double* bsin, const double* bcos, double line_scale, int line_number, int line_offset,
                  int line_len, const double* sos, int sections, bool f32, double filter_scale) {
  const int line_start = (line_number - line_offset) * line_len;  // GetLineRange (:4295-4309)
  const int pad = burst_end - burst_start;
  const int padded_start = std::max(0, line_start + burst_start - pad);
  const int padded_end = std::min(n, padded_start + burst_end + pad);
  std::vector<double> w(std::max(0, padded_end - padded_start));
  for (size_t i = 0; i < w.size(); i++) {
    const int src = padded_start + (int)i;
    const float hs = (float)het[src];
    w[i] = f32 ? (double)(float)(hs * (float)chroma[src]) : (double)hs * chroma[src];
  }
  if (sos) burst_filter(sos, sections, f32, w);
  if (filter_scale != 0.0)
    for (double& x : w) x = x * filter_scale;
  const int len = (int)w.size();
  const int fs = std::min(pad, len);
  const int fe = std::max(fs, len - pad);
  return demod_burst(w.data() + fs, fe - fs, line_scale, line_start, padded_start + pad, bsin, bcos, f32);
}

struct PhaseLine {
  int line, rot;
  Burst b;
};

// ComputeLineScale (:4152-4162).
double line_scale(const double* loc, int line, int input_line_length, int last_line) {
  return line < last_line - 1 ? (loc[line + 1] - loc[line]) / input_line_length : 1.0;
}

// GetNumpySliceRange (:4310-4340).
void numpy_slice(int n, int line_offset, int line_len, int line, int& s, int& e) {
  s = (line - line_offset) * line_len;
  e = s + line_len;
  auto norm = [&](int i) {
    if (i < 0) i += n;
    return std::clamp(i, 0, n);
  };
  s = norm(s);
  e = norm(e);
  if (e <= s) e = s;
}

}  // namespace

extern "C" {

// BuildHeterodyneTable (:1967-2007): out4[p * n + i] = (float)-cos(tau * scale * i + p pi/2 + drift).
void oracle_chroma_heterodyne(int n, double fsc_mhz, double cuc_mhz, double rate_mhz, double drift, double* out4) {
  const double het_scale = (fsc_mhz + cuc_mhz) / rate_mhz;
  for (int p = 0; p < 4; p++) {
    const double off = ((M_PI / 2.0) * p) + drift;
    for (int i = 0; i < n; i++) out4[(size_t)p * n + i] = (double)(float)-std::cos(((2.0 * M_PI) * het_scale * i) + off);
  }
}

// BuildCarrierTables (:3793-3832).
void oracle_chroma_carrier(int n, double carrier_mhz, double rate_mhz, double* bsin, double* bcos) {
  const double scale = carrier_mhz / rate_mhz;
  for (int i = 0; i < n; i++) {
    const float theta = (float)((2.0 * M_PI) * scale * i);
    bsin[i] = numpy_trig_f32(theta, false);
    bcos[i] = numpy_trig_f32(theta, true);
  }
}

float oracle_numpy_trig_f32(float x, int cosine) { return numpy_trig_f32(x, cosine != 0); }

// DemodBurst (:1421-1436): out = phase, offset, magnitude, I, Q.
void oracle_chroma_demod_burst(const double* burst, int len, double line_scale, int line_start, int burst_start,
                               const double* bsin, const double* bcos, int f32, double* out5) {
  const Burst b = demod_burst(burst, len, line_scale, line_start, burst_start, bsin, bcos, f32 != 0);
  out5[0] = b.phase;
  out5[1] = b.offset;
  out5[2] = b.mag;
  out5[3] = b.i;
  out5[4] = b.q;
}

// ProbeUpconvertedBurst (:1631-1713) with the final SOS filter (sos may be NULL).
void oracle_chroma_probe(const double* chroma, int n, const double* het, int burst_start, int burst_end,
                         const double* bsin, const double* bcos, double scale, int line_number, int line_offset,
                         int line_len, const double* sos, int sections, int f32, double filter_scale, double* out5) {
  const Burst b = probe_burst(chroma, n, het, burst_start, burst_end, bsin, bcos, scale, line_number, line_offset,
                              line_len, sos, sections, f32 != 0, filter_scale);
  out5[0] = b.phase;
  out5[1] = b.offset;
  out5[2] = b.mag;
  out5[3] = b.i;
  out5[4] = b.q;
}

// AnalyzeFieldPhaseWithWorkspace (:574-702) -> GetPhaseRotationSequenceCore (:1906-1965):
// BuildPhaseSequence (:3927-4036; the track-phase check off, one worker), ShouldFlipTrackPhase
// (:4164-4217) and SummarizePhaseSequence (:4219-4293).  `probe_table` (optional) replaces
// the probe with given per-line results [line][rotation][5] (tests of the sequence logic).
// Outputs per line (lines_out): rot, burst (phase, offset, mag, I, Q); summary[4] = mean
// magnitude, phase average, even average, odd average; ints[2] = next rotation index,
// burst-detected line.
int oracle_chroma_phase(const double* chroma, int n, const double* loc, int n_loc, int line_offset, int lines_out,
                        int input_line_length, int out_line_len, const double* het4, const double* bsin,
                        const double* bcos, const double* sos, int sections, int burst_start, int burst_end,
                        const int* rotation, int rotation_count, int rotation_index_in, int has_rotation_index,
                        int f32, int color_killer, int prev_burst_line, int is_ntsc, const double* probe_table,
                        int* rot_out, double* burst_out, double* summary, int* ints) {
  const int end = lines_out + line_offset;
  if (n_loc <= end) return -1;
  auto probe = [&](int line, int rot) -> Burst {
    if (probe_table) {
      const double* t = probe_table + ((size_t)line * 4 + rot) * 5;
      return Burst{t[0], t[1], t[2], t[3], t[4]};
    }
    return probe_burst(chroma, n, het4 + (size_t)rot * n, burst_start, burst_end, bsin, bcos,
                       line_scale(loc, line, input_line_length, end), line, line_offset, out_line_len, sos,
                       sections, f32 != 0, 0.0);
  };
  const bool has_rot = rotation_count > 0;
  auto build = [&](int start_index, std::vector<PhaseLine>& seq) {
    int idx, track;
    if (has_rot) {
      idx = start_index;
      track = rotation[idx];
    } else {
      idx = 0;
      track = start_index;
    }
    seq.assign(std::max(0, end - line_offset), PhaseLine{});
    int cur = 0;
    for (int line = line_offset; line < end; line++) {
      cur = positive_modulo(cur + track, 4);
      seq[line - line_offset] = PhaseLine{line, cur, probe(line, cur)};
    }
    if (has_rot && idx == start_index) idx = positive_modulo(idx + 1, 2);
    return idx;
  };
  const int start = has_rotation_index ? rotation_index_in : 0;
  std::vector<PhaseLine> seq;
  int next = build(start, seq);
  // ShouldFlipTrackPhase
  bool flip = false;
  if (has_rot) {
    int d[4] = {0, 0, 0, 0};
    const int check_end = end - kBurstCheckSkipLines;
    for (size_t i = 1; i < seq.size(); i++) {
      const PhaseLine& prev = seq[i - 1];
      const PhaseLine& cur = seq[i];
      if (cur.line <= kBurstCheckSkipLines || cur.line >= check_end) continue;
      const double delta = positive_degrees(cur.b.phase - prev.b.phase);
      const int bucket = (int)std::floor((delta + 45.0) / 90.0) % 4;
      d[bucket]++;
    }
    flip = is_ntsc ? d[0] < d[2] : (d[1] + d[3]) < (d[0] + d[2]);
  }
  if (flip) next = build(next, seq);
  // SummarizePhaseSequence
  const int check_end = end - kBurstCheckSkipLines;
  int burst_line = 0;
  double ei = 0.0, eq = 0.0, oi = 0.0, oq = 0.0, mag_avg = 0.0;
  int count = 0;
  for (const PhaseLine& pl : seq) {
    if (pl.line <= kBurstCheckSkipLines || pl.line >= check_end) continue;
    if (pl.b.mag == 0.0) continue;
    const double ni = pl.b.i / pl.b.mag, nq = pl.b.q / pl.b.mag;
    count++;
    mag_avg += pl.b.mag;
    if (color_killer && prev_burst_line == -1 && burst_line == 0 && pl.b.mag > kBurstMagnitudeThreshold)
      burst_line = pl.line;
    if ((pl.line & 1) == 1) {
      oi += ni;
      oq += nq;
    } else {
      ei += ni;
      eq += nq;
    }
  }
  if (count == 0) return -2;
  mag_avg /= count;
  if (color_killer && mag_avg < kBurstMagnitudeThreshold) burst_line = -1;
  for (size_t i = 0; i < seq.size(); i++) {
    rot_out[i] = seq[i].rot;
    double* b = burst_out + 5 * i;
    b[0] = seq[i].b.phase;
    b[1] = seq[i].b.offset;
    b[2] = seq[i].b.mag;
    b[3] = seq[i].b.i;
    b[4] = seq[i].b.q;
  }
  summary[0] = mag_avg;
  summary[1] = positive_degrees(std::atan2(eq + oq, ei + oi) * 180.0 / M_PI);
  summary[2] = positive_degrees(std::atan2(eq, ei) * 180.0 / M_PI);
  summary[3] = positive_degrees(std::atan2(oq, oi) * 180.0 / M_PI);
  ints[0] = next;
  ints[1] = burst_line;
  return 0;
}

// RefineLineLocationsFromBurst (:2342-2428), v0.4.0 form (no frequency-drift term).
// Phase line k is line seq_first_line + k with burst phase phase[k] and offset offset[k].
void oracle_chroma_refine_lines(const double* loc, int n_loc, int out_line_len, double fsc_ratio, int seq_first_line,
                                int seq_count, const double* phase, const double* offset, int burst_line,
                                double avg_all, double avg_even, double avg_odd, int is_ntsc, double* out) {
  for (int i = 0; i < n_loc; i++) out[i] = loc[i];
  if (burst_line == -1) return;
  for (int k = std::max(9, burst_line); k < seq_count; k++) {
    const int line = seq_first_line + k;
    if (line < 0 || line + 1 >= n_loc) continue;
    const double target = is_ntsc ? avg_all : ((line & 1) == 1 ? avg_odd : avg_even);
    const double delta = positive_degrees(((target - phase[k]) + offset[k]) + 180.0) - 180.0;
    const double ll = out[line + 1] - out[line];
    const double scale = ll / out_line_len;
    out[line] += ((delta / 360.0) * fsc_ratio) * scale;
  }
}

// ApplyBurstDeemphasisCore (:2695-2719).
void oracle_chroma_burst_deemphasis(double* v, int line_offset, int lines_out, int line_len, int burst_end,
                                    int after) {
  const int first = burst_end + after;
  if (first >= line_len) return;
  for (int line = line_offset; line < lines_out + line_offset; line++) {
    const int s = (line - line_offset) * line_len;
    for (int i = s + first; i < s + line_len; i++) v[i] *= 2.0;
  }
}

// UpconvertChroma (:2009-2045): lines from GetNumpySliceRange, zero elsewhere.
void oracle_chroma_upconvert(const double* chroma, int n, int line_offset, int line_len, const int* lines,
                             const int* rots, int count, const double* het4, int het_len, double* out) {
  std::fill(out, out + n, 0.0);
  for (int k = 0; k < count; k++) {
    int s, e;
    numpy_slice(n, line_offset, line_len, lines[k], s, e);
    const double* het = het4 + (size_t)rots[k] * het_len;
    for (int i = s; i < e; i++) out[i] = (double)(float)(chroma[i] * het[i]);
  }
}

// UpconvertChromaPhaseCompensated (:2167-2249): per line a running FP64 phase in four
// lanes (the numba vectorised loop), cos of it times the sample, to float.
void oracle_chroma_upconvert_compensated(const double* chroma, int n, int line_offset, int line_len, const int* lines,
                                         const int* rots, const double* burst_phase, int count, double cuc_hz,
                                         double fsc_mhz, double target_even, double target_odd, double* out) {
  const double half_pi = 1.5707963267948966;
  const double deg2rad = 0.017453292519943295;
  const double half_pi_per_mhz = 1.5707963267948965e-6;
  std::fill(out, out + n, 0.0);
  const double carrier_term = cuc_hz * half_pi_per_mhz;
  const double carrier_ratio = carrier_term / fsc_mhz;
  const double het = carrier_ratio + half_pi;
  const double twice = het * 2.0, thrice = het * 3.0, four = het * 4.0;
  for (int k = 0; k < count; k++) {
    const int s = (lines[k] - line_offset) * line_len;
    const int e = s + line_len;
    const double target = lines[k] % 2 == 0 ? target_even : target_odd;
    const double phase_term = rots[k] * half_pi;
    double theta = std::fma((double)s, het, phase_term);
    theta = std::fma(target + burst_phase[k], deg2rad, theta);
    if (s < 0 || e > n) {
      for (int idx = s; idx < e; idx++) {
        int si = idx < 0 ? idx + n : idx;
        if (si < 0 || si >= n) continue;  // the reference throws here
        out[si] = (double)(float)(chroma[si] * -std::cos(theta));
        theta += het;
      }
      continue;
    }
    const int vc = (e - s) & ~3;
    const int ve = s + vc;
    double t0 = theta + (het * 0.0), t1 = theta + het, t2 = theta + twice, t3 = theta + thrice;
    int i = s;
    for (; i < ve; i += 4) {
      out[i] = (double)(float)(chroma[i] * -std::cos(t0));
      out[i + 1] = (double)(float)(chroma[i + 1] * -std::cos(t1));
      out[i + 2] = (double)(float)(chroma[i + 2] * -std::cos(t2));
      out[i + 3] = (double)(float)(chroma[i + 3] * -std::cos(t3));
      t0 += four;
      t1 += four;
      t2 += four;
      t3 += four;
    }
    theta += het * vc;
    for (; i < e; i++) {
      out[i] = (double)(float)(chroma[i] * -std::cos(theta));
      theta += het;
    }
  }
}

// RmsFloat32 (:4399-4421) and Rms (:4376-4397).
double oracle_chroma_rms(const double* v, int n, int f32) {
  if (n == 0) return 0.0;
  if (f32) {
    float mean = 0.0f;
    for (int i = 0; i < n; i++) mean += (float)v[i];
    mean /= n;
    float ss = 0.0f;
    for (int i = 0; i < n; i++) {
      const float c = (float)v[i] - mean;
      const float sq = c * c;
      ss += sq;
    }
    return std::sqrt(ss / n);
  }
  double mean = 0.0;
  for (int i = 0; i < n; i++) mean += v[i];
  mean /= n;
  double ss = 0.0;
  for (int i = 0; i < n; i++) {
    const double c = v[i] - mean;
    ss += c * c;
  }
  return std::sqrt(ss / n);
}

// ApplyAutomaticChromaGain (:2721-2763): returns the mean burst RMS.
double oracle_chroma_acc(const double* chroma, int n, double burst_abs_ref, int burst_start, int burst_end,
                         int line_len, int lines, int burst_line, int f32_rms, double* out) {
  std::fill(out, out + n, 0.0);
  double acc = 0.0;
  for (int line = kStartingLine; line < lines; line++) {
    const int s = line * line_len;
    if (line < burst_line) continue;
    const double rms = oracle_chroma_rms(chroma + s + burst_start, burst_end - burst_start, f32_rms);
    const double scale = rms != 0.0 ? burst_abs_ref / rms : 1.0;
    for (int i = 0; i < line_len; i++) out[s + i] = chroma[s + i] * scale;
    acc += rms;
  }
  return acc / (lines - kStartingLine);
}

// ApplyAutomaticChromaGainWithCombToU16 (:3504-3583) with WriteScaledChromaToU16
// (:3605-3656): comb of each line with the lines `distance` away (the last two lines
// uncombed), burst RMS, gain to the reference level, uint16 around 32767.
void oracle_chroma_acc_comb_u16(const double* chroma, int n, double burst_abs_ref, int burst_start, int burst_end,
                                int line_len, int lines, int burst_line, int distance, int retain_f32, int f32_rms,
                                uint16_t* out) {
  const int first = std::min(lines, std::max(kStartingLine, burst_line));
  const int configured = lines * line_len;
  for (int i = 0; i < first * line_len; i++) out[i] = (uint16_t)kS16AbsMax;
  for (int i = configured; i < n; i++) out[i] = (uint16_t)kS16AbsMax;
  const int vec_len = n & ~3;
  std::vector<double> comb(line_len);
  auto write = [&](const double* src, int s, double scale) {
    const int sat = std::clamp(vec_len - s, 0, line_len);
    for (int i = 0; i < line_len; i++) {
      const double scaled = src[i] * scale;
      const double shifted = scaled + kS16AbsMax;
      uint16_t o;
      if (i < sat) {
        o = (!std::isfinite(shifted) || shifted <= 0.0) ? 0 : shifted >= 65535.0 ? 65535 : (uint16_t)shifted;
      } else {
        o = (!std::isfinite(shifted) || shifted < -9.2233720368547758e18 || shifted > 9.2233720368547758e18)
                ? 0
                : (uint16_t)(int64_t)shifted;
      }
      out[s + i] = o;
    }
  };
  for (int line = first; line < lines; line++) {
    const int s = line * line_len;
    if (line < lines - 2) {
      const double* cur = chroma + s;
      const double* adv = chroma + (line + distance) * line_len;
      const double* del = chroma + (line - distance) * line_len;
      for (int i = 0; i < line_len; i++) {
        const double c = (!retain_f32 && distance == 2) ? (((cur[i] * 2.0) - del[i]) - adv[i]) / 4.0
                                                         : (((cur[i] * 2.0) - adv[i]) - del[i]) / 4.0;
        comb[i] = retain_f32 ? (double)(float)c : c;
      }
      const double rms = oracle_chroma_rms(comb.data() + burst_start, burst_end - burst_start, f32_rms);
      write(comb.data(), s, rms != 0.0 ? burst_abs_ref / rms : 1.0);
    } else {
      const double rms = oracle_chroma_rms(chroma + s + burst_start, burst_end - burst_start, f32_rms);
      write(chroma + s, s, rms != 0.0 ? burst_abs_ref / rms : 1.0);
    }
  }
}

// ApplyCombInPlace (:3658-3704) (ApplyNtscComb distance 1, ApplyPalComb distance 2).
void oracle_chroma_comb(double* v, int n, int line_len, int distance, int retain_f32) {
  const int count = n / line_len;
  if (count <= kStartingLine + 2) return;
  std::vector<double> delayed(v + (kStartingLine - distance) * line_len,
                              v + (kStartingLine - distance) * line_len + distance * line_len);
  for (int line = kStartingLine; line < count - 2; line++) {
    const int s = line * line_len, a = (line + distance) * line_len;
    double* d = delayed.data() + ((line - kStartingLine) % distance) * line_len;
    for (int i = 0; i < line_len; i++) {
      const double cur = v[s + i], adv = v[a + i], del = d[i];
      const double c = (!retain_f32 && distance == 2) ? (((cur * 2.0) - del) - adv) / 4.0
                                                       : (((cur * 2.0) - adv) - del) / 4.0;
      v[s + i] = retain_f32 ? (double)(float)c : c;
      d[i] = cur;
    }
  }
}

// NtscFieldPhaseTarget (:3915-3925): target phase in degrees; *phase_id receives the id.
double oracle_chroma_ntsc_target(int is_first_field, int field_number, int* phase_id) {
  const bool second = ((field_number / 2) & 1) == 1;
  int id;
  double t;
  if (is_first_field && !second) {
    id = 1;
    t = -33.0;
  } else if (!is_first_field && second) {
    id = 2;
    t = 147.0;
  } else if (is_first_field && second) {
    id = 3;
    t = 147.0;
  } else {
    id = 4;
    t = -33.0;
  }
  if (phase_id) *phase_id = id;
  return t;
}

// CompleteFieldWithPhaseCore (:953-1251), v0.4.0 profile without the AFC pre-filter or the
// chroma de-emphasis filter: NTSC burst de-emphasis and phase-compensated upconversion
// (the field's parity known), PAL heterodyne upconversion; the final SOS band-pass forward
// and backward in float32 over the whole field; comb + ACC to uint16.  `field` (lines x
// line_len, from the render resample) is modified.  A killed burst (burst_line == -1)
// gives the neutral field.
void oracle_chroma_decode(double* field, int lines, int line_len, int line_offset, const int* seq_lines,
                          const int* seq_rots, const double* seq_phase, int seq_count, int is_ntsc, int parity_known,
                          int is_first_field, int field_number, double cuc_hz, double fsc_mhz, const double* het4,
                          const double* sos, int sections, int burst_start, int burst_end, double burst_abs_ref,
                          int burst_line, uint16_t* out) {
  const int n = lines * line_len;
  if (burst_line == -1) {
    for (int i = 0; i < n; i++) out[i] = (uint16_t)kS16AbsMax;
    return;
  }
  if (is_ntsc) oracle_chroma_burst_deemphasis(field, line_offset, lines, line_len, burst_end, 5);
  std::vector<double> up(n);
  if (is_ntsc && parity_known) {
    const double t = oracle_chroma_ntsc_target(is_first_field, field_number, nullptr);
    oracle_chroma_upconvert_compensated(field, n, line_offset, line_len, seq_lines, seq_rots, seq_phase, seq_count,
                                        cuc_hz, fsc_mhz, t, t, up.data());
  } else {
    oracle_chroma_upconvert(field, n, line_offset, line_len, seq_lines, seq_rots, seq_count, het4, n, up.data());
  }
  {
    const int pad = oracle_shared::default_pad_length(sos, sections);
    std::vector<float> f(n);
    oracle_shared::sosfiltfilt_f32(sos, sections, up.data(), n, pad, f.data());
    for (int i = 0; i < n; i++) up[i] = f[i];
  }
  oracle_chroma_acc_comb_u16(up.data(), n, burst_abs_ref, burst_start, burst_end, line_len, lines, burst_line,
                             is_ntsc ? 1 : 2, 1, 1, out);
}

}  // extern "C"
