// Oracle: CPU FP64 restatement of the reference's VHS (v0.4.0 profile) field sync path:
// per-field level detection, sync pulses, pulse classification, field parity, line 0,
// line locations and the read position of the next field.  TEST INFRASTRUCTURE ONLY:
// linked into liboracle.so, which tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg load as the checker; the shipped library never links or calls it.
//
// Reference files (paths relative to the reference checkout):
//  - src/VHSDecode.Core/Decode/TbcFieldDecodePipeline.cs  DecodePrepared :1071-1290,
//    PrepareSyncSpan :2776-3190, PrepareSyncSpanFromLevels :3197-3234, FindRawPulses
//    :2375-2384, ComputeNextFieldOffsetSamples :2245-2325, ResolveLine0Location :5452-5651,
//    level selection helpers :5679-5782, ThrowIfInsufficientFieldData :5961-6008,
//    SelectSingleVBlankEstimate :6010-6031, UpdateSyncHistory / EstimateFirstHSyncFromPrevious
//    / AlignLine0Anchor / AlignFirstHSyncToValidPulses / TryEstimateLine0FromPrevious
//    :6119-6375, BuildRecoveryException :5936-5959.
//  - src/VHSDecode.Core/Dsp/SyncAnalysis.cs  (SyncAnalyzer) :88-785.
//  - src/VHSDecode.Core/Dsp/VsyncSerrationDetector.cs :1-755.
//  - src/VHSDecode.Core/Dsp/LevelDetection.cs :1-595, VhsFieldLevelState.cs :1-83.
//  - src/VHSDecode.Core/Dsp/FieldParityDetector.cs :16-268.
//  - src/VHSDecode.Core/Dsp/VBlankSyncResolver.cs :317-747.
//  - src/VHSDecode.Core/Decode/TbcFieldSequenceDecodeEngine.cs :161-171, 1936-1941 and
//    Decode/DecodeReadWindowPlanner.cs :15-63 (read windows of the field sequence).
//  - src/VHSDecode.Core/Dsp/NumpyReduction.cs, NumbaReduction.cs (reductions).
//
// Every FP64 step keeps the reference's operation order (-ffp-contract=off).
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

namespace {

using std::vector;

// Math.Round(x, MidpointRounding.ToEven)
double round_even(double x) { return std::nearbyint(x); }
// Math.Round(x, MidpointRounding.AwayFromZero)
double round_away(double x) { return std::round(x); }

// ------------------------------------------------------------------ reductions
// NumpyReduction.PairwiseSumFloat64 (Dsp/NumpyReduction.cs:400-451).
double pairwise_sum_f64(const double* v, int64_t n) {
  if (n < 8) {
    double s = -0.0;
    for (int64_t i = 0; i < n; i++) s += v[i];
    return s;
  }
  if (n > 128) {
    int64_t split = n / 2;
    split -= split % 8;
    return pairwise_sum_f64(v, split) + pairwise_sum_f64(v + split, n - split);
  }
  double r[8];
  for (int k = 0; k < 8; k++) r[k] = v[k];
  int64_t i = 8;
  const int64_t end = n - (n % 8);
  for (; i < end; i += 8)
    for (int k = 0; k < 8; k++) r[k] += v[i + k];
  double s = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
  for (; i < n; i++) s += v[i];
  return s;
}

// NumpyReduction.MeanFloat64 (Dsp/NumpyReduction.cs:226-227).
double mean_f64(const double* v, int64_t n) { return pairwise_sum_f64(v, n) / (double)n; }

// NumbaReduction.SumFloat64FastMath (Dsp/NumbaReduction.cs:227-283): 16 accumulators.
double mean_f64_fastmath(const double* v, int64_t n) {
  double acc[16] = {0};
  int64_t i = 0;
  const int64_t vend = n - (n % 16);
  for (; i < vend; i += 16)
    for (int k = 0; k < 16; k++) acc[k] += v[i + k];
  double lanes[4];
  for (int l = 0; l < 4; l++) {
    const double left = acc[l] + acc[4 + l];
    const double right = acc[8 + l] + acc[12 + l];
    lanes[l] = left + right;
  }
  double sum = (lanes[0] + lanes[2]) + (lanes[1] + lanes[3]);
  const int64_t four_end = n - ((n - i) % 4);
  if (i < four_end) {
    lanes[0] = sum;
    lanes[1] = lanes[2] = lanes[3] = 0.0;
    for (; i < four_end; i += 4)
      for (int l = 0; l < 4; l++) lanes[l] += v[i + l];
    sum = (lanes[0] + lanes[2]) + (lanes[1] + lanes[3]);
  }
  for (; i < n; i++) sum += v[i];
  return sum / (double)n;
}

// NumpyReduction.MedianFloat64 (Dsp/NumpyReduction.cs:12-104): NaN propagates; the
// middle order statistic, or the mean of the two middle ones for an even count.
double median_f64(const double* v, int64_t n) {
  if (n <= 0) return std::numeric_limits<double>::quiet_NaN();
  vector<double> w(v, v + n);
  for (double x : w)
    if (std::isnan(x)) return x;
  std::sort(w.begin(), w.end());
  const int64_t mid = n / 2;
  return (n & 1) ? w[mid] : (w[mid - 1] + w[mid]) / 2.0;
}
double median_f64(const vector<double>& v) { return median_f64(v.data(), (int64_t)v.size()); }

// VsyncSerrationDetector.MinimumFloat64 (:698-724).
double minimum_f64(const double* v, int64_t n) {
  double m = v[0];
  if (std::isnan(m)) return m;
  for (int64_t i = 1; i < n; i++) {
    if (std::isnan(v[i])) return v[i];
    if (v[i] < m) m = v[i];
  }
  return m;
}

bool in_range(double v, double lo, double hi) { return v >= lo && v <= hi; }

// ------------------------------------------------------------------ pulses
struct Pulse {
  int start;
  int length;
};

// PulseDetection.FindPulses (Dsp/PulseDetection.cs:45-222).
vector<Pulse> find_pulses(const double* x, int64_t n, double high, int min_len, int max_len) {
  vector<Pulse> out;
  if (n <= 0) return out;
  bool in_pulse = x[0] <= high;
  int64_t current = 0;
  for (int64_t p = 0; p < n; p++) {
    const double v = x[p];
    if (in_pulse) {
      if (v > high) {
        const int64_t len = p - current;
        if (in_range((double)len, min_len, max_len) && current != 0) out.push_back({(int)current, (int)len});
        in_pulse = false;
      }
    } else if (v <= high) {
      current = p;
      in_pulse = true;
    }
  }
  return out;
}

// ------------------------------------------------------------------ IIR (transfer function)
struct Tf {
  vector<double> b, a;  // a[0] == 1
};

// IirFilter.ApplyForwardBackwardInPlace (Dsp/IirFilter.cs:78-115, 187-373, 375-420, 476-498).
void tf_filtfilt(const Tf& f, double* x, int64_t n) {
  const int order = (int)f.a.size() - 1;
  const double* b = f.b.data();
  const double* a = f.a.data();
  const int64_t m = order + 1;
  int64_t edge = std::min<int64_t>(3 * m, n - 1);
  auto sum_numpy = [](const double* v, int count) {
    if (count < 8) {
      double s = -0.0;
      for (int i = 0; i < count; i++) s += v[i];
      return s;
    }
    double r[8];
    for (int k = 0; k < 8; k++) r[k] = v[k];
    int i = 8;
    const int end = count - (count % 8);
    for (; i < end; i += 8)
      for (int k = 0; k < 8; k++) r[k] += v[i + k];
    double s = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
    for (; i < count; i++) s += v[i];
    return s;
  };
  const double steady = sum_numpy(b, (int)m) / sum_numpy(a, (int)m);
  vector<double> zi(order);
  double cum = 0.0;
  for (int i = order; i >= 1; i--) {
    cum += b[i] - (steady * a[i]);
    zi[i - 1] = cum;
  }
  const int64_t L = n + 2 * edge;
  vector<double> e(L);
  if (edge > 0) {
    for (int64_t i = 0; i < edge; i++) e[i] = (2.0 * x[0]) - x[edge - i];
    for (int64_t i = 0; i < n; i++) e[edge + i] = x[i];
    for (int64_t i = 0; i < edge; i++) e[edge + n + i] = (2.0 * x[n - 1]) - x[n - 2 - i];
  } else {
    for (int64_t i = 0; i < n; i++) e[i] = x[i];
  }
  auto pass = [&]() {
    vector<double> s(order);
    for (int i = 0; i < order; i++) s[i] = zi[i] * e[0];
    for (int64_t t = 0; t < L; t++) {
      const double xv = e[t];
      const double y = (b[0] * xv) + (order > 0 ? s[0] : 0.0);
      for (int i = 1; i < order; i++) s[i - 1] = ((b[i] * xv) + s[i]) - (a[i] * y);
      if (order > 0) s[order - 1] = (b[order] * xv) - (a[order] * y);
      e[t] = y;
    }
  };
  pass();
  std::reverse(e.begin(), e.end());
  pass();
  std::reverse(e.begin(), e.end());
  for (int64_t i = 0; i < n; i++) x[i] = e[edge + i];
}

// The same Butterworth filter as second-order sections, forward-backward with odd
// extension of the transfer function's pad length and each section's steady state
// scaled by the pass's first sample (the well-conditioned evaluation the device uses).
struct Sos {
  vector<double> c;  // [sections][6] b0 b1 b2 a0 a1 a2, a0 == 1
  int pad;
};

void sos_filtfilt(const Sos& f, double* x, int64_t n) {
  const int ns = (int)f.c.size() / 6;
  int64_t edge = std::min<int64_t>(f.pad, n - 1);
  const int64_t L = n + 2 * edge;
  vector<double> e(L);
  for (int64_t i = 0; i < edge; i++) e[i] = (2.0 * x[0]) - x[edge - i];
  for (int64_t i = 0; i < n; i++) e[edge + i] = x[i];
  for (int64_t i = 0; i < edge; i++) e[edge + n + i] = (2.0 * x[n - 1]) - x[n - 2 - i];
  // steady state of the cascade for a unit step, section by section (scipy sosfilt_zi)
  vector<double> zi(2 * ns);
  double scale = 1.0;
  for (int s = 0; s < ns; s++) {
    const double* c = &f.c[6 * s];
    const double b0 = c[0], b1 = c[1], b2 = c[2], a1 = c[4], a2 = c[5];
    // DF2T steady state: z1 = (b1 + b2 - (a1 + a2) * G) ... solved for unit input
    const double G = (b0 + b1 + b2) / (1.0 + a1 + a2);
    const double z2 = b2 - a2 * G;
    const double z1 = (b1 - a1 * G) + z2;
    zi[2 * s] = z1 * scale;
    zi[2 * s + 1] = z2 * scale;
    scale *= G;
  }
  auto pass = [&]() {
    for (int s = 0; s < ns; s++) {
      const double* c = &f.c[6 * s];
      const double b0 = c[0], b1 = c[1], b2 = c[2], a1 = c[4], a2 = c[5];
      double z1 = zi[2 * s] * e[0], z2 = zi[2 * s + 1] * e[0];
      for (int64_t t = 0; t < L; t++) {
        const double xv = e[t];
        const double y = b0 * xv + z1;
        z1 = b1 * xv - a1 * y + z2;
        z2 = b2 * xv - a2 * y;
        e[t] = y;
      }
    }
  };
  pass();
  std::reverse(e.begin(), e.end());
  pass();
  std::reverse(e.begin(), e.end());
  for (int64_t i = 0; i < n; i++) x[i] = e[edge + i];
}

// ------------------------------------------------------------------ configuration
}  // namespace

extern "C" {
// Sync configuration of one tape format at one sample rate (mirrored by oracle.py).
typedef struct OracleSyncConfig {
  double sample_rate_hz;
  int32_t is_pal;
  double line_period_us;      // sysparams line_period
  double hsync_us, eq_us, vsync_us;
  int32_t num_pulses;
  double hsync_tolerance_us;  // 0.7 for vhs (DecodeSession.cs:424-428)
  double eq_tolerance_us;     // 0.9 for vhs
  double ire0, hz_ire, vsync_ire;
  double fps, frame_lines;
  int32_t output_lines;       // TbcFrameSpec.OutputLineCount (NTSC 263, PAL 313)
  int64_t block_length, block_cut, block_stride;
  int64_t read_samples;       // DecodeReadWindowPlanner.EstimateReadSampleCount
  int32_t level_divisor;      // level_detect_divisor (1 unless the rate is low)
  int32_t serration_mode;     // 0: the reference's transfer-function recurrences; 1: SOS evaluation;
                              // 2: no serration search (Analyze measures nothing, as on the device)
  // four filters of VsyncSerrationDetector: vsync envelope low-pass, serration high-pass,
  // serration low-pass, serration envelope low-pass; each as TF (order, b[order+1], a[order+1])
  // and as SOS (sections, [sections][6])
  int32_t tf_order[4];
  double tf_b[4][21];
  double tf_a[4][21];
  int32_t sos_sections[4];
  double sos[4][10][6];
} OracleSyncConfig;

// One field of the sequence.
typedef struct OracleSyncField {
  int64_t begin;              // intended start sample of the read
  int64_t window_start;       // DecodeReadWindowPlanner.Resolve span start (global)
  int64_t window_length;
  int32_t status;             // 0 decoded; 1 no sync pulses, 2 no first hsync, 3 insufficient data
  int32_t level_source;       // 0 default threshold, 1 serration levels, 2 refined levels, 3 saved, 4 reversed defaults
  double sync_level, blank_level, dc_offset, threshold;
  int32_t serration_candidates, serration_measurements;
  int32_t raw_pulses, refined_pulses;
  double hsync_median;
  double mean_line_length;
  int32_t is_first_field;
  int32_t parity_confidence;
  double line0;               // window-relative
  double first_hsync;         // window-relative
  double first_hsync_line;
  int32_t processed_lines;
  double next_offset;         // ComputeNextFieldOffsetSamples (window-relative)
  int64_t next_begin;
} OracleSyncField;
}

namespace {

enum Kind { kHSync = 0, kEq = 1, kVSync = 2, kEq2 = 3 };

struct CPulse {
  Kind kind;
  Pulse p;
  bool in_order;
};

struct Range {
  double lo, hi;
  bool contains(double v) const { return v >= lo && v <= hi; }
};

struct Timing {
  double nominal, hsync_median, offset;
  Range hsync, eq, vsync;
};

// SyncAnalyzer (Dsp/SyncAnalysis.cs:31-785) built as FromParameters does (:88-114).
struct Analyzer {
  double fs_hz, fs_mhz, line_us, hsync_us, eq_us, vsync_us;
  int num_pulses;
  double hsync_tol, eq_tol, nominal;

  // new SyncAnalyzer(...) (:35-66): the constructor keeps the line period as given.
  void init_direct(const double* a) {
    fs_hz = a[0];
    fs_mhz = fs_hz / 1000000.0;
    line_us = a[1];
    hsync_us = a[2];
    eq_us = a[3];
    vsync_us = a[4];
    num_pulses = (int)a[5];
    hsync_tol = a[6];
    eq_tol = a[7];
    nominal = us(line_us);
  }

  void init(const OracleSyncConfig& c) {
    fs_hz = c.sample_rate_hz;
    fs_mhz = fs_hz / 1000000.0;
    double lp = c.line_period_us;
    if (fs_mhz > 0.0) {
      const double nl = round_even(lp * fs_mhz);
      lp = nl / fs_mhz;
    }
    line_us = lp;
    hsync_us = c.hsync_us;
    eq_us = c.eq_us;
    vsync_us = c.vsync_us;
    num_pulses = c.num_pulses;
    hsync_tol = c.hsync_tolerance_us;
    eq_tol = c.eq_tolerance_us;
    nominal = us(line_us);
  }
  double us(double u) const { return u * fs_mhz; }

  // EstimateTiming (:131-166), preserveVhsEmptyHSyncMedianUnits = true for vhs.
  Timing estimate_timing(const vector<Pulse>& raw, bool preserve_units = true) const {
    const double lo = us(hsync_us - 1.75), hi = us(hsync_us + 2.0);
    vector<double> lens;
    for (const Pulse& p : raw)
      if (in_range(p.length, lo, hi)) lens.push_back(p.length);
    const double med = !lens.empty() ? median_f64(lens) : (preserve_units ? hsync_us : us(hsync_us));
    const double typical = us(hsync_us);
    const double off = med - typical;
    Timing t;
    t.nominal = nominal;
    t.hsync_median = med;
    t.offset = off;
    t.hsync = {med - us(hsync_tol), med + us(hsync_tol)};
    t.eq = {us(eq_us - eq_tol) + off, us(eq_us + eq_tol) + off};
    t.vsync = {us(vsync_us * 0.5) + off, us(vsync_us + 1.0) + off};
    return t;
  }

  static int classify(const Pulse& p, const Timing& t) {
    if (t.hsync.contains(p.length)) return kHSync;
    if (t.eq.contains(p.length)) return kEq;
    return t.vsync.contains(p.length) ? kVSync : -1;
  }

  // PulseQualityCheck (:712-722)
  bool quality(const CPulse& prev, Kind kind, const Pulse& p) const {
    double lo, hi;
    if (prev.kind != kHSync && kind != kHSync) {
      lo = 0.4;
      hi = 0.6;
    } else if (prev.kind == kHSync && kind == kHSync) {
      lo = 0.9;
      hi = 1.1;
    } else {
      lo = 0.4;
      hi = 1.1;
    }
    const double ll = (p.start - prev.p.start) / nominal;
    return in_range(ll, lo, hi);
  }

  // TryRescueLongHSync (:316-338)
  bool rescue(const Pulse& p, const double* ref, int64_t n, double step, Pulse* out) const {
    const int64_t s = std::clamp<int64_t>(p.start, 0, n);
    const int64_t e = std::clamp<int64_t>((int64_t)p.start + p.length, s, n);
    if (s >= e) return false;
    const double thr = ref[s] - step;
    const vector<Pulse> c = find_pulses(ref + s, e - s, thr, std::max(0, (int)std::ceil(us(eq_us) / 8.0)),
                                        std::max(1, (int)(nominal * 5.0)));
    if (c.empty()) return false;
    *out = {(int)(s + c[0].start), c[0].length};
    return true;
  }

  // TryReadVBlank (:340-471)
  bool read_vblank(const vector<Pulse>& raw, const Timing& t, int start, int end_ex, vector<CPulse>& dst,
                   int* count) const {
    vector<CPulse> valid;
    int state = -1;
    double state_end = 0.0;
    bool has_len = false;
    double state_len = 0.0;
    for (int i = std::max(0, start); i < end_ex; i++) {
      const Pulse& p = raw[i];
      const int rk = classify(p, t);
      if (rk < 0) continue;
      int acc = -1;
      bool done = false;
      switch (state) {
        case -1:
          if (rk == kHSync) acc = kHSync;
          break;
        case kHSync:
          if (rk == kHSync) acc = kHSync;
          else if (rk == kEq) {
            acc = kEq;
            has_len = true;
            state_len = num_pulses / 2.0;
          } else if (rk == kVSync)
            acc = kVSync;
          break;
        case kEq:
          if (rk == kEq) acc = kEq;
          else if (rk == kVSync) {
            acc = kVSync;
            has_len = true;
            state_len = num_pulses / 2.0;
          } else if (rk == kHSync)
            acc = kHSync;
          break;
        case kVSync:
          if (rk == kEq) {
            acc = kEq2;
            has_len = true;
            state_len = num_pulses / 2.0;
          } else if (rk == kVSync)
            acc = kVSync;
          else if (rk == kHSync && p.start > state_end)
            acc = kHSync;
          break;
        case kEq2:
          if (rk == kEq) acc = kEq2;
          else if (rk == kHSync) {
            acc = kHSync;
            done = true;
          }
          break;
      }
      if (acc < 0) continue;
      if (acc != state) {
        if (p.start < state_end) {
          acc = -1;
        } else if (has_len) {
          state_end = p.start + ((state_len - 0.1) * nominal);
          has_len = false;
        }
      }
      if (acc >= 0) {
        const bool io = !valid.empty() && quality(valid.back(), (Kind)acc, p);
        valid.push_back({(Kind)acc, p, io});
        state = acc;
      }
      if (done) {
        for (size_t k = 2; k < valid.size(); k++) dst.push_back(valid[k]);
        *count = (int)valid.size();
        return true;
      }
    }
    *count = 0;
    return false;
  }

  // RefinePulses (:243-314); `raw` is updated in place by the long-hsync rescue.
  vector<CPulse> refine(vector<Pulse>& raw, const Timing& t, const double* ref, int64_t n, double step) const {
    vector<CPulse> dst;
    int index = 0;
    const int count = (int)raw.size();
    while (index < count) {
      const Pulse p = raw[index];
      if (t.hsync.contains(p.length)) {
        const bool io = !dst.empty() && quality(dst.back(), kHSync, p);
        dst.push_back({kHSync, p, io});
        index++;
        continue;
      }
      if (n > 0 && step > 0.0 && in_range(p.length, t.hsync.hi, t.hsync.hi * 3.0)) {
        Pulse r;
        if (rescue(p, ref, n, step, &r) && (r.start != p.start || r.length != p.length)) {
          raw[index] = r;
          continue;
        }
        index++;
        continue;
      }
      const bool can_start = index > 2 && t.eq.contains(p.length) && !dst.empty() && dst.back().kind == kHSync;
      if (!can_start) {
        index++;
        continue;
      }
      int vb = 0;
      if (!read_vblank(raw, t, index - 2, std::min(count, index + 24), dst, &vb)) {
        index++;
        continue;
      }
      index += std::max(1, vb - 2);
    }
    return dst;
  }

  // ComputeMeanLineLength (:490-538)
  double mean_line_length(const vector<CPulse>& v) const {
    int run_start = -1, run_len = -1, cur_start = -1, cur_len = 0;
    for (int i = 0; i < (int)v.size(); i++) {
      if (v[i].kind != kHSync) {
        if (cur_start >= 0 && cur_len > run_len) {
          run_start = cur_start;
          run_len = cur_len;
        }
        cur_start = -1;
        cur_len = 0;
      } else if (cur_start < 0) {
        cur_start = i;
        cur_len = 0;
      } else {
        cur_len++;
      }
    }
    if (cur_start >= 0 && cur_len > run_len) {
      run_start = cur_start;
      run_len = cur_len;
    }
    vector<double> lens;
    for (int i = run_start + 1; i < run_start + run_len; i++) {
      const double ll = (double)(v[i].p.start - v[i - 1].p.start);
      if (in_range(ll / nominal, 0.95, 1.05)) lens.push_back(ll);
    }
    if (lens.empty()) return nominal;
    double s = 0.0;
    for (double x : lens) s += x;
    return s / (double)lens.size();
  }

  // BuildUpstreamLineLocations (:585-657)
  vector<double> upstream_lines(const vector<CPulse>& v, double ref_pulse, int ref_line, double mll,
                                int lines) const {
    vector<double> locs;
    for (const CPulse& c : v) locs.push_back((double)c.p.start);
    std::stable_sort(locs.begin(), locs.end());
    vector<double> out(lines);
    int cur = 0;
    const double maxd = mll / 1.5;
    for (int line = 0; line < lines; line++) {
      const double expected = ref_pulse + (mll * (line - ref_line));
      out[line] = expected;
      if (cur >= (int)locs.size()) continue;
      double cd = std::fabs(locs[cur] - expected);
      double smallest = maxd;
      double sel = -1.0;
      int si = cur;
      while (si < (int)locs.size() - 1) {
        if (cd <= smallest) {
          smallest = cd;
          cur = si;
          sel = locs[si];
        }
        const double nd = std::fabs(locs[si + 1] - expected);
        if (nd > cd) break;
        cd = nd;
        si++;
      }
      if (sel >= 0.0) {
        out[line] = sel;
        cur++;
      }
    }
    return out;
  }
};

// FieldParityDetector.MeasureVBlankBoundaryConsensus / ResolveCadence (Dsp/FieldParityDetector.cs:16-268).
struct Boundary {
  int consensus, detected, first_conf, second_conf;
};

double round_half_line(double v) {
  const double h = round_away(v / 0.5) * 0.5;
  return round_away(h * 10.0) / 10.0;
}

Boundary boundary_consensus(const vector<CPulse>& v, double mll, bool pal, const int field_lines[2]) {
  Boundary r{0, 0, 0, 0};
  if (v.empty()) return r;
  double lens[4] = {-1.0, -1.0, -1.0, -1.0};
  int group = 0, last = -1;
  const double second = v[0].p.start + (field_lines[0] * mll);
  for (int i = 0; i < (int)v.size(); i++) {
    const CPulse& cur = v[i];
    if (last >= 0 && cur.in_order) {
      const CPulse& prev = v[last];
      if (group == 0 && cur.p.start > second) group = 2;
      if (prev.kind == kHSync && cur.kind != kHSync)
        lens[group] = round_half_line((cur.p.start - prev.p.start) / mll);
      else if (prev.kind != kHSync && cur.kind == kHSync)
        lens[group + 1] = round_half_line((cur.p.start - prev.p.start) / mll);
    }
    last = i;
  }
  const double ff_ntsc[4] = {1.0, 0.5, 0.5, 1.0}, sf_ntsc[4] = {0.5, 1.0, 1.0, 0.5};
  const double ff_pal[4] = {0.5, 0.5, 1.0, 1.0}, sf_pal[4] = {1.0, 1.0, 0.5, 0.5};
  const double* ff = pal ? ff_pal : ff_ntsc;
  const double* sf = pal ? sf_pal : sf_ntsc;
  int consensus = 0, detected = 0;
  for (int i = 0; i < 4; i++) {
    if (lens[i] < 0.0) continue;
    if (std::fabs(lens[i] - ff[i]) < 1e-9) {
      consensus++;
      detected++;
    } else if (std::fabs(lens[i] - sf[i]) < 1e-9) {
      detected++;
    }
  }
  if (detected == 0) return r;
  const double w = detected / 4.0;
  r.consensus = consensus;
  r.detected = detected;
  r.first_conf = (int)round_away((consensus / (double)detected) * w * 100.0);
  r.second_conf = (int)round_away(((detected - consensus) / (double)detected) * w * 100.0);
  return r;
}

struct Parity {
  bool first;
  int confidence;
};

Parity resolve_cadence(const vector<CPulse>& v, double mll, bool pal, const int field_lines[2], int prev_first,
                       int min_conf, bool has_prev_hsync, bool has_fallback_line = false, int fallback_first = -1,
                       int fallback_conf = -1) {
  const Boundary b = boundary_consensus(v, mll, pal, field_lines);
  bool first;
  int conf = 0;
  if (prev_first >= 0) {
    first = !prev_first;
  } else {
    first = b.detected == 0 || round_away(b.consensus / (double)b.detected) == 1.0 || fallback_first == 1;
  }
  int eff = min_conf;
  if (b.detected > 0 && !has_fallback_line && !has_prev_hsync && eff > 50) eff = 50;
  if (b.first_conf >= eff && b.first_conf > b.second_conf) {
    first = true;
    conf = b.first_conf;
  } else if (b.second_conf >= eff && b.first_conf < b.second_conf) {
    first = false;
    conf = b.second_conf;
  }
  if ((has_fallback_line || fallback_first >= 0) && fallback_conf > b.first_conf && fallback_conf > b.second_conf) {
    first = fallback_first == 1;
    conf = fallback_conf;
  }
  return {first, conf};
}

// VBlankSyncResolver boundary lengths and FirstHSyncLine (Dsp/VBlankSyncResolver.cs:447-465, 734-746).
struct FieldBoundaries {
  double first_hsync, first_eq2, last_hsync, last_eq2;
};

FieldBoundaries field_boundaries(bool pal, bool first) {
  if (!pal) return first ? FieldBoundaries{1.0, 0.5, 0.5, 1.0} : FieldBoundaries{0.5, 1.0, 1.0, 0.5};
  return first ? FieldBoundaries{0.5, 0.5, 1.0, 1.0} : FieldBoundaries{1.0, 1.0, 0.5, 0.5};
}

double first_hsync_line(bool pal, int num_pulses, bool first) {
  const FieldBoundaries b = field_boundaries(pal, first);
  const double section = num_pulses / 2.0;
  return b.first_hsync + (section * 3.0) - 0.5 + b.first_eq2;
}

struct Marker {
  double sample, line;
};
struct Acc {
  double loc_sum = 0.0, off_sum = 0.0;
  int count = 0;
  Acc operator+(const Acc& o) const {
    Acc r;
    r.loc_sum = loc_sum + o.loc_sum;
    r.off_sum = off_sum + o.off_sum;
    r.count = count + o.count;
    return r;
  }
};
struct Estimate {
  bool ok = false;
  double line0, first_hsync, first_hsync_line;
  int valid_count;
  double unaligned;
};

Acc accumulate_pair(const Marker& f, const Marker& s, double mll, double fhl) {
  const double actual = (f.sample - s.sample) / mll;
  const double expected = f.line - s.line;
  if (actual <= expected - 0.5 || actual >= expected + 0.5) return Acc();
  Acc a;
  a.loc_sum = s.sample + (mll * (fhl - s.line));
  a.off_sum = actual - expected;
  a.count = 1;
  return a;
}

vector<Marker> partial_markers(const double* samples, const bool* has, double eq1_line, double section) {
  const double lines[4] = {eq1_line, eq1_line + section, eq1_line + (section * 2.0), eq1_line + (section * 3.0) - 0.5};
  vector<Marker> m;
  for (int i = 0; i < 4; i++)
    if (has[i]) m.push_back({samples[i], lines[i]});
  return m;
}

Acc accumulate_within(const vector<Marker>& m, double mll, double fhl) {
  Acc a;
  for (size_t i = 0; i < m.size(); i++)
    for (size_t j = i + 1; j < m.size(); j++) a = a + accumulate_pair(m[i], m[j], mll, fhl);
  return a;
}

// FinalizeEstimate (:666-718)
Estimate finalize(const vector<CPulse>& v, const Acc& a, double mll, double fhl, int field_lines) {
  Estimate e;
  if (a.count == 0) return e;
  double fh = round_away((a.loc_sum + a.off_sum) / a.count);
  const double unaligned = fh;
  double off = 0.0;
  int cnt = 0;
  for (const CPulse& c : v) {
    if (c.kind != kHSync || !c.in_order) continue;
    const double line = ((c.p.start - fh) / mll) + fhl;
    const int rl = (int)round_away(line);
    if (rl > field_lines) break;
    if (rl >= fhl) {
      off += fh + (mll * (rl - fhl)) - c.p.start;
      cnt++;
    }
  }
  if (cnt > 0) fh -= off / cnt;
  e.ok = true;
  e.line0 = fh - (mll * fhl);
  e.first_hsync = fh;
  e.first_hsync_line = fhl;
  e.valid_count = a.count;
  e.unaligned = unaligned;
  return e;
}

struct Consensus {
  Estimate first, last, combined;
};

// EstimateLine0FromTransitions (:317-445)
Consensus line0_from_transitions(const vector<CPulse>& v, double mll, bool pal, int num_pulses, bool first_field,
                                 int cur_lines, int first_lines) {
  Consensus r;
  if (v.empty()) return r;
  double fs[4], ls[4];
  bool fh[4] = {false, false, false, false}, lh[4] = {false, false, false, false};
  double* as = fs;
  bool* ah = fh;
  int last = -1;
  for (int i = 0; i < (int)v.size(); i++) {
    const CPulse& cur = v[i];
    if (last >= 0 && cur.in_order) {
      if (as == fs && cur.p.start > v[0].p.start + (first_lines * mll)) {
        as = ls;
        ah = lh;
      }
      const CPulse& prev = v[last];
      if (prev.kind == kHSync && cur.kind != kHSync) {
        as[0] = cur.p.start;
        ah[0] = true;
      } else if (prev.kind == kEq && cur.kind == kVSync) {
        as[1] = cur.p.start;
        ah[1] = true;
      } else if (prev.kind == kVSync && cur.kind == kEq2) {
        as[2] = cur.p.start;
        ah[2] = true;
      } else if (prev.kind != kHSync && cur.kind == kHSync) {
        as[3] = prev.p.start;
        ah[3] = true;
      }
    }
    last = i;
  }
  const FieldBoundaries b = field_boundaries(pal, first_field);
  const double section = num_pulses / 2.0;
  const double fhl = b.first_hsync + (section * 3.0) - 0.5 + b.first_eq2;
  const vector<Marker> fm = partial_markers(fs, fh, b.first_hsync, section);
  const vector<Marker> lm = partial_markers(ls, lh, cur_lines + b.last_hsync, section);
  const Acc fa = accumulate_within(fm, mll, fhl);
  const Acc la = accumulate_within(lm, mll, fhl);
  r.first = finalize(v, fa, mll, fhl, cur_lines);
  r.last = finalize(v, la, mll, fhl, cur_lines);
  if (fa.count > 0 && la.count > 0) {
    const double fr = fa.loc_sum / fa.count;
    const double lr = la.loc_sum / la.count;
    if (std::fabs(fr - lr) < 0.5 * mll) {
      Acc cross;
      for (const Marker& a : fm)
        for (const Marker& c : lm) cross = cross + accumulate_pair(a, c, mll, fhl);
      r.combined = finalize(v, fa + la + cross, mll, fhl, cur_lines);
    }
  }
  return r;
}

// AlignFirstHSyncToValidPulses (TbcFieldDecodePipeline.cs:6207-6242)
double align_first_hsync(const vector<CPulse>& v, double fh, double mll, double fhl, int cur_lines) {
  double off = 0.0;
  int cnt = 0;
  for (const CPulse& c : v) {
    if (c.kind != kHSync || !c.in_order) continue;
    const double line = ((c.p.start - fh) / mll) + fhl;
    const int rl = (int)round_away(line);
    if (rl > cur_lines) break;
    if (rl >= fhl) {
      off += fh + (mll * (rl - fhl)) - c.p.start;
      cnt++;
    }
  }
  return cnt > 0 ? fh - (off / cnt) : fh;
}

// MovingAverageWindow of VsyncSerrationDetector (:727-754) and VhsFieldLevelState (Dsp/VhsFieldLevelState.cs:50-82).
struct MovingAverage {
  int window;
  int watermark;  // HasValues: count > watermark
  vector<double> values;
  bool has_values() const { return (int)values.size() > watermark; }
  void push(double v, bool bounded) {
    if (bounded && (int)values.size() == window) values.erase(values.begin());
    values.push_back(v);
  }
  bool pull(double* out) {
    if (values.empty()) return false;
    if ((int)values.size() >= window) values.erase(values.begin(), values.end() - window);
    double s = 0.0;
    for (double x : values) s += x;
    *out = s / (double)values.size();
    return true;
  }
};

struct Levels {
  double sync, blank;
};

enum Failure { kNone = 0, kMissing = 1, kNonFinite = 2, kCheckFailed = 3 };

// ------------------------------------------------------------------ the decoder
struct Decoder {
  OracleSyncConfig cfg;
  Analyzer an;
  bool pal;
  int field_lines[2];  // FieldLinesForParityDetection
  // VsyncSerrationDetector
  int divisor;
  double srate;
  int eq_len, vsync_len, line_len;
  double vbi_min, vbi_max;
  Tf tf[4];
  Sos sos[4];
  MovingAverage ser_sync{2, 1, {}}, ser_blank{2, 1, {}};
  int ser_field_count = 0;
  // VhsFieldLevelState
  MovingAverage fs_sync{10, 0, {}}, fs_blank{10, 0, {}};
  bool has_last_levels = false;
  Levels last_levels{0, 0};
  // sync history
  int prev_first = -1;
  bool has_prev_first_hsync = false;
  double prev_first_hsync = 0.0;
  int64_t prev_first_hsync_read = 0;
  double prev_hsync_difference = -1.0;
  bool pulse_search_initialized = false;

  void init(const OracleSyncConfig& c) {
    cfg = c;
    an.init(c);
    pal = c.is_pal != 0;
    if (!pal) {
      field_lines[0] = c.output_lines;
      field_lines[1] = c.output_lines - 1;
    } else {
      field_lines[0] = c.output_lines - 1;
      field_lines[1] = c.output_lines;
    }
    divisor = c.level_divisor > 0 ? c.level_divisor : 1;
    srate = c.sample_rate_hz / divisor;
    const double vf = c.fps * 2.0, hf = c.fps * c.frame_lines;
    eq_len = (int)round_even(srate * c.eq_us * 1e-6);
    vsync_len = (int)round_even(srate / vf);
    line_len = (int)round_even(srate / hf);
    const double vbi = 6.5 * line_len;
    vbi_min = vbi * 0.75;
    vbi_max = vbi * 1.25;
    for (int k = 0; k < 4; k++) {
      const int o = c.tf_order[k];
      tf[k].b.assign(c.tf_b[k], c.tf_b[k] + o + 1);
      tf[k].a.assign(c.tf_a[k], c.tf_a[k] + o + 1);
      sos[k].c.assign(&c.sos[k][0][0], &c.sos[k][0][0] + 6 * c.sos_sections[k]);
      sos[k].pad = 3 * (o + 1);
    }
    const double fps = c.fps;
    const double field_rate = fps * 2.0;
    const int w = (int)round_even(field_rate < 60.0 ? field_rate / 5.0 : field_rate / 6.0);
    fs_sync.window = fs_blank.window = w;
  }

  double ire_to_hz(double ire) const { return cfg.ire0 + (cfg.hz_ire * ire); }

  void filt(int k, double* x, int64_t n) const {
    if (cfg.serration_mode == 1) sos_filtfilt(sos[k], x, n);
    else tf_filtfilt(tf[k], x, n);
  }

  bool ser_pull(Levels* out) {
    double s = 0.0, b = 0.0;
    const bool hs = ser_sync.pull(&s);
    const bool hb = ser_blank.pull(&b);
    if (hs && hb) {
      *out = {s, b};
      return true;
    }
    return false;
  }
  void ser_push(double s, double b) {
    ser_sync.push(s, false);
    ser_blank.push(b, false);
  }
  bool ser_has_levels() const { return ser_sync.has_values() && ser_blank.has_values(); }

  bool fs_has_levels() const { return fs_sync.has_values() && fs_blank.has_values(); }
  void fs_push_sync(double s) { fs_sync.push(s, true); }
  void fs_push(double s, double b) {
    fs_blank.push(b, true);
    fs_push_sync(s);
  }
  bool fs_pull_sync(double* s) { return fs_sync.pull(s); }
  bool fs_pull(Levels* out) {
    double b;
    if (!fs_blank.pull(&b)) return false;
    double s;
    if (!fs_pull_sync(&s)) return false;
    *out = {s, b};
    return true;
  }

  // VsyncSerrationDetector.LocalMinimaIndices (:261-277)
  static vector<int> local_minima(const double* d, int64_t n) {
    vector<int> m;
    if (n < 3) return m;
    for (int64_t i = 1; i < n - 1; i++)
      if (d[i] < d[i - 1] && d[i] <= d[i + 1]) m.push_back((int)i);
    return m;
  }

  // ArbitrateVsync (:279-326)
  static vector<int> arbitrate(int vl, const vector<int>& env, const vector<int>& ser, int64_t data_len) {
    vector<int> out;
    if (env.size() > 1) {
      vector<int> valid;
      for (size_t i = 0; i < ser.size(); i++) {
        const int edge = ser[i];
        const int next = ser[std::min(i + 1, ser.size() - 1)];
        for (int m : env)
          if (edge <= m && m <= next) valid.push_back(edge);
      }
      for (int s : valid)
        if (s - vl >= 0 || s + vl <= data_len - 1) out.push_back(s);
      return out;
    }
    if (env.size() == 1) {
      const int m = env[0];
      if (m + vl < data_len - 1) return {m, m + vl};
      return {m, std::max(m - vl, 0)};
    }
    return out;
  }

  static int py_sign(double v) { return v > 0.0 ? 1 : (v < 0.0 ? -1 : 0); }

  // GetSerrationSyncLevels (:449-473)
  static bool serration_levels(const double* s, int64_t n, Levels* out) {
    const double half = mean_f64_fastmath(s, n);
    vector<double> peaks, valleys;
    for (int64_t i = 0; i < n; i++) {
      if (s[i] > half) peaks.push_back(s[i]);
      else valleys.push_back(s[i]);
    }
    if (peaks.empty() || valleys.empty()) return false;
    *out = {median_f64(valleys), median_f64(peaks)};
    return true;
  }

  // TryMeasureSerration (:328-447)
  bool measure_serration(const double* data, int64_t n, int pos, Levels* out) const {
    if (n <= 0 || line_len <= 0 || eq_len <= 0) return false;
    const int64_t start = std::max<int64_t>(0, (int64_t)pos - (int64_t)line_len * 30);
    const int64_t end = std::min<int64_t>(n - 1, (int64_t)pos + (int64_t)line_len * 30);
    if (end <= start) return false;
    const double* block = data + start;
    const int64_t bn = end - start;
    const double mn = minimum_f64(block, bn);
    const double med = median_f64(block, bn);
    const double level = ((med - mn) / 2.0) + mn;
    vector<int> cross;
    int ps = py_sign(block[0] - level);
    for (int64_t i = 1; i < bn; i++) {
      const int s = py_sign(block[i] - level);
      if (s != ps) cross.push_back((int)(i - 1));
      ps = s;
    }
    vector<int> vd;
    for (int i = 0; i + 1 < (int)cross.size(); i++) {
      const int d = cross[i + 1] - cross[i];
      if ((eq_len * 0.2) < d && d < (eq_len * 5.0 / 4.0)) vd.push_back(i);
    }
    if (vd.size() < 9 || vd.size() > 12) return false;
    const int64_t ss = cross[vd[0]] + start;
    const int64_t se = std::min<int64_t>((int64_t)(cross[vd.back()] + (eq_len / 2.0)) + start, n - 1);
    const int64_t len = se - ss;
    if (!(vbi_min < len && len < vbi_max)) return false;
    return serration_levels(data + ss, se - ss, out);
  }

  struct SerrationResult {
    bool found, has_levels, levels_ok;
    Levels levels;
    int candidates, measurements;
    vector<int> env_minima, harmonic_minima, candidate_list;
    vector<Levels> measured;
  };

  // VsyncSerrationDetector.Analyze (:145-242)
  SerrationResult analyze(const double* data, int64_t n) {
    SerrationResult r{false, false, false, {0, 0}, 0, 0, {}, {}, {}, {}};
    const int64_t rl = (n + divisor - 1) / divisor;
    vector<double> reduced(rl);
    for (int64_t i = 0, s = 0; i < rl; i++, s += divisor) reduced[i] = data[s];
    const int64_t pad = std::min<int64_t>(1024, rl);
    const int64_t pl = rl + pad;
    int min_filtered = 0;
    for (int k = 0; k < 4; k++) min_filtered = std::max(min_filtered, 3 * (int)tf[k].a.size());
    min_filtered += 1;
    if (cfg.serration_mode == 2 || pl < min_filtered) {
      ser_field_count++;
      r.levels_ok = ser_pull(&r.levels);
      r.has_levels = ser_has_levels();
      return r;
    }
    vector<double> padded(pl);
    for (int64_t i = 0; i < pad; i++) padded[i] = reduced[pad - 1 - i];
    for (int64_t i = 0; i < rl; i++) padded[pad + i] = reduced[i];
    // BuildVsyncEnvelope (:538-566)
    vector<double> fwd(pl);
    for (int64_t i = 0; i < pl; i++) fwd[i] = std::max(0.0, padded[i]);
    double bias = fwd[0];
    for (int64_t i = 1; i < pl; i++) bias = std::min(bias, fwd[i]);
    vector<double> rev(fwd.rbegin(), fwd.rend());
    filt(0, fwd.data(), pl);
    filt(0, rev.data(), pl);
    std::reverse(rev.begin(), rev.end());
    const int64_t half = pl / 2;
    vector<double> env(rl);
    for (int64_t i = 0; i < rl; i++) {
      const int64_t src = pad + i;
      env[i] = (src < half ? rev[src] : fwd[src]) - bias;
    }
    const vector<int> env_min = local_minima(env.data(), rl);
    // PowerRatioSearch (:568-580)
    vector<double> h(padded);
    filt(1, h.data(), pl);
    filt(2, h.data(), pl);
    for (int64_t i = 0; i < pl; i++) h[i] *= h[i];
    filt(3, h.data(), pl);
    const vector<int> harm_min = local_minima(h.data(), pl);
    const vector<int> cands = arbitrate(vsync_len, env_min, harm_min, rl + std::min<int64_t>(1024, rl));
    r.candidates = (int)cands.size();
    r.env_minima = env_min;
    r.harmonic_minima = harm_min;
    r.candidate_list = cands;
    int meas = 0;
    for (int c : cands) {
      Levels m;
      if (measure_serration(reduced.data(), rl, c, &m)) {
        meas++;
        r.measured.push_back(m);
        ser_push(m.sync, m.blank);
      }
    }
    ser_field_count++;
    r.measurements = meas;
    r.found = meas > 0;
    r.levels_ok = ser_pull(&r.levels);
    r.has_levels = ser_has_levels();
    return r;
  }

  // VsyncSerrationDetector.CheckLevels (:475-536)
  static bool check_levels(const double* d, int64_t n, double old_sync, double new_sync, double new_blank,
                           double ref_sync, double hz_ire) {
    if (n <= 0 || hz_ire == 0.0) return false;
    const double diff = (new_blank - new_sync) / hz_ire;
    if ((ref_sync - new_sync) > (hz_ire * 15.0) || diff > 47.0) return false;
    if (new_sync - old_sync < (hz_ire * 5.0)) return true;
    int64_t below_sync = 0, below_blank = 0;
    for (int64_t i = 0; i < n; i++) {
      if (d[i] < new_sync) below_sync++;
      if (d[i] < new_blank) below_blank++;
    }
    const double a = (double)below_sync / n;
    const double b = (double)below_blank / n;
    if (a > 0.07 || b < 0.005) return false;
    return true;
  }

  // LevelDetection.RefineSerrationLevelsFromPulses (Dsp/LevelDetection.cs:172-235) with
  // FallbackVsyncLocationMeans (:22-53) and PulsesBlackLevelMeans (:445-524).
  bool refine_from_pulses(const double* d, int64_t n, const vector<Pulse>& pulses, double ref_sync, Levels* out,
                          Failure* fail, bool* has_measured, double* measured) const {
    *fail = kMissing;
    *has_measured = false;
    const double vl = an.us(an.vsync_us);
    const double lo = vl * 0.8, hi = vl * 1.2;
    const int off = (int)an.fs_mhz;
    vector<int> locs;
    vector<double> means;
    for (int i = 0; i < (int)pulses.size(); i++) {
      const Pulse& p = pulses[i];
      if (p.length <= lo || p.length >= hi) continue;
      const int64_t s = (int64_t)p.start + off;
      const int64_t e = (int64_t)p.start + p.length - off;
      if (s < 0 || e > n || e <= s) continue;  // the reference throws here
      locs.push_back(i);
      means.push_back(mean_f64(d + s, e - s));
    }
    if (means.empty()) return false;
    const double sync = median_f64(means);
    *has_measured = true;
    *measured = sync;
    // PulsesBlackLevelMeans
    vector<double> black;
    if (!locs.empty() && !pulses.empty()) {
      int before = locs[0];
      int after = locs.back();
      const int last_index = (int)pulses.size() - 1;
      if (locs.size() != 12) {
        while (before > 1 && pulses[before].start - pulses[before - 1].start < 600) before--;
        while (after < last_index && pulses[after].start - pulses[after + 1].start < 600) after++;
      }
      vector<int> idx;
      if (before > 1)
        for (int i = std::max(before - 5, 1); i < before; i++) idx.push_back(i);
      if (after < last_index - 1)
        for (int i = after + 1; i < std::max(after + 6, last_index); i++) idx.push_back(i);
      for (int i : idx) {
        if (i < 0 || i >= (int)pulses.size()) continue;
        const Pulse& p = pulses[i];
        if (!in_range(p.length, an.fs_mhz * 0.75, an.fs_mhz * 3.0)) continue;
        const int64_t s = std::clamp<int64_t>((int64_t)(p.start + (an.fs_mhz * 5.0)), 0, n);
        const int64_t e = std::clamp<int64_t>((int64_t)(p.start + (an.fs_mhz * 20.0)), s, n);
        if (e > s) black.push_back(mean_f64(d + s, e - s));
      }
    }
    if (black.empty()) {
      *fail = kNonFinite;
      return false;
    }
    const double blank = median_f64(black);
    if (!std::isfinite(sync) || !std::isfinite(blank) || blank < sync) {
      *fail = kNonFinite;
      return false;
    }
    if (means.size() <= 3 || !check_levels(d, n, ref_sync, sync, blank, ref_sync, cfg.hz_ire)) {
      *fail = kCheckFailed;
      return false;
    }
    *fail = kNone;
    *out = {sync, blank};
    return true;
  }

  // LevelDetection.RefineSerrationLevels (:98-170)
  bool refine_levels(const double* d, int64_t n, double s0, double b0, double ref_sync, Levels* out, Failure* fail,
                     bool* has_measured, double* measured) const {
    *fail = kMissing;
    *has_measured = false;
    if (n <= 0) return false;
    const double thr = (s0 + b0) / 2.0;
    const double eql = (double)(eq_len * divisor);
    const double ll = (double)(line_len * divisor);
    const int mn = std::max(0, (int)std::ceil(eql / 8.0));
    const int mx = std::max(1, (int)std::floor(ll * 5.0));
    const vector<Pulse> p = find_pulses(d, n, thr, mn, mx);
    return refine_from_pulses(d, n, p, ref_sync, out, fail, has_measured, measured);
  }

  // LevelDetection.SearchFallbackSerrationLevels (:265-396)
  bool search_fallback(const double* d, int64_t n, double blank_level, double ref_sync, bool check_long, Levels* out,
                       Failure* fail, bool* has_measured, double* measured) const {
    *fail = kMissing;
    *has_measured = false;
    if (n <= 0 || cfg.hz_ire == 0.0) return false;
    double min_sync = d[0];
    for (int64_t i = 1; i < n; i++) min_sync = std::min(min_sync, d[i]);
    const double eql = (double)(eq_len * divisor);
    const double ll = (double)(line_len * divisor);
    const double min_vsync = an.us(an.vsync_us) * 0.8;
    const double min_long = an.us(an.vsync_us) * 2.6;
    const double max_long = ll * 5.0;
    int prev_vsync = 0, prev_long = 0;
    double prev_min_sync = min_sync;
    bool found = false, check_next = true;
    vector<Pulse> pulses;
    auto find_reduced = [&](double thr, int div) {
      const int mn = std::max(0, (int)std::ceil((eql / 8.0) / div));
      const int mx = std::max(1, (int)std::floor((ll * 5.0) / div));
      // divisor 1 only (the reduced data equals the input)
      pulses = find_pulses(d, n, thr, mn, mx);
    };
    for (int retry = 0; retry < 30; retry++) {
      const double thr = (min_sync + blank_level) / 2.0;
      find_reduced(thr, divisor);
      if (pulses.size() > 200) {
        int vc = 0, lc = 0;
        for (const Pulse& p : pulses)
          if (p.length > min_vsync) vc++;
        if (check_long && vc <= 2)
          for (const Pulse& p : pulses)
            if (in_range(p.length, min_long, max_long)) lc++;
        if (vc > 4 || lc >= 1) {
          if ((vc == 12 || lc == 2) && !check_next) break;
          if (!found || vc > prev_vsync || lc > prev_long) {
            found = true;
            prev_vsync = vc;
            prev_long = lc;
            prev_min_sync = min_sync;
            check_next = true;
          } else if (vc < prev_vsync || lc < prev_long || !check_next) {
            min_sync = prev_min_sync;
            find_reduced((min_sync + blank_level) / 2.0, 1);
            break;
          } else {
            check_next = false;
          }
        }
      }
      // AdvanceFallbackSyncSearchLevel (:398-407)
      const double sync_ire = (min_sync - blank_level) / cfg.hz_ire;
      const double next_ire = sync_ire + 5.0;
      const double scaled = cfg.hz_ire * next_ire;
      min_sync = blank_level + scaled;
    }
    return refine_from_pulses(d, n, pulses, ref_sync, out, fail, has_measured, measured);
  }

  // TryResolveVhsSerrationRefinementFallback (TbcFieldDecodePipeline.cs:5679-5704)
  bool try_fieldstate_fallback(Failure f, bool* has_levels, Levels* levels) {
    *has_levels = false;
    if (f != kNonFinite && f != kMissing) return false;
    double fs_sync_pulled = 0.0;
    bool has_sync = false;
    if (f == kMissing) has_sync = fs_pull_sync(&fs_sync_pulled);
    *has_levels = fs_pull(levels);
    if (f == kMissing && !has_sync) return false;
    return true;
  }

  // SelectUsableVhsLevels (:5706-5747)
  bool select_usable(const double* d, int64_t n, double ref_sync, bool has, Levels lv, Levels* out) {
    if (!has) return false;
    if (check_levels(d, n, ref_sync, lv.sync, lv.blank, ref_sync, cfg.hz_ire)) {
      *out = lv;
      return true;
    }
    return fs_pull(out);
  }

  struct Prepared {
    bool shifted;       // DC offset applied to the sync reference
    double dc;
    double threshold;
    int source;
    Levels levels;
    int candidates, measurements;
  };

  // PrepareSyncSpanFromLevels (:3197-3234) for vhs: DC-shift the low-pass sync reference
  // so the blank level lands on ire0; threshold halfway between the shifted levels.
  Prepared from_levels(Levels lv, int source, const SerrationResult& sr) const {
    Prepared p{false, 0.0, 0.0, source, lv, sr.candidates, sr.measurements};
    const double dc = cfg.ire0 - lv.blank;
    if (!std::isfinite(dc) || dc == 0.0) {
      p.threshold = (lv.sync + lv.blank) / 2.0;
      return p;
    }
    const double as = lv.sync + dc, ab = lv.blank + dc;
    p.shifted = true;
    p.dc = dc;
    p.threshold = (as + ab) / 2.0;
    return p;
  }

  // PrepareSyncSpan, vhs branch with DetectLevels (:2776-3190)
  Prepared prepare(const double* d, int64_t n) {
    const double default_threshold = ire_to_hz(cfg.vsync_ire / 2.0);
    const double ref_sync = ire_to_hz(cfg.vsync_ire);
    SerrationResult sr = analyze(d, n);
    Prepared def{false, 0.0, default_threshold, 0, {0, 0}, sr.candidates, sr.measurements};
    bool has_pending = false;
    Levels pending{0, 0};
    double ssync = sr.levels.sync, sblank = sr.levels.blank;
    bool s_ok = sr.levels_ok;
    if (sr.found && s_ok) {
      Levels rl;
      Failure f;
      bool hm;
      double ms;
      const bool ok = refine_levels(d, n, ssync, sblank, ref_sync, &rl, &f, &hm, &ms);
      if (hm) fs_push_sync(ms);
      if (ok) {
        fs_push(rl.sync, rl.blank);
        ser_push(rl.sync, rl.blank);
      } else {
        bool hl;
        Levels fl;
        try_fieldstate_fallback(f, &hl, &fl);
        if (hl) ser_push(fl.sync, fl.blank);
      }
      Levels avg{0.0, 0.0};
      s_ok = ser_pull(&avg);
      ssync = avg.sync;
      sblank = avg.blank;
    }
    if ((sr.found || sr.has_levels) && s_ok) {
      has_pending = true;
      pending = {ssync, sblank};
      if (sr.found) {
        Levels sel;
        if (select_usable(d, n, ref_sync, true, pending, &sel)) {
          last_levels = sel;
          has_last_levels = true;
          return from_levels(sel, 1, sr);
        }
        if (ser_has_levels() && !fs_has_levels()) return from_levels({cfg.ire0, ref_sync}, 4, sr);
        return def;
      }
    }
    // fallback search (vhs)
    Levels first;
    Failure f;
    bool hm;
    double ms;
    bool has_first = search_fallback(d, n, ire_to_hz(0.0), ref_sync, false, &first, &f, &hm, &ms);
    if (hm) fs_push_sync(ms);
    bool first_handled = false;
    if (!has_first) {
      bool hl;
      Levels fl;
      first_handled = try_fieldstate_fallback(f, &hl, &fl);
      if (first_handled) {
        has_first = hl;
        first = fl;
      }
    }
    bool pushed = false;
    if (has_first) {
      Levels sel;
      Failure f2;
      bool hm2;
      double ms2;
      const bool ok = refine_levels(d, n, first.sync, first.blank, ref_sync, &sel, &f2, &hm2, &ms2);
      if (hm2) fs_push_sync(ms2);
      if (ok) {
        fs_push(sel.sync, sel.blank);
        ser_push(sel.sync, sel.blank);
        pushed = true;
      } else {
        bool hl;
        Levels fl;
        if (try_fieldstate_fallback(f2, &hl, &fl)) {
          if (hl) {
            ser_push(fl.sync, fl.blank);
            pushed = true;
          }
        }
      }
    }
    if (pushed) {
      // SelectVhsLevelsAfterRefinement (:5762-5782)
      Levels sel;
      bool ok;
      if (ser_has_levels()) {
        Levels pulled;
        const bool hp = ser_pull(&pulled);
        ok = select_usable(d, n, ref_sync, hp, pulled, &sel);
      } else {
        ok = fs_pull(&sel);
      }
      if (ok) {
        last_levels = sel;
        has_last_levels = true;
        return from_levels(sel, 2, sr);
      }
    }
    if (has_pending) {
      Levels sel;
      if (select_usable(d, n, ref_sync, true, pending, &sel)) {
        last_levels = sel;
        has_last_levels = true;
        return from_levels(sel, 1, sr);
      }
      if (ser_has_levels() && !fs_has_levels()) return from_levels({cfg.ire0, ref_sync}, 4, sr);
    }
    // SelectTerminalVhsLevels (:5749-5760); fallbackToSavedLevels = true
    Levels term;
    bool has_term = false;
    if (fs_has_levels() && fs_pull(&term)) has_term = true;
    else if (has_last_levels) {
      term = last_levels;
      has_term = true;
    }
    if (has_term) {
      last_levels = term;
      has_last_levels = true;
      return from_levels(term, 3, sr);
    }
    return def;
  }

  // EstimateFirstHSyncFromPrevious (:6146-6173)
  bool estimate_first_hsync_prev(int64_t span_start, double mll, bool first, double* out) const {
    if (!has_prev_first_hsync || prev_first_hsync <= 0.0 || mll <= 0.0) return false;
    const double cur = field_lines[first ? 0 : 1];
    const double prev = field_lines[first ? 1 : 0];
    const double est_lines = pal ? cur : prev;
    const double last_off = (prev_first_hsync_read - span_start) / mll;
    *out = round_away((last_off + est_lines + (prev_first_hsync / mll)) * mll);
    return true;
  }

  struct Line0 {
    bool ok;
    double location, first_hsync, unaligned;
    bool used_previous;
  };

  // TryEstimateLine0FromPrevious, vhs (:6288-6375)
  Line0 line0_from_previous(const vector<CPulse>& v, int64_t span_start, double mll, int cur_lines,
                            bool first) const {
    Line0 r{false, 0, 0, 0, true};
    double est;
    if (!estimate_first_hsync_prev(span_start, mll, first, &est)) return r;
    const double fhl = first_hsync_line(pal, an.num_pulses, first);
    if (prev_hsync_difference >= -0.5 && prev_hsync_difference <= 0.5) est += mll * prev_hsync_difference;
    if (est <= 0.0) est = !v.empty() ? v[0].p.start : 0.0;
    const double unaligned = round_away(est);
    const double fh = align_first_hsync(v, unaligned, mll, fhl, cur_lines);
    r.ok = true;
    r.location = fh - (fhl * mll);
    r.first_hsync = fh;
    r.unaligned = unaligned;
    return r;
  }

  // ResolveLine0Location for vhs (:5452-5651): the transition consensus replaces the
  // vblank-group consensus; combined, then a single vblank estimate, then the
  // previous-field projection.
  Line0 resolve_line0(const vector<CPulse>& v, int64_t span_start, double mll, bool first) const {
    const int cur_lines = field_lines[first ? 0 : 1];
    const Consensus c = line0_from_transitions(v, mll, pal, an.num_pulses, first, cur_lines, field_lines[0]);
    const Line0 prev = line0_from_previous(v, span_start, mll, cur_lines, first);
    if (c.combined.ok) return {true, c.combined.line0, c.combined.first_hsync, c.combined.unaligned, false};
    // SelectSingleVBlankEstimate (:6010-6031)
    const int fc = c.first.ok ? c.first.valid_count : 0;
    const int lc = c.last.ok ? c.last.valid_count : 0;
    const bool has_prev = has_prev_first_hsync && prev_first_hsync > 0.0;
    const Estimate* single = nullptr;
    if (fc == 6 || (!has_prev && fc > lc)) single = c.first.ok ? &c.first : nullptr;
    else if (lc == 6 || (!has_prev && lc > fc)) single = c.last.ok ? &c.last : nullptr;
    if (single) return {true, single->line0, single->first_hsync, single->unaligned, false};
    if (prev.ok) return prev;
    return {false, 0, 0, 0, false};
  }

  // UpdateSyncHistory (:6119-6144)
  void update_history(int64_t span_start, const Line0& l0, double mll, bool first) {
    double est;
    if (!l0.used_previous && estimate_first_hsync_prev(span_start, mll, first, &est) && std::isfinite(l0.unaligned))
      prev_hsync_difference = (l0.unaligned - est) / mll;
    if (std::isfinite(l0.first_hsync)) {
      prev_first_hsync_read = span_start;
      prev_first_hsync = l0.first_hsync;
      has_prev_first_hsync = true;
    }
    prev_first = first ? 1 : 0;
  }

  // One field: DecodePrepared up to the line locations (:1071-1290) and the next read offset.
  // `d` is the window's FP64 sync reference; `shifted` receives it with the DC offset
  // applied (the reference the hsync refinement reads).
  int decode(const double* d, int64_t n, int64_t span_start, OracleSyncField* f, vector<double>* locs,
             vector<double>* shifted) {
    const Prepared pr = prepare(d, n);
    f->level_source = pr.source;
    f->sync_level = pr.levels.sync;
    f->blank_level = pr.levels.blank;
    f->dc_offset = pr.dc;
    f->threshold = pr.threshold;
    f->serration_candidates = pr.candidates;
    f->serration_measurements = pr.measurements;
    shifted->assign(d, d + n);
    if (pr.shifted)
      for (int64_t i = 0; i < n; i++) (*shifted)[i] = d[i] + pr.dc;
    const double* ref = shifted->data();
    const int mn = std::max(0, (int)std::ceil(an.us(an.eq_us) / 8.0));
    const int mx = std::max(1, (int)(an.nominal * 5.0));
    vector<Pulse> raw = find_pulses(ref, n, pr.threshold, mn, mx);
    f->raw_pulses = (int)raw.size();
    if (raw.empty()) {
      // BuildRecoveryException (:5936-5959), tape offsets
      const bool missing_start = pulse_search_initialized;
      f->status = missing_start ? 2 : 1;
      f->next_offset = missing_start ? an.nominal * 100.0 : an.fs_hz / 10.0;
      return f->status;
    }
    pulse_search_initialized = true;
    const Timing t = an.estimate_timing(raw);
    f->hsync_median = t.hsync_median;
    const vector<CPulse> refined = an.refine(raw, t, ref, n, 10.0 * cfg.hz_ire);
    f->refined_pulses = (int)refined.size();
    const double mll = an.mean_line_length(refined);
    f->mean_line_length = mll;
    const Parity par =
        resolve_cadence(refined, mll, pal, field_lines, prev_first, 100, has_prev_first_hsync);
    f->is_first_field = par.first ? 1 : 0;
    f->parity_confidence = par.confidence;
    prev_first = par.first ? 1 : 0;
    const Line0 l0 = resolve_line0(refined, span_start, mll, par.first);
    if (!l0.ok) {
      f->status = 2;
      f->next_offset = an.nominal * 100.0;
      return f->status;
    }
    update_history(span_start, l0, mll, par.first);
    const int cur_lines = field_lines[par.first ? 0 : 1];
    const int processed = cfg.output_lines + 10;
    f->processed_lines = processed;
    f->line0 = l0.location;
    f->first_hsync = l0.first_hsync;
    const double fhl = first_hsync_line(pal, an.num_pulses, par.first);
    f->first_hsync_line = fhl;
    // ThrowIfInsufficientFieldData (:5961-6008), tape
    const double last_line = (((double)n - l0.location) / mll) - 1.0;
    if (last_line < processed) {
      double suggested = l0.location - (mll * 20.0);
      suggested = std::max(suggested, an.nominal);
      f->status = 3;
      f->next_offset = suggested;
      return f->status;
    }
    *locs = an.upstream_lines(refined, std::trunc(l0.first_hsync), (int)fhl, mll, processed);
    // ComputeVhsNextFieldOffsetSamples (:2299-2325) via VBlankSyncResolver.NextFieldLocation
    const FieldBoundaries b = field_boundaries(pal, par.first);
    const double next_eq1 = cur_lines + b.last_hsync;
    const double next_loc = l0.first_hsync + (mll * (next_eq1 - fhl));
    f->next_offset = next_loc - (8.0 * an.nominal) + (double)cfg.block_cut;
    if (!(std::isfinite(f->next_offset) && f->next_offset > 0.0)) {
      // the line-location fallbacks of ComputeNextFieldOffsetSamples (:2286-2297); the
      // FindNextVBlankLine0 candidate before them is not restated (unreached when the
      // resolved first hsync is finite and past 8 lines)
      const int fb = std::max(0, cfg.output_lines - 7);
      if ((int)locs->size() > fb && std::isfinite((*locs)[fb]) && (*locs)[fb] > 0.0)
        f->next_offset = (*locs)[fb] + (double)cfg.block_cut;
      else
        f->next_offset = std::max(1.0, l0.location + (fb * mll) + (double)cfg.block_cut);
    }
    f->status = 0;
    return 0;
  }
};

// DecodeReadWindowPlanner.Resolve (Decode/DecodeReadWindowPlanner.cs:37-63)
void resolve_window(const OracleSyncConfig& c, int64_t begin, int64_t* start, int64_t* len) {
  const int64_t read = std::max<int64_t>(0, begin - c.block_cut);
  const int64_t a0 = (read / c.block_length) * c.block_length;
  const int64_t a1 = a0 + c.read_samples;
  const int64_t fb = a0 / c.block_stride;
  const int64_t lb = a1 / c.block_stride;
  *start = fb * c.block_stride;
  *len = (lb - fb + 1) * c.block_stride;
}

}  // namespace

extern "C" {

int oracle_sync_config_size() { return (int)sizeof(OracleSyncConfig); }
int oracle_sync_field_size() { return (int)sizeof(OracleSyncField); }

// The VHS field sequence over a demodulated span whose sample 0 is global sample 0 of the
// stride grid: fields are read from `first_begin` on, each from its read window
// (DecodeReadWindowPlanner), the next one at the window start plus the rounded next-field
// offset (TbcFieldSequenceDecodeEngine.cs:161-171, 1936-1941), recovery offsets added to
// the intended start as the sequence engine does (:846).  Stops when a window would run
// past `n` or after `max_fields` attempts.  `locs` receives [max_fields][max_lines]
// window-relative line locations of decoded fields; `shifted_out` (optional,
// [max_fields][shifted_stride]) receives each decoded field's DC-shifted sync reference
// (the data its hsync refinement reads).  Returns the number of records written.
int oracle_vhs_field_sequence(const double* vlp, int64_t n, const OracleSyncConfig* cfg, int64_t first_begin,
                              int max_fields, int max_lines, OracleSyncField* fields, double* locs,
                              double* shifted_out, int64_t shifted_stride) {
  Decoder dec;
  dec.init(*cfg);
  int64_t begin = first_begin;
  int out = 0;
  vector<double> l, sh;
  while (out < max_fields) {
    int64_t ws, wl;
    resolve_window(*cfg, begin, &ws, &wl);
    if (ws + wl > n) break;
    OracleSyncField f;
    std::memset(&f, 0, sizeof(f));
    f.begin = begin;
    f.window_start = ws;
    f.window_length = wl;
    l.clear();
    const int st = dec.decode(vlp + ws, wl, ws, &f, &l, &sh);
    if (st == 0) {
      const int nl = std::min<int>(max_lines, (int)l.size());
      for (int i = 0; i < nl; i++) locs[(int64_t)out * max_lines + i] = l[i];
      const int64_t advance = std::max<int64_t>(1, (int64_t)round_away(f.next_offset));
      f.next_begin = ws + advance;
      if (shifted_out && shifted_stride >= wl)
        std::memcpy(shifted_out + (int64_t)out * shifted_stride, sh.data(), sizeof(double) * wl);
    } else {
      f.next_begin = std::max<int64_t>(0, begin + (int64_t)f.next_offset);
    }
    fields[out++] = f;
    begin = f.next_begin;
  }
  return out;
}

// VsyncSerrationDetector.Analyze on a fresh detector (vectors of CompatibilityTests.cs:7138-7260):
// counts[0..3] = envelope minima, harmonic minima, candidates, measurements (each list written
// up to `cap`); levels[0..1] = the first measurement's sync and blank levels.
void oracle_serration_analyze(const double* data, int64_t n, const OracleSyncConfig* cfg, int* env_minima,
                              int* harmonic_minima, int* candidates, int cap, int* counts, double* levels) {
  Decoder dec;
  dec.init(*cfg);
  const Decoder::SerrationResult r = dec.analyze(data, n);
  const vector<int>* lists[3] = {&r.env_minima, &r.harmonic_minima, &r.candidate_list};
  int* outs[3] = {env_minima, harmonic_minima, candidates};
  for (int k = 0; k < 3; k++) {
    counts[k] = (int)lists[k]->size();
    for (int i = 0; i < (int)lists[k]->size() && i < cap; i++) outs[k][i] = (*lists[k])[i];
  }
  counts[3] = r.measurements;
  levels[0] = r.measured.empty() ? 0.0 : r.measured[0].sync;
  levels[1] = r.measured.empty() ? 0.0 : r.measured[0].blank;
}

// VsyncSerrationDetector.TryMeasureSerration (:328-447) with the detector geometry of `cfg`.
int oracle_serration_measure(const double* data, int64_t n, const OracleSyncConfig* cfg, int position,
                             double* levels) {
  Decoder dec;
  dec.init(*cfg);
  Levels m;
  if (!dec.measure_serration(data, n, position, &m)) return 0;
  levels[0] = m.sync;
  levels[1] = m.blank;
  return 1;
}

// Pieces for isolated checks against the reference's own test vectors.
int oracle_serration_levels(const double* s, int64_t n, double* sync, double* blank) {
  Levels lv;
  if (!Decoder::serration_levels(s, n, &lv)) return 0;
  *sync = lv.sync;
  *blank = lv.blank;
  return 1;
}

double oracle_mean_f64(const double* v, int64_t n) { return mean_f64(v, n); }
double oracle_mean_f64_fastmath(const double* v, int64_t n) { return mean_f64_fastmath(v, n); }
double oracle_median_f64(const double* v, int64_t n) { return median_f64(v, n); }

void oracle_filtfilt_tf(const double* b, const double* a, int order, double* x, int64_t n) {
  Tf f;
  f.b.assign(b, b + order + 1);
  f.a.assign(a, a + order + 1);
  tf_filtfilt(f, x, n);
}

void oracle_filtfilt_sos(const double* sos, int sections, int pad, double* x, int64_t n) {
  Sos f;
  f.c.assign(sos, sos + 6 * sections);
  f.pad = pad;
  sos_filtfilt(f, x, n);
}

// LevelDetection.RefineSerrationLevels with the analyzer of `cfg` (vectors of
// SyncLevelNumericsCompatibilityTests.cs:74-111).  Returns 1 on success.
int oracle_refine_serration_levels(const double* d, int64_t n, const OracleSyncConfig* cfg, double s0, double b0,
                                   double ref_sync, double hz_ire, double* sync, double* blank, int* vsync_count,
                                   int* pulse_count) {
  Decoder dec;
  dec.init(*cfg);
  dec.cfg.hz_ire = hz_ire;
  const double thr = (s0 + b0) / 2.0;
  const double eql = dec.an.us(dec.an.eq_us);
  const double ll = dec.an.nominal;
  const int mn = std::max(0, (int)std::ceil(eql / 8.0));
  const int mx = std::max(1, (int)std::floor(ll * 5.0));
  const vector<Pulse> p = find_pulses(d, n, thr, mn, mx);
  *pulse_count = (int)p.size();
  Levels lv;
  Failure f;
  bool hm;
  double ms;
  if (!dec.refine_from_pulses(d, n, p, ref_sync, &lv, &f, &hm, &ms)) return -(int)f;
  *sync = lv.sync;
  *blank = lv.blank;
  int vc = 0;
  const double vl = dec.an.us(dec.an.vsync_us);
  for (const Pulse& q : p)
    if (!(q.length <= vl * 0.8 || q.length >= vl * 1.2)) vc++;
  *vsync_count = vc;
  return 1;
}

// ---- pieces of SyncAnalyzer / FieldParityDetector / VBlankSyncResolver on caller data, for the
// reference's own vectors (CompatibilityTests.cs:8164-8338, 8339-8588, 8827-8918, 9146-9390).
// `analyzer` = {sample_rate_hz, line_period_us, hsync_us, eq_us, vsync_us, num_pulses,
// hsync_tolerance_us, eq_tolerance_us} as the SyncAnalyzer constructor takes them.

static vector<CPulse> make_pulses(const int* kinds, const int* starts, const int* lens, const int* in_order, int n) {
  vector<CPulse> v(n);
  for (int i = 0; i < n; i++) v[i] = {(Kind)kinds[i], {starts[i], lens[i]}, in_order[i] != 0};
  return v;
}

// timing_out: hsync_median, hsync lo/hi, eq lo/hi, vsync lo/hi, nominal
void oracle_sync_timing(const double* analyzer, const int* starts, const int* lens, int n, int preserve_units,
                        double* timing_out) {
  Analyzer an;
  an.init_direct(analyzer);
  vector<Pulse> raw(n);
  for (int i = 0; i < n; i++) raw[i] = {starts[i], lens[i]};
  const Timing t = an.estimate_timing(raw, preserve_units != 0);
  const double v[8] = {t.hsync_median, t.hsync.lo, t.hsync.hi, t.eq.lo, t.eq.hi, t.vsync.lo, t.vsync.hi, an.nominal};
  std::memcpy(timing_out, v, sizeof(v));
}

// ClassifyPulses (:168-195) or, with refine != 0, RefinePulses (:243-314; rescue when ref != NULL).
// Timing from EstimateTiming(raw, preserve_units).  Returns the classified count.
int oracle_sync_classify(const double* analyzer, const int* starts, const int* lens, int n, int preserve_units,
                         int refine, const double* ref, int64_t ref_n, double rescue_step, int* kinds, int* pstarts,
                         int* plens, int* in_order, int cap) {
  Analyzer an;
  an.init_direct(analyzer);
  vector<Pulse> raw(n);
  for (int i = 0; i < n; i++) raw[i] = {starts[i], lens[i]};
  const Timing t = an.estimate_timing(raw, preserve_units != 0);
  vector<CPulse> r;
  if (refine) {
    r = an.refine(raw, t, ref, ref ? ref_n : 0, rescue_step);
  } else {
    for (const Pulse& p : raw) {
      const int k = Analyzer::classify(p, t);
      if (k < 0) continue;
      const bool io = !r.empty() && an.quality(r.back(), (Kind)k, p);
      r.push_back({(Kind)k, p, io});
    }
  }
  for (int i = 0; i < (int)r.size() && i < cap; i++) {
    kinds[i] = r[i].kind;
    pstarts[i] = r[i].p.start;
    plens[i] = r[i].p.length;
    in_order[i] = r[i].in_order ? 1 : 0;
  }
  return (int)r.size();
}

double oracle_sync_mean_line_length(const double* analyzer, const int* kinds, const int* starts, const int* lens,
                                    const int* in_order, int n) {
  Analyzer an;
  an.init_direct(analyzer);
  return an.mean_line_length(make_pulses(kinds, starts, lens, in_order, n));
}

void oracle_sync_upstream_lines(const int* kinds, const int* starts, const int* lens, const int* in_order, int n,
                                double ref_pulse, int ref_line, double mll, int lines, double* out) {
  Analyzer an;
  const vector<double> r = an.upstream_lines(make_pulses(kinds, starts, lens, in_order, n), ref_pulse, ref_line, mll,
                                             lines);
  std::memcpy(out, r.data(), sizeof(double) * lines);
}

// FieldParityDetector.ResolveCadence; prev_first / fallback_first: -1 null, 0 false, 1 true.
int oracle_sync_resolve_cadence(const int* kinds, const int* starts, const int* lens, const int* in_order, int n,
                                double mll, int is_pal, int field_lines0, int field_lines1, int prev_first,
                                int min_conf, int has_prev_hsync, int has_fallback_line, int fallback_first,
                                int fallback_conf, int* confidence) {
  const int fl[2] = {field_lines0, field_lines1};
  const Parity p = resolve_cadence(make_pulses(kinds, starts, lens, in_order, n), mll, is_pal != 0, fl, prev_first,
                                   min_conf, has_prev_hsync != 0, has_fallback_line != 0, fallback_first,
                                   fallback_conf);
  *confidence = p.confidence;
  return p.first ? 1 : 0;
}

// VBlankSyncResolver.EstimateLine0FromTransitions: out[3][5] = {ok, line0, first_hsync,
// first_hsync_line, valid_count} for first, last, combined.
void oracle_sync_line0_transitions(const int* kinds, const int* starts, const int* lens, const int* in_order, int n,
                                   double mll, int is_pal, int num_pulses, int first_field, int cur_lines,
                                   int first_lines, double* out) {
  const Consensus c = line0_from_transitions(make_pulses(kinds, starts, lens, in_order, n), mll, is_pal != 0,
                                             num_pulses, first_field != 0, cur_lines, first_lines);
  const Estimate* e[3] = {&c.first, &c.last, &c.combined};
  for (int k = 0; k < 3; k++) {
    out[5 * k] = e[k]->ok ? 1.0 : 0.0;
    out[5 * k + 1] = e[k]->ok ? e[k]->line0 : 0.0;
    out[5 * k + 2] = e[k]->ok ? e[k]->first_hsync : 0.0;
    out[5 * k + 3] = e[k]->ok ? e[k]->first_hsync_line : 0.0;
    out[5 * k + 4] = e[k]->ok ? e[k]->valid_count : 0.0;
  }
}

// VBlankSyncResolver.NextFieldLocation (:467-497) and, with nominal > 0,
// TbcFieldDecodePipeline.ComputeVhsNextFieldOffsetSamples (:2299-2325).
double oracle_sync_next_field(int is_pal, int num_pulses, int first_field, int cur_lines, double mll,
                              double first_hsync, double nominal, int cut) {
  const FieldBoundaries b = field_boundaries(is_pal != 0, first_field != 0);
  const double fhl = first_hsync_line(is_pal != 0, num_pulses, first_field != 0);
  const double next = first_hsync + (mll * ((cur_lines + b.last_hsync) - fhl));
  return nominal > 0.0 ? next - (8.0 * nominal) + cut : next;
}

}  // extern "C"
