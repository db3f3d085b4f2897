// Colour-under chroma of the vhs field sequence (the reference's default `vhs` decode of a
// colour-under format, v0.4.0 profile); the stages are listed in chroma.hpp.
//
// Every kernel takes the round's decoded-field count from the device (the sequence's
// control block) and returns at once for absent fields, so the host issues a fixed
// sequence.  Arithmetic restates the reference's operation order: FP32 where it computes
// in float (burst products, band-pass filters, comb, burst RMS), FP64 elsewhere, with
// contraction off.  The transcendentals are the device library's (cos, atan2, hypot).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "chroma.hpp"
#include "tbc_sinc.hpp"
#include "vhs_sync.hpp"
#include "vhsd_device.hpp"

namespace vhsd {
namespace chroma {

constexpr int kBurstCheckSkipLines = 16;  // VhsChromaDecoder.BurstCheckSkipLines (:400)
constexpr int kStartingLine = 16;         // VhsChromaDecoder.StartingLine (:398)

__device__ __forceinline__ int pmod(int v, int m) {
  const int r = v % m;
  return r < 0 ? r + m : r;
}

// PositiveDegrees (VhsChromaDecoder.cs:4430-4434)
__device__ __forceinline__ double positive_degrees(double d) {
  const double r = fmod(d, 360.0);
  return r < 0.0 ? r + 360.0 : r;
}

// The resample position and wow level adjust of field sample d (counted from line 0) on
// the line table `loc` (LinearLineLocationInterpolator.FillOutputPositions and the level
// adjusts, Tbc/TbcLineResampler.cs:732-878, 1495-1525), as the luma TBC kernel forms them.
__device__ __forceinline__ void position_and_level(const double* loc, int d, int line_len, double nominal,
                                                   const double* lx, const double* ly0, const double* pow_table,
                                                   int smooth, double* pos, double* adj) {
#pragma clang fp contract(off)
  const int left = d / line_len;
  const int smp = d - left * line_len;
  const double ll = loc[left];
  const double delta = loc[left + 1] - ll;
  const double scale = nominal / line_len;
  const double sp = (double)d * scale;
  double frac = (sp - (left * nominal)) / nominal;
  frac = frac < 0.0 ? 0.0 : (frac > 1.0 ? 1.0 : frac);
  *pos = ll + (delta * frac);
  const double x = lx[left];
  *adj = smooth ? x + (ly0[left] - x) * pow_table[smp] : x;
}

// Probe window of phase line q (GetLineRange + ProbeUpconvertedBurstValue's padding,
// VhsChromaDecoder.cs:1680-1690): [ps, pe) of the analysis field.
struct Window {
  int ps, pe;
};
__device__ __forceinline__ Window probe_window(const ChromaCfg& cc, int q) {
  const int n = cc.lines * cc.line_len;
  const int pad = cc.burst_end - cc.burst_start;
  const int ps = max(0, q * cc.line_len + cc.burst_start - pad);
  return Window{ps, min(n, ps + cc.burst_end + pad)};
}

// 1. The analysis resample (ResampleChromaBurst from line 1) at the probe windows only:
// every probe sample is one output sample of the field resample.
__global__ void __launch_bounds__(256) chroma_window_kernel(ChromaCfg cc, ChromaWork cw, const double* refined,
                                                            int line_stride, const long long* origins,
                                                            const double* pow_table, double nominal, int smooth,
                                                            const float* __restrict__ lut, const int* decoded) {
  const int f = blockIdx.y;
  if (f >= *decoded) return;
  const int q = blockIdx.x;
  const Window w = probe_window(cc, q);
  const double* loc = refined + (long long)f * line_stride;
  const double* lx = cw.an_x + (long long)f * cw.lvl_stride;
  const double* ly0 = cw.an_y0 + (long long)f * cw.lvl_stride;
  const long long org = origins[f];
  const float* src = cw.plane + org;
  const long long len = cw.plane_len - org;
  float* dst = cw.win + ((long long)f * cc.lines + q) * cc.win_max;
  for (int i = threadIdx.x; i < w.pe - w.ps; i += blockDim.x) {
    double pos, adj;
    position_and_level(loc, w.ps + i + cc.line_len, cc.line_len, nominal, lx, ly0, pow_table, smooth, &pos, &adj);
    {
#pragma clang fp contract(off)
      dst[i] = (float)(sample_sinc(src, len, 0.0, pos, lut) * adj);
    }
  }
}

// One sample through the band-pass's cascade of S biquads (transposed direct form II, the
// reference's per-sample float32 order), states st[s] = (z1, z2).
template <int S>
__device__ __forceinline__ void iir_step(const float (&c)[S][6], float (&st)[S][2], float& v) {
#pragma clang fp contract(off)
#pragma unroll
  for (int s = 0; s < S; s++) {
    const float fo = (c[s][0] * v) + st[s][0];
    st[s][0] = ((c[s][1] * v) - (c[s][4] * fo)) + st[s][1];
    st[s][1] = (c[s][2] * v) - (c[s][5] * fo);
    v = fo;
  }
}

// Section s of the cascade on one sample: returns its output and updates its state.
template <int S>
__device__ __forceinline__ float iir_section(const float (&c)[S][6], float (&st)[S][2], int s, float v) {
#pragma clang fp contract(off)
  const float fo = (c[s][0] * v) + st[s][0];
  st[s][0] = ((c[s][1] * v) - (c[s][4] * fo)) + st[s][1];
  st[s][1] = (c[s][2] * v) - (c[s][5] * fo);
  return fo;
}

// `count` samples of a strided sequence (sample j at src[j * step]) through the cascade from
// `st`; with kStore the outputs go to dst[j * step] (dst may be src).  Every section sees
// the same samples in the same order as in the sample-by-sample cascade, so the values are
// identical; only the schedule is skewed: at time k section s takes sample k - s, so the S
// section updates of one time step are independent.  The inputs are loaded one batch of B
// time steps ahead; a sample is read before its own output (S - 1 time steps later) or any
// later output is written.
template <int S, int B, bool kStore, typename P>
__device__ __forceinline__ void iir_run(const float (&c)[S][6], float (&st)[S][2], const P* src, P* dst, int step,
                                        int count) {
  if (count < S) {
    for (int j = 0; j < count; j++) {
      float v = src[j * step];
      iir_step<S>(c, st, v);
      if (kStore) dst[j * step] = v;
    }
    return;
  }
  float pipe[S];  // pipe[s]: section s's latest output
  // times 0 .. S-2: sections 0 .. k
#pragma unroll
  for (int k = 0; k < S - 1; k++) {
#pragma unroll
    for (int s = S - 1; s >= 0; s--)
      if (s <= k) pipe[s] = iir_section<S>(c, st, s, s == 0 ? (float)src[k * step] : pipe[s > 0 ? s - 1 : 0]);
  }
  // times S-1 .. count-1: every section; the output of sample k - (S - 1)
  const int k0 = S - 1;
  const int steady = count - k0;
  const int full = steady - steady % B;
  float cur[B], nxt[B];
  // the loads past the last sample read the last sample again (unused)
#pragma unroll
  for (int j = 0; j < B; j++) cur[j] = src[min(k0 + j, count - 1) * step];
  for (int base = 0; base < full; base += B) {
#pragma unroll
    for (int j = 0; j < B; j++) nxt[j] = src[min(k0 + base + B + j, count - 1) * step];
#pragma unroll
    for (int j = 0; j < B; j++) {
#pragma unroll
      for (int s = S - 1; s >= 0; s--) pipe[s] = iir_section<S>(c, st, s, s == 0 ? cur[j] : pipe[s > 0 ? s - 1 : 0]);
      if (kStore) dst[(base + j) * step] = pipe[S - 1];
    }
#pragma unroll
    for (int j = 0; j < B; j++) cur[j] = nxt[j];
  }
  for (int j = full; j < steady; j++) {
    const float x = src[(k0 + j) * step];
#pragma unroll
    for (int s = S - 1; s >= 0; s--) pipe[s] = iir_section<S>(c, st, s, s == 0 ? x : pipe[s > 0 ? s - 1 : 0]);
    if (kStore) dst[j * step] = pipe[S - 1];
  }
  // times count .. count+S-2: sections e .. S-1 finish the last samples
#pragma unroll
  for (int e = 1; e < S; e++) {
#pragma unroll
    for (int s = S - 1; s >= 1; s--)
      if (s >= e) pipe[s] = iir_section<S>(c, st, s, pipe[s - 1]);
    if (kStore) dst[(count - S + e) * step] = pipe[S - 1];
  }
}

// 2. The burst probe of phase line q for track-phase start `cand`: heterodyne, the final
// band-pass forward and backward in float32 over the padded window
// (SosFilter.ApplyForwardBackwardFloat32InPlace: odd extension, steady-state initial
// conditions, cascaded biquads), and DemodBurstFloat32 (VhsChromaDecoder.cs:1521-1626).
// One thread per (line, start); its filter buffer is a column of LDS (element k at k * 64).
// S = the band-pass's section count.
template <int S>
__global__ void __launch_bounds__(kProbeThreads) chroma_probe_kernel(ChromaCfg cc, ChromaWork cw,
                                                                     const double* refined, int line_stride,
                                                                     const int* decoded) {
#pragma clang fp contract(off)
  extern __shared__ float probe_lds[];
  const int f = blockIdx.y;
  if (f >= *decoded) return;
  const int t = blockIdx.x * kProbeThreads + threadIdx.x;
  if (t >= 2 * cc.lines) return;  // no barriers below
  const int cand = t / cc.lines;
  const int q = t - cand * cc.lines;
  const int n = cc.lines * cc.line_len;
  const int rot = pmod(cc.rot[cand] * (q + 1), 4);
  const Window w = probe_window(cc, q);
  const int nw = w.pe - w.ps;
  const int P = cc.pad;
  float* e = probe_lds + threadIdx.x;
  const float* win = cw.win + ((long long)f * cc.lines + q) * cc.win_max;
  const float* het = cc.het + (long long)rot * n;
#pragma unroll 16
  for (int i = 0; i < nw; i++) e[(P + i) * kProbeThreads] = het[w.ps + i] * win[i];
  if (nw > P) {
    const float first = e[P * kProbeThreads];
    const float last = e[(P + nw - 1) * kProbeThreads];
    for (int i = 0; i < P; i++) e[i * kProbeThreads] = (2.0f * first) - e[(2 * P - i) * kProbeThreads];
    for (int i = 0; i < P; i++)
      e[(P + nw + i) * kProbeThreads] = (2.0f * last) - e[(P + nw - 2 - i) * kProbeThreads];
    const int m = nw + 2 * P;
    float c[S][6];
    for (int s = 0; s < S; s++)
      for (int k = 0; k < 6; k++) c[s][k] = cc.sos[s][k];
    float sz[S][2];
    for (int dir = 0; dir < 2; dir++) {
      const float x0 = e[(dir == 0 ? 0 : m - 1) * kProbeThreads];
      for (int s = 0; s < S; s++) {
        sz[s][0] = cc.zi[s][0] * x0;
        sz[s][1] = cc.zi[s][1] * x0;
      }
      float* p = e + (dir == 0 ? 0 : (m - 1) * kProbeThreads);
      iir_run<S, 8, true>(c, sz, p, p, dir == 0 ? kProbeThreads : -kProbeThreads, m);
    }
  }
  // DemodBurstFloat32 over the filtered window less the padding on either side
  const int pad_b = cc.burst_end - cc.burst_start;
  const int fs = min(pad_b, nw);
  const int fe = max(fs, nw - pad_b);
  const int len = fe - fs;
  const int bstart = w.ps + pad_b;
  const float* b = e + (P + fs) * kProbeThreads;
  auto acc = [&](int index, double& ia, double& qa) {
    const float sample = b[index * kProbeThreads];
    const float ip = sample * cc.bcos[bstart + index];
    const float qp = sample * cc.bsin[bstart + index];
    ia += (double)ip;
    qa += (double)qp;
  };
  double ia[4][4], qa[4][4];
  for (int g = 0; g < 4; g++)
    for (int l = 0; l < 4; l++) ia[g][l] = qa[g][l] = 0.0;
  const int main_end = len & ~15;
  for (int index = 0; index < main_end; index += 16)
    for (int lane = 0; lane < 4; lane++)
      for (int g = 0; g < 4; g++) acc(index + lane + 4 * g, ia[g][lane], qa[g][lane]);
  double lanes_i[4], lanes_q[4];
  for (int k = 0; k < 4; k++) {
    lanes_i[k] = ((ia[1][k] + ia[0][k]) + ia[2][k]) + ia[3][k];
    lanes_q[k] = ((qa[1][k] + qa[0][k]) + qa[2][k]) + qa[3][k];
  }
  double ic = (lanes_i[0] + lanes_i[2]) + (lanes_i[1] + lanes_i[3]);
  double qc = (lanes_q[0] + lanes_q[2]) + (lanes_q[1] + lanes_q[3]);
  const int epi_end = len & ~3;
  if (epi_end > main_end) {
    double ei[4] = {ic, 0.0, 0.0, 0.0}, eq[4] = {qc, 0.0, 0.0, 0.0};
    for (int index = main_end; index < epi_end; index += 4)
      for (int lane = 0; lane < 4; lane++) acc(index + lane, ei[lane], eq[lane]);
    ic = (ei[0] + ei[2]) + (ei[1] + ei[3]);
    qc = (eq[0] + eq[2]) + (eq[1] + eq[3]);
  }
  for (int index = epi_end; index < len; index++) acc(index, ic, qc);
  // ComputeLineScale (:4152-4162) on the analysis line numbers (line q + 1 of N + 1)
  const double* loc = refined + (long long)f * line_stride;
  const double scale = q + 1 < cc.lines ? (loc[q + 2] - loc[q + 1]) / cc.input_line_length : 1.0;
  double* out = cw.probe + (((long long)f * 2 + cand) * cc.lines + q) * 5;
  out[0] = positive_degrees(atan2(qc, ic) * (180.0 / M_PI));
  out[1] = positive_degrees((double)(bstart - q * cc.line_len) * __fma_rn(-scale, 90.0, 90.0));
  out[2] = hypot(ic, qc);
  out[3] = ic;
  out[4] = qc;
}

// 3. Per field and track-phase start: ShouldFlipTrackPhase (:4164-4217) and the averages of
// SummarizePhaseSequence (:4219-4293; colour killer off: the burst-detected line is 0).  One
// wave per start: the lanes take the lines (delta buckets, normalised I/Q), the bucket counts
// are integer sums in any order, and lane 0 adds the FP64 terms in line order.
constexpr int kSummaryLines = 512;
__global__ void __launch_bounds__(128) chroma_summary_kernel(ChromaCfg cc, ChromaWork cw, const int* decoded) {
#pragma clang fp contract(off)
  __shared__ double ni_s[2][kSummaryLines], nq_s[2][kSummaryLines], mag_s[2][kSummaryLines];
  __shared__ unsigned char use_s[2][kSummaryLines];
  __shared__ int buckets[2][4];
  const int f = blockIdx.x;
  if (f >= *decoded) return;
  const int cand = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const double* pr = cw.probe + ((long long)f * 2 + cand) * cc.lines * 5;
  const int end = cc.lines + 1;  // lines_out + line offset
  const int check_end = end - kBurstCheckSkipLines;
  if (lane < 4) buckets[cand][lane] = 0;
  __syncthreads();
  for (int k = lane; k < cc.lines; k += 64) {
    const int line = k + 1;
    const bool inside = !(line <= kBurstCheckSkipLines || line >= check_end);
    if (k >= 1 && inside) {
      const double delta = positive_degrees(pr[k * 5] - pr[(k - 1) * 5]);
      const int bucket = (int)floor((delta + 45.0) / 90.0) % 4;
      atomicAdd(&buckets[cand][bucket], 1);
    }
    const double m = pr[k * 5 + 2];
    const bool used = inside && m != 0.0;
    use_s[cand][k] = used ? 1 : 0;
    mag_s[cand][k] = m;
    ni_s[cand][k] = used ? pr[k * 5 + 3] / m : 0.0;
    nq_s[cand][k] = used ? pr[k * 5 + 4] / m : 0.0;
  }
  __syncthreads();
  if (lane != 0) return;
  const int* d = buckets[cand];
  cw.flip[f * 2 + cand] = cc.is_ntsc ? (d[0] < d[2]) : ((d[1] + d[3]) < (d[0] + d[2]));
  double ei = 0.0, eq = 0.0, oi = 0.0, oq = 0.0, mag = 0.0;
  int count = 0;
  // branch-free, so the loads of several lines issue together: a skipped line adds +0.0,
  // which leaves a sum started at +0.0 unchanged bit for bit (round to nearest never
  // produces -0.0 from it)
#pragma unroll 8
  for (int k = 0; k < cc.lines; k++) {
    const bool u = use_s[cand][k] != 0;
    const bool odd = ((k + 1) & 1) == 1;
    const double m = mag_s[cand][k], a = ni_s[cand][k], b = nq_s[cand][k];
    count += u ? 1 : 0;
    mag += u ? m : 0.0;
    oi += u && odd ? a : 0.0;
    oq += u && odd ? b : 0.0;
    ei += u && !odd ? a : 0.0;
    eq += u && !odd ? b : 0.0;
  }
  double* sm = cw.summ + ((long long)f * 2 + cand) * 4;
  sm[0] = count > 0 ? mag / count : 0.0;
  sm[1] = positive_degrees(atan2(eq + oq, ei + oi) * 180.0 / M_PI);
  sm[2] = positive_degrees(atan2(eq, ei) * 180.0 / M_PI);
  sm[3] = positive_degrees(atan2(oq, oi) * 180.0 / M_PI);
}

// 4. The track-phase start carried through the round's fields in order
// (GetPhaseRotationSequenceCore :1906-1965 with BuildPhaseSequence's next index, then
// CommitChromaStateCore): a start s uses its own sequence and hands 1 - s on, unless the
// flip test fires, when the field uses start 1 - s and hands s on.
__global__ void chroma_chain_kernel(ChromaWork cw, const int* decoded) {
  if (threadIdx.x != 0) return;
  const int nf = *decoded;
  ChromaState st = *cw.state;
  for (int f = 0; f < nf; f++) {
    const int s = st.rot_index;
    const bool fl = cw.flip[f * 2 + s] != 0;
    cw.chosen[f] = fl ? 1 - s : s;
    st.rot_index = fl ? s : 1 - s;
    cw.field_no[f] = st.field_number++;
  }
  *cw.state = st;
}

// 5. RefineLineLocationsFromBurst (:2342-2428, v0.4.0 form) on the hsync-refined lines, then
// the NTSC fsc phase shift: the render locations.  Each line moves by its own phase and
// the original distance to the next line, so lines are independent.
__global__ void __launch_bounds__(256) chroma_render_kernel(ChromaCfg cc, ChromaWork cw, const double* refined,
                                                            double* render, int line_stride, int nloc,
                                                            double fsc_shift, const int* decoded) {
#pragma clang fp contract(off)
  const int f = blockIdx.y;
  if (f >= *decoded) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nloc) return;
  const double* loc = refined + (long long)f * line_stride;
  double v = loc[i];
  const int k = i - 1;  // phase-sequence index of line i (the sequence starts at line 1)
  if (k >= 9 && k < cc.lines && i + 1 < nloc) {
    const int c = cw.chosen[f];
    const double* s = cw.summ + ((long long)f * 2 + c) * 4;
    const double* pr = cw.probe + (((long long)f * 2 + c) * cc.lines + k) * 5;
    const double target = cc.is_ntsc ? s[1] : ((i & 1) == 1 ? s[3] : s[2]);
    const double delta = positive_degrees(((target - pr[0]) + pr[1]) + 180.0) - 180.0;
    const double ll = loc[i + 1] - v;
    const double scale = ll / cc.line_len;
    v += ((delta / 360.0) * cc.fsc_ratio) * scale;
  }
  render[(long long)f * line_stride + i] = fsc_shift != 0.0 ? v - fsc_shift : v;
}

// Phase line of output line j when the phase sequence (lines 1..N) is applied from output
// line offset `lo` through numpy slices (GetNumpySliceRange :4310-4340 in UpconvertChroma):
// -1 for a line no slice covers.
__device__ __forceinline__ int phase_line_of(int j, int lines, int lo) {
  if (j <= lines - lo) return j + lo - 1;
  const int k = j - lines + lo - 1;  // a wrapped (negative-start) slice
  return (k >= 0 && k <= lo - 3) ? k : -1;
}

// 6. The render resample of the chroma plane (same positions and level adjusts as the
// luma), NTSC burst de-emphasis (ApplyBurstDeemphasisCore :2695-2719, five samples after
// the burst) and the upconversion: NTSC phase compensated (the cosine of the sample's phase
// from the phase kernel), PAL heterodyne (UpconvertChroma :2009-2045).
__global__ void __launch_bounds__(256) chroma_field_kernel(ChromaCfg cc, ChromaWork cw, const double* render,
                                                           int line_stride, const long long* origins,
                                                           const int* first_lines, const double* line_x,
                                                           const double* line_y0, const double* pow_table,
                                                           double nominal, int smooth,
                                                           const float* __restrict__ lut, const int* decoded) {
#pragma clang fp contract(off)
  __shared__ float4 rows[256 / 64][kTbcStage];
  const int f = blockIdx.y;
  if (f >= *decoded) return;
  const int n = cc.lines * cc.line_len;
  const int d0 = blockIdx.x * blockDim.x + threadIdx.x;
  // every lane takes part in the wave's table-row fetch; lanes past the field repeat its
  // last sample
  const bool valid = d0 < n;
  const int d = valid ? d0 : n - 1;
  const int first = first_lines[f];
  const long long org = origins[f];
  double pos, adj;
  position_and_level(render + (long long)f * line_stride, d + first * cc.line_len, cc.line_len, nominal,
                     line_x + (long long)f * cw.lvl_stride, line_y0 + (long long)f * cw.lvl_stride, pow_table, smooth,
                     &pos, &adj);
  const double sv = sample_sinc_wave(cw.plane + org, cw.plane_len - org, 0.0, pos, lut, rows[threadIdx.x >> 6]);
  if (!valid) return;
  double v = (double)(float)(sv * adj);
  const int j = d / cc.line_len;
  const int smp = d - j * cc.line_len;
  float o;
  if (cc.is_ntsc) {
    if (smp >= cc.burst_end + 5) v *= 2.0;
    o = (float)((double)(float)v * -cos(cw.theta[(long long)f * n + d]));
  } else {
    const int k = phase_line_of(j, cc.lines, first);
    if (k < 0) {
      o = 0.0f;
    } else {
      const int rot = pmod(cc.rot[cw.chosen[f]] * (k + 1), 4);
      o = (float)(v * (double)cc.het[(long long)rot * n + d]);
    }
  }
  cw.field[(long long)f * n + d] = o;
}

// 7. NTSC: the phases of UpconvertChromaPhaseCompensated (:2167-2249) for the field's
// colour-frame target (NtscFieldPhaseTarget :3915-3925): per line, four lanes each carry an
// FP64 phase in steps of four samples (the vectorised loop), the remainder carried from the
// line's start.  Only the running sums are sequential; the field kernel takes each sample's
// cosine.  Runs before the field kernel.
__global__ void __launch_bounds__(256) chroma_ntsc_phase_kernel(ChromaCfg cc, ChromaWork cw, const int* decoded) {
#pragma clang fp contract(off)
  const int f = blockIdx.y;
  if (f >= *decoded) return;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 4 * cc.lines) return;
  const int k = t >> 2, lane = t & 3;
  const int n = cc.lines * cc.line_len;
  const double half_pi = 1.5707963267948966;
  const double deg2rad = 0.017453292519943295;
  const double half_pi_per_mhz = 1.5707963267948965e-6;
  const double carrier_term = cc.cuc_hz * half_pi_per_mhz;
  const double carrier_ratio = carrier_term / cc.fsc_mhz;
  const double het = carrier_ratio + half_pi;
  const double coef[4] = {het * 0.0, het, het * 2.0, het * 3.0};
  const double four = het * 4.0;
  const int c = cw.chosen[f];
  // the target depends on the colour frame only: (first, second frame) -> 147 or -33
  const bool second_frame = ((cw.field_no[f] / 2) & 1) == 1;
  const double target = second_frame ? 147.0 : -33.0;
  const int rot = pmod(cc.rot[c] * (k + 1), 4);
  const double burst_phase = cw.probe[(((long long)f * 2 + c) * cc.lines + k) * 5];
  const int s = k * cc.line_len;
  const int e = s + cc.line_len;
  double theta = __fma_rn((double)s, het, rot * half_pi);
  theta = __fma_rn(target + burst_phase, deg2rad, theta);
  const int vc = cc.line_len & ~3;
  const int ve = s + vc;
  double* th = cw.theta + (long long)f * n;
  double tl = theta + coef[lane];
  for (int i = s + lane; i < ve; i += 4) {
    th[i] = tl;
    tl += four;
  }
  if (lane == 0) {
    double tt = theta + het * vc;
    for (int i = ve; i < e; i++) {
      th[i] = tt;
      tt += het;
    }
  }
}

// 8. The final band-pass forward and backward in float32 over the whole field
// (SosFilter.ApplyForwardBackwardFloat32InPlace), exactly, spread over the chip: the
// odd-extended sequence is cut into chunks of `iir_chunk` samples, one thread each.  A
// chunk's thread runs the section cascade sample by sample (the values of the reference's
// per-sample cascade) from a zero state `iir_warm` samples before its chunk (the band-pass
// forgets its state well within them) and keeps its state on entering the chunk and on
// leaving it; the chunk processed first starts from the true initial state.  The check
// kernel then compares every chunk's entry state with its predecessor's end state bit for
// bit: when all agree, every chunk started from the sequential state (induction from the
// first), so every value is the sequential one; a chunk that disagrees is run again from
// its predecessor's end state, in chunk order, until the states agree again.
__device__ __forceinline__ long long iir_field(const ChromaWork& cw, int f) {
  return (long long)f * cw.iir_chunks * kIirChunk;
}

// Chunk-major layout: element idx of chunk idx / C at position idx % C, so that the threads
// of consecutive chunks read consecutive words at every step.
__device__ __forceinline__ int iir_at(const ChromaWork& cw, int idx) {
  return (idx % kIirChunk) * cw.iir_chunks + idx / kIirChunk;
}

// Runs chunk t of a field (src -> dst) from `st`, which holds the state entering it; on
// return `st` holds the state leaving it.
template <int S>
__device__ __forceinline__ void iir_chunk(const ChromaWork& cw, const float (&c)[S][6], const float* src, float* dst,
                                          int m, int t, bool forward, float (&st)[S][2]) {
  const int count = min(m - t * kIirChunk, kIirChunk);
  const int stride = cw.iir_chunks;
  const int at = t + (forward ? 0 : (count - 1) * stride);
  iir_run<S, 16, true>(c, st, src + at, dst + at, forward ? stride : -stride, count);
}

// The odd-extended field into fa (chunk-major) and, after the backward pass, fa's middle
// back into the field: tiles of 64 chunks x 64 samples, transposed through LDS so that both
// sides move whole rows.
__global__ void __launch_bounds__(256) chroma_ext_kernel(ChromaCfg cc, ChromaWork cw, const int* decoded) {
#pragma clang fp contract(off)
  __shared__ float tile[kIirChunk][kIirChunk + 1];
  const int f = blockIdx.y;
  if (f >= *decoded) return;
  const int n = cc.lines * cc.line_len;
  const int P = cc.pad;
  const int m = n + 2 * P;
  const int t0 = blockIdx.x * kIirChunk;
  const float* fld = cw.field + (long long)f * n;
  for (int r = threadIdx.x; r < kIirChunk * kIirChunk; r += blockDim.x) {
    const int idx = t0 * kIirChunk + r;
    float v = 0.0f;
    if (idx < P)
      v = (2.0f * fld[0]) - fld[P - idx];
    else if (idx < P + n)
      v = fld[idx - P];
    else if (idx < m)
      v = (2.0f * fld[n - 1]) - fld[n - 2 - (idx - P - n)];
    tile[r / kIirChunk][r % kIirChunk] = v;
  }
  __syncthreads();
  float* a = cw.fa + iir_field(cw, f);
  for (int r = threadIdx.x; r < kIirChunk * kIirChunk; r += blockDim.x) {
    const int k = r / kIirChunk, tc = r % kIirChunk;
    if (t0 + tc < cw.iir_chunks) a[(long long)k * cw.iir_chunks + t0 + tc] = tile[tc][k];
  }
}

__global__ void __launch_bounds__(256) chroma_out_kernel(ChromaCfg cc, ChromaWork cw, const int* decoded) {
  __shared__ float tile[kIirChunk][kIirChunk + 1];
  const int f = blockIdx.y;
  if (f >= *decoded) return;
  const int n = cc.lines * cc.line_len;
  const int P = cc.pad;
  const int t0 = blockIdx.x * kIirChunk;
  const float* a = cw.fa + iir_field(cw, f);
  for (int r = threadIdx.x; r < kIirChunk * kIirChunk; r += blockDim.x) {
    const int k = r / kIirChunk, tc = r % kIirChunk;
    tile[tc][k] = t0 + tc < cw.iir_chunks ? a[(long long)k * cw.iir_chunks + t0 + tc] : 0.0f;
  }
  __syncthreads();
  float* fld = cw.field + (long long)f * n;
  for (int r = threadIdx.x; r < kIirChunk * kIirChunk; r += blockDim.x) {
    const int i = t0 * kIirChunk + r - P;
    if (i >= 0 && i < n) fld[i] = tile[r / kIirChunk][r % kIirChunk];
  }
}

// One pass: forward reads fa and writes fb, backward reads fb and writes fa.
template <int S>
__global__ void __launch_bounds__(256) chroma_iir_pass_kernel(ChromaCfg cc, ChromaWork cw, int forward,
                                                              const int* decoded) {
#pragma clang fp contract(off)
  const int f = blockIdx.y;
  if (f >= *decoded) return;
  const int n = cc.lines * cc.line_len;
  const int m = n + 2 * cc.pad;
  const int C = kIirChunk;
  const int chunks = (m + C - 1) / C;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= chunks) return;
  const float* src = (forward ? cw.fa : cw.fb) + iir_field(cw, f);
  float* dst = (forward ? cw.fb : cw.fa) + iir_field(cw, f);
  float c[S][6];
  for (int s = 0; s < S; s++)
    for (int q = 0; q < 6; q++) c[s][q] = cc.sos[s][q];
  const float x0 = forward ? src[iir_at(cw, 0)] : src[iir_at(cw, m - 1)];
  float st[S][2];
  const int lo = t * C, hi = min(m, lo + C);
  // warm-up: from the true initial state when it reaches the sequence's first sample; chunk
  // by chunk (each a column of the chunk-major buffer)
  const int stride = cw.iir_chunks;
  if (forward) {
    const int w0 = lo - cc.iir_warm;
    const bool exact = w0 <= 0;
    for (int s = 0; s < S; s++) {
      st[s][0] = exact ? cc.zi[s][0] * x0 : 0.0f;
      st[s][1] = exact ? cc.zi[s][1] * x0 : 0.0f;
    }
    for (int i = exact ? 0 : w0; i < lo;) {
      const int u = i / kIirChunk, k = i % kIirChunk;
      iir_run<S, 16, false>(c, st, src + u + k * stride, (float*)nullptr, stride, kIirChunk - k);
      i += kIirChunk - k;
    }
  } else {
    const int w1 = hi - 1 + cc.iir_warm;
    const bool exact = w1 >= m - 1;
    for (int s = 0; s < S; s++) {
      st[s][0] = exact ? cc.zi[s][0] * x0 : 0.0f;
      st[s][1] = exact ? cc.zi[s][1] * x0 : 0.0f;
    }
    for (int i = exact ? m - 1 : w1; i >= hi;) {
      const int u = i / kIirChunk, k = i % kIirChunk;
      iir_run<S, 16, false>(c, st, src + u + k * stride, (float*)nullptr, -stride, k + 1);
      i -= k + 1;
    }
  }
  float* ent = cw.iir_state + (((long long)f * cw.iir_chunks + t) * 2) * (2 * kMaxSections);
  float* ext = ent + 2 * kMaxSections;
  for (int s = 0; s < S; s++) {
    ent[2 * s] = st[s][0];
    ent[2 * s + 1] = st[s][1];
  }
  iir_chunk<S>(cw, c, src, dst, m, t, forward != 0, st);
  for (int s = 0; s < S; s++) {
    ext[2 * s] = st[s][0];
    ext[2 * s + 1] = st[s][1];
  }
}

// The check after a pass (one workgroup per field): entry states against the predecessors'
// end states.  A round takes the chunks that disagree with their predecessors' end states,
// gives each that state as its entry state and runs them again, all at once; the first
// disagreeing chunk's predecessor is exact (every chunk before it agrees), so each round
// makes at least that chunk exact and the rounds end; a warm-up that merged leaves none to
// run.
template <int S>
__global__ void __launch_bounds__(1024) chroma_iir_check_kernel(ChromaCfg cc, ChromaWork cw, int forward,
                                                                const int* decoded) {
#pragma clang fp contract(off)
  __shared__ int any_bad;
  const int f = blockIdx.x;
  if (f >= *decoded) return;
  const int n = cc.lines * cc.line_len;
  const int m = n + 2 * cc.pad;
  const int chunks = (m + kIirChunk - 1) / kIirChunk;
  float* states = cw.iir_state + (long long)f * cw.iir_chunks * 2 * (2 * kMaxSections);
  auto entry = [&](int t) { return states + (long long)t * 2 * (2 * kMaxSections); };
  auto leave = [&](int t) { return entry(t) + 2 * kMaxSections; };
  const float* src = (forward ? cw.fa : cw.fb) + iir_field(cw, f);
  float* dst = (forward ? cw.fb : cw.fa) + iir_field(cw, f);
  float c[S][6];
  for (int s = 0; s < S; s++)
    for (int q = 0; q < 6; q++) c[s][q] = cc.sos[s][q];
  for (int round = 0; round <= chunks; round++) {
    if (threadIdx.x == 0) any_bad = 0;
    __syncthreads();
    // entry(t) is read and written by chunk t's thread only in this loop, leave(p) only read
    unsigned mine = 0;
    for (int t = threadIdx.x, k = 0; t < chunks; t += blockDim.x, k++) {
      const int p = forward ? t - 1 : t + 1;
      if (p < 0 || p >= chunks) continue;
      bool same = true;
      for (int q = 0; q < 2 * S; q++) same &= __float_as_uint(entry(t)[q]) == __float_as_uint(leave(p)[q]);
      if (same) continue;
      for (int q = 0; q < 2 * S; q++) entry(t)[q] = leave(p)[q];
      mine |= 1u << k;
      any_bad = 1;
    }
    __syncthreads();
    if (!any_bad) break;
    for (int t = threadIdx.x, k = 0; t < chunks; t += blockDim.x, k++) {
      if (!((mine >> k) & 1u)) continue;
      float st[S][2];
      for (int s = 0; s < S; s++) {
        st[s][0] = entry(t)[2 * s];
        st[s][1] = entry(t)[2 * s + 1];
      }
      iir_chunk<S>(cw, c, src, dst, m, t, forward != 0, st);
      for (int s = 0; s < S; s++) {
        leave(t)[2 * s] = st[s][0];
        leave(t)[2 * s + 1] = st[s][1];
      }
    }
    __syncthreads();
  }
}

// 9. ApplyAutomaticChromaGainWithCombToU16 (:3504-3583) with WriteScaledChromaToU16
// (:3605-3656), float32 comb retained and float32 burst RMS (RmsFloat32 :4399-4421): one
// workgroup per line.
__global__ void __launch_bounds__(256) chroma_acc_kernel(ChromaCfg cc, ChromaWork cw, const int* out_index,
                                                         const int* decoded) {
#pragma clang fp contract(off)
  extern __shared__ float comb[];
  __shared__ double scale_s;
  const int f = blockIdx.y;
  if (f >= *decoded) return;
  const int line = blockIdx.x;
  const int L = cc.line_len;
  const int N = cc.lines;
  const int n = N * L;
  const int s = line * L;
  unsigned short* out = cw.out + (long long)out_index[f] * n + s;
  const int first = min(N, max(kStartingLine, 0));
  if (line < first) {
    for (int i = threadIdx.x; i < L; i += blockDim.x) out[i] = 32767;
    return;
  }
  const float* fld = cw.field + (long long)f * n;
  const int dist = cc.is_ntsc ? 1 : 2;
  for (int i = threadIdx.x; i < L; i += blockDim.x) {
    float v;
    if (line < N - 2) {
      const double cur = fld[s + i], adv = fld[(line + dist) * L + i], del = fld[(line - dist) * L + i];
      v = (float)((((cur * 2.0) - adv) - del) / 4.0);
    } else {
      v = fld[s + i];
    }
    comb[i] = v;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nb = cc.burst_end - cc.burst_start;
    float rms = 0.0f;
    if (nb > 0) {
      float mean = 0.0f;
      for (int i = 0; i < nb; i++) mean += comb[cc.burst_start + i];
      mean /= (float)nb;
      float ss = 0.0f;
      for (int i = 0; i < nb; i++) {
        const float c = comb[cc.burst_start + i] - mean;
        const float sq = c * c;
        ss += sq;
      }
      rms = sqrtf(ss / (float)nb);
    }
    scale_s = rms != 0.0f ? cc.burst_abs_ref / (double)rms : 1.0;
  }
  __syncthreads();
  const double scale = scale_s;
  const int sat = min(max((n & ~3) - s, 0), L);
  for (int i = threadIdx.x; i < L; i += blockDim.x) {
    const double scaled = (double)comb[i] * scale;
    const double shifted = scaled + 32767.0;
    unsigned short o;
    if (i < sat) {
      o = (!isfinite(shifted) || shifted <= 0.0) ? 0 : shifted >= 65535.0 ? 65535 : (unsigned short)shifted;
    } else {
      o = (!isfinite(shifted) || shifted < -9.2233720368547758e18 || shifted > 9.2233720368547758e18)
              ? 0
              : (unsigned short)(long long)shifted;
    }
    out[i] = o;
  }
}

// reset: a new sequence (track-phase start 0); field_number >= 0 sets the next field's number
__global__ void chroma_state_kernel(ChromaState* st, int reset, int field_number) {
  if (threadIdx.x != 0) return;
  if (reset) st->rot_index = 0;
  if (field_number >= 0) st->field_number = field_number;
}

__global__ void fill_ones_kernel(int* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 1;
}

}  // namespace chroma
}  // namespace vhsd

using namespace vhsd::chroma;

extern "C" {

hipError_t vhsd_launch_chroma_state(ChromaState* st, int reset, int field_number, int* ones, int max_fields,
                                    hipStream_t stream) {
  hipLaunchKernelGGL(chroma_state_kernel, dim3(1), dim3(64), 0, stream, st, reset, field_number);
  if (ones && max_fields > 0)
    hipLaunchKernelGGL(fill_ones_kernel, dim3((max_fields + 63) / 64), dim3(64), 0, stream, ones, max_fields);
  return hipGetLastError();
}

// Stages 1-5 (before the luma render): analysis resample, probes, summaries, chain and
// the render locations.  an_x / an_y0 must hold the analysis level adjusts.
hipError_t vhsd_launch_chroma_analysis(const ChromaCfg* cc, const ChromaWork* cw, const double* refined,
                                       double* render, int line_stride, int nloc, const long long* origins,
                                       const double* pow_table, double nominal, int smooth, const float* lut,
                                       double fsc_shift, int max_fields, const int* decoded, hipStream_t stream) {
  if (max_fields <= 0) return hipSuccess;
  hipLaunchKernelGGL(chroma_window_kernel, dim3(cc->lines, max_fields), dim3(256), 0, stream, *cc, *cw, refined,
                     line_stride, origins, pow_table, nominal, smooth, lut, decoded);
  const size_t lds = (size_t)cc->probe_ext * kProbeThreads * sizeof(float);
  const dim3 pgrid((2 * cc->lines + kProbeThreads - 1) / kProbeThreads, max_fields);
  hipError_t e = hipSuccess;
  switch (cc->sections) {
#define VHSD_PROBE_CASE(S)                                                                                case S:                                                                                                   e = vhsd::lds_limit_once((const void*)chroma_probe_kernel<S>, (int)lds);                                if (e != hipSuccess) return e;                                                                          hipLaunchKernelGGL(chroma_probe_kernel<S>, pgrid, dim3(kProbeThreads), lds, stream, *cc, *cw, refined,                        line_stride, decoded);                                                               break;
    VHSD_PROBE_CASE(1)
    VHSD_PROBE_CASE(2)
    VHSD_PROBE_CASE(3)
    VHSD_PROBE_CASE(4)
    VHSD_PROBE_CASE(5)
    VHSD_PROBE_CASE(6)
    VHSD_PROBE_CASE(7)
    VHSD_PROBE_CASE(8)
#undef VHSD_PROBE_CASE
    default:
      return hipErrorInvalidValue;
  }
  if (cc->lines > kSummaryLines) return hipErrorInvalidValue;
  hipLaunchKernelGGL(chroma_summary_kernel, dim3(max_fields), dim3(128), 0, stream, *cc, *cw, decoded);
  hipLaunchKernelGGL(chroma_chain_kernel, dim3(1), dim3(64), 0, stream, *cw, decoded);
  hipLaunchKernelGGL(chroma_render_kernel, dim3((nloc + 255) / 256, max_fields), dim3(256), 0, stream, *cc, *cw,
                     refined, render, line_stride, nloc, fsc_shift, decoded);
  return hipGetLastError();
}

// Stages 6-9 (after the luma render, whose level adjusts line_x / line_y0 they reuse).
hipError_t vhsd_launch_chroma_fields(const ChromaCfg* cc, const ChromaWork* cw, const double* render,
                                     int line_stride, const long long* origins, const int* first_lines,
                                     const double* line_x, const double* line_y0, const double* pow_table,
                                     double nominal, int smooth, const float* lut,
                                     const vhsd::vsync::FieldRecord* records, const vhsd::vsync::SeqCtl* ctl,
                                     const int* out_index, int max_fields, const int* decoded, hipStream_t stream) {
  if (max_fields <= 0) return hipSuccess;
  const int n = cc->lines * cc->line_len;
  if (cc->is_ntsc)
    hipLaunchKernelGGL(chroma_ntsc_phase_kernel, dim3((4 * cc->lines + 255) / 256, max_fields), dim3(256), 0, stream,
                       *cc, *cw, decoded);
  hipLaunchKernelGGL(chroma_field_kernel, dim3((n + 255) / 256, max_fields), dim3(256), 0, stream, *cc, *cw, render,
                     line_stride, origins, first_lines, line_x, line_y0, pow_table, nominal, smooth, lut, decoded);
  (void)records;
  (void)ctl;
  {
    const int m = n + 2 * cc->pad;
    const int chunks = (m + kIirChunk - 1) / kIirChunk;
    // the check kernel's 1024 threads take at most 32 chunks each (a bit mask)
    if (cc->iir_chunk != kIirChunk || chunks > cw->iir_chunks || chunks > 32 * 1024) return hipErrorInvalidValue;
    const dim3 tiles((chunks + kIirChunk - 1) / kIirChunk, max_fields);
    hipLaunchKernelGGL(chroma_ext_kernel, tiles, dim3(256), 0, stream, *cc, *cw, decoded);
    switch (cc->sections) {
#define VHSD_IIR_CASE(S)                                                                                        \
  case S:                                                                                                       \
    for (int dir = 1; dir >= 0; dir--) {                                                                        \
      hipLaunchKernelGGL(chroma_iir_pass_kernel<S>, dim3((chunks + 255) / 256, max_fields), dim3(256), 0, stream, \
                         *cc, *cw, dir, decoded);                                                               \
      hipLaunchKernelGGL(chroma_iir_check_kernel<S>, dim3(max_fields), dim3(1024), 0, stream, *cc, *cw, dir,   \
                         decoded);                                                                              \
    }                                                                                                           \
    break;
      VHSD_IIR_CASE(1)
      VHSD_IIR_CASE(2)
      VHSD_IIR_CASE(3)
      VHSD_IIR_CASE(4)
      VHSD_IIR_CASE(5)
      VHSD_IIR_CASE(6)
      VHSD_IIR_CASE(7)
      VHSD_IIR_CASE(8)
#undef VHSD_IIR_CASE
      default:
        return hipErrorInvalidValue;
    }
    hipLaunchKernelGGL(chroma_out_kernel, tiles, dim3(256), 0, stream, *cc, *cw, decoded);
  }
  hipLaunchKernelGGL(chroma_acc_kernel, dim3(cc->lines, max_fields), dim3(256),
                     (size_t)cc->line_len * sizeof(float), stream, *cc, *cw, out_index, decoded);
  return hipGetLastError();
}

}  // extern "C"
