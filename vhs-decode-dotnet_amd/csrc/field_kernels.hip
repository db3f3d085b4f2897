// Field-level kernels: sync pulse detection, hsync edge refinement and TBC
// line resampling, all on the stitched field span that the block kernels write.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vhsd_args.hpp"
#include "tbc_sinc.hpp"
#include "vhsd_device.hpp"

namespace vhsd {

// ------------------------------------------------------------------ pulses
// PulseDetection.FindPulsesScalar (Dsp/PulseDetection.cs:195-222): a two-state
// machine over `value <= high`, NaN samples keep the state.  Runs of samples are
// processed independently from the state after the sample before the run (the last
// non-NaN sample at or before it), edges are compacted in order (per-tile counts,
// then each tile writes at the sum of the earlier tiles' counts), then pulses pair
// each end edge with the preceding start edge.

constexpr int kPulseTile = 4096;   // samples per workgroup
constexpr int kPulseThreads = 256;

__device__ __forceinline__ int in_pulse_at(const float* s, long long i, float bound) {
  // state after processing sample i (bound: pass_bound_f32 of the bias and threshold)
  for (long long k = i; k >= 0; k--) {
    const float v = s[k];
    if (v == v) return v <= bound ? 1 : 0;
  }
  return 0;  // all NaN from the start: the initial state `s[0] <= high` is false
}

// Each thread walks kPulseRun consecutive samples carrying the state after its
// previous sample (the state after sample i0 - 1 comes from in_pulse_at), so a
// sample's edge needs one load; NaN samples keep the state, sample 0 sets it
// without an edge, and samples past the span count as NaN.
constexpr int kPulseRun = kPulseTile / kPulseThreads;  // 16
static_assert(kPulseRun == 16, "the run loads are four float4");

__device__ __forceinline__ void load_run(const float* s, long long i0, long long n, bool vec, float (&x)[kPulseRun]) {
  if (vec && i0 + kPulseRun <= n) {
    const float4* p4 = reinterpret_cast<const float4*>(s + i0);
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const float4 v = p4[q];
      x[4 * q] = v.x; x[4 * q + 1] = v.y; x[4 * q + 2] = v.z; x[4 * q + 3] = v.w;
    }
  } else {
#pragma unroll
    for (int k = 0; k < kPulseRun; k++) x[k] = i0 + k < n ? s[i0 + k] : __builtin_nanf("");
  }
}

// Edge codes of the run (0 none, 1 start, 2 end) as two bits per sample.
__device__ __forceinline__ unsigned run_edges(const float* s, long long i0, long long n, float bound, bool vec) {
  float x[kPulseRun];
  load_run(s, i0, n, vec, x);
  int st = i0 > 0 && i0 <= n ? in_pulse_at(s, i0 - 1, bound) : 0;
  unsigned codes = 0;
#pragma unroll
  for (int k = 0; k < kPulseRun; k++) {
    const float v = x[k];
    if (v == v) {
      const int cur = v <= bound ? 1 : 0;
      if (i0 + k >= 1) {
        const unsigned e = (!st && cur) ? 1u : ((st && !cur) ? 2u : 0u);
        codes |= e << (2 * k);
      }
      st = cur;
    }
  }
  return codes;
}

__global__ void __launch_bounds__(kPulseThreads) pulse_count_kernel(const float* s, long long n, double bias,
                                                                    double high, int* tile_counts) {
  __shared__ int wave_counts[kPulseThreads / 64];
  const bool vec = (reinterpret_cast<uintptr_t>(s) & 15) == 0;
  const long long i0 = (long long)blockIdx.x * kPulseTile + (long long)threadIdx.x * kPulseRun;
  int c = __popc(run_edges(s, i0, n, pass_bound_f32(bias, 0.0, high), vec));  // a sample's two bits are never both set
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
  if ((threadIdx.x & 63) == 0) wave_counts[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    int tot = 0;
    for (int w = 0; w < kPulseThreads / 64; w++) tot += wave_counts[w];
    tile_counts[blockIdx.x] = tot;
  }
}

// Sum of counts[0, m) by the whole workgroup (every thread gets it).
__device__ __forceinline__ int block_total(const int* counts, int m, int* red) {
  int v = 0;
  for (int i = threadIdx.x; i < m; i += blockDim.x) v += counts[i];
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  int tot = 0;
  for (int w = 0; w < (int)(blockDim.x >> 6); w++) tot += red[w];
  __syncthreads();
  return tot;
}

// Ordered compaction of edges: edge_pos[k], edge_kind[k].  The tile's offset is the
// sum of the earlier tiles' counts.
__global__ void __launch_bounds__(kPulseThreads) pulse_write_kernel(const float* s, long long n, double bias,
                                                                    double high, const int* tile_counts,
                                                                    long long* edge_pos, unsigned char* edge_kind,
                                                                    int edge_cap) {
  __shared__ int red[kPulseThreads / 64];
  __shared__ int wave_tot[kPulseThreads / 64];
  const bool vec = (reinterpret_cast<uintptr_t>(s) & 15) == 0;
  const int tile_off = block_total(tile_counts, blockIdx.x, red);
  const long long i0 = (long long)blockIdx.x * kPulseTile + (long long)threadIdx.x * kPulseRun;
  const unsigned codes = run_edges(s, i0, n, pass_bound_f32(bias, 0.0, high), vec);
  const int c = __popc(codes);
  // exclusive scan of the per-thread counts
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int incl = c;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_up(incl, o);
    if (lane >= o) incl += u;
  }
  if (lane == 63) wave_tot[wave] = incl;
  __syncthreads();
  int idx = tile_off + incl - c;
  for (int w = 0; w < wave; w++) idx += wave_tot[w];
  unsigned rest = codes;
  while (rest) {
    const int k = (__ffs(rest) - 1) >> 1;  // sample k of the run
    if (idx < edge_cap) {
      edge_pos[idx] = i0 + k;
      edge_kind[idx] = (unsigned char)((rest >> (2 * k)) & 3u);
    }
    idx++;
    rest &= ~(3u << (2 * k));
  }
}

// Pair end edges with their start edge, filter by length and start != 0, compact
// in order.  Single workgroup.  counters[0] receives the number of edges,
// counters[1] the number of pulses found; only the first `capacity` are stored.
__global__ void __launch_bounds__(1024) pulse_pair_kernel(const long long* edge_pos, const unsigned char* edge_kind,
                                                          const int* tile_counts, int tiles, int edge_cap, int min_len,
                                                          int max_len, int* starts, int* lengths, int capacity,
                                                          int* counters) {
  __shared__ int wave_counts[16];
  __shared__ int running;
  __shared__ int red[16];
  const int total = block_total(tile_counts, tiles, red);
  const int ne = total < edge_cap ? total : edge_cap;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  if (threadIdx.x == 0) running = 0;
  __syncthreads();
  for (int b0 = 0; b0 < ne; b0 += 1024) {
    const int k = b0 + threadIdx.x;
    bool ok = false;
    long long st = 0, len = 0;
    if (k < ne && edge_kind[k] == 2) {
      const long long e = edge_pos[k];
      st = (k > 0) ? edge_pos[k - 1] : 0;  // the edge before an end edge is its start
      len = e - st;
      ok = len >= min_len && len <= max_len && st != 0;
    }
    const unsigned long long m = __ballot(ok);
    if (lane == 0) wave_counts[wave] = __popcll(m);
    __syncthreads();
    int before = running;
    for (int w = 0; w < wave; w++) before += wave_counts[w];
    if (ok) {
      const int idx = before + __popcll(m & ((1ull << lane) - 1ull));
      if (idx < capacity) {
        starts[idx] = (int)st;
        lengths[idx] = (int)len;
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      int tot = 0;
      for (int w = 0; w < 16; w++) tot += wave_counts[w];
      running += tot;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    counters[0] = total;
    counters[1] = running;
  }
}

// ------------------------------------------------------------------ hsync refine
// TbcFieldDecodePipeline.RefineLineLocationsFromHSync for vhs
// (Decode/TbcFieldDecodePipeline.cs:4570-4684, helpers :4807-5052) in FP64.
// Sequential over lines (the porch level carries from line to line); one
// wavefront per field, lanes cooperate on searches and extrema.  Sums are
// accumulated left to right on lane 0 so results are bit-identical to the
// sequential reference on identical input.
struct WaveCtx {
  const float* s;
  long long n;
  int lane;
  double bias;
  double dc;  // DC offset added to the sync reference (PrepareSyncSpanFromLevels); 0 adds nothing
  __device__ __forceinline__ double at(long long i) const {
#pragma clang fp contract(off)
    return ((double)s[i] + bias) + dc;
  }
};

__device__ __forceinline__ int iround(double v) { return (int)round(v); }

// CalculateZeroCrossing forward (PulseDetection.cs:224-274)
__device__ bool zero_crossing(const WaveCtx& w, long long start, double target, int edge, int count, double* out) {
  if (start < 0 || start >= w.n) return false;
  int e = edge;
  if (e == 0) e = (w.at(start) < target) ? 1 : -1;
  long long end = w.n < start + count + 1 ? w.n : start + count + 1;
  const long long span = end - start;
  for (long long b = 1; b < span; b += 64) {
    const long long i = b + w.lane;
    bool hit = false;
    if (i < span) {
      const double a = w.at(start + i - 1);
      const double c = w.at(start + i);
      hit = (e == 1) ? (a < target && c >= target) : (a > target && c <= target);
    }
    const unsigned long long m = __ballot(hit);
    if (m) {
      const long long x = start + b + (long long)(__ffsll((long long)m) - 1);
      const double a = w.at(x - 1) - target;
      const double c = w.at(x) - target;
      const double y = (c - a) != 0.0 ? -a / (-a + c) : 0.0;
      *out = (double)(x - 1) + y;
      return true;
    }
  }
  return false;
}

__device__ bool area_valid(const WaveCtx& w, double crossing, int one_us, double end_us, double min_hz,
                           double max_hz) {
  long long st = iround(crossing - 0.75 * one_us);
  st = st < 0 ? 0 : (st > w.n ? w.n : st);
  long long en = iround(crossing + end_us * one_us);
  en = en < st ? st : (en > w.n ? w.n : en);
  if (en <= st) return false;
  double mn = INFINITY, mx = -INFINITY;
  for (long long i = st + w.lane; i < en; i += 64) {
    const double v = w.at(i);
    mn = fmin(mn, v);
    mx = fmax(mx, v);
  }
  for (int o = 32; o > 0; o >>= 1) {
    mn = fmin(mn, __shfl_xor(mn, o));
    mx = fmax(mx, __shfl_xor(mx, o));
  }
  return mn >= min_hz && mx <= max_hz;
}

// NaN-aware note: Math.Min/Math.Max propagate NaN; fmin/fmax drop it.  A NaN in
// the window makes the reference comparison false; mirror that explicitly.
__device__ bool area_has_nan(const WaveCtx& w, double crossing, int one_us, double end_us) {
  long long st = iround(crossing - 0.75 * one_us);
  st = st < 0 ? 0 : (st > w.n ? w.n : st);
  long long en = iround(crossing + end_us * one_us);
  en = en < st ? st : (en > w.n ? w.n : en);
  bool nan = false;
  for (long long i = st + w.lane; i < en; i += 64) nan |= (w.s[i] != w.s[i]);
  return __ballot(nan) != 0;
}

// Broadcast lane k's double to the whole wave (k uniform).
__device__ __forceinline__ double wave_read(double v, int k) {
  const long long bits = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(bits & 0xffffffffll), k);
  const int hi = __builtin_amdgcn_readlane((int)(bits >> 32), k);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// TryMeanRange: the sum runs left to right in FP64 like the reference; lanes load a
// chunk in parallel and every lane accumulates the broadcast values in order.
__device__ bool mean_range(const WaveCtx& w, double start, double end, double* mean) {
  long long a = iround(start);
  a = a < 0 ? 0 : (a > w.n ? w.n : a);
  long long b = iround(end);
  b = b < a ? a : (b > w.n ? w.n : b);
  if (b <= a) {
    *mean = 0.0;
    return false;
  }
  double sum = 0.0;
  for (long long base = a; base < b; base += 64) {
    const long long i = base + w.lane;
    const double v = i < b ? w.at(i) : 0.0;
    const int cnt = (int)(b - base < 64 ? b - base : 64);
    for (int k = 0; k < cnt; k++) sum += wave_read(v, k);
  }
  *mean = sum / (double)(b - a);
  return true;
}

// Geometry shared by every line of a refinement.
struct HsyncGeom {
  int one_us;
  int normal_len;
  int search_count;
};

__device__ __forceinline__ HsyncGeom hsync_geom(const HsyncParams& p) {
  HsyncGeom g;
  g.one_us = (int)p.fs_mhz > 1 ? (int)p.fs_mhz : 1;
  g.normal_len = (int)(p.hsync_us * p.fs_mhz) > 1 ? (int)(p.hsync_us * p.fs_mhz) : 1;
  g.search_count = g.one_us * 2 > 1 ? g.one_us * 2 : 1;
  return g;
}

// Lines the refinement leaves alone (vsync lines are marked filled).  Returns
// 0 = refine, 1 = vsync line, 2 = other skip.
__device__ __forceinline__ int line_kind(const HsyncParams& p, const HsyncGeom& g, long long n, int line, double cur,
                                         long long* search_start) {
  if ((line >= 3 && line <= 6) || (p.is_pal && line >= 1 && line <= 2)) return 1;
  if (!isfinite(cur)) return 2;
  const long long ss = iround(cur - g.one_us);
  if (ss < 0 || ss >= n - 1) return 2;
  *search_start = ss;
  return 0;
}

// Right-edge candidate of TryRefineFromRightHSync / TryConfirmRightHSync without
// the final distance test against the left result (TbcFieldDecodePipeline.cs:4807-4868,
// 4950-4987).  Independent of the carried porch level.
__device__ bool right_prelim(const WaveCtx& w, const HsyncParams& p, const HsyncGeom& g, long long search_start,
                             double* cand, double* rporch) {
  const long long n = w.n;
  auto ire_hz = [&](double ire) { return p.ire0 + p.hz_ire * ire; };
  const long long rs = search_start + g.normal_len - g.one_us;
  if (rs < 0 || rs >= n - 1) return false;
  if (w.at(rs) > p.threshold) return false;
  const int rcount = g.normal_len * 2 > 1 ? g.normal_len * 2 : 1;
  const int rc_n = (int)(rcount < n - rs - 1 ? rcount : n - rs - 1);
  double rcross = 0.0;
  if (!zero_crossing(w, rs, p.threshold, 1, rc_n, &rcross)) return false;
  const double rds = rcross - g.normal_len;
  double sync_level = 0.0;
  if (!(area_valid(w, rds, g.one_us, 8.0, ire_hz(-65.0), ire_hz(30.0)) && !area_has_nan(w, rds, g.one_us, 8.0)))
    return false;
  if (!mean_range(w, rds + g.one_us, rds + 2.5 * g.one_us, &sync_level)) return false;
  if (!mean_range(w, rds + g.normal_len + g.one_us, rds + g.normal_len + 2.0 * g.one_us, rporch)) return false;
  const double mid = (*rporch + sync_level) / 2.0;
  const int c4 = (int)(400 < n - rs - 1 ? 400 : n - rs - 1);
  double rrc = 0.0;
  if (!zero_crossing(w, rs, mid, 0, c4, &rrc)) return false;
  if (!(fabs(rrc - rcross) < g.one_us / 2.0)) return false;
  const double sr = (double)(float)p.fs_mhz;
  *cand = rcross - g.normal_len + (2.25 * (sr / 40.0));
  return true;
}

// One refinable line given the porch level carried in (`p_in`), as in the loop body
// of RefineLineLocationsFromHSync with TryRefineLeftHSync (:4870-4948).  Writes the
// refined location, filled flag and the porch level carried out.
__device__ void refine_line(const WaveCtx& w, const HsyncParams& p, const HsyncGeom& g, long long search_start,
                            double cur, unsigned char fin, double p_in, bool right_ok, double rcand, double rporch,
                            double* out_loc, unsigned char* out_filled, double* p_out) {
  const long long n = w.n;
  auto ire_hz = [&](double ire) { return p.ire0 + p.hz_ire * ire; };
  double prev_porch = p_in;
  double crossing = 0.0;
  const int cnt = (int)(g.search_count < n - search_start - 1 ? g.search_count : n - search_start - 1);
  const bool has = zero_crossing(w, search_start, p.threshold, 0, cnt, &crossing);
  const double original = cur;
  double right_ref = original;
  double loc = original;
  unsigned char fil = fin;
  if (has && !fin) {
    double refined_loc = crossing, porch = 0.0, sync_level = 0.0;
    bool preserve = false, left_ok = false;
    const bool valid = area_valid(w, crossing, g.one_us, 3.5, ire_hz(-65.0), ire_hz(110.0)) &&
                       !area_has_nan(w, crossing, g.one_us, 3.5);
    if (valid && mean_range(w, crossing + g.one_us, crossing + 2.5 * g.one_us, &sync_level)) {
      preserve = true;
      bool porch_ok = true;
      if (prev_porch > 0.0)
        porch = prev_porch;
      else if (!mean_range(w, crossing - g.one_us, crossing - 0.5 * g.one_us, &porch))
        porch_ok = false;
      if (porch_ok) {
        const double mid = (porch + sync_level) / 2.0;
        const int c4 = (int)(400 < n - search_start - 1 ? 400 : n - search_start - 1);
        double rc = 0.0;
        bool good = zero_crossing(w, search_start, mid, 0, c4, &rc) && fabs(rc - crossing) < g.one_us / 2.0;
        if (!good && prev_porch > 0.0) {
          good = zero_crossing(w, search_start, (prev_porch + sync_level) / 2.0, 0, c4, &rc) &&
                 fabs(rc - crossing) < g.one_us / 2.0;
          if (good) porch = prev_porch;
        }
        if (good) {
          refined_loc = rc;
          left_ok = true;
        }
      }
    }
    if (left_ok) {
      loc = refined_loc;
      right_ref = refined_loc;
      fil = fin;
      if (porch > 0.0) prev_porch = porch;
    } else {
      fil = 1;
      loc = original;
      if (preserve) right_ref = refined_loc;
    }
  } else {
    fil = 1;
    loc = original;
  }
  if (p.use_right && right_ok && !(fabs(rcand - right_ref) >= g.one_us * 2.0)) {
    loc = rcand;
    fil = 0;
    if (rporch > 0.0) prev_porch = rporch;
  }
  *out_loc = loc;
  *out_filled = fil;
  *p_out = prev_porch;
}

// Per-(field, line) refinement state kept between the passes.
struct HsyncWork {
  unsigned char* kind;   // line_kind
  unsigned char* rt_ok;  // right prelim succeeded
  double* rt_cand;
  double* rt_porch;
  double* p_used;        // porch level the speculative pass assumed
  double* p_out;         // porch level the speculative pass carried out
  int* fb_line;          // per field: number of lines the ordered walk refined again
  double* fb_porch;      // reserved
  int stride;            // entries per field (max lines)
  const double* field_threshold;  // per field sync threshold, or NULL for HsyncParams.threshold
  const double* field_dc;         // per field DC offset of the sync reference, or NULL for none
};

__device__ __forceinline__ HsyncParams field_params(HsyncParams p, const HsyncWork& wk, int f) {
  if (wk.field_threshold) p.threshold = wk.field_threshold[f];
  return p;
}

// Pass 1: line kinds and right-edge candidates, one wavefront per line.
__global__ void __launch_bounds__(64) hsync_prelim_kernel(const float* const* spans, const long long* span_len,
                                                          const double* const* locs, const int* line_counts,
                                                          HsyncParams p, HsyncWork wk, double bias) {
  const int f = blockIdx.y;
  const int line = blockIdx.x;
  if (line >= line_counts[f]) return;
  WaveCtx w{spans[f], span_len[f], (int)threadIdx.x, bias, wk.field_dc ? wk.field_dc[f] : 0.0};
  p = field_params(p, wk, f);
  const HsyncGeom g = hsync_geom(p);
  long long ss = 0;
  const int k = line_kind(p, g, w.n, line, locs[f][line], &ss);
  double cand = 0.0, rp = 0.0;
  const bool ok = (k == 0 && p.use_right) ? right_prelim(w, p, g, ss, &cand, &rp) : false;
  if (w.lane == 0) {
    const int i = f * wk.stride + line;
    wk.kind[i] = (unsigned char)k;
    wk.rt_ok[i] = ok ? 1 : 0;
    wk.rt_cand[i] = cand;
    wk.rt_porch[i] = rp;
  }
}

// Pass 2: every line refined in parallel, assuming the porch level carried in is the
// right-edge porch of the nearest earlier line with a right candidate (-1 if none).
__global__ void __launch_bounds__(64) hsync_spec_kernel(const float* const* spans, const long long* span_len,
                                                        const double* const* locs, const unsigned char* const* filled,
                                                        const int* line_counts, HsyncParams p, HsyncWork wk,
                                                        double* const* refined_out, unsigned char* const* filled_out,
                                                        double bias) {
  const int f = blockIdx.y;
  const int line = blockIdx.x;
  if (line >= line_counts[f]) return;
  WaveCtx w{spans[f], span_len[f], (int)threadIdx.x, bias, wk.field_dc ? wk.field_dc[f] : 0.0};
  p = field_params(p, wk, f);
  const HsyncGeom g = hsync_geom(p);
  const int base = f * wk.stride;
  const int k = wk.kind[base + line];
  const double cur = locs[f][line];
  if (k != 0) {
    if (w.lane == 0) {
      refined_out[f][line] = cur;
      filled_out[f][line] = k == 1 ? 1 : filled[f][line];
    }
    return;
  }
  double p_in = -1.0;
  for (int l = line - 1; l >= 0; l--) {
    const int i = base + l;
    if (wk.kind[i] == 0 && wk.rt_ok[i] && wk.rt_porch[i] > 0.0) {
      p_in = wk.rt_porch[i];
      break;
    }
  }
  long long ss = iround(cur - g.one_us);
  double loc = 0.0, pout = 0.0;
  unsigned char fil = 0;
  refine_line(w, p, g, ss, cur, filled[f][line], p_in, wk.rt_ok[base + line] != 0, wk.rt_cand[base + line],
              wk.rt_porch[base + line], &loc, &fil, &pout);
  if (w.lane == 0) {
    refined_out[f][line] = loc;
    filled_out[f][line] = fil;
    wk.p_used[base + line] = p_in;
    wk.p_out[base + line] = pout;
  }
}

// Pass 3: one wavefront per field checks the carried porch level in line order.  A
// line whose speculative porch matches the true one keeps its parallel result.  The
// check runs 64 lines at a time: the true porch carried into a line is the porch
// carried out of the previous refined line, so every line before the first mismatch
// is verified at once; from the first mismatch on the wave walks the lines in order,
// refining again (with the true porch, exactly as the sequential reference would)
// every line whose speculation was wrong.
__global__ void __launch_bounds__(64) hsync_verify_kernel(const float* const* spans, const long long* span_len,
                                                          const double* const* locs, const unsigned char* const* filled,
                                                          const int* line_counts, HsyncParams p, HsyncWork wk,
                                                          double* const* refined_out,
                                                          unsigned char* const* filled_out, double bias) {
  const int f = blockIdx.x;
  const int lines = line_counts[f];
  WaveCtx w{spans[f], span_len[f], (int)threadIdx.x, bias, wk.field_dc ? wk.field_dc[f] : 0.0};
  p = field_params(p, wk, f);
  const HsyncGeom g = hsync_geom(p);
  const int base = f * wk.stride;
  // Alternate a parallel check (64 lines at a time: the first line at or after `from`
  // whose speculative porch differs from the carried one) with an in-order refinement
  // of that line, which fixes the porch carried on.
  double carry = -1.0;  // porch carried into line `from`
  int from = 0;
  int redone = 0;
  while (from < lines) {
    int bad_line = lines;
    double bad_pc = 0.0;
    for (int c0 = from; c0 < lines; c0 += 64) {
      const int l = c0 + w.lane;
      const bool valid = l < lines;
      const bool is0 = valid && wk.kind[base + l] == 0;
      const double pu = is0 ? wk.p_used[base + l] : 0.0;
      const double po = is0 ? wk.p_out[base + l] : 0.0;
      const unsigned long long mask = __ballot(is0);
      const unsigned long long below = mask & ((1ull << w.lane) - 1ull);
      const double prev_out = __shfl(po, below ? 63 - __clzll((long long)below) : 0);
      const double expected = below ? prev_out : carry;
      const bool bad = is0 && __double_as_longlong(pu) != __double_as_longlong(expected);
      const unsigned long long bm = __ballot(bad);
      if (bm) {
        const int first = __ffsll((long long)bm) - 1;
        bad_line = c0 + first;
        bad_pc = __shfl(expected, first);
        break;
      }
      if (mask) carry = __shfl(po, 63 - __clzll((long long)mask));
    }
    if (bad_line >= lines) break;
    const double cur = locs[f][bad_line];
    const long long ss = iround(cur - g.one_us);
    double loc = 0.0, pout = 0.0;
    unsigned char fil = 0;
    refine_line(w, p, g, ss, cur, filled[f][bad_line], bad_pc, wk.rt_ok[base + bad_line] != 0,
                wk.rt_cand[base + bad_line], wk.rt_porch[base + bad_line], &loc, &fil, &pout);
    redone++;
    if (w.lane == 0) {
      refined_out[f][bad_line] = loc;
      filled_out[f][bad_line] = fil;
    }
    carry = pout;
    from = bad_line + 1;
  }
  if (w.lane == 0) wk.fb_line[f] = redone;
}

// ------------------------------------------------------------------ TBC resample
// (SampleSinc helpers in tbc_sinc.hpp)
__global__ void __launch_bounds__(256) tbc_resample_kernel(const float* src, long long src_len, double bias,
                                                           const double* locs,
                                                           const double* level_adjust, int level_per_line,
                                                           TbcParams p, const float* __restrict__ lut,
                                                           unsigned short* out_u16, double* out_f64) {
  const long long d = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long dest_len = (long long)p.out_len * p.line_count;
  if (d >= dest_len) return;
  const long long prefix = (long long)p.first_line * p.out_len;
  const long long si = prefix + d;
  const int left = (int)(si / p.out_len);
  const double ll = locs[left];
  const double delta = locs[left + 1] - ll;
  const double scale = p.nominal / p.out_len;
  const double sp = (double)si * scale;
  double frac = (sp - (left * p.nominal)) / p.nominal;
  frac = frac < 0.0 ? 0.0 : (frac > 1.0 ? 1.0 : frac);
  double pos;
  {
#pragma clang fp contract(off)
    pos = ll + (delta * frac);
  }
  const double adj = level_per_line ? level_adjust[left] : level_adjust[si];
  double r;
  {
#pragma clang fp contract(off)
    r = (double)(float)(sample_sinc(src, src_len, bias, pos, lut) * adj);
  }
  if (out_f64) out_f64[d] = r;
  if (p.has_conv && out_u16) {
    double sd;
    {
#pragma clang fp contract(off)
      sd = p.out_scale * (r - p.ire0);
    }
    const double v = __fma_rn(sd, p.inv_hz_ire, p.conv_offset);
    unsigned short o;
    if (v <= 0.0)
      o = 0;
    else if (v >= 65535.0)
      o = 65535;
    else
      o = (unsigned short)v;
    out_u16[d] = o;
  }
}

// ------------------------------------------------------------------ batched field TBC
// Per-field wow level adjusts (TbcLineResampler.BuildLinearLevelAdjusts +
// SmoothLevelAdjusts, Tbc/TbcLineResampler.cs:818-878, 997-1011): per-line
// derivative factors, median/MAD outlier clamp, then the exponential smoothing
// y[i] = fma(x[i] - y[i-1], alpha, y[i-1]) over all output samples.  Within a line
// x is constant, so the smoothed value is x + (y_before - x) * (1 - alpha)^(s+1);
// the per-line entry values y_before are carried sequentially (one step per line)
// and the per-sample powers come from a host table.
constexpr int kLevelThreads = 512;

__device__ double block_median_512(double* buf, int n, double pad) {
  // bitonic sort of 512 entries (n <= 512 valid, the rest padded with +inf)
  const int t = threadIdx.x;
  __syncthreads();
  buf[t] = t < n ? buf[t] : pad;
  __syncthreads();
  for (int k = 2; k <= kLevelThreads; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      const int ixj = t ^ j;
      if (ixj > t) {
        const double a = buf[t], b = buf[ixj];
        const bool up = (t & k) == 0;
        if ((a > b) == up) {
          buf[t] = b;
          buf[ixj] = a;
        }
      }
      __syncthreads();
    }
  }
  const int m = n / 2;
  const double med = (n & 1) ? buf[m] : (buf[m - 1] + buf[m]) / 2.0;
  __syncthreads();
  return med;
}

// line_decay = (1 - alpha)^out_len, the smoothing carry over one whole line; 0 disables smoothing.
// Lines: the output lines plus, with first_lines, the field's lines before its first
// output line (TbcLineResampler's level adjusts run from line 0 of the locations,
// Tbc/TbcLineResampler.cs:818-878).  line_x / line_y0 hold line_stride entries per field.
__global__ void __launch_bounds__(kLevelThreads) tbc_levels_kernel(const double* locs, int nloc, int out_lines,
                                                                   const int* first_lines, int line_stride,
                                                                   double nominal, double line_decay, double* line_x,
                                                                   double* line_y0, const int* field_count) {
  __shared__ double buf[kLevelThreads];
  __shared__ double fac[kLevelThreads];
  __shared__ int nan_flag;
  const int f = blockIdx.x;
  if (field_count && f >= *field_count) return;  // uniform across the workgroup
  const int lines = out_lines + (first_lines ? first_lines[f] : 0);
  const int t = threadIdx.x;
  const double* loc = locs + (long long)f * nloc;
  const double inv = 1.0 / nominal;
  if (t == 0) nan_flag = 0;
  __syncthreads();
  double v = 0.0;
  if (t < lines) {
    int left = t < nloc - 2 ? t : nloc - 2;
    {
#pragma clang fp contract(off)
      v = (loc[left + 1] * inv) - (loc[left] * inv);
    }
    if (v != v) atomicOr(&nan_flag, 1);
    fac[t] = v;
  }
  buf[t] = v;
  const double med = block_median_512(buf, lines, INFINITY);
  double dv = 0.0;
  if (t < lines) dv = fabs(fac[t] - med);
  buf[t] = dv;
  const double mad = block_median_512(buf, lines, INFINITY);
  const bool any_nan = nan_flag != 0;
  const double m0 = any_nan ? NAN : med;
  const double mad0 = any_nan ? NAN : mad;
  const double thr = mad0 > 0.0 ? 15.0 * mad0 : 0.001;
  if (t < lines) {
    const double x = fabs(fac[t] - m0) > thr ? m0 : fac[t];
    line_x[(long long)f * line_stride + t] = x;
    fac[t] = x;
  }
  __syncthreads();
  if (t == 0) {
    double* y0 = line_y0 + (long long)f * line_stride;
    if (line_decay > 0.0) {
      // entry value of each line; line 0's first sample is the unsmoothed factor
      double prev = fac[0];
      for (int l = 0; l < lines; l++) {
        y0[l] = prev;
        prev = fac[l] + (prev - fac[l]) * line_decay;
      }
    } else {
      for (int l = 0; l < lines; l++) y0[l] = fac[l];
    }
  }
}

struct FieldTbcParams {
  int out_len;
  int lines;        // output lines per field
  int nloc;         // line locations per field (at least first line + lines + 1)
  int line_stride;  // level entries per field
  int relative;     // 1: locations count from the field's origin (no origin subtraction)
  double nominal;
  int smooth;       // 1: apply the smoothing powers table
  double ire0;
  double out_scale;
  double inv_hz_ire;
  double conv_offset;
  float inv_out_len;  // 1 / out_len (FP32), the estimate of each sample's output line
};

// Stage for the cooperative table-row fetch: per wave, 64 pixels x 8 float4 (two rows),
// slot s of pixel p at p * 8 + (s ^ (p & 7)) so that the 16-byte accesses of eight
// consecutive pixels fall on distinct banks; then the wave's source window (256 floats).

__global__ void __launch_bounds__(256) tbc_fields_kernel(const float* src, long long src_len, double bias,
                                                         const long long* origins, const double* locs,
                                                         const int* first_lines, const double* line_x,
                                                         const double* line_y0, const double* pow_table,
                                                         FieldTbcParams p, const float* __restrict__ lut,
                                                         unsigned short* out, const int* out_index,
                                                         const int* field_count) {
  __shared__ float4 rows[256 / 64][kTbcStage];
  const int f = blockIdx.y;
  if (field_count && f >= *field_count) return;
  const int dest_len = p.out_len * p.lines;
  const int d0 = blockIdx.x * blockDim.x + threadIdx.x;  // a field has < 2^31 output samples
  // every lane takes part in the row fetch; lanes past the field repeat its last sample
  const bool valid = d0 < dest_len;
  const int dd = valid ? d0 : dest_len - 1;
  // output sample dd is resampled sample d of the field, counted from its line 0
  // (LinearLineLocationInterpolator.FillOutputPositions with the output's first line)
  const int d = dd + (first_lines ? first_lines[f] * p.out_len : 0);
  const double* loc = locs + (long long)f * p.nloc;
  // line of sample d: an FP32 estimate (d < 2^24 is exact in FP32, so it is off by at
  // most one) and one correction
  int left = (int)((float)d * p.inv_out_len);
  int smp = d - left * p.out_len;
  if (smp < 0) {
    left -= 1;
    smp += p.out_len;
  } else if (smp >= p.out_len) {
    left += 1;
    smp -= p.out_len;
  }
  const double ll = loc[left];
  const double delta = loc[left + 1] - ll;
  const double scale = p.nominal / p.out_len;
  const double sp = (double)d * scale;
  double frac = (sp - (left * p.nominal)) / p.nominal;
  frac = frac < 0.0 ? 0.0 : (frac > 1.0 ? 1.0 : frac);
  double pos, adj;
  const double x = line_x[(long long)f * p.line_stride + left];
  const long long org = origins ? origins[f] : 0;
  {
#pragma clang fp contract(off)
    pos = ll + (delta * frac);
    if (!p.relative) pos -= (double)org;
    adj = p.smooth ? x + (line_y0[(long long)f * p.line_stride + left] - x) * pow_table[smp] : x;
  }
  const bool finite = isfinite(pos);
  const SincPhase ph = finite ? sinc_phase(pos) : SincPhase{0, 0.0f, 0, 0};
  // SampleSinc's two table rows, fetched cooperatively: in each load, four lanes read
  // the four 16-byte quarters of one pixel's row, so a wave load touches 16 rows
  // instead of 64; the rows then reach their pixels' lanes through LDS.
  const int lane = threadIdx.x & 63;
  float4* stage = rows[threadIdx.x >> 6];
  const float4* lut4 = reinterpret_cast<const float4*>(lut);
  {
    const int quarter = lane & 3, group = lane >> 2;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int pix = group + 16 * k;
      const int r0 = __shfl(ph.a0, pix), r1 = __shfl(ph.a1, pix);
      const float4 c0 = lut4[r0 * 4 + quarter];
      const float4 c1 = lut4[r1 * 4 + quarter];
      stage[tbc_slot(pix, quarter)] = c0;
      stage[tbc_slot(pix, 4 + quarter)] = c1;
    }
  }
  // The wave's 64 pixels take their 16 taps from a few hundred consecutive source
  // samples: one coalesced load of the 256 samples from the first lane's first tap
  // replaces most of the per-lane tap loads.  A lane whose taps leave that window (span
  // edges, a jump between line locations) reads the plane directly.
  const float* sp_src = src + org;
  const long long sp_len = src_len - org;
  const long long s0 = (long long)ph.ci - 7;
  const long long base = __shfl(s0, 0);
  float* win = reinterpret_cast<float*>(stage + 64 * 8);
#pragma unroll
  for (int i = 0; i < kTbcWindow / 64; i++) {
    const int w = lane + 64 * i;
    const long long g = base + w;
    win[w] = (g >= 0 && g < sp_len) ? sp_src[g] : 0.0f;
  }
  __syncthreads();
  float ws[16], we[16];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const float4 a = stage[tbc_slot(lane, q)], b = stage[tbc_slot(lane, 4 + q)];
    ws[4 * q] = a.x; ws[4 * q + 1] = a.y; ws[4 * q + 2] = a.z; ws[4 * q + 3] = a.w;
    we[4 * q] = b.x; we[4 * q + 1] = b.y; we[4 * q + 2] = b.z; we[4 * q + 3] = b.w;
  }
  if (!valid) return;
  double r;
  {
#pragma clang fp contract(off)
    const long long rel = s0 - base;
    const bool in_window = sp_len >= 16 && s0 >= 0 && s0 <= sp_len - 16 && rel >= 0 && rel <= kTbcWindow - 16;
    const double sv = !finite ? 0.0
                      : in_window ? sinc_sum_window(win + (int)rel, bias, ph, ws, we)
                                  : sinc_sum(sp_src, sp_len, bias, ph, ws, we);
    r = (double)(float)(sv * adj);
  }
  double sd;
  {
#pragma clang fp contract(off)
    sd = p.out_scale * (r - p.ire0);
  }
  const double v = __fma_rn(sd, p.inv_hz_ire, p.conv_offset);
  unsigned short o;
  if (v <= 0.0)
    o = 0;
  else if (v >= 65535.0)
    o = 65535;
  else
    o = (unsigned short)v;
  out[(long long)(out_index ? out_index[f] : f) * dest_len + dd] = o;
}

}  // namespace vhsd

extern "C" {

hipError_t vhsd_launch_pulses(const float* s, long long n, double bias, double high, int min_len, int max_len,
                              int* tile_counts,
                              int n_tiles_cap, long long* edge_pos, unsigned char* edge_kind, int edge_cap,
                              int* counters, int* starts, int* lengths, int capacity, hipStream_t stream) {
  const int tiles = (int)((n + vhsd::kPulseTile - 1) / vhsd::kPulseTile);
  if (tiles > n_tiles_cap) return hipErrorInvalidValue;
  if (tiles == 0) return hipMemsetAsync(counters, 0, 2 * sizeof(int), stream);
  hipLaunchKernelGGL(vhsd::pulse_count_kernel, dim3(tiles), dim3(vhsd::kPulseThreads), 0, stream, s, n, bias, high,
                     tile_counts);
  hipLaunchKernelGGL(vhsd::pulse_write_kernel, dim3(tiles), dim3(vhsd::kPulseThreads), 0, stream, s, n, bias, high,
                     tile_counts, edge_pos, edge_kind, edge_cap);
  hipLaunchKernelGGL(vhsd::pulse_pair_kernel, dim3(1), dim3(1024), 0, stream, edge_pos, edge_kind, tile_counts, tiles,
                     edge_cap, min_len, max_len, starts, lengths, capacity, counters);
  return hipGetLastError();
}

hipError_t vhsd_launch_hsync(const float* const* spans, const long long* span_len, const double* const* locs,
                             const unsigned char* const* filled, const int* line_counts, const vhsd::HsyncParams* p,
                             double* const* refined_out, unsigned char* const* filled_out, int fields, int max_lines,
                             double bias, void* work, const double* field_threshold, const double* field_dc,
                             hipStream_t stream) {
  if (fields <= 0 || max_lines <= 0) return hipSuccess;
  // work layout: kind, rt_ok [fields*max_lines bytes each], rt_cand, rt_porch, p_used, p_out
  // [fields*max_lines doubles each], fb_line [fields ints], fb_porch [fields doubles]
  const size_t nl = (size_t)fields * max_lines;
  unsigned char* base = reinterpret_cast<unsigned char*>(work);
  vhsd::HsyncWork wk;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    unsigned char* q = base + off;
    off += (bytes + 255) & ~(size_t)255;
    return q;
  };
  wk.kind = take(nl);
  wk.rt_ok = take(nl);
  wk.rt_cand = reinterpret_cast<double*>(take(nl * 8));
  wk.rt_porch = reinterpret_cast<double*>(take(nl * 8));
  wk.p_used = reinterpret_cast<double*>(take(nl * 8));
  wk.p_out = reinterpret_cast<double*>(take(nl * 8));
  wk.fb_line = reinterpret_cast<int*>(take((size_t)fields * 4));
  wk.fb_porch = reinterpret_cast<double*>(take((size_t)fields * 8));
  wk.stride = max_lines;
  wk.field_threshold = field_threshold;
  wk.field_dc = field_dc;
  const dim3 lg(max_lines, fields);
  hipLaunchKernelGGL(vhsd::hsync_prelim_kernel, lg, dim3(64), 0, stream, spans, span_len, locs, line_counts, *p, wk,
                     bias);
  hipLaunchKernelGGL(vhsd::hsync_spec_kernel, lg, dim3(64), 0, stream, spans, span_len, locs, filled, line_counts, *p,
                     wk, refined_out, filled_out, bias);
  hipLaunchKernelGGL(vhsd::hsync_verify_kernel, dim3(fields), dim3(64), 0, stream, spans, span_len, locs, filled,
                     line_counts, *p, wk, refined_out, filled_out, bias);
  return hipGetLastError();
}

size_t vhsd_hsync_work_bytes(int fields, int max_lines) {
  const size_t nl = (size_t)fields * max_lines;
  auto r = [](size_t b) { return (b + 255) & ~(size_t)255; };
  return 2 * r(nl) + 4 * r(nl * 8) + r((size_t)fields * 4) + r((size_t)fields * 8);
}

// The per-field wow level adjusts alone (tbc_levels_kernel), from each field's first line.
hipError_t vhsd_launch_tbc_levels(const double* locs, int nloc, int lines, const int* first_lines, int line_stride,
                                  double nominal, double line_decay, double* line_x, double* line_y0,
                                  const int* field_count, int fields, hipStream_t stream) {
  if (fields <= 0) return hipSuccess;
  if (line_stride > vhsd::kLevelThreads || nloc < 2 || line_stride > nloc - 1) return hipErrorInvalidValue;
  hipLaunchKernelGGL(vhsd::tbc_levels_kernel, dim3(fields), dim3(vhsd::kLevelThreads), 0, stream, locs, nloc, lines,
                     first_lines, line_stride, nominal, line_decay, line_x, line_y0, field_count);
  return hipGetLastError();
}

hipError_t vhsd_launch_tbc_fields(const float* src, long long src_len, double bias, const long long* origins,
                                  const double* refined, int nloc, int lines, const int* first_lines,
                                  int max_first_line, int relative,
                                  int fields, double nominal, double line_decay, const double* pow_table,
                                  double* line_x, double* line_y0, int out_len, int smooth, double ire0,
                                  double out_scale, double inv_hz_ire, double conv_offset, const float* lut,
                                  unsigned short* out, const int* out_index, const int* field_count,
                                  hipStream_t stream) {
  if (fields <= 0) return hipSuccess;
  const int line_stride = lines + (first_lines ? max_first_line : 0);
  if (line_stride > vhsd::kLevelThreads || nloc < 2 || line_stride > nloc - 1 || max_first_line < 0)
    return hipErrorInvalidValue;
  hipLaunchKernelGGL(vhsd::tbc_levels_kernel, dim3(fields), dim3(vhsd::kLevelThreads), 0, stream, refined, nloc,
                     lines, first_lines, line_stride, nominal, smooth ? line_decay : 0.0, line_x, line_y0,
                     field_count);
  vhsd::FieldTbcParams p{out_len,    lines,      nloc,        line_stride, relative,  nominal,
                         smooth,     ire0,       out_scale,   inv_hz_ire,  conv_offset, 1.0f / (float)out_len};
  const long long dest = (long long)out_len * lines;
  hipLaunchKernelGGL(vhsd::tbc_fields_kernel, dim3((unsigned)((dest + 255) / 256), fields), dim3(256), 0, stream, src,
                     src_len, bias, origins, refined, first_lines, line_x, line_y0, pow_table, p, lut, out,
                     out_index, field_count);
  return hipGetLastError();
}

hipError_t vhsd_launch_tbc(const float* src, long long src_len, double bias, const double* locs, const double* level_adjust,
                           int level_per_line, const vhsd::TbcParams* p, const float* lut, unsigned short* out_u16,
                           double* out_f64, hipStream_t stream) {
  const long long dest = (long long)p->out_len * p->line_count;
  if (dest <= 0) return hipSuccess;
  const int blocks = (int)((dest + 255) / 256);
  hipLaunchKernelGGL(vhsd::tbc_resample_kernel, dim3(blocks), dim3(256), 0, stream, src, src_len, bias, locs,
                     level_adjust,
                     level_per_line, *p, lut, out_u16, out_f64);
  return hipGetLastError();
}

}  // extern "C"

// ------------------------------------------------------------------ pulses without allocation
// vhsdecode_hip_find_pulses_noalloc: the pulse kernels above with their tile counts and
// edge list in a file-scope device array instead of context memory, for callers that
// cannot allocate (code under graph capture, tight serving loops).  The array is one
// per device; its device address comes from hipGetSymbolAddress once per device.
#include <mutex>

#include "host_common.hpp"
#include "vhsdecode_hip.h"

namespace {
constexpr size_t kNoallocTileBytes = (size_t)(VHSDECODE_HIP_NOALLOC_MAX_SAMPLES / vhsd::kPulseTile) * sizeof(int);
constexpr size_t kNoallocEdgePosBytes = (size_t)VHSDECODE_HIP_NOALLOC_MAX_EDGES * sizeof(long long);
constexpr size_t kNoallocEdgeKindBytes = (size_t)VHSDECODE_HIP_NOALLOC_MAX_EDGES;
constexpr size_t kNoallocScratchBytes = kNoallocTileBytes + kNoallocEdgePosBytes + kNoallocEdgeKindBytes;
static_assert(VHSDECODE_HIP_NOALLOC_MAX_SAMPLES % vhsd::kPulseTile == 0, "whole tiles");
static_assert(kNoallocTileBytes % 256 == 0 && kNoallocEdgePosBytes % 256 == 0, "aligned sub-buffers");
}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

namespace {
constexpr int kNoallocMaxDevices = 64;
std::mutex g_noalloc_mu;
void* g_noalloc_ptr[kNoallocMaxDevices] = {};

// The device address of `scratch` on the current device (looked up once per device).
hipError_t noalloc_scratch(int device, unsigned char** out) {
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  if (!g_noalloc_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_noalloc_ptr[device] = p;
  }
  *out = static_cast<unsigned char*>(g_noalloc_ptr[device]);
  return hipSuccess;
}
}  // namespace

extern "C" int32_t vhsdecode_hip_find_pulses_noalloc(const float* sync_ref, int64_t length, double plane_bias_hz,
                                                     double threshold, int32_t min_length, int32_t max_length,
                                                     int32_t* starts, int32_t* lengths, int32_t capacity,
                                                     int32_t* counts, void* stream) {
  if (length < 0 || length > VHSDECODE_HIP_NOALLOC_MAX_SAMPLES)
    return vhsd_host::set_error(VHSDECODE_HIP_E_CAPACITY,
                                "find_pulses_noalloc: length " + std::to_string(length) + " exceeds the maximum of " +
                                    std::to_string((long long)VHSDECODE_HIP_NOALLOC_MAX_SAMPLES) + " samples");
  if ((!sync_ref && length > 0) || !counts || capacity < 0 || ((!starts || !lengths) && capacity > 0))
    return vhsd_host::set_error(VHSDECODE_HIP_E_INVALID_ARGUMENT, "find_pulses_noalloc: invalid argument");
  int device = 0;
  hipError_t e = hipGetDevice(&device);
  if (e == hipSuccess && (device < 0 || device >= kNoallocMaxDevices)) e = hipErrorInvalidDevice;
  unsigned char* base = nullptr;
  if (e == hipSuccess) e = noalloc_scratch(device, &base);
  if (e != hipSuccess)
    return vhsd_host::set_error(VHSDECODE_HIP_E_HIP, std::string("find_pulses_noalloc: ") + hipGetErrorString(e));
  int* tile_counts = reinterpret_cast<int*>(base);
  long long* edge_pos = reinterpret_cast<long long*>(base + kNoallocTileBytes);
  unsigned char* edge_kind = base + kNoallocTileBytes + kNoallocEdgePosBytes;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int tiles = (int)((length + vhsd::kPulseTile - 1) / vhsd::kPulseTile);
  // Every launch rewrites what the next one reads (tile counts, edges, both counters), so
  // nothing is cleared between calls.  With no tiles the pair kernel alone writes zeros.
  if (tiles > 0) {
    hipLaunchKernelGGL(vhsd::pulse_count_kernel, dim3(tiles), dim3(vhsd::kPulseThreads), 0, st, sync_ref, length,
                       plane_bias_hz, threshold, tile_counts);
    hipLaunchKernelGGL(vhsd::pulse_write_kernel, dim3(tiles), dim3(vhsd::kPulseThreads), 0, st, sync_ref, length,
                       plane_bias_hz, threshold, tile_counts, edge_pos, edge_kind,
                       (int)VHSDECODE_HIP_NOALLOC_MAX_EDGES);
  }
  hipLaunchKernelGGL(vhsd::pulse_pair_kernel, dim3(1), dim3(1024), 0, st, edge_pos, edge_kind, tile_counts, tiles,
                     (int)VHSDECODE_HIP_NOALLOC_MAX_EDGES, min_length, max_length, starts, lengths, capacity, counts);
  e = hipGetLastError();
  if (e != hipSuccess)
    return vhsd_host::set_error(VHSDECODE_HIP_E_HIP, std::string("find_pulses_noalloc: ") + hipGetErrorString(e));
  return VHSDECODE_HIP_OK;
}
