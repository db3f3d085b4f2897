// The reference's VHS field sync path (v0.4.0 profile) for the fields of a demodulated
// span, on the device: per-field level detection (the fallback serration-level search and
// its refinement), the levels' moving-average chain across fields, sync pulses on the
// DC-shifted sync reference, pulse classification and refinement, field parity, line 0,
// the read window of the next field and the upstream line locations.
//
// Reference: src/VHSDecode.Core/Decode/TbcFieldDecodePipeline.cs DecodePrepared
// :1071-1290, PrepareSyncSpan :2776-3190, PrepareSyncSpanFromLevels :3197-3234,
// ResolveLine0Location :5452-5651, level selection :5679-5782, sync history :6119-6375;
// Dsp/SyncAnalysis.cs :131-657; Dsp/LevelDetection.cs :98-407, 445-524;
// Dsp/VsyncSerrationDetector.cs :475-536 (CheckLevels), :727-754 (moving average);
// Dsp/VhsFieldLevelState.cs; Dsp/FieldParityDetector.cs :16-268;
// Dsp/VBlankSyncResolver.cs :317-747; Decode/DecodeReadWindowPlanner.cs :37-63.
//
// Fields run in parallel on read windows speculated from the vertical-sync groups of the
// span (a window is the reference's DecodeReadWindowPlanner.Resolve of the previous
// field's next-field offset, aligned down to a 32768-sample block, so a prediction within
// a few samples of the true offset gives the same window); a one-thread chain kernel then
// walks the fields in order, carries every cross-field state (level moving averages,
// parity cadence, sync history) exactly as the reference does and checks each window
// against the one the previous field implies.  The first field whose window was
// mispredicted ends the span's field list; the caller resumes there.
//
// The VsyncSerrationDetector candidate search is not run: it filters each ~1M-sample
// window with four transfer-function IIR recurrences whose candidates are set by the
// recurrences' own rounding (an accurate evaluation of the same filters picks other
// candidates), so no parallel evaluation reproduces it.  Its measurements are taken as
// absent, which is what the reference measures on a 0.5 MHz sync reference whenever the
// equalizing pulses widen past 1.25 eqPulseUS at the serration level.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <climits>
#include <cmath>

#include "vhs_sync.hpp"
#include "vhsd_device.hpp"

namespace vhsd {
namespace vsync {

constexpr int kT = 1024;   // threads of the per-field workgroups
constexpr int kRun = 16;   // samples per thread per pulse tile
constexpr int kMaxMeans = 256;    // pulse means per level refinement

__device__ __forceinline__ double round_away(double x) { return round(x); }
__device__ __forceinline__ double round_even(double x) { return rint(x); }
__device__ __forceinline__ bool in_range(double v, double lo, double hi) { return v >= lo && v <= hi; }

struct Plane {
  const float* p;
  long long n;
  double bias;
  double dc;
  __device__ __forceinline__ double at(long long i) const {
#pragma clang fp contract(off)
    return ((double)p[i] + bias) + dc;
  }
};

// ------------------------------------------------------------------ workgroup helpers
__device__ __forceinline__ int wave_scan_incl(int v, int lane) {
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_up(v, o);
    if (lane >= o) v += u;
  }
  return v;
}

// Exclusive prefix of v over the workgroup; *total receives the sum.  red: LDS int[16 + 1].
__device__ int wg_exclusive(int v, int* red, int* total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int incl = wave_scan_incl(v, lane);
  if (lane == 63) red[wave] = incl;
  __syncthreads();
  int before = 0, tot = 0;
  for (int w = 0; w < (int)(blockDim.x >> 6); w++) {
    const int x = red[w];
    if (w < wave) before += x;
    tot += x;
  }
  __syncthreads();
  *total = tot;
  return before + incl - v;
}

__device__ double wg_min(double v, double* dred) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o));
  if (lane == 0) dred[wave] = v;
  __syncthreads();
  double m = dred[0];
  for (int w = 1; w < (int)(blockDim.x >> 6); w++) m = fmin(m, dred[w]);
  __syncthreads();
  return m;
}

__device__ long long wg_sum_ll(long long v, long long* lred) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  if (lane == 0) lred[wave] = v;
  __syncthreads();
  long long s = 0;
  for (int w = 0; w < (int)(blockDim.x >> 6); w++) s += lred[w];
  __syncthreads();
  return s;
}

// ------------------------------------------------------------------ pulses
// PulseDetection.FindPulses (Dsp/PulseDetection.cs:45-222) over the plane: a two-state
// machine over `value <= high` (NaN keeps the state), in tiles of kRun samples per
// thread.  The transitions alternate (pulse start, pulse end), so a pulse is an end
// transition paired with the transition before it: within a thread's run that is a local
// transition, and before the run's first one it is the last transition of the earlier
// threads (a workgroup max-scan of each run's last transition) or of the earlier tiles.
// Pulses starting at sample 0 (or with no start transition) are dropped.  Pulses are
// compacted in order; returns the pulse count (every thread) and stores the first `cap`
// into starts/lens; *vsync receives the number of pulses longer than vs_min.
struct PulseScratch {
  int red[kT / 64 + 4];
  int wlast[kT / 64];
  int offset;
  int carry;  // position of the last transition before this tile, -1 for none
};

__device__ __forceinline__ int state_before(const Plane& x, long long i, double high) {
  for (long long k = i; k >= 0; k--) {
    const float v = x.p[k];
    if (v == v) return x.at(k) <= high ? 1 : 0;
  }
  return 0;
}

// Largest v over the workgroup's threads before this one (exclusive), -1 for none.
__device__ int wg_exclusive_max(int v, PulseScratch* ps) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int incl = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_up(incl, o);
    if (lane >= o) incl = incl > u ? incl : u;
  }
  if (lane == 63) ps->wlast[wave] = incl;
  __syncthreads();
  int before = -1;
  for (int w = 0; w < wave; w++) before = before > ps->wlast[w] ? before : ps->wlast[w];
  const int up = __shfl_up(incl, 1);
  const int excl = lane > 0 ? (before > up ? before : up) : before;
  return excl;
}

__device__ int wg_find_pulses(const Plane& x, double high, int min_len, int max_len, double vs_min, int* starts,
                              int* lens, int cap, PulseScratch* ps, int* vsync) {
  const int tid = threadIdx.x;
  if (tid == 0) {
    ps->offset = 0;
    ps->carry = -1;
  }
  __syncthreads();
  int vs_count = 0;
  for (long long base = 0; base < x.n; base += (long long)blockDim.x * kRun) {
    const long long i0 = base + (long long)tid * kRun;
    unsigned codes = 0;
    if (i0 < x.n) {
      int st = i0 > 0 ? state_before(x, i0 - 1, high) : 0;
#pragma unroll
      for (int k = 0; k < kRun; k++) {
        const long long i = i0 + k;
        if (i < x.n) {
          const float v = x.p[i];
          if (v == v) {
            const int cur = x.at(i) <= high ? 1 : 0;
            if (i >= 1) codes |= ((!st && cur) ? 1u : ((st && !cur) ? 2u : 0u)) << (2 * k);
            st = cur;
          }
        }
      }
    }
    // this run's last transition, and the last one before the run
    const int mine = codes ? (int)(i0 + ((31 - __clz(codes)) >> 1)) : -1;
    int prev = wg_exclusive_max(mine, ps);
    if (prev < 0) prev = ps->carry;
    // pair each end transition with the transition before it
    int cnt = 0;
    unsigned rest = codes;
    int pv = prev;
    while (rest) {
      const int k = (__ffs(rest) - 1) >> 1;
      const int pos = (int)(i0 + k);
      if (((rest >> (2 * k)) & 3u) == 2u && pv > 0) {
        const int len = pos - pv;
        if (in_range((double)len, min_len, max_len)) {
          cnt++;
          if ((double)len > vs_min) vs_count++;
        }
      }
      pv = pos;
      rest &= ~(3u << (2 * k));
    }
    int tot;
    int idx = wg_exclusive(cnt, ps->red, &tot) + ps->offset;
    rest = codes;
    pv = prev;
    while (rest && cnt) {
      const int k = (__ffs(rest) - 1) >> 1;
      const int pos = (int)(i0 + k);
      if (((rest >> (2 * k)) & 3u) == 2u && pv > 0) {
        const int len = pos - pv;
        if (in_range((double)len, min_len, max_len)) {
          if (idx < cap) {
            starts[idx] = pv;
            lens[idx] = len;
          }
          idx++;
        }
      }
      pv = pos;
      rest &= ~(3u << (2 * k));
    }
    // the tile's last transition carries into the next tile
    int last_all;
    {
      const int lane = tid & 63;
      int m = mine;
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        const int u = __shfl_xor(m, o);
        m = m > u ? m : u;
      }
      if (lane == 0) ps->wlast[tid >> 6] = m;
      __syncthreads();
      last_all = -1;
      for (int w = 0; w < (int)(blockDim.x >> 6); w++) last_all = last_all > ps->wlast[w] ? last_all : ps->wlast[w];
    }
    __syncthreads();
    if (tid == 0) {
      ps->offset += tot;
      if (last_all >= 0) ps->carry = last_all;
    }
    __syncthreads();
  }
  const int count = ps->offset;
  int vt;
  (void)wg_exclusive(vs_count, ps->red, &vt);
  if (vsync) *vsync = vt;
  __syncthreads();
  return count;
}

// ------------------------------------------------------------------ pairwise means
// NumpyReduction.MeanFloat64 (Dsp/NumpyReduction.cs:226-227, 400-451) of several ranges:
// numpy's pairwise sum splits a block longer than 128 at half its length rounded down to
// a multiple of 8, sums a leaf with eight accumulators, and adds the halves in order.  One
// wave per range: lane l sums the leaf at path l of the tree cut at depth kMeanDepth (the
// leftmost path under a shallower leaf sums it), then the lanes combine the tree level by
// level in the recursion's order.  Ranges whose tree is deeper than kMeanDepth (longer
// than ~7000 samples) are summed by one lane in recursion order.
constexpr int kMeanDepth = 6;
struct MeanScratch {
  int unused;
};

// Node of the pairwise tree of n values along path bits `path` at `level`; `owner` is true
// for the leftmost path under a leaf at or above that level (the one that sums it).
__device__ __forceinline__ void pw_path(int n, int level, unsigned path, int& off, int& len, bool& owner,
                                        bool& exists) {
  off = 0;
  len = n;
  owner = true;
  exists = true;
  for (int l = 0; l < level; l++) {
    if (len <= 128) {
      owner = (path & ((1u << (level - l)) - 1)) == 0;
      exists = false;
      return;
    }
    const int split = (len >> 1) & ~7;
    if ((path >> (level - 1 - l)) & 1u) {
      off += split;
      len -= split;
    } else {
      len = split;
    }
  }
}

__device__ double leaf_sum(const Plane& x, int s, int n) {
#pragma clang fp contract(off)
  if (n < 8) {
    double r = -0.0;
    for (int i = 0; i < n; i++) r += x.at(s + i);
    return r;
  }
  double r[8];
  for (int k = 0; k < 8; k++) r[k] = x.at(s + k);
  int i = 8;
  const int end = n - (n % 8);
  for (; i < end; i += 8)
    for (int k = 0; k < 8; k++) r[k] += x.at(s + i + k);
  double t = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
  for (; i < n; i++) t += x.at(s + i);
  return t;
}

// the recursion itself, for one lane (deep trees only)
__device__ double pairwise_serial(const Plane& x, int s, int n) {
#pragma clang fp contract(off)
  // explicit stack of pending right halves; partial sums combine as the recursion does
  int st_s[40], st_n[40], st_state[40];
  double st_left[40];
  int sp = 0;
  int cs = s, cn = n;
  double result = 0.0;
  for (;;) {
    while (cn > 128) {
      int split = cn / 2;
      split -= split % 8;
      st_s[sp] = cs + split;
      st_n[sp] = cn - split;
      st_state[sp] = 0;
      sp++;
      cn = split;
    }
    result = leaf_sum(x, cs, cn);
    // unwind: a pending right half gets summed next; a finished one adds to its left
    while (sp > 0 && st_state[sp - 1] == 1) {
      sp--;
      result = st_left[sp] + result;
    }
    if (sp == 0) return result;
    st_state[sp - 1] = 1;
    st_left[sp - 1] = result;
    cs = st_s[sp - 1];
    cn = st_n[sp - 1];
  }
}

// means[k] = mean of x[rs[k], re[k]) for k < count (all threads call)
__device__ void wg_means(const Plane& x, const int* rs, const int* re, int count, double* means, MeanScratch* m) {
#pragma clang fp contract(off)
  (void)m;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int k = wave; k < count; k += (int)(blockDim.x >> 6)) {
    const int s = rs[k], n = re[k] - rs[k];
    double total;
    if (n > 7000) {
      total = lane == 0 ? pairwise_serial(x, s, n) : 0.0;
      total = __shfl(total, 0);
    } else {
      int off, len;
      bool owner, exists;
      pw_path(n, kMeanDepth, (unsigned)lane, off, len, owner, exists);
      // lane l holds the value of the tree node whose slot it is (slot of a node at
      // level L along path t: t << (kMeanDepth - L))
      double val = owner ? leaf_sum(x, s + off, len) : 0.0;
      for (int level = kMeanDepth - 1; level >= 0; level--) {
        const int sh = kMeanDepth - level;
        const double right = __shfl(val, (lane + (1 << (sh - 1))) & 63);
        if ((lane & ((1 << sh) - 1)) == 0) {
          int o2, l2;
          bool ow2, ex2;
          pw_path(n, level, (unsigned)(lane >> sh), o2, l2, ow2, ex2);
          if (ex2 && l2 > 128) val = val + right;
        }
      }
      total = __shfl(val, 0);
    }
    if (lane == 0) means[k] = total / (double)n;
  }
  __syncthreads();
}

// NumpyReduction.MedianFloat64 of v[0, n) (n small; one thread; sorts in place)
__device__ double median_small(double* v, int n) {
  if (n <= 0) return __builtin_nan("");
  for (int i = 0; i < n; i++)
    if (v[i] != v[i]) return v[i];
  for (int i = 1; i < n; i++) {
    const double x = v[i];
    int j = i - 1;
    while (j >= 0 && v[j] > x) {
      v[j + 1] = v[j];
      j--;
    }
    v[j + 1] = x;
  }
  const int mid = n / 2;
  return (n & 1) ? v[mid] : (v[mid - 1] + v[mid]) / 2.0;
}

// ------------------------------------------------------------------ level refinement
struct RefineScratch {
  PulseScratch ps;
  MeanScratch ms;
  int rs[kMaxMeans], re[kMaxMeans];
  double means[kMaxMeans];
  double black[kMaxMeans];
  int nv, nb;
  int loc_first, loc_last, n_locs;
  int idx_count;
  long long lred[kT / 64];
  int flag;
};

// VsyncSerrationDetector.CheckLevels (Dsp/VsyncSerrationDetector.cs:475-536), the window
// counts taken by the workgroup when the early returns do not decide.
__device__ bool wg_check_levels(const Plane& x, double old_sync, double new_sync, double new_blank, double ref_sync,
                                double hz_ire, RefineScratch* r) {
#pragma clang fp contract(off)
  if (x.n <= 0 || hz_ire == 0.0) return false;
  const double diff = (new_blank - new_sync) / hz_ire;
  if ((ref_sync - new_sync) > (hz_ire * 15.0) || diff > 47.0) return false;
  if (new_sync - old_sync < (hz_ire * 5.0)) return true;
  long long bs = 0, bb = 0;
  for (long long i = threadIdx.x; i < x.n; i += blockDim.x) {
    const double v = x.at(i);
    if (v < new_sync) bs++;
    if (v < new_blank) bb++;
  }
  bs = wg_sum_ll(bs, r->lred);
  bb = wg_sum_ll(bb, r->lred);
  const double a = (double)bs / x.n;
  const double b = (double)bb / x.n;
  return !(a > 0.07 || b < 0.005);
}

enum Fail { kNone = 0, kMissing = 1, kNonFinite = 2, kCheckFailed = 3 };

struct RefineOut {
  int ok;
  double sync, blank;
  int fail;
  int has_measured;
  double measured;
};

// LevelDetection.RefineSerrationLevelsFromPulses (Dsp/LevelDetection.cs:172-235) with
// FallbackVsyncLocationMeans (:22-53) and PulsesBlackLevelMeans (:445-524) on a pulse
// list in order (starts strictly increasing).
__device__ RefineOut wg_refine_from_pulses(const Plane& x, const int* starts, const int* lens, int npulses,
                                           const SyncCfg& c, double ref_sync, RefineScratch* r) {
#pragma clang fp contract(off)
  RefineOut o{0, 0.0, 0.0, kMissing, 0, 0.0};
  const double vl = c.vsync_us * c.fs_mhz;
  const double lo = vl * 0.8, hi = vl * 1.2;
  const int off = (int)c.fs_mhz;
  if (threadIdx.x == 0) r->nv = 0;
  __syncthreads();
  // vsync-length pulses, in order: ordered compaction
  for (int b0 = 0; b0 < npulses; b0 += blockDim.x) {
    const int k = b0 + threadIdx.x;
    bool ok = false;
    int s = 0, e = 0;
    if (k < npulses) {
      const int len = lens[k];
      if (!(len <= lo || len >= hi)) {
        s = starts[k] + off;
        e = starts[k] + len - off;
        ok = !(s < 0 || e > x.n || e <= s);
      }
    }
    int tot;
    const int idx = wg_exclusive(ok ? 1 : 0, r->ps.red, &tot) + r->nv;
    if (ok && idx < kMaxMeans) {
      r->rs[idx] = s;
      r->re[idx] = e;
      if (idx == 0) r->loc_first = k;
    }
    __syncthreads();
    if (threadIdx.x == 0) r->nv += tot;
    __syncthreads();
  }
  // the last vsync pulse: the largest k written (recompute exactly)
  if (threadIdx.x == 0) r->flag = -1;
  __syncthreads();
  for (int k = threadIdx.x; k < npulses; k += blockDim.x) {
    const int len = lens[k];
    if (!(len <= lo || len >= hi)) {
      const int s = starts[k] + off, e = starts[k] + len - off;
      if (!(s < 0 || e > x.n || e <= s)) atomicMax(&r->flag, k);
    }
  }
  __syncthreads();
  const int nv = r->nv < kMaxMeans ? r->nv : kMaxMeans;
  if (nv == 0) return o;
  const int loc_last = r->flag;
  wg_means(x, r->rs, r->re, nv, r->means, &r->ms);
  double sync = 0.0;
  if (threadIdx.x == 0) r->black[0] = median_small(r->means, nv);
  __syncthreads();
  sync = r->black[0];
  o.has_measured = 1;
  o.measured = sync;
  // PulsesBlackLevelMeans: indices around the vsync pulses
  const int last_index = npulses - 1;
  int before = r->loc_first, after = loc_last;
  if (nv != 12) {
    while (before > 1 && starts[before] - starts[before - 1] < 600) before--;
    // pulse starts increase, so starts[after] - starts[after + 1] < 600 holds to the end
    after = last_index;
  }
  if (threadIdx.x == 0) r->nb = 0;
  __syncthreads();
  const double fs = c.fs_mhz;
  // candidates: [max(before - 5, 1), before) if before > 1, then [after + 1, max(after + 6, last_index))
  const int a0 = before > 1 ? (before - 5 > 1 ? before - 5 : 1) : 0;
  const int a1 = before > 1 ? before : 0;
  const int b0i = after < last_index - 1 ? after + 1 : 0;
  const int b1i = after < last_index - 1 ? (after + 6 > last_index ? after + 6 : last_index) : 0;
  const int na = a1 - a0, nb = b1i - b0i;
  for (int base = 0; base < na + nb; base += blockDim.x) {
    const int q = base + threadIdx.x;
    bool ok = false;
    int s = 0, e = 0;
    if (q < na + nb) {
      const int i = q < na ? a0 + q : b0i + (q - na);
      if (i >= 0 && i < npulses && in_range((double)lens[i], fs * 0.75, fs * 3.0)) {
        long long ss = (long long)(starts[i] + (fs * 5.0));
        ss = ss < 0 ? 0 : (ss > x.n ? x.n : ss);
        long long ee = (long long)(starts[i] + (fs * 20.0));
        ee = ee < ss ? ss : (ee > x.n ? x.n : ee);
        ok = ee > ss;
        s = (int)ss;
        e = (int)ee;
      }
    }
    int tot;
    const int idx = wg_exclusive(ok ? 1 : 0, r->ps.red, &tot) + r->nb;
    if (ok && idx < kMaxMeans) {
      r->rs[idx] = s;
      r->re[idx] = e;
    }
    __syncthreads();
    if (threadIdx.x == 0) r->nb += tot;
    __syncthreads();
  }
  const int nbm = r->nb < kMaxMeans ? r->nb : kMaxMeans;
  if (nbm == 0) {
    o.fail = kNonFinite;
    return o;
  }
  wg_means(x, r->rs, r->re, nbm, r->black, &r->ms);
  if (threadIdx.x == 0) r->means[0] = median_small(r->black, nbm);
  __syncthreads();
  const double blank = r->means[0];
  __syncthreads();
  if (!isfinite(sync) || !isfinite(blank) || blank < sync) {
    o.fail = kNonFinite;
    return o;
  }
  if (nv <= 3 || !wg_check_levels(x, ref_sync, sync, blank, ref_sync, c.hz_ire, r)) {
    o.fail = kCheckFailed;
    return o;
  }
  o.ok = 1;
  o.fail = kNone;
  o.sync = sync;
  o.blank = blank;
  return o;
}

// LevelDetection.RefineSerrationLevels (:98-170): pulses at the initial levels' midpoint
__device__ RefineOut wg_refine_levels(const Plane& x, double s0, double b0, const SyncCfg& c, double ref_sync,
                                      int* pstart, int* plen, int pcap, RefineScratch* r, int* overflow) {
#pragma clang fp contract(off)
  if (x.n <= 0) return RefineOut{0, 0.0, 0.0, kMissing, 0, 0.0};
  const double thr = (s0 + b0) / 2.0;
  const double eql = (double)c.ser_eq_len, ll = (double)c.ser_line_len;
  const int mn = (int)ceil(eql / 8.0) > 0 ? (int)ceil(eql / 8.0) : 0;
  const int mx = (int)floor(ll * 5.0) > 1 ? (int)floor(ll * 5.0) : 1;
  const int np = wg_find_pulses(x, thr, mn, mx, 1e300, pstart, plen, pcap, &r->ps, nullptr);
  if (np > pcap && threadIdx.x == 0) *overflow = 1;
  return wg_refine_from_pulses(x, pstart, plen, np < pcap ? np : pcap, c, ref_sync, r);
}

// ------------------------------------------------------------------ segmented pulse search
// The pulse searches over whole read windows split into segments of kSegLen samples, one
// 256-thread workgroup each (one tile of kRun samples a thread), so the fields' windows
// keep every CU busy and a workgroup waits for one round of loads.  A segment's
// summary holds its first transition (and whether it ends a pulse), its last transition
// and the pulses (in range) whose start and end both lie in it; combining the summaries
// in order adds the pulses that cross a segment boundary (a segment's first transition
// ending a pulse whose start is the last transition before the segment) and gives each
// segment its output offset and the transition carried into it, with which a second
// pass writes the pulses in order.  Counts, lists and pairings are those of
// wg_find_pulses over the whole window.
constexpr int kSegT = 256;
constexpr int kSegTiles = 1;
static_assert(kSegLen == kSegT * kRun * kSegTiles, "segment length");

struct SegScratch {
  int red[kSegT / 64 + 4];
  int wlast[kSegT / 64];
  int wfirst[kSegT / 64];
};

template <int NT>
__device__ int wgt_exclusive(int v, int* red, int* total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int incl = wave_scan_incl(v, lane);
  if (lane == 63) red[wave] = incl;
  __syncthreads();
  int before = 0, tot = 0;
  for (int w = 0; w < NT / 64; w++) {
    const int x = red[w];
    if (w < wave) before += x;
    tot += x;
  }
  __syncthreads();
  *total = tot;
  return before + incl - v;
}

// One thread's kRun samples of a segment and the last sample before them that is not
// NaN (NaN when there is none): loaded once for every threshold the segment is searched at.
struct SegRun {
  float v[kRun];
  float before;
};

__device__ __forceinline__ void seg_load(const Plane& x, long long i0, long long s1, SegRun* run) {
  run->before = __builtin_nanf("");
  if (i0 >= s1) return;
  if (i0 + kRun <= s1 && (reinterpret_cast<uintptr_t>(x.p + i0) & 15) == 0) {
    const float4* p4 = reinterpret_cast<const float4*>(x.p + i0);
#pragma unroll
    for (int q = 0; q < kRun / 4; q++) {
      const float4 v = p4[q];
      run->v[4 * q] = v.x;
      run->v[4 * q + 1] = v.y;
      run->v[4 * q + 2] = v.z;
      run->v[4 * q + 3] = v.w;
    }
  } else {
#pragma unroll
    for (int k = 0; k < kRun; k++) run->v[k] = i0 + k < s1 ? x.p[i0 + k] : __builtin_nanf("");
  }
  for (long long k = i0 - 1; k >= 0; k--) {
    const float v = x.p[k];
    if (v == v) {
      run->before = v;
      break;
    }
  }
}

// Segment [s0, s1) (one tile: s1 - s0 <= kSegT * kRun) of the window at threshold
// `high`: its summary (every thread), and with `starts` the pulses paired from
// `carry_in` on, written from `out_base` (below cap).  The samples come from seg_load.
__device__ SegSummary seg_scan(const Plane& x, const SegRun& run, long long s0, long long s1, double high,
                               int min_len, int max_len, double vs_min, int carry_in, int* starts, int* lens, int cap,
                               int out_base, SegScratch* ps) {
#pragma clang fp contract(off)
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  int n_in = 0, n_vs = 0;
  int first_end = 0;
  const long long i0 = s0 + (long long)tid * kRun;
  unsigned codes = 0;
  if (i0 < s1) {
    // x.at(i) <= high as one FP32 comparison per sample
    const float bound = pass_bound_f32(x.bias, x.dc, high);
    const float bv = run.before;
    int st = (i0 > 0 && bv == bv) ? (bv <= bound ? 1 : 0) : 0;
#pragma unroll
    for (int k = 0; k < kRun; k++) {
      const long long i = i0 + k;
      const float v = run.v[k];
      if (i < s1 && v == v) {
        const int cur = v <= bound ? 1 : 0;
        if (i >= 1) codes |= ((!st && cur) ? 1u : ((st && !cur) ? 2u : 0u)) << (2 * k);
        st = cur;
      }
    }
  }
  const int mine = codes ? (int)(i0 + ((31 - __clz(codes)) >> 1)) : -1;
  const int myfirst = codes ? (int)(i0 + ((__ffs(codes) - 1) >> 1)) : INT_MAX;
  // the transition before this run: the latest of the earlier threads', or the carry
  int incl = mine;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_up(incl, o);
    if (lane >= o) incl = incl > u ? incl : u;
  }
  int fmin = myfirst;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const int u = __shfl_xor(fmin, o);
    fmin = fmin < u ? fmin : u;
  }
  if (lane == 63) ps->wlast[wave] = incl;
  if (lane == 0) ps->wfirst[wave] = fmin;
  __syncthreads();
  int before = carry_in;
  for (int w = 0; w < wave; w++) before = before > ps->wlast[w] ? before : ps->wlast[w];
  const int up = __shfl_up(incl, 1);
  const int prev = lane > 0 ? (before > up ? before : up) : before;
  int tile_last = carry_in, tile_first = INT_MAX;
  for (int w = 0; w < kSegT / 64; w++) {
    tile_last = tile_last > ps->wlast[w] ? tile_last : ps->wlast[w];
    tile_first = tile_first < ps->wfirst[w] ? tile_first : ps->wfirst[w];
  }
  int cnt = 0;
  unsigned rest = codes;
  int pv = prev;
  while (rest) {
    const int k = (__ffs(rest) - 1) >> 1;
    const int pos = (int)(i0 + k);
    const bool end = ((rest >> (2 * k)) & 3u) == 2u;
    if (pos == tile_first) first_end = end ? 1 : 0;
    if (end && pv > 0) {
      const int len = pos - pv;
      if (in_range((double)len, min_len, max_len)) {
        cnt++;
        if (!starts && (double)len > vs_min) n_vs++;
      }
    }
    pv = pos;
    rest &= ~(3u << (2 * k));
  }
  n_in += cnt;
  if (starts) {
    int tot = 0;
    int idx = wgt_exclusive<kSegT>(cnt, ps->red, &tot) + out_base;
    rest = codes;
    pv = prev;
    while (rest && cnt) {
      const int k = (__ffs(rest) - 1) >> 1;
      const int pos = (int)(i0 + k);
      if (((rest >> (2 * k)) & 3u) == 2u && pv > 0) {
        const int len = pos - pv;
        if (in_range((double)len, min_len, max_len)) {
          if (idx < cap) {
            starts[idx] = pv;
            lens[idx] = len;
          }
          idx++;
        }
      }
      pv = pos;
      rest &= ~(3u << (2 * k));
    }
  }
  // reduce the per-thread counts (count mode) and the first transition's kind
  int tot_in, tot_vs, tot_fe;
  (void)wgt_exclusive<kSegT>(n_in, ps->red, &tot_in);
  (void)wgt_exclusive<kSegT>(n_vs, ps->red, &tot_vs);
  (void)wgt_exclusive<kSegT>(first_end, ps->red, &tot_fe);
  SegSummary sm;
  sm.first_pos = tile_first == INT_MAX ? -1 : tile_first;
  sm.first_end = tot_fe > 0 ? 1 : 0;
  sm.last_pos = tile_first == INT_MAX ? -1 : tile_last;
  sm.n_in = tot_in;
  sm.n_vs = tot_vs;
  __syncthreads();
  return sm;
}

// Pulse search jobs over the fields' windows: thresholds [field][R], the DC added to the
// plane, length bounds and summaries [field][R][segments].
struct SegJob {
  const double* thr;
  int R;
  int r0, nr;              // thresholds [r0, r0 + nr) of a count launch
  const double* dc;        // per field, or null
  int min_len, max_len;
  double vs_min;
  const int* active;       // per field: run the job (null: every field)
  SegSummary* sum;
  int segs;                // segments per window
  const int* choice;       // per field: the threshold index the list is written at
  int* seg_carry;          // [field][segments]
  int* seg_off;
  int* count;              // per field: pulses of the list
  int* raw_count;          // per field, or null: the count capped at the list capacity (overflow flagged)
  int* starts;             // [field][pulse_cap]
  int* lens;
};

__device__ __forceinline__ Plane job_plane(const SyncCfg& c, const SyncWork& w, const SegJob& j, int f) {
  return Plane{w.sync_ref + w.win_start[f], w.win_len[f], c.bias, j.dc ? j.dc[f] : 0.0};
}

// A segment's summaries at thresholds [j.r0, j.r0 + j.nr), its samples loaded once.
__global__ void __launch_bounds__(kSegT) seg_count_kernel(SyncCfg c, SyncWork w, SegJob j) {
  __shared__ SegScratch ps;
  const int sg = blockIdx.x, f = blockIdx.z;
  if (f >= *w.spec_count || (j.active && !j.active[f])) return;
  const Plane x = job_plane(c, w, j, f);
  const long long s0 = (long long)sg * kSegLen;
  const long long s1 = s0 + kSegLen < x.n ? s0 + kSegLen : x.n;
  SegRun run;
  if (s0 < x.n) seg_load(x, s0 + (long long)threadIdx.x * kRun, s1, &run);
  for (int r = j.r0; r < j.r0 + j.nr; r++) {
    SegSummary sm{-1, 0, -1, 0, 0};
    if (s0 < x.n)
      sm = seg_scan(x, run, s0, s1, j.thr[(long long)f * j.R + r], j.min_len, j.max_len, j.vs_min, -1, nullptr,
                    nullptr, 0, 0, &ps);
    if (threadIdx.x == 0) j.sum[((long long)f * j.R + r) * j.segs + sg] = sm;
  }
}

// Segment summaries of one threshold, in order, on one wave: the pulse count, the
// long-pulse count and (with seg_off) each segment's offset and the transition carried
// into it.  A segment's boundary pulse depends only on the last transition of the
// nearest earlier segment that has one, so that transition comes from a max-scan over
// (segment, last transition) and the counts from a prefix sum, 64 segments at a time.
__device__ void wave_seg_combine(const SegJob& j, int f, int r, int* np, int* nvs, int* seg_carry, int* seg_off) {
  const int lane = threadIdx.x & 63;
  long long carry_key = -1;  // (segment << 32) | last transition of the latest segment with one
  int count = 0, vs = 0;
  const SegSummary* sm = j.sum + ((long long)f * j.R + r) * j.segs;
  for (int b0 = 0; b0 < j.segs; b0 += 64) {
    const int sg = b0 + lane;
    SegSummary a{-1, 0, -1, 0, 0};
    if (sg < j.segs) a = sm[sg];
    const bool has = a.first_pos >= 0;
    long long incl = has ? (((long long)sg << 32) | (unsigned)a.last_pos) : -1;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const long long u = __shfl_up(incl, o);
      if (lane >= o) incl = incl > u ? incl : u;
    }
    long long excl = __shfl_up(incl, 1);
    if (lane == 0) excl = -1;
    excl = excl > carry_key ? excl : carry_key;
    const int last = excl < 0 ? -1 : (int)(excl & 0xffffffffll);
    int c = 0, v = 0;
    if (has) {
      if (a.first_end && last > 0) {
        const int len = a.first_pos - last;
        if (in_range((double)len, j.min_len, j.max_len)) {
          c = 1;
          if ((double)len > j.vs_min) v = 1;
        }
      }
      c += a.n_in;
      v += a.n_vs;
    }
    const int ci = wave_scan_incl(c, lane), vi = wave_scan_incl(v, lane);
    if (seg_off && sg < j.segs) {
      seg_carry[sg] = last;
      seg_off[sg] = count + ci - c;
    }
    count += __shfl(ci, 63);
    vs += __shfl(vi, 63);
    const long long top = __shfl(incl, 63);
    carry_key = top > carry_key ? top : carry_key;
  }
  *np = count;
  if (nvs) *nvs = vs;
}

// One threshold per field (j.choice, or 0), one wave a field: the list's segment offsets
// and its count.
__global__ void __launch_bounds__(64) seg_plan_kernel(SyncCfg c, SyncWork w, SegJob j, int max_fields) {
  const int f = blockIdx.x;
  if (f >= max_fields || f >= *w.spec_count || (j.active && !j.active[f])) return;
  const int r = j.choice ? j.choice[f] : 0;
  if (r < 0) return;
  int np;
  wave_seg_combine(j, f, r, &np, nullptr, j.seg_carry + (long long)f * j.segs, j.seg_off + (long long)f * j.segs);
  if (threadIdx.x == 0) {
    j.count[f] = np;
    if (j.raw_count) {
      if (np > c.pulse_cap) w.overflow[f] = 1;
      j.raw_count[f] = np < c.pulse_cap ? np : c.pulse_cap;
    }
  }
}

__global__ void __launch_bounds__(kSegT) seg_write_kernel(SyncCfg c, SyncWork w, SegJob j) {
  __shared__ SegScratch ps;
  const int sg = blockIdx.x, f = blockIdx.y;
  if (f >= *w.spec_count || (j.active && !j.active[f])) return;
  const int r = j.choice ? j.choice[f] : 0;
  if (r < 0) return;
  const Plane x = job_plane(c, w, j, f);
  const long long s0 = (long long)sg * kSegLen;
  if (s0 >= x.n) return;
  const long long s1 = s0 + kSegLen < x.n ? s0 + kSegLen : x.n;
  const long long fo = (long long)f * j.segs + sg;
  SegRun run;
  seg_load(x, s0 + (long long)threadIdx.x * kRun, s1, &run);
  (void)seg_scan(x, run, s0, s1, j.thr[(long long)f * j.R + r], j.min_len, j.max_len, 1e300, j.seg_carry[fo],
                 j.starts + (long long)f * c.pulse_cap, j.lens + (long long)f * c.pulse_cap, c.pulse_cap,
                 j.seg_off[fo], &ps);
}

// ------------------------------------------------------------------ S1: per-field level search
// LevelDetection.SearchFallbackSerrationLevels (:265-396, divisor 1, no long-pulse check)
// then RefineSerrationLevels at the found levels: both depend only on the window.
// The whole search of a field on one workgroup: the fields whose threshold sequence runs
// past the kSpecR thresholds evaluated at once (it runs to 30 on windows without usable
// sync).
__global__ void __launch_bounds__(kT) level_search_kernel(SyncCfg c, SyncWork w) {
#pragma clang fp contract(off)
  __shared__ RefineScratch r;
  __shared__ double dred[kT / 64];
  const int f = blockIdx.x;
  if (f >= *w.spec_count || !w.seg.fallback[f]) return;
  const Plane x{w.sync_ref + w.win_start[f], w.win_len[f], c.bias, 0.0};
  int* pst = w.pulse_start + (long long)f * c.pulse_cap;
  int* pln = w.pulse_len + (long long)f * c.pulse_cap;
  int* ovf = w.overflow + f;
  LevelSearch out{};
  // window minimum (Math.Min keeps a NaN)
  double mn = INFINITY;
  bool nan = false;
  for (long long i = threadIdx.x; i < x.n; i += kT) {
    const double v = x.at(i);
    if (v != v) nan = true;
    mn = fmin(mn, v);
  }
  mn = wg_min(mn, dred);
  if (__syncthreads_or(nan ? 1 : 0)) mn = __builtin_nan("");
  const double blank_level = c.ire0;  // IreToHz(0)
  const double ref_sync = c.ire0 + (c.hz_ire * c.vsync_ire);
  const double eql = (double)c.ser_eq_len, ll = (double)c.ser_line_len;
  const double min_vsync = (c.vsync_us * c.fs_mhz) * 0.8;
  const int pmn = (int)ceil(eql / 8.0) > 0 ? (int)ceil(eql / 8.0) : 0;
  const int pmx = (int)floor(ll * 5.0) > 1 ? (int)floor(ll * 5.0) : 1;
  double min_sync = mn, prev_min_sync = mn;
  int prev_vsync = 0;
  bool found = false, check_next = true;
  double list_thr = 0.0;  // threshold of the pulse list the refinement reads
  for (int retry = 0; retry < 30; retry++) {
    const double thr = (min_sync + blank_level) / 2.0;
    int vc = 0;
    const int np = wg_find_pulses(x, thr, pmn, pmx, min_vsync, pst, pln, 0, &r.ps, &vc);
    list_thr = thr;
    bool stop = false;
    if (np > 200) {
      if (vc > 4) {
        if (vc == 12 && !check_next) {
          stop = true;
        } else if (!found || vc > prev_vsync) {
          found = true;
          prev_vsync = vc;
          prev_min_sync = min_sync;
          check_next = true;
        } else if (vc < prev_vsync || !check_next) {
          min_sync = prev_min_sync;
          list_thr = (min_sync + blank_level) / 2.0;
          stop = true;
        } else {
          check_next = false;
        }
      }
    }
    if (stop) break;
    // AdvanceFallbackSyncSearchLevel (:398-407)
    const double sync_ire = (min_sync - blank_level) / c.hz_ire;
    const double next_ire = sync_ire + 5.0;
    const double scaled = c.hz_ire * next_ire;
    min_sync = blank_level + scaled;
  }
  const int np = wg_find_pulses(x, list_thr, pmn, pmx, 1e300, pst, pln, c.pulse_cap, &r.ps, nullptr);
  if (np > c.pulse_cap && threadIdx.x == 0) *ovf = 1;
  const RefineOut first = wg_refine_from_pulses(x, pst, pln, np < c.pulse_cap ? np : c.pulse_cap, c, ref_sync, &r);
  out.first_ok = first.ok;
  out.first_sync = first.sync;
  out.first_blank = first.blank;
  out.first_fail = first.fail;
  out.first_has_measured = first.has_measured;
  out.first_measured = first.measured;
  if (first.ok) {
    const RefineOut sec = wg_refine_levels(x, first.sync, first.blank, c, ref_sync, pst, pln, c.pulse_cap, &r, ovf);
    out.ref_ok = sec.ok;
    out.ref_sync = sec.sync;
    out.ref_blank = sec.blank;
    out.ref_fail = sec.fail;
    out.ref_has_measured = sec.has_measured;
    out.ref_measured = sec.measured;
  }
  if (threadIdx.x == 0) w.search[f] = out;
}

// The level search of every field through segment workgroups: the window minimum, the
// counts at the first kSpecR thresholds of the search sequence at once, the sequence's
// decisions (a field whose sequence needs more thresholds goes to level_search_kernel),
// the list at the chosen threshold, the first refinement, the list at its midpoint and
// the second refinement.
__global__ void __launch_bounds__(kSegT) seg_min_kernel(SyncCfg c, SyncWork w) {
  __shared__ double dred[kSegT / 64];
  const int sg = blockIdx.x, f = blockIdx.y;
  if (f >= *w.spec_count) return;
  const Plane x{w.sync_ref + w.win_start[f], w.win_len[f], c.bias, 0.0};
  const long long s0 = (long long)sg * kSegLen;
  const long long s1 = s0 + kSegLen < x.n ? s0 + kSegLen : x.n;
  double mn = INFINITY;
  bool nan = false;
  for (long long i = s0 + threadIdx.x; i < s1; i += kSegT) {
    const double v = x.at(i);
    if (v != v) nan = true;
    mn = fmin(mn, v);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int o = 32; o > 0; o >>= 1) mn = fmin(mn, __shfl_xor(mn, o));
  if (lane == 0) dred[wave] = mn;
  const int any_nan = __syncthreads_or(nan ? 1 : 0);
  if (threadIdx.x == 0) {
    double m = dred[0];
    for (int k = 1; k < kSegT / 64; k++) m = fmin(m, dred[k]);
    w.seg.seg_min[(long long)f * w.seg.segs + sg] = m;
    w.seg.seg_nan[(long long)f * w.seg.segs + sg] = any_nan;
  }
}

__device__ __forceinline__ double next_min_sync(const SyncCfg& c, double min_sync) {
#pragma clang fp contract(off)
  // AdvanceFallbackSyncSearchLevel (LevelDetection.cs:398-407)
  const double blank = c.ire0;
  const double sync_ire = (min_sync - blank) / c.hz_ire;
  const double next_ire = sync_ire + 5.0;
  const double scaled = c.hz_ire * next_ire;
  return blank + scaled;
}

// The window minimum from its segments' (one wave a field; the minimum of finite values
// does not depend on the order, and a NaN anywhere makes it NaN as Math.Min does), then
// the level search's threshold sequence.
__global__ void __launch_bounds__(64) search_thresholds_kernel(SyncCfg c, SyncWork w, int max_fields) {
#pragma clang fp contract(off)
  const int f = blockIdx.x;
  if (f >= max_fields || f >= *w.spec_count) return;
  const long long n = w.win_len[f];
  const int segs = (int)((n + kSegLen - 1) / kSegLen);
  double mn = INFINITY;
  bool nan = false;
  for (int sg = threadIdx.x; sg < segs; sg += 64) {
    mn = fmin(mn, w.seg.seg_min[(long long)f * w.seg.segs + sg]);
    nan |= w.seg.seg_nan[(long long)f * w.seg.segs + sg] != 0;
  }
  for (int o = 32; o > 0; o >>= 1) mn = fmin(mn, __shfl_xor(mn, o));
  nan = __ballot(nan) != 0;
  if (threadIdx.x != 0) return;
  if (nan) mn = __builtin_nan("");
  double min_sync = mn;
  for (int r = 0; r < kSpecR; r++) {
    w.seg.thr[(long long)f * kSpecR + r] = (min_sync + c.ire0) / 2.0;
    min_sync = next_min_sync(c, min_sync);
  }
  w.seg.fallback[f] = 0;
  w.seg.pending[f] = 0;
}

// SearchFallbackSerrationLevels' retry decisions (LevelDetection.cs:265-396) over the
// counts of the first `avail` thresholds, one wave a threshold's counts.  A search that
// has not decided within them is pending (its next thresholds are counted and it is
// decided again over all kSpecR); one that has not decided within kSpecR falls back to
// the whole search on one workgroup.
__global__ void __launch_bounds__(64 * kSpecR) search_decide_kernel(SyncCfg c, SyncWork w, SegJob j, int avail,
                                                                     int only_pending) {
  __shared__ int cnt[kSpecR], vcs[kSpecR];
  const int f = blockIdx.x;
  if (f >= *w.spec_count || (only_pending && !w.seg.pending[f])) return;
  const int r = threadIdx.x >> 6;
  if (r < avail) {
    int np, vc;
    wave_seg_combine(j, f, r, &np, &vc, nullptr, nullptr);
    if ((threadIdx.x & 63) == 0) {
      cnt[r] = np;
      vcs[r] = vc;
    }
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  int prev_vsync = 0, prev_r = 0;
  bool found = false, check_next = true;
  int choice = -1;
  for (int k = 0; k < avail && choice < 0; k++) {
    const int np = cnt[k], vc = vcs[k];
    if (np > 200 && vc > 4) {
      if (vc == 12 && !check_next) {
        choice = k;
      } else if (!found || vc > prev_vsync) {
        found = true;
        prev_vsync = vc;
        prev_r = k;
        check_next = true;
      } else if (vc < prev_vsync || !check_next) {
        choice = prev_r;
      } else {
        check_next = false;
      }
    }
  }
  const bool more = choice < 0 && avail < kSpecR;
  w.seg.pending[f] = more ? 1 : 0;
  w.seg.choice[f] = choice;
  w.seg.fallback[f] = (choice < 0 && !more) ? 1 : 0;
  w.seg.active[f] = choice < 0 ? 0 : 1;
  w.seg.active2[f] = 0;
}

// The first refinement on the chosen list, then the second refinement's threshold.
__global__ void __launch_bounds__(kT) refine_first_kernel(SyncCfg c, SyncWork w) {
#pragma clang fp contract(off)
  __shared__ RefineScratch r;
  const int f = blockIdx.x;
  if (f >= *w.spec_count || !w.seg.active[f]) return;
  const Plane x{w.sync_ref + w.win_start[f], w.win_len[f], c.bias, 0.0};
  const int* pst = w.pulse_start + (long long)f * c.pulse_cap;
  const int* pln = w.pulse_len + (long long)f * c.pulse_cap;
  const int np = w.seg.count[f] < c.pulse_cap ? w.seg.count[f] : c.pulse_cap;
  const double ref_sync = c.ire0 + (c.hz_ire * c.vsync_ire);
  const RefineOut first = wg_refine_from_pulses(x, pst, pln, np, c, ref_sync, &r);
  if (threadIdx.x == 0) {
    LevelSearch out{};
    out.first_ok = first.ok;
    out.first_sync = first.sync;
    out.first_blank = first.blank;
    out.first_fail = first.fail;
    out.first_has_measured = first.has_measured;
    out.first_measured = first.measured;
    w.search[f] = out;
    // RefineSerrationLevels (:98-170): pulses at the first levels' midpoint
    w.seg.thr2[f] = (first.sync + first.blank) / 2.0;
    w.seg.active2[f] = first.ok ? 1 : 0;
  }
}

__global__ void __launch_bounds__(kT) refine_second_kernel(SyncCfg c, SyncWork w) {
#pragma clang fp contract(off)
  __shared__ RefineScratch r;
  const int f = blockIdx.x;
  if (f >= *w.spec_count || !w.seg.active2[f]) return;
  const Plane x{w.sync_ref + w.win_start[f], w.win_len[f], c.bias, 0.0};
  const int* pst = w.pulse_start + (long long)f * c.pulse_cap;
  const int* pln = w.pulse_len + (long long)f * c.pulse_cap;
  const int np = w.seg.count2[f] < c.pulse_cap ? w.seg.count2[f] : c.pulse_cap;
  const double ref_sync = c.ire0 + (c.hz_ire * c.vsync_ire);
  const RefineOut sec = wg_refine_from_pulses(x, pst, pln, np, c, ref_sync, &r);
  if (threadIdx.x == 0) {
    LevelSearch out = w.search[f];
    out.ref_ok = sec.ok;
    out.ref_sync = sec.sync;
    out.ref_blank = sec.blank;
    out.ref_fail = sec.fail;
    out.ref_has_measured = sec.has_measured;
    out.ref_measured = sec.measured;
    w.search[f] = out;
  }
}


// ------------------------------------------------------------------ S2: level chain
// The moving averages of VsyncSerrationDetector (window 2, watermark 1, unbounded pushes
// trimmed on pull) and VhsFieldLevelState (window 10, bounded pushes), kept by every
// thread of the one workgroup (identical inputs, identical branches) so the rare
// refinement from field-state levels can run on the whole workgroup.
struct Avg {
  // values in registers: every loop runs over the whole capacity with compile-time indices
  double v[kAvgCap];
  int n;
  int window;
  int watermark;
  bool bounded;
  __device__ bool has_values() const { return n > watermark; }
  __device__ void shift_out() {  // drop the oldest value
#pragma unroll
    for (int i = 0; i + 1 < kAvgCap; i++) v[i] = v[i + 1];
    n--;
  }
  __device__ void push(double x) {
    if (bounded && n == window) shift_out();
    if (n == kAvgCap) shift_out();  // unbounded lists never exceed a few values between pulls
#pragma unroll
    for (int i = 0; i < kAvgCap; i++)
      if (i == n) v[i] = x;
    n++;
  }
  __device__ bool pull(double* out) {
#pragma clang fp contract(off)
    if (n == 0) return false;
    while (n > window) shift_out();
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < kAvgCap; i++)
      if (i < n) s += v[i];
    *out = s / (double)n;
    return true;
  }
};

struct LevelChain {
  Avg ser_s, ser_b, fs_s, fs_b;
  int has_last;
  double last_s, last_b;

  __device__ void load(const LevelStateDev& s) {
    ser_s.n = s.ser_n;
    ser_b.n = s.ser_n;
    fs_s.n = s.fs_sync_n;
    fs_b.n = s.fs_blank_n;
#pragma unroll
    for (int i = 0; i < kAvgCap; i++) {
      ser_s.v[i] = s.ser_sync[i];
      ser_b.v[i] = s.ser_blank[i];
      fs_s.v[i] = s.fs_sync[i];
      fs_b.v[i] = s.fs_blank[i];
    }
    has_last = s.has_last;
    last_s = s.last_sync;
    last_b = s.last_blank;
  }
  __device__ void store(LevelStateDev* s) const {
    s->ser_n = ser_s.n;
    s->fs_sync_n = fs_s.n;
    s->fs_blank_n = fs_b.n;
#pragma unroll
    for (int i = 0; i < kAvgCap; i++) {
      s->ser_sync[i] = ser_s.v[i];
      s->ser_blank[i] = ser_b.v[i];
      s->fs_sync[i] = fs_s.v[i];
      s->fs_blank[i] = fs_b.v[i];
    }
    s->has_last = has_last;
    s->last_sync = last_s;
    s->last_blank = last_b;
  }
  __device__ bool ser_pull(double* s, double* b) {
    double a = 0.0, c2 = 0.0;
    const bool hs = ser_s.pull(&a);
    const bool hb = ser_b.pull(&c2);
    *s = a;
    *b = c2;
    return hs && hb;
  }
  __device__ void ser_push(double s, double b) {
    ser_s.push(s);
    ser_b.push(b);
  }
  __device__ bool ser_has() const { return ser_s.has_values() && ser_b.has_values(); }
  __device__ bool fs_has() const { return fs_s.has_values() && fs_b.has_values(); }
  __device__ void fs_push(double s, double b) {
    fs_b.push(b);
    fs_s.push(s);
  }
  __device__ bool fs_pull(double* s, double* b) {
    double bb;
    if (!fs_b.pull(&bb)) return false;
    double ss;
    if (!fs_s.pull(&ss)) return false;
    *s = ss;
    *b = bb;
    return true;
  }
  // TryResolveVhsSerrationRefinementFallback (TbcFieldDecodePipeline.cs:5679-5704)
  __device__ bool fieldstate_fallback(int fail, bool* has, double* s, double* b) {
    *has = false;
    if (fail != kNonFinite && fail != kMissing) return false;
    double tmp;
    bool has_sync = false;
    if (fail == kMissing) has_sync = fs_s.pull(&tmp);
    *has = fs_pull(s, b);
    if (fail == kMissing && !has_sync) return false;
    return true;
  }
};

// 256 threads: the moving averages stay in registers (every thread keeps them); the
// workgroup helpers run at any workgroup size.
constexpr int kChainT = 256;
__global__ void __launch_bounds__(kChainT) level_chain_kernel(SyncCfg c, SyncWork w) {
#pragma clang fp contract(off)
  __shared__ RefineScratch r;
  LevelChain ch;
  ch.ser_s = Avg{{}, 0, 2, 1, false};
  ch.ser_b = Avg{{}, 0, 2, 1, false};
  ch.fs_s = Avg{{}, 0, c.fs_window, 0, true};
  ch.fs_b = Avg{{}, 0, c.fs_window, 0, true};
  ch.load(*w.level_state_in);
  const int nf = *w.spec_count;
  const double ref_sync = c.ire0 + (c.hz_ire * c.vsync_ire);
  for (int f = 0; f < nf; f++) {
    if (threadIdx.x == 0) ch.store(&w.level_before[f]);
    const Plane x{w.sync_ref + w.win_start[f], w.win_len[f], c.bias, 0.0};
    const LevelSearch sr = w.search[f];
    // VsyncSerrationDetector.Analyze with no measurement: pull (trims), HasLevels
    double ps = 0.0, pb = 0.0;
    const bool s_ok = ch.ser_pull(&ps, &pb);
    const bool has_levels = ch.ser_has();
    const bool has_pending = has_levels && s_ok;
    // fallback search (first pass from the level-search kernel)
    bool has_first = sr.first_ok != 0;
    double fs_ = sr.first_sync, fb = sr.first_blank;
    if (sr.first_has_measured) ch.fs_s.push(sr.first_measured);
    bool from_search = has_first;
    if (!has_first) {
      bool hl;
      double ls, lb;
      if (ch.fieldstate_fallback(sr.first_fail, &hl, &ls, &lb)) {
        has_first = hl;
        fs_ = ls;
        fb = lb;
      }
    }
    bool pushed = false;
    if (has_first) {
      RefineOut rr;
      if (from_search) {
        rr = RefineOut{sr.ref_ok, sr.ref_sync, sr.ref_blank, sr.ref_fail, sr.ref_has_measured, sr.ref_measured};
      } else {
        int* pst = w.pulse_start + (long long)f * c.pulse_cap;
        int* pln = w.pulse_len + (long long)f * c.pulse_cap;
        rr = wg_refine_levels(x, fs_, fb, c, ref_sync, pst, pln, c.pulse_cap, &r, w.overflow + f);
      }
      if (rr.has_measured) ch.fs_s.push(rr.measured);
      if (rr.ok) {
        ch.fs_push(rr.sync, rr.blank);
        ch.ser_push(rr.sync, rr.blank);
        pushed = true;
      } else {
        bool hl;
        double ls, lb;
        if (ch.fieldstate_fallback(rr.fail, &hl, &ls, &lb) && hl) {
          ch.ser_push(ls, lb);
          pushed = true;
        }
      }
    }
    int source = 0;
    double ss = 0.0, sb = 0.0;
    bool done = false;
    if (pushed) {
      bool ok;
      if (ch.ser_has()) {
        double a, b;
        const bool hp = ch.ser_pull(&a, &b);
        ok = false;
        if (hp) {
          if (wg_check_levels(x, ref_sync, a, b, ref_sync, c.hz_ire, &r)) {
            ss = a;
            sb = b;
            ok = true;
          } else {
            ok = ch.fs_pull(&ss, &sb);
          }
        }
      } else {
        ok = ch.fs_pull(&ss, &sb);
      }
      if (ok) {
        source = 2;
        done = true;
      }
    }
    if (!done && has_pending) {
      if (wg_check_levels(x, ref_sync, ps, pb, ref_sync, c.hz_ire, &r)) {
        ss = ps;
        sb = pb;
        source = 1;
        done = true;
      } else if (ch.fs_pull(&ss, &sb)) {
        source = 1;
        done = true;
      } else if (ch.ser_has() && !ch.fs_has()) {
        ss = c.ire0;
        sb = ref_sync;
        source = 4;
        done = true;
      }
    }
    if (!done) {
      double ts, tb;
      bool has_term = false;
      if (ch.fs_has() && ch.fs_pull(&ts, &tb)) has_term = true;
      else if (ch.has_last) {
        ts = ch.last_s;
        tb = ch.last_b;
        has_term = true;
      }
      if (has_term) {
        ss = ts;
        sb = tb;
        source = 3;
        done = true;
      }
    }
    FieldLevels lv{};
    if (done) {
      if (source != 4) {
        ch.has_last = 1;
        ch.last_s = ss;
        ch.last_b = sb;
      }
      lv.source = source;
      lv.sync = ss;
      lv.blank = sb;
      const double dc = c.ire0 - sb;
      if (!isfinite(dc) || dc == 0.0) {
        lv.dc = 0.0;
        lv.threshold = (ss + sb) / 2.0;
      } else {
        lv.dc = dc;
        lv.threshold = ((ss + dc) + (sb + dc)) / 2.0;
      }
    } else {
      lv.source = 0;
      lv.dc = 0.0;
      lv.threshold = c.ire0 + (c.hz_ire * (c.vsync_ire / 2.0));
    }
    if (threadIdx.x == 0) {
      w.levels[f] = lv;
      // the field's pulse search runs at this threshold on the reference shifted by dc
      w.seg.thr3[f] = lv.threshold;
      w.seg.dc3[f] = lv.dc;
    }
  }
  if (threadIdx.x == 0) ch.store(w.level_state_after_all);
}

// ------------------------------------------------------------------ S4: per-field classification
enum Kind { kHSync = 0, kEq = 1, kVSync = 2, kEq2 = 3 };

struct Timing {
  double hsync_lo, hsync_hi, eq_lo, eq_hi, vs_lo, vs_hi, median;
};

// The order-quality windows of RefinePulses (SyncAnalysis.cs:286-300) as whole sample
// distances: the quotient distance / line length rounds monotonically in the distance, so
// each bound is the first (last) distance whose quotient passes it.  Windows: 0 neither
// pulse hsync, 1 both hsync, 2 one hsync.
struct QualityWindows {
  int lo[3], hi[3];
};

__device__ long long first_distance_at_least(double lo, double nominal) {
#pragma clang fp contract(off)
  long long d = (long long)floor(lo * nominal);
  while ((double)(d - 1) / nominal >= lo) d--;
  while (!((double)d / nominal >= lo)) d++;
  return d;
}

__device__ long long last_distance_at_most(double hi, double nominal) {
#pragma clang fp contract(off)
  long long d = (long long)ceil(hi * nominal);
  while (!((double)d / nominal <= hi)) d--;
  while ((double)(d + 1) / nominal <= hi) d++;
  return d;
}

__device__ QualityWindows quality_windows(double nominal) {
  QualityWindows q;
  const double lo[3] = {0.4, 0.9, 0.4}, hi[3] = {0.6, 1.1, 1.1};
  const bool finite = nominal > 0.0 && nominal < 1e9;
  for (int k = 0; k < 3; k++) {
    // the start distances are whole ints, so the bounds clamp to the int range
    const long long a = finite ? first_distance_at_least(lo[k], nominal) : LLONG_MAX;
    const long long b = finite ? last_distance_at_most(hi[k], nominal) : LLONG_MIN;
    q.lo[k] = a > INT_MAX ? INT_MAX : (a < INT_MIN ? INT_MIN : (int)a);
    q.hi[k] = b > INT_MAX ? INT_MAX : (b < INT_MIN ? INT_MIN : (int)b);
  }
  return q;
}

__device__ __forceinline__ bool quality_int(const QualityWindows& q, int prev_kind, int prev_start, int kind,
                                            int start) {
  const bool ph = prev_kind == kHSync, h = kind == kHSync;
  const int lo = (!ph && !h) ? q.lo[0] : ((ph && h) ? q.lo[1] : q.lo[2]);
  const int hi = (!ph && !h) ? q.hi[0] : ((ph && h) ? q.hi[1] : q.hi[2]);
  const int d = start - prev_start;
  return d >= lo && d <= hi;
}

__device__ __forceinline__ double round_half_line(double v) {
#pragma clang fp contract(off)
  const double h = round_away(v / 0.5) * 0.5;
  return round_away(h * 10.0) / 10.0;
}

struct Bounds {
  double first_hsync, first_eq2, last_hsync, last_eq2;
};

__device__ Bounds field_bounds(int pal, bool first) {
  if (!pal) return first ? Bounds{1.0, 0.5, 0.5, 1.0} : Bounds{0.5, 1.0, 1.0, 0.5};
  return first ? Bounds{0.5, 0.5, 1.0, 1.0} : Bounds{1.0, 1.0, 0.5, 0.5};
}

__device__ double first_hsync_line(int pal, int num_pulses, bool first) {
#pragma clang fp contract(off)
  const Bounds b = field_bounds(pal, first);
  const double section = num_pulses / 2.0;
  return b.first_hsync + (section * 3.0) - 0.5 + b.first_eq2;
}

struct Acc {
  double loc, off;
  int n;
};

__device__ Acc acc_pair(double fs_, double fl, double ss, double sl, double mll, double fhl) {
#pragma clang fp contract(off)
  const double actual = (fs_ - ss) / mll;
  const double expected = fl - sl;
  if (actual <= expected - 0.5 || actual >= expected + 0.5) return Acc{0.0, 0.0, 0};
  return Acc{ss + (mll * (fhl - sl)), actual - expected, 1};
}

__device__ Acc acc_add(Acc a, Acc b) {
#pragma clang fp contract(off)
  return Acc{a.loc + b.loc, a.off + b.off, a.n + b.n};
}

// The markers found, in marker order, with their lines (partial marker list of
// EstimateLine0FromTransitions).  Held in registers: every access names its slot.
struct MarkerList {
  double s0, s1, s2, s3, l0, l1, l2, l3;
  int n;
  __device__ __forceinline__ void push(double smp, double line) {
    switch (n) {
      case 0: s0 = smp; l0 = line; break;
      case 1: s1 = smp; l1 = line; break;
      case 2: s2 = smp; l2 = line; break;
      default: s3 = smp; l3 = line; break;
    }
    n++;
  }
  __device__ __forceinline__ double smp(int i) const { return i == 0 ? s0 : (i == 1 ? s1 : (i == 2 ? s2 : s3)); }
  __device__ __forceinline__ double line(int i) const { return i == 0 ? l0 : (i == 1 ? l1 : (i == 2 ? l2 : l3)); }
};

__device__ __forceinline__ MarkerList partial_markers(const double* smp, const int* has, double eq1_line,
                                                      double section) {
#pragma clang fp contract(off)
  const double lines[4] = {eq1_line, eq1_line + section, eq1_line + (section * 2.0),
                           eq1_line + (section * 3.0) - 0.5};
  MarkerList m{0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0};
#pragma unroll
  for (int i = 0; i < 4; i++)
    if (has[i]) m.push(smp[i], lines[i]);
  return m;
}

// The k-th pair (i < j) of n items in the order of the nested loops over i, then j.
__device__ __forceinline__ void pair_ij(int k, int n, int* i, int* j) {
  int a = 0;
  while (k >= n - 1 - a) {
    k -= n - 1 - a;
    a++;
  }
  *i = a;
  *j = a + 1 + k;
}

__device__ __forceinline__ double readlane_d(double v, int l) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
  return __hiloint2double(hi, lo);
}

// acc_add over lanes [from, to) of the wave, in lane order, from zero
__device__ __forceinline__ Acc acc_lanes(const Acc& p, int from, int to) {
#pragma clang fp contract(off)
  Acc a{0.0, 0.0, 0};
  for (int l = from; l < to; l++) {
    a.loc = a.loc + readlane_d(p.loc, l);
    a.off = a.off + readlane_d(p.off, l);
    a.n = a.n + __builtin_amdgcn_readlane(p.n, l);
  }
  return a;
}

// The pulse lists of one field: in LDS up to kPulseLds pulses, else in the field's
// global lists (a window full of noise).  eff_*: a raw pulse after TryRescueLongHSync as
// often as it applies; craw/ceff: ClassifyPulses' kind of the raw and the rescued length.
struct LogicScratch {
  int* rs;            // raw pulses
  int* rl;
  int* es;            // rescued pulses
  int* el;
  signed char* craw;  // classify(raw length)
  signed char* ceff;  // classify(rescued length)
  unsigned char* eflag;    // bit 0: hsync length, bit 1: rescue length (rescue failed), bit 2: eq length
  int* cs;            // refined pulses
  int* cl;
  unsigned char* ck;  // kind
  unsigned char* co;  // in order
  int lrs[kPulseLds], lrl[kPulseLds], les[kPulseLds], lel[kPulseLds], lcs[kPulseLds], lcl[kPulseLds];
  signed char lcraw[kPulseLds], lceff[kPulseLds];
  unsigned char leflag[kPulseLds], lck[kPulseLds], lco[kPulseLds];
  int hist[kHistBins];
  int nref;
  Timing t;
  double mll;
  // parallel reductions
  int red[kT / 64];
  long long lred[kT / 64];
  int g0;                   // first refined pulse of the second group
  int slot[4];              // consensus transitions: last refined index per (group, kind)
  int mark[8];              // line-0 markers: last refined index per (group, marker)
  unsigned long long run;   // (size << 32) | (~start): the longest hsync run, first on ties
  // FinalizeEstimate jobs: parity x (first, last, combined)
  Acc acc[6];
  double fhl[2];
  Estimate est[6];
};

__device__ __forceinline__ int classify(int len, const Timing& t) {
  if (in_range(len, t.hsync_lo, t.hsync_hi)) return kHSync;
  if (in_range(len, t.eq_lo, t.eq_hi)) return kEq;
  return in_range(len, t.vs_lo, t.vs_hi) ? kVSync : -1;
}


// TryRescueLongHSync (:316-338) on one wave: the first run below (first sample - step)
// inside [s, e) that ends inside it, has a length the search keeps and does not start at
// s.  The samples come 64 at a time; the run boundaries are walked on the ballots.
__device__ bool wave_rescue(const Plane& x, long long s, long long e, double step, int mn, int mx, int* rs,
                            int* rl) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x & 63;
  const double thr = x.at(s) - step;
  bool inp = x.at(s) <= thr;
  long long cur = 0;
  for (long long base = 0; base < e - s; base += 64) {
    const long long p = base + lane;
    const bool valid = p < e - s;
    const double v = valid ? x.at(s + p) : 0.0;
    const unsigned long long below = __ballot(valid && v <= thr);
    const unsigned long long above = __ballot(valid && v > thr);
    int pos = 0;
    for (;;) {
      unsigned long long m = inp ? above : below;
      m = pos >= 64 ? 0ull : (m & (~0ull << pos));
      if (!m) break;
      const int b = __ffsll((long long)m) - 1;
      if (inp) {
        const long long len = base + b - cur;
        if (in_range((double)len, mn, mx) && cur != 0) {
          *rs = (int)(s + cur);
          *rl = (int)len;
          return true;
        }
        inp = false;
      } else {
        cur = base + b;
        inp = true;
      }
      pos = b + 1;
    }
  }
  return false;
}

// RefinePulses (SyncAnalysis.cs:243-314) on one wave: its control flow is uniform, 64
// pulses' flags and rescued pulses sit in the lanes and are read by lane index, lane 0
// writes the refined pulses.
template <bool kLds>
__device__ __forceinline__ void refine_walk(LogicScratch& L, const SyncCfg& c, int np, double nominal) {
#pragma clang fp contract(off)
  {
    const int lane = threadIdx.x & 63;
    const QualityWindows qw = quality_windows(nominal);
    // the lists in LDS are named directly, so their accesses are LDS accesses
    const unsigned char* eflag = kLds ? L.leflag : L.eflag;
    const int* es = kLds ? L.les : L.es;
    const int* el = kLds ? L.lel : L.el;
    const int* rs = kLds ? L.lrs : L.rs;
    const int* rl = kLds ? L.lrl : L.rl;
    const signed char* craw = kLds ? L.lcraw : L.craw;
    const signed char* ceff = kLds ? L.lceff : L.ceff;
    int* cs = kLds ? L.lcs : L.cs;
    int* cl = kLds ? L.lcl : L.cl;
    unsigned char* ck = kLds ? L.lck : L.ck;
    unsigned char* co = kLds ? L.lco : L.co;
    const int cap = c.pulse_cap;
    int nref = 0, pk = -1, pstart = 0;
    int pv1 = -1, pv2 = -1;  // the last two pulses the walk reached before this one
    int cb = -64;            // the first of the 64 pulses held in the lanes
    int cfl = 0, ces = 0, cel = 0;
    unsigned long long hmask = 0;  // the lanes' hsync-length flags
    int index = 0;
    while (index < np) {
      // the walk's state is the same on every lane
      index = __builtin_amdgcn_readfirstlane(index);
      nref = __builtin_amdgcn_readfirstlane(nref);
      pk = __builtin_amdgcn_readfirstlane(pk);
      pstart = __builtin_amdgcn_readfirstlane(pstart);
      pv1 = __builtin_amdgcn_readfirstlane(pv1);
      pv2 = __builtin_amdgcn_readfirstlane(pv2);
      cb = __builtin_amdgcn_readfirstlane(cb);
      if (index >= cb + 64) {
        cb = index;
        const int i = cb + lane;
        cfl = i < np ? eflag[i] : 0;
        ces = i < np ? es[i] : 0;
        cel = i < np ? el[i] : 0;
        // the loads land here, so the walk's later steps wait for nothing
        __builtin_amdgcn_s_waitcnt(0);
        hmask = __ballot((cfl & 1) != 0);
      }
      const int k = index - cb;
      const int fl = __builtin_amdgcn_readlane(cfl, k);
      if (fl & 1) {
        // a run of hsync-length pulses is appended as it stands, one lane a pulse: each is
        // judged in order against the pulse before it (the refined list never outruns
        // the walk's index, which stays below the list capacity)
        const unsigned long long rest = ~(hmask >> k);
        const int run = rest == 0 ? 64 - k : (__builtin_ctzll(rest) < 64 - k ? __builtin_ctzll(rest) : 64 - k);
        const int j = lane - k;
        const int prev_s = __shfl_up(ces, 1);
        if (j >= 0 && j < run) {
          const bool io = j == 0 ? (nref > 0 && quality_int(qw, pk, pstart, kHSync, ces))
                                 : quality_int(qw, kHSync, prev_s, kHSync, ces);
          const int q = nref + j;
          if (q < cap) {
            ck[q] = (unsigned char)kHSync;
            cs[q] = ces;
            cl[q] = cel;
            co[q] = io ? 1 : 0;
          }
        }
        nref += run;
        pk = kHSync;
        pstart = __builtin_amdgcn_readlane(ces, k + run - 1);
        pv2 = run >= 2 ? index + run - 2 : pv1;
        pv1 = index + run - 1;
        index += run;
        continue;
      }
      // a rescue-length pulse the rescue does not shorten is passed over
      if ((fl & 2) || !(index > 2 && (fl & 4) && nref > 0 && pk == kHSync)) {
        pv2 = pv1;
        pv1 = index;
        index++;
        continue;
      }
      // TryReadVBlank (:340-471): pulses the walk has reached are read rescued.  The
      // accepted pulses after the first two go straight behind the refined list and
      // count when the read completes (the list never outruns the walk's index, so they
      // stay inside it).
      const int start = index - 2 > 0 ? index - 2 : 0;
      const int end_ex = np < index + 24 ? np : index + 24;
      int nv = 0, state = -1, lastk = -1, lasts = 0, written = 0, wk = -1, ws = 0;
      double state_end = 0.0, state_len = 0.0;
      bool has_len = false, done = false;
      // the read's pulses, one a lane (at most 26)
      int vrk = -1, vst = 0, vln = 0;
      {
        const int i = start + lane;
        if (i < end_ex) {
          const bool seen = i == index || i == pv1 || i == pv2;
          vrk = seen ? ceff[i] : craw[i];
          vst = seen ? es[i] : rs[i];
          vln = seen ? el[i] : rl[i];
        }
        __builtin_amdgcn_s_waitcnt(0);
      }
      for (int i = start; i < end_ex && !done; i++) {
        const int rk = __builtin_amdgcn_readlane(vrk, i - start);
        if (rk < 0) continue;
        const int pst = __builtin_amdgcn_readlane(vst, i - start);
        int acc = -1;
        bool fin = false;
        switch (state) {
          case -1:
            if (rk == kHSync) acc = kHSync;
            break;
          case kHSync:
            if (rk == kHSync) acc = kHSync;
            else if (rk == kEq) {
              acc = kEq;
              has_len = true;
              state_len = c.num_pulses / 2.0;
            } else if (rk == kVSync)
              acc = kVSync;
            break;
          case kEq:
            if (rk == kEq) acc = kEq;
            else if (rk == kVSync) {
              acc = kVSync;
              has_len = true;
              state_len = c.num_pulses / 2.0;
            } else if (rk == kHSync)
              acc = kHSync;
            break;
          case kVSync:
            if (rk == kEq) {
              acc = kEq2;
              has_len = true;
              state_len = c.num_pulses / 2.0;
            } else if (rk == kVSync)
              acc = kVSync;
            else if (rk == kHSync && pst > state_end)
              acc = kHSync;
            break;
          case kEq2:
            if (rk == kEq) acc = kEq2;
            else if (rk == kHSync) {
              acc = kHSync;
              fin = true;
            }
            break;
        }
        if (acc < 0) continue;
        if (acc != state) {
          if (pst < state_end) acc = -1;
          else if (has_len) {
            state_end = pst + ((state_len - 0.1) * nominal);
            has_len = false;
          }
        }
        if (acc >= 0) {
          const bool io = nv > 0 && quality_int(qw, lastk, lasts, acc, pst);
          if (nv < 32) {
            if (nv >= 2 && nref + (nv - 2) < cap) {
              if (lane == 0) {
                const int q = nref + (nv - 2);
                ck[q] = (unsigned char)acc;
                cs[q] = pst;
                cl[q] = __builtin_amdgcn_readlane(vln, i - start);
                co[q] = io ? 1 : 0;
              }
              written++;
              wk = acc;
              ws = pst;
            }
            lastk = acc;
            lasts = pst;
            nv++;
          }
          state = acc;
        }
        if (fin) done = true;
      }
      pv2 = pv1;
      pv1 = index;
      if (!done) {
        index++;
        continue;
      }
      if (written > 0) {
        nref += written;
        pk = wk;
        pstart = ws;
      }
      index += (nv - 2 > 1 ? nv - 2 : 1);
    }
    if (lane == 0) L.nref = nref;
  }
}

__global__ void __launch_bounds__(kT) field_logic_kernel(SyncCfg c, SyncWork w) {
#pragma clang fp contract(off)
  __shared__ LogicScratch L;
  const int f = blockIdx.x;
  if (f >= *w.spec_count) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int np = w.raw_count[f];
  int* gst = w.pulse_start + (long long)f * c.pulse_cap;
  int* gln = w.pulse_len + (long long)f * c.pulse_cap;
  const bool lds = np <= kPulseLds;
  if (tid == 0) {
    unsigned char* aux = w.logic_aux + (long long)f * kLogicAux * c.pulse_cap;
    const long long pc = c.pulse_cap;
    L.rs = lds ? L.lrs : gst;
    L.rl = lds ? L.lrl : gln;
    L.es = lds ? L.les : reinterpret_cast<int*>(aux);
    L.el = lds ? L.lel : reinterpret_cast<int*>(aux + 4 * pc);
    L.cl = lds ? L.lcl : reinterpret_cast<int*>(aux + 8 * pc);
    L.craw = lds ? L.lcraw : reinterpret_cast<signed char*>(aux + 12 * pc);
    L.ceff = lds ? L.lceff : reinterpret_cast<signed char*>(aux + 13 * pc);
    L.eflag = lds ? L.leflag : aux + 14 * pc;
    L.ck = lds ? L.lck : aux + 15 * pc;
    L.co = lds ? L.lco : aux + 16 * pc;
    L.cs = lds ? L.lcs : w.ref_start + (long long)f * c.pulse_cap;
    L.g0 = INT_MAX;
    L.run = 0;
    for (int k = 0; k < 4; k++) L.slot[k] = -1;
    for (int k = 0; k < 8; k++) L.mark[k] = -1;
  }
  if (lds)
    for (int i = tid; i < np; i += kT) {
      L.lrs[i] = gst[i];
      L.lrl[i] = gln[i];
    }
  for (int i = tid; i < kHistBins; i += kT) L.hist[i] = 0;
  __syncthreads();
  // EstimateTiming (:131-166): median of the hsync-like lengths by a histogram
  const double hs_lo = (c.hsync_us - 1.75) * c.fs_mhz, hs_hi = (c.hsync_us + 2.0) * c.fs_mhz;
  const int hbase = (int)ceil(hs_lo);
  for (int i = tid; i < np; i += kT) {
    const int len = L.rl[i];
    if (in_range(len, hs_lo, hs_hi)) {
      const int b = len - hbase;
      if (b >= 0 && b < kHistBins) atomicAdd(&L.hist[b], 1);
    }
  }
  __syncthreads();
  if (wave == 0) {
    constexpr int kPer = kHistBins / 64;
    int loc[kPer];
    int lsum = 0;
#pragma unroll
    for (int k = 0; k < kPer; k++) {
      loc[k] = L.hist[lane * kPer + k];
      lsum += loc[k];
    }
    const int incl = wave_scan_incl(lsum, lane);
    const int nh = __shfl(incl, 63);
    const int before = incl - lsum;
    // the bin holding sorted position k (k < nh)
    auto kth = [&](int k) {
      int bin = 0;
      if (k >= before && k < incl) {
        int acc = before;
#pragma unroll
        for (int q = 0; q < kPer; q++) {
          acc += loc[q];
          if (k < acc) {
            bin = lane * kPer + q;
            break;
          }
        }
      }
      const unsigned long long owner = __ballot(k >= before && k < incl);
      return (double)(__shfl(bin, __ffsll((long long)owner) - 1) + hbase);
    };
    double med;
    if (nh > 0) {
      const int mid = nh / 2;
      const double a = kth(mid);
      med = (nh & 1) ? a : (kth(mid - 1) + a) / 2.0;
    } else {
      med = c.hsync_us;  // v0.4.0 keeps the empty-list fallback in microseconds
    }
    if (lane == 0) {
      const double typical = c.hsync_us * c.fs_mhz;
      const double toff = med - typical;
      Timing t;
      t.median = med;
      t.hsync_lo = med - (c.hsync_tol * c.fs_mhz);
      t.hsync_hi = med + (c.hsync_tol * c.fs_mhz);
      t.eq_lo = ((c.eq_us - c.eq_tol) * c.fs_mhz) + toff;
      t.eq_hi = ((c.eq_us + c.eq_tol) * c.fs_mhz) + toff;
      t.vs_lo = ((c.vsync_us * 0.5) * c.fs_mhz) + toff;
      t.vs_hi = ((c.vsync_us + 1.0) * c.fs_mhz) + toff;
      L.t = t;
    }
  }
  __syncthreads();
  const Timing t = L.t;
  // RefinePulses (:243-314) with the long-hsync rescue on the DC-shifted reference
  const FieldLevels lv = w.levels[f];
  const Plane x{w.sync_ref + w.win_start[f], w.win_len[f], c.bias, lv.dc};
  const double nominal = c.nominal;
  const double step = 10.0 * c.hz_ire;
  const bool can_rescue = x.n > 0 && step > 0.0;
  auto is_h = [&](int len) { return in_range(len, t.hsync_lo, t.hsync_hi); };
  auto is_r = [&](int len) { return can_rescue && in_range(len, t.hsync_hi, t.hsync_hi * 3.0); };
  for (int i = tid; i < np; i += kT) {
    const int len = L.rl[i];
    L.craw[i] = (signed char)classify(len, t);
    L.es[i] = L.rs[i];
    L.el[i] = len;
  }
  __syncthreads();
  // each rescue-length pulse on one wave, rescued as often as the walk would rescue it
  // (the walk examines a rewritten pulse again)
  {
    const int mn = (int)ceil((c.eq_us * c.fs_mhz) / 8.0) > 0 ? (int)ceil((c.eq_us * c.fs_mhz) / 8.0) : 0;
    const int mx = (int)(nominal * 5.0) > 1 ? (int)(nominal * 5.0) : 1;
    for (int i = wave; i < np; i += kT / 64) {
      int ps = L.rs[i], pl = L.rl[i];
      if (is_h(pl) || !is_r(pl)) continue;
      for (;;) {
        long long s = ps < 0 ? 0 : (ps > x.n ? x.n : ps);
        long long e = (long long)ps + pl;
        e = e < s ? s : (e > x.n ? x.n : e);
        int r_s = 0, r_l = 0;
        const bool ok = s < e && wave_rescue(x, s, e, step, mn, mx, &r_s, &r_l);
        if (!ok || (r_s == ps && r_l == pl)) break;
        ps = r_s;
        pl = r_l;
        if (is_h(pl) || !is_r(pl)) break;
      }
      if (lane == 0) {
        L.es[i] = ps;
        L.el[i] = pl;
      }
    }
  }
  __syncthreads();
  for (int i = tid; i < np; i += kT) {
    const int len = L.el[i];
    L.ceff[i] = (signed char)classify(len, t);
    L.eflag[i] = (unsigned char)((is_h(len) ? 1 : 0) | (!is_h(len) && is_r(len) ? 2 : 0) |
                                 (in_range(len, t.eq_lo, t.eq_hi) ? 4 : 0));
  }
  __syncthreads();
  // the walk on wave 0
  if (__builtin_amdgcn_readfirstlane(wave) == 0) {
    if (lds) refine_walk<true>(L, c, np, nominal);
    else refine_walk<false>(L, c, np, nominal);
  }
  __syncthreads();
  const int nref = L.nref;
  // ComputeMeanLineLength (:490-538): the first longest run of hsync pulses
  {
    int carry = -1;  // last non-hsync index before the block
    for (int b0 = 0; b0 < nref; b0 += kT) {
      const int i = b0 + tid;
      const bool in = i < nref;
      const bool h = in && L.ck[i] == kHSync;
      const int brk = in && !h ? i : -1;
      // inclusive max-scan of the break positions
      int incl = brk;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const int u = __shfl_up(incl, o);
        if (lane >= o) incl = incl > u ? incl : u;
      }
      if (lane == 63) L.red[wave] = incl;
      __syncthreads();
      int before = carry;
      for (int ww = 0; ww < wave; ww++) before = before > L.red[ww] ? before : L.red[ww];
      const int last_brk = incl > before ? incl : before;
      if (h && (i + 1 == nref || L.ck[i + 1] != kHSync)) {
        const int st = last_brk + 1;
        const unsigned long long key = ((unsigned long long)(i - st + 1) << 32) | (0xFFFFFFFFull - (unsigned)st);
        atomicMax(&L.run, key);
      }
      int blk = carry;
      for (int ww = 0; ww < kT / 64; ww++) blk = blk > L.red[ww] ? blk : L.red[ww];
      __syncthreads();
      carry = blk;
    }
    __syncthreads();
    const unsigned long long best = L.run;
    const int run_start = best ? (int)(0xFFFFFFFFull - (best & 0xFFFFFFFFull)) : -1;
    const int run_len = best ? (int)(best >> 32) - 1 : -1;
    long long sum = 0;
    int cnt = 0;
    for (int i = run_start + 1 + tid; i < run_start + run_len; i += kT) {
      const int d = L.cs[i] - L.cs[i - 1];
      if (in_range((double)d / nominal, 0.95, 1.05)) {
        sum += d;
        cnt++;
      }
    }
    sum = wg_sum_ll(sum, L.lred);
    const long long c2 = wg_sum_ll(cnt, L.lred);
    // the sum of whole sample counts is exact in FP64, so its order does not matter
    if (tid == 0) L.mll = c2 > 0 ? (double)sum / (double)c2 : nominal;
    __syncthreads();
  }
  const double mll = L.mll;
  // MeasureVBlankBoundaryConsensus (FieldParityDetector.cs:172-253) and the transition
  // markers of EstimateLine0FromTransitions (VBlankSyncResolver.cs:317-445): the refined
  // pulses after the first in-order one past the first field's lines count to the second
  // group; the last transition of each kind in a group is the one kept
  const double second = nref > 0 ? L.cs[0] + (c.field_lines0 * mll) : 0.0;
  for (int i = 1 + tid; i < nref; i += kT)
    if (L.co[i] && L.cs[i] > second) atomicMin(&L.g0, i);
  __syncthreads();
  {
    const int g0 = L.g0;
    for (int i = 1 + tid; i < nref; i += kT) {
      if (!L.co[i]) continue;
      const int grp = i >= g0 ? 1 : 0;
      const int pk = L.ck[i - 1], k = L.ck[i];
      if (pk == kHSync && k != kHSync) {
        atomicMax(&L.slot[grp * 2], i);
        atomicMax(&L.mark[grp * 4], i);
      } else if (pk != kHSync && k == kHSync) {
        atomicMax(&L.slot[grp * 2 + 1], i);
        atomicMax(&L.mark[grp * 4 + 3], i);
      } else if (pk == kEq && k == kVSync) {
        atomicMax(&L.mark[grp * 4 + 1], i);
      } else if (pk == kVSync && k == kEq2) {
        atomicMax(&L.mark[grp * 4 + 2], i);
      }
    }
  }
  __syncthreads();
  // consensus and the line-0 marker pairs on wave 0: one lane a pair (32 lanes a parity),
  // the pairs added up in the nested loops' order
  if (__builtin_amdgcn_readfirstlane(wave) == 0) {
    double lens[4];
    for (int k = 0; k < 4; k++) {
      const int i = L.slot[k];
      lens[k] = i < 0 ? -1.0 : round_half_line((L.cs[i] - L.cs[i - 1]) / mll);
    }
    const double ffn[4] = {1.0, 0.5, 0.5, 1.0}, sfn[4] = {0.5, 1.0, 1.0, 0.5};
    const double ffp[4] = {0.5, 0.5, 1.0, 1.0}, sfp[4] = {1.0, 1.0, 0.5, 0.5};
    int cons = 0, det = 0;
    for (int i = 0; i < 4; i++) {
      if (lens[i] < 0.0) continue;
      if (fabs(lens[i] - (c.is_pal ? ffp[i] : ffn[i])) < 1e-9) {
        cons++;
        det++;
      } else if (fabs(lens[i] - (c.is_pal ? sfp[i] : sfn[i])) < 1e-9) {
        det++;
      }
    }
    if (lane == 0) {
      FieldLogic& out = w.logic[f];
      out.consensus = cons;
      out.detected = det;
      out.first_conf = 0;
      out.second_conf = 0;
      if (det > 0) {
        const double wgt = det / 4.0;
        out.first_conf = (int)round_away((cons / (double)det) * wgt * 100.0);
        out.second_conf = (int)round_away(((det - cons) / (double)det) * wgt * 100.0);
      }
    }
    double fsm[4], lsm[4];
    int fh[4], lh[4];
    for (int k = 0; k < 4; k++) {
      const int a = L.mark[k], b = L.mark[4 + k];
      fh[k] = a >= 0;
      lh[k] = b >= 0;
      fsm[k] = a < 0 ? 0.0 : (double)(k == 3 ? L.cs[a - 1] : L.cs[a]);
      lsm[k] = b < 0 ? 0.0 : (double)(k == 3 ? L.cs[b - 1] : L.cs[b]);
    }
    const int par = lane >> 5, idx = lane & 31;
    const bool first = par == 0;
    const int cur_lines = first ? c.field_lines0 : c.field_lines1;
    const Bounds b = field_bounds(c.is_pal, first);
    const double section = c.num_pulses / 2.0;
    const double fhl = b.first_hsync + (section * 3.0) - 0.5 + b.first_eq2;
    const MarkerList fm = partial_markers(fsm, fh, b.first_hsync, section);
    const MarkerList lm = partial_markers(lsm, lh, cur_lines + b.last_hsync, section);
    // which markers were found does not depend on the parity
    const int nf2 = fm.n * (fm.n - 1) / 2, nl2 = lm.n * (lm.n - 1) / 2, nx = fm.n * lm.n;
    Acc p{0.0, 0.0, 0};
    if (idx < nf2 + nl2 + nx) {
      int i, j;
      if (idx < nf2) {
        pair_ij(idx, fm.n, &i, &j);
        p = acc_pair(fm.smp(i), fm.line(i), fm.smp(j), fm.line(j), mll, fhl);
      } else if (idx < nf2 + nl2) {
        pair_ij(idx - nf2, lm.n, &i, &j);
        p = acc_pair(lm.smp(i), lm.line(i), lm.smp(j), lm.line(j), mll, fhl);
      } else {
        const int q = idx - nf2 - nl2;
        i = q / lm.n;
        j = q - i * lm.n;
        p = acc_pair(fm.smp(i), fm.line(i), lm.smp(j), lm.line(j), mll, fhl);
      }
    }
    for (int pr = 0; pr < 2; pr++) {
      const int base = pr * 32;
      const Acc fa = acc_lanes(p, base, base + nf2);
      const Acc la = acc_lanes(p, base + nf2, base + nf2 + nl2);
      Acc comb{0.0, 0.0, 0};
      if (fa.n > 0 && la.n > 0) {
        const double fr = fa.loc / fa.n, lr = la.loc / la.n;
        if (fabs(fr - lr) < 0.5 * mll) comb = acc_add(acc_add(fa, la), acc_lanes(p, base + nf2 + nl2, base + nf2 + nl2 + nx));
      }
      if (lane == 0) {
        L.acc[pr * 3] = fa;
        L.acc[pr * 3 + 1] = la;
        L.acc[pr * 3 + 2] = comb;
      }
    }
    if (idx == 0) L.fhl[par] = fhl;
  }
  __syncthreads();
  // FinalizeEstimate (Dsp/VBlankSyncResolver.cs:666-718): one wave per estimate; the
  // alignment offsets add up in pulse order
  if (__builtin_amdgcn_readfirstlane(wave) < 6) {
    const Acc a = L.acc[wave];
    const int par = wave / 3;
    const double fhl = L.fhl[par];
    const int cur_lines = par == 0 ? c.field_lines0 : c.field_lines1;
    Estimate e{};
    if (a.n > 0) {
      double fh = round_away((a.loc + a.off) / a.n);
      const double unaligned = fh;
      int brk = nref;
      for (int b0 = 0; b0 < nref; b0 += 64) {
        const int i = b0 + lane;
        bool bad = false;
        if (i < nref && L.ck[i] == kHSync && L.co[i]) {
          const int rl = (int)round_away(((L.cs[i] - fh) / mll) + fhl);
          bad = rl > cur_lines;
        }
        const unsigned long long m = __ballot(bad);
        if (m) {
          brk = b0 + __ffsll((long long)m) - 1;
          break;
        }
      }
      double off = 0.0;
      int cnt = 0;
      for (int b0 = 0; b0 < brk; b0 += 64) {
        const int i = b0 + lane;
        bool use = false;
        double term = 0.0;
        if (i < brk && L.ck[i] == kHSync && L.co[i]) {
          const int rl = (int)round_away(((L.cs[i] - fh) / mll) + fhl);
          if (rl >= fhl) {
            use = true;
            term = fh + (mll * (rl - fhl)) - L.cs[i];
          }
        }
        unsigned long long um = __ballot(use);
        while (um) {
          const int k = __builtin_ctzll(um);
          um &= um - 1;
          off += readlane_d(term, k);
          cnt++;
        }
      }
      if (cnt > 0) fh -= off / cnt;
      e.ok = 1;
      e.line0 = fh - (mll * fhl);
      e.first_hsync = fh;
      e.unaligned = unaligned;
      e.valid_count = a.n;
    }
    if (lane == 0) L.est[wave] = e;
  }
  __syncthreads();
  if (tid == 0) {
    FieldLogic& out = w.logic[f];
    out.raw = np;
    out.refined = nref;
    out.hsync_median = t.median;
    out.mll = mll;
    for (int par = 0; par < 2; par++) {
      out.tr[par].first = L.est[par * 3];
      out.tr[par].last = L.est[par * 3 + 1];
      out.tr[par].combined = L.est[par * 3 + 2];
    }
  }
  // the refined pulses for the line locations and the rare previous-field path
  int* ks = w.ref_kind + (long long)f * c.pulse_cap;
  int* ss = w.ref_start + (long long)f * c.pulse_cap;
  for (int i = tid; i < nref; i += kT) {
    ks[i] = L.ck[i] | (L.co[i] << 8);
    ss[i] = L.cs[i];
  }
}

// ------------------------------------------------------------------ S5: field chain
// AlignFirstHSyncToValidPulses (TbcFieldDecodePipeline.cs:6207-6242) over the stored refined pulses
__device__ double align_first_hsync(const int* ks, const int* ss, int n, double fh, double mll, double fhl,
                                    int cur_lines) {
#pragma clang fp contract(off)
  double off = 0.0;
  int cnt = 0;
  for (int i = 0; i < n; i++) {
    if ((ks[i] & 0xff) != kHSync || !(ks[i] >> 8)) continue;
    const double line = ((ss[i] - fh) / mll) + fhl;
    const int rl = (int)round_away(line);
    if (rl > cur_lines) break;
    if (rl >= fhl) {
      off += fh + (mll * (rl - fhl)) - ss[i];
      cnt++;
    }
  }
  return cnt > 0 ? fh - (off / cnt) : fh;
}

__device__ void resolve_window(const SyncCfg& c, long long begin, long long* ws, long long* wl) {
  long long read = begin - c.block_cut;
  read = read > 0 ? read : 0;
  const long long a0 = (read / c.block_length) * c.block_length;
  const long long a1 = a0 + c.read_samples;
  const long long fb = a0 / c.stride, lb = a1 / c.stride;
  *ws = fb * c.stride;
  *wl = (lb - fb + 1) * c.stride;
}

// The sequence walk over the round's fields on one thread, their inputs first staged in
// LDS by the wave (so the walk waits on no global load) and the records, the sync
// history and the level state written as it goes (stores only; the level state and the
// history once, as the last field processed leaves them).
constexpr int kChainLds = 64;

__global__ void __launch_bounds__(64) field_chain_kernel(SyncCfg c, SyncWork w) {
#pragma clang fp contract(off)
  __shared__ FieldLogic s_logic[kChainLds];
  __shared__ FieldLevels s_levels[kChainLds];
  __shared__ long long s_ws[kChainLds], s_wl[kChainLds];
  __shared__ int s_ovf[kChainLds];
  SeqCtl& ctl = *w.ctl;
  if (ctl.halted) return;
  const int nf = *w.spec_count;
  const bool staged = nf <= kChainLds;
  if (staged)
    for (int i = threadIdx.x; i < nf; i += blockDim.x) {
      s_logic[i] = w.logic[i];
      s_levels[i] = w.levels[i];
      s_ws[i] = w.win_start[i];
      s_wl[i] = w.win_len[i];
      s_ovf[i] = w.overflow[i];
    }
  __syncthreads();
  if (threadIdx.x != 0) return;
  const FieldLogic* in_logic = staged ? s_logic : w.logic;
  const FieldLevels* in_levels = staged ? s_levels : w.levels;
  const long long* in_ws = staged ? s_ws : w.win_start;
  const long long* in_wl = staged ? s_wl : w.win_len;
  const int* in_ovf = staged ? s_ovf : w.overflow;
  const int base = ctl.base;
  SyncHistoryDev h = *w.history_in;
  h.prev_first_hsync_read -= w.state_shift;
  // the state the next call continues from: as the last field (decoded or failed) left it
  int last_done = -1;
  long long begin = ctl.begin;
  int out_count = 0;
  int stop_status = kStatusNotRead;
  int f = 0;
  // A field the reference fails (no sync pulses, no first hsync, too little data) is
  // recorded with its recovery offset (BuildRecoveryException, TbcFieldDecodePipeline.cs
  // :5936-5959; tape offsets, never a stop) and the sequence state it leaves; the call
  // ends there, since later windows were speculated from this field decoding.
  auto recover = [&](FieldRecord& rec, int status, double offset) {
    rec.status = status;
    rec.next_offset = offset;
    const long long nb = begin + (long long)offset;
    rec.next_begin = nb > 0 ? nb : 0;
    begin = rec.next_begin;
    stop_status = status;
    last_done = f;
  };
  for (; f < nf; f++) {
    FieldRecord rec{};
    rec.struct_size = sizeof(FieldRecord);
    bool stop = false;
    do {
      long long ws, wl;
      resolve_window(c, begin, &ws, &wl);
      rec.begin = begin;
      rec.window_start = ws;
      rec.window_length = wl;
      rec.next_begin = begin;
      if (ws != in_ws[f] || wl != in_wl[f]) {
        rec.status = kStatusMispredicted;
        stop_status = rec.status;
        stop = true;
        break;
      }
      const FieldLevels lv = in_levels[f];
      const FieldLogic lg = in_logic[f];
      rec.level_source = lv.source;
      rec.sync_level = lv.sync;
      rec.blank_level = lv.blank;
      rec.dc_offset = lv.dc;
      rec.threshold = lv.threshold;
      rec.raw_pulses = lg.raw;
      rec.refined_pulses = lg.refined;
      rec.hsync_median = lg.hsync_median;
      rec.mean_line_length = lg.mll;
      if (in_ovf[f]) {
        rec.status = kStatusCapacity;
        stop_status = rec.status;
        stop = true;
        break;
      }
      if (lg.raw == 0) {
        const bool missing_start = h.pulse_search_initialized != 0;
        recover(rec, missing_start ? 2 : 1, missing_start ? c.nominal * 100.0 : (c.fs_mhz * 1000000.0) / 10.0);
        stop = true;
        break;
      }
      h.pulse_search_initialized = 1;
      const double mll = lg.mll;
      // FieldParityDetector.ResolveCadence, no fallback line
      bool first;
      int conf = 0;
      if (h.prev_first >= 0) first = !h.prev_first;
      else first = lg.detected == 0 || round_away(lg.consensus / (double)lg.detected) == 1.0;
      int eff = 100;
      if (lg.detected > 0 && !h.has_prev_first_hsync && eff > 50) eff = 50;
      if (lg.first_conf >= eff && lg.first_conf > lg.second_conf) {
        first = true;
        conf = lg.first_conf;
      } else if (lg.second_conf >= eff && lg.first_conf < lg.second_conf) {
        first = false;
        conf = lg.second_conf;
      }
      rec.is_first_field = first ? 1 : 0;
      rec.parity_confidence = conf;
      h.prev_first = first ? 1 : 0;
      const int cur_lines = first ? c.field_lines0 : c.field_lines1;
      const double fhl = first_hsync_line(c.is_pal, c.num_pulses, first);
      // EstimateFirstHSyncFromPrevious (:6146-6173)
      bool has_est = false;
      double est = 0.0;
      if (h.has_prev_first_hsync && h.prev_first_hsync > 0.0 && mll > 0.0) {
        const double cur = first ? c.field_lines0 : c.field_lines1;
        const double prev = first ? c.field_lines1 : c.field_lines0;
        const double est_lines = c.is_pal ? cur : prev;
        const double last_off = (h.prev_first_hsync_read - ws) / mll;
        est = round_away((last_off + est_lines + (h.prev_first_hsync / mll)) * mll);
        has_est = true;
      }
      // ResolveLine0Location, vhs (:5452-5651)
      const Transitions& tr = lg.tr[first ? 0 : 1];
      bool ok = false, used_prev = false;
      double line0 = 0.0, fh = 0.0, unaligned = 0.0;
      if (tr.combined.ok) {
        ok = true;
        line0 = tr.combined.line0;
        fh = tr.combined.first_hsync;
        unaligned = tr.combined.unaligned;
      } else {
        const int fc = tr.first.ok ? tr.first.valid_count : 0;
        const int lc = tr.last.ok ? tr.last.valid_count : 0;
        const bool has_prev = h.has_prev_first_hsync && h.prev_first_hsync > 0.0;
        const Estimate* single = nullptr;
        if (fc == 6 || (!has_prev && fc > lc)) single = tr.first.ok ? &tr.first : nullptr;
        else if (lc == 6 || (!has_prev && lc > fc)) single = tr.last.ok ? &tr.last : nullptr;
        if (single) {
          ok = true;
          line0 = single->line0;
          fh = single->first_hsync;
          unaligned = single->unaligned;
        } else if (has_est) {
          double e = est;
          if (h.prev_hsync_difference >= -0.5 && h.prev_hsync_difference <= 0.5) e += mll * h.prev_hsync_difference;
          const int* ks = w.ref_kind + (long long)f * c.pulse_cap;
          const int* ss = w.ref_start + (long long)f * c.pulse_cap;
          if (e <= 0.0) e = lg.refined > 0 ? ss[0] : 0.0;
          unaligned = round_away(e);
          fh = align_first_hsync(ks, ss, lg.refined, unaligned, mll, fhl, cur_lines);
          line0 = fh - (fhl * mll);
          ok = true;
          used_prev = true;
        }
      }
      if (!ok) {
        recover(rec, 2, c.nominal * 100.0);
        stop = true;
        break;
      }
      // UpdateSyncHistory (:6119-6144)
      if (!used_prev && has_est && isfinite(unaligned)) h.prev_hsync_difference = (unaligned - est) / mll;
      if (isfinite(fh)) {
        h.prev_first_hsync_read = ws;
        h.prev_first_hsync = fh;
        h.has_prev_first_hsync = 1;
      }
      rec.line0 = line0;
      rec.first_hsync = fh;
      rec.first_hsync_line = fhl;
      const int processed = c.output_lines + 10;
      const double last_line = (((double)wl - line0) / mll) - 1.0;
      if (last_line < processed) {
        // ThrowIfInsufficientFieldData (:5961-6008), tape
        const double suggested = line0 - (mll * 20.0);
        recover(rec, 3, suggested > c.nominal ? suggested : c.nominal);
        stop = true;
        break;
      }
      // ComputeVhsNextFieldOffsetSamples (:2299-2325)
      const Bounds b = field_bounds(c.is_pal, first);
      const double next_loc = fh + (mll * ((cur_lines + b.last_hsync) - fhl));
      double next_off = next_loc - (8.0 * c.nominal) + (double)c.block_cut;
      if (!(isfinite(next_off) && next_off > 0.0)) {
        const int fbl = c.output_lines - 7 > 0 ? c.output_lines - 7 : 0;
        next_off = line0 + (fbl * mll) + (double)c.block_cut;  // line-location fallback, projected
        next_off = next_off > 1.0 ? next_off : 1.0;
      }
      rec.next_offset = next_off;
      const long long adv = (long long)round_away(next_off);
      rec.next_begin = ws + (adv > 1 ? adv : 1);
      rec.status = 0;
      begin = rec.next_begin;
      out_count++;
      last_done = f;
    } while (false);
    w.records[base + f] = rec;
    if (stop) break;
  }
  *w.history_out = h;
  w.level_state_out[0] = last_done < 0 ? *w.level_state_in
                                        : (last_done + 1 < nf ? w.level_before[last_done + 1] : *w.level_state_after_all);
  // a failed field takes a record; the stop record after the decoded fields does not
  const bool failed = stop_status >= 1 && stop_status <= 3;
  const int used = base + out_count + (failed ? 1 : 0);
  long long nws, nwl;
  resolve_window(c, begin, &nws, &nwl);
  if (f == nf && base + f < c.max_records) {
    // the next field's window was not speculated (it runs past the span, or no vertical
    // sync in the span places it)
    FieldRecord& rec = w.records[base + f];
    rec = FieldRecord{};
    rec.struct_size = sizeof(FieldRecord);
    rec.status = kStatusNotRead;
    rec.begin = begin;
    rec.window_start = nws;
    rec.window_length = nwl;
    rec.next_begin = begin;
  }
  ctl.begin = begin;
  ctl.records = used;
  ctl.decoded = out_count;
  if (stop_status == kStatusCapacity || used >= c.max_records ||
      (stop_status == kStatusNotRead && nws + nwl > w.span_len))
    ctl.halted = 1;
}

// ------------------------------------------------------------------ S6: line locations
constexpr int kLineLds = 512;  // processed lines per field (the line stride is at most 511)

// BuildUpstreamLineLocations (SyncAnalysis.cs:585-657) from trunc(first hsync) at line
// FirstHSyncLine, written window-relative for every processed line.  A line's search walks
// the sorted starts from the current pulse while the next start is not farther from the
// expected position, and takes where it stops when that is within mll / 1.5.  Over sorted
// starts the distance falls up to the expected position, so a walk that begins at or
// before J (the first index from one below the lower bound whose successor is strictly
// farther, at most n - 2) stops at J: J and its distance are found for every line at
// once, and the pass over the lines that carries the current pulse walks only where the
// current pulse has passed J.
__global__ void __launch_bounds__(kT) field_lines_kernel(SyncCfg c, SyncWork w) {
#pragma clang fp contract(off)
  __shared__ int lsrc[kPulseLds];
  __shared__ int lsrt[kPulseLds];
  __shared__ int lj[kLineLds];
  __shared__ int lsel[kLineLds];
  __shared__ unsigned char lok[kLineLds];
  __shared__ int unsorted;
  const int f = blockIdx.x;
  const int tid = threadIdx.x;
  if (f >= w.ctl->decoded) return;
  const FieldRecord rec = w.records[w.ctl->base + f];
  const int n = rec.refined_pulses;
  int* ss = w.ref_start + (long long)f * c.pulse_cap;
  const double mll = rec.mean_line_length;
  const double ref_pulse = trunc(rec.first_hsync);
  const int ref_line = (int)rec.first_hsync_line;
  const int lines = c.output_lines + 10;
  double* out = w.lines + (long long)f * c.line_stride;
  const double maxd = mll / 1.5;
  if (n > kPulseLds) {
    // a window full of noise: sorted in place (the field chain has read the list) and
    // walked line by line
    if (tid != 0) return;
    for (int i = 1; i < n; i++) {  // nearly sorted: insertion sort (equal keys are equal values)
      const int v = ss[i];
      int j = i - 1;
      while (j >= 0 && ss[j] > v) {
        ss[j + 1] = ss[j];
        j--;
      }
      ss[j + 1] = v;
    }
    int cur = 0;
    for (int line = 0; line < lines; line++) {
      const double expected = ref_pulse + (mll * (line - ref_line));
      out[line] = expected;
      if (cur >= n) continue;
      double cd = fabs(ss[cur] - expected);
      double smallest = maxd;
      double sel = -1.0;
      int si = cur;
      while (si < n - 1) {
        if (cd <= smallest) {
          smallest = cd;
          cur = si;
          sel = ss[si];
        }
        const double nd = fabs(ss[si + 1] - expected);
        if (nd > cd) break;
        cd = nd;
        si++;
      }
      if (sel >= 0.0) {
        out[line] = sel;
        cur++;
      }
    }
    return;
  }
  if (tid == 0) unsorted = 0;
  for (int i = tid; i < n; i += kT) lsrc[i] = ss[i];
  __syncthreads();
  for (int i = tid; i + 1 < n; i += kT)
    if (lsrc[i] > lsrc[i + 1]) unsorted = 1;
  __syncthreads();
  if (unsorted) {
    // rank sort (OrderBy is stable, and equal keys are equal values)
    for (int i = tid; i < n; i += kT) {
      const int v = lsrc[i];
      int r = 0;
      for (int j = 0; j < n; j++) {
        const int u = lsrc[j];
        r += (u < v || (u == v && j < i)) ? 1 : 0;
      }
      lsrt[r] = v;
    }
  } else {
    for (int i = tid; i < n; i += kT) lsrt[i] = lsrc[i];
  }
  __syncthreads();
  const int* srt = lsrt;
  // J and the acceptance at J for every line
  for (int line = tid; line < lines; line += kT) {
    if (n < 2) break;
    const double e = ref_pulse + (mll * (line - ref_line));
    int lo = 0, hi = n;  // lower bound: the first start >= e
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if ((double)srt[mid] < e) lo = mid + 1;
      else hi = mid;
    }
    int j = lo - 1 > 0 ? lo - 1 : 0;
    while (j < n - 2 && !(fabs(srt[j + 1] - e) > fabs(srt[j] - e))) j++;
    if (j > n - 2) j = n - 2;
    lj[line] = j;
    lok[line] = (unsigned char)((fabs(srt[j] - e) <= maxd ? 1 : 0) | ((double)srt[j] >= 0.0 ? 2 : 0));
  }
  __syncthreads();
  // the pass that carries the current pulse, on wave 0: 64 lines' J and acceptance sit
  // in the lanes and are read by lane index
  if (tid < 64) {
    const int lane = tid;
    int cur = 0;
    for (int l0 = 0; l0 < lines; l0 += 64) {
      const int l = l0 + lane;
      const int myj = l < lines ? lj[l] : 0;
      const int myok = l < lines ? lok[l] : 0;
      const int cnt = lines - l0 < 64 ? lines - l0 : 64;
      int mysel = -1;
      for (int k = 0; k < cnt; k++) {
        const int jl = __builtin_amdgcn_readlane(myj, k), okl = __builtin_amdgcn_readlane(myok, k);
        int sel = -1;
        if (cur < n - 1) {
          int t, ok;
          if (cur <= jl) {
            t = jl;
            ok = okl;
          } else {
            const double e = ref_pulse + (mll * ((l0 + k) - ref_line));
            t = cur;
            while (t < n - 2 && !(fabs(srt[t + 1] - e) > fabs(srt[t] - e))) t++;
            ok = (fabs(srt[t] - e) <= maxd ? 1 : 0) | ((double)srt[t] >= 0.0 ? 2 : 0);
          }
          // the walk's pick moves the current pulse; a non-negative pick is taken and passed
          if (ok & 1) {
            cur = t;
            if (ok & 2) {
              sel = t;
              cur = t + 1;
            }
          }
        }
        if (lane == k) mysel = sel;
      }
      if (l < lines) lsel[l] = mysel;
    }
  }
  __syncthreads();
  for (int line = tid; line < lines; line += kT) {
    const int sel = lsel[line];
    out[line] = sel >= 0 ? (double)srt[sel] : ref_pulse + (mll * (line - ref_line));
  }
}

// ------------------------------------------------------------------ S0: speculated windows
// The read windows of the span's fields: the first from the caller's start, each next one
// three-and-a-bit lines (the equalizing section) before the next vertical-sync group of
// broad pulses, less eight lines, plus the block cut, as the next-field offset places it.
__global__ void __launch_bounds__(kT) spec_windows_kernel(SyncCfg c, SyncWork w, const int* starts, const int* lens,
                                                          const int* pulse_count, int capacity, long long span_len,
                                                          int max_fields) {
#pragma clang fp contract(off)
  __shared__ int groups[kMaxGroups];
  __shared__ int ng;
  __shared__ int red[kT / 64 + 4];
  SeqCtl& ctl = *w.ctl;
  if (ctl.halted) {
    // a halted call: this round decodes nothing
    if (threadIdx.x == 0) {
      *w.spec_count = 0;
      ctl.decoded = 0;
    }
    return;
  }
  // this round's list-capacity flags start clear
  for (int i = threadIdx.x; i < max_fields; i += kT) w.overflow[i] = 0;
  const int pc = *pulse_count < capacity ? *pulse_count : capacity;
  const double broad = 0.5 * ((c.hsync_us + c.vsync_us) * c.fs_mhz);
  if (threadIdx.x == 0) ng = 0;
  __syncthreads();
  for (int b0 = 0; b0 < pc; b0 += kT) {
    const int k = b0 + threadIdx.x;
    const bool g = k < pc && lens[k] > broad && (k == 0 || lens[k - 1] <= broad);
    int tot;
    const int idx = wg_exclusive(g ? 1 : 0, red, &tot) + ng;
    if (g && idx < kMaxGroups) groups[idx] = starts[k];
    __syncthreads();
    if (threadIdx.x == 0) ng += tot;
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  const int n_groups = ng < kMaxGroups ? ng : kMaxGroups;
  const double line = c.nominal;
  const double eq_lines = c.num_pulses / 2.0;
  long long begin = ctl.begin;
  ctl.base = ctl.records;
  ctl.decoded = 0;
  const int room = c.max_records - ctl.records;
  const int lim = room < max_fields ? room : max_fields;
  int nf = 0;
  int g = 0;
  for (int f = 0; f < lim; f++) {
    long long ws, wl;
    resolve_window(c, begin, &ws, &wl);
    if (ws + wl > span_len) break;
    w.win_start[f] = ws;
    w.win_len[f] = wl;
    nf++;
    // this field's vertical sync: the first group at least two lines into its window
    while (g < n_groups && groups[g] < ws + 2.0 * line) g++;
    if (g + 1 >= n_groups) break;
    const double next_eq1 = groups[g + 1] - eq_lines * line;
    begin = (long long)round_away(next_eq1 - 8.0 * line + (double)c.block_cut);
    g++;
  }
  *w.spec_count = nf;
}

// ------------------------------------------------------------------ sequence state and tails
__global__ void reset_kernel(LevelStateDev* level, SyncHistoryDev* hist) {
  if (threadIdx.x != 0) return;
  *level = LevelStateDev{};
  SyncHistoryDev h{};
  h.prev_first = -1;
  h.prev_hsync_difference = -1.0;
  *hist = h;
}

__global__ void tail_tables_kernel(SyncCfg c, SyncWork w, FieldTail t, int max_fields) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= max_fields) return;
  const int n = w.ctl->decoded;
  const int lines = c.output_lines + 10;
  if (f >= n) {
    t.line_counts[f] = 0;
    return;
  }
  const FieldRecord rec = w.records[w.ctl->base + f];
  t.spans[f] = w.sync_ref + rec.window_start;
  t.span_len[f] = rec.window_length;
  t.locs[f] = w.lines + (long long)f * c.line_stride;
  t.filled[f] = t.zeros;
  t.refined[f] = t.refined_base + (long long)f * c.line_stride;
  t.filled_out[f] = t.filled_out_base + (long long)f * c.line_stride;
  t.line_counts[f] = lines;
  t.origins[f] = rec.window_start;
  t.first_lines[f] = t.out_first_line[rec.is_first_field ? 0 : 1];
  t.threshold[f] = rec.threshold;
  t.dc[f] = rec.dc_offset;
  t.out_index[f] = w.ctl->base + f;
}

// ApplyNtscFscPhaseShiftCore (TbcFieldDecodePipeline.cs:4022-4038): every location less the shift
__global__ void render_kernel(SyncCfg c, SyncWork w, FieldTail t, int max_fields) {
#pragma clang fp contract(off)
  const int f = blockIdx.y;
  if (f >= w.ctl->decoded) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int lines = c.output_lines + 10;
  if (i >= lines) return;
  const double v = t.refined_base[(long long)f * c.line_stride + i];
  t.render[(long long)f * c.line_stride + i] = t.fsc_shift != 0.0 ? v - t.fsc_shift : v;
}

// The round's tables into the call's, by record: line locations, refined and render
// locations (window coordinates), and the render locations from the field's output first
// line in span coordinates (entries past the last location are NaN).
__global__ void export_kernel(SyncCfg c, SyncWork w, FieldTail t, int max_fields) {
#pragma clang fp contract(off)
  const int f = blockIdx.y;
  if (f >= w.ctl->decoded) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int lines = c.output_lines + 10;
  if (i >= lines) return;
  const long long src = (long long)f * c.line_stride + i;
  const long long dst = (long long)(w.ctl->base + f) * c.line_stride + i;
  t.acc_lines[dst] = w.lines[src];
  t.acc_refined[dst] = t.refined_base[src];
  t.acc_render[dst] = t.render[src];
  const int first = t.first_lines[f];
  t.acc_span[dst] = i + first < lines ? t.render[src + first] + (double)t.origins[f] : __builtin_nan("");
}

__global__ void results_kernel(const SeqCtl* ctl, int* record_count, long long* next_begin) {
  if (threadIdx.x != 0) return;
  *record_count = ctl->records;
  if (next_begin) *next_begin = ctl->begin;
}

__global__ void ctl_init_kernel(SeqCtl* ctl, long long first_begin, int first_record) {
  if (threadIdx.x != 0) return;
  *ctl = SeqCtl{first_begin, first_record, first_record, 0, 0};
}

}  // namespace vsync
}  // namespace vhsd

extern "C" hipError_t vhsd_launch_vhs_ctl_init(vhsd::vsync::SeqCtl* ctl, long long first_begin, int first_record,
                                               hipStream_t stream) {
  hipLaunchKernelGGL(vhsd::vsync::ctl_init_kernel, dim3(1), dim3(64), 0, stream, ctl, first_begin, first_record);
  return hipGetLastError();
}

extern "C" hipError_t vhsd_launch_vhs_results(const vhsd::vsync::SeqCtl* ctl, int* record_count, int64_t* next_begin,
                                              hipStream_t stream) {
  hipLaunchKernelGGL(vhsd::vsync::results_kernel, dim3(1), dim3(64), 0, stream, ctl, record_count,
                     reinterpret_cast<long long*>(next_begin));
  return hipGetLastError();
}

extern "C" hipError_t vhsd_launch_vhs_export(const vhsd::vsync::SyncCfg* c, const vhsd::vsync::SyncWork* w,
                                             const vhsd::vsync::FieldTail* t, int max_fields, hipStream_t stream) {
  const int lines = c->output_lines + 10;
  hipLaunchKernelGGL(vhsd::vsync::export_kernel, dim3((lines + 255) / 256, max_fields), dim3(256), 0, stream, *c, *w,
                     *t, max_fields);
  return hipGetLastError();
}

extern "C" hipError_t vhsd_launch_vhs_sync_reset(vhsd::vsync::LevelStateDev* level, vhsd::vsync::SyncHistoryDev* hist,
                                                 hipStream_t stream) {
  hipLaunchKernelGGL(vhsd::vsync::reset_kernel, dim3(1), dim3(64), 0, stream, level, hist);
  return hipGetLastError();
}

extern "C" hipError_t vhsd_launch_vhs_tail_tables(const vhsd::vsync::SyncCfg* c, const vhsd::vsync::SyncWork* w,
                                                  const vhsd::vsync::FieldTail* t, int max_fields,
                                                  hipStream_t stream) {
  hipLaunchKernelGGL(vhsd::vsync::tail_tables_kernel, dim3((max_fields + 63) / 64), dim3(64), 0, stream, *c, *w, *t,
                     max_fields);
  return hipGetLastError();
}

extern "C" hipError_t vhsd_launch_vhs_render(const vhsd::vsync::SyncCfg* c, const vhsd::vsync::SyncWork* w,
                                             const vhsd::vsync::FieldTail* t, int max_fields, hipStream_t stream) {
  const int lines = c->output_lines + 10;
  hipLaunchKernelGGL(vhsd::vsync::render_kernel, dim3((lines + 255) / 256, max_fields), dim3(256), 0, stream, *c, *w,
                     *t, max_fields);
  return hipGetLastError();
}

extern "C" hipError_t vhsd_launch_vhs_sync(const vhsd::vsync::SyncCfg* c, const vhsd::vsync::SyncWork* w,
                                           const int* starts, const int* lens, const int* pulse_count, int capacity,
                                           long long span_len, int max_fields, hipStream_t stream) {
  using namespace vhsd::vsync;
  if (max_fields <= 0) return hipSuccess;
  if (c->pulse_cap < 1 || c->output_lines + 10 > c->line_stride) return hipErrorInvalidValue;
  hipLaunchKernelGGL(spec_windows_kernel, dim3(1), dim3(kT), 0, stream, *c, *w, starts, lens, pulse_count, capacity,
                     span_len, max_fields);
  const SegBuffers& b = w->seg;
  const int segs = b.segs;
  // S1: level search
  const double eql = (double)c->ser_eq_len, ll = (double)c->ser_line_len;
  const int pmn = (int)std::ceil(eql / 8.0) > 0 ? (int)std::ceil(eql / 8.0) : 0;
  const int pmx = (int)std::floor(ll * 5.0) > 1 ? (int)std::floor(ll * 5.0) : 1;
  const double min_vsync = (c->vsync_us * c->fs_mhz) * 0.8;
  hipLaunchKernelGGL(seg_min_kernel, dim3(segs, max_fields), dim3(kSegT), 0, stream, *c, *w);
  hipLaunchKernelGGL(search_thresholds_kernel, dim3(max_fields), dim3(64), 0, stream, *c, *w, max_fields);
  SegJob j1{b.thr, kSpecR, 0, kSpecHalf, nullptr, pmn, pmx, min_vsync, nullptr, b.sum, segs, b.choice, b.seg_carry,
            b.seg_off, b.count, nullptr, w->pulse_start, w->pulse_len};
  hipLaunchKernelGGL(seg_count_kernel, dim3(segs, 1, max_fields), dim3(kSegT), 0, stream, *c, *w, j1);
  hipLaunchKernelGGL(search_decide_kernel, dim3(max_fields), dim3(64 * kSpecR), 0, stream, *c, *w, j1, kSpecHalf, 0);
  SegJob j1b = j1;
  j1b.r0 = kSpecHalf;
  j1b.nr = kSpecR - kSpecHalf;
  j1b.active = b.pending;
  hipLaunchKernelGGL(seg_count_kernel, dim3(segs, 1, max_fields), dim3(kSegT), 0, stream, *c, *w, j1b);
  hipLaunchKernelGGL(search_decide_kernel, dim3(max_fields), dim3(64 * kSpecR), 0, stream, *c, *w, j1, kSpecR, 1);
  hipLaunchKernelGGL(level_search_kernel, dim3(max_fields), dim3(kT), 0, stream, *c, *w);
  j1.active = b.active;
  hipLaunchKernelGGL(seg_plan_kernel, dim3(max_fields), dim3(64), 0, stream, *c, *w, j1, max_fields);
  hipLaunchKernelGGL(seg_write_kernel, dim3(segs, max_fields), dim3(kSegT), 0, stream, *c, *w, j1);
  hipLaunchKernelGGL(refine_first_kernel, dim3(max_fields), dim3(kT), 0, stream, *c, *w);
  SegJob j2{b.thr2, 1, 0, 1, nullptr, pmn, pmx, 1e300, b.active2, b.sum, segs, nullptr, b.seg_carry, b.seg_off,
            b.count2, nullptr, w->pulse_start, w->pulse_len};
  hipLaunchKernelGGL(seg_count_kernel, dim3(segs, 1, max_fields), dim3(kSegT), 0, stream, *c, *w, j2);
  hipLaunchKernelGGL(seg_plan_kernel, dim3(max_fields), dim3(64), 0, stream, *c, *w, j2, max_fields);
  hipLaunchKernelGGL(seg_write_kernel, dim3(segs, max_fields), dim3(kSegT), 0, stream, *c, *w, j2);
  hipLaunchKernelGGL(refine_second_kernel, dim3(max_fields), dim3(kT), 0, stream, *c, *w);
  // S2: the level chain
  hipLaunchKernelGGL(level_chain_kernel, dim3(1), dim3(kChainT), 0, stream, *c, *w);
  // S3: the fields' pulses at their thresholds on the DC-shifted reference
  SegJob j3{b.thr3, 1, 0, 1, b.dc3, c->pulse_min, c->pulse_max, 1e300, nullptr, b.sum, segs, nullptr, b.seg_carry,
            b.seg_off, b.count3, w->raw_count, w->pulse_start, w->pulse_len};
  hipLaunchKernelGGL(seg_count_kernel, dim3(segs, 1, max_fields), dim3(kSegT), 0, stream, *c, *w, j3);
  hipLaunchKernelGGL(seg_plan_kernel, dim3(max_fields), dim3(64), 0, stream, *c, *w, j3, max_fields);
  hipLaunchKernelGGL(seg_write_kernel, dim3(segs, max_fields), dim3(kSegT), 0, stream, *c, *w, j3);
  hipLaunchKernelGGL(field_logic_kernel, dim3(max_fields), dim3(kT), 0, stream, *c, *w);
  hipLaunchKernelGGL(field_chain_kernel, dim3(1), dim3(64), 0, stream, *c, *w);
  hipLaunchKernelGGL(field_lines_kernel, dim3(max_fields), dim3(kT), 0, stream, *c, *w);
  return hipGetLastError();
}
