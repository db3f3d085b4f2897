// RF-envelope dropout detection of tape fields
// (TbcFieldDecodePipeline.DetectTapeDropouts, Decode/TbcFieldDecodePipeline.cs:6593-6642):
//
//   threshold  = float32 pairwise mean of the envelope over the field's read window
//                (NumpyReduction.MeanFloat32, Dsp/NumpyReduction.cs:196-204, 253-339)
//                x (float)fraction, or the absolute threshold
//   ranges     = RfDropoutDetector.FindDropouts (Decode/TbcDropoutMapping.cs:52-107)
//                between the field's first and last detection line
//   map rows   = TbcDropoutMapper.MapTapeRfToTbc (Decode/TbcDropoutMapping.cs:288-366)
//
// Four launches:
//  * dropout_mean_kernel, 128 workgroups per field: workgroup g covers node g of
//    depth 7 of the pairwise tree; its threads are the depth-15 slots below, and the
//    leftmost slot under each leaf sums the leaf in the reference's order (eight
//    lanes, one accumulator each, loads issued up front); the workgroup combines its
//    slots up to depth 7 and records the smallest value under the node.  dropout_threshold_kernel (one workgroup per
//    field) completes the tree to the root and writes the threshold and the
//    detection range.
//  * dropout_masks_kernel, 32 workgroups per field: every sample of the detection
//    range is classified low (<= threshold) or high (>= threshold * hysteresis) into
//    two bit masks, one wave ballot per 64 samples, with one summary bit per mask word.
//  * dropout_walk_kernel, one workgroup per field: the summaries and line locations
//    go to LDS and one wave walks the hysteresis chain (range start = first low at
//    or after the merge point, range end = first high after it) with next-set-bit
//    queries on the masks, mapping each kept range to TBC rows as it goes.
//
// Line convention: output line k of a field starts at its location k (the reference's
// location index outputFirstLine + k, TbcFieldDecodePipeline.cs:2475-2483, which is its
// detection start line LaserDiscLineOffset + 1, :6087-6095), so the detection runs over
// locations 0 .. min(count - 1, line_count + 1) and the rows carry field line k + 1.
// The read window and parity of each field come from its sequence record.
#include <hip/hip_runtime.h>

#include "vhsd_args.hpp"

namespace vhsd {
namespace {

constexpr int kSlotDepth = 15;        // leaves are at depth <= 15 (windows <= 3.6 M samples)
constexpr int kGroupDepth = 7;        // mean workgroups per field: 2^7
constexpr int kGroups = 1 << kGroupDepth;
constexpr int kSlotThreads = 1 << (kSlotDepth - kGroupDepth);
constexpr int kMaskGroups = 32;       // mask workgroups per field
constexpr int kMaskThreads = 1024;
constexpr int kWalkThreads = 256;
static_assert(kGroups == 128, "the threshold kernel reduces the group minima over two waves");

constexpr long long kMaxWindow = 3600000;  // every depth-15 node of a window's pairwise tree is a leaf

// Field f's read window [w0, w1) in span coordinates, as its record states it (the
// window DecodeReadWindowPlanner.Resolve placed, Decode/DecodeReadWindowPlanner.cs:37-63);
// false when the window is longer than the pairwise tree the kernels walk.
__device__ inline bool read_window(const DropoutArgs& a, int f, long long& w0, long long& w1) {
  const long long s = a.fields[f].window_start, l = a.fields[f].window_length;
  w0 = s < 0 ? 0 : (s > a.span_len ? a.span_len : s);
  const long long e = s + (l > 0 ? l : 0);
  w1 = e < w0 ? w0 : (e > a.span_len ? a.span_len : e);
  return w1 - w0 <= kMaxWindow;
}

// Node of the pairwise tree of n values at `level` along path bits `path` (most
// significant first): a node splits at half its length rounded down to a multiple of
// 8 while it is longer than 128.  `exists` is false below a leaf.
__device__ inline void pw_node(int n, int level, unsigned path, int& off, int& len, bool& exists) {
  off = 0;
  len = n;
  exists = true;
  for (int l = 0; l < level; l++) {
    if (len <= 128) {
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

// Like pw_node, but a path under a shallower leaf ends at that leaf; `owner` is true
// when the path is the leftmost one under it (the one that sums it).
__device__ inline void pw_descend(int n, int level, unsigned path, int& off, int& len, bool& owner) {
  off = 0;
  len = n;
  owner = true;
  for (int l = 0; l < level; l++) {
    if (len <= 128) {
      owner = (path & ((1u << (level - l)) - 1)) == 0;
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

// A leaf of NumpyReduction.PairwiseSumFloat32 (Dsp/NumpyReduction.cs:253-339).
__device__ float pw_leaf(const float* v, int len) {
#pragma clang fp contract(off)
  if (len < 8) {
    float s = -0.0f;
    for (int i = 0; i < len; i++) s += v[i];
    return s;
  }
  float r0 = v[0], r1 = v[1], r2 = v[2], r3 = v[3], r4 = v[4], r5 = v[5], r6 = v[6], r7 = v[7];
  const int vend = len - len % 8;
  for (int i = 8; i < vend; i += 8) {
    r0 += v[i];
    r1 += v[i + 1];
    r2 += v[i + 2];
    r3 += v[i + 3];
    r4 += v[i + 4];
    r5 += v[i + 5];
    r6 += v[i + 6];
    r7 += v[i + 7];
  }
  float s = ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7));
  for (int i = vend; i < len; i++) s += v[i];
  return s;
}

__global__ void __launch_bounds__(kSlotThreads) dropout_mean_kernel(DropoutArgs a) {
#pragma clang fp contract(off)
  __shared__ float v[kSlotThreads];
  __shared__ float wmin[kSlotThreads / 64];
  const int f = blockIdx.y;
  if (f >= *a.field_count) return;
  const int g = blockIdx.x, t = threadIdx.x;
  long long w0, w1;
  const bool fits = read_window(a, f, w0, w1);
  const int n = fits && a.fields[f].status == 0 ? (int)(w1 - w0) : 0;
  int goff, glen;
  bool gown;
  pw_descend(n, kGroupDepth, (unsigned)g, goff, glen, gown);
  // groups under a leaf another group owns contribute nothing (their value is never read)
  if (!(gown && n > 0)) {
    if (t == 0) {
      a.partials[(size_t)f * kGroups + g] = 0.0f;
      a.group_min[(size_t)f * kGroups + g] = __builtin_inff();
    }
    return;
  }
  const float* src = a.envelope + w0 + goff;
  float m = __builtin_inff();  // smallest value this thread loaded (fminf skips NaN)
  float gval = 0.0f;
  if (glen <= 128) {
    // the group's node is itself a leaf (at depth 7 or above)
    for (int j = t; j < glen; j += kSlotThreads) m = fminf(m, src[j]);
    if (t == 0) gval = pw_leaf(src, glen);
  } else {
    // leaves: eight lanes per leaf, lane k running accumulator k (elements k, k+8, ...)
    // with all of its loads issued first; the three xor-shuffles combine the
    // accumulators as ((r0+r1)+(r2+r3))+((r4+r5)+(r6+r7))
    const int oct = t >> 3, k8 = t & 7;
    for (int q = 0; q < 8; q++) {
      const int slot = oct * 8 + q;
      int off, len;
      bool own;
      pw_descend(glen, kSlotDepth - kGroupDepth, (unsigned)slot, off, len, own);
      const bool vec = own && len >= 8 && len <= 128;
      const int vend = len - len % 8;
      float x[16];
#pragma unroll
      for (int it = 0; it < 16; it++) {
        const int i = k8 + 8 * it;
        x[it] = (vec && i < vend) ? src[off + i] : 0.0f;
      }
      float r = x[0];
#pragma unroll
      for (int it = 1; it < 16; it++) {
        if (k8 + 8 * it < vend) r += x[it];
      }
      if (vec) {
#pragma unroll
        for (int it = 0; it < 16; it++) {
          if (k8 + 8 * it < vend) m = fminf(m, x[it]);
        }
      }
      r = r + __shfl_xor(r, 1);
      r = r + __shfl_xor(r, 2);
      r = r + __shfl_xor(r, 4);
      if (k8 == 0) {
        float sum = 0.0f;
        if (own && len <= 128) {
          if (len < 8) {
            sum = -0.0f;
            for (int i = 0; i < len; i++) {
              const float e = src[off + i];
              sum += e;
              m = fminf(m, e);
            }
          } else {
            sum = r;
            for (int i = vend; i < len; i++) {
              const float e = src[off + i];
              sum += e;
              m = fminf(m, e);
            }
          }
        }
        v[slot] = sum;
      }
    }
    __syncthreads();
    for (int level = kSlotDepth - 1; level >= kGroupDepth; level--) {
      const int cnt = 1 << (level - kGroupDepth);
      if (t < cnt) {
        int o2, l2;
        bool ex;
        pw_node(glen, level - kGroupDepth, (unsigned)t, o2, l2, ex);
        if (ex && l2 > 128) {
          const int left = t << (kSlotDepth - level);
          const int right = (2 * t + 1) << (kSlotDepth - 1 - level);
          v[left] = v[left] + v[right];
        }
      }
      __syncthreads();
    }
    gval = v[0];
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) m = fminf(m, __shfl_xor(m, d));
  if ((t & 63) == 0) wmin[t >> 6] = m;
  __syncthreads();
  if (t == 0) {
    float gm = wmin[0];
    for (int w = 1; w < kSlotThreads / 64; w++) gm = fminf(gm, wmin[w]);
    a.group_min[(size_t)f * kGroups + g] = gm;
    a.partials[(size_t)f * kGroups + g] = gval;
  }
}

// One workgroup per field: the pairwise tree from depth 7 to the root, then the
// threshold and the detection range.
__global__ void __launch_bounds__(kGroups) dropout_threshold_kernel(DropoutArgs a) {
#pragma clang fp contract(off)
  __shared__ float v[kGroups];
  __shared__ float wm[kGroups / 64];
  const int f = blockIdx.x;
  if (f >= *a.field_count) return;
  const int t = threadIdx.x;
  long long w0, w1;
  const bool fits = read_window(a, f, w0, w1);
  const long long n = fits ? w1 - w0 : 0;
  v[t] = a.partials[(size_t)f * kGroups + t];
  float m = a.group_min[(size_t)f * kGroups + t];
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) m = fminf(m, __shfl_xor(m, d));
  if ((t & 63) == 0) wm[t >> 6] = m;
  __syncthreads();
  for (int level = kGroupDepth - 1; level >= 0; level--) {
    const int cnt = 1 << level;
    if (t < cnt) {
      int o2, l2;
      bool ex;
      pw_node((int)n, level, (unsigned)t, o2, l2, ex);
      if (ex && l2 > 128) {
        const int left = t << (kGroupDepth - level);
        const int right = (2 * t + 1) << (kGroupDepth - 1 - level);
        v[left] = v[left] + v[right];
      }
    }
    __syncthreads();
  }
  if (t != 0) return;
  // NumpyReduction.MeanFloat32: float sum / count (NaN for an empty window)
  const float mean = n > 0 ? v[0] / (float)n : __builtin_nanf("");
  const double thr = a.absolute_threshold >= 0.0 ? a.absolute_threshold : (double)(mean * (float)a.fraction);
  const int nloc = a.nloc;
  const double* loc = a.locations + (size_t)f * nloc;
  const int line_count = a.fields[f].is_first_field ? a.first_lines : a.second_lines;
  const int end_line = nloc - 1 < line_count + 1 ? nloc - 1 : line_count + 1;
  long long a0 = 0, b0 = 0;
  if (end_line > 0) {
    a0 = (long long)floor(loc[0]);
    a0 = a0 < w0 ? w0 : (a0 > w1 ? w1 : a0);
    b0 = (long long)ceil(loc[end_line]);
    b0 = b0 < a0 ? a0 : (b0 > w1 ? w1 : b0);
  }
  const long long span = b0 - a0;
  const int nwords = (int)((span + 63) >> 6);
  DropoutField inf;
  inf.threshold = thr;
  inf.up = thr * a.hysteresis;
  inf.a0 = a0;
  inf.b0 = b0;
  inf.nwords = nwords;
  inf.end_line = end_line;
  // a failed record (no field decoded) has no rows
  inf.status = a.fields[f].status != 0 ? 0
               : !fits                 ? -1
               : span <= 0             ? 0
               : ((nwords > a.words_cap || nwords > kDropSummaryWords * 32) ? -1 : 1);
  // no window sample at or below the threshold: no range can open, the masks and the
  // walk have nothing to do
  const float window_min = fminf(wm[0], wm[1]);
  if (inf.status == 1 && (double)window_min > thr) inf.status = 0;
  a.info[f] = inf;
}

__global__ void __launch_bounds__(kMaskThreads) dropout_masks_kernel(DropoutArgs a) {
#pragma clang fp contract(off)
  __shared__ unsigned lsum[kDropSummaryWords / kMaskGroups];
  __shared__ unsigned hsum[kDropSummaryWords / kMaskGroups];
  const int f = blockIdx.y;
  if (f >= *a.field_count) return;
  const int g = blockIdx.x, t = threadIdx.x;
  const DropoutField inf = a.info[f];
  if (inf.status != 1) return;
  const double thr = inf.threshold, up = inf.up;
  const long long a0 = inf.a0, b0 = inf.b0;
  const int nwords = inf.nwords;
  const int nsw = (nwords + 31) >> 5;
  const int per = (nsw + kMaskGroups - 1) / kMaskGroups;  // summary words per workgroup (<= 128)
  const int sw0 = g * per;
  const int sw1 = sw0 + per < nsw ? sw0 + per : nsw;
  if (sw0 >= sw1) return;
  if (t < per) {
    lsum[t] = 0u;
    hsum[t] = 0u;
  }
  __syncthreads();
  unsigned long long* lowm = a.masks + (size_t)f * 2 * a.words_cap;
  unsigned long long* highm = lowm + a.words_cap;
  const int lane = t & 63, wave = t >> 6;
  const int wbeg = sw0 << 5;
  const int wend = (sw1 << 5) < nwords ? (sw1 << 5) : nwords;
  // each wave takes eight words at a time, the eight loads in flight together
  constexpr int kWaves = kMaskThreads / 64;
  for (int w8 = wbeg + 8 * wave; w8 < wend; w8 += 8 * kWaves) {
    float x[8];
    bool in[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const long long i = a0 + ((long long)(w8 + k) << 6) + lane;
      in[k] = w8 + k < wend && i < b0;
      x[k] = in[k] ? a.envelope[i] : 0.0f;
    }
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const int w = w8 + k;
      const double xd = (double)x[k];
      const unsigned long long lo = __ballot(in[k] && xd <= thr);
      const unsigned long long hi = __ballot(in[k] && xd >= up);
      if (lane == 0 && w < wend) {
        lowm[w] = lo;
        highm[w] = hi;
        if (lo) atomicOr(&lsum[(w >> 5) - sw0], 1u << (w & 31));
        if (hi) atomicOr(&hsum[(w >> 5) - sw0], 1u << (w & 31));
      }
    }
  }
  __syncthreads();
  unsigned* gs = a.summaries + (size_t)f * 2 * kDropSummaryWords;
  if (t < sw1 - sw0) {
    gs[sw0 + t] = lsum[t];
    gs[kDropSummaryWords + sw0 + t] = hsum[t];
  }
}

// First set bit at relative position >= rel in a bit mask of nwords 64-bit words with
// one summary bit per word (set when the word is nonzero), or -1.  Called by all 64
// lanes of one wave with the same arguments.
__device__ long long next_set(const unsigned long long* mask, const unsigned* sum, int nwords, long long rel) {
  if (rel >= (long long)nwords * 64) return -1;
  const int w = (int)(rel >> 6);
  const unsigned long long bits = mask[w] & (~0ull << (rel & 63));
  if (bits) return ((long long)w << 6) + __ffsll(bits) - 1;
  const int q = w + 1;
  if (q >= nwords) return -1;
  const int nsw = (nwords + 31) >> 5;
  const int sw = q >> 5;
  const int lane = threadIdx.x & 63;
  for (int base = sw; base < nsw; base += 64) {
    const int idx = base + lane;
    unsigned s = idx < nsw ? sum[idx] : 0u;
    if (idx == sw) s &= ~0u << (q & 31);
    const unsigned long long b = __ballot(s != 0u);
    if (b) {
      const int l = __ffsll(b) - 1;
      const unsigned sv = __shfl(s, l);
      const int word = ((base + l) << 5) + __ffs(sv) - 1;
      return ((long long)word << 6) + __ffsll(mask[word]) - 1;
    }
  }
  return -1;
}

__global__ void __launch_bounds__(kWalkThreads) dropout_walk_kernel(DropoutArgs a) {
#pragma clang fp contract(off)
  __shared__ double loc[kDropMaxLocations];
  __shared__ unsigned lsum[kDropSummaryWords];
  __shared__ unsigned hsum[kDropSummaryWords];
  const int f = blockIdx.x;
  if (f >= *a.field_count) return;
  const int t = threadIdx.x;
  const DropoutField inf = a.info[f];
  if (inf.status != 1) {
    if (t == 0) {
      a.row_count[f] = inf.status;
      if (a.range_count) a.range_count[f] = inf.status;
    }
    return;
  }
  const int nloc = a.nloc;
  for (int i = t; i < nloc; i += kWalkThreads) loc[i] = a.locations[(size_t)f * nloc + i];
  const int nwords = inf.nwords;
  const int nsw = (nwords + 31) >> 5;
  const unsigned* gs = a.summaries + (size_t)f * 2 * kDropSummaryWords;
  for (int i = t; i < nsw; i += kWalkThreads) {
    lsum[i] = gs[i];
    hsum[i] = gs[kDropSummaryWords + i];
  }
  __syncthreads();
  if (t >= 64) return;
  const int lane = t;
  const unsigned long long* lowm = a.masks + (size_t)f * 2 * a.words_cap;
  const unsigned long long* highm = lowm + a.words_cap;
  const long long a0 = inf.a0, span = inf.b0 - inf.a0;
  const int end_line = inf.end_line;
  const int cap = a.capacity;
  const int out_len = a.out_len;
  const size_t rbase = (size_t)f * cap;
  int nr = 0, rows = 0;
  bool mapping = true;
  int li = 0;
  double ls = loc[0], le = loc[1];
  int pend_line = 0, pend_x = 0;
  auto emit = [&](int ex) {
    if (rows < cap && lane == 0) {
      a.field_line[rbase + rows] = pend_line;
      a.start_x[rbase + rows] = pend_x;
      a.end_x[rbase + rows] = ex;
    }
    rows++;
  };
  long long p = 0;
  while (true) {
    const long long s = next_set(lowm, lsum, nwords, p);
    if (s < 0) break;
    long long e = next_set(highm, hsum, nwords, s + 1);
    const bool open_end = e < 0;
    if (open_end) e = span;
    if (e - s > a.min_len) {
      if (a.range_start && nr < cap && lane == 0) {
        a.range_start[rbase + nr] = (int)(a0 + s);
        a.range_end[rbase + nr] = (int)(a0 + e);
      }
      nr++;
      if (mapping) {
        // TbcDropoutMapper.MapTapeRfToTbc: the start line is searched forward from the
        // previous range's end line; rows run line by line to the line holding the end
        const double S = (double)(a0 + s), E = (double)(a0 + e);
        bool found = false;
        while (li < end_line) {
          const double len = le - ls;
          if ((S >= ls || li == 0) && S < le && len > 0.0) {
            const int px = (int)floor(((S - ls) / len) * out_len);
            pend_line = li + 1;
            pend_x = px > 0 ? px : 0;
            found = true;
            break;
          }
          li++;
          if (li < end_line) {
            ls = loc[li];
            le = loc[li + 1];
          }
        }
        if (!found) {
          mapping = false;
        } else {
          while (li < end_line) {
            const double len = le - ls;
            if (E < le && len > 0.0) {
              const int px = (int)ceil(((E - ls) / len) * out_len);
              emit(px < out_len ? px : out_len);
              break;
            }
            emit(out_len);
            li++;
            if (li < end_line) {
              ls = loc[li];
              le = loc[li + 1];
              pend_line = li + 1;
              pend_x = 0;
            }
          }
        }
      }
    }
    if (open_end) break;
    p = e + a.merge + 1;
  }
  if (lane == 0) {
    a.row_count[f] = rows;
    if (a.range_count) a.range_count[f] = nr;
  }
}

}  // namespace
}  // namespace vhsd

hipError_t vhsd_launch_dropouts(const vhsd::DropoutArgs* a, int max_fields, hipStream_t stream) {
  if (max_fields <= 0) return hipSuccess;
  if (a->nloc < 2 || a->nloc > vhsd::kDropMaxLocations || a->capacity <= 0 || a->words_cap <= 0)
    return hipErrorInvalidValue;
  hipLaunchKernelGGL(vhsd::dropout_mean_kernel, dim3(vhsd::kGroups, max_fields), dim3(vhsd::kSlotThreads), 0,
                     stream, *a);
  hipLaunchKernelGGL(vhsd::dropout_threshold_kernel, dim3(max_fields), dim3(vhsd::kGroups), 0, stream, *a);
  hipLaunchKernelGGL(vhsd::dropout_masks_kernel, dim3(vhsd::kMaskGroups, max_fields), dim3(vhsd::kMaskThreads), 0,
                     stream, *a);
  hipLaunchKernelGGL(vhsd::dropout_walk_kernel, dim3(max_fields), dim3(vhsd::kWalkThreads), 0, stream, *a);
  return hipGetLastError();
}
