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
// Two launches:
//  * dropout_mean_slots_kernel, 64 workgroups per field: the pairwise tree of the
//    window is cut at depth 15; every depth-15 slot (or the leftmost slot under a
//    shallower leaf) sums its leaf in the reference's order (eight accumulators), and
//    each workgroup combines its 512 slots up to depth 6 in LDS.
//  * dropout_fields_kernel, one workgroup per field: finishes the tree (depth 6 to
//    the root), classifies every sample of the detection range as low (<= threshold)
//    or high (>= threshold * hysteresis) into two bit masks (one wave ballot per 64
//    samples, plus one summary bit per mask word in LDS), then one wave walks the
//    hysteresis chain (range start = first low at or after the merge point, range
//    end = first high after it) with next-set-bit queries on the masks and maps
//    each kept range to TBC rows against the field's line locations in LDS.
//
// Line convention: output line k of a field starts at its location k (the reference's
// location index outputFirstLine + k, TbcFieldDecodePipeline.cs:2475-2483, which is its
// detection start line LaserDiscLineOffset + 1, :6087-6095), so the detection runs over
// locations 0 .. min(count - 1, line_count + 1) and the rows carry field line k + 1.
#include <hip/hip_runtime.h>

#include "vhsd_args.hpp"

namespace vhsd {
namespace {

constexpr int kMeanDepth = 15;     // slots per field: 2^15
constexpr int kMeanGroupDepth = 6; // workgroups per field: 2^6
constexpr int kMeanThreads = 1 << (kMeanDepth - kMeanGroupDepth);
constexpr int kDropThreads = 1024;

// Field f's read window [w0, w1) in span coordinates: DecodeReadWindowPlanner.Resolve
// (Decode/DecodeReadWindowPlanner.cs:37-63) of the intended start lead_lines before
// line 0, with the requested read length; the same arithmetic places the TBC origin.
__device__ inline void read_window(const DropoutArgs& a, double line0, long long& w0, long long& w1) {
  double intended = line0 - a.lead_lines * a.line_len;
  intended = intended > 0.0 ? intended : 0.0;
  long long read = (long long)intended - a.block_cut;
  read = read > 0 ? read : 0;
  const long long aligned = (read / a.block_length) * a.block_length;
  const long long end = aligned + a.read_samples;
  w0 = (aligned / a.stride) * a.stride;
  w1 = (end / a.stride + 1) * a.stride;
  w0 = w0 < a.span_len ? w0 : a.span_len;
  w1 = w1 < a.span_len ? w1 : a.span_len;
}

// Node of the pairwise tree of n values at `level` along path bits `path` (most
// significant first): a node splits at half its length rounded down to a multiple of
// 8 while it is longer than 128.  `exists` is false below a leaf.
__device__ inline void pw_node(long long n, int level, unsigned path, long long& off, long long& len, bool& exists) {
  off = 0;
  len = n;
  exists = true;
  for (int l = 0; l < level; l++) {
    if (len <= 128) {
      exists = false;
      return;
    }
    const long long half = len / 2;
    const long long split = half - half % 8;
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

__global__ void __launch_bounds__(kMeanThreads) dropout_mean_slots_kernel(DropoutArgs a) {
#pragma clang fp contract(off)
  const int f = blockIdx.y;
  if (f >= *a.field_count) return;
  __shared__ float v[kMeanThreads];
  long long w0, w1;
  read_window(a, a.line0[f], w0, w1);
  const long long n = w1 - w0;
  const int t = threadIdx.x;
  const unsigned slot = (blockIdx.x << (kMeanDepth - kMeanGroupDepth)) | t;
  // descend to this slot's leaf; only the leftmost slot under a leaf sums it
  long long off = 0, len = n;
  bool mine = true;
  for (int l = 0; l < kMeanDepth; l++) {
    if (len <= 128) {
      mine = (slot & ((1u << (kMeanDepth - l)) - 1)) == 0;
      break;
    }
    const long long half = len / 2;
    const long long split = half - half % 8;
    if ((slot >> (kMeanDepth - 1 - l)) & 1u) {
      off += split;
      len -= split;
    } else {
      len = split;
    }
  }
  // windows are limited on the host so that every depth-15 node is a leaf (len <= 128)
  v[t] = (mine && n > 0 && len <= 128) ? pw_leaf(a.envelope + w0 + off, (int)len) : 0.0f;
  __syncthreads();
  for (int level = kMeanDepth - 1; level >= kMeanGroupDepth; level--) {
    const int cnt = 1 << (level - kMeanGroupDepth);
    if (t < cnt) {
      const unsigned j = (blockIdx.x << (level - kMeanGroupDepth)) | t;
      long long o2, l2;
      bool ex;
      pw_node(n, level, j, o2, l2, ex);
      if (ex && l2 > 128) {
        const int left = t << (kMeanDepth - level);
        const int right = (2 * t + 1) << (kMeanDepth - 1 - level);
        v[left] = v[left] + v[right];
      }
    }
    __syncthreads();
  }
  if (t == 0) a.partials[(size_t)f * (1 << kMeanGroupDepth) + blockIdx.x] = v[0];
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

__global__ void __launch_bounds__(kDropThreads) dropout_fields_kernel(DropoutArgs a) {
#pragma clang fp contract(off)
  const int f = blockIdx.x;
  if (f >= *a.field_count) return;
  __shared__ double loc[kDropMaxLocations];
  __shared__ float pv[1 << kMeanGroupDepth];
  __shared__ unsigned lsum[kDropSummaryWords];
  __shared__ unsigned hsum[kDropSummaryWords];
  const int t = threadIdx.x;
  const int nloc = a.nloc;
  for (int i = t; i < nloc; i += kDropThreads) loc[i] = a.locations[(size_t)f * nloc + i];
  if (t < (1 << kMeanGroupDepth)) pv[t] = a.partials[(size_t)f * (1 << kMeanGroupDepth) + t];
  long long w0, w1;
  read_window(a, a.line0[f], w0, w1);
  const long long n = w1 - w0;
  __syncthreads();
  for (int level = kMeanGroupDepth - 1; level >= 0; level--) {
    const int cnt = 1 << level;
    if (t < cnt) {
      long long o2, l2;
      bool ex;
      pw_node(n, level, (unsigned)t, o2, l2, ex);
      if (ex && l2 > 128) {
        const int left = t << (kMeanGroupDepth - level);
        const int right = (2 * t + 1) << (kMeanGroupDepth - 1 - level);
        pv[left] = pv[left] + pv[right];
      }
    }
    __syncthreads();
  }
  // NumpyReduction.MeanFloat32: float sum / count (NaN for an empty window)
  const float mean = n > 0 ? pv[0] / (float)n : __builtin_nanf("");
  const double thr = a.absolute_threshold >= 0.0 ? a.absolute_threshold : (double)(mean * (float)a.fraction);
  const double up = thr * a.hysteresis;
  const int line_count = a.field_first[f] ? a.first_lines : a.second_lines;
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
  if (span <= 0 || nwords > a.words_cap || nwords > kDropSummaryWords * 32) {
    if (t == 0) {
      // a detection range longer than the mask capacity is reported as -1
      const int rc = span <= 0 ? 0 : -1;
      a.row_count[f] = rc;
      if (a.range_count) a.range_count[f] = rc;
    }
    return;
  }
  const int nsw = (nwords + 31) >> 5;
  for (int i = t; i < nsw; i += kDropThreads) {
    lsum[i] = 0u;
    hsum[i] = 0u;
  }
  __syncthreads();
  unsigned long long* lowm = a.masks + (size_t)f * 2 * a.words_cap;
  unsigned long long* highm = lowm + a.words_cap;
  const int lane = t & 63, wave = t >> 6;
  for (int w = wave; w < nwords; w += kDropThreads / 64) {
    const long long i = a0 + ((long long)w << 6) + lane;
    const bool in = i < b0;
    const double x = in ? (double)a.envelope[i] : 0.0;
    const unsigned long long lo = __ballot(in && x <= thr);
    const unsigned long long hi = __ballot(in && x >= up);
    if (lane == 0) {
      lowm[w] = lo;
      highm[w] = hi;
      if (lo) atomicOr(&lsum[w >> 5], 1u << (w & 31));
      if (hi) atomicOr(&hsum[w >> 5], 1u << (w & 31));
    }
  }
  __syncthreads();
  if (wave != 0) return;
  // ---- one wave: hysteresis chain, then the TBC rows of each kept range
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
  hipLaunchKernelGGL(vhsd::dropout_mean_slots_kernel, dim3(1 << vhsd::kMeanGroupDepth, max_fields),
                     dim3(vhsd::kMeanThreads), 0, stream, *a);
  hipLaunchKernelGGL(vhsd::dropout_fields_kernel, dim3(max_fields), dim3(vhsd::kDropThreads), 0, stream, *a);
  return hipGetLastError();
}
