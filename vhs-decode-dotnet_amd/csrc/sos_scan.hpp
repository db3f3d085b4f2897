// Workgroup-parallel FP32 forward-backward SOS filtering of one block.
//
// Restates SosFilter.ApplyForwardBackwardFloat32 (Dsp/SosFilter.cs:300-521,
// 584-833, 924-966, 1034-1063): odd extension of `pad` samples at both ends,
// steady-state initial conditions scaled by the first (forward) or last
// (backward) value, direct-form-II-transposed biquads evaluated with the same
// FP32 operation order.  The sequential recurrence is split into 1024 chunks,
// one per thread: each chunk is run from a zero state, the chunk-entry states
// come from a Hillis-Steele scan over the affine state maps (transition matrix
// (A^C)^(2^d) per level, precomputed in FP64 on the host), and each chunk is then
// re-run from its true entry state.  Only the injected entry states differ in
// rounding from the sequential reference (relative error ~1e-7).
#pragma once
#include <hip/hip_runtime.h>

#include "vhsd_device.hpp"

namespace vhsd {

constexpr int kMaxSections = 4;
constexpr int kScanThreads = 1024;
constexpr int kScanLevels = 10;  // log2(kScanThreads)
constexpr int kChunk = 33;       // supports blocks of up to 1024*33 - 2*pad samples

struct SosScanArgs {
  int sections;                          // 1..kMaxSections
  int pad;                               // odd-extension length (DefaultPadLength)
  int chunk;                             // samples per thread chunk, <= kChunk
  float coef[kMaxSections][5];           // b0, b1, b2, a1, a2 (FP32 cast of the FP64 design)
  float zi[2 * kMaxSections];            // steady-state initial conditions (FP32)
  const float* trans;                    // [kScanLevels][2S][2S] row-major, A^(chunk * 2^d)
};

template <int S>
__device__ __forceinline__ void sos_step(const float (&c)[kMaxSections][5], float* z, float& v) {
#pragma clang fp contract(off)
#pragma unroll
  for (int k = 0; k < S; k++) {
    const float f = (c[k][0] * v) + z[2 * k];
    z[2 * k] = ((c[k][1] * v) - (c[k][3] * f)) + z[2 * k + 1];
    z[2 * k + 1] = (c[k][2] * v) - (c[k][4] * f);
    v = f;
  }
}

// Exclusive scan of chunk end-states: returns in `st` the state entering this
// thread's chunk, given `st` = end state of the chunk when started from zero
// (thread 0: from the true initial state).  `ex` is LDS space of 1024*2S floats.
template <int S>
__device__ __forceinline__ void scan_states(float* st, float* ex, const float* __restrict__ trans) {
  const int j = threadIdx.x;
  constexpr int W = 2 * S;
#pragma unroll 1
  for (int d = 0; d < kScanLevels; d++) {
#pragma unroll
    for (int q = 0; q < W; q++) ex[j * W + q] = st[q];
    __syncthreads();
    const int off = 1 << d;
    if (j >= off) {
      float p[W];
#pragma unroll
      for (int q = 0; q < W; q++) p[q] = ex[(j - off) * W + q];
      const float* m = trans + d * W * W;
#pragma unroll
      for (int r = 0; r < W; r++) {
        float acc = 0.0f;
#pragma unroll
        for (int q = 0; q < W; q++) acc += m[r * W + q] * p[q];
        st[r] += acc;
      }
    }
    __syncthreads();
  }
  // inclusive -> exclusive
#pragma unroll
  for (int q = 0; q < W; q++) ex[j * W + q] = st[q];
  __syncthreads();
  if (j > 0) {
#pragma unroll
    for (int q = 0; q < W; q++) st[q] = ex[(j - 1) * W + q];
  }
  __syncthreads();
}

// One direction over this thread's chunk values `val[0..len)` (in processing
// order).  Thread 0 starts from zi * first value of the whole sequence.
template <int S>
__device__ __forceinline__ void run_direction(const SosScanArgs& p, float* val, int len, float first_value,
                                              float* ex) {
  const int j = threadIdx.x;
  float z[2 * S];
  float init[2 * S];
#pragma unroll
  for (int q = 0; q < 2 * S; q++) {
    init[q] = p.zi[q] * first_value;
    z[q] = (j == 0) ? init[q] : 0.0f;
  }
#pragma unroll
  for (int t = 0; t < kChunk; t++) {
    if (t < len) {
      float v = val[t];
      sos_step<S>(p.coef, z, v);
    }
  }
  scan_states<S>(z, ex, p.trans);
  if (j == 0) {
#pragma unroll
    for (int q = 0; q < 2 * S; q++) z[q] = init[q];
  }
#pragma unroll
  for (int t = 0; t < kChunk; t++) {
    if (t < len) {
      float v = val[t];
      sos_step<S>(p.coef, z, v);
      val[t] = v;
    }
  }
}

// Filter N samples held in LDS `raw` (FP32).  Writes the kept window of the
// result, rolled by `roll` (numpy.roll semantics), optionally mean-subtracted,
// to out[0 .. out_count) = result[out_begin ..).  Returns true (on every
// thread) when any output sample is exactly zero.
template <int S>
__device__ bool sos_filtfilt_block_t(const SosScanArgs& p, float* raw, int n, float* scratch, float* out,
                                     int out_begin, int out_count, int roll, bool subtract_mean) {
  const int j = threadIdx.x;
  const int pad = p.pad;
  const int L = n + 2 * pad;
  const int C = p.chunk;
  float* head = scratch;          // [pad] extension values, then results
  float* tail = scratch + 64;     // [pad]
  float* red = scratch + 128;     // reductions
  float val[kChunk];
  // ---- forward: chunk j covers ext indices [j*C, j*C + len)
  const int fbeg = j * C;
  int len = L - fbeg;
  len = len < 0 ? 0 : (len > C ? C : len);
  const float first = raw[0];
  const float last = raw[n - 1];
#pragma unroll
  for (int t = 0; t < kChunk; t++) {
    const int i = fbeg + t;
    float e = 0.0f;
    if (t < len) {
      if (i < pad)
        e = (2.0f * first) - raw[pad - i];
      else if (i < pad + n)
        e = raw[i - pad];
      else
        e = (2.0f * last) - raw[n - 2 - (i - pad - n)];
    }
    val[t] = e;
  }
  const float ext0 = (2.0f * first) - raw[pad];
  const float ext_first = pad > 0 ? ext0 : first;
  __syncthreads();  // raw no longer needed: its space becomes the scan exchange
  run_direction<S>(p, val, len, ext_first, raw);
  // ---- exchange forward results into LDS (interior) and scratch (extensions)
#pragma unroll
  for (int t = 0; t < kChunk; t++) {
    if (t < len) {
      const int i = fbeg + t;
      if (i < pad)
        head[i] = val[t];
      else if (i < pad + n)
        raw[i - pad] = val[t];
      else
        tail[i - pad - n] = val[t];
    }
  }
  __syncthreads();
  // ---- backward: reversed chunk j covers ext indices i = L-1-r, r in [j*C, j*C + len)
  int blen = L - j * C;
  blen = blen < 0 ? 0 : (blen > C ? C : blen);
  const float y_last = pad > 0 ? tail[pad - 1] : raw[n - 1];
#pragma unroll
  for (int t = 0; t < kChunk; t++) {
    const int i = L - 1 - (j * C + t);
    float e = 0.0f;
    if (t < blen) {
      if (i < pad)
        e = head[i];
      else if (i < pad + n)
        e = raw[i - pad];
      else
        e = tail[i - pad - n];
    }
    val[t] = e;
  }
  __syncthreads();
  run_direction<S>(p, val, blen, y_last, raw);
  // ---- zero test and mean over the N interior outputs
  bool zero = false;
  float sum = 0.0f;
#pragma unroll
  for (int t = 0; t < kChunk; t++) {
    const int i = L - 1 - (j * C + t);
    if (t < blen && i >= pad && i < pad + n) {
      zero |= (val[t] == 0.0f);
      sum += val[t];
    }
  }
  const int wave = j >> 6;
  const unsigned long long zb = __ballot(zero);
  float ws = sum;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) ws += __shfl_xor(ws, o);
  if ((j & 63) == 0) {
    red[wave] = ws;
    red[32 + wave] = zb ? 1.0f : 0.0f;
  }
  __syncthreads();
  float total = 0.0f;
  bool any_zero = false;
  for (int w = 0; w < kScanThreads / 64; w++) {
    total += red[w];
    any_zero |= red[32 + w] != 0.0f;
  }
  const float mean = total / (float)n;
  // ---- write the kept window
#pragma unroll
  for (int t = 0; t < kChunk; t++) {
    const int i = L - 1 - (j * C + t);
    if (t < blen && i >= pad && i < pad + n) {
      int s = i - pad + roll;
      s %= n;
      if (s < 0) s += n;
      const int o = s - out_begin;
      if (o >= 0 && o < out_count) out[o] = subtract_mean ? (val[t] - mean) : val[t];
    }
  }
  __syncthreads();
  return any_zero;
}

__device__ __forceinline__ bool sos_filtfilt_block(const SosScanArgs& p, float* raw, int n, float* scratch,
                                                   float* out, int out_begin, int out_count, int roll,
                                                   bool subtract_mean = false) {
  switch (p.sections) {
    case 1: return sos_filtfilt_block_t<1>(p, raw, n, scratch, out, out_begin, out_count, roll, subtract_mean);
    case 2: return sos_filtfilt_block_t<2>(p, raw, n, scratch, out, out_begin, out_count, roll, subtract_mean);
    case 3: return sos_filtfilt_block_t<3>(p, raw, n, scratch, out, out_begin, out_count, roll, subtract_mean);
    default: return sos_filtfilt_block_t<4>(p, raw, n, scratch, out, out_begin, out_count, roll, subtract_mean);
  }
}

}  // namespace vhsd
