// Workgroup-parallel FP32 forward-backward SOS filtering of one block.
//
// Restates SosFilter.ApplyForwardBackwardFloat32 (Dsp/SosFilter.cs:300-521,
// 584-833, 924-966, 1034-1063): odd extension of `pad` samples at both ends,
// steady-state initial conditions scaled by the first (forward) or last
// (backward) value, direct-form-II-transposed biquads evaluated with the same
// FP32 operation order.
//
// The extended sequence (length L = n + 2 pad) lives in LDS and is cut into
// chunks of C samples, one per thread: forward chunks start at index 0, backward
// chunks at index L-1, so in both directions thread 0 holds the chunk processed
// first and only the last non-empty chunk can be partial.  A partial chunk runs
// over LDS margin cells past the end of the sequence in its processing order, so
// no sample loop carries a per-sample condition and every LDS access is one base
// address plus a compile-time offset.  Each chunk is run from a zero state, the
// chunk-entry states come from a Hillis-Steele scan over the affine state maps
// (transition matrix (A^C)^(2^d) per level, precomputed in FP64 on the host),
// and each chunk is then re-run from its true entry state.  Only the injected
// entry states differ in rounding from the sequential reference (relative error
// ~1e-7).
#pragma once
#include <hip/hip_runtime.h>

#include "vhsd_device.hpp"

namespace vhsd {

constexpr int kMaxSections = 10;  // chroma band-passes up to order 10 (Video8)
constexpr int kScanThreads = 1024;
constexpr int kScanLevels = 10;  // log2(kScanThreads)
constexpr int kChunk = 33;       // supports blocks of up to 1024*33 - 2*pad samples
// LDS layout: the sequence starts kSosMargin floats into the buffer, which must hold
// kSosMargin + L + kChunk floats (partial chunks overrun by less than kChunk on
// either side).
constexpr int kSosMargin = 64;

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

// Copies the transition-matrix table to LDS (kScanLevels * (2S)^2 floats); the
// caller's next barrier publishes it.
template <int S>
__device__ __forceinline__ void load_trans(const SosScanArgs& p, float* dst) {
  constexpr int T = kScanLevels * 4 * S * S;
  for (int u = threadIdx.x; u < T; u += kScanThreads) dst[u] = p.trans[u];
}

// Exclusive scan of chunk end-states in thread order: on entry `st` is the end
// state of this thread's chunk when started from zero (thread 0: from the true
// initial state); on return it is the state entering the chunk.  `ex` is LDS space
// of 1024*2S floats, laid out component-major (conflict-free); `trans` the LDS copy
// of the level matrices.
template <int S>
__device__ __forceinline__ void scan_states(float* st, float* ex, const float* trans) {
  constexpr int W = 2 * S;
  const int q = threadIdx.x;
#pragma unroll 1
  for (int d = 0; d < kScanLevels; d++) {
#pragma unroll
    for (int c = 0; c < W; c++) ex[c * kScanThreads + q] = st[c];
    __syncthreads();
    const int off = 1 << d;
    if (q >= off) {
      float p[W];
#pragma unroll
      for (int c = 0; c < W; c++) p[c] = ex[c * kScanThreads + q - off];
      const float* m = trans + d * W * W;
#pragma unroll
      for (int r = 0; r < W; r++) {
        float acc = 0.0f;
#pragma unroll
        for (int c = 0; c < W; c++) acc += m[r * W + c] * p[c];
        st[r] += acc;
      }
    }
    __syncthreads();
  }
  // inclusive -> exclusive
#pragma unroll
  for (int c = 0; c < W; c++) ex[c * kScanThreads + q] = st[c];
  __syncthreads();
  if (q > 0) {
#pragma unroll
    for (int c = 0; c < W; c++) st[c] = ex[c * kScanThreads + q - 1];
  }
  __syncthreads();
}

// One direction over this thread's chunk val[0..kChunk) in processing order; the
// chunk of thread 0 starts from zi * first_value.
template <int S>
__device__ __forceinline__ void run_direction(const SosScanArgs& p, float (&val)[kChunk], float first_value,
                                              float* ex, const float* trans) {
  const bool lead = threadIdx.x == 0;
  float z[2 * S];
#pragma unroll
  for (int c = 0; c < 2 * S; c++) z[c] = lead ? p.zi[c] * first_value : 0.0f;
#pragma unroll
  for (int t = 0; t < kChunk; t++) {
    float v = val[t];
    sos_step<S>(p.coef, z, v);
  }
  scan_states<S>(z, ex, trans);
  if (lead) {
#pragma unroll
    for (int c = 0; c < 2 * S; c++) z[c] = p.zi[c] * first_value;  // recomputed: not kept live across the scan
  }
#pragma unroll
  for (int t = 0; t < kChunk; t++) {
    float v = val[t];
    sos_step<S>(p.coef, z, v);
    val[t] = v;
  }
}

// Filters the n samples the caller stored at ext[pad .. pad + n) (LDS, FP32; ext
// starts kSosMargin floats into a buffer of at least kSosMargin + n + 2 pad + kChunk
// floats, all of which are overwritten).  Writes the kept window of the result,
// rolled by `roll` (numpy.roll semantics), optionally mean-subtracted, to
// out[0 .. out_count) = result[out_begin ..).  Returns true (on every thread) when
// any of the n output samples is exactly zero.  `red` is 64 floats of LDS, `trans`
// the LDS copy made by load_trans.  The caller's stores to ext must be followed by
// a barrier before the call.
template <int S>
__device__ __forceinline__ bool sos_filtfilt_block_t(const SosScanArgs& p, float* ext, int n, float* red,
                                                     const float* trans, float* out, int out_begin, int out_count,
                                                     int roll, bool subtract_mean) {
  const int j = threadIdx.x;
  const int pad = p.pad;
  const int L = n + 2 * pad;
  const int C = p.chunk;
  float* ex = ext - kSosMargin;  // scan exchange: the whole buffer, free while chunks sit in registers
  // ---- odd extension: ext[i] = 2 x[0] - x[pad - i], ext[pad + n + k] = 2 x[n-1] - x[n - 2 - k]
  {
    const float first = ext[pad];
    const float last = ext[pad + n - 1];
    float e = 0.0f;
    if (j < pad) e = (2.0f * first) - ext[2 * pad - j];
    else if (j < 2 * pad) e = (2.0f * last) - ext[pad + n - 2 - (j - pad)];
    __syncthreads();
    if (j < pad) ext[j] = e;
    else if (j < 2 * pad) ext[pad + n + (j - pad)] = e;
    __syncthreads();
  }
  const bool active = j * C < L;  // chunk holds at least one sample
  float val[kChunk];
  // ---- forward: thread j owns ext[jC .. jC + C)
  float* fwd = ext + j * C;
#pragma unroll
  for (int t = 0; t < kChunk; t++) val[t] = active ? fwd[t] : 0.0f;
  const float ext_first = ext[0];
  __syncthreads();
  run_direction<S>(p, val, ext_first, ex, trans);
  if (active) {
#pragma unroll
    for (int t = 0; t < kChunk; t++) fwd[t] = val[t];
  }
  __syncthreads();
  // ---- backward: thread j owns ext[L-1-jC-t], t = 0 .. C-1 (descending)
  float* bwd = ext + (L - 1 - j * C);
#pragma unroll
  for (int t = 0; t < kChunk; t++) val[t] = active ? bwd[-t] : 0.0f;
  const float y_last = ext[L - 1];
  __syncthreads();
  run_direction<S>(p, val, y_last, ex, trans);
  // ---- zero test and sum over the n interior outputs; results back to LDS
  bool zero = false;
  float sum = 0.0f;
  if (active) {
#pragma unroll
    for (int t = 0; t < kChunk; t++) {
      bwd[-t] = val[t];
      const int i = L - 1 - j * C - t;
      const bool in = (i >= pad) && (i < pad + n);
      zero |= in && (val[t] == 0.0f);
      sum += in ? val[t] : 0.0f;
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
  const float mean = subtract_mean ? total / (float)n : 0.0f;
  // ---- write the kept window, coalesced: out[o] = y[(out_begin + o - roll) mod n]
  const float* y = ext + pad;
  for (int o = j; o < out_count; o += kScanThreads) {
    int i = out_begin + o - roll;
    if (i < 0) i += n;
    if (i >= n) i -= n;
    out[o] = y[i] - mean;
  }
  __syncthreads();
  return any_zero;
}

}  // namespace vhsd
