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
// address plus a compile-time offset.
//
// A cascade processed sample by sample gives the same values as its sections
// applied one after another over the whole sequence (each section's step reads
// only its own state and its input), so each direction runs the sections in turn
// on the chunk held in registers.  Per section, each chunk runs once from a zero
// state (thread 0 from the section's true initial state); the chunk-entry states
// come from a scan over the section's affine state maps (2 x 2 matrices
// (A^C)^(2^d), precomputed in FP64 on the host and read as uniform scalars): a
// shuffle scan inside each wave, a scan of the 16 wave totals, and a second
// shuffle scan with each wave's entry state injected.  The zero-input response of
// the section to its entry state (a precomputed C x 2 table) is then added to the
// chunk.  Relative to the sequential reference the outputs differ by rounding
// (relative error ~1e-6).
#pragma once
#include <hip/hip_runtime.h>

#include "vhsd_device.hpp"

namespace vhsd {

constexpr int kMaxSections = 10;  // chroma band-passes up to order 10 (Video8)
constexpr int kScanThreads = 1024;
constexpr int kScanLevels = 10;  // log2(kScanThreads)
constexpr int kChunk = 33;       // supports blocks of up to 1024*33 - 2*pad samples
// LDS layout: the sequence starts kSosMargin floats into the buffer, which must hold
// kSosMargin + L + kChunk + 64 floats (partial chunks overrun by less than kChunk on
// either side; the scan exchange follows).
constexpr int kSosMargin = 64;

struct SosScanArgs {
  int sections;                          // 1..kMaxSections
  int pad;                               // odd-extension length (DefaultPadLength)
  int chunk;                             // samples per thread chunk, <= kChunk
  float coef[kMaxSections][5];           // b0, b1, b2, a1, a2 (FP32 cast of the FP64 design)
  float zi[2 * kMaxSections];            // steady-state initial conditions (FP32)
  const float* trans;                    // per section: [kScanLevels][2][2] row-major, A^(chunk * 2^d);
                                         // then per section [kChunk][2]: output t of a chunk per
                                         // unit entry state
};

// y += M x for a 2 x 2 section transition.
__device__ __forceinline__ void add_matvec2(float* y, const float* __restrict__ m, const float* x) {
  const float y0 = fmaf(m[0], x[0], m[1] * x[1]);
  const float y1 = fmaf(m[2], x[0], m[3] * x[1]);
  y[0] += y0;
  y[1] += y1;
}

// Inclusive scan of affine chunk maps across the lanes of one wave with a zero state
// entering lane 0: x := x + M_d x[lane - 2^d] for d = 0..5.
__device__ __forceinline__ void wave_scan2(float* x, const float* __restrict__ trans, int lane) {
#pragma unroll
  for (int d = 0; d < 6; d++) {
    float pv[2];
    pv[0] = __shfl_up(x[0], 1u << d);
    pv[1] = __shfl_up(x[1], 1u << d);
    if (lane >= (1 << d)) add_matvec2(x, trans + d * 4, pv);
  }
}

// Exclusive scan of one section's chunk end-states in thread order: on entry `st` is
// the end state of this thread's chunk when started from zero (thread 0: from the
// true initial state); on return it is the state entering the chunk.  `ex` is LDS
// space of at least 64 floats; `trans` the section's level matrices (global memory).
__device__ __forceinline__ void scan_states2(float* st, float* ex, const float* __restrict__ trans) {
  constexpr int kWaves = kScanThreads / 64;
  static_assert(kWaves == 16, "the wave totals are scanned by 16 lanes in four levels");
  const int q = threadIdx.x, lane = q & 63, wave = q >> 6;
  float x[2] = {st[0], st[1]};
  wave_scan2(x, trans, lane);
  if (lane == 63) {
    ex[wave * 2] = x[0];
    ex[wave * 2 + 1] = x[1];
  }
  __syncthreads();
  if (wave == 0) {
    // wave totals: one wave step spans 64 chunks, levels 6..9
    float tot[2];
    tot[0] = lane < kWaves ? ex[lane * 2] : 0.0f;
    tot[1] = lane < kWaves ? ex[lane * 2 + 1] : 0.0f;
#pragma unroll
    for (int e = 0; e < 4; e++) {
      float pv[2];
      pv[0] = __shfl_up(tot[0], 1u << e);
      pv[1] = __shfl_up(tot[1], 1u << e);
      if (lane >= (1 << e)) add_matvec2(tot, trans + (6 + e) * 4, pv);
    }
    const float p0 = __shfl_up(tot[0], 1u), p1 = __shfl_up(tot[1], 1u);
    if (lane < kWaves) {
      ex[2 * kWaves + lane * 2] = lane == 0 ? 0.0f : p0;
      ex[2 * kWaves + lane * 2 + 1] = lane == 0 ? 0.0f : p1;
    }
  }
  __syncthreads();
  // the wave's entry state, injected into its lane 0 (end state = zero-start end + A^C entry)
  const float e_in[2] = {ex[2 * kWaves + wave * 2], ex[2 * kWaves + wave * 2 + 1]};
  x[0] = st[0];
  x[1] = st[1];
  if (lane == 0) add_matvec2(x, trans, e_in);
  wave_scan2(x, trans, lane);
  const float p0 = __shfl_up(x[0], 1u), p1 = __shfl_up(x[1], 1u);
  st[0] = lane == 0 ? e_in[0] : p0;
  st[1] = lane == 0 ? e_in[1] : p1;
  __syncthreads();
}

// One direction over this thread's chunk val[0..kChunk) in processing order, one
// section after another (direct form II transposed, fused multiply-adds).  The chunk
// of thread 0 starts each section from zi * first_value, the others from zero and
// then receive the section's response to their entry state.
template <int S>
__device__ __forceinline__ void run_direction(const SosScanArgs& p, float (&val)[kChunk], float first_value,
                                              float* ex) {
  const bool lead = threadIdx.x == 0;
#pragma unroll 1
  for (int k = 0; k < S; k++) {
    const float b0 = p.coef[k][0], b1 = p.coef[k][1], b2 = p.coef[k][2], a1 = p.coef[k][3], a2 = p.coef[k][4];
    float z[2];
    z[0] = lead ? p.zi[2 * k] * first_value : 0.0f;
    z[1] = lead ? p.zi[2 * k + 1] * first_value : 0.0f;
#pragma unroll
    for (int t = 0; t < kChunk; t++) {
      const float v = val[t];
      const float f = fmaf(b0, v, z[0]);
      z[0] = fmaf(b1, v, fmaf(-a1, f, z[1]));
      z[1] = fmaf(b2, v, -(a2 * f));
      val[t] = f;
    }
    // the level matrices are read from global memory with uniform addresses (scalar
    // loads into SGPRs, which the FMAs take directly)
    scan_states2(z, ex, p.trans + k * kScanLevels * 4);
    if (!lead) {
      const float* __restrict__ resp = p.trans + S * kScanLevels * 4 + k * kChunk * 2;
#pragma unroll
      for (int t = 0; t < kChunk; t++) val[t] = fmaf(resp[2 * t], z[0], fmaf(resp[2 * t + 1], z[1], val[t]));
    }
  }
}

// Filters the n samples the caller stored at ext[pad .. pad + n) (LDS, FP32; ext
// starts kSosMargin floats into a buffer of at least kSosMargin + n + 2 pad + kChunk
// + 64 floats, all of which are overwritten).  Writes the kept window of the result,
// rolled by `roll` (numpy.roll semantics), optionally mean-subtracted, to
// out[0 .. out_count) = result[out_begin ..).  Returns true (on every thread) when
// any of the n output samples is exactly zero.  `red` is 64 floats of LDS.  The
// caller's stores to ext must be followed by a barrier before the call.
template <int S>
__device__ __forceinline__ bool sos_filtfilt_block_t(const SosScanArgs& p, float* ext, int n, float* red, float* out,
                                                     int out_begin, int out_count, int roll, bool subtract_mean) {
  const int j = threadIdx.x;
  const int pad = p.pad;
  const int L = n + 2 * pad;
  const int C = p.chunk;
  float* ex = ext + L + kChunk;  // scan exchange (64 floats) past the sequence and its overrun
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
  run_direction<S>(p, val, ext_first, ex);
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
  run_direction<S>(p, val, y_last, ex);
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
