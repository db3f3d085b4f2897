// SampleSinc of the TBC resampler on the device, shared by the luma and chroma kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace vhsd {

// ------------------------------------------------------------------ TBC resample
// TbcLineResampler linear positions + SampleSinc + VideoOutputConverter
// (Tbc/TbcLineResampler.cs:732-804, 1167-1234, 1495-1525;
// Tbc/VideoOutputConverter.cs:74-90).  One thread per output sample.
// Source samples are plane offsets; the reference reads (float)video with video in
// FP64 Hz, reproduced here as (float)((double)offset + bias).
// Table rows and interpolation weight of SampleSinc for a finite position.
struct SincPhase {
  int ci;       // integer part of the FP32 coordinate
  float alpha;  // weight of the second row
  int a0, a1;   // table rows (each 16 floats, 64-byte aligned)
};

__device__ __forceinline__ SincPhase sinc_phase(double position) {
  const int phases = 65536;
  SincPhase ph;
  const float coord = (float)position;
  ph.ci = (int)coord;
  const float frac = coord - (float)ph.ci;
  const float pp = frac * (float)phases;
  const int ps = (int)pp;
  const int pe = ps + 1;
  ph.alpha = pp - (float)ps;
  int a0 = ps < 0 ? ps + phases + 1 : ps;
  ph.a0 = a0 < 0 ? 0 : (a0 > phases ? phases : a0);
  int a1 = pe < 0 ? pe + phases + 1 : pe;
  ph.a1 = a1 < 0 ? 0 : (a1 > phases ? phases : a1);
  return ph;
}

// The 16-tap sum of SampleSinc given the two table rows.
__device__ __forceinline__ double sinc_sum(const float* src, long long len, double bias, const SincPhase& ph,
                                           const float (&ws)[16], const float (&we)[16]) {
  const int taps = 16;
  const long long s0 = (long long)ph.ci - (taps / 2 - 1);
  const bool interior = len >= taps && s0 >= 0 && s0 <= len - taps;
  double result = 0.0;
#pragma unroll
  for (int t = 0; t < taps; t++) {
    const float sw = ws[t];
    const float wgt = __builtin_fmaf(ph.alpha, we[t] - sw, sw);
    long long si = s0 + t;
    if (!interior) {
      if (si < 0) si += len;
      si = si < 0 ? 0 : (si > len - 1 ? len - 1 : si);
    }
    float prod;
    {
#pragma clang fp contract(off)
      prod = (float)((double)src[si] + bias) * wgt;
    }
    result = (double)prod + result;
  }
  return result;
}

// sinc_sum for taps read from a wave's source window staged in LDS (taps[t], t = 0..15,
// the lane's first tap onwards).  Same operations in the same order.
__device__ __forceinline__ double sinc_sum_window(const float* taps, double bias, const SincPhase& ph,
                                                  const float (&ws)[16], const float (&we)[16]) {
  double result = 0.0;
#pragma unroll
  for (int t = 0; t < 16; t++) {
    const float sw = ws[t];
    const float wgt = __builtin_fmaf(ph.alpha, we[t] - sw, sw);
    const float v = taps[t];
    float prod;
    {
#pragma clang fp contract(off)
      prod = (float)((double)v + bias) * wgt;
    }
    result = (double)prod + result;
  }
  return result;
}

__device__ __forceinline__ double sample_sinc(const float* src, long long len, double bias, double position,
                                              const float* __restrict__ lut) {
  if (!isfinite(position)) return 0.0;
  const SincPhase ph = sinc_phase(position);
  // table rows are 64-byte aligned: fetch each as four 16-byte loads
  const float4* ws4 = reinterpret_cast<const float4*>(lut + ph.a0 * 16);
  const float4* we4 = reinterpret_cast<const float4*>(lut + ph.a1 * 16);
  float ws[16], we[16];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const float4 a = ws4[q], b = we4[q];
    ws[4 * q] = a.x; ws[4 * q + 1] = a.y; ws[4 * q + 2] = a.z; ws[4 * q + 3] = a.w;
    we[4 * q] = b.x; we[4 * q + 1] = b.y; we[4 * q + 2] = b.z; we[4 * q + 3] = b.w;
  }
  return sinc_sum(src, len, bias, ph, ws, we);
}

// A wave's row stage: 64 pixels x two 64-byte rows (XOR-swizzled slots), then the wave's
// source window of kTbcWindow samples.
constexpr int kTbcWindow = 256;
constexpr int kTbcStage = 64 * 8 + kTbcWindow / 4;
__device__ __forceinline__ int tbc_slot(int pix, int s) { return pix * 8 + (s ^ (pix & 7)); }

// sample_sinc for the 64 lanes of a wave together (the same values): the two table rows
// of every lane are fetched cooperatively (four lanes read the four 16-byte quarters of
// one row, so a wave load touches 16 rows instead of 64) and handed over through the
// wave's `stage` (kTbcStage float4 of LDS); the wave's taps come from one coalesced load
// of kTbcWindow source samples from lane 0's first tap, and a lane whose taps leave that
// window reads the plane.  Every thread of the workgroup calls it (one barrier inside);
// lanes without an output pass any position.
__device__ __forceinline__ double sample_sinc_wave(const float* src, long long len, double bias, double position,
                                                   const float* __restrict__ lut, float4* stage) {
  const bool finite = isfinite(position);
  const SincPhase ph = finite ? sinc_phase(position) : SincPhase{0, 0.0f, 0, 0};
  const int lane = threadIdx.x & 63;
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
  const long long s0 = (long long)ph.ci - 7;
  const long long base = __shfl(s0, 0);
  float* win = reinterpret_cast<float*>(stage + 64 * 8);
#pragma unroll
  for (int i = 0; i < kTbcWindow / 64; i++) {
    const int w = lane + 64 * i;
    const long long g = base + w;
    win[w] = (g >= 0 && g < len) ? src[g] : 0.0f;
  }
  __syncthreads();
  float ws[16], we[16];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const float4 a = stage[tbc_slot(lane, q)], b = stage[tbc_slot(lane, 4 + q)];
    ws[4 * q] = a.x; ws[4 * q + 1] = a.y; ws[4 * q + 2] = a.z; ws[4 * q + 3] = a.w;
    we[4 * q] = b.x; we[4 * q + 1] = b.y; we[4 * q + 2] = b.z; we[4 * q + 3] = b.w;
  }
  const long long rel = s0 - base;
  const bool in_window = len >= 16 && s0 >= 0 && s0 <= len - 16 && rel >= 0 && rel <= kTbcWindow - 16;
  return !finite ? 0.0
         : in_window ? sinc_sum_window(win + (int)rel, bias, ph, ws, we)
                     : sinc_sum(src, len, bias, ph, ws, we);
}

}  // namespace vhsd
