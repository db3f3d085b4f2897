// Per-block RF demodulation kernels for gfx950 (VHS real-FFT path).
//
// One 1024-thread workgroup demodulates one 32768-sample RF block entirely in
// registers + LDS: the int16 block is read once from HBM and only the trimmed
// output window (samples [cut, cut + stride) of each plane) is written, straight
// into the stitched field-span layout, so the overlap-save stitch of
// RfBlockStreamDecoder.CopyTrimmedWindow (Decode/RfBlockStreamDecoder.cs:497-549,
// 1579-1612) costs nothing extra.
//
// Chain (Dsp/RfDemodulator.cs:288-918, VHS real path, exact backend semantics):
//   rfft(x) -> x RfVideo -> analytic (one-sided spectrum, complex inverse as two
//   16384-point passes for even and odd output samples) -> FP32 polynomial atan2
//   FM discriminator (PortedMath.cs:519-551) -> envelope: |float(Re)| rolled +4,
//   FP32 sosfiltfilt (RfDemodulator.cs:1755-1839) -> rfft(demod) -> x Video and
//   x VideoLowPass05 -> two inverse real FFTs (video, sync reference rolled -32).
// The chroma plane (FP32 sosfiltfilt of the input with the colour-under band-pass,
// RfBlockDecodePipeline.cs:539-573) is a second kernel; the diff-demod spike
// repair (RfDemodulator.cs:2527-2581, :966-1001) is a third, run only by blocks
// whose demod exceeded the repair threshold.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "fft16k.hpp"
#include "sos_scan.hpp"
#include "vhsd_args.hpp"
#include "vhsd_device.hpp"

namespace vhsd {

// LDS layout shared by the block kernels (bytes): 128 KiB exchange buffer,
// 16 KiB twiddle octant, small scratch.
struct BlockLds {
  c32 buf[kFftBufEntries];
  float2 oct[kOctantEntries];
  float scratch[kScratchFloats];
  float tree[256];  // pairwise-sum levels
};

// Strided element k = j + 1024 m of a global array through a uniform row pointer, so
// the address is an SGPR base plus the lane offset instead of a 64-bit VGPR per m.
template <typename T>
__device__ __forceinline__ T& row_at(T* base, int m) {
  return (base + (long long)m * kFftThreads)[threadIdx.x];
}

__device__ __forceinline__ void load_octant(BlockLds& s, const float2* __restrict__ oct_g) {
  for (int u = threadIdx.x; u < kOctantEntries; u += blockDim.x) s.oct[u] = oct_g[u];
}

// rfft post-processing: Z = FFT16384(x[2m] + i x[2m+1]) held in strided
// distribution -> X[k] for k = j + 1024 m (and X[16384] returned via xm on thread 0).
// Strided slot m of thread j in the padded exchange buffer, and the slot holding
// entry (M - k) mod M for k = j + 1024 m: thread 1024 - j, slot 15 - m (thread 0:
// slot (16 - m) mod 16 of thread 0).
constexpr int kPadRow = kFftThreads + kFftThreads / 32;
__device__ __forceinline__ int mirror_slot(int j, int m) {
  return j > 0 ? lds_pad(kFftThreads - j) + (15 - m) * kPadRow : ((16 - m) & 15) * kPadRow;
}

__device__ __forceinline__ void rfft_post(c32* v, BlockLds& s, c32& xm) {
  const int j = threadIdx.x;
  const int pj = lds_pad(j);
#pragma unroll
  for (int m = 0; m < 16; m++) s.buf[pj + m * kPadRow] = v[m];
  __syncthreads();
  c32 z0 = s.buf[0];
  const c32 wj32 = tw32k(s.oct, j);  // roots of k = j + 1024 m: rot32(wj32, m)
#pragma unroll
  for (int m = 0; m < 16; m++) {
    const c32 zk = v[m];
    const c32 zc = conjc(s.buf[mirror_slot(j, m)]);
    const c32 e = add(zk, zc);
    const c32 o = sub(zk, zc);
    // X[k] = 0.5*(e - i*W^k*o)
    const c32 wo = mul(rot32(wj32, m), o);
    v[m] = scale(add(e, mul_mi(wo)), 0.5f);
  }
  xm = c32{z0.x - z0.y, 0.0f};
  __syncthreads();
}

// Inverse real FFT pre-processing of a half spectrum V[k] (strided, plus V[M] in vm):
// Y[k] = (V[k] + conj(V[M-k])) + i * (V[k] - conj(V[M-k])) * exp(+2 pi i k / N).
// numpy.fft.irfft semantics: imaginary parts of V[0] and V[M] are ignored.
__device__ __forceinline__ void irfft_pre(c32* v, c32 vm, BlockLds& s) {
  const int j = opaque_tid();
  const int pj = lds_pad(j);
#pragma unroll
  for (int m = 0; m < 16; m++) s.buf[pj + m * kPadRow] = v[m];
  __syncthreads();
  const c32 wj32 = tw32k(s.oct, j);  // roots of k = j + 1024 m: rot32(wj32, m)
#pragma unroll
  for (int m = 0; m < 16; m++) {
    const int k = j + m * kFftThreads;
    c32 y;
    if (k == 0) {
      const float a = v[m].x, b = vm.x;
      y = c32{a + b, a - b};
    } else {
      const c32 vk = v[m];
      const c32 vc = conjc(s.buf[mirror_slot(j, m)]);
      const c32 e = add(vk, vc);
      const c32 o = mul(sub(vk, vc), conjc(rot32(wj32, m)));
      y = add(e, mul_pi(o));
    }
    v[m] = y;
  }
  __syncthreads();
}

// Stage B+: analytic signal from the filtered half spectrum F (strided, F[M] in fm)
// and the FM discriminator.  Writes the demod to stash[k] = float2(sample 2k,
// sample 2k+1) for k = j + 1024 m.  With kEnvRaw the envelope source
// raw[(n + 4) mod N] = |float(Re analytic[n])| goes straight to LDS at `raw` (inside
// the transform buffer, free after the last transform).  State that must outlive a
// transform goes through the block's stash (written and read back by this workgroup
// only) instead of registers.  Returns whether this thread saw a demod sample in
// [20, N-20) above diff_max.
template <bool kEnvRaw>
__device__ __forceinline__ bool analytic_demod(c32* F, c32 fm, BlockLds& s, float fs, double diff_max,
                                               float2* __restrict__ stash, float* raw) {
  const int j = threadIdx.x;
  const int lane = j & 63, wave = j >> 6;
  const float inv_n = 1.0f / 32768.0f;
  float2* st_go = stash;            // odd-pass input spectrum
  float2* st_th = stash + kFftM;    // (angle, real part) of the even samples
  // odd output samples' input, parked until the even pass is done
  const c32 wj32 = tw32k(s.oct, j);  // roots of k = j + 1024 m: rot32(wj32, m)
#pragma unroll
  for (int m = 0; m < 16; m++) {
    const int k = j + m * kFftThreads;
    const c32 g = (k == 0) ? sub(F[m], fm) : mul(scale(F[m], 2.0f), conjc(rot32(wj32, m)));
    row_at(st_go, m) = make_float2(g.x, g.y);
  }
  // even output samples
#pragma unroll
  for (int m = 0; m < 16; m++) {
    const int k = j + m * kFftThreads;
    F[m] = (k == 0) ? add(F[m], fm) : scale(F[m], 2.0f);
  }
  fft16k_inverse(F, s.buf, s.oct);
#pragma unroll
  for (int m = 0; m < 16; m++) {
    const c32 a = scale(F[m], inv_n);
    row_at(st_th, m) = make_float2(vhs_atan2(a.y, a.x), a.x);
  }
#pragma unroll
  for (int m = 0; m < 16; m++) {
    const float2 g = row_at(st_go, m);
    F[m] = c32{g.x, g.y};
  }
  fft16k_inverse(F, s.buf, s.oct);
  // odd sample 2k+1 against even sample 2k in the same thread; even sample 2k against
  // odd sample 2k-1, which is the previous lane's (shuffle) or, across waves, the
  // last lane of the previous wave (through `edge`, 16 angles per wave)
  float* edge = s.scratch;
  float theta_e[16];
  float theta_o[16];
  float dem_o[16];
#pragma unroll
  for (int m = 0; m < 16; m++) {
    const float2 te = row_at(st_th, m);
    theta_e[m] = te.x;
    const c32 a = scale(F[m], inv_n);
    const float t = vhs_atan2(a.y, a.x);
    theta_o[m] = t;
    dem_o[m] = vhs_freq_diff(t, te.x, fs);
    if (lane == 63) edge[wave * 16 + m] = t;
    if constexpr (kEnvRaw) {
      const int n0 = 2 * (j + m * kFftThreads);
      raw[(n0 + 4) & 32767] = fabsf(te.y);
      raw[(n0 + 5) & 32767] = fabsf(a.x);
    }
  }
  __syncthreads();
  bool spike = false;
#pragma unroll
  for (int m = 0; m < 16; m++) {
    // previous odd sample of even sample 2*(j + 1024 m) is odd index (j-1) + 1024 m
    const float up = __shfl_up(theta_o[m], 1u);
    float prev;
    if (lane > 0)
      prev = up;
    else if (wave > 0)
      prev = edge[(wave - 1) * 16 + m];
    else
      prev = (m > 0) ? edge[15 * 16 + (m - 1)] : 0.0f;
    const float de = (j == 0 && m == 0) ? 0.0f : vhs_freq_diff(theta_e[m], prev, fs);
    row_at(st_go, m) = make_float2(de, dem_o[m]);
    if (diff_max > 0.0) {
      const int n0 = 2 * (j + m * kFftThreads);
      spike |= (n0 >= 20 && n0 < 32768 - 20 && (double)de > diff_max);
      spike |= (n0 + 1 >= 20 && n0 + 1 < 32768 - 20 && (double)dem_o[m] > diff_max);
    }
  }
  __syncthreads();
  return spike;
}

// Video and sync-reference planes from the demodulated signal held in the stash as
// stash[k] = float2(sample 2k, sample 2k+1), k = j + 1024 m.
__device__ __forceinline__ void video_stage(BlockLds& s, const DemodArgs& a, long long out_base,
                                            float2* __restrict__ stash) {
  const int j = threadIdx.x;
  // The ~4 MHz carrier level would dominate the FP32 rounding of the transforms; remove
  // the plane bias first (exact: both are near the same binade) and restore the DC the
  // filters pass as a constant correction, leaving offsets from the bias in the planes.
  c32 D[16];
  const bool clip = a.clip_hi > a.clip_lo;
#pragma unroll
  for (int m = 0; m < 16; m++) {
    const float2 d = row_at(stash, m);
    float x0 = d.x, x1 = d.y;
    if (clip) {
      x0 = fminf(fmaxf(x0, a.clip_lo), a.clip_hi);
      x1 = fminf(fmaxf(x1, a.clip_lo), a.clip_hi);
    }
    D[m] = c32{x0 - a.stash_sub, x1 - a.stash_sub};
  }
  fft16k_forward(D, s.buf, s.oct);
  c32 dm;
  rfft_post(D, s, dm);
  const float inv_n = 1.0f / 32768.0f;
  // the spectrum feeds two inverse transforms (and the sub-deemphasis kernel); park it
  // rather than hold it in registers
#pragma unroll
  for (int m = 0; m < 16; m++) row_at(stash, m) = make_float2(D[m].x, D[m].y);
  if (j == 0) stash[2 * kFftM] = make_float2(dm.x, dm.y);
  // passes: video, sync reference, then up to two reference planes (LaserDisc burst, pilot)
#pragma unroll 1
  for (int pass = 0; pass < 2 + a.nref; pass++) {
    const bool ref1 = pass == 3;
    const float2* filt = pass == 0 ? a.video : pass == 1 ? a.vlp05 : (ref1 ? a.ref_filt[1] : a.ref_filt[0]);
    float* out = pass == 0 ? a.out_video : pass == 1 ? a.out_vlp : (ref1 ? a.ref_out[1] : a.ref_out[0]);
    const int shift = pass == 0 ? 0 : pass == 1 ? a.vlp_offset : (ref1 ? a.ref_shift[1] : a.ref_shift[0]);
    const float corr = pass == 0 ? a.video_corr : pass == 1 ? a.vlp_corr : (ref1 ? a.ref_corr[1] : a.ref_corr[0]);
    c32 v[16];
    if (pass == 0) {
      // the first pass takes the spectrum from registers; later passes read it back
#pragma unroll
      for (int m = 0; m < 16; m++) {
        const float2 f = row_at(filt, m);
        v[m] = mul(D[m], c32{f.x, f.y});
      }
    } else {
#pragma unroll
      for (int m = 0; m < 16; m++) {
        const float2 f = row_at(filt, m);
        const float2 d = row_at(stash, m);
        v[m] = mul(c32{d.x, d.y}, c32{f.x, f.y});
      }
    }
    const float2 fM = filt[kFftM];
    const c32 vm = mul(dm, c32{fM.x, fM.y});
    irfft_pre(v, vm, s);
    fft16k_inverse(v, s.buf, s.oct);
    // v[m] = N * (y[2k] + i y[2k+1]); output sample n = 2k lands at n - shift (roll by -shift)
    // position o = n - out_begin (even: out_begin, shift and out_base are even)
    float2* orow = reinterpret_cast<float2*>(out + out_base - shift - a.out_begin);
#pragma unroll
    for (int m = 0; m < 16; m++) {
      const int k = j + m * kFftThreads;
      const int o = 2 * k - shift - a.out_begin;
      if (o >= 0 && o < a.out_count) {
        const float2 val = make_float2(v[m].x * inv_n + corr, v[m].y * inv_n + corr);
        if (o + 1 < a.out_count)
          row_at(orow, m) = val;
        else
          out[out_base + o] = val.x;
      }
    }
  }
}

__device__ __forceinline__ void vhs_block_demod_body(const DemodArgs& a, const int blk) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  BlockLds& s = *reinterpret_cast<BlockLds*>(lds_raw);
  const int j = threadIdx.x;
  const long long start = a.block_start[blk];
  const long long out_base = a.out_offset[blk];
  load_octant(s, a.oct);
  // ---- stage A: rfft of the int16 block
  c32 v[16];
  const int16_t* x = a.rf + start;
#pragma unroll
  for (int m = 0; m < 16; m++) {
    const int16_t* xr = x + 2 * m * kFftThreads;
    v[m] = c32{(float)xr[2 * j], (float)xr[2 * j + 1]};
  }
  __syncthreads();
  fft16k_forward(v, s.buf, s.oct);
  c32 xm;
  rfft_post(v, s, xm);
  // ---- filter by RfVideo (RfMtf is all ones for tape)
#pragma unroll
  for (int m = 0; m < 16; m++) {
    const float2 f = row_at(a.rf_video, m);
    v[m] = mul(v[m], c32{f.x, f.y});
  }
  const float2 fM = a.rf_video[kFftM];
  const c32 fm = mul(xm, c32{fM.x, fM.y});
  c32 fm_b = c32{0.0f, 0.0f};
  // X[M] lives in thread 0 only; broadcast through scratch
  if (j == 0) {
    s.scratch[0] = fm.x;
    s.scratch[1] = fm.y;
  }
  __syncthreads();
  fm_b = c32{s.scratch[0], s.scratch[1]};
  __syncthreads();
  // ---- stage B: analytic signal + discriminator
  float2* stash = a.stash + (long long)blk * kStashEntries;
  // demod to the stash; spike check over demod[20 .. N-20); the envelope source
  // raw[(n + 4) mod N] = |float(re[n])| to LDS for the FP32 sosfiltfilt
  float* raw = reinterpret_cast<float*>(s.buf) + kSosMargin + a.env.pad;
  const bool spike = analytic_demod<true>(v, fm_b, s, a.fs, a.diff_max, stash, raw);
  // the VHS envelope filter is first order (DecodeFilterSet.BuildVhsEnvelopeSos): one section
  const bool weak = sos_filtfilt_block_t<1>(a.env, reinterpret_cast<float*>(s.buf) + kSosMargin, 32768, s.scratch,
                                            a.out_env + out_base, a.out_begin,
                                            a.out_count, /*roll=*/0, false);
  // ---- flags
  const unsigned long long any_spike = __ballot(spike);
  if ((j & 63) == 0 && any_spike) atomicOr(&a.flags[blk], 2u);
  if (j == 0 && weak) atomicOr(&a.flags[blk], 1u);
}

__global__ void __launch_bounds__(1024) vhs_block_demod_kernel(DemodArgs a) { vhs_block_demod_body(a, blockIdx.x); }


// numpy's pairwise float32 sum of the 32768 values at y (NumpyReduction.PairwiseSumFloat32,
// Dsp/NumpyReduction.cs:253-339) in its exact order: 256 runs of 128 summed with eight
// accumulators each (accumulator k adds elements k, k+8, ... in order; then
// ((r0+r1)+(r2+r3))+((r4+r5)+(r6+r7))), and the run sums added pairwise up a binary
// tree.  Every thread returns the total.
__device__ float pairwise_sum_32768(const float* y, float* tree) {
  const int t = threadIdx.x;
  float run[2];
#pragma unroll
  for (int h = 0; h < 2; h++) {
    const int c = t + h * kFftThreads;  // chain: run c >> 3, accumulator c & 7
    const float* base = y + ((c >> 3) << 7) + (c & 7);
    float r = base[0];
#pragma unroll
    for (int i = 1; i < 16; i++) r += base[8 * i];
    // accumulators of a run sit in 8 consecutive lanes
    r = r + __shfl_xor(r, 1);
    r = r + __shfl_xor(r, 2);
    r = r + __shfl_xor(r, 4);
    run[h] = r;
  }
  if ((t & 7) == 0) {
    tree[t >> 3] = run[0];
    tree[128 + (t >> 3)] = run[1];
  }
  __syncthreads();
  for (int width = 128; width >= 1; width >>= 1) {
    float v = 0.0f;
    if (t < width) v = tree[2 * t] + tree[2 * t + 1];
    __syncthreads();
    if (t < width) tree[t] = v;
    __syncthreads();
  }
  const float total = tree[0];
  __syncthreads();
  return total;
}

// Block demod with the RF high boost (Betamax; RfDemodulator.cs:356-392, 635-663,
// 2374-2449): rfft -> x RfVideo (kept in aux) -> filtered RF by irfft -> envelope ->
// unless weak, high = FP32 forward-backward RF top band-pass of the filtered RF, scaled
// by 0.9 mean(envelope) / envelope and the multiplier, whose spectrum is added to the
// filtered spectrum -> analytic signal and discriminator as in the plain kernel.
__device__ __forceinline__ void vhs_block_demod_boost_body(const DemodArgs& a, const int blk) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  BlockLds& s = *reinterpret_cast<BlockLds*>(lds_raw);
  const int j = threadIdx.x;
  const long long start = a.block_start[blk];
  const long long out_base = a.out_offset[blk];
  float2* stash = a.stash + (long long)blk * kStashEntries;
  float2* aux = a.aux + (long long)blk * kAuxEntries;
  load_octant(s, a.oct);
  c32 v[16];
  const int16_t* x = a.rf + start;
#pragma unroll
  for (int m = 0; m < 16; m++) {
    const int16_t* xr = x + 2 * m * kFftThreads;
    v[m] = c32{(float)xr[2 * j], (float)xr[2 * j + 1]};
  }
  __syncthreads();
  fft16k_forward(v, s.buf, s.oct);
  c32 xm;
  rfft_post(v, s, xm);
#pragma unroll
  for (int m = 0; m < 16; m++) {
    const float2 f = row_at(a.rf_video, m);
    v[m] = mul(v[m], c32{f.x, f.y});
    row_at(aux, m) = make_float2(v[m].x, v[m].y);
  }
  const float2 fM = a.rf_video[kFftM];
  const c32 fm = mul(xm, c32{fM.x, fM.y});  // every thread computed xm from the same LDS words
  if (j == 0) aux[kFftM] = make_float2(fm.x, fm.y);
  // filtered real RF (the envelope source and the boost input), as sample pairs
  const float inv_n = 1.0f / 32768.0f;
  irfft_pre(v, fm, s);
  fft16k_inverse(v, s.buf, s.oct);
  float2* st_x = stash + kFftM;
#pragma unroll
  for (int m = 0; m < 16; m++) row_at(st_x, m) = make_float2(v[m].x * inv_n, v[m].y * inv_n);
  // envelope: raw[(n + 4) mod N] = |float(x[n])|
  float* ext = reinterpret_cast<float*>(s.buf) + kSosMargin;
  {
    float* raw = ext + a.env.pad;
#pragma unroll
    for (int m = 0; m < 16; m++) {
      const int n0 = 2 * (j + m * kFftThreads);
      raw[(n0 + 4) & 32767] = fabsf(v[m].x * inv_n);
      raw[(n0 + 5) & 32767] = fabsf(v[m].y * inv_n);
    }
  }
  __syncthreads();
  const bool weak = sos_filtfilt_block_t<1>(a.env, ext, 32768, s.scratch,
                                            a.out_env + out_base, a.out_begin, a.out_count, /*roll=*/0, false);
  c32 fb;
  if (!weak) {  // uniform: every thread holds the same flag
    const float* env = ext + a.env.pad;
    const float mean = pairwise_sum_32768(env, s.tree) / 32768.0f;
    const float numer = mean * 0.9f;
    // per-sample scale, parked in the stash while the RF top filter uses the LDS
#pragma unroll
    for (int m = 0; m < 16; m++) {
      const int k = j + m * kFftThreads;
      row_at(stash, m) = make_float2(numer / env[2 * k], numer / env[2 * k + 1]);
    }
    __syncthreads();
    float* hx = ext + a.boost.pad;
#pragma unroll
    for (int m = 0; m < 16; m++) {
      const int k = j + m * kFftThreads;
      const float2 xv = row_at(st_x, m);
      hx[2 * k] = xv.x;
      hx[2 * k + 1] = xv.y;
    }
    __syncthreads();
    sos_filtfilt_block_t<1>(a.boost, ext, 32768, s.scratch, nullptr, 0, 0, 0, false);
#pragma unroll
    for (int m = 0; m < 16; m++) {
      const int k = j + m * kFftThreads;
      const float2 sc = row_at(stash, m);
      v[m] = c32{(hx[2 * k] * sc.x) * a.boost_mult, (hx[2 * k + 1] * sc.y) * a.boost_mult};
    }
    __syncthreads();
    fft16k_forward(v, s.buf, s.oct);
    c32 hm;
    rfft_post(v, s, hm);
#pragma unroll
    for (int m = 0; m < 16; m++) {
      const float2 f = row_at(aux, m);
      v[m] = c32{f.x + v[m].x, f.y + v[m].y};
    }
    const float2 f0 = aux[kFftM];
    fb = c32{f0.x + hm.x, f0.y + hm.y};
  } else {
#pragma unroll
    for (int m = 0; m < 16; m++) {
      const float2 f = row_at(aux, m);
      v[m] = c32{f.x, f.y};
    }
    const float2 f0 = aux[kFftM];
    fb = c32{f0.x, f0.y};
    if (j == 0) atomicOr(&a.flags[blk], 1u);
  }
  const bool spike = analytic_demod<false>(v, fb, s, a.fs, a.diff_max, stash, nullptr);
  const unsigned long long any_spike = __ballot(spike);
  if ((j & 63) == 0 && any_spike) atomicOr(&a.flags[blk], 2u);
}

__global__ void __launch_bounds__(1024) vhs_block_demod_boost_kernel(DemodArgs a) { vhs_block_demod_boost_body(a, blockIdx.x); }


// ------------------------------------------------------------------ LaserDisc
// PortedMath.UnwrapHilbertConjugateProduct (Dsp/PortedMath.cs:39-68) for one pair of
// FP32 analytic samples: the products are exact in FP64, then atan2 wrapped to
// [0, 2 pi), times fs / 2 pi.
__device__ __forceinline__ double conj_disc(float2 c, float2 p, double scale) {
  const double re = (double)c.x * p.x + (double)c.y * p.y;
  const double im = (double)c.y * p.x - (double)c.x * p.y;
  double d = atan2(im, re);
  if (d < 0.0) d += 6.283185307179586;
  return d * scale;
}

__device__ __forceinline__ float mag32(float2 v) { return sqrtf(v.x * v.x + v.y * v.y); }

// Analytic signal and conjugate-product discriminator of the LaserDisc complex path
// (RfDemodulator.cs:564-605, 722-743).  As analytic_demod, with the analytic samples
// themselves kept: even samples in the stash, odd samples in LDS.  The demod goes to
// stash[k] = (sample 2k, sample 2k+1) as offsets from the plane bias; the envelope
// |analytic| (BuildAnalyticMagnitudeEnvelope, RfDemodulator.cs:1709-1727) is written
// straight to the kept window of the envelope plane.
__device__ __forceinline__ void analytic_demod_ld(c32* F, c32 fm, BlockLds& s, const DemodArgs& a,
                                                  long long out_base, float2* __restrict__ stash) {
  const int j = threadIdx.x;
  const float inv_n = 1.0f / 32768.0f;
  float2* st_go = stash;
  float2* st_ev = stash + kFftM;
  const c32 wj32 = tw32k(s.oct, j);  // roots of k = j + 1024 m: rot32(wj32, m)
#pragma unroll
  for (int m = 0; m < 16; m++) {
    const int k = j + m * kFftThreads;
    const c32 g = (k == 0) ? sub(F[m], fm) : mul(scale(F[m], 2.0f), conjc(rot32(wj32, m)));
    row_at(st_go, m) = make_float2(g.x, g.y);
  }
#pragma unroll
  for (int m = 0; m < 16; m++) {
    const int k = j + m * kFftThreads;
    F[m] = (k == 0) ? add(F[m], fm) : scale(F[m], 2.0f);
  }
  fft16k_inverse(F, s.buf, s.oct);
#pragma unroll
  for (int m = 0; m < 16; m++) row_at(st_ev, m) = make_float2(F[m].x * inv_n, F[m].y * inv_n);
#pragma unroll
  for (int m = 0; m < 16; m++) {
    const float2 g = row_at(st_go, m);
    F[m] = c32{g.x, g.y};
  }
  fft16k_inverse(F, s.buf, s.oct);
  float2* odd = reinterpret_cast<float2*>(s.buf);
#pragma unroll
  for (int m = 0; m < 16; m++) odd[j + m * kFftThreads] = make_float2(F[m].x * inv_n, F[m].y * inv_n);
  __syncthreads();
  const double scale_hz = a.fs_d / 6.283185307179586;
  const double bias = (double)a.bias;
  float* env = a.out_env + out_base - a.out_begin;
#pragma unroll 1
  for (int m = 0; m < 16; m++) {
    const int k = j + m * kFftThreads;
    const float2 ae = row_at(st_ev, m);
    const float2 ao = odd[k];
    const double de = k == 0 ? 0.0 : conj_disc(ae, odd[k - 1], scale_hz);
    const double dd = conj_disc(ao, ae, scale_hz);
    row_at(st_go, m) = make_float2((float)(de - bias), (float)(dd - bias));
    const int o = 2 * k - a.out_begin;
    if (o >= 0 && o < a.out_count) env[2 * k] = mag32(ae);
    if (o + 1 >= 0 && o + 1 < a.out_count) env[2 * k + 1] = mag32(ao);
  }
  __syncthreads();
}

// One 2^log2n-point inverse complex FFT per channel, both channels at once, radix-2
// Stockham between two LDS buffers; 1024 threads, one butterfly per thread per channel
// slot.  Returns the buffer holding the (unnormalised) result.
__device__ __forceinline__ float2* audio_ifft(float2* x, float2* y, int n, int log2n, const float2* __restrict__ oct) {
  const int half = n >> 1;
  for (int p = 0; p < log2n; p++) {
    const int ns = 1 << p;
    for (int t = threadIdx.x; t < 2 * half; t += blockDim.x) {
      const int ch = t / half;
      const int b = t - ch * half;
      const float2* xi = x + ch * n;
      float2* yo = y + ch * n;
      const int k = b & (ns - 1);
      const c32 u = c32{xi[b].x, xi[b].y};
      c32 v = c32{xi[b + half].x, xi[b + half].y};
      // exp(+2 pi i k / (2 ns)) = conj(tw16k(k * 16384 / (2 ns)))
      v = mul(v, conjc(tw16k(oct, (k * (kFftM >> p)) >> 1)));
      const int idx = ((b - k) << 1) + k;
      const c32 s0 = add(u, v), s1 = sub(u, v);
      yo[idx] = make_float2(s0.x, s0.y);
      yo[idx + ns] = make_float2(s1.x, s1.y);
    }
    __syncthreads();
    float2* t = x;
    x = y;
    y = t;
  }
  return x;
}

// LaserDisc block demod (RfDemodulator.DemodulateCore complex path, RfDemodulator.cs:564-605,
// 621-633, 722-743): rfft -> x (RfVideo . RfMtf) -> analytic -> conjugate-product
// discriminator, |analytic| envelope.  With analog audio the two carrier slices of the
// input spectrum (SliceSpectrum, DecodeFilterSet.cs:991-1017) are multiplied by their
// stage-1 responses, inverse transformed and demodulated at the slice rate
// (RfBlockDecodePipeline.DecodeAnalogAudioChannel, RfBlockDecodePipeline.cs:732-752);
// each block writes its kept audio window (RfBlockStreamDecoder.cs:1690-1709).
__device__ __forceinline__ void vhs_block_demod_ld_body(const DemodArgs& a, const int blk) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  BlockLds& s = *reinterpret_cast<BlockLds*>(lds_raw);
  const int j = threadIdx.x;
  const long long start = a.block_start[blk];
  const long long out_base = a.out_offset[blk];
  load_octant(s, a.oct);
  c32 v[16];
  const int16_t* x = a.rf + start;
#pragma unroll
  for (int m = 0; m < 16; m++) {
    const int16_t* xr = x + 2 * m * kFftThreads;
    v[m] = c32{(float)xr[2 * j], (float)xr[2 * j + 1]};
  }
  __syncthreads();
  fft16k_forward(v, s.buf, s.oct);
  c32 xm;
  rfft_post(v, s, xm);
  float2* slices = a.aux + (long long)blk * kAuxEntries;
  if (a.audio) {
    // slice[i] = X[low + i] for i < half; slice[half + i] = conj(X[low + half - i])
    const int n = a.audio_bins, half = n >> 1;
#pragma unroll
    for (int m = 0; m < 16; m++) {
      const int k = j + m * kFftThreads;
#pragma unroll
      for (int ch = 0; ch < 2; ch++) {
        const int low = a.audio_low[ch];
        const float2* st1 = ch ? a.audio_stage1[1] : a.audio_stage1[0];
        float2* sl = slices + ch * n;
        const int i = k - low;
        if (i >= 0 && i < half) {
          const c32 y = mul(v[m], c32{st1[i].x, st1[i].y});
          sl[i] = make_float2(y.x, y.y);
        }
        const int r = low + half - k;
        if (r >= 0 && r < half) {
          const c32 y = mul(conjc(v[m]), c32{st1[half + r].x, st1[half + r].y});
          sl[half + r] = make_float2(y.x, y.y);
        }
      }
    }
  }
#pragma unroll
  for (int m = 0; m < 16; m++) {
    const float2 f = row_at(a.rf_video, m);
    v[m] = mul(v[m], c32{f.x, f.y});
  }
  const float2 fM = a.rf_video[kFftM];
  const c32 fm = mul(xm, c32{fM.x, fM.y});
  if (j == 0) {
    s.scratch[0] = fm.x;
    s.scratch[1] = fm.y;
  }
  __syncthreads();
  const c32 fm_b = c32{s.scratch[0], s.scratch[1]};
  __syncthreads();
  float2* stash = a.stash + (long long)blk * kStashEntries;
  analytic_demod_ld(v, fm_b, s, a, out_base, stash);
  if (!a.audio) return;
  // ---- analog audio: both slices through LDS, inverse FFT, discriminator, kept window
  const int n = a.audio_bins;
  float2* b0 = reinterpret_cast<float2*>(s.buf);
  float2* b1 = b0 + 2 * n;
  for (int t = j; t < 2 * n; t += kFftThreads) b0[t] = slices[t];
  __syncthreads();
  const float2* res = audio_ifft(b0, b1, n, a.audio_log2, s.oct);
  const long long ob = out_base / a.audio_decimation;
  for (int t = j; t < 2 * n; t += kFftThreads) {
    const int ch = t >= n;
    const int i = t - ch * n;
    const float2* r = res + ch * n;
    const double d = i == 0 ? 0.0 : conj_disc(r[i], r[i - 1], a.audio_scale);
    const float val = (float)(d + (ch ? a.audio_low_freq[1] : a.audio_low_freq[0]));
    const int o = i - a.audio_cut;
    if (o >= 0 && o < a.audio_stride) (ch ? a.out_audio[1] : a.out_audio[0])[ob + o] = val;
  }
}

__global__ void __launch_bounds__(1024) vhs_block_demod_ld_kernel(DemodArgs a) { vhs_block_demod_ld_body(a, blockIdx.x); }


// Stage C: video and sync-reference planes from the demod the block demod (or spike
// repair) kernel left in each block's stash.
__device__ __forceinline__ void vhs_block_video_body(const DemodArgs& a, const int blk) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  BlockLds& s = *reinterpret_cast<BlockLds*>(lds_raw);
  load_octant(s, a.oct);
  __syncthreads();
  video_stage(s, a, a.out_offset[blk], a.stash + (long long)blk * kStashEntries);
}

__global__ void __launch_bounds__(1024) vhs_block_video_kernel(DemodArgs a) { vhs_block_video_body(a, blockIdx.x); }


// ------------------------------------------------------------------ chroma
template <int S>
__device__ __forceinline__ void vhs_block_chroma_body(const ChromaArgs& a, const int blk) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  BlockLds& s = *reinterpret_cast<BlockLds*>(lds_raw);
  const int16_t* x = a.rf + a.block_start[blk];
  float* ext = reinterpret_cast<float*>(s.buf) + kSosMargin;
  float* raw = ext + a.sos.pad;
  int16_t xv[32];
#pragma unroll
  for (int m = 0; m < 32; m++) xv[m] = row_at(x, m);
#pragma unroll
  for (int m = 0; m < 32; m++) row_at(raw, m) = (float)xv[m];
  __syncthreads();
  sos_filtfilt_block_t<S>(a.sos, ext, 32768, s.scratch, a.out_chroma + a.out_offset[blk], a.out_begin,
                          a.out_count, a.chroma_offset, /*subtract_mean=*/true);
}

template <int S>
__global__ void __launch_bounds__(1024) vhs_block_chroma_kernel(ChromaArgs a) { vhs_block_chroma_body<S>(a, blockIdx.x); }


// ------------------------------------------------------------------ sub-deemphasis
// RfDemodulator.ApplySubDeemphasisRealCore (Dsp/RfDemodulator.cs:2897-2999) on one block:
// H = D x Video x HighPass; the analytic signal of irfft(H) (one-sided H, inverse as
// even/odd 16384-point passes) gives both the high-passed video (its real part) and
// its magnitude; the magnitude / (deviation / 2) is low-pass filtered forward-backward,
// mapped through max(0, .) * scaling1, pow, * scaling2 and the optional logistic, and
// video -= high * (1 - gain) + high * static_factor over the kept window.
// |z| the way RfDemodulator.NumpyComplexMagnitude forms it (RfDemodulator.cs:3216-3242):
// larger * sqrt(1 + (smaller / larger)^2).
__device__ __forceinline__ float numpy_abs(c32 z) {
  const float re = fabsf(z.x), im = fabsf(z.y);
  const float lg = fmaxf(re, im), sm = fminf(re, im);
  if (lg == 0.0f) return 0.0f;
  const float r = sm / lg;
  return sqrtf(fmaf(r, r, 1.0f)) * lg;
}

__device__ __forceinline__ void vhs_block_subdeemph_body(const SubDeemphArgs& a, const int blk) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  BlockLds& s = *reinterpret_cast<BlockLds*>(lds_raw);
  const int j = threadIdx.x;
  float2* stash = a.stash + (long long)blk * kStashEntries;
  float2* aux = a.aux + (long long)blk * kAuxEntries;
  const float inv_n = 1.0f / 32768.0f;
  load_octant(s, a.oct);
  if (a.notch) {
    // the whole block's video, as the video kernel formed it: irfft(D x Video) / N + corr
    c32 w[16];
#pragma unroll
    for (int m = 0; m < 16; m++) {
      const float2 d = row_at(stash, m);
      const float2 f = row_at(a.video, m);
      w[m] = mul(c32{d.x, d.y}, c32{f.x, f.y});
    }
    const float2 dM0 = stash[2 * kFftM];
    const float2 fV = a.video[kFftM];
    const c32 vm = mul(c32{dM0.x, dM0.y}, c32{fV.x, fV.y});
    __syncthreads();  // twiddles in LDS
    irfft_pre(w, vm, s);
    fft16k_inverse(w, s.buf, s.oct);
#pragma unroll
    for (int m = 0; m < 16; m++)
      row_at(aux, m) = make_float2(w[m].x * inv_n + a.video_corr, w[m].y * inv_n + a.video_corr);
  }
  c32 F[16];
#pragma unroll
  for (int m = 0; m < 16; m++) {
    const float2 d = row_at(stash, m);
    const float2 f = row_at(a.vh, m);
    F[m] = mul(c32{d.x, d.y}, c32{f.x, f.y});
  }
  const float2 dM = stash[2 * kFftM];
  const float2 fM = a.vh[kFftM];
  const c32 fm = mul(c32{dM.x, dM.y}, c32{fM.x, fM.y});
  __syncthreads();  // twiddles and transition matrices in LDS; D read by every thread
  float2* st_go = stash;          // odd-pass input
  float2* st_ev = stash + kFftM;  // even samples (real part, magnitude), then real parts (even, odd)
  const c32 wj32 = tw32k(s.oct, j);  // roots of k = j + 1024 m: rot32(wj32, m)
#pragma unroll
  for (int m = 0; m < 16; m++) {
    const int k = j + m * kFftThreads;
    const c32 g = (k == 0) ? sub(F[m], fm) : mul(scale(F[m], 2.0f), conjc(rot32(wj32, m)));
    row_at(st_go, m) = make_float2(g.x, g.y);
  }
#pragma unroll
  for (int m = 0; m < 16; m++) {
    const int k = j + m * kFftThreads;
    F[m] = (k == 0) ? add(F[m], fm) : scale(F[m], 2.0f);
  }
  fft16k_inverse(F, s.buf, s.oct);
#pragma unroll
  for (int m = 0; m < 16; m++) {
    const c32 e = scale(F[m], inv_n);
    row_at(st_ev, m) = make_float2(e.x, numpy_abs(e));
  }
#pragma unroll
  for (int m = 0; m < 16; m++) {
    const float2 g = row_at(st_go, m);
    F[m] = c32{g.x, g.y};
  }
  fft16k_inverse(F, s.buf, s.oct);
  float* ext = reinterpret_cast<float*>(s.buf) + kSosMargin;
  float* amp = ext + a.amp.pad;
  const float hd = a.half_deviation;
#pragma unroll
  for (int m = 0; m < 16; m++) {
    const int k = j + m * kFftThreads;
    const c32 o = scale(F[m], inv_n);
    const float2 ev = row_at(st_ev, m);
    amp[2 * k] = ev.y / hd;
    amp[2 * k + 1] = numpy_abs(o) / hd;
    row_at(st_ev, m) = make_float2(ev.x, o.x);
  }
  __syncthreads();
  sos_filtfilt_block_t<1>(a.amp, ext, 32768, s.scratch, nullptr, 0, 0, 0, false);
  const float* y = ext + a.amp.pad;
  float* vout = a.out_video + a.out_offset[blk];
  auto gain = [&](float amp) {
    float v = fmaxf(0.0f, amp) * a.scaling1;
    v = powf(v, a.exp_scaling) * a.scaling2;
    if (a.logistic_rate > 0.0f) v *= 1.0f / (1.0f + expf(-a.logistic_rate * (v - a.logistic_mid)));
    return v;
  };
  if (!a.notch) {
    for (int o = j; o < a.out_count; o += kFftThreads) {
      const int i = a.out_begin + o;
      const float v = gain(y[i]);
      const float2 r = st_ev[i >> 1];
      const float hp = (i & 1) ? r.y : r.x;
      vout[o] -= (hp * (1.0f - v)) + (hp * a.static_factor);
    }
    return;
  }
  // fsc notch (RfDemodulator.ApplyBetamaxFscNotchInPlace, RfDemodulator.cs:1018-1033) over the
  // whole corrected block, then the kept window
#pragma unroll
  for (int m = 0; m < 16; m++) {
    const int k = j + m * kFftThreads;
    const float g0 = gain(y[2 * k]), g1 = gain(y[2 * k + 1]);
    const float2 r = row_at(st_ev, m);
    const float2 vv = row_at(aux, m);
    row_at(aux, m) = make_float2(vv.x - ((r.x * (1.0f - g0)) + (r.x * a.static_factor)),
                                 vv.y - ((r.y * (1.0f - g1)) + (r.y * a.static_factor)));
  }
  __syncthreads();
  float* nx = ext + a.notch_sos.pad;
#pragma unroll
  for (int m = 0; m < 16; m++) {
    const int k = j + m * kFftThreads;
    const float2 c = row_at(aux, m);
    nx[2 * k] = c.x;
    nx[2 * k + 1] = c.y;
  }
  __syncthreads();
  sos_filtfilt_block_t<1>(a.notch_sos, ext, 32768, s.scratch, vout, a.out_begin,
                          a.out_count, /*roll=*/0, false);
}

__global__ void __launch_bounds__(1024) vhs_block_subdeemph_kernel(SubDeemphArgs a) { vhs_block_subdeemph_body(a, blockIdx.x); }


// ------------------------------------------------------------------ spike repair
// ReplaceSpikes (RfDemodulator.cs:966-1001) for one 32768-sample block: spike indices
// come from the unmodified demod dm (global), then windows [i-8, min(i+30, N-1)) are
// processed in ascending order against the progressively repaired demod.  The order
// is inherently sequential; one wave walks the spike list with a lane per window
// sample (a window is at most 38 samples) on the demod held in LDS.  `list` is
// 32768 ints of global scratch.  On return the repaired demod is in s.buf (as floats)
// and every thread has passed a barrier.
__device__ __forceinline__ void replace_spikes_block(const float* __restrict__ dm, const float* __restrict__ dd,
                                                     int* __restrict__ list, double mx, BlockLds& s) {
  const int j = threadIdx.x;
  // ascending spike list: thread j owns samples [32 j, 32 j + 32)
  unsigned mask = 0;
#pragma unroll 4
  for (int q = 0; q < 32; q++) mask |= ((double)dm[32 * j + q] > mx ? 1u : 0u) << q;
  const int cnt = __popc(mask);
  int incl = cnt;
  const int lane = j & 63, wave = j >> 6;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_up(incl, o);
    if (lane >= o) incl += u;
  }
  int* wsum = reinterpret_cast<int*>(s.scratch);
  if (lane == 63) wsum[wave] = incl;
  float* dms = reinterpret_cast<float*>(s.buf);  // 32768 floats
  for (int n = j; n < 32768; n += kFftThreads) dms[n] = dm[n];
  __syncthreads();
  int base = incl - cnt, total = 0;
  for (int w = 0; w < kFftThreads / 64; w++) {
    const int v = wsum[w];
    if (w < wave) base += v;
    total += v;
  }
  while (mask) {
    const int q = __ffs(mask) - 1;
    mask &= mask - 1;
    list[base++] = 32 * j + q;
  }
  __threadfence_block();
  __syncthreads();
  if (j < 64) {
    for (int q0 = 0; q0 < total; q0 += 64) {
      const int mine = q0 + lane < total ? list[q0 + lane] : 0;
      const int nq = total - q0 < 64 ? total - q0 : 64;
      for (int q = 0; q < nq; q++) {
        const int i = __shfl(mine, q);
        const int st = i - 8 > 0 ? i - 8 : 0;
        const int en = i + 30 < 32767 ? i + 30 : 32767;
        if (st >= en) continue;
        const int k = st + lane;
        const bool valid = k < en;
        const float d = valid ? dd[k] : -INFINITY;
        float md = d, mo = valid ? dms[k] : -INFINITY;
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
          md = fmaxf(md, __shfl_xor(md, o));
          mo = fmaxf(mo, __shfl_xor(mo, o));
        }
        if (md < mo && valid) dms[k] = d;
      }
    }
  }
  __syncthreads();
}

__device__ __forceinline__ void vhs_block_repair_body(const RepairArgs& r, const int blk) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  BlockLds& s = *reinterpret_cast<BlockLds*>(lds_raw);
  const DemodArgs& a = r.d;
  const int j = threadIdx.x;
  if ((a.flags[blk] & 2u) == 0) return;  // uniform across the workgroup
  const long long start = a.block_start[blk];
  float2* A = r.scratch_analytic + (long long)blk * 32768;
  float* dm = r.scratch_demod + (long long)blk * 65536;
  float* dd = dm + 32768;
  load_octant(s, a.oct);
  c32 v[16];
  const int16_t* x = a.rf + start;
#pragma unroll
  for (int m = 0; m < 16; m++) {
    const int k = j + m * kFftThreads;
    v[m] = c32{(float)x[2 * k], (float)x[2 * k + 1]};
  }
  __syncthreads();
  fft16k_forward(v, s.buf, s.oct);
  c32 xm;
  rfft_post(v, s, xm);
#pragma unroll
  for (int m = 0; m < 16; m++) {
    const float2 f = a.rf_video[j + m * kFftThreads];
    v[m] = mul(v[m], c32{f.x, f.y});
  }
  const float2 fM = a.rf_video[kFftM];
  if (j == 0) {
    const c32 fm = mul(xm, c32{fM.x, fM.y});
    s.scratch[0] = fm.x;
    s.scratch[1] = fm.y;
  }
  __syncthreads();
  const c32 fm = c32{s.scratch[0], s.scratch[1]};
  __syncthreads();
  // analytic signal, materialised
  const float inv_n = 1.0f / 32768.0f;
  c32 w[16];
#pragma unroll
  for (int m = 0; m < 16; m++) {
    const int k = j + m * kFftThreads;
    w[m] = (k == 0) ? add(v[m], fm) : scale(v[m], 2.0f);
  }
  fft16k_inverse(w, s.buf, s.oct);
#pragma unroll
  for (int m = 0; m < 16; m++) {
    const int k = j + m * kFftThreads;
    A[2 * k] = make_float2(w[m].x * inv_n, w[m].y * inv_n);
  }
  const c32 wj32 = tw32k(s.oct, j);  // roots of k = j + 1024 m: rot32(wj32, m)
#pragma unroll
  for (int m = 0; m < 16; m++) {
    const int k = j + m * kFftThreads;
    w[m] = (k == 0) ? sub(v[m], fm) : mul(scale(v[m], 2.0f), conjc(rot32(wj32, m)));
  }
  fft16k_inverse(w, s.buf, s.oct);
#pragma unroll
  for (int m = 0; m < 16; m++) {
    const int k = j + m * kFftThreads;
    A[2 * k + 1] = make_float2(w[m].x * inv_n, w[m].y * inv_n);
  }
  __threadfence_block();
  __syncthreads();
  // demod and diffed-analytic demod, element-wise (angles recomputed per sample)
  for (int n = j; n < 32768; n += kFftThreads) {
    const float2 c = A[n];
    const float tc = vhs_atan2(c.y, c.x);
    float2 p = n > 0 ? A[n - 1] : make_float2(0.0f, 0.0f);
    const float tp = vhs_atan2(p.y, p.x);
    dm[n] = n == 0 ? 0.0f : vhs_freq_diff(tc, tp, a.fs);
    // diffed[n] = A[n] - A[n-1], diffed[0] = 0
    float2 dc = n > 0 ? make_float2(c.x - p.x, c.y - p.y) : make_float2(0.0f, 0.0f);
    float2 dp = make_float2(0.0f, 0.0f);
    if (n > 1) {
      const float2 pp = A[n - 2];
      dp = make_float2(p.x - pp.x, p.y - pp.y);
    }
    const float tdc = vhs_atan2(dc.y, dc.x);
    const float tdp = vhs_atan2(dp.y, dp.x);
    dd[n] = n == 0 ? 0.0f : vhs_freq_diff(tdc, tdp, a.fs);
  }
  __threadfence_block();
  __syncthreads();
  // ReplaceSpikes (RfDemodulator.cs:966-1001): spike indices are collected from the
  // unmodified demod, then windows [i-8, min(i+30, N-1)) are processed in ascending
  // order against the progressively repaired demod.  The order is inherently
  // sequential; one wave walks the spike list with a lane per window sample (the
  // window is at most 38 samples), the demod held in LDS.
  replace_spikes_block(dm, dd, reinterpret_cast<int*>(A), a.diff_max, s);  // the analytic scratch is free now
  const float* dms = reinterpret_cast<const float*>(s.buf);
  // the repaired demod replaces the block's stashed demod; the video kernel runs next
  float2* stash = r.d.stash + (long long)blk * kStashEntries;
#pragma unroll
  for (int m = 0; m < 16; m++) {
    const int k = j + m * kFftThreads;
    row_at(stash, m) = make_float2(dms[2 * k], dms[2 * k + 1]);
  }
}

__global__ void __launch_bounds__(1024) vhs_block_repair_kernel(RepairArgs r) { vhs_block_repair_body(r, blockIdx.x); }


// Fills n words (a kernel rather than hipMemsetAsync, so the chain can be captured
// into a graph and replayed).
__global__ void __launch_bounds__(256) fill_u32_kernel(unsigned* p, long long n, unsigned value) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = value;
}

// ------------------------------------------------------------------ diagnostics
// One block's ReplaceSpikes on caller-supplied demod and diffed demod.
__global__ void __launch_bounds__(1024) debug_replace_spikes_kernel(const float* dm, const float* dd, int* list,
                                                                    double mx, float* out) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  BlockLds& s = *reinterpret_cast<BlockLds*>(lds_raw);
  replace_spikes_block(dm, dd, list, mx, s);
  const float* dms = reinterpret_cast<const float*>(s.buf);
  for (int n = threadIdx.x; n < 32768; n += kFftThreads) out[n] = dms[n];
}

__global__ void __launch_bounds__(1024) debug_fft_kernel(float2* data, int inverse, const float2* oct_g) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  BlockLds& s = *reinterpret_cast<BlockLds*>(lds_raw);
  const int j = threadIdx.x;
  float2* d = data + (long long)blockIdx.x * kFftM;
  load_octant(s, oct_g);
  c32 v[16];
#pragma unroll
  for (int m = 0; m < 16; m++) {
    const float2 t = d[j + m * kFftThreads];
    v[m] = c32{t.x, t.y};
  }
  __syncthreads();
  if (inverse)
    fft16k_inverse(v, s.buf, s.oct);
  else
    fft16k_forward(v, s.buf, s.oct);
#pragma unroll
  for (int m = 0; m < 16; m++) d[j + m * kFftThreads] = make_float2(v[m].x, v[m].y);
}

}  // namespace vhsd

// LaserDisc analog audio stage 2 (LaserDiscAnalogAudioPhase2.Apply,
// Decode/LaserDiscAnalogAudioPhase2.cs:12-220) over the assembled audio of a read: blocks
// of the stage-2 length N = 32768 overlapping by 512 samples (the first at 0, the middle
// ones every N - 512, the last at count - N - 1), one workgroup per (block, channel).  Per
// block: the float32 offsets from the float32 carrier (PrepareChannel); the left channel's
// peaks (values from 500 kHz up; a peak at i - 1 wherever sample i exceeds the block's last
// value and sample i + 1 exceeds sample i) zero samples [p - 8, p + 8) of both channels;
// the zero-padded block's real transform times the stage-2 response, the inverse
// transform plus the carrier.  Each output sample is written by the one block whose copy
// the reference keeps (the last block's copy is placed one sample late, as there).
namespace vhsd {

constexpr int kLdAudio2Skip = 512;
__device__ __forceinline__ float ld2_level(float v) { return v < 500000.0f ? 0.0f : v; }

__global__ void __launch_bounds__(1024) ld_audio_phase2_kernel(LdAudio2Args a) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  BlockLds& s = *reinterpret_cast<BlockLds*>(lds_raw);
  const int j = threadIdx.x;
  const int b = blockIdx.x;
  const int ch = blockIdx.y;
  const long long count = a.count;
  const int N = 2 * kFftM;
  const long long J = N - kLdAudio2Skip;
  const bool single = count <= N;
  // 0 first, 1 middle, 2 last
  const int kind = single || b == 0 ? 0 : (b <= a.nmid ? 1 : 2);
  const long long start = kind == 0 ? 0 : (kind == 1 ? (long long)b * J : count - N - 1);
  const int cnt = (int)min((long long)N, count - start);
  load_octant(s, a.oct);
  float* raw = reinterpret_cast<float*>(s.buf);
  // the left channel's offsets, for the peaks
  const float cl = (float)a.center[0];
  for (int i = j; i < N; i += kFftThreads) raw[i] = i < cnt ? a.in[0][start + i] - cl : 0.0f;
  __syncthreads();
  // this thread's samples [32 j, 32 j + 32): zeroed where a peak i in [q - 6, q + 9] lies
  unsigned zero = 0;
  {
    const float last = ld2_level(raw[cnt - 1]);
    const int q0 = 32 * j;
    float cur = 0.0f;
    int i = max(0, q0 - 6);
    if (i < cnt) cur = ld2_level(raw[i]);
    for (; i <= q0 + 40 && i < cnt - 1; i++) {
      const float nxt = ld2_level(raw[i + 1]);
      if (cur > last && nxt > cur) {
        const int p = i - 1;
        const int lo = max(max(0, p - 8), q0), hi = min(min(p + 8, cnt), q0 + 32);
        for (int q = lo; q < hi; q++) zero |= 1u << (q - q0);
      }
      cur = nxt;
    }
  }
  __syncthreads();
  if (ch == 1) {
    const float cr = (float)a.center[1];
    for (int i = j; i < N; i += kFftThreads) raw[i] = i < cnt ? a.in[1][start + i] - cr : 0.0f;
    __syncthreads();
  }
  for (int r = 0; r < 32; r++)
    if ((zero >> r) & 1u) raw[32 * j + r] = 0.0f;
  __syncthreads();
  c32 v[16];
#pragma unroll
  for (int m = 0; m < 16; m++) {
    const int k = j + m * kFftThreads;
    v[m] = c32{raw[2 * k], raw[2 * k + 1]};
  }
  __syncthreads();
  fft16k_forward(v, s.buf, s.oct);
  c32 xm;
  rfft_post(v, s, xm);
  const float2* filt = a.filt[ch];
#pragma unroll
  for (int m = 0; m < 16; m++) {
    const float2 f = row_at(filt, m);
    v[m] = mul(v[m], c32{f.x, f.y});
  }
  const float2 fM = filt[kFftM];
  const c32 vm = mul(xm, c32{fM.x, fM.y});
  irfft_pre(v, vm, s);
  fft16k_inverse(v, s.buf, s.oct);
  const float inv_n = 1.0f / 32768.0f;
  const double center = a.center[ch];
  const long long final_start = count - (N - kLdAudio2Skip);
  float* out = a.out[ch];
#pragma unroll
  for (int m = 0; m < 16; m++) {
    const int k = j + m * kFftThreads;
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const int li = 2 * k + h;
      const float y = (h == 0 ? v[m].x : v[m].y) * inv_n;
      long long o;
      bool own;
      if (single) {
        o = li;
        own = li < cnt;
      } else if (kind == 0) {
        o = li;
        own = o < final_start;
      } else if (kind == 1) {
        o = start + li;
        own = li >= kLdAudio2Skip && li < cnt && o < final_start;
      } else {
        o = start + 1 + li;
        own = li >= kLdAudio2Skip;
      }
      if (own) out[o] = (float)((double)y + center);
    }
  }
}

}  // namespace vhsd

// ------------------------------------------------------------------ launchers
template <typename K, typename... Args>
static hipError_t launch_block_kernel(K kernel, int count, hipStream_t stream, Args... args) {
  const size_t lds = sizeof(vhsd::BlockLds);
  hipError_t e = vhsd::lds_limit_once((const void*)kernel, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(kernel, dim3(count), dim3(1024), lds, stream, args...);
  return hipGetLastError();
}


template <int S>
static hipError_t launch_chroma_s(const vhsd::ChromaArgs* a, int nblocks, hipStream_t stream) {
  const size_t lds = sizeof(vhsd::BlockLds);
  hipError_t e = vhsd::lds_limit_once((const void*)vhsd::vhs_block_chroma_kernel<S>, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(vhsd::vhs_block_chroma_kernel<S>, dim3(nblocks), dim3(1024), lds, stream, *a);
  return hipGetLastError();
}

extern "C" {

size_t vhsd_block_lds_bytes() { return sizeof(vhsd::BlockLds); }

hipError_t vhsd_launch_block_demod(const vhsd::DemodArgs* a, int nblocks, hipStream_t stream) {
  const size_t lds = sizeof(vhsd::BlockLds);
  hipError_t e = vhsd::lds_limit_once((const void*)vhsd::vhs_block_demod_kernel, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(vhsd::vhs_block_demod_kernel, dim3(nblocks), dim3(1024), lds, stream, *a);
  return hipGetLastError();
}

hipError_t vhsd_launch_block_chroma(const vhsd::ChromaArgs* a, int nblocks, hipStream_t stream) {
  switch (a->sos.sections) {
    case 1: return launch_chroma_s<1>(a, nblocks, stream);
    case 2: return launch_chroma_s<2>(a, nblocks, stream);
    case 3: return launch_chroma_s<3>(a, nblocks, stream);
    case 4: return launch_chroma_s<4>(a, nblocks, stream);
    case 10: return launch_chroma_s<10>(a, nblocks, stream);
    default: return hipErrorInvalidValue;
  }
}

hipError_t vhsd_launch_debug_replace_spikes(const float* dm, const float* dd, int* list, double mx, float* out,
                                           hipStream_t stream) {
  return launch_block_kernel(vhsd::debug_replace_spikes_kernel, 1, stream, dm, dd, list, mx, out);
}

hipError_t vhsd_launch_ld_audio_phase2(const vhsd::LdAudio2Args* a, int nblocks, hipStream_t stream) {
  const size_t lds = sizeof(vhsd::BlockLds);
  hipError_t e = vhsd::lds_limit_once((const void*)vhsd::ld_audio_phase2_kernel, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(vhsd::ld_audio_phase2_kernel, dim3(nblocks, 2), dim3(1024), lds, stream, *a);
  return hipGetLastError();
}

hipError_t vhsd_launch_debug_fft(float* data, int count, int inverse, const float2* oct, hipStream_t stream) {
  const size_t lds = sizeof(vhsd::BlockLds);
  hipError_t e = vhsd::lds_limit_once((const void*)vhsd::debug_fft_kernel, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(vhsd::debug_fft_kernel, dim3(count), dim3(1024), lds, stream, reinterpret_cast<float2*>(data),
                     inverse, oct);
  return hipGetLastError();
}

hipError_t vhsd_launch_block_demod_boost(const vhsd::DemodArgs* a, int nblocks, hipStream_t stream) {
  const size_t lds = sizeof(vhsd::BlockLds);
  hipError_t e = vhsd::lds_limit_once((const void*)vhsd::vhs_block_demod_boost_kernel, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(vhsd::vhs_block_demod_boost_kernel, dim3(nblocks), dim3(1024), lds, stream, *a);
  return hipGetLastError();
}


namespace vhsd {
// The first stage of a batch: every demod kind and every chroma section count in one
// launch, so blocks of all formats run together (the kind is uniform per workgroup).
__global__ void __launch_bounds__(1024) vhs_block_stage1_table_kernel(const DemodArgs* __restrict__ drecs,
                                                                      const ChromaArgs* __restrict__ crecs,
                                                                      StageTable t) {
  const int i = blockIdx.x;
  if (i < t.n_demod) {
    const int blk = t.demod_ids[i];
    const int g = t.block_rec[blk];
    switch (t.rec_kind[g]) {
      case kTableDemodBoost: vhs_block_demod_boost_body(drecs[g], blk); break;
      case kTableDemodLd: vhs_block_demod_ld_body(drecs[g], blk); break;
      default: vhs_block_demod_body(drecs[g], blk); break;
    }
  } else {
    const int blk = t.chroma_ids[i - t.n_demod];
    const int g = t.block_rec[blk];
    switch (t.rec_sections[g]) {
      case 1: vhs_block_chroma_body<1>(crecs[g], blk); break;
      case 2: vhs_block_chroma_body<2>(crecs[g], blk); break;
      case 3: vhs_block_chroma_body<3>(crecs[g], blk); break;
      case 4: vhs_block_chroma_body<4>(crecs[g], blk); break;
      default: vhs_block_chroma_body<10>(crecs[g], blk); break;
    }
  }
}

// The second stage of a batch: repair, video and sub-deemphasis of one block in one
// workgroup, in that order (each body starts from the block's stash and planes the
// previous one wrote; the barrier between them orders the workgroup's global and LDS
// accesses).
__global__ void __launch_bounds__(1024) vhs_block_stage2_table_kernel(const RepairArgs* __restrict__ rrecs,
                                                                      const DemodArgs* __restrict__ drecs,
                                                                      const SubDeemphArgs* __restrict__ srecs,
                                                                      StageTable2 t) {
  const int blk = blockIdx.x;
  const int g = t.block_rec[blk];
  if (t.rec_repair[g]) {
    vhs_block_repair_body(rrecs[g], blk);
    __syncthreads();
  }
  vhs_block_video_body(drecs[g], blk);
  if (t.rec_subdeemph[g]) {
    __syncthreads();
    vhs_block_subdeemph_body(srecs[g], blk);
  }
}
}  // namespace vhsd

hipError_t vhsd_launch_stage2_table(const vhsd::RepairArgs* rrecs, const vhsd::DemodArgs* drecs,
                                    const vhsd::SubDeemphArgs* srecs, vhsd::StageTable2 t, int blocks,
                                    hipStream_t stream) {
  if (blocks <= 0) return hipSuccess;
  return launch_block_kernel(vhsd::vhs_block_stage2_table_kernel, blocks, stream, rrecs, drecs, srecs, t);
}

hipError_t vhsd_launch_stage1_table(const vhsd::DemodArgs* drecs, const vhsd::ChromaArgs* crecs,
                                    vhsd::StageTable t, int n_chroma, hipStream_t stream) {
  const int count = t.n_demod + n_chroma;
  if (count <= 0) return hipSuccess;
  return launch_block_kernel(vhsd::vhs_block_stage1_table_kernel, count, stream, drecs, crecs, t);
}

hipError_t vhsd_launch_fill_u32(unsigned* p, long long n, unsigned value, hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  const int blocks = (int)std::min<long long>((n + 255) / 256, 1024);
  hipLaunchKernelGGL(vhsd::fill_u32_kernel, dim3(blocks), dim3(256), 0, stream, p, n, value);
  return hipGetLastError();
}

hipError_t vhsd_launch_block_demod_ld(const vhsd::DemodArgs* a, int nblocks, hipStream_t stream) {
  const size_t lds = sizeof(vhsd::BlockLds);
  hipError_t e = vhsd::lds_limit_once((const void*)vhsd::vhs_block_demod_ld_kernel, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(vhsd::vhs_block_demod_ld_kernel, dim3(nblocks), dim3(1024), lds, stream, *a);
  return hipGetLastError();
}

hipError_t vhsd_launch_block_video(const vhsd::DemodArgs* a, int nblocks, hipStream_t stream) {
  const size_t lds = sizeof(vhsd::BlockLds);
  hipError_t e = vhsd::lds_limit_once((const void*)vhsd::vhs_block_video_kernel, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(vhsd::vhs_block_video_kernel, dim3(nblocks), dim3(1024), lds, stream, *a);
  return hipGetLastError();
}

hipError_t vhsd_launch_block_subdeemph(const vhsd::SubDeemphArgs* a, int nblocks, hipStream_t stream) {
  const size_t lds = sizeof(vhsd::BlockLds);
  hipError_t e = vhsd::lds_limit_once((const void*)vhsd::vhs_block_subdeemph_kernel, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(vhsd::vhs_block_subdeemph_kernel, dim3(nblocks), dim3(1024), lds, stream, *a);
  return hipGetLastError();
}

hipError_t vhsd_launch_block_repair(const vhsd::RepairArgs* a, int nblocks, hipStream_t stream) {
  const size_t lds = sizeof(vhsd::BlockLds);
  hipError_t e = vhsd::lds_limit_once((const void*)vhsd::vhs_block_repair_kernel, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(vhsd::vhs_block_repair_kernel, dim3(nblocks), dim3(1024), lds, stream, *a);
  return hipGetLastError();
}

}  // extern "C"
