// HiFi Hilbert FM discriminator for gfx950 (HiFiHilbertDiscriminator.Demodulate,
// src/VHSDecode.Core/HiFi/HiFiHilbertDiscriminator.cs:42-156).
//
// The reference takes the analytic signal of a whole IF block (2^22 samples at the
// 2^23 Hz Hilbert IF rate, two blocks a second) with one circular FFT pair, keeps the
// float phase and turns phase steps into frequency.  Here the N = N1 * 16384 point
// transforms are four-step FFTs over an N1 x 16384 matrix (n = 16384 n1 + n2,
// k = k1 + N1 k2):
//   1. columns: for each n2, an N1-point DFT over n1, times W_N^(n2 k1), written as
//      rows k1 (hifi_columns_forward_kernel);
//   2. rows: for each k1, the 16384-point FFT of the workgroup-resident kernel gives
//      X[k1 + N1 k2]; the Hilbert multiplier (1, 2 ... 2, 1, 0 ... 0) is applied and the
//      row goes straight back through the inverse 16384-point FFT and the conjugate
//      twiddle (hifi_rows_kernel) - the spectrum never leaves the workgroup;
//   3. columns again: the inverse N1-point DFTs give the analytic samples in natural
//      order; their FP64 atan2 rounded to float is the phase (hifi_columns_phase_kernel);
//   4. the discriminator, one output per phase step (hifi_discriminator_kernel).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "fft16k.hpp"

namespace vhsd {

constexpr int kHifiColTile = 32;          // columns per workgroup of the column kernels
constexpr int kHifiColThreads = 256;
constexpr int kHifiMaxN1 = 256;
constexpr int kHifiColStride = kHifiMaxN1 + 1;  // padded LDS column

struct HifiArgs {
  const float* in;        // [N] real IF samples
  float2* work;           // [N] complex: rows k1 of 16384
  float* phase;           // [N]
  float* out;             // [N]; out[N-1] is left untouched
  const float2* oct;      // 16384-point twiddle octant
  const float2* fine;     // exp(-2 pi i r / N), r < N / 16384
  int n1;                 // N / 16384 (power of two, <= 256)
  int log2n1;
  long long n;
  double carrier_scale;   // carrier / deviation
  double rate;            // (double)(float)sample_rate
  double deviation;
};

// exp(-2 pi i t / N) for N = 16384 * n1: the 16384-th root of t / n1 from the octant
// table times the fine root of t mod n1.
__device__ __forceinline__ c32 twiddle_n(const HifiArgs& a, long long t) {
  t &= a.n - 1;
  c32 w = tw16k(a.oct, (int)(t >> a.log2n1));
  const int r = (int)(t & (a.n1 - 1));
  if (r) {
    const float2 f = a.fine[r];
    w = mul(w, c32{f.x, f.y});
  }
  return w;
}

// Radix-2 Stockham DFT of length n1 down each of the tile's columns (column c at
// buf[c * kHifiColStride]); `inverse` uses exp(+i...).  Returns the result buffer.
__device__ __forceinline__ c32* column_dft(c32* x, c32* y, int n1, int log2n1, bool inverse,
                                           const float2* __restrict__ oct) {
  const int half = n1 >> 1;
  for (int p = 0; p < log2n1; p++) {
    const int ns = 1 << p;
    for (int t = threadIdx.x; t < kHifiColTile * half; t += blockDim.x) {
      const int c = t >> (log2n1 - 1);
      const int b = t & (half - 1);
      const int k = b & (ns - 1);
      const c32* xi = x + c * kHifiColStride;
      c32* yo = y + c * kHifiColStride;
      c32 w = tw16k(oct, (k * (kFftM >> p)) >> 1);
      if (inverse) w = conjc(w);
      const c32 u = xi[b];
      const c32 v = mul(xi[b + half], w);
      const int idx = ((b - k) << 1) + k;
      yo[idx] = add(u, v);
      yo[idx + ns] = sub(u, v);
    }
    __syncthreads();
    c32* t = x;
    x = y;
    y = t;
  }
  return x;
}

constexpr size_t kHifiColLds = 2 * sizeof(c32) * kHifiColTile * kHifiColStride;

__global__ void __launch_bounds__(kHifiColThreads) hifi_columns_forward_kernel(HifiArgs a) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  c32* b0 = reinterpret_cast<c32*>(lds_raw);
  c32* b1 = b0 + kHifiColTile * kHifiColStride;
  const int c0 = blockIdx.x * kHifiColTile;
  const int tile = kHifiColTile * a.n1;
  for (int e = threadIdx.x; e < tile; e += blockDim.x) {
    const int r = e / kHifiColTile, c = e % kHifiColTile;
    b0[c * kHifiColStride + r] = c32{a.in[(long long)r * kFftM + c0 + c], 0.0f};
  }
  __syncthreads();
  const c32* res = column_dft(b0, b1, a.n1, a.log2n1, false, a.oct);
  for (int e = threadIdx.x; e < tile; e += blockDim.x) {
    const int k1 = e / kHifiColTile, c = e % kHifiColTile;
    const c32 v = mul(res[c * kHifiColStride + k1], twiddle_n(a, (long long)(c0 + c) * k1));
    a.work[(long long)k1 * kFftM + c0 + c] = make_float2(v.x, v.y);
  }
}

struct HifiRowLds {
  c32 buf[kFftBufEntries];
  float2 oct[kOctantEntries];
};

__global__ void __launch_bounds__(1024) hifi_rows_kernel(HifiArgs a) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  HifiRowLds& s = *reinterpret_cast<HifiRowLds*>(lds_raw);
  c32* buf = s.buf;
  float2* oct = s.oct;
  const int j = threadIdx.x;
  const int k1 = blockIdx.x;
  for (int u = j; u < kOctantEntries; u += kFftThreads) oct[u] = a.oct[u];
  float2* row = a.work + (long long)k1 * kFftM;
  c32 v[16];
#pragma unroll
  for (int m = 0; m < 16; m++) {
    const float2 t = row[j + m * kFftThreads];
    v[m] = c32{t.x, t.y};
  }
  __syncthreads();
  fft16k_forward(v, buf, oct);
  // Hilbert multiplier of k = k1 + n1 k2 (PortedMath.BuildHilbertMultiplier semantics)
  const long long half = a.n >> 1;
#pragma unroll
  for (int m = 0; m < 16; m++) {
    const long long k = k1 + (long long)a.n1 * (j + m * kFftThreads);
    const float h = (k == 0 || k == half) ? 1.0f : (k < half ? 2.0f : 0.0f);
    v[m] = scale(v[m], h);
  }
  fft16k_inverse(v, buf, oct);
#pragma unroll
  for (int m = 0; m < 16; m++) {
    const int n2 = j + m * kFftThreads;
    const c32 w = conjc(twiddle_n(a, (long long)n2 * k1));
    const c32 y = mul(v[m], w);
    row[n2] = make_float2(y.x, y.y);
  }
}

__global__ void __launch_bounds__(kHifiColThreads) hifi_columns_phase_kernel(HifiArgs a) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  c32* b0 = reinterpret_cast<c32*>(lds_raw);
  c32* b1 = b0 + kHifiColTile * kHifiColStride;
  const int c0 = blockIdx.x * kHifiColTile;
  const int tile = kHifiColTile * a.n1;
  for (int e = threadIdx.x; e < tile; e += blockDim.x) {
    const int r = e / kHifiColTile, c = e % kHifiColTile;
    const float2 t = a.work[(long long)r * kFftM + c0 + c];
    b0[c * kHifiColStride + r] = c32{t.x, t.y};
  }
  __syncthreads();
  const c32* res = column_dft(b0, b1, a.n1, a.log2n1, true, a.oct);
  // the 1/N of the inverse transform is a power of two: atan2 is unchanged by it
  for (int e = threadIdx.x; e < tile; e += blockDim.x) {
    const int n1 = e / kHifiColTile, c = e % kHifiColTile;
    const c32 z = res[c * kHifiColStride + n1];
    a.phase[(long long)n1 * kFftM + c0 + c] = (float)atan2((double)z.y, (double)z.x);
  }
}

// HiFiHilbertDiscriminator.DemodulatePhase (HiFiHilbertDiscriminator.cs:88-156), one
// phase step per output: the step wrapped to [-pi, pi) gives the correction
// wrapped - step when |step| >= pi, and the output is correction / 2 pi * rate /
// deviation - carrier / deviation.  The reference forms the correction as the
// difference of two running sums; here it is the step's own correction.
__global__ void __launch_bounds__(256) hifi_discriminator_kernel(HifiArgs a) {
  const double pi = 3.141592653589793, tau = 6.283185307179586;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < a.n - 1;
       i += (long long)gridDim.x * blockDim.x) {
    const double cur = (double)a.phase[i];
    const double prev = i > 0 ? (double)a.phase[i - 1] : 0.0;
    const double diff = cur - prev;
    double m = fmod(diff + pi, tau);
    if (m < 0.0) m += tau;
    double wrapped = m - pi;
    if (wrapped == -pi && diff > 0.0) wrapped = pi;
    double corr = wrapped - diff;
    if (fabs(diff) < pi) corr = 0.0;
    const double value = corr / tau * a.rate / a.deviation;
    a.out[i] = (float)fmin(fmax(value - a.carrier_scale, -3.4028234663852886e38), 3.4028234663852886e38);
  }
}

}  // namespace vhsd

extern "C" {

size_t vhsd_hifi_args_size() { return sizeof(vhsd::HifiArgs); }

// Runs the four steps for one block; the caller has checked n = 16384 * n1 with n1 a
// power of two <= 256 and sized work/phase to n.
hipError_t vhsd_launch_hifi(const float* in, float* out, float2* work, float* phase, const float2* oct,
                            const float2* fine, int n1, double carrier_scale, double rate, double deviation,
                            hipStream_t stream) {
  vhsd::HifiArgs a{};
  a.in = in;
  a.work = work;
  a.phase = phase;
  a.out = out;
  a.oct = oct;
  a.fine = fine;
  a.n1 = n1;
  int lg = 0;
  while ((1 << lg) < n1) lg++;
  a.log2n1 = lg;
  a.n = (long long)n1 * vhsd::kFftM;
  a.carrier_scale = carrier_scale;
  a.rate = rate;
  a.deviation = deviation;
  const int col_blocks = vhsd::kFftM / vhsd::kHifiColTile;
  const int col_lds = (int)vhsd::kHifiColLds;
  const int row_lds = (int)sizeof(vhsd::HifiRowLds);
  hipError_t e;
  if ((e = vhsd::lds_limit_once((const void*)vhsd::hifi_columns_forward_kernel, col_lds)) != hipSuccess)
    return e;
  if ((e = vhsd::lds_limit_once((const void*)vhsd::hifi_columns_phase_kernel, col_lds)) != hipSuccess)
    return e;
  if ((e = vhsd::lds_limit_once((const void*)vhsd::hifi_rows_kernel, row_lds)) != hipSuccess)
    return e;
  hipLaunchKernelGGL(vhsd::hifi_columns_forward_kernel, dim3(col_blocks), dim3(vhsd::kHifiColThreads), col_lds,
                     stream, a);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(vhsd::hifi_rows_kernel, dim3(n1), dim3(1024), row_lds, stream, a);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(vhsd::hifi_columns_phase_kernel, dim3(col_blocks), dim3(vhsd::kHifiColThreads), col_lds,
                     stream, a);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  const long long n = a.n;
  const int disc_blocks = (int)std::min<long long>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(vhsd::hifi_discriminator_kernel, dim3(disc_blocks), dim3(256), 0, stream, a);
  return hipGetLastError();
}

}  // extern "C"
