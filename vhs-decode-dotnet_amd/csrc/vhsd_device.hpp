// Device helpers shared by the decode kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <mutex>
#include <utility>
#include <vector>

namespace vhsd {

// threadIdx.x through an opaque register copy.  A kernel inlines several transforms;
// without this the compiler computes each pass's per-thread LDS addresses and twiddle
// indices once and keeps them live across every transform (and the code between
// them), which exhausts the 128-VGPR budget of a 1024-thread workgroup.  Recomputing
// them per transform costs a few integer ops.
__device__ __forceinline__ int opaque_tid() {
  int j = threadIdx.x;
  asm volatile("v_mov_b32 %0, %0" : "+v"(j));
  return j;
}


constexpr int kScratchFloats = 512;

// Raises kernel `fn`'s dynamic LDS limit to `bytes` on the current device, once: the
// attribute stays with the function on that device, so later launches skip the call
// (it took a measurable share of the host's enqueue time when made per launch).
inline hipError_t lds_limit_once(const void* fn, int bytes) {
  static std::mutex mu;
  static std::vector<std::pair<const void*, int>> done;  // (kernel, device) pairs set
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> lock(mu);
  for (const auto& d : done)
    if (d.first == fn && d.second == dev) return hipSuccess;
  e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  if (e == hipSuccess) done.emplace_back(fn, dev);
  return e;
}

// The pulse searches' test ((double)v + bias) + dc <= high of an FP32 sample v as one FP32
// comparison v <= T: the two FP64 additions round monotonically, so the samples that pass
// are those up to the largest float T that passes (found from the nearest float to
// high - dc - bias and a step or two either way).  NaN when no sample passes, +inf when
// every one does; NaN samples fail both forms.
__device__ __forceinline__ float pass_bound_f32(double bias, double dc, double high) {
#pragma clang fp contract(off)
  auto pass = [&](float v) { return (((double)v + bias) + dc) <= high; };
  if (!pass(-INFINITY)) return __builtin_nanf("");
  if (pass(INFINITY)) return INFINITY;
  float t = (float)((high - dc) - bias);
  if (!(t == t)) t = 0.0f;
  if (t == INFINITY) t = 3.40282347e38f;
  if (t == -INFINITY) t = -3.40282347e38f;
  while (!pass(t)) t = nextafterf(t, -INFINITY);
  while (pass(nextafterf(t, INFINITY))) t = nextafterf(t, INFINITY);
  return t;
}

// PortedMath.VhsRustAtan2Approximation (PortedMath.cs:519-544): FP32 polynomial
// atan2, evaluated with separately rounded operations (no contraction) in the
// reference's order.  The quotient is the hardware reciprocal's product with one FMA
// correction step, which gives the IEEE quotient but for rare last-bit ties: the
// analytic samples it divides are FP32 transform outputs whose own rounding is ~1e-6
// relative, so the discriminator is judged by tolerance and the 9-instruction IEEE
// division sequence would buy nothing.
__device__ __forceinline__ float vhs_atan2(float y, float x) {
#pragma clang fp contract(off)
  const float min_normal = 1.17549435E-38f;
  x += copysignf(min_normal, x);
  const bool swap = fabsf(x) < fabsf(y);
  // one division of selected operands; |den| >= min_normal and |num| <= |den|
  const float num = swap ? x : y, den = swap ? y : x;
  const float rc = __builtin_amdgcn_rcpf(den);
  const float q0 = num * rc;
  const float input = __builtin_fmaf(__builtin_fmaf(-q0, den, num), rc, q0);
  const float sq = input * input;
  float r = input * (0.99997726f +
                     (sq * (-0.33262347f +
                            (sq * (0.19354346f +
                                   (sq * (-0.11643287f + (sq * (0.05265332f + (sq * -0.01172120f))))))))));
  const float pi = 3.14159265358979323846f;
  const float half_pi = input >= 0.0f ? pi / 2.0f : -pi / 2.0f;
  r = swap ? half_pi - r : r;
  // quadrant fix without branches: x < 0 adds pi (y >= 0) or -pi (y < 0)
  const bool xp = x >= 0.0f, yp = y >= 0.0f;
  const float shifted = (yp ? pi : -pi) + r;
  return xp ? r : shifted;
}

// x / (2 pi) in FP32, correctly rounded: a multiply by the correctly rounded
// reciprocal and one FMA correction step give the IEEE quotient for every FP32 x with
// 2^-100 <= |x| <= 2^120 (checked exhaustively over all such inputs); zeros, tiny,
// huge and non-finite inputs take the IEEE division.
__device__ __forceinline__ float div_tau(float x) {
#pragma clang fp contract(off)
  const float tau = 6.28318530717958647692f;
  const float rcp = 1.0f / tau;
  const float ax = fabsf(x);
  if (!(ax >= 0x1p-100f && ax <= 0x1p+120f)) return x / tau;
  const float q = x * rcp;
  return __builtin_fmaf(__builtin_fmaf(-q, tau, x), rcp, q);
}

// PortedMath.VhsRustFrequencyDifference (PortedMath.cs:546-551): the two divisions by
// 2 pi are the IEEE quotients (div_tau).
__device__ __forceinline__ float vhs_freq_diff(float cur, float prev, float freq) {
#pragma clang fp contract(off)
  const float tau = 6.28318530717958647692f;
  float d = cur - prev;
  d -= floorf(div_tau(d)) * tau;
  return div_tau(d * freq);
}

}  // namespace vhsd
