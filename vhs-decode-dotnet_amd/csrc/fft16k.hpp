// Workgroup-resident 16384-point complex FFT for gfx950.
//
// One workgroup of 1024 threads (16 wave64s) owns one transform.  Each thread
// keeps 16 complex values in registers in "strided" distribution: thread j holds
// element j + 1024*m in v[m], m = 0..15.  The transform is a Stockham
// autosort FFT of radix 16, 16, 16, 4; each of the three radix-16 passes
// exchanges through a 128 KiB LDS buffer, the final radix-4 pass works on the
// thread's own registers, so the output comes back in the same strided
// distribution and natural order (v[m] = X[j + 1024*m]).
//
// Twiddles come from an LDS table of the first octant of the 16384-th roots of
// unity (oct[u] = (cos, sin)(2*pi*u/16384), u = 0..2048), expanded by octant
// symmetry, so every 16384-th root is a table value up to exact sign/swap
// operations.  A radix-16 pass looks up w, w^2, w^4 and w^8 and forms the other
// eleven powers as products of at most three table values; the radix-4 pass looks
// up the thread's W^j and W^2j once and rotates them by compile-time constants for its
// four butterflies (rot32).  Odd 32768-th roots take one extra multiply by
// exp(-i*pi/16384).
#pragma once
#include <hip/hip_runtime.h>

#include "vhsd_device.hpp"

namespace vhsd {

constexpr int kFftThreads = 1024;
constexpr int kFftM = 16384;          // complex points per transform
constexpr int kFftE = 16;             // complex values per thread
constexpr int kOctant = 2048;         // octant table covers u in [0, 2048] of the 16384-th roots
constexpr int kOctantEntries = kOctant + 1;
// The exchange buffer is padded by one entry every 32 so the radix-16 passes'
// strided writes spread over the LDS banks.
constexpr int kFftBufEntries = kFftM + kFftM / 32;

__device__ __forceinline__ int lds_pad(int i) { return i + (i >> 5); }


struct c32 {
  float x, y;
};

__device__ __forceinline__ c32 mk(float x, float y) { return c32{x, y}; }
__device__ __forceinline__ c32 add(c32 a, c32 b) { return c32{a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ c32 sub(c32 a, c32 b) { return c32{a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ c32 mul(c32 a, c32 b) { return c32{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ __forceinline__ c32 conjc(c32 a) { return c32{a.x, -a.y}; }
__device__ __forceinline__ c32 scale(c32 a, float s) { return c32{a.x * s, a.y * s}; }
// multiply by -i and +i
__device__ __forceinline__ c32 mul_mi(c32 a) { return c32{a.y, -a.x}; }
__device__ __forceinline__ c32 mul_pi(c32 a) { return c32{-a.y, a.x}; }

// exp(-2*pi*i*t/16384) for any t (taken mod 16384), from the octant table, without
// branches: the octant reflection is a select, the quadrant rotation by (-i)^q a
// swap select plus sign flips.
__device__ __forceinline__ c32 tw16k(const float2* __restrict__ oct, int t) {
  t &= 16383;
  const int q = t >> 12;  // quadrant of 4096 steps
  const int r = t & 4095;
  const bool lo = r <= kOctant;
  const float2 e = oct[lo ? r : 4096 - r];
  const float c = lo ? e.x : e.y;
  const float s = lo ? e.y : e.x;
  // exp(-i*theta) (-i)^q: q=0 (c,-s), q=1 (-s,-c), q=2 (-c,s), q=3 (s,c)
  const bool odd = q & 1;
  const float a = odd ? s : c;
  const float b = odd ? c : s;
  const unsigned fx = ((unsigned)((q + 1) >> 1) & 1u) << 31;  // negate x for q = 1, 2
  const unsigned fy = (unsigned)(q < 2) << 31;                 // negate y for q = 0, 1
  return c32{__uint_as_float(__float_as_uint(a) ^ fx), __uint_as_float(__float_as_uint(b) ^ fy)};
}

// exp(-2*pi*i*t/32768) for t in [0, 32768)
__device__ __forceinline__ c32 tw32k(const float2* __restrict__ oct, int t) {
  c32 w = tw16k(oct, t >> 1);
  if (t & 1) {
    const c32 w1 = c32{0.99999998161642933f, -1.9174759731070330e-4f};
    w = mul(w, w1);
  }
  return w;
}

// exp(-2*pi*i*m/32) rounded to float, m = 0..15 (m a compile-time constant after
// unrolling, so the switch folds to two literals).
__device__ __forceinline__ c32 w32c(int m) {
  switch (m & 15) {
    case 0: return c32{1.0f, 0.0f};
    case 1: return c32{0.98078528040323044f, -0.19509032201612826f};
    case 2: return c32{0.92387953251128674f, -0.38268343236508978f};
    case 3: return c32{0.83146961230254524f, -0.55557023301960218f};
    case 4: return c32{0.70710678118654752f, -0.70710678118654752f};
    case 5: return c32{0.55557023301960218f, -0.83146961230254524f};
    case 6: return c32{0.38268343236508978f, -0.92387953251128674f};
    case 7: return c32{0.19509032201612826f, -0.98078528040323044f};
    case 8: return c32{0.0f, -1.0f};
    case 9: return c32{-0.19509032201612826f, -0.98078528040323044f};
    case 10: return c32{-0.38268343236508978f, -0.92387953251128674f};
    case 11: return c32{-0.55557023301960218f, -0.83146961230254524f};
    case 12: return c32{-0.70710678118654752f, -0.70710678118654752f};
    case 13: return c32{-0.83146961230254524f, -0.55557023301960218f};
    case 14: return c32{-0.92387953251128674f, -0.38268343236508978f};
    default: return c32{-0.98078528040323044f, -0.19509032201612826f};
  }
}

// w * exp(-2*pi*i*m/32) for a compile-time m: the root of index k + 1024 m of the
// 32768-th roots from the root of k (m = 8 is the exact multiply by -i).  One table
// lookup per thread then serves all sixteen strided slots; the product adds about one
// ulp to the table value's rounding.
__device__ __forceinline__ c32 rot32(c32 w, int m) {
  if (m == 0) return w;
  if (m == 8) return mul_mi(w);
  return mul(w, w32c(m));
}

__device__ __forceinline__ void dft4(c32& a0, c32& a1, c32& a2, c32& a3) {
  c32 s02 = add(a0, a2), d02 = sub(a0, a2);
  c32 s13 = add(a1, a3), d13 = sub(a1, a3);
  a0 = add(s02, s13);
  a2 = sub(s02, s13);
  a1 = add(d02, mul_mi(d13));
  a3 = sub(d02, mul_mi(d13));
}

// 16-point forward DFT in place, natural order in and out.
__device__ __forceinline__ void dft16(c32* v) {
  const float c1 = 0.92387953251128674f;  // cos(pi/8)
  const float s1 = 0.38268343236508978f;  // sin(pi/8)
  const float r2 = 0.70710678118654752f;
  // step 1: DFT4 over a for each b: elements 4a+b
#pragma unroll
  for (int b = 0; b < 4; b++) dft4(v[b], v[4 + b], v[8 + b], v[12 + b]);
  // now v[4c + b] = Y[b][c]; step 2: twiddle W16^(b*c)
  // b=1: c=1..3 -> W^1, W^2, W^3
  v[4 + 1] = mul(v[4 + 1], mk(c1, -s1));
  v[8 + 1] = mul(v[8 + 1], mk(r2, -r2));
  v[12 + 1] = mul(v[12 + 1], mk(s1, -c1));
  // b=2: W^2, W^4, W^6
  v[4 + 2] = mul(v[4 + 2], mk(r2, -r2));
  v[8 + 2] = mul_mi(v[8 + 2]);
  v[12 + 2] = mul(v[12 + 2], mk(-r2, -r2));
  // b=3: W^3, W^6, W^9
  v[4 + 3] = mul(v[4 + 3], mk(s1, -c1));
  v[8 + 3] = mul(v[8 + 3], mk(-r2, -r2));
  v[12 + 3] = mul(v[12 + 3], mk(-c1, s1));
  // step 3: DFT4 over b for each c: X[c + 4d] from Y[0..3][c] = v[4c + 0..3]
  c32 t[16];
#pragma unroll
  for (int c = 0; c < 4; c++) {
    c32 a0 = v[4 * c + 0], a1 = v[4 * c + 1], a2 = v[4 * c + 2], a3 = v[4 * c + 3];
    dft4(a0, a1, a2, a3);
    t[c + 0] = a0;
    t[c + 4] = a1;
    t[c + 8] = a2;
    t[c + 12] = a3;
  }
#pragma unroll
  for (int k = 0; k < 16; k++) v[k] = t[k];
}

// Forward FFT of 16384 points held in strided distribution.  `buf` is the
// padded LDS exchange buffer (kFftBufEntries, indexed through lds_pad); the caller guarantees it is free and that
// all threads reach this call.  Output in the same distribution.
__device__ __forceinline__ void fft16k_forward(c32* v, c32* __restrict__ buf, const float2* __restrict__ oct) {
  const int j = opaque_tid();
  // radix-16 passes with Ns = 1, 16, 256
#pragma unroll
  for (int pass = 0; pass < 3; pass++) {
    const int ns = 1 << (4 * pass);
    const int k = j & (ns - 1);
    if (pass > 0) {
      // twiddle W_{16 ns}^{k r} = exp(-2 pi i k r / (16 ns)) = tw16k(k r * 16384 / (16 ns)):
      // table values for r = 1, 2, 4, 8, the others as products of at most three of them
      const int t1 = k * (kFftM / (16 * ns));
      const c32 w1 = tw16k(oct, t1), w2 = tw16k(oct, 2 * t1);
      const c32 w3 = mul(w2, w1);
      v[1] = mul(v[1], w1);
      v[2] = mul(v[2], w2);
      v[3] = mul(v[3], w3);
      {
        const c32 w4 = tw16k(oct, 4 * t1);
        v[4] = mul(v[4], w4);
        v[5] = mul(v[5], mul(w4, w1));
        v[6] = mul(v[6], mul(w4, w2));
        v[7] = mul(v[7], mul(w4, w3));
        const c32 w8 = tw16k(oct, 8 * t1);
        v[8] = mul(v[8], w8);
        v[9] = mul(v[9], mul(w8, w1));
        v[10] = mul(v[10], mul(w8, w2));
        v[11] = mul(v[11], mul(w8, w3));
        const c32 w12 = mul(w8, w4);
        v[12] = mul(v[12], w12);
        v[13] = mul(v[13], mul(w12, w1));
        v[14] = mul(v[14], mul(w12, w2));
        v[15] = mul(v[15], mul(w12, w3));
      }
    }
    dft16(v);
    // lds_pad(base + c*ns) == lds_pad(base) + c*ns + ((c*ns) >> 5) in every pass (the low
    // five bits of base and c*ns never carry into each other), so the per-c offsets
    // are compile-time constants; likewise for the strided reads.
    const int base = lds_pad((j >> (4 * pass)) * ns * 16 + k);
#pragma unroll
    for (int c = 0; c < 16; c++) buf[base + c * ns + ((c * ns) >> 5)] = v[c];
    __syncthreads();
    const int rbase = lds_pad(j);
#pragma unroll
    for (int m = 0; m < 16; m++) v[m] = buf[rbase + m * (kFftThreads + kFftThreads / 32)];
    __syncthreads();
  }
  // radix-4 pass, Ns = 4096: butterfly b = j + 1024 q on elements v[q + 4 r]; its
  // twiddles W^b and W^2b (W the 16384-th root) are the thread's W^j and W^2j rotated by
  // the constants exp(-2 pi i q/16) and exp(-2 pi i q/8)
  const c32 wj1 = tw16k(oct, j), wj2 = tw16k(oct, 2 * j);
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const c32 w1 = rot32(wj1, 2 * q), w2 = rot32(wj2, 4 * q);
    c32 a0 = v[q], a1 = mul(v[q + 4], w1), a2 = mul(v[q + 8], w2), a3 = mul(v[q + 12], mul(w2, w1));
    dft4(a0, a1, a2, a3);
    v[q] = a0;
    v[q + 4] = a1;
    v[q + 8] = a2;
    v[q + 12] = a3;
  }
}

// Unnormalised inverse FFT (exp(+i...)) via conjugation.
__device__ __forceinline__ void fft16k_inverse(c32* v, c32* __restrict__ buf, const float2* __restrict__ oct) {
#pragma unroll
  for (int m = 0; m < 16; m++) v[m].y = -v[m].y;
  fft16k_forward(v, buf, oct);
#pragma unroll
  for (int m = 0; m < 16; m++) v[m].y = -v[m].y;
}

}  // namespace vhsd
