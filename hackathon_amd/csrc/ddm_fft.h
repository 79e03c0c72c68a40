// Device FFT building blocks for gfx950 (CDNA4, wave64).
//
// Stockham autosort FFTs staged through LDS, radix 8 where possible, then 4/2
// (n = 2^LOGN).  Natural-order input and output, unnormalised both ways, the
// same sign and twiddle convention as the reference's radix-2 Stockham
// (motility-ddm src/fft.js:59-93, src/ddm.wgsl:78-101): forward uses
// exp(-2 pi i m / n), inverse its conjugate.  Twiddles come from a host FP64
// table rounded once to f32 (no device sin/cos), as src/ddm.wgsl:10-13 does.
//
// Stockham pass with radix R after a sub-transform length Ns:
//   for butterfly j in [0, n/R):  k = j mod Ns
//     a[r] = x[j + r n/R] * w^(r k n/(Ns R))      r = 0..R-1
//     X[(j-k) R + k + s Ns] = sum_r a[r] exp(-+2 pi i r s / R)
#pragma once
#include <hip/hip_runtime.h>

namespace ddm {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
// multiply by -i (forward) or +i (inverse)
template <bool INV> __device__ __forceinline__ float2 mul_mi(float2 a) {
  return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

template <int R, bool INV> __device__ __forceinline__ void dft_small(float2* a);

template <> __device__ __forceinline__ void dft_small<2, false>(float2* a) {
  float2 t = a[0];
  a[0] = cadd(t, a[1]);
  a[1] = csub(t, a[1]);
}
template <> __device__ __forceinline__ void dft_small<2, true>(float2* a) { dft_small<2, false>(a); }

template <bool INV> __device__ __forceinline__ void dft4(float2& a0, float2& a1, float2& a2, float2& a3) {
  float2 t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3), t3 = mul_mi<INV>(csub(a1, a3));
  a0 = cadd(t0, t2);
  a2 = csub(t0, t2);
  a1 = cadd(t1, t3);
  a3 = csub(t1, t3);
}
template <> __device__ __forceinline__ void dft_small<4, false>(float2* a) { dft4<false>(a[0], a[1], a[2], a[3]); }
template <> __device__ __forceinline__ void dft_small<4, true>(float2* a) { dft4<true>(a[0], a[1], a[2], a[3]); }

template <bool INV> __device__ __forceinline__ void dft8(float2* a) {
  // even/odd split: two 4-point DFTs, then W8^s combine
  float2 e0 = a[0], e1 = a[2], e2 = a[4], e3 = a[6];
  float2 o0 = a[1], o1 = a[3], o2 = a[5], o3 = a[7];
  dft4<INV>(e0, e1, e2, e3);
  dft4<INV>(o0, o1, o2, o3);
  const float r = 0.70710678118654752440f;
  // W8^1 = (1 - i)/sqrt2 forward, (1 + i)/sqrt2 inverse
  float2 w1o1 = INV ? make_float2(r * (o1.x - o1.y), r * (o1.x + o1.y)) : make_float2(r * (o1.x + o1.y), r * (o1.y - o1.x));
  float2 w2o2 = mul_mi<INV>(o2);
  // W8^3 = (-1 - i)/sqrt2 forward, (-1 + i)/sqrt2 inverse
  float2 w3o3 = INV ? make_float2(-r * (o3.x + o3.y), r * (o3.x - o3.y)) : make_float2(r * (o3.y - o3.x), -r * (o3.x + o3.y));
  a[0] = cadd(e0, o0);
  a[4] = csub(e0, o0);
  a[1] = cadd(e1, w1o1);
  a[5] = csub(e1, w1o1);
  a[2] = cadd(e2, w2o2);
  a[6] = csub(e2, w2o2);
  a[3] = cadd(e3, w3o3);
  a[7] = csub(e3, w3o3);
}
template <> __device__ __forceinline__ void dft_small<8, false>(float2* a) { dft8<false>(a); }
template <> __device__ __forceinline__ void dft_small<8, true>(float2* a) { dft8<true>(a); }

// log2 of the radix of the next pass given REM = log2(n) - log2(Ns): radix 8
// while possible, finishing with 4x4 or 4 rather than a lone radix 2.
__host__ __device__ constexpr int next_lr(int rem) { return (rem == 4 || rem == 2) ? 2 : (rem == 1 ? 1 : 3); }

// One in-place Stockham pass over B contiguous transforms of length 2^LOGN in
// LDS `buf`, executed by NT threads.  `tw[m] = exp(-2 pi i m / n)`, m < n.
// All threads read their butterflies' inputs, barrier, then write.
template <int LOGN, int B, int NT, int LR, int NSLOG, bool INV>
__device__ __forceinline__ void stockham_pass(float2* buf, const float2* tw, int tid) {
  constexpr int n = 1 << LOGN;
  constexpr int R = 1 << LR;
  constexpr int Ns = 1 << NSLOG;
  constexpr int NBF = B * (n / R);                 // butterflies in the pass
  constexpr int ITEMS = (NBF + NT - 1) / NT;       // per thread
  float2 a[ITEMS][R];
#pragma unroll
  for (int it = 0; it < ITEMS; it++) {
    int g = tid + it * NT;
    if (NBF % NT == 0 || g < NBF) {
      int f = g >> (LOGN - LR);
      int j = g & ((n / R) - 1);
      int k = j & (Ns - 1);
      const float2* src = buf + f * n + j;
#pragma unroll
      for (int r = 0; r < R; r++) a[it][r] = src[r * (n / R)];
      if (NSLOG > 0) {
#pragma unroll
        for (int r = 1; r < R; r++) {
          float2 w = tw[(r * k) << (LOGN - NSLOG - LR)];
          if (INV) w.y = -w.y;
          a[it][r] = cmul(a[it][r], w);
        }
      }
      dft_small<R, INV>(a[it]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < ITEMS; it++) {
    int g = tid + it * NT;
    if (NBF % NT == 0 || g < NBF) {
      int f = g >> (LOGN - LR);
      int j = g & ((n / R) - 1);
      int k = j & (Ns - 1);
      float2* dst = buf + f * n + ((j - k) << LR) + k;
#pragma unroll
      for (int s = 0; s < R; s++) dst[s * Ns] = a[it][s];
    }
  }
  __syncthreads();
}

template <int LOGN, int B, int NT, bool INV, int NSLOG = 0>
__device__ __forceinline__ void lds_fft(float2* buf, const float2* tw, int tid) {
  if constexpr (NSLOG < LOGN) {
    constexpr int LR = next_lr(LOGN - NSLOG);
    stockham_pass<LOGN, B, NT, LR, NSLOG, INV>(buf, tw, tid);
    lds_fft<LOGN, B, NT, INV, NSLOG + LR>(buf, tw, tid);
  }
}

// Error-free transformation accumulate: (s, c) += x with Neumaier's
// compensation; the value is s + c.
__device__ __forceinline__ void comp_add(float& s, float& c, float x) {
  float t = s + x;
  float e = (fabsf(s) >= fabsf(x)) ? ((s - t) + x) : ((x - t) + s);
  s = t;
  c += e;
}

}  // namespace ddm
