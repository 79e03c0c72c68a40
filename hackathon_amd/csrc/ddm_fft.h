// Device FFT building blocks for gfx950 (CDNA4, wave64).
//
// Stockham autosort FFTs staged through LDS, radix 8 where possible, then 4/2
// (n = 2^LOGN), and a radix-3 pass for n = 3 * 2^m.  Natural-order input and output, unnormalised both ways, the
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

// Threads per workgroup of the DDM kernels (four wavefronts).
constexpr int NT = 256;

// Sum over aligned segments of SEG lanes (SEG a power of two <= 64).
template <int SEG, typename T> __device__ __forceinline__ T seg_sum(T v) {
#pragma unroll
  for (int o = SEG / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

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

// 3-point DFT: X0 = a0 + (a1 + a2), X1,2 = a0 - (a1 + a2)/2 -+ i (sqrt3/2)(a1 - a2) (forward; the
// inverse flips the sign of the i term)
template <bool INV> __device__ __forceinline__ void dft3(float2* a) {
  const float h = 0.86602540378443864676f;
  const float2 t1 = cadd(a[1], a[2]), t2 = csub(a[1], a[2]);
  const float2 m = make_float2(a[0].x - 0.5f * t1.x, a[0].y - 0.5f * t1.y);
  const float2 r = INV ? make_float2(-h * t2.y, h * t2.x) : make_float2(h * t2.y, -h * t2.x);
  a[0] = cadd(a[0], t1);
  a[1] = cadd(m, r);
  a[2] = csub(m, r);
}
template <> __device__ __forceinline__ void dft_small<3, false>(float2* a) { dft3<false>(a); }
template <> __device__ __forceinline__ void dft_small<3, true>(float2* a) { dft3<true>(a); }

// log2 of the radix of the next pass given REM = log2(n) - log2(Ns): radix 8
// while possible, finishing with 4x4 or 4 rather than a lone radix 2.
__host__ __device__ constexpr int next_lr(int rem) { return (rem == 4 || rem == 2) ? 2 : (rem == 1 ? 1 : 3); }
// The same with the radix capped at 2^MAXLR: MAXLR 3 is the policy above;
// MAXLR 2 takes radix-4 passes, a radix-2 one last when REM is odd (the
// autotuner's FFT radix choice).
__host__ __device__ constexpr int next_lr_cap(int rem, int maxlr) {
  return maxlr >= 3 ? next_lr(rem) : (rem >= 2 ? 2 : rem);
}

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

template <int LOGN, int B, int NT, bool INV, int NSLOG = 0, int MAXLR = 3>
__device__ __forceinline__ void lds_fft(float2* buf, const float2* tw, int tid) {
  if constexpr (NSLOG < LOGN) {
    constexpr int LR = next_lr_cap(LOGN - NSLOG, MAXLR);
    stockham_pass<LOGN, B, NT, LR, NSLOG, INV>(buf, tw, tid);
    lds_fft<LOGN, B, NT, INV, NSLOG + LR, MAXLR>(buf, tw, tid);
  }
}

// ----------------------------------------------------------------- n = 3 * 2^m
//
// The same autosort passes for transform lengths with one factor 3 (frame
// sides such as 384 = 3 * 128): a radix-3 pass first (sub-transform length
// Ns = 1, no twiddles), then radix 8/4/2 passes on sub-transform lengths 3,
// 3R, ...; the pass formulas above hold for any radix whose product is n, with
// twiddle w^(r k n/(Ns R)) from the n-point table.
template <int n, int B, int NT, int R, int Ns, bool INV>
__device__ __forceinline__ void stockham_pass_n(float2* buf, const float2* tw, int tid) {
  constexpr int NBF = B * (n / R);
  constexpr int ITEMS = (NBF + NT - 1) / NT;
  float2 a[ITEMS][R];
#pragma unroll
  for (int it = 0; it < ITEMS; it++) {
    const int g = tid + it * NT;
    if (NBF % NT == 0 || g < NBF) {
      const int f = g / (n / R), j = g % (n / R), k = j % Ns;
      const float2* src = buf + f * n + j;
#pragma unroll
      for (int r = 0; r < R; r++) a[it][r] = src[r * (n / R)];
      if (Ns > 1) {
#pragma unroll
        for (int r = 1; r < R; r++) {
          float2 w = tw[r * k * (n / (Ns * R))];
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
    const int g = tid + it * NT;
    if (NBF % NT == 0 || g < NBF) {
      const int f = g / (n / R), j = g % (n / R), k = j % Ns;
      float2* dst = buf + f * n + (j - k) * R + k;
#pragma unroll
      for (int s = 0; s < R; s++) dst[s * Ns] = a[it][s];
    }
  }
  __syncthreads();
}

__host__ __device__ constexpr int ilog2n(int n) { return n <= 1 ? 0 : 1 + ilog2n(n >> 1); }

template <int n, int B, int NT, bool INV, int Ns = 1, int MAXLR = 3>
__device__ __forceinline__ void lds_fft_n(float2* buf, const float2* tw, int tid) {
  if constexpr (Ns < n) {
    constexpr int rem = n / Ns;
    constexpr int R = rem % 3 == 0 ? 3 : (1 << next_lr_cap(ilog2n(rem), MAXLR));
    stockham_pass_n<n, B, NT, R, Ns, INV>(buf, tw, tid);
    lds_fft_n<n, B, NT, INV, Ns * R, MAXLR>(buf, tw, tid);
  }
}

// Any supported length: the power-of-two passes above, or the radix-3 form;
// radix capped at 2^MAXLR.
template <int n, int B, int NT, bool INV, int MAXLR = 3>
__device__ __forceinline__ void lds_fft_any(float2* buf, const float2* tw, int tid) {
  if constexpr ((n & (n - 1)) == 0) lds_fft<ilog2n(n), B, NT, INV, 0, MAXLR>(buf, tw, tid);
  else lds_fft_n<n, B, NT, INV, 1, MAXLR>(buf, tw, tid);
}

// ----------------------------------------------------------------- FP64 Stockham
//
// The same autosort pass structure in FP64 for the few bin-level transforms
// whose outputs are differenced (G(0) - G(tau)): radix 4, a last radix 2 when
// LOGN is odd, one transform of length 2^LOGN in LDS, NT threads.  Twiddles
// tw[m] = exp(-2 pi i m / n), m < n, come from an FP64 global table read
// through the cache.
__device__ __forceinline__ double2 dcadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 dcsub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 dcmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

template <int LOGN, int NT, bool INV, int NSLOG = 0>
__device__ __forceinline__ void lds_fft64(double2* buf, const double2* __restrict__ tw, int tid) {
  if constexpr (NSLOG < LOGN) {
    constexpr int LR = (LOGN - NSLOG) >= 2 ? 2 : 1;
    constexpr int n = 1 << LOGN, R = 1 << LR, Ns = 1 << NSLOG;
    constexpr int NBF = n / R;
    constexpr int ITEMS = (NBF + NT - 1) / NT;
    double2 a[ITEMS][R];
#pragma unroll
    for (int it = 0; it < ITEMS; it++) {
      const int j = tid + it * NT;
      if (NBF % NT == 0 || j < NBF) {
        const int k = j & (Ns - 1);
#pragma unroll
        for (int r = 0; r < R; r++) a[it][r] = buf[j + r * NBF];
        if (NSLOG > 0) {
#pragma unroll
          for (int r = 1; r < R; r++) {
            double2 w = tw[(r * k) << (LOGN - NSLOG - LR)];
            if (INV) w.y = -w.y;
            a[it][r] = dcmul(a[it][r], w);
          }
        }
        if constexpr (R == 2) {
          const double2 t = a[it][0];
          a[it][0] = dcadd(t, a[it][1]);
          a[it][1] = dcsub(t, a[it][1]);
        } else {
          const double2 t0 = dcadd(a[it][0], a[it][2]), t1 = dcsub(a[it][0], a[it][2]);
          const double2 t2 = dcadd(a[it][1], a[it][3]), d = dcsub(a[it][1], a[it][3]);
          const double2 t3 = INV ? make_double2(-d.y, d.x) : make_double2(d.y, -d.x);
          a[it][0] = dcadd(t0, t2);
          a[it][2] = dcsub(t0, t2);
          a[it][1] = dcadd(t1, t3);
          a[it][3] = dcsub(t1, t3);
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < ITEMS; it++) {
      const int j = tid + it * NT;
      if (NBF % NT == 0 || j < NBF) {
        const int k = j & (Ns - 1);
        double2* dst = buf + ((j - k) << LR) + k;
#pragma unroll
        for (int s = 0; s < R; s++) dst[s * Ns] = a[it][s];
      }
    }
    __syncthreads();
    lds_fft64<LOGN, NT, INV, NSLOG + LR>(buf, tw, tid);
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

namespace ddm {

// ----------------------------------------------------------------- wavefront FFTs
//
// One 64-lane wavefront computes one n-point FFT (n = 2^LOGN, n >= 256).  Lane
// l holds P = n/64 points in the "canonical" distribution v[q] = x[l + 64 q],
// which is both what coalesced global loads produce and what every Stockham
// pass reads (butterfly j = l + 64 b reads x[j + r n/R] = v[b + r n/(64 R)]).
// Between passes the outputs are permuted back to canonical through a
// wave-private LDS buffer; no workgroup barrier is involved.  The last pass
// writes X[j + s n/R], which is canonical again, straight into registers.

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Padded LDS index for the exchange buffers: one spare slot per 16 entries
// spreads the strided Stockham writes over the banks.
__device__ __forceinline__ int xpad(int i) { return i + (i >> 4); }
__host__ __device__ constexpr int xch_size(int n) { return n + n / 16; }

// First-pass DFTs whose upper half of inputs is zero (zero-padded data).
template <int R, bool INV> __device__ __forceinline__ void dft_half(float2* a);
template <> __device__ __forceinline__ void dft_half<4, false>(float2* a) {
  float2 x0 = a[0], x1 = a[1];
  a[0] = cadd(x0, x1);
  a[2] = csub(x0, x1);
  float2 t = mul_mi<false>(x1);
  a[1] = cadd(x0, t);
  a[3] = csub(x0, t);
}
template <> __device__ __forceinline__ void dft_half<4, true>(float2* a) {
  float2 x0 = a[0], x1 = a[1];
  a[0] = cadd(x0, x1);
  a[2] = csub(x0, x1);
  float2 t = mul_mi<true>(x1);
  a[1] = cadd(x0, t);
  a[3] = csub(x0, t);
}
template <bool INV> __device__ __forceinline__ void dft8_half(float2* a) {
  // inputs a[0..3], a[4..7] == 0: E = DFT4(a0, a2, 0, 0), O = DFT4(a1, a3, 0, 0)
  float2 a0 = a[0], a1 = a[1], a2 = a[2], a3 = a[3];
  float2 e0 = cadd(a0, a2), e2 = csub(a0, a2), t = mul_mi<INV>(a2);
  float2 e1 = cadd(a0, t), e3 = csub(a0, t);
  float2 o0 = cadd(a1, a3), o2 = csub(a1, a3), u = mul_mi<INV>(a3);
  float2 o1 = cadd(a1, u), o3 = csub(a1, u);
  const float r = 0.70710678118654752440f;
  float2 w1o1 = INV ? make_float2(r * (o1.x - o1.y), r * (o1.x + o1.y)) : make_float2(r * (o1.x + o1.y), r * (o1.y - o1.x));
  float2 w2o2 = mul_mi<INV>(o2);
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
template <> __device__ __forceinline__ void dft_half<8, false>(float2* a) { dft8_half<false>(a); }
template <> __device__ __forceinline__ void dft_half<8, true>(float2* a) { dft8_half<true>(a); }

__host__ __device__ constexpr int ilog2c(int n) { return n <= 1 ? 0 : 1 + ilog2c(n >> 1); }
// radix (log2) of a wavefront pass: the usual choice, capped so every lane owns
// at least one butterfly (R <= n/64)
__host__ __device__ constexpr int wave_lr(int logn, int nslog, int maxlr = 3) {
  return next_lr_cap(logn - nslog, maxlr) < (logn - 6) ? next_lr_cap(logn - nslog, maxlr) : (logn - 6);
}

// Twiddle-table slot of entry m in an LDS-staged table: one spare slot per 16
// entries.  The radix-8 passes read entries k 2^s and 4 k 2^s for k = lane mod
// Ns; in a plain table those land on a few bank pairs (up to 8-way conflicts),
// padded they spread over all 16 pairs of ds_read_b64.  The slot is an
// affine function of k within each pass, so the loads keep immediate offsets.
__host__ __device__ constexpr int twpad(int m) { return m + (m >> 4); }

// Twiddles w^r, r = 1..R-1, of one butterfly.  TWP (an LDS table stored at
// twpad() slots): every w^r is read from the table, an exact f32 value (the
// product forms below carry 1-3 ulp, which measurably doubled the error of
// the temporal transforms' longest FFT lags: 8192 points, max 1.11e-5 ->
// 8.6e-6 relative, at no cost in time).  A table in global memory: w^1 and
// w^4 are read, the others are at most two products away from them.
template <int R, bool INV, bool TWP = false>
__device__ __forceinline__ void butterfly_twiddles(const float2* tw, int m1, float2* w) {
  if constexpr (TWP) {
#pragma unroll
    for (int r = 1; r < R; r++) {
      float2 x = tw[twpad(r * m1)];
      if (INV) x.y = -x.y;
      w[r] = x;
    }
    return;
  }
  float2 w1 = tw[m1];
  if (INV) w1.y = -w1.y;
  w[1] = w1;
  if (R >= 4) {
    w[2] = cmul(w1, w1);
    w[3] = cmul(w[2], w1);
  }
  if (R >= 8) {
    float2 w4 = tw[4 * m1];
    if (INV) w4.y = -w4.y;
    w[4] = w4;
    w[5] = cmul(w4, w1);
    w[6] = cmul(w4, w[2]);
    w[7] = cmul(w4, w[3]);
  }
}

template <int LOGN, int NSLOG, bool INV, bool HALF, bool TWP = false, bool WL = false, int MAXLR = 3,
          bool W2 = false>
__device__ __forceinline__ void wave_pass(float2 (&v)[(1 << LOGN) / 64], float2* xch, const float2* tw, int lane,
                                          const float2* wl = nullptr, const float2* w2 = nullptr) {
  constexpr int n = 1 << LOGN;
  constexpr int LR = wave_lr(LOGN, NSLOG, MAXLR);
  constexpr int R = 1 << LR;
  constexpr int Ns = 1 << NSLOG;
  constexpr int NBL = n / (64 * R);          // butterflies per lane
  constexpr int QS = n / (64 * R);           // canonical index step between a butterfly's inputs
  constexpr bool LAST = (NSLOG + LR == LOGN);
  // No scheduling barrier here: the next pass's twiddle loads and products
  // may be issued while this pass's exchange is in flight (measured: the
  // 2048-point temporal transform 3.5 % faster, no VGPR growth).
  if constexpr (!LAST) wave_sync();          // the previous pass's reads of xch are done
#pragma unroll
  for (int b = 0; b < NBL; b++) {
    const int j = lane + 64 * b;
    const int k = j & (Ns - 1);
    float2 a[R];
#pragma unroll
    for (int r = 0; r < R; r++) a[r] = v[b + r * QS];
    if (NSLOG > 0) {
      float2 w[R];
      if constexpr (WL && LAST && R == 8 && NBL == 1) {
#pragma unroll
        for (int r = 1; r < R; r++) w[r] = wl[r];
      } else if constexpr (W2 && NSLOG == 3 && R == 8 && NBL == 1) {
#pragma unroll
        for (int r = 1; r < R; r++) w[r] = w2[r];
      } else {
        butterfly_twiddles<R, INV, TWP>(tw, k << (LOGN - NSLOG - LR), w);
      }
#pragma unroll
      for (int r = 1; r < R; r++) a[r] = cmul(a[r], w[r]);
    }
    if (HALF && NSLOG == 0 && R >= 4) {
      dft_half<R, INV>(a);
    } else {
      dft_small<R, INV>(a);
    }
    if constexpr (LAST) {
#pragma unroll
      for (int s = 0; s < R; s++) v[b + s * QS] = a[s];
    } else {
      const int d0 = ((j - k) << LR) + k;
#pragma unroll
      for (int s = 0; s < R; s++) xch[xpad(d0 + s * Ns)] = a[s];
    }
  }
  if constexpr (!LAST) {
    wave_sync();
#pragma unroll
    for (int q = 0; q < n / 64; q++) v[q] = xch[xpad(lane + 64 * q)];
    wave_pass<LOGN, NSLOG + LR, INV, false, TWP, WL, MAXLR, W2>(v, xch, tw, lane, wl, w2);
  }
}

// In-place canonical-distribution FFT of one wavefront.  HALF: inputs with
// index >= n/2 are known to be zero (v[q] for q >= n/128 are ignored).  TWP:
// tw is an LDS table stored at twpad() slots.
template <int LOGN, bool INV, bool HALF, bool TWP = false, int MAXLR = 3>
__device__ __forceinline__ void wave_fft(float2 (&v)[(1 << LOGN) / 64], float2* xch, const float2* tw, int lane) {
  wave_pass<LOGN, 0, INV, HALF, TWP, false, MAXLR>(v, xch, tw, lane);
}
// The same with the last pass's twiddles supplied by the caller (wl[1..7]; a
// radix-8 last pass with one butterfly per lane, k = lane): a caller that runs
// many transforms keeps them in registers instead of rebuilding them each time.
template <int LOGN, bool INV, bool HALF, bool TWP>
__device__ __forceinline__ void wave_fft_wl(float2 (&v)[(1 << LOGN) / 64], float2* xch, const float2* tw, int lane,
                                            const float2* wl) {
  wave_pass<LOGN, 0, INV, HALF, TWP, true>(v, xch, tw, lane, wl);
}
// The same with the second pass's twiddles supplied too (w2[1..7]; a 512-point
// transform's middle radix-8 pass, k = lane mod 8).
template <int LOGN, bool INV, bool HALF, bool TWP>
__device__ __forceinline__ void wave_fft_wl2(float2 (&v)[(1 << LOGN) / 64], float2* xch, const float2* tw, int lane,
                                             const float2* w2, const float2* wl) {
  wave_pass<LOGN, 0, INV, HALF, TWP, true, 3, true>(v, xch, tw, lane, wl, w2);
}

// 2Sum accumulate (branch-free): (s, c) += x, value s + c.
__device__ __forceinline__ void two_sum_add(float& s, float& c, float x) {
  float t = s + x;
  float z = t - s;
  c += (s - (t - z)) + (x - z);
  s = t;
}

// Cross-lane sums on the DPP lane network (no LDS traffic, no per-step
// address arithmetic).  dpp<CTRL, ROWS>(v): v of the lane CTRL selects, in
// the rows of 16 lanes that ROWS enables (0 elsewhere).  Controls used:
// quad_perm [1,0,3,2] = 0xb1 and [2,3,0,1] = 0x4e, row_half_mirror = 0x141,
// row_mirror = 0x140, row_bcast15 = 0x142, row_bcast31 = 0x143.
template <int CTRL, int ROWS = 0xf>
__device__ __forceinline__ int dpp(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, ROWS, 0xf, false);
}
template <int CTRL, int ROWS = 0xf>
__device__ __forceinline__ float dpp(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWS, 0xf, false));
}

// Sum over aligned segments of SEG lanes (SEG in 1, 2, 4, 8, 16); every lane
// of a segment gets its sum: xor 1, xor 2 within quads, then the half-row and
// row mirrors pair the quads and the half-rows.
template <int SEG, typename T> __device__ __forceinline__ T seg_sum_dpp(T v) {
  static_assert(SEG == 1 || SEG == 2 || SEG == 4 || SEG == 8 || SEG == 16, "segment");
  if constexpr (SEG >= 2) v += dpp<0xb1>(v);
  if constexpr (SEG >= 4) v += dpp<0x4e>(v);
  if constexpr (SEG >= 8) v += dpp<0x141>(v);
  if constexpr (SEG >= 16) v += dpp<0x140>(v);
  return v;
}

// Wavefront total, returned to every lane (a wave-uniform value): row sums,
// then row 1 += row 0 and row 3 += row 2 (bcast15), rows 2, 3 += lane 31
// (bcast31), so lane 63 holds the total.
template <typename T> __device__ __forceinline__ T wave_total(T v) {
  v = seg_sum_dpp<16>(v);
  v += dpp<0x142, 0xa>(v);
  v += dpp<0x143, 0xc>(v);
  if constexpr (sizeof(T) == 4 && !__is_same(T, float)) {
    return (T)__builtin_amdgcn_readlane((int)v, 63);
  } else {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
  }
}

// Wavefront sum of a float (all lanes get the result).
__device__ __forceinline__ float wave_sum(float v) { return wave_total(v); }

}  // namespace ddm
