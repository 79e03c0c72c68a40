// Workgroup-cooperative temporal kernel of the DDM hot path for gfx950, for
// zero-padded transform lengths n = 2^LOGT >= 2048 (series of T <= n/2
// frames).
//
//  k_temporal_c  one work item = up to 64 modes of ONE radial bin, as for
//                k_temporal / k_temporal_w, with the same outputs: per item
//                the |Y|^2 sums (r-major, see below), the p(t) sums, P, and
//                the direct sums of the last KT lags.
//
// The four waves of a workgroup work on one mode at a time.  The n-point
// transform of a series that is zero for t >= n/2 splits exactly into
// m = n/512 transforms of 512 points (first decimation-in-frequency stage):
//   Y(m k + r) = FFT_512[u_r](k),
//   u_r(t1) = w_n^{r t1} sum_{t2 < m/2} y(t1 + 512 t2) w_m^{r t2}.
// Thread tid owns the two columns t1 = tid and tid + 256: it loads exactly the
// samples y(t1 + 512 t2) of both (t = tid + 256 i, i < m), so the m-point
// first stage (the "fold") runs in its registers; the folded rows u_r go to
// LDS, where wave w transforms rows r = w, w + 4, ... in place (each row is
// that wave's exchange buffer) and adds |Y|^2 into registers.  The waves meet
// at two barriers per mode (the masked mean; the folded rows), and no |Y|^2
// accumulator lives in LDS, so the workgroup needs only the m folded rows:
// 17 KB at n = 2048, 70 KB at n = 8192 (two workgroups per CU).
//
// Reference: src/ddm.wgsl (temporal :202-259) and its FP64 twin
// src/ddm_ref.js:229-288.
#include "ddm_internal.h"
#include "ddm_fft.h"

namespace ddm {

typedef float v2f __attribute__((ext_vector_type(2)));

// 16-point forward DFT of inputs a[0..7] (a[8..15] zero): even outputs are the
// 8-point DFT of a, odd outputs the 8-point DFT of a[t] w16^t.
__device__ __forceinline__ void dft16_half(float2* a) {
  float2 b[8];
  // w16^t, t = 0..7: cos/sin of t pi / 8
  const float c1 = 0.92387953251128675613f, s1 = 0.38268343236508977173f, h = 0.70710678118654752440f;
  b[0] = a[0];
  b[1] = cmul(a[1], make_float2(c1, -s1));
  b[2] = make_float2(h * (a[2].x + a[2].y), h * (a[2].y - a[2].x));
  b[3] = cmul(a[3], make_float2(s1, -c1));
  b[4] = make_float2(a[4].y, -a[4].x);
  b[5] = cmul(a[5], make_float2(-s1, -c1));
  b[6] = make_float2(h * (a[6].y - a[6].x), -h * (a[6].x + a[6].y));
  b[7] = cmul(a[7], make_float2(-c1, -s1));
  float2 e[8];
#pragma unroll
  for (int i = 0; i < 8; i++) e[i] = a[i];
  dft8<false>(e);
  dft8<false>(b);
#pragma unroll
  for (int s = 0; s < 8; s++) {
    a[2 * s] = e[s];
    a[2 * s + 1] = b[s];
  }
}

// m-point forward DFT with the upper half of the inputs zero (a[0 .. m/2 - 1]
// given), natural-order outputs a[0 .. m - 1].
template <int M> __device__ __forceinline__ void fold_dft(float2* a);
template <> __device__ __forceinline__ void fold_dft<4>(float2* a) { dft_half<4, false>(a); }
template <> __device__ __forceinline__ void fold_dft<8>(float2* a) { dft_half<8, false>(a); }
template <> __device__ __forceinline__ void fold_dft<16>(float2* a) { dft16_half(a); }

// LDS layout of k_temporal_c, in float2 slots
template <int LOGT> struct TcLayout {
  static constexpr int m = 1 << (LOGT - 9);
  static constexpr int RS = xch_size(512);          // one folded row, xpad-indexed (the wave's exchange buffer)
  static constexpr int U = 0;                       // m rows
  static constexpr int TW = m * RS;                 // 512-point twiddle table at twpad() slots
  static constexpr int TAIL = TW + twpad(512);      // 2 x 64 (y, mask) as float4: head, tail
  static constexpr int RED = TAIL + 256;            // 2 parities x 4 waves x (sx, sy); then P
  static constexpr int ORD = RED + 8 + 4;           // m <= 8: the frame order (n/2 entries) of a gathered series
  static constexpr int SIZE = ORD + (m <= 8 ? (512 * m) / 4 : 0);
};

// Modes summed plainly per wave before the sums are folded into a second
// register set (two-level summation: every term added at most 16 + 4 times).
constexpr int TC_GROUP = 16;

// Pre-twiddle w_n^{r t1} of one column t1 from the exact table values
// b = (w_n^{t1}, w_n^{2 t1}, w_n^{3 t1}, w_n^{4 t1}, w_n^{8 t1}): at most two
// products, r = 8 r2 + 4 r1 + r0.
template <int R>
__device__ __forceinline__ float2 column_twiddle(const float2 (&b)[5]) {
  static_assert(R >= 1 && R < 16, "r");
  constexpr int r0 = R & 3, r1 = (R >> 2) & 1, r2 = R >> 3;
  float2 w = r0 ? b[r0 - 1] : (r1 ? b[3] : b[4]);
  if (r0 && r1) w = cmul(w, b[3]);
  if ((r0 || r1) && r2) w = cmul(w, b[4]);
  return w;
}

template <int M, int R = 1>
__device__ __forceinline__ void store_folded(float2* U, int rs, int t1x, const float2* a, const float2 (&b)[5]) {
  if constexpr (R < M) {
    U[R * rs + t1x] = cmul(a[R], column_twiddle<R>(b));
    store_folded<M, R + 1>(U, rs, t1x, a, b);
  }
}

// DENSE: every sample t < n/2 is a valid frame (T = n/2, a full mask), known
// at the launch: the per-sample mask tests and selects drop out.
template <int LOGT, int MAXLR, bool TWO_LEVEL, bool DENSE = false>
__device__ __forceinline__ void temporal_c_item(const TemporalArgs& A, const int item, float2* lds,
                                                const float2 (&twb)[2][5], const float2 (&wlast)[8],
                                                const float2 (&wmid)[8]) {
  using L = TcLayout<LOGT>;
  constexpr int m = L::m;          // sub-transforms (and samples per thread)
  constexpr int MH = m / 2;        // nonzero rows of the fold
  constexpr int SW = m / 4;        // sub-transforms per wave
  static_assert(m >= 4 && m <= 16, "n = 2048 .. 8192");
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const float2* __restrict__ hist = A.hist;
  const int T = A.T, hs = A.hs, KT = A.KT;
  const float inv_n0 = A.inv_n0;
  const int4 it = A.items[item];
  const int start = it.y, count = it.z;
  float2* U = lds + L::U;
  const float2* twl = lds + L::TW;
  float4* head = reinterpret_cast<float4*>(lds + L::TAIL);   // (y, mask, 0) of t < KT
  float4* tailv = head + 64;                                  // of T - KT <= t < T
  float* red = reinterpret_cast<float*>(lds + L::RED);

  // a gathered series reads its slots through the order: staged in LDS where
  // it fits (m <= 8), else re-read from the cache with each mode
  const bool identity = A.identity != 0;
  uint32_t* ordl = reinterpret_cast<uint32_t*>(lds + L::ORD);
  if constexpr (m <= 8) {
    if (!identity)
      for (int i = tid; i < 256 * m; i += NT) ordl[i] = A.order[i];
  }
  // sample i of this thread: t = tid + 256 i = t1 + 512 t2 with t1 = tid + 256 (i & 1), t2 = i >> 1
  uint32_t mbits = (1u << m) - 1u;
  bool dense = true;
  if constexpr (DENSE) {
    __syncthreads();   // the order's staging
  } else {
    mbits = 0;
#pragma unroll
    for (int i = 0; i < m; i++) {
      const int t = tid + 256 * i;
      if (t < T && A.mask[t] != 0.f) mbits |= 1u << i;
    }
    // (also the barrier behind the order's staging)
    dense = __syncthreads_and(mbits == ((1u << m) - 1u));
  }

  float yacc[SW][8], yacc2[TWO_LEVEL ? SW : 1][8];
  float pacc[m], pacc2[TWO_LEVEL ? m : 1];
#pragma unroll
  for (int s = 0; s < SW; s++)
#pragma unroll
    for (int q = 0; q < 8; q++) yacc[s][q] = 0.f;
#pragma unroll
  for (int i = 0; i < m; i++) pacc[i] = 0.f;
  if constexpr (TWO_LEVEL) {
#pragma unroll
    for (int s = 0; s < SW; s++)
#pragma unroll
      for (int q = 0; q < 8; q++) yacc2[s][q] = 0.f;
#pragma unroll
    for (int i = 0; i < m; i++) pacc2[i] = 0.f;
  }
  float aP = 0.f, cP = 0.f;
  // tail lags T - KT + j (KT <= 64): a group of G threads per lag for
  // KT <= 32; above, lag j and lag KT - 1 - j share a group (KT + 1 pairs
  // between them, so the groups are balanced and the pass takes ~2/3 of the
  // one-lag groups' iterations; at KT <= 32 the one-lag groups do less work
  // in all, measured); thread G gi + s sums the pairs t = s, s + G, ...
  const bool paired = KT > 32;
  const int ngrp = paired ? (KT + 1) / 2 : KT;
  const int G = (paired || KT > 16) ? 8 : 16;
  const int gi = tid / G, gs = tid % G;
  const int ja = gi, jb = paired ? KT - 1 - gi : gi;   // jb == ja: one lag (the middle one of an odd KT)
  float Gt = 0.f, Gtc = 0.f, Gt2 = 0.f, Gtc2 = 0.f;

  // series loads without guards: t < n/2 stays inside the history (rows of
  // hs >= T slots, the last one followed by n/2 - hs spare samples; samples
  // t >= T are masked off); a non-identity order gathers sample t from slot
  // order[t] (order's entries past T are 0)
  // (one 8-byte non-temporal load per sample, a wave-uniform base plus a
  // 32-bit per-sample offset; the order is re-read with each mode, from LDS
  // or the cache: m more registers held across the loop would spill at
  // n = 8192)
  auto load_series = [&](float2 (&dst)[m], int slot) {
    const char* x = reinterpret_cast<const char*>(hist + (long)A.bin_modes[start + slot] * hs);
    const uint32_t* op = A.order;
    asm volatile("" : "+s"(op));
#pragma unroll
    for (int i = 0; i < m; i++) {
      const uint32_t off =
          8u * (identity ? (uint32_t)(tid + 256 * i) : (m <= 8 ? ordl[tid + 256 * i] : op[tid + 256 * i]));
      const v2f v = __builtin_nontemporal_load(reinterpret_cast<const v2f*>(x + off));
      dst[i] = make_float2(v.x, v.y);
    }
  };
  // the next mode's series is loaded once this mode's samples are folded
  // (during its sub-transforms)
  float2 nx[m];
  load_series(nx, 0);
  for (int mi = 0; mi < count; mi++) {
    float2 y[m];
#pragma unroll
    for (int i = 0; i < m; i++) y[i] = nx[i];
    // the mask bits (and at m = 16 the pre-twiddle bases) are re-read from
    // opaque copies each mode: hoisted out of the loop, their per-sample
    // selects and the 2 (m - 1) pre-twiddles would stay live in registers for
    // the whole item
    uint32_t mb = mbits;
    if constexpr (!DENSE) asm volatile("" : "+v"(mb));
    // masked sums: the mean, and sum |y|^2 for P
    float sx = 0.f, sy = 0.f, ps = 0.f;
    if (!dense) {
#pragma unroll
      for (int i = 0; i < m; i++) {
        if (!((mb >> i) & 1u)) y[i] = make_float2(0.f, 0.f);
      }
    }
#pragma unroll
    for (int i = 0; i < m; i++) {
      sx += y[i].x;
      sy += y[i].y;
      ps += y[i].x * y[i].x + y[i].y * y[i].y;
    }
    two_sum_add(aP, cP, ps);
    sx = wave_sum(sx);
    sy = wave_sum(sy);
    float* rp = red + 8 * (mi & 1);
    if (lane == 0) {
      rp[2 * wave] = sx;
      rp[2 * wave + 1] = sy;
    }
    // the sums are visible; every wave is done with the previous mode's rows
    // and tail samples
    __syncthreads();
    const float mx = ((rp[0] + rp[2]) + (rp[4] + rp[6])) * inv_n0;
    const float my = ((rp[1] + rp[3]) + (rp[5] + rp[7])) * inv_n0;
#pragma unroll
    for (int i = 0; i < m; i++) {
      y[i] = make_float2(y[i].x - mx, y[i].y - my);
      if (!dense && !((mb >> i) & 1u)) y[i] = make_float2(0.f, 0.f);
      pacc[i] += y[i].x * y[i].x + y[i].y * y[i].y;
    }
    // samples of the direct tail: y(t < KT) (t = tid, i = 0: KT <= 64) and
    // y(T - KT <= t < T) (i in [i0, i1], at most two values), with masks
    // (tv: an opaque copy of tid, so that the per-sample tail tests and
    // addresses are formed here, for the one or two i taken, not hoisted)
    int tv = tid;
    asm volatile("" : "+v"(tv));
    if (tv < KT) head[tv] = make_float4(y[0].x, y[0].y, (mb & 1u) ? 1.f : 0.f, 0.f);
    {
      const int i0 = (T - KT) >> 8, i1 = (T - 1) >> 8;
#pragma unroll
      for (int i = 0; i < m; i++) {
        if (i >= i0 && i <= i1) {
          const int t = tv + 256 * i;
          if (t >= T - KT && t < T) tailv[t - (T - KT)] = make_float4(y[i].x, y[i].y, ((mb >> i) & 1u) ? 1.f : 0.f, 0.f);
        }
      }
    }
    // the fold of both columns, pre-twiddled, into the rows
#pragma unroll
    for (int c = 0; c < 2; c++) {
      __builtin_amdgcn_sched_barrier(0);   // one column at a time (register pressure at m = 16)
      float2 a[m];
#pragma unroll
      for (int t2 = 0; t2 < MH; t2++) a[t2] = y[2 * t2 + c];
      fold_dft<m>(a);
      float2 b[5];
#pragma unroll
      for (int j = 0; j < 5; j++) {
        b[j] = twb[c][j];
        if constexpr (m == 16) asm volatile("" : "+v"(b[j].x), "+v"(b[j].y));
      }
      const int t1x = xpad(tid + 256 * c);
      U[t1x] = a[0];
      store_folded<m>(U, L::RS, t1x, a, b);
    }
    __syncthreads();
    if (mi + 1 < count) load_series(nx, mi + 1);
    // direct tail sums, difference form sum m(t) m(t + tau) |y(t + tau) - y(t)|^2
    {
      float g = 0.f, g2 = 0.f;
      if (gi < ngrp) {
        for (int t = gs; t < KT - ja; t += G) {
          const float4 a = head[t], b = tailv[t + ja];
          const float dx = b.x - a.x, dy = b.y - a.y;
          g += (a.z * b.z) * (dx * dx + dy * dy);
        }
        if (jb > ja)
          for (int t = gs; t < KT - jb; t += G) {
            const float4 a = head[t], b = tailv[t + jb];
            const float dx = b.x - a.x, dy = b.y - a.y;
            g2 += (a.z * b.z) * (dx * dx + dy * dy);
          }
      }
      if (G == 16) {
        g = seg_sum_dpp<16>(g);
        g2 = seg_sum_dpp<16>(g2);
      } else {
        g = seg_sum_dpp<8>(g);
        g2 = seg_sum_dpp<8>(g2);
      }
      if (gs == 0 && gi < ngrp) {
        two_sum_add(Gt, Gtc, g);
        two_sum_add(Gt2, Gtc2, g2);
      }
    }
    // the wave's sub-transforms r = wave + 4 s
#pragma unroll
    for (int s = 0; s < SW; s++) {
      __builtin_amdgcn_sched_barrier(0);   // one sub-transform at a time
      float2* row = U + (wave + 4 * s) * L::RS;
      float2 v[8];
#pragma unroll
      for (int q = 0; q < 8; q++) v[q] = row[xpad(lane + 64 * q)];
      // (at m = 16 the middle pass's twiddles are formed per mode: 14 more
      // registers held across the loop would spill there)
      if constexpr (MAXLR == 3 && m <= 8) wave_fft_wl2<9, false, false, true>(v, row, twl, lane, wmid, wlast);
      else if constexpr (MAXLR == 3) wave_fft_wl<9, false, false, true>(v, row, twl, lane, wlast);
      else wave_fft<9, false, false, true, MAXLR>(v, row, twl, lane);
#pragma unroll
      for (int q = 0; q < 8; q++) yacc[s][q] += v[q].x * v[q].x + v[q].y * v[q].y;
    }
    if constexpr (TWO_LEVEL) {
      if ((mi % TC_GROUP) == TC_GROUP - 1) {
#pragma unroll
        for (int s = 0; s < SW; s++)
#pragma unroll
          for (int q = 0; q < 8; q++) {
            yacc2[s][q] += yacc[s][q];
            yacc[s][q] = 0.f;
          }
#pragma unroll
        for (int i = 0; i < m; i++) {
          pacc2[i] += pacc[i];
          pacc[i] = 0.f;
        }
      }
    }
  }
  // outputs: |Y|^2 sums r-major (index r 512 + j holds k = m j + r), p(t), P, tail lags
  float* Sp = A.S_part + (long)item * (512 * m);
#pragma unroll
  for (int s = 0; s < SW; s++)
#pragma unroll
    for (int q = 0; q < 8; q++)
      Sp[(wave + 4 * s) * 512 + lane + 64 * q] = TWO_LEVEL ? yacc2[s][q] + yacc[s][q] : yacc[s][q];
  float* pp = A.p_part + (long)item * T;
#pragma unroll
  for (int i = 0; i < m; i++) {
    const int t = tid + 256 * i;
    if (t < T) pp[t] = TWO_LEVEL ? pacc2[i] + pacc[i] : pacc[i];
  }
  const float pw = wave_sum(aP + cP);
  __syncthreads();   // the last mode's reads of the rows, tail samples and sums are done
  if (lane == 0) red[16 + wave] = pw;
  __syncthreads();
  if (tid == 0) A.P_part[item] = (red[16] + red[17]) + (red[18] + red[19]);
  if (gs == 0 && gi < ngrp) {
    A.G_part[(long)item * KT + ja] = Gt + Gtc;
    if (jb > ja) A.G_part[(long)item * KT + jb] = Gt2 + Gtc2;
  }
}

// Per-workgroup set-up: the 512-point twiddle table in LDS, the last pass's
// twiddles, and the bases w_n^{j t1} (j = 1, 2, 3, 4, 8) of this thread's two
// columns' pre-twiddles (the same for every mode; exact f32 table values).
template <int LOGT, int MAXLR>
__device__ __forceinline__ void temporal_c_setup(float2* lds, const float2* __restrict__ tw_g,
                                                 const float2* __restrict__ tw_f, float2 (&twb)[2][5],
                                                 float2 (&wlast)[8], float2 (&wmid)[8]) {
  using L = TcLayout<LOGT>;
  constexpr int n = 1 << LOGT;
  const int tid = threadIdx.x;
  float2* twl = lds + L::TW;
  for (int i = tid; i < 512; i += NT) twl[twpad(i)] = tw_f[i];
#pragma unroll
  for (int c = 0; c < 2; c++) {
    const int t1 = tid + 256 * c;
    twb[c][0] = tw_g[t1];
    twb[c][1] = tw_g[(2 * t1) & (n - 1)];
    twb[c][2] = tw_g[(3 * t1) & (n - 1)];
    twb[c][3] = tw_g[(4 * t1) & (n - 1)];
    twb[c][4] = tw_g[(8 * t1) & (n - 1)];
  }
  __syncthreads();
  // the 512-point radix-8 passes' twiddles that are the same for every mode:
  // the last pass's (k = lane) and the middle pass's (k = lane mod 8)
  if constexpr (MAXLR == 3) {
    butterfly_twiddles<8, false, true>(twl, tid & 63, wlast);
    if constexpr (LOGT <= 12) butterfly_twiddles<8, false, true>(twl, (tid & 7) << 3, wmid);
  }
}

template <int LOGT, int MAXLR, bool TWO_LEVEL, bool DENSE>
__global__ __launch_bounds__(NT, LOGT >= 12 ? 2 : 3) void k_temporal_c(
    const float2* __restrict__ hist, int T, int hs, const int4* __restrict__ items, const int* __restrict__ bin_modes,
    const uint32_t* __restrict__ order, const float* __restrict__ mask, int identity, float inv_n0,
    const float2* __restrict__ tw_g, const float2* __restrict__ tw_f, float* __restrict__ S_part,
    float* __restrict__ p_part, float* __restrict__ P_part, int KT, float* __restrict__ G_part) {
  extern __shared__ float2 lds[];
  float2 twb[2][5], wlast[8], wmid[8];
  temporal_c_setup<LOGT, MAXLR>(lds, tw_g, tw_f, twb, wlast, wmid);
  const TemporalArgs A{hist, items, bin_modes, order, mask, S_part, p_part, P_part, G_part, T, hs, identity, KT, 0, inv_n0};
  temporal_c_item<LOGT, MAXLR, TWO_LEVEL, DENSE>(A, blockIdx.x, lds, twb, wlast, wmid);
}

// The batch form: a persistent grid over the work items of many videos of
// one transform length, taken from a queue (as k_temporal_wq does for the
// one-mode-per-wave kernel): vids[v] holds video v's arguments and item0, the
// index of its first item in the queue (increasing in v); nwork items in all;
// *counter is zero before the launch.  Every workgroup reads the same index
// after the barrier, so all of them leave the loop together, and an item is
// computed exactly as k_temporal_c computes it (bit-identical partials).
template <int LOGT, int MAXLR, bool TWO_LEVEL>
__global__ __launch_bounds__(NT, LOGT >= 12 ? 2 : 3) void k_temporal_cq(const TemporalArgs* __restrict__ vids,
                                                                       int nvid, int nwork, int* __restrict__ counter,
                                                                       const float2* __restrict__ tw_g,
                                                                       const float2* __restrict__ tw_f) {
  extern __shared__ float2 lds[];
  __shared__ int next;
  float2 twb[2][5], wlast[8], wmid[8];
  temporal_c_setup<LOGT, MAXLR>(lds, tw_g, tw_f, twb, wlast, wmid);
  for (;;) {
    if (threadIdx.x == 0) next = atomicAdd(counter, 1);
    __syncthreads();
    const int w = next;
    if (w >= nwork) break;
    int lo = 0, hi = nvid - 1;   // the last video whose first item is <= w
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (vids[mid].item0 <= w) lo = mid;
      else hi = mid - 1;
    }
    const TemporalArgs A = vids[lo];
    temporal_c_item<LOGT, MAXLR, TWO_LEVEL>(A, w - A.item0, lds, twb, wlast, wmid);
    __syncthreads();   // the item's last reads of `next` and of the LDS rows before the next item
  }
}

template <int LOGT, int MAXLR, bool TWO_LEVEL>
static hipError_t launch_temporal_c_t(const float2* hist, int T, int hs, const int4* items, int nitems, const int* bin_modes,
                                      const uint32_t* order, const float* mask, int identity, int dense,
                                      float inv_n0, const float2* tw, const float2* tw_f, float* S_part,
                                      float* p_part, float* P_part, int KT, float* G_part, hipStream_t st) {
  const size_t lds = sizeof(float2) * TcLayout<LOGT>::SIZE;
  auto kern = dense ? k_temporal_c<LOGT, MAXLR, TWO_LEVEL, true> : k_temporal_c<LOGT, MAXLR, TWO_LEVEL, false>;
  hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(kern, dim3(nitems), dim3(NT), lds, st, hist, T, hs, items, bin_modes, order, mask, identity, inv_n0,
                     tw, tw_f, S_part, p_part, P_part, KT, G_part);
  return hipGetLastError();
}

// n = 2^logt in [2048, 8192]; items of at most 64 modes (at most 16 at
// logt = 13, whose sums are one-level); S_part rows r-major with
// lm = logt - 9.
hipError_t launch_temporal_c(int logt, int maxlr, const float2* hist, int T, int hs, const int4* items, int nitems,
                             const int* bin_modes, const uint32_t* order, const float* mask, int identity, int dense,
                             float inv_n0, const float2* tw, const float2* tw_f, float* S_part, float* p_part,
                             float* P_part, int KT, float* G_part, hipStream_t st) {
  if (KT < 1 || KT > 64 || KT > T || hs < T || 2 * T > (1 << logt) || (dense && 2 * T != (1 << logt)))
    return hipErrorInvalidValue;
  switch (logt * 4 + (maxlr == 2 ? 2 : 3)) {
#define CASE(LG_, R_, TL_)                                                                                          \
  case LG_ * 4 + R_:                                                                                               \
    return launch_temporal_c_t<LG_, R_, TL_>(hist, T, hs, items, nitems, bin_modes, order, mask, identity, dense, \
                                             inv_n0, tw, tw_f, S_part, p_part, P_part, KT, G_part, st);
    CASE(11, 2, true) CASE(12, 2, true) CASE(13, 2, false)
    CASE(11, 3, true) CASE(12, 3, true) CASE(13, 3, false)
#undef CASE
  }
  return hipErrorInvalidValue;
}

__global__ void k_fill_counter(int* __restrict__ p) { *p = 0; }

// One video's items through a persistent grid smaller than the item count
// (the analysis's arguments by value; *counter zero before the launch): a
// grid of a few workgroups per CU leaves room on every CU for another
// analysis's memory-bound spatial kernels.
template <int LOGT, int MAXLR, bool TWO_LEVEL>
__global__ __launch_bounds__(NT, LOGT >= 12 ? 2 : 3) void k_temporal_cp(const TemporalArgs A, int nwork,
                                                                       int* __restrict__ counter,
                                                                       const float2* __restrict__ tw_g,
                                                                       const float2* __restrict__ tw_f) {
  extern __shared__ float2 lds[];
  __shared__ int next;
  float2 twb[2][5], wlast[8], wmid[8];
  temporal_c_setup<LOGT, MAXLR>(lds, tw_g, tw_f, twb, wlast, wmid);
  for (;;) {
    if (threadIdx.x == 0) next = atomicAdd(counter, 1);
    __syncthreads();
    const int w = next;
    if (w >= nwork) break;
    temporal_c_item<LOGT, MAXLR, TWO_LEVEL>(A, w, lds, twb, wlast, wmid);
    __syncthreads();
  }
}

template <int LOGT, int MAXLR, bool TWO_LEVEL>
static hipError_t launch_temporal_cp_t(const TemporalArgs& a, int nwork, int* counter, const float2* tw,
                                       const float2* tw_f, int grid, hipStream_t st) {
  const size_t lds = sizeof(float2) * TcLayout<LOGT>::SIZE;
  auto kern = k_temporal_cp<LOGT, MAXLR, TWO_LEVEL>;
  hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_fill_counter, dim3(1), dim3(1), 0, st, counter);
  hipLaunchKernelGGL(kern, dim3(grid), dim3(NT), lds, st, a, nwork, counter, tw, tw_f);
  return hipGetLastError();
}

hipError_t launch_temporal_cp(int logt, int maxlr, const TemporalArgs& a, int nwork, int* counter, const float2* tw,
                              const float2* tw_f, int grid, hipStream_t st) {
  if (nwork < 1 || grid < 1 || a.KT < 1 || a.KT > 64 || a.KT > a.T || a.hs < a.T || 2 * a.T > (1 << logt))
    return hipErrorInvalidValue;
  switch (logt * 4 + (maxlr == 2 ? 2 : 3)) {
#define CASE(LG_, R_, TL_) \
  case LG_ * 4 + R_: return launch_temporal_cp_t<LG_, R_, TL_>(a, nwork, counter, tw, tw_f, grid, st);
    CASE(11, 2, true) CASE(12, 2, true) CASE(13, 2, false)
    CASE(11, 3, true) CASE(12, 3, true) CASE(13, 3, false)
#undef CASE
  }
  return hipErrorInvalidValue;
}

template <int LOGT, int MAXLR, bool TWO_LEVEL>
static hipError_t launch_temporal_cq_t(const TemporalArgs* vids, int nvid, int nwork, int* counter, const float2* tw,
                                       const float2* tw_f, int grid, hipStream_t st) {
  const size_t lds = sizeof(float2) * TcLayout<LOGT>::SIZE;
  auto kern = k_temporal_cq<LOGT, MAXLR, TWO_LEVEL>;
  hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  // the counter is cleared by a kernel: a memset node would not do in a captured graph
  hipLaunchKernelGGL(k_fill_counter, dim3(1), dim3(1), 0, st, counter);
  hipLaunchKernelGGL(kern, dim3(grid), dim3(NT), lds, st, vids, nvid, nwork, counter, tw, tw_f);
  return hipGetLastError();
}

template <int LOGT, int MAXLR, bool TWO_LEVEL>
static int temporal_cq_grid_t() {
  const size_t lds = sizeof(float2) * TcLayout<LOGT>::SIZE;
  auto kern = k_temporal_cq<LOGT, MAXLR, TWO_LEVEL>;
  int dev = 0, cus = 0, per_cu = 0;
  if (hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess ||
      hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)kern, NT, lds) != hipSuccess)
    return 0;
  return cus * (per_cu > 0 ? per_cu : 1);
}

int temporal_cq_grid(int logt, int maxlr) {
  switch (logt * 4 + (maxlr == 2 ? 2 : 3)) {
#define CASE(LG_, R_, TL_) \
  case LG_ * 4 + R_: return temporal_cq_grid_t<LG_, R_, TL_>();
    CASE(11, 2, true) CASE(12, 2, true) CASE(13, 2, false)
    CASE(11, 3, true) CASE(12, 3, true) CASE(13, 3, false)
#undef CASE
  }
  return 0;
}

hipError_t launch_temporal_cq(int logt, int maxlr, const TemporalArgs* vids, int nvid, int nwork, int* counter,
                              const float2* tw, const float2* tw_f, int grid, hipStream_t st) {
  if (nvid < 1 || nwork < 1 || grid < 1) return hipErrorInvalidValue;
  switch (logt * 4 + (maxlr == 2 ? 2 : 3)) {
#define CASE(LG_, R_, TL_) \
  case LG_ * 4 + R_: return launch_temporal_cq_t<LG_, R_, TL_>(vids, nvid, nwork, counter, tw, tw_f, grid, st);
    CASE(11, 2, true) CASE(12, 2, true) CASE(13, 2, false)
    CASE(11, 3, true) CASE(12, 3, true) CASE(13, 3, false)
#undef CASE
  }
  return hipErrorInvalidValue;
}

}  // namespace ddm
