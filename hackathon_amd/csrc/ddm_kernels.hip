// Temporal and radial HIP kernels of the DDM hot path for gfx950 (the spatial
// stage is in ddm_spatial.hip).
//
//  k_temporal  one work item = up to M modes of ONE radial bin.  Per mode:
//              gather (order, mask), masked-mean removal, p = m|y|^2, zero pad
//              to TPAD, forward FFT, |Y|^2; p and |Y|^2 are summed over the
//              item's modes (compensated) and written as a per-item partial.
//  k_bin_*     per bin: sum the items' partials, FP64 inverse FFT of the
//              summed |Y|^2 (= sum of the modes' Re G by linearity) taken
//              relative to its zero lag, FP64 masked sums A, B of the summed
//              p, D = (A + B - 2 sum p - 2 (G(tau) - G(0)))/N(tau)/n; the
//              tail lags from the temporal kernels' direct sums.
//
// Reference: src/ddm.wgsl (temporal :202-259, radial :265-322) and its FP64
// twin src/ddm_ref.js:229-288,360-412.
#include "ddm_internal.h"
#include "ddm_fft.h"

namespace ddm {

// float2 slots per wave of k_temporal_w: the sub-transform exchange buffer,
// then the wave's accumulators (n floats of sum |Y|^2, n/2 floats of sum p).
__host__ __device__ constexpr int temporal_wave_lds(int logt, int lm) {
  return xch_size(1 << (logt - lm)) + (3 << logt) / 4;
}

// ----------------------------------------------------------------- k_temporal

// Block-wide sum of a complex value (NT threads), result to all threads.
__device__ __forceinline__ float2 block_sum2(float2 v, float* red, int tid) {
  v.x = seg_sum<64>(v.x);
  v.y = seg_sum<64>(v.y);
  const int w = tid >> 6;
  if ((tid & 63) == 0) {
    red[2 * w] = v.x;
    red[2 * w + 1] = v.y;
  }
  __syncthreads();
  float2 r = make_float2(0.f, 0.f);
#pragma unroll
  for (int i = 0; i < NT / 64; i++) {
    r.x += red[2 * i];
    r.y += red[2 * i + 1];
  }
  __syncthreads();
  return r;
}

template <int LOGT>
__global__ __launch_bounds__(NT) void k_temporal(const float2* __restrict__ hist, int T, int hs, const int4* __restrict__ items,
                                                 const int* __restrict__ bin_modes, const uint32_t* __restrict__ order,
                                                 const float* __restrict__ mask, int identity, float inv_n0,
                                                 const float2* __restrict__ tw_g, float* __restrict__ S_part,
                                                 float* __restrict__ p_part, float* __restrict__ P_part, int KT,
                                                 float* __restrict__ G_part) {
  constexpr int n = 1 << LOGT;
  constexpr int KJ = n / NT;           // frequency bins per thread (n >= 256 enforced by the host: n >= 2T, T >= 2,
                                       // and the host pads n to at least NT)
  constexpr int TJ = (n / 2 + NT - 1) / NT;  // time samples per thread (T <= n/2)
  extern __shared__ float2 lds[];
  float2* buf = lds;                   // n
  float2* tw = lds + n;                // n
  float* red = reinterpret_cast<float*>(tw + n);  // 2 * NT/64
  const int tid = threadIdx.x;
  const int4 it = items[blockIdx.x];
  const int start = it.y, count = it.z;
  for (int i = tid; i < n; i += NT) tw[i] = tw_g[i];

  float mk[TJ];
  uint32_t od[TJ];
#pragma unroll
  for (int i = 0; i < TJ; i++) {
    int j = tid + i * NT;
    mk[i] = (j < T && mask[j] != 0.f) ? 1.f : 0.f;
    od[i] = (j < T) ? order[j] : 0u;
  }
  float ap[TJ], cp[TJ];
  float aS[KJ], cS[KJ];
#pragma unroll
  for (int i = 0; i < TJ; i++) ap[i] = cp[i] = 0.f;
#pragma unroll
  for (int i = 0; i < KJ; i++) aS[i] = cS[i] = 0.f;
  float aP = 0.f, cP = 0.f;
  float Gt = 0.f, Gtc = 0.f;   // tail lag T - KT + tid / 4 (threads 4 j, j < KT), summed directly
  __syncthreads();

  for (int mi = 0; mi < count; mi++) {
    const long mode = bin_modes[start + mi];
    const float2* x = hist + mode * hs;
    float2 y[TJ];
#pragma unroll
    for (int i = 0; i < TJ; i++) {
      int j = tid + i * NT;
      y[i] = (j < T) ? x[j] : make_float2(0.f, 0.f);
    }
    if (!identity) {
#pragma unroll
      for (int i = 0; i < TJ; i++) {
        int j = tid + i * NT;
        if (j < T) buf[j] = y[i];
      }
      __syncthreads();
#pragma unroll
      for (int i = 0; i < TJ; i++) {
        int j = tid + i * NT;
        y[i] = (j < T) ? buf[od[i]] : make_float2(0.f, 0.f);
      }
      __syncthreads();
    }
    float2 sy = make_float2(0.f, 0.f);
#pragma unroll
    for (int i = 0; i < TJ; i++) {
      y[i] = make_float2(y[i].x * mk[i], y[i].y * mk[i]);
      sy = cadd(sy, y[i]);
      comp_add(aP, cP, y[i].x * y[i].x + y[i].y * y[i].y);
    }
    float2 mean = block_sum2(sy, red, tid);
    mean.x *= inv_n0;
    mean.y *= inv_n0;
#pragma unroll
    for (int i = 0; i < TJ; i++) {
      int j = tid + i * NT;
      float2 v = make_float2((y[i].x - mean.x) * mk[i], (y[i].y - mean.y) * mk[i]);
      comp_add(ap[i], cp[i], v.x * v.x + v.y * v.y);
      if (j < n / 2) buf[j] = v;
    }
    for (int j = tid; j < n / 2; j += NT) buf[n / 2 + j] = make_float2(0.f, 0.f);
    __syncthreads();
    {
      // tail lags in the difference form sum m m |y(t + tau) - y(t)|^2 (see
      // k_temporal_w); 4 threads per lag (KT <= 64): thread 4 j + s sums
      // pairs t = s, s + 4, ...
      const int j = tid >> 2, s4 = tid & 3;
      float g = 0.f;
      if (j < KT)
        for (int t = s4; t < KT - j; t += 4) {
          const float2 a = buf[t], b = buf[T - KT + j + t];
          const float w = (mask[t] != 0.f && mask[T - KT + j + t] != 0.f) ? 1.f : 0.f;
          const float dx = b.x - a.x, dy = b.y - a.y;
          g += w * (dx * dx + dy * dy);
        }
      g += __shfl_xor(g, 1, 64);
      g += __shfl_xor(g, 2, 64);
      if (s4 == 0 && j < KT) two_sum_add(Gt, Gtc, g);
    }
    __syncthreads();
    lds_fft<LOGT, 1, NT, false>(buf, tw, tid);
#pragma unroll
    for (int i = 0; i < KJ; i++) {
      float2 v = buf[tid + i * NT];
      comp_add(aS[i], cS[i], v.x * v.x + v.y * v.y);
    }
    __syncthreads();
  }
  float* Sp = S_part + (long)blockIdx.x * n;
#pragma unroll
  for (int i = 0; i < KJ; i++) Sp[tid + i * NT] = aS[i] + cS[i];
  float* pp = p_part + (long)blockIdx.x * T;
#pragma unroll
  for (int i = 0; i < TJ; i++) {
    int j = tid + i * NT;
    if (j < T) pp[j] = ap[i] + cp[i];
  }
  float2 tp = block_sum2(make_float2(aP + cP, 0.f), red, tid);
  if (tid == 0) P_part[blockIdx.x] = tp.x;
  if ((tid & 3) == 0 && (tid >> 2) < KT) G_part[(long)blockIdx.x * KT + (tid >> 2)] = Gt + Gtc;
}

// ----------------------------------------------------------------- k_finalize


// ----------------------------------------------------------------- k_temporal_w
//
// Wavefront-per-mode version of k_temporal.  The item's modes are processed in
// rounds of 4, one per wave.  The wave keeps the mode's series y(t), t < n/2,
// in registers (canonical distribution t = lane + 64 q).  Because y is zero
// for t >= n/2, the n-point transform splits exactly into m = 2^LM independent
// (n/m)-point transforms (first decimation-in-frequency stage):
//   Y(m k + r) = FFT_{n/m}[ u_r ](k),  u_r(t1) = w_n^{r t1} sum_{t2} y(t1 + t2 n/m) w_m^{r t2}
// (only t2 < m/2 contribute), so a wave never holds more than n/m points of a
// transform (m = 1 for n <= 512, whose first pass skips the zero half
// instead; 512-point sub-transforms above).  Each sub-transform uses
// wave-private LDS exchanges.  |Y|^2 and p are added into accumulators
// private to the wave (plain LDS read-modify-write: one lane owns each
// address, sums in mode order), so the waves never wait for each other inside
// the mode loop; the 4 waves' sums are added pairwise once, at the end.  (The
// same adds as LDS float atomics made the kernel 3.7x slower on gfx950.)  The next round's
// series is loaded while the current one is transformed.
// Per-workgroup set-up shared by every item a workgroup takes: the
// sub-transform twiddle table staged in LDS, the per-lane DIF pre-twiddle
// bases and the last pass's twiddles (the same for every mode and video of
// one transform length).
template <int LOGT, int LM, int MAXLR>
__device__ __forceinline__ void temporal_w_setup(float2* lds, const float2* __restrict__ tw_g,
                                                 const float2* __restrict__ tw_f, float2 (&tbase)[(1 << LM) > 1 ? (1 << LM) : 1],
                                                 float2 (&wlast)[8]) {
  constexpr int m = 1 << LM;
  constexpr int LOGF = LOGT - LM;
  constexpr int NF = 1 << LOGF;
  constexpr int WS = temporal_wave_lds(LOGT, LM);
  const int tid = threadIdx.x, lane = tid & 63;
  // the sub-transform's twiddle table, staged once per workgroup at twpad()
  // slots: the FFT passes read it from LDS instead of waiting on global loads
  float2* twl = lds + 4 * WS + 2;
  for (int i = tid; i < NF; i += NT) twl[twpad(i)] = tw_f[i];
  // DIF pre-twiddles w_n^{r (lane + 64 q)} = w_n^{r lane} w_n^{64 r q}: the
  // per-lane factor is loaded once here, the second is wave-uniform
#pragma unroll
  for (int r = 0; r < (m > 1 ? m : 1); r++) tbase[r] = tw_g[r * lane];
  __syncthreads();
  // last-pass twiddles w_NF^{r lane} of the 512-point sub-transforms, the same
  // for every mode: kept in registers
  if constexpr (LOGF == 9 && MAXLR == 3) butterfly_twiddles<8, false, true>(twl, lane, wlast);
}

// Register budget switches of the temporal kernel (transform lengths 2^LOGT up
// to these keep the |Y|^2 sums in registers / prefetch the next series).
#ifndef DDM_YREG_MAX_LOGT
#define DDM_YREG_MAX_LOGT 10
#endif
#ifndef DDM_PREFETCH_MAX_LOGT
#define DDM_PREFETCH_MAX_LOGT 11
#endif

// One work item (up to 64 modes of one bin) of one video's temporal stage;
// A.items[item] names it and its partials go to row `item` of A's outputs.
template <int LOGT, int LM, int MAXLR>
__device__ __forceinline__ void temporal_w_item(const TemporalArgs& A, const int item, float2* lds,
                                                const float2 (&tbase)[(1 << LM) > 1 ? (1 << LM) : 1],
                                                const float2 (&wlast)[8]) {
  constexpr int n = 1 << LOGT;
  constexpr int m = 1 << LM;
  constexpr int LOGF = LOGT - LM;  // sub-transform size
  constexpr int NF = 1 << LOGF;
  constexpr int PF = NF / 64;      // points per lane of a sub-transform
  constexpr int PT = n / 128;      // time samples per lane (T <= n/2)
  constexpr int WS = temporal_wave_lds(LOGT, LM);  // float2 per wave
  constexpr int KS = n / NT;       // frequencies per thread in the accumulation
  constexpr int TS = (n / 2) / NT > 0 ? (n / 2) / NT : 1;
  static_assert(PT <= 32, "mask bits");
  const float2* __restrict__ hist = A.hist;
  const int4* __restrict__ items = A.items;
  const int* __restrict__ bin_modes = A.bin_modes;
  const uint32_t* __restrict__ order = A.order;
  const float* __restrict__ mask = A.mask;
  float* __restrict__ S_part = A.S_part;
  float* __restrict__ p_part = A.p_part;
  float* __restrict__ P_part = A.P_part;
  float* __restrict__ G_part = A.G_part;
  const int T = A.T, hs = A.hs, identity = A.identity, KT = A.KT;
  const float inv_n0 = A.inv_n0;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  float2* xch = lds + wave * WS;
  float* red = reinterpret_cast<float*>(lds + 4 * WS);  // 4
  const float2* twl = lds + 4 * WS + 2;
  const int4 it = items[item];
  const int start = it.y, count = it.z;
  uint32_t mbits = 0;
#pragma unroll
  for (int q = 0; q < PT; q++) {
    int t = lane + 64 * q;
    if (t < T && mask[t] != 0.f) mbits |= 1u << q;
  }
  constexpr uint32_t all_bits = PT == 32 ? 0xffffffffu : ((1u << PT) - 1u);
  // wave-uniform: every lane's PT samples valid (full mask and T = n/2)
  const bool dense = __all(mbits == all_bits);
  // the wave's accumulators behind its exchange buffer: sum |Y(k)|^2 at
  // [0, n) (r-major for LM > 0: index r n/m + j holds k = m j + r; k_bin_sum
  // undoes the order) and, at the end, sum p(t) at [n, 3n/2).  A wave sums at most 16
  // modes of an item plainly (positive terms: relative error <= 15 ulp); the
  // items' partials are added with compensation in k_bin_sum.
  float* acc = reinterpret_cast<float*>(xch + xch_size(NF));
  // sum |Y|^2 in registers too where they fit (lane-owned frequencies: the
  // sub-transform outputs are canonical), else read-modify-write in LDS
  constexpr bool YREG = LOGT <= DDM_YREG_MAX_LOGT;
  float yacc[YREG ? n / 64 : 1];
  if constexpr (YREG) {
#pragma unroll
    for (int i = 0; i < n / 64; i++) yacc[i] = 0.f;
  } else {
    for (int i = lane; i < n; i += 64) acc[i] = 0.f;
  }
  // sum p(t) of the wave's modes: lane-owned samples t = lane + 64 q, kept in
  // registers (stored behind the |Y|^2 sums once the item is done)
  float pacc[PT];
#pragma unroll
  for (int q = 0; q < PT; q++) pacc[q] = 0.f;
  float aP = 0.f, cP = 0.f;
  const int lpl = KT <= 16 ? 4 : (KT <= 32 ? 2 : 1);   // lanes per tail lag (wave-uniform; KT <= 64)
  float Gt = 0.f, Gtc = 0.f;   // tail lag T - KT + lane / lpl (lanes lpl j, j < KT)
  const int rounds = (count + 3) / 4;
  // Series loads without guards: a wave past the item's last mode re-reads
  // that mode (its contributions are dropped through `active`); samples
  // t in [T, n/2) read the row's hs - T padding slots, the next mode's head
  // or the n/2 - hs spare samples the history buffer carries behind its last
  // row (they exist only when the series is not dense, and the mask multiply
  // zeroes them).  One base
  // address with immediate offsets; guarded or clamped loads would cost a
  // branch region or an address pair per sample.
  // A non-identity order is applied by the loads themselves: sample t is
  // gathered from slot order[t] (the n/2 entries of order[] past T are 0,
  // and their samples are masked off).
  uint32_t ord[PT];
#pragma unroll
  for (int q = 0; q < PT; q++) ord[q] = identity ? (uint32_t)(lane + 64 * q) : order[lane + 64 * q];
  auto load_series = [&](float2 (&dst)[PT], int slot) {
    const float2* x = hist + (long)bin_modes[start + (slot < count ? slot : count - 1)] * hs;
    if (identity) {
#pragma unroll
      for (int q = 0; q < PT; q++) {   // each series is read once: non-temporal loads
        const float2* src = &x[lane + 64 * q];
        dst[q] = make_float2(__builtin_nontemporal_load(&src->x), __builtin_nontemporal_load(&src->y));
      }
    } else {
#pragma unroll
      for (int q = 0; q < PT; q++) {
        const float2* src = &x[ord[q]];
        dst[q] = make_float2(__builtin_nontemporal_load(&src->x), __builtin_nontemporal_load(&src->y));
      }
    }
  };
  // next round's series, loaded while the current one is transformed (up to
  // n = 2048; above that the registers are needed by the split transforms)
  constexpr bool PREFETCH = LOGT <= DDM_PREFETCH_MAX_LOGT;
  float2 nx[PREFETCH ? PT : 1];
  if constexpr (PREFETCH) load_series(nx, wave);
  wave_sync();
  for (int rd = 0; rd < rounds; rd++) {
    const int mi = rd * 4 + wave;
    const bool active = mi < count;
    float2 y[PT];
    if constexpr (PREFETCH) {
#pragma unroll
      for (int q = 0; q < PT; q++) y[q] = nx[q];
      load_series(nx, mi + 4);
    } else {
      load_series(y, mi);
    }
    float sx = 0.f, sy = 0.f, ps = 0.f;
    if (dense) {   // every sample of every lane valid: no masking
#pragma unroll
      for (int q = 0; q < PT; q++) {
        sx += y[q].x;
        sy += y[q].y;
        ps += y[q].x * y[q].x + y[q].y * y[q].y;
      }
    } else {
#pragma unroll
      for (int q = 0; q < PT; q++) {
        const float mq = ((mbits >> q) & 1u) ? 1.f : 0.f;
        y[q] = make_float2(y[q].x * mq, y[q].y * mq);
        sx += y[q].x;
        sy += y[q].y;
        ps += y[q].x * y[q].x + y[q].y * y[q].y;
      }
    }
    // one compensated add per mode and lane: the lane's 16 positive terms are
    // summed plainly (relative error <= 15 ulp of this mode's share)
    two_sum_add(aP, cP, active ? ps : 0.f);
    const float mx = wave_sum(sx) * inv_n0, my = wave_sum(sy) * inv_n0;
    if (dense) {
#pragma unroll
      for (int q = 0; q < PT; q++) y[q] = make_float2(y[q].x - mx, y[q].y - my);
    } else {
#pragma unroll
      for (int q = 0; q < PT; q++) {
        const float mq = ((mbits >> q) & 1u) ? 1.f : 0.f;
        y[q] = make_float2((y[q].x - mx) * mq, (y[q].y - my) * mq);
      }
    }
    // Tail lags tau = T - KT + j (few valid pairs) directly, in the
    // difference form sum_t m(t) m(t + tau) |y(t + tau) - y(t)|^2: the FFT
    // route's error is ~ulp x sum_t |y|^2, far above D N(tau) there, and
    // products y*(t) y(t + tau) would still cancel in slow modes.  y(t < KT)
    // and y(t >= T - KT) meet in the exchange buffer with their mask values
    // (xm, behind the first 128 slots): the head is in register 0 (KT <= 64),
    // the tail in the registers q0..q1 (wave-uniform, at most two).  The
    // previous mode's last exchange reads are done first.
    float* xm = reinterpret_cast<float*>(xch + 128);
    wave_sync();
    if (lane < KT) {
      xch[lane] = y[0];
      xm[lane] = (mbits & 1u) ? 1.f : 0.f;
    }
    {
      const int q0 = (T - KT) >> 6, q1 = (T - 1) >> 6;
#pragma unroll
      for (int q = 0; q < PT; q++) {
        if (q >= q0 && q <= q1) {
          const int t = lane + 64 * q;
          if (t >= T - KT && t < T) {
            xch[KT + t - (T - KT)] = y[q];
            xm[KT + t - (T - KT)] = ((mbits >> q) & 1u) ? 1.f : 0.f;
          }
        }
      }
    }
    wave_sync();
    {
      // lpl lanes per lag (4 for KT <= 16, 2 for KT <= 32, else 1): lane lpl j + s
      // sums pairs t = s, s + lpl, ...
      const int j = lane / lpl, s4 = lane % lpl;
      float g = 0.f;
      if (j < KT)
        for (int t = s4; t < KT - j; t += lpl) {
          const float2 a = xch[t], b = xch[KT + t + j];
          const float dx = b.x - a.x, dy = b.y - a.y;
          g += (xm[t] * xm[KT + t + j]) * (dx * dx + dy * dy);
        }
      g = lpl == 4 ? seg_sum_dpp<4>(g) : (lpl == 2 ? seg_sum_dpp<2>(g) : g);
      if (s4 == 0 && j < KT && active) two_sum_add(Gt, Gtc, g);
    }
    if constexpr (LM == 0) {
      float2 v[PF];
#pragma unroll
      for (int q = 0; q < PT; q++) v[q] = y[q];
#pragma unroll
      for (int q = PT; q < PF; q++) v[q] = make_float2(0.f, 0.f);
      wave_fft<LOGF, false, true, true, MAXLR>(v, xch, twl, lane);
      if (active) {
#pragma unroll
        for (int q = 0; q < PF; q++) {
          const float a2 = v[q].x * v[q].x + v[q].y * v[q].y;
          if constexpr (YREG) yacc[q] += a2;
          else acc[lane + 64 * q] += a2;
        }
#pragma unroll
        for (int q = 0; q < PT; q++) pacc[q] += y[q].x * y[q].x + y[q].y * y[q].y;
      }
    } else {
      // u_r(t1) = w_n^{r t1} sum_{t2 < m/2} y(t1 + t2 n/m) w_m^{r t2}: for
      // t1 = lane + 64 q the samples t1 + t2 n/m = lane + 64 (q + t2 PF) are
      // in this lane's own registers.  |Y(m k + r)|^2 is added after each
      // sub-transform (consecutive lanes, consecutive words: no bank conflicts).
      // Unrolled: with r a constant, the (-i)^r rotation, the pre-twiddle base
      // and the table offsets fold away (about 300 fewer VALU ops per mode at
      // m = 4 than a loop over r).
#pragma unroll
      for (int r = 0; r < m; r++) {
        float2 v[PF];
        __builtin_amdgcn_sched_barrier(0);
        float2 tb = tbase[0];
#pragma unroll
        for (int k = 1; k < m; k++)
          if (r == k) tb = tbase[k];
        // the wave-uniform factors w_m^{r t2} and w_n^{64 r q} are entries of
        // the sub-transform table (w_NF^k) read as LDS broadcasts; w_m^{r m/4}
        // = (-i)^r is applied as a swap and a sign
#pragma unroll
        for (int q = 0; q < PF; q++) {
          float2 u = y[q];
#pragma unroll
          for (int t2 = 1; t2 < m / 2; t2++) {
            const float2 b = y[q + t2 * PF];
            if (4 * t2 == m) {
              const float2 rb = (r & 1) ? make_float2(b.y, -b.x) : b;
              u = (r & 2) ? csub(u, rb) : cadd(u, rb);
            } else {
              u = cadd(u, cmul(b, twl[twpad(((r * t2) & (m - 1)) * (NF / m))]));
            }
          }
          v[q] = r ? cmul(u, cmul(tb, twl[twpad(((64 / m) * r * q) & (NF - 1))])) : u;
        }
        if constexpr (LOGF == 9 && MAXLR == 3) wave_fft_wl<LOGF, false, false, true>(v, xch, twl, lane, wlast);
        else wave_fft<LOGF, false, false, true, MAXLR>(v, xch, twl, lane);
        if (active) {
#pragma unroll
          for (int q = 0; q < PF; q++) {
            const float a2 = v[q].x * v[q].x + v[q].y * v[q].y;
            if constexpr (YREG) yacc[r * PF + q] += a2;
            else acc[r * NF + lane + 64 * q] += a2;
          }
        }
      }
      if (active) {
#pragma unroll
        for (int q = 0; q < PT; q++) pacc[q] += y[q].x * y[q].x + y[q].y * y[q].y;
      }
    }
  }
#pragma unroll
  for (int q = 0; q < PT; q++) acc[n + lane + 64 * q] = pacc[q];
  if constexpr (YREG) {
#pragma unroll
    for (int i = 0; i < n / 64; i++) acc[(i / PF) * NF + lane + 64 * (i % PF)] = yacc[i];
  }
  __syncthreads();
  {
    // the 4 waves' sums, added pairwise (positive terms: relative error <= 2 ulp)
    const float* p0 = reinterpret_cast<const float*>(lds) + 2 * xch_size(NF);
    constexpr int WF = 2 * WS;                     // wave stride, in floats
    float* Sp = S_part + (long)item * n;
#pragma unroll
    for (int i = 0; i < KS; i++) {
      const int k = tid + NT * i;
      Sp[k] = (p0[k] + p0[WF + k]) + (p0[2 * WF + k] + p0[3 * WF + k]);
    }
    float* pp = p_part + (long)item * T;
#pragma unroll
    for (int i = 0; i < TS; i++) {
      const int t = tid + NT * i;
      if (t < T) pp[t] = (p0[n + t] + p0[WF + n + t]) + (p0[2 * WF + n + t] + p0[3 * WF + n + t]);
    }
  }
  const float pw = wave_sum(aP + cP);
  if (lane == 0) red[wave] = pw;
  // the 4 waves' tail lags meet in their (now idle) exchange buffers
  if (lane % lpl == 0 && lane / lpl < KT) reinterpret_cast<float*>(xch)[lane / lpl] = Gt + Gtc;
  __syncthreads();
  if (tid == 0) P_part[item] = (red[0] + red[1]) + (red[2] + red[3]);
  if (tid < KT) {
    const float* g0 = reinterpret_cast<const float*>(lds);
    constexpr int WF = 2 * WS;
    G_part[(long)item * KT + tid] = (g0[tid] + g0[WF + tid]) + (g0[2 * WF + tid] + g0[3 * WF + tid]);
  }
}

template <int LOGT, int LM, int MAXLR>
// TPAD = 4096: one workgroup per CU by LDS (the per-wave accumulators), so
// up to 512 registers per lane instead of spilling at 256
__global__ __launch_bounds__(NT, LOGT >= 12 ? 1 : 2) void k_temporal_w(const float2* __restrict__ hist, int T, int hs,
                                                      const int4* __restrict__ items,
                                                      const int* __restrict__ bin_modes,
                                                      const uint32_t* __restrict__ order,
                                                      const float* __restrict__ mask, int identity, float inv_n0,
                                                      const float2* __restrict__ tw_g, const float2* __restrict__ tw_f,
                                                      float* __restrict__ S_part, float* __restrict__ p_part,
                                                      float* __restrict__ P_part, int KT, float* __restrict__ G_part) {
  constexpr int n = 1 << LOGT;
  constexpr int m = 1 << LM;
  constexpr int LOGF = LOGT - LM;  // sub-transform size
  constexpr int NF = 1 << LOGF;
  constexpr int PF = NF / 64;      // points per lane of a sub-transform
  constexpr int PT = n / 128;      // time samples per lane (T <= n/2)
  constexpr int WS = temporal_wave_lds(LOGT, LM);  // float2 per wave
  constexpr int KS = n / NT;       // frequencies per thread in the accumulation
  constexpr int TS = (n / 2) / NT > 0 ? (n / 2) / NT : 1;
  static_assert(PT <= 32, "mask bits");
  extern __shared__ float2 lds[];
  float2 tbase[(1 << LM) > 1 ? (1 << LM) : 1], wlast[8];
  temporal_w_setup<LOGT, LM, MAXLR>(lds, tw_g, tw_f, tbase, wlast);
  const TemporalArgs A{hist, items, bin_modes, order, mask, S_part, p_part, P_part, G_part, T, hs, identity, KT, 0, inv_n0};
  temporal_w_item<LOGT, LM, MAXLR>(A, blockIdx.x, lds, tbase, wlast);
}

// The batch form: a persistent grid over the work items of many videos (one
// transform length), taken from a queue.  vids[v] holds video v's arguments
// and item0, the index of its first item in the queue (increasing in v);
// nwork items in all.  A workgroup takes the next index from *counter (zero
// before the launch) until the queue is empty -- every workgroup sees the
// same index, so all of them leave the loop -- and computes the item exactly
// as k_temporal_w does, so the partials are bit-identical.
template <int LOGT, int LM, int MAXLR>
__global__ __launch_bounds__(NT, LOGT >= 12 ? 1 : 2) void k_temporal_wq(const TemporalArgs* __restrict__ vids,
                                                                       int nvid, int nwork, int* __restrict__ counter,
                                                                       const float2* __restrict__ tw_g,
                                                                       const float2* __restrict__ tw_f) {
  extern __shared__ float2 lds[];
  __shared__ int next;
  float2 tbase[(1 << LM) > 1 ? (1 << LM) : 1], wlast[8];
  temporal_w_setup<LOGT, LM, MAXLR>(lds, tw_g, tw_f, tbase, wlast);
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
    temporal_w_item<LOGT, LM, MAXLR>(A, w - A.item0, lds, tbase, wlast);
    __syncthreads();   // the item's last LDS reads before the next item's writes
  }
}

// ----------------------------------------------------------------- k_temporal_full
//
// Per-mode outputs of readTemporalFull (opt-in, not part of computeTemporal):
// for modes m0 .. m0 + gridDim.x - 1 (every mode, binned or not), one
// workgroup per mode:
//   C(tau) = G(tau) / N(tau), G the masked autocorrelation of the gathered
//            series y = m x (no mean removal: C is not shift-invariant);
//   D(tau) = (A'(tau) + B'(tau) - 2 Re G'(tau)) / N(tau) on the mean-removed
//            series y' = m (y - mean) (D is shift-invariant; the removal keeps
//            the cancellation small), A'(tau) = sum_t m(t) p'(t + tau),
//            B'(tau) = A'(-tau), p' = m |y'|^2;
//   P = sum_t m |x|^2 / N(0).
// Three zero-padded TPAD transforms pairs: y -> |Y|^2 -> G; y' -> |Y'|^2 -> G';
// z = m + i s p' -> conj(M) P -> A' (s = T / sum p' brings p' to the mask's
// scale so the packed transform does not bury it).  The last KT lags take G,
// G' and A' + B' from direct sums (the transforms' error there is ~ulp x T /
// N(tau) of the value).
template <int LOGT>
__global__ __launch_bounds__(NT) void k_temporal_full(const float2* __restrict__ hist, int T, int hs, int L, long m0,
                                                      const uint32_t* __restrict__ order,
                                                      const float* __restrict__ mask, int identity, float inv_n0,
                                                      const float* __restrict__ npairs, const float2* __restrict__ tw_g,
                                                      int KT, const int* __restrict__ mode_list,
                                                      float* __restrict__ outD, float2* __restrict__ outC,
                                                      float* __restrict__ outP) {
  constexpr int n = 1 << LOGT;
  constexpr int TJ = (n / 2 + NT - 1) / NT;   // time samples per thread (T <= n/2)
  constexpr int LJ = (n / 2 + NT - 1) / NT;   // lags per thread (L <= T <= n/2)
  extern __shared__ float2 lds[];
  float2* buf = lds;                           // n
  float2* tw = lds + n;                        // n
  float2* tailC = tw + n;                      // KT (complex G of the tail lags)
  float* tailG = reinterpret_cast<float*>(tailC + 64);   // KT
  float* tailAB = tailG + 64;                  // KT
  float* red = tailAB + 64;                    // 2 * NT / 64
  const int tid = threadIdx.x;
  // modes m0 + blockIdx.x, or entries m0 + blockIdx.x of mode_list
  const long mode = mode_list ? (long)mode_list[m0 + blockIdx.x] : m0 + blockIdx.x;
  for (int i = tid; i < n; i += NT) tw[i] = tw_g[i];
  const float2* x = hist + mode * hs;
  float mk[TJ];
  float2 y[TJ];
#pragma unroll
  for (int i = 0; i < TJ; i++) {
    const int j = tid + i * NT;
    mk[i] = (j < T && mask[j] != 0.f) ? 1.f : 0.f;
    y[i] = j < T ? x[j] : make_float2(0.f, 0.f);
  }
  if (!identity) {
#pragma unroll
    for (int i = 0; i < TJ; i++) {
      const int j = tid + i * NT;
      if (j < T) buf[j] = y[i];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < TJ; i++) {
      const int j = tid + i * NT;
      y[i] = j < T ? buf[order[j]] : make_float2(0.f, 0.f);
    }
    __syncthreads();
  }
  float aP = 0.f, cP = 0.f;
  float2 sy = make_float2(0.f, 0.f);
#pragma unroll
  for (int i = 0; i < TJ; i++) {
    y[i] = make_float2(y[i].x * mk[i], y[i].y * mk[i]);
    sy = cadd(sy, y[i]);
    two_sum_add(aP, cP, y[i].x * y[i].x + y[i].y * y[i].y);
  }
  const float2 Ptot = block_sum2(make_float2(aP + cP, 0.f), red, tid);
  float2 mean = block_sum2(sy, red, tid);
  mean.x *= inv_n0;
  mean.y *= inv_n0;
  if (tid == 0 && outP) outP[blockIdx.x] = Ptot.x * inv_n0;
  const int j4 = tid >> 2, s4 = tid & 3;      // 4 threads per tail lag

  // ---- pass 1: G of y -> C
#pragma unroll
  for (int i = 0; i < TJ; i++) {
    const int j = tid + i * NT;
    if (j < n / 2) buf[j] = y[i];
  }
  for (int j = tid; j < n / 2; j += NT) buf[n / 2 + j] = make_float2(0.f, 0.f);
  __syncthreads();
  {
    float gr = 0.f, gi = 0.f;
    if (j4 < KT)
      for (int t = s4; t < KT - j4; t += 4) {
        const float2 a = buf[t], b = buf[T - KT + j4 + t];   // conj(y(t)) y(t + tau), as IFFT(|Y|^2)
        gr += a.x * b.x + a.y * b.y;
        gi += a.x * b.y - a.y * b.x;
      }
    gr += __shfl_xor(gr, 1, 64);
    gr += __shfl_xor(gr, 2, 64);
    gi += __shfl_xor(gi, 1, 64);
    gi += __shfl_xor(gi, 2, 64);
    if (s4 == 0 && j4 < KT) tailC[j4] = make_float2(gr, gi);
  }
  __syncthreads();
  lds_fft<LOGT, 1, NT, false>(buf, tw, tid);
  for (int i = tid; i < n; i += NT) {
    const float2 v = buf[i];
    buf[i] = make_float2(v.x * v.x + v.y * v.y, 0.f);
  }
  __syncthreads();
  lds_fft<LOGT, 1, NT, true>(buf, tw, tid);
  const float invn = 1.0f / n;
#pragma unroll
  for (int k = 0; k < LJ; k++) {
    const int tau = tid + k * NT;
    if (tau < L) {
      float2 g = make_float2(buf[tau].x * invn, buf[tau].y * invn);
      if (tau >= T - KT) g = tailC[tau - (T - KT)];
      const float np = npairs[tau];
      outC[(long)blockIdx.x * L + tau] = np > 0.5f ? make_float2(g.x / np, g.y / np) : make_float2(0.f, 0.f);
    }
  }
  __syncthreads();

  // ---- pass 2: G' of the mean-removed series
  float psum = 0.f;
#pragma unroll
  for (int i = 0; i < TJ; i++) {
    y[i] = make_float2((y[i].x - mean.x) * mk[i], (y[i].y - mean.y) * mk[i]);
    psum += y[i].x * y[i].x + y[i].y * y[i].y;
    const int j = tid + i * NT;
    if (j < n / 2) buf[j] = y[i];
  }
  for (int j = tid; j < n / 2; j += NT) buf[n / 2 + j] = make_float2(0.f, 0.f);
  const float2 pt = block_sum2(make_float2(psum, 0.f), red, tid);   // barriers inside
  {
    float g = 0.f, ab = 0.f;
    if (j4 < KT)
      for (int t = s4; t < KT - j4; t += 4) {
        const float2 a = buf[t], b = buf[T - KT + j4 + t];
        g += a.x * b.x + a.y * b.y;
        // A' + B' pairs (t, t + tau): m(t) p'(t + tau) + p'(t) m(t + tau); y' = 0 where m = 0
        const float ma = mask[t] != 0.f ? 1.f : 0.f, mb = mask[T - KT + j4 + t] != 0.f ? 1.f : 0.f;
        ab += ma * (b.x * b.x + b.y * b.y) + mb * (a.x * a.x + a.y * a.y);
      }
    g += __shfl_xor(g, 1, 64);
    g += __shfl_xor(g, 2, 64);
    ab += __shfl_xor(ab, 1, 64);
    ab += __shfl_xor(ab, 2, 64);
    if (s4 == 0 && j4 < KT) {
      tailG[j4] = g;
      tailAB[j4] = ab;
    }
  }
  __syncthreads();
  lds_fft<LOGT, 1, NT, false>(buf, tw, tid);
  for (int i = tid; i < n; i += NT) {
    const float2 v = buf[i];
    buf[i] = make_float2(v.x * v.x + v.y * v.y, 0.f);
  }
  __syncthreads();
  lds_fft<LOGT, 1, NT, true>(buf, tw, tid);
  float gp[LJ];
#pragma unroll
  for (int k = 0; k < LJ; k++) {
    const int tau = tid + k * NT;
    gp[k] = tau < L ? (tau >= T - KT ? tailG[tau - (T - KT)] : buf[tau].x * invn) : 0.f;
  }
  __syncthreads();

  // ---- pass 3: A' + B' from z = m + i s p'
  const float sc = pt.x > 0.f ? (float)T / pt.x : 1.f;
#pragma unroll
  for (int i = 0; i < TJ; i++) {
    const int j = tid + i * NT;
    if (j < n / 2) buf[j] = make_float2(mk[i], sc * mk[i] * (y[i].x * y[i].x + y[i].y * y[i].y));
  }
  for (int j = tid; j < n / 2; j += NT) buf[n / 2 + j] = make_float2(0.f, 0.f);
  __syncthreads();
  lds_fft<LOGT, 1, NT, false>(buf, tw, tid);
  // S(f) = conj(M(f)) P(f) with M = (Z(f) + conj Z(-f)) / 2, P = (Z(f) - conj Z(-f)) / 2i,
  // pairs (f, n - f) by one thread
  float2 sf[(n / 2 + 1 + NT - 1) / NT], sg[(n / 2 + 1 + NT - 1) / NT];
#pragma unroll
  for (int k = 0; k < (n / 2 + 1 + NT - 1) / NT; k++) {
    const int f = tid + k * NT;
    if (f <= n / 2) {
      const int g = (n - f) & (n - 1);
      const float2 zf = buf[f], zg = buf[g];
      // at f: M = (zf + conj zg)/2, P = (zf - conj zg)/(2i); at g: M' = (zg + conj zf)/2 = conj M, P' = conj P
      const float2 M = make_float2(0.5f * (zf.x + zg.x), 0.5f * (zf.y - zg.y));
      const float2 Pp = make_float2(0.5f * (zf.y + zg.y), -0.5f * (zf.x - zg.x));
      // conj(M) P at f and conj(M') P' = M conj(P) at g
      sf[k] = make_float2(M.x * Pp.x + M.y * Pp.y, M.x * Pp.y - M.y * Pp.x);
      sg[k] = make_float2(sf[k].x, -sf[k].y);
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < (n / 2 + 1 + NT - 1) / NT; k++) {
    const int f = tid + k * NT;
    if (f <= n / 2) {
      buf[f] = sf[k];
      if (f != 0 && f != n / 2) buf[n - f] = sg[k];
    }
  }
  __syncthreads();
  lds_fft<LOGT, 1, NT, true>(buf, tw, tid);
  const float inv_sn = invn / sc;
#pragma unroll
  for (int k = 0; k < LJ; k++) {
    const int tau = tid + k * NT;
    if (tau < L) {
      const float ab = tau >= T - KT ? tailAB[tau - (T - KT)]
                                     : (buf[tau].x + buf[(n - tau) & (n - 1)].x) * inv_sn;
      const float np = npairs[tau];
      if (outD) outD[(long)blockIdx.x * L + tau] = np > 0.5f ? (ab - 2.f * gp[k]) / np : 0.f;
    }
  }
}

// ----------------------------------------------------------------- k_drift
//
// Phase-drift least-squares sums of the radial record (accumulateMode,
// src/ddm_ref.js:360-381): for every bin b and lag tau < L-1, over the bin's
// modes (kx, ky), z = C(tau+1) conj C(tau), w = |z|, phi = arg z:
//   Sw, Sxx = sum w kx^2, Sxy, Syy, Rx = sum w kx phi, Ry, Rpp = sum w phi^2, count(w > 0).
// One thread per (bin, lag) over the bin's modes that lie in the current chunk
// [c0, c0 + nc) of the binned-mode list; FP64 accumulation in acc, chunk after
// chunk in stream order (deterministic).  C rows are the chunk's.
__global__ __launch_bounds__(NT) void k_drift(const float2* __restrict__ C, int L, int c0, int nc,
                                              const int* __restrict__ bin_offsets, const int2* __restrict__ kxy,
                                              int b0, double* __restrict__ acc) {
  const int b = b0 + blockIdx.y;
  const int tau = blockIdx.x * NT + threadIdx.x;
  if (tau >= L - 1) return;
  const int i0 = max(bin_offsets[b], c0), i1 = min(bin_offsets[b + 1], c0 + nc);
  double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int i = i0; i < i1; i++) {
    const float2 a = C[(long)(i - c0) * L + tau], bq = C[(long)(i - c0) * L + tau + 1];
    const double zr = (double)bq.x * a.x + (double)bq.y * a.y;
    const double zi = (double)bq.y * a.x - (double)bq.x * a.y;
    const double w = sqrt(zr * zr + zi * zi), phi = atan2(zi, zr);
    const double kx = kxy[i].x, ky = kxy[i].y;
    s[0] += w;
    s[1] += w * kx * kx;
    s[2] += w * kx * ky;
    s[3] += w * ky * ky;
    s[4] += w * kx * phi;
    s[5] += w * ky * phi;
    s[6] += w * phi * phi;
    s[7] += w > 0 ? 1.0 : 0.0;
  }
  double* o = acc + ((long)b * L + tau) * 8;
#pragma unroll
  for (int c = 0; c < 8; c++) o[c] += s[c];
}

// ----------------------------------------------------------------- radial finalisation
//
// Three launches.  k_bin_sum adds the work items' partials of each bin
// (|Y|^2 per frequency, p per time sample, P); k_bin_ifft inverse-FFTs the
// bin's summed |Y|^2 to its summed Re G(tau); k_bin_lags (any mask) or
// k_bin_lags_full (every frame valid) forms the masked correlations
// A(tau) = sum_t m(t) p(t+tau), B(tau) = sum_t p(t) m(t+tau) and
// D = (A + B - 2G) / N(tau) / nModes.
// k_bin_sum: grid (NB, n/NT + ceil(T/NT)): one thread per frequency (first
// n/NT block rows) or per time sample (the rest) sums the bin's item partials
// with 2Sum, four independent loads in flight; frequency sums are written in
// natural order (undoing the r-major interleave, see above).  Block (b, 0) also
// sums the items' P.
template <int LOGT>
__global__ __launch_bounds__(NT) void k_bin_sum(const float* __restrict__ S_part, const float* __restrict__ p_part,
                                                const float* __restrict__ P_part, const int* __restrict__ bin_items,
                                                int T, int lm, double* __restrict__ Sbin, double* __restrict__ pbin,
                                                float* __restrict__ Pbin) {
  constexpr int n = 1 << LOGT;
  extern __shared__ float2 lds[];
  float* red = reinterpret_cast<float*>(lds);   // 2 x (NT / 64)
  const int tid = threadIdx.x, b = blockIdx.x;
  const int i0 = bin_items[b], i1 = bin_items[b + 1];
  const int fb = n / NT;
  const float* src;
  long stride;
  int idx, lim;
  if ((int)blockIdx.y < fb) {
    idx = blockIdx.y * NT + tid;
    src = S_part + idx;
    stride = n;
    lim = n;
  } else {
    idx = (blockIdx.y - fb) * NT + tid;
    src = p_part + idx;
    stride = T;
    lim = T;
  }
  if (idx < lim) {
    float s = 0.f, c = 0.f;
    int it = i0;
    for (; it + 4 <= i1; it += 4) {
      const float a0 = src[(long)it * stride], a1 = src[(long)(it + 1) * stride];
      const float a2 = src[(long)(it + 2) * stride], a3 = src[(long)(it + 3) * stride];
      two_sum_add(s, c, a0);
      two_sum_add(s, c, a1);
      two_sum_add(s, c, a2);
      two_sum_add(s, c, a3);
    }
    for (; it < i1; it++) two_sum_add(s, c, src[(long)it * stride]);
    if ((int)blockIdx.y < fb) {
      const int k = ((idx & ((n >> lm) - 1)) << lm) | (idx >> (LOGT - lm));
      Sbin[(long)b * n + k] = (double)s + (double)c;
    } else {
      pbin[(long)b * T + idx] = (double)s + (double)c;
    }
  }
  if (blockIdx.y == 0) {
    float ps = 0.f, pc = 0.f;
    for (int it = i0 + tid; it < i1; it += NT) two_sum_add(ps, pc, P_part[it]);
    const float2 Ptot = block_sum2(make_float2(ps + pc, 0.f), red, tid);  // barriers inside
    if (tid == 0) Pbin[b] = Ptot.x;
  }
}

// FP64 block sum (NT threads; every thread gets the result; red: NT/64 doubles).
__device__ __forceinline__ double block_sum_d(double v, double* red, int tid) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((tid & 63) == 0) red[tid >> 6] = v;
  __syncthreads();
  double r = 0.0;
#pragma unroll
  for (int i = 0; i < NT / 64; i++) r += red[i];
  __syncthreads();
  return r;
}

// k_bin_ifft: one workgroup per bin, in FP64.
//   transform lags tau < T - KT: Gd(tau) = Re G(tau) - G(0) =
//       [IFFT(S)(tau) - IFFT(S)(0)] / n = -(1/n) sum_f S(f) (1 - cos(2 pi f tau / n)),
//       S = sum |Y|^2 over the bin's modes: a positively weighted sum of S,
//       whose FP32 rounding stays relative to D itself instead of to 2 G(0)
//       (slow bins at short lags have D << 2 P, where an FP32 transform of S
//       loses the digits).  The lag kernels form
//       D N(tau) = (A + B - 2 sum p) - 2 Gd, whose first term holds only the
//       samples a lag leaves unpaired (FP64 sums of the bin's p);
//   tail lags tau >= T - KT: D N(tau) itself, the items' direct sums.
template <int LOGT>
__global__ __launch_bounds__(NT) void k_bin_ifft(const double* __restrict__ Sbin, int L,
                                                 const double2* __restrict__ tw64, const float* __restrict__ G_part,
                                                 const int* __restrict__ bin_items, int T, int KT,
                                                 double* __restrict__ Gbin) {
  constexpr int n = 1 << LOGT;
  extern __shared__ float2 lds[];
  double2* buf = reinterpret_cast<double2*>(lds);
  const int tid = threadIdx.x, b = blockIdx.x;
  for (int i = tid; i < n; i += NT) buf[i] = make_double2(Sbin[(long)b * n + i], 0.0);
  __syncthreads();
  lds_fft64<LOGT, NT, true>(buf, tw64, tid);
  const double g0 = buf[0].x, invn = 1.0 / n;
  const int i0 = bin_items[b], i1 = bin_items[b + 1];
  for (int tau = tid; tau < L; tau += NT) {
    double v;
    if (tau >= T - KT) {   // tail lags: D N(tau) from the temporal kernel's direct sums
      v = 0.0;
      for (int it = i0; it < i1; it++) v += (double)G_part[(long)it * KT + tau - (T - KT)];
    } else {
      v = tau == 0 ? 0.0 : (buf[tau].x - g0) * invn;
    }
    Gbin[(long)b * L + tau] = v;
  }
}

// k_bin_lags_full: every frame valid (m = 1 on [0, T)).  Then
// A(tau) + B(tau) - 2 sum_t p(t) = -S(tau) - (S(T) - S(T - tau))
// with S the exclusive prefix sum of the bin's p, scanned in FP64 (one
// workgroup per bin); D = (A + B - 2 sum p - 2 Gd) / N(tau) / nModes, and
// D = DN / N(tau) / nModes at the directly summed tail lags tau >= T - KT.
__global__ __launch_bounds__(NT) void k_bin_lags_full(const double* __restrict__ Gbin, const double* __restrict__ pbin,
                                                      const float* __restrict__ Pbin,
                                                      const uint32_t* __restrict__ nmodes,
                                                      const float* __restrict__ npairs, int T, int L, int KT, float n0,
                                                      float* __restrict__ radD, float* __restrict__ radP) {
  extern __shared__ float2 lds[];
  double* S = reinterpret_cast<double*>(lds);   // T + 1
  double* wsum = S + T + 1;                      // NT / 64
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, b = blockIdx.x;
  // blocked exclusive scan: thread tid owns the contiguous run [c0, c1)
  const int per = (T + NT - 1) / NT;
  const int c0 = min(T, tid * per), c1 = min(T, c0 + per);
  double loc = 0;
  for (int t = c0; t < c1; t++) loc += pbin[(long)b * T + t];
  // inclusive scan of the per-thread totals across the block
  double v = loc;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const double u = __shfl_up(v, o, 64);
    if (lane >= o) v += u;
  }
  if (lane == 63) wsum[wave] = v;
  __syncthreads();
  double base = 0;
  for (int w = 0; w < wave; w++) base += wsum[w];
  double run = base + v - loc;                   // exclusive prefix at c0
  for (int t = c0; t < c1; t++) {
    S[t] = run;
    run += pbin[(long)b * T + t];
  }
  if (tid == NT - 1) S[T] = run;
  __syncthreads();
  const uint32_t nm = nmodes[b];
  for (int tau = tid; tau < L; tau += NT) {
    const double gd = Gbin[(long)b * L + tau];
    const double dn = tau >= T - KT ? gd : (-S[tau] - (S[T] - S[T - tau])) - 2.0 * gd;
    const float np = npairs[tau];
    float d = 0.f;
    if (np > 0.5f && nm > 0) d = (float)(dn / (double)np / (double)nm);
    radD[(long)b * L + tau] = d;
  }
  if (tid == 0) radP[b] = (nm > 0 && n0 > 0.5f) ? Pbin[b] / n0 / (float)nm : 0.f;
}

// k_bin_lags: any mask.  One workgroup per (bin, 64 lags): lane = lag, the
// four waves split t.  A(tau) + B(tau) = sum_t m(t) p(t + tau) + p(t) m(t + tau)
// in FP64 (m is 0/1), less 2 sum_t p(t); D = (A + B - 2 sum p - 2 Gd) /
// N(tau) / nModes, and D = DN / N(tau) / nModes at the directly summed tail
// lags tau >= T - KT.
__global__ __launch_bounds__(NT) void k_bin_lags(const double* __restrict__ Gbin, const double* __restrict__ pbin,
                                                 const float* __restrict__ Pbin, const uint32_t* __restrict__ nmodes,
                                                 const float* __restrict__ mask01, const float* __restrict__ npairs,
                                                 int T, int L, int KT, float n0, float* __restrict__ radD,
                                                 float* __restrict__ radP) {
  extern __shared__ float2 lds[];
  double* part = reinterpret_cast<double*>(lds);  // 4 waves x 64 lanes, then NT / 64 for the p sum
  double* red = part + NT;
  double* p = red + NT / 64;                             // T
  float* m = reinterpret_cast<float*>(p + T);            // T
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int b = blockIdx.x;
  const int tau = blockIdx.y * 64 + lane;
  double ps = 0.0;
  for (int t = tid; t < T; t += NT) {
    const double pv = pbin[(long)b * T + t];
    p[t] = pv;
    m[t] = mask01[t];
    ps += pv;
  }
  const double psum = block_sum_d(ps, red, tid);   // barriers inside (p, m staged)
  double s = 0.0;
  if (tau < L) {
    const int tend = T - tau;                 // t + tau < T
    const int q = (tend + 3) / 4;
    const int t0 = wave * q, t1 = min(tend, t0 + q);
    for (int t = t0; t < t1; t++) s += (double)m[t] * p[t + tau] + p[t] * (double)m[t + tau];
  }
  part[wave * 64 + lane] = s;
  __syncthreads();
  if (wave == 0 && tau < L) {
    const double ab = (part[lane] + part[64 + lane]) + (part[128 + lane] + part[192 + lane]);
    const float np = npairs[tau];
    const uint32_t nm = nmodes[b];
    const double gd = Gbin[(long)b * L + tau];
    const double dn = tau >= T - KT ? gd : (ab - 2.0 * psum) - 2.0 * gd;
    float d = 0.f;
    if (np > 0.5f && nm > 0) d = (float)(dn / (double)np / (double)nm);
    radD[(long)b * L + tau] = d;
  }
  if (tid == 0 && blockIdx.y == 0) {
    const uint32_t nm = nmodes[b];
    radP[b] = (nm > 0 && n0 > 0.5f) ? Pbin[b] / n0 / (float)nm : 0.f;
  }
}

// ----------------------------------------------------------------- launchers

template <int LOGT>
static hipError_t launch_temporal_t(const float2* hist, int T, int hs, const int4* items, int nitems, const int* bin_modes,
                                    const uint32_t* order, const float* mask, int identity, float inv_n0,
                                    const float2* tw, float* S_part, float* p_part, float* P_part, int KT,
                                    float* G_part, hipStream_t st) {
  constexpr int n = 1 << LOGT;
  size_t lds = sizeof(float2) * 2 * n + sizeof(float) * 2 * (NT / 64);
  auto kern = k_temporal<LOGT>;
  hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(kern, dim3(nitems), dim3(NT), lds, st, hist, T, hs, items, bin_modes, order, mask, identity, inv_n0,
                     tw, S_part, p_part, P_part, KT, G_part);
  return hipGetLastError();
}

hipError_t launch_temporal(int logt, const float2* hist, int T, int hs, const int4* items, int nitems, const int* bin_modes,
                           const uint32_t* order, const float* mask, int identity, float inv_n0, const float2* tw,
                           float* S_part, float* p_part, float* P_part, int KT, float* G_part, hipStream_t st) {
  if (KT < 1 || KT > NT / 4 || KT > T || hs < T) return hipErrorInvalidValue;
  switch (logt) {
#define CASE(LG_)                                                                                                    \
  case LG_:                                                                                                        \
    return launch_temporal_t<LG_>(hist, T, hs, items, nitems, bin_modes, order, mask, identity, inv_n0, tw, S_part, p_part, \
                                P_part, KT, G_part, st);
    CASE(8) CASE(9) CASE(10) CASE(11) CASE(12) CASE(13)
#undef CASE
  }
  return hipErrorInvalidValue;
}


template <int LOGT, int LM>
static size_t temporal_w_lds() {
  return sizeof(float2) * (4 * temporal_wave_lds(LOGT, LM) + 2 + twpad((1 << LOGT) >> LM));
}

template <int LOGT, int LM, int MAXLR>
static hipError_t launch_temporal_w_t(const float2* hist, int T, int hs, const int4* items, int nitems, const int* bin_modes,
                                      const uint32_t* order, const float* mask, int identity, float inv_n0,
                                      const float2* tw, const float2* tw_f, float* S_part, float* p_part,
                                      float* P_part, int KT, float* G_part, hipStream_t st) {
  const size_t lds = temporal_w_lds<LOGT, LM>();
  auto kern = k_temporal_w<LOGT, LM, MAXLR>;
  hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(kern, dim3(nitems), dim3(NT), lds, st, hist, T, hs, items, bin_modes, order, mask, identity, inv_n0,
                     tw, tw_f, S_part, p_part, P_part, KT, G_part);
  return hipGetLastError();
}

__global__ void k_fill_int(int* __restrict__ p, int v) { p[threadIdx.x] = v; }

template <int LOGT, int LM, int MAXLR>
static hipError_t launch_temporal_wq_t(const TemporalArgs* vids, int nvid, int nwork, int* counter, const float2* tw,
                                       const float2* tw_f, int grid, hipStream_t st) {
  const size_t lds = temporal_w_lds<LOGT, LM>();
  auto kern = k_temporal_wq<LOGT, LM, MAXLR>;
  hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  // the queue's counter is cleared by a kernel (a memset node would not do
  // in a captured graph)
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(1), 0, st, counter, 0);
  hipLaunchKernelGGL(kern, dim3(grid), dim3(NT), lds, st, vids, nvid, nwork, counter, tw, tw_f);
  return hipGetLastError();
}

template <int LOGT, int LM, int MAXLR>
static int temporal_wq_grid_t() {
  const size_t lds = temporal_w_lds<LOGT, LM>();
  auto kern = k_temporal_wq<LOGT, LM, MAXLR>;
  int dev = 0, cus = 0, per_cu = 0;
  if (hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess ||
      hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)kern, NT, lds) != hipSuccess)
    return 0;
  return cus * (per_cu > 0 ? per_cu : 1);
}

// Workgroups of the persistent temporal kernel: as many as fit the device at
// once (0 on a query failure).
int temporal_wq_grid(int logt, int maxlr) {
  switch (logt * 4 + (maxlr == 2 ? 2 : 3)) {
#define CASE(LG_, LM_, R_) \
  case LG_ * 4 + R_: return temporal_wq_grid_t<LG_, LM_, R_>();
    CASE(8, 0, 2) CASE(9, 0, 2) CASE(10, 1, 2) CASE(11, 2, 2) CASE(12, 3, 2)
    CASE(8, 0, 3) CASE(9, 0, 3) CASE(10, 1, 3) CASE(11, 2, 3) CASE(12, 3, 3)
#undef CASE
  }
  return 0;
}

hipError_t launch_temporal_wq(int logt, int maxlr, const TemporalArgs* vids, int nvid, int nwork, int* counter,
                              const float2* tw, const float2* tw_f, int grid, hipStream_t st) {
  if (nvid < 1 || nwork < 1 || grid < 1) return hipErrorInvalidValue;
  switch (logt * 4 + (maxlr == 2 ? 2 : 3)) {
#define CASE(LG_, LM_, R_) \
  case LG_ * 4 + R_: return launch_temporal_wq_t<LG_, LM_, R_>(vids, nvid, nwork, counter, tw, tw_f, grid, st);
    CASE(8, 0, 2) CASE(9, 0, 2) CASE(10, 1, 2) CASE(11, 2, 2) CASE(12, 3, 2)
    CASE(8, 0, 3) CASE(9, 0, 3) CASE(10, 1, 3) CASE(11, 2, 3) CASE(12, 3, 3)
#undef CASE
  }
  return hipErrorInvalidValue;
}

// tw: exp(-2 pi i j/n) for the TPAD length n; tw_f: the same for the
// sub-transform length n/m (m = 1 for n <= 512, n/512 above); maxlr: the
// FFT radix cap (3: radix 8, 2: radix 4).
hipError_t launch_temporal_w(int logt, int maxlr, const float2* hist, int T, int hs, const int4* items, int nitems,
                             const int* bin_modes, const uint32_t* order, const float* mask, int identity,
                             float inv_n0, const float2* tw, const float2* tw_f, float* S_part, float* p_part,
                             float* P_part, int KT, float* G_part, hipStream_t st) {
  // KT tail lags: four lanes each, and their samples must fit the exchange buffer
  if (KT < 1 || KT > 64 || KT > T || hs < T || 2 * KT > xch_size(std::min(1 << logt, 512)))
    return hipErrorInvalidValue;
  switch (logt * 4 + (maxlr == 2 ? 2 : 3)) {
#define CASE(LG_, LM_, R_)                                                                                           \
  case LG_ * 4 + R_:                                                                                                \
    return launch_temporal_w_t<LG_, LM_, R_>(hist, T, hs, items, nitems, bin_modes, order, mask, identity, inv_n0, tw,  \
                                             tw_f, S_part, p_part, P_part, KT, G_part, st);
    CASE(8, 0, 2) CASE(9, 0, 2) CASE(10, 1, 2) CASE(11, 2, 2) CASE(12, 3, 2)
    CASE(8, 0, 3) CASE(9, 0, 3) CASE(10, 1, 3) CASE(11, 2, 3) CASE(12, 3, 3)
#undef CASE
  }
  return hipErrorInvalidValue;
}

template <int LOGT>
static hipError_t launch_temporal_full_t(const float2* hist, int T, int hs, int L, long m0, int nm, const uint32_t* order,
                                         const float* mask, int identity, float inv_n0, const float* npairs,
                                         const float2* tw, int KT, const int* mode_list, float* outD, float2* outC,
                                         float* outP, hipStream_t st) {
  constexpr int n = 1 << LOGT;
  size_t lds = sizeof(float2) * (2 * n + 64) + sizeof(float) * (2 * 64 + 2 * (NT / 64));
  auto kern = k_temporal_full<LOGT>;
  hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(kern, dim3(nm), dim3(NT), lds, st, hist, T, hs, L, m0, order, mask, identity, inv_n0, npairs, tw, KT,
                     mode_list, outD, outC, outP);
  return hipGetLastError();
}

hipError_t launch_temporal_full(int logt, const float2* hist, int T, int hs, int L, long m0, int nm, const uint32_t* order,
                                const float* mask, int identity, float inv_n0, const float* npairs, const float2* tw,
                                int KT, const int* mode_list, float* outD, float2* outC, float* outP, hipStream_t st) {
  if (KT < 1 || KT > 64 || KT > T || hs < T || nm < 1) return hipErrorInvalidValue;
  switch (logt) {
#define CASE(LG_)                                                                                                   \
  case LG_:                                                                                                       \
    return launch_temporal_full_t<LG_>(hist, T, hs, L, m0, nm, order, mask, identity, inv_n0, npairs, tw, KT,         \
                                       mode_list, outD, outC, outP, st);
    CASE(8) CASE(9) CASE(10) CASE(11) CASE(12) CASE(13)
#undef CASE
  }
  return hipErrorInvalidValue;
}

hipError_t launch_drift(const float2* C, int L, int c0, int nc, const int* bin_offsets, const int2* kxy, int b0,
                        int nb, double* acc, hipStream_t st) {
  if (L < 2 || nb < 1) return hipSuccess;
  hipLaunchKernelGGL(k_drift, dim3((L - 1 + NT - 1) / NT, nb), dim3(NT), 0, st, C, L, c0, nc, bin_offsets, kxy, b0, acc);
  return hipGetLastError();
}

template <int LOGT>
static hipError_t launch_bin_sum_ifft_t(const float* S_part, const float* p_part, const float* P_part,
                                       const int* bin_items, int NB, int T, int L, int lm, const double2* tw64,
                                       const float* G_part, int KT, double* Sbin, double* Gbin, double* pbin,
                                       float* Pbin, hipStream_t st) {
  constexpr int n = 1 << LOGT;
  hipLaunchKernelGGL(k_bin_sum<LOGT>, dim3(NB, n / NT + (T + NT - 1) / NT), dim3(NT), sizeof(float) * 2 * (NT / 64), st,
                     S_part, p_part, P_part, bin_items, T, lm, Sbin, pbin, Pbin);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  size_t lds = sizeof(double2) * n;
  auto kern = k_bin_ifft<LOGT>;
  e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(kern, dim3(NB), dim3(NT), lds, st, (const double*)Sbin, L, tw64, G_part, bin_items, T, KT,
                     Gbin);
  return hipGetLastError();
}

hipError_t launch_bin_finalize(int logt, int lm, const float* S_part, const float* p_part, const float* P_part,
                               const float* G_part, int KT, const int* bin_items, int NB, const uint32_t* nmodes,
                               const float* mask, int full, const float* npairs, int T, int L, float n0,
                               const double2* tw64, double* Sbin, double* Gbin, double* pbin, float* Pbin, float* radD,
                               float* radP, hipStream_t st) {
  hipError_t e = hipErrorInvalidValue;
  switch (logt) {
#define CASE(LG_)                                                                                                \
  case LG_:                                                                                                    \
    e = launch_bin_sum_ifft_t<LG_>(S_part, p_part, P_part, bin_items, NB, T, L, lm, tw64, G_part, KT, Sbin, Gbin, \
                                   pbin, Pbin, st);                                                            \
    break;
    CASE(8) CASE(9) CASE(10) CASE(11) CASE(12) CASE(13)
#undef CASE
  }
  if (e != hipSuccess) return e;
  if (full) {
    size_t lds = sizeof(double) * (T + 1 + NT / 64);
    e = hipFuncSetAttribute((const void*)k_bin_lags_full, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_bin_lags_full, dim3(NB), dim3(NT), lds, st, (const double*)Gbin, (const double*)pbin,
                       (const float*)Pbin, nmodes, npairs, T, L, KT, n0, radD, radP);
    return hipGetLastError();
  }
  size_t lds = sizeof(double) * (NT + NT / 64 + T) + sizeof(float) * T;
  e = hipFuncSetAttribute((const void*)k_bin_lags, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_bin_lags, dim3(NB, (L + 63) / 64), dim3(NT), lds, st, (const double*)Gbin, (const double*)pbin,
                     (const float*)Pbin, nmodes, mask, npairs, T, L, KT, n0, radD, radP);
  return hipGetLastError();
}

}  // namespace ddm
