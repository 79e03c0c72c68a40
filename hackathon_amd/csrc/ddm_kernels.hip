// HIP kernels of the DDM hot path for gfx950.
//
//  k_rows      frame stack -> row spectra.  Decode u8/u16/f32, exact per-row
//              mean removal, separable Blackman-Harris window, R2C FFT along x
//              (two real rows packed into one complex FFT), store kx-major so
//              the column pass reads contiguous columns.
//  k_cols      row spectra -> history.  Mean correction on the kx <= 3 columns
//              (turns the per-row means into the frame mean, see below), FFT
//              along y, x 1/N, corner turn into hist[(m*T + slot)] with F frames
//              per workgroup so every store run is F*8 contiguous bytes.
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
// Reference: src/ddm.wgsl (reduce_mean/fft_rows/fft_cols :115-192, temporal
// :202-259, radial :265-322) and its FP64 twin src/ddm_ref.js:229-288,360-412.
#include "ddm_internal.h"
#include "ddm_fft.h"

namespace ddm {

constexpr int NT = 256;

// Real rows per spatial-row workgroup = 2 * rows_pairs_n(N): about 32 KB of
// complex LDS, at most N rows, and 2B dividing N (a power of two B: for
// N = 3 * 2^m every power of two up to 2^(m-1) qualifies).
__host__ __device__ constexpr int rows_pairs_n(int n) {
  int b = 1;
  while (2 * b * n <= 4096 && 2 * b <= 64 && 4 * b <= n && n % (4 * b) == 0) b *= 2;
  return n >= 2048 ? 4 : b;
}
// lanes per row segment of the row-sum reduction: the largest power of two
// <= 64 dividing the row's vector loads
__host__ __device__ constexpr int seg_lanes(int nv) { return (nv & -nv) < 64 ? (nv & -nv) : 64; }
// frame sides the spatial kernels take: powers of two in [16, 2048] and
// 3 * 2^m in [48, 1536] (radix-3 first pass)
__host__ __device__ constexpr bool spatial_size_ok(int n) {
  return (n >= 16 && n <= 2048 && (n & (n - 1)) == 0) ||
         (n >= 48 && n <= 1536 && n % 3 == 0 && ((n / 3) & (n / 3 - 1)) == 0);
}

// float2 slots per wave of k_temporal_w: the sub-transform exchange buffer,
// then the wave's accumulators (n floats of sum |Y|^2, n/2 floats of sum p).
__host__ __device__ constexpr int temporal_wave_lds(int logt, int lm) {
  return xch_size(1 << (logt - lm)) + (3 << logt) / 4;
}

// ----------------------------------------------------------------- helpers

template <typename T> struct vec8b;
template <> struct vec8b<uint8_t> { static constexpr int n = 8; using type = uint2; };
template <> struct vec8b<uint16_t> { static constexpr int n = 4; using type = uint2; };
template <> struct vec8b<float> { static constexpr int n = 2; using type = uint2; };

template <typename T> __device__ __forceinline__ void unpack8(uint2 v, T* out);
template <> __device__ __forceinline__ void unpack8<uint8_t>(uint2 v, uint8_t* o) {
#pragma unroll
  for (int i = 0; i < 4; i++) {
    o[i] = (v.x >> (8 * i)) & 0xff;
    o[4 + i] = (v.y >> (8 * i)) & 0xff;
  }
}
template <> __device__ __forceinline__ void unpack8<uint16_t>(uint2 v, uint16_t* o) {
  o[0] = v.x & 0xffff;
  o[1] = v.x >> 16;
  o[2] = v.y & 0xffff;
  o[3] = v.y >> 16;
}
template <> __device__ __forceinline__ void unpack8<float>(uint2 v, float* o) {
  o[0] = __uint_as_float(v.x);
  o[1] = __uint_as_float(v.y);
}

// Sum over aligned segments of SEG lanes (SEG a power of two <= 64).
template <int SEG, typename T> __device__ __forceinline__ T seg_sum(T v) {
#pragma unroll
  for (int o = SEG / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// ----------------------------------------------------------------- k_rows
//
// Workgroup = RW = 2B consecutive rows of one frame.  Integer pixels: the row
// sum S_y is exact, and the value fed to the FFT is (I*N - S_y) * (s/N), i.e.
// the row mean is removed exactly.  f32 pixels: the row mean is removed as an
// unevaluated (hi, lo) pair, like reduce_mean does per frame (src/ddm.wgsl:129-150).
// rowinfo[y] keeps (S_y, 0) or (hi_y, lo_y) for the column pass.
template <int N, int B, typename Tin>
__global__ __launch_bounds__(NT) void k_rows(const Tin* __restrict__ frames, int nframes, float scale,
                                             const float* __restrict__ win, const float2* __restrict__ tw_g,
                                             float2* __restrict__ scratch, int2* __restrict__ rowinfo) {
  constexpr int NH = N / 2 + 1, RW = 2 * B;
  static_assert(N % RW == 0, "rows per workgroup must divide N");
  constexpr int VEC = vec8b<Tin>::n;              // pixels per 8-byte load
  constexpr int NV = RW * N / VEC;                // vector loads per workgroup
  constexpr int LPT = (NV + NT - 1) / NT;         // loads per thread
  constexpr int SEG = seg_lanes(N / VEC);         // lanes per row segment within a wave
  constexpr bool IS_INT = !std::is_same<Tin, float>::value;
  extern __shared__ float2 lds[];
  float2* buf = lds;                               // B * N
  float2* tw = lds + B * N;                        // N
  constexpr int NSEG = (N / VEC) / SEG;           // segment partials per row
  float* rowsum_f = reinterpret_cast<float*>(tw + N);      // RW (f32 path: hi sums)
  float* rowsum_lo = rowsum_f + RW;                        // RW (f32 path: residual sums)
  int* rowsum_i = reinterpret_cast<int*>(rowsum_lo + RW);  // RW (integer path)
  float* part_f = reinterpret_cast<float*>(rowsum_i + RW); // RW * NSEG
  int* part_i = reinterpret_cast<int*>(part_f + RW * NSEG);  // RW * NSEG
  const int tid = threadIdx.x;
  const int groups = N / RW;
  const int f = blockIdx.x / groups;
  const int y0 = (blockIdx.x % groups) * RW;
  if (f >= nframes) return;

  for (int i = tid; i < N; i += NT) tw[i] = tw_g[i];
  const uint2* src = reinterpret_cast<const uint2*>(frames + ((long)f * N + y0) * N);
  uint2 raw[LPT];
#pragma unroll
  for (int i = 0; i < LPT; i++) {
    int v = tid + i * NT;
    raw[i] = (NV % NT == 0 || v < NV) ? src[v] : make_uint2(0, 0);
  }
  // Row sums in a fixed order (deterministic): lane segments -> per-segment
  // partials in LDS -> one thread per row adds its partials.
#pragma unroll
  for (int i = 0; i < LPT; i++) {
    int v = tid + i * NT;
    Tin px[VEC];
    unpack8<Tin>(raw[i], px);
    if constexpr (IS_INT) {
      int s = 0;
#pragma unroll
      for (int q = 0; q < VEC; q++) s += px[q];
      s = seg_sum<SEG>(s);
      if ((NV % NT == 0 || v < NV) && (tid % SEG) == 0) part_i[(v / (N / VEC)) * NSEG + (v % (N / VEC)) / SEG] = s;
    } else {
      float s = 0.f;
#pragma unroll
      for (int q = 0; q < VEC; q++) s += px[q];
      s = seg_sum<SEG>(s);
      if ((NV % NT == 0 || v < NV) && (tid % SEG) == 0) part_f[(v / (N / VEC)) * NSEG + (v % (N / VEC)) / SEG] = s;
    }
  }
  __syncthreads();
  if (tid < RW) {
    if constexpr (IS_INT) {
      int s = 0;
      for (int q = 0; q < NSEG; q++) s += part_i[tid * NSEG + q];
      rowsum_i[tid] = s;
    } else {
      float s = 0.f;
      for (int q = 0; q < NSEG; q++) s += part_f[tid * NSEG + q];
      rowsum_f[tid] = s;
    }
  }
  __syncthreads();
  if constexpr (!IS_INT) {
    // second pass: residual about the first-pass mean (hi, lo pair)
#pragma unroll
    for (int i = 0; i < LPT; i++) {
      int v = tid + i * NT;
      float px[VEC];
      unpack8<float>(raw[i], px);
      int r = (v * VEC) / N;
      float hi = (NV % NT == 0 || v < NV) ? rowsum_f[r] * (1.0f / N) : 0.f;
      float s = 0.f;
#pragma unroll
      for (int q = 0; q < VEC; q++) s += px[q] - hi;
      s = seg_sum<SEG>(s);
      if ((NV % NT == 0 || v < NV) && (tid % SEG) == 0) part_f[r * NSEG + (v % (N / VEC)) / SEG] = s;
    }
    __syncthreads();
    if (tid < RW) {
      float s = 0.f;
      for (int q = 0; q < NSEG; q++) s += part_f[tid * NSEG + q];
      rowsum_lo[tid] = s;
    }
    __syncthreads();
  }
  // decode, subtract the row mean, window, pack rows (2p, 2p+1) as (re, im)
  float* bufs = reinterpret_cast<float*>(buf);
#pragma unroll
  for (int i = 0; i < LPT; i++) {
    int v = tid + i * NT;
    if (NV % NT == 0 || v < NV) {
      Tin px[VEC];
      unpack8<Tin>(raw[i], px);
      int e0 = v * VEC;
      int r = e0 / N;
      int x0 = e0 % N;
      float wy = win[y0 + r];
      float* dst = bufs + 2 * ((r >> 1) * N) + (r & 1);
#pragma unroll
      for (int q = 0; q < VEC; q++) {
        float d;
        if constexpr (IS_INT) {
          d = (float)((int)px[q] * N - rowsum_i[r]) * (scale / N);
        } else {
          float hi = rowsum_f[r] * (1.0f / N);
          float lo = rowsum_lo[r] * (1.0f / N);
          d = ((px[q] - hi) - lo) * scale;
        }
        dst[2 * (x0 + q)] = d * win[x0 + q] * wy;
      }
    }
  }
  if constexpr (IS_INT) {
    if (tid < RW) rowinfo[(long)f * N + y0 + tid] = make_int2(rowsum_i[tid], 0);
  } else {
    if (tid < RW)
      rowinfo[(long)f * N + y0 + tid] =
          make_int2(__float_as_int(rowsum_f[tid] * (1.0f / N)), __float_as_int(rowsum_lo[tid] * (1.0f / N)));
  }
  __syncthreads();
  lds_fft_any<N, B, NT, false>(buf, tw, tid);
  // untangle the packed real transforms, store kx-major: scratch[(f*NH + kx)*N + y]
  constexpr int NOUT = NH * RW;
  for (int e = tid; e < NOUT; e += NT) {
    int r = e % RW;
    int kx = e / RW;
    int p = r >> 1;
    float2 z = buf[p * N + kx];
    float2 zc = cconj(buf[p * N + (N - kx) % N]);
    float2 o;
    if (r & 1) {
      float2 u = csub(z, zc);
      o = make_float2(0.5f * u.y, -0.5f * u.x);
    } else {
      o = make_float2(0.5f * (z.x + zc.x), 0.5f * (z.y + zc.y));
    }
    scratch[((long)f * NH + kx) * N + y0 + r] = o;
  }
}

// ----------------------------------------------------------------- k_cols
//
// Mean correction.  The rows were centred on their own means c_y; the
// reference centres the whole frame on its mean mu (src/fft.js:152-160).  The
// difference (c_y - mu) is constant along each row, so after the row FFT it
// only touches the columns where the window's DFT Wx(kx) is nonzero, kx <= 3
// for the 4-term window: row spectrum += (c_y - mu) * s * Wx(kx) * w(y).
// Integer frames: c_y - mu = (N*S_y - sum_y S_y) / N^2, exact in int64.
template <int N, int F>
__global__ __launch_bounds__(NT) void k_cols(const float2* __restrict__ scratch, const int2* __restrict__ rowinfo,
                                             int nframes, int is_int, float scale, const float* __restrict__ win,
                                             const float2* __restrict__ tw_g, const float2* __restrict__ wx,
                                             float2* __restrict__ hist, int T, int first_slot) {
  constexpr int NH = N / 2 + 1;
  extern __shared__ float2 lds[];
  float2* buf = lds;                  // F * N
  float2* tw = lds + F * N;           // N
  float* mu = reinterpret_cast<float*>(tw + N);   // 4F: frame mean as (hi sum, its compensation, lo, pad)
  long long* tot = reinterpret_cast<long long*>(mu + 4 * F);  // F
  const int tid = threadIdx.x;
  const int kx = blockIdx.x;
  const int f0 = blockIdx.y * F;
  for (int i = tid; i < N; i += NT) tw[i] = tw_g[i];
  for (int e = tid; e < F * N; e += NT) {
    int f = e / N, y = e % N;
    buf[e] = (f0 + f < nframes) ? scratch[((long)(f0 + f) * NH + kx) * N + y] : make_float2(0.f, 0.f);
  }
  if (kx <= 3) {
    if (tid < F && f0 + tid < nframes) {
      const int2* ri = rowinfo + (long)(f0 + tid) * N;
      if (is_int) {
        long long s = 0;
        for (int y = 0; y < N; y++) s += ri[y].x;
        tot[tid] = s;
      } else {
        float sh = 0.f, ch = 0.f, sl = 0.f, cl = 0.f;
        for (int y = 0; y < N; y++) {
          comp_add(sh, ch, __int_as_float(ri[y].x));
          comp_add(sl, cl, __int_as_float(ri[y].y));
        }
        // the frame mean as m0 + (ch + r0)/N + (sl + cl)/N with m0 = sh/N
        // rounded and r0 = sh - N m0 its exact residual (0 when N is a power
        // of two), so the leading term's rounding does not reach the mean
        const float m0 = sh * (1.0f / N);
        const float r0 = fmaf(-m0, (float)N, sh);
        mu[tid] = m0;
        mu[F + tid] = (ch + r0) * (1.0f / N);
        mu[2 * F + tid] = (sl + cl) * (1.0f / N);
      }
    }
    __syncthreads();
    float2 w = wx[kx];
    for (int e = tid; e < F * N; e += NT) {
      int f = e / N, y = e % N;
      if (f0 + f >= nframes) continue;
      int2 ri = rowinfo[(long)(f0 + f) * N + y];
      float d;
      if (is_int) {
        long long num = (long long)N * ri.x - tot[f];
        d = (float)num * (scale / ((float)N * (float)N));
      } else {
        d = (((__int_as_float(ri.x) - mu[f]) - mu[F + f]) + (__int_as_float(ri.y) - mu[2 * F + f])) * scale;
      }
      float c = d * win[y];
      buf[e] = cadd(buf[e], make_float2(c * w.x, c * w.y));
    }
  }
  __syncthreads();
  lds_fft_any<N, F, NT, false>(buf, tw, tid);
  const float s = 1.0f / N;
  for (int e = tid; e < F * N; e += NT) {
    int f = e % F, ky = e / F;
    if (f0 + f >= nframes) continue;
    int slot = (first_slot + f0 + f) % T;
    float2 v = buf[f * N + ky];
    hist[((long)ky * NH + kx) * T + slot] = make_float2(v.x * s, v.y * s);
  }
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
__global__ __launch_bounds__(NT) void k_temporal(const float2* __restrict__ hist, int T, const int4* __restrict__ items,
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
    const float2* x = hist + mode * T;
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
template <int LOGT, int LM>
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
  if constexpr (LOGF == 9) butterfly_twiddles<8, false, true>(twl, lane, wlast);
}

// One work item (up to 64 modes of one bin) of one video's temporal stage;
// A.items[item] names it and its partials go to row `item` of A's outputs.
template <int LOGT, int LM>
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
  const int T = A.T, identity = A.identity, KT = A.KT;
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
  // undoes the order) and sum p(t) at [n, 3n/2).  A wave sums at most 16
  // modes of an item plainly (positive terms: relative error <= 15 ulp); the
  // items' partials are added with compensation in k_bin_sum.
  float* acc = reinterpret_cast<float*>(xch + xch_size(NF));
  for (int i = lane; i < n + n / 2; i += 64) acc[i] = 0.f;
  float aP = 0.f, cP = 0.f;
  const int lpl = KT <= 16 ? 4 : (KT <= 32 ? 2 : 1);   // lanes per tail lag (wave-uniform; KT <= 64)
  float Gt = 0.f, Gtc = 0.f;   // tail lag T - KT + lane / lpl (lanes lpl j, j < KT)
  const int rounds = (count + 3) / 4;
  // Series loads without guards: a wave past the item's last mode re-reads
  // that mode (its contributions are dropped through `active`); samples
  // t in [T, n/2) read the next mode's head or the n/2 - T spare samples the
  // history buffer carries behind its last mode (they exist only when the
  // series is not dense, and the mask multiply zeroes them).  One base
  // address with immediate offsets; guarded or clamped loads would cost a
  // branch region or an address pair per sample.
  // A non-identity order is applied by the loads themselves: sample t is
  // gathered from slot order[t] (the n/2 entries of order[] past T are 0,
  // and their samples are masked off).
  uint32_t ord[PT];
#pragma unroll
  for (int q = 0; q < PT; q++) ord[q] = identity ? (uint32_t)(lane + 64 * q) : order[lane + 64 * q];
  auto load_series = [&](float2 (&dst)[PT], int slot) {
    const float2* x = hist + (long)bin_modes[start + (slot < count ? slot : count - 1)] * T;
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
  constexpr bool PREFETCH = LOGT <= 11;
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
      if (lpl >= 2) g += __shfl_xor(g, 1, 64);
      if (lpl == 4) g += __shfl_xor(g, 2, 64);
      if (s4 == 0 && j < KT && active) two_sum_add(Gt, Gtc, g);
    }
    if constexpr (LM == 0) {
      float2 v[PF];
#pragma unroll
      for (int q = 0; q < PT; q++) v[q] = y[q];
#pragma unroll
      for (int q = PT; q < PF; q++) v[q] = make_float2(0.f, 0.f);
      wave_fft<LOGF, false, true, true>(v, xch, twl, lane);
      if (active) {
#pragma unroll
        for (int q = 0; q < PF; q++) acc[lane + 64 * q] += v[q].x * v[q].x + v[q].y * v[q].y;
#pragma unroll
        for (int q = 0; q < PT; q++) acc[n + lane + 64 * q] += y[q].x * y[q].x + y[q].y * y[q].y;
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
        if constexpr (LOGF == 9) wave_fft_wl<LOGF, false, false, true>(v, xch, twl, lane, wlast);
        else wave_fft<LOGF, false, false, true>(v, xch, twl, lane);
        if (active) {
#pragma unroll
          for (int q = 0; q < PF; q++) acc[r * NF + lane + 64 * q] += v[q].x * v[q].x + v[q].y * v[q].y;
        }
      }
      if (active) {
#pragma unroll
        for (int q = 0; q < PT; q++) acc[n + lane + 64 * q] += y[q].x * y[q].x + y[q].y * y[q].y;
      }
    }
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

template <int LOGT, int LM>
// TPAD = 4096: one workgroup per CU by LDS (the per-wave accumulators), so
// up to 512 registers per lane instead of spilling at 256
__global__ __launch_bounds__(NT, LOGT >= 12 ? 1 : 2) void k_temporal_w(const float2* __restrict__ hist, int T,
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
  temporal_w_setup<LOGT, LM>(lds, tw_g, tw_f, tbase, wlast);
  const TemporalArgs A{hist, items, bin_modes, order, mask, S_part, p_part, P_part, G_part, T, identity, KT, 0, inv_n0};
  temporal_w_item<LOGT, LM>(A, blockIdx.x, lds, tbase, wlast);
}

// The batch form: a persistent grid over the work items of many videos (one
// transform length), taken from a queue.  vids[v] holds video v's arguments
// and item0, the index of its first item in the queue (increasing in v);
// nwork items in all.  A workgroup takes the next index from *counter (zero
// before the launch) until the queue is empty -- every workgroup sees the
// same index, so all of them leave the loop -- and computes the item exactly
// as k_temporal_w does, so the partials are bit-identical.
template <int LOGT, int LM>
__global__ __launch_bounds__(NT, LOGT >= 12 ? 1 : 2) void k_temporal_wq(const TemporalArgs* __restrict__ vids,
                                                                       int nvid, int nwork, int* __restrict__ counter,
                                                                       const float2* __restrict__ tw_g,
                                                                       const float2* __restrict__ tw_f) {
  extern __shared__ float2 lds[];
  __shared__ int next;
  float2 tbase[(1 << LM) > 1 ? (1 << LM) : 1], wlast[8];
  temporal_w_setup<LOGT, LM>(lds, tw_g, tw_f, tbase, wlast);
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
    temporal_w_item<LOGT, LM>(A, w - A.item0, lds, tbase, wlast);
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
__global__ __launch_bounds__(NT) void k_temporal_full(const float2* __restrict__ hist, int T, int L, long m0,
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
  const float2* x = hist + mode * T;
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

// ----------------------------------------------------------------- k_rows_w / k_cols_w
//
// Wavefront versions of the spatial passes for N >= 512.  k_rows_w: each wave
// transforms two real rows at a time (packed as one complex FFT, canonical
// x = lane + 64 q), untangles them through its LDS exchange buffer and parks
// the kx values of its rows in an LDS tile; the tile [kx][16 rows] is stored
// with 16-row (128 B) runs.  k_cols_w: each wave transforms F/4 frames' column
// kx in registers; the results meet in an LDS tile [ky][F] that is stored as
// F-frame runs of the history (the corner turn).

template <typename T> __device__ __forceinline__ float px_to_float(T v) { return (float)v; }

template <int LOGN, typename Tin>
__global__ __launch_bounds__(NT) void k_rows_w(const Tin* __restrict__ frames, int nframes, float scale,
                                               const float* __restrict__ win, const float2* __restrict__ tw,
                                               float2* __restrict__ scratch, int2* __restrict__ rowinfo) {
  constexpr int N = 1 << LOGN, NH = N / 2 + 1, P = N / 64;
  constexpr int RW = 16;                       // rows per workgroup (4 per wave)
  constexpr int XS = xch_size(N);
  constexpr int TS = RW + 1;                   // padded tile row
  constexpr bool IS_INT = !std::is_same<Tin, float>::value;
  extern __shared__ float2 lds[];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  float2* xch = lds + wave * XS;
  float2* tile = lds + 4 * XS;                 // NH x TS
  const int groups = N / RW;
  const int f = blockIdx.x / groups;
  const int y0 = (blockIdx.x % groups) * RW;
  if (f >= nframes) return;
#pragma unroll 1
  for (int pr = 0; pr < 2; pr++) {
    const int ra = wave * 4 + 2 * pr;          // rows ra, ra+1 of the tile
    const Tin* rowa = frames + ((long)f * N + y0 + ra) * N;
    const Tin* rowb = rowa + N;
    float xa[P], xb[P];
#pragma unroll
    for (int q = 0; q < P; q++) {
      // frames are read once: non-temporal loads
      xa[q] = px_to_float(__builtin_nontemporal_load(&rowa[lane + 64 * q]));
      xb[q] = px_to_float(__builtin_nontemporal_load(&rowb[lane + 64 * q]));
    }
    float ha = 0.f, hb = 0.f, la = 0.f, lb = 0.f;
    int Sa = 0, Sb = 0;
    if constexpr (IS_INT) {
      // exact integer row sums
      int ia = 0, ib = 0;
#pragma unroll
      for (int q = 0; q < P; q++) {
        ia += (int)xa[q];
        ib += (int)xb[q];
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        ia += __shfl_xor(ia, o, 64);
        ib += __shfl_xor(ib, o, 64);
      }
      Sa = ia;
      Sb = ib;
    } else {
      // row mean as an unevaluated (hi, lo) pair: hi from the sum, lo from the residuals
      float sa = 0.f, sb = 0.f;
#pragma unroll
      for (int q = 0; q < P; q++) {
        sa += xa[q];
        sb += xb[q];
      }
      ha = wave_sum(sa) * (1.0f / N);
      hb = wave_sum(sb) * (1.0f / N);
      float ra2 = 0.f, rb2 = 0.f;
#pragma unroll
      for (int q = 0; q < P; q++) {
        ra2 += xa[q] - ha;
        rb2 += xb[q] - hb;
      }
      la = wave_sum(ra2) * (1.0f / N);
      lb = wave_sum(rb2) * (1.0f / N);
    }
    const float wya = win[y0 + ra], wyb = win[y0 + ra + 1];
    float2 v[P];
#pragma unroll
    for (int q = 0; q < P; q++) {
      const int x = lane + 64 * q;
      const float wx = win[x];
      float da, db;
      if constexpr (IS_INT) {
        da = (float)((int)xa[q] * N - Sa) * (scale / N);
        db = (float)((int)xb[q] * N - Sb) * (scale / N);
      } else {
        da = ((xa[q] - ha) - la) * scale;
        db = ((xb[q] - hb) - lb) * scale;
      }
      v[q] = make_float2(da * wx * wya, db * wx * wyb);
    }
    if (lane == 0) {
      if constexpr (IS_INT) {
        rowinfo[(long)f * N + y0 + ra] = make_int2(Sa, 0);
        rowinfo[(long)f * N + y0 + ra + 1] = make_int2(Sb, 0);
      } else {
        rowinfo[(long)f * N + y0 + ra] = make_int2(__float_as_int(ha), __float_as_int(la));
        rowinfo[(long)f * N + y0 + ra + 1] = make_int2(__float_as_int(hb), __float_as_int(lb));
      }
    }
    wave_fft<LOGN, false, false>(v, xch, tw, lane);
    // untangle: X_a(k) = (Z(k) + conj Z(N-k))/2, X_b(k) = (Z(k) - conj Z(N-k))/(2i)
    wave_sync();
#pragma unroll
    for (int q = 0; q < P; q++) xch[xpad(lane + 64 * q)] = v[q];
    wave_sync();
#pragma unroll
    for (int q = 0; q <= P / 2; q++) {
      const int k = lane + 64 * q;
      if (k <= N / 2) {
        const float2 z = v[q];
        const float2 zc = cconj(xch[xpad((N - k) & (N - 1))]);
        const float2 u = csub(z, zc);
        tile[k * TS + ra] = make_float2(0.5f * (z.x + zc.x), 0.5f * (z.y + zc.y));
        tile[k * TS + ra + 1] = make_float2(0.5f * u.y, -0.5f * u.x);
      }
    }
  }
  __syncthreads();
  for (int e = tid; e < NH * RW; e += NT) {
    const int r = e % RW, kx = e / RW;
    scratch[((long)f * NH + kx) * N + y0 + r] = tile[kx * TS + r];
  }
}

template <int LOGN, int F>
__global__ __launch_bounds__(NT) void k_cols_w(const float2* __restrict__ scratch, const int2* __restrict__ rowinfo,
                                               int nframes, int is_int, float scale, const float* __restrict__ win,
                                               const float2* __restrict__ tw, const float2* __restrict__ wx,
                                               float2* __restrict__ hist, int T, int first_slot) {
  constexpr int N = 1 << LOGN, NH = N / 2 + 1, P = N / 64;
  constexpr int FW = F / 4;                    // frames per wave
  constexpr int XS = xch_size(N);
  constexpr int TS = F + 1;
  extern __shared__ float2 lds[];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  float2* xch = lds + wave * XS;
  float2* tile = lds;                          // (N/2) x TS, reused after every wave's FFTs
  const int kx = blockIdx.x;
  const int f0 = blockIdx.y * F;
  // N <= 512: the column FFT's twiddle table in LDS (twpad slots, behind the
  // tile/exchange area), read there instead of waiting on global loads; at
  // N = 1024 the extra 8.5 KB would cost a resident workgroup
  constexpr bool TWL = LOGN <= 9;
  float2* twl = lds + ((N / 2) * TS > 4 * XS ? (N / 2) * TS : 4 * XS);
  if constexpr (TWL) {
    for (int i = tid; i < N; i += NT) twl[twpad(i)] = tw[i];
    __syncthreads();
  }
  float2 res[FW][P];
#pragma unroll
  for (int i = 0; i < FW; i++) {
    const int fl = wave * FW + i;
    const int fr = f0 + fl;
    float2* v = res[i];
    if (fr < nframes) {
      const float2* col = scratch + ((long)fr * NH + kx) * N;
#pragma unroll
      for (int q = 0; q < P; q++) v[q] = col[lane + 64 * q];
      if (kx <= 3) {
        const int2* ri = rowinfo + (long)fr * N;
        const float2 w = wx[kx];
        if (is_int) {
          long long s = 0;
#pragma unroll
          for (int q = 0; q < P; q++) s += ri[lane + 64 * q].x;
#pragma unroll
          for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
#pragma unroll
          for (int q = 0; q < P; q++) {
            const int y = lane + 64 * q;
            const float d = (float)((long long)N * ri[y].x - s) * (scale / ((float)N * (float)N));
            const float c = d * win[y];
            v[q] = cadd(v[q], make_float2(c * w.x, c * w.y));
          }
        } else {
          float sh = 0.f, ch = 0.f, sl = 0.f, cl = 0.f;
#pragma unroll
          for (int q = 0; q < P; q++) {
            two_sum_add(sh, ch, __int_as_float(ri[lane + 64 * q].x));
            two_sum_add(sl, cl, __int_as_float(ri[lane + 64 * q].y));
          }
          // wave-wide compensated sums: reduce (s, c) pairs lane by lane
#pragma unroll
          for (int o = 32; o > 0; o >>= 1) {
            float osh = __shfl_xor(sh, o, 64), och = __shfl_xor(ch, o, 64);
            float osl = __shfl_xor(sl, o, 64), ocl = __shfl_xor(cl, o, 64);
            two_sum_add(sh, ch, osh);
            ch += och;
            two_sum_add(sl, cl, osl);
            cl += ocl;
          }
          const float mh = sh * (1.0f / N), mc = ch * (1.0f / N), ml = (sl + cl) * (1.0f / N);
#pragma unroll
          for (int q = 0; q < P; q++) {
            const int y = lane + 64 * q;
            const float d = (((__int_as_float(ri[y].x) - mh) - mc) + (__int_as_float(ri[y].y) - ml)) * scale;
            const float c = d * win[y];
            v[q] = cadd(v[q], make_float2(c * w.x, c * w.y));
          }
        }
      }
      wave_fft<LOGN, false, false, TWL>(res[i], xch, TWL ? twl : tw, lane);
    }
  }
  // The corner turn in two halves of ky: a half-size tile lets four
  // workgroups share a CU (the full N x (F+1) tile allowed two).
  const float sc = 1.0f / N;
#pragma unroll
  for (int h = 0; h < 2; h++) {
    __syncthreads();   // h = 0: every wave is done with its exchange buffer; 1: the first half is stored
#pragma unroll
    for (int i = 0; i < FW; i++) {
      const int fl = wave * FW + i;
#pragma unroll
      for (int q = h * P / 2; q < (h + 1) * P / 2; q++) {
        const float2 z = res[i][q];
        tile[(lane + 64 * (q - h * P / 2)) * TS + fl] = make_float2(z.x * sc, z.y * sc);
      }
    }
    __syncthreads();
    for (int e = tid; e < (N / 2) * F; e += NT) {
      const int fl = e % F, ky = h * (N / 2) + e / F;
      if (f0 + fl >= nframes) continue;
      const int slot = (first_slot + f0 + fl) % T;
      // non-temporal: the history is read back only by the temporal pass, after
      // the whole window; keeping it out of the caches leaves them to the row spectra
      const float2 z = tile[(e / F) * TS + fl];
      float2* dst = &hist[((long)ky * NH + kx) * T + slot];
      __builtin_nontemporal_store(z.x, &dst->x);
      __builtin_nontemporal_store(z.y, &dst->y);
    }
  }
}

// ----------------------------------------------------------------- launchers

template <typename Tin, int N>
static hipError_t launch_rows_t(const void* frames, int nframes, float scale, const float* win, const float2* tw,
                                float2* scratch, int2* rowinfo, hipStream_t st) {
  constexpr int B = rows_pairs_n(N);
  constexpr int RW = 2 * B;
  static_assert(RW <= N && N % RW == 0, "rows per workgroup must divide N");
  constexpr int VEC = vec8b<Tin>::n;
  constexpr int SEG = seg_lanes(N / VEC);
  constexpr int NSEG = (N / VEC) / SEG;
  size_t lds = sizeof(float2) * (B * N + N) + sizeof(float) * 2 * RW + sizeof(int) * RW +
               (sizeof(float) + sizeof(int)) * RW * NSEG;
  auto kern = k_rows<N, B, Tin>;
  hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  dim3 grid(nframes * (N / RW));
  hipLaunchKernelGGL(kern, grid, dim3(NT), lds, st, (const Tin*)frames, nframes, scale, win, tw, scratch, rowinfo);
  return hipGetLastError();
}

template <typename Tin>
static hipError_t launch_rows_dt(int n, const void* frames, int nframes, float scale, const float* win,
                                 const float2* tw, float2* scratch, int2* rowinfo, hipStream_t st) {
  switch (n) {
#define CASE(N_) \
  case N_: return launch_rows_t<Tin, N_>(frames, nframes, scale, win, tw, scratch, rowinfo, st);
    CASE(16) CASE(32) CASE(64) CASE(128) CASE(256) CASE(512) CASE(1024) CASE(2048)
    CASE(48) CASE(96) CASE(192) CASE(384) CASE(768) CASE(1536)
#undef CASE
  }
  return hipErrorInvalidValue;
}

hipError_t launch_rows(int dtype, int n, const void* frames, int nframes, float scale, const float* win,
                       const float2* tw, float2* scratch, int2* rowinfo, hipStream_t st) {
  if (dtype == 0) return launch_rows_dt<uint8_t>(n, frames, nframes, scale, win, tw, scratch, rowinfo, st);
  if (dtype == 1) return launch_rows_dt<uint16_t>(n, frames, nframes, scale, win, tw, scratch, rowinfo, st);
  return launch_rows_dt<float>(n, frames, nframes, scale, win, tw, scratch, rowinfo, st);
}

// Frames per column workgroup: the corner turn writes runs of this many slots
// per mode.  At N = 2048 eight (64 B runs, one 147 KB workgroup per CU)
// measured 20 % faster than four (32 B runs, two per CU).
int cols_frames_per_wg(int n) { return n >= 1024 ? 8 : 16; }

template <int N>
static hipError_t launch_cols_t(const float2* scratch, const int2* rowinfo, int nframes, int is_int, float scale,
                                const float* win, const float2* tw, const float2* wx, float2* hist, int T,
                                int first_slot, hipStream_t st) {
  constexpr int NH = N / 2 + 1;
  constexpr int F = N >= 1024 ? 8 : 16;
  size_t lds = sizeof(float2) * (F * N + N) + sizeof(float) * 4 * F + sizeof(long long) * F;
  auto kern = k_cols<N, F>;
  hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  dim3 grid(NH, (nframes + F - 1) / F);
  hipLaunchKernelGGL(kern, grid, dim3(NT), lds, st, scratch, rowinfo, nframes, is_int, scale, win, tw, wx, hist, T,
                     first_slot);
  return hipGetLastError();
}

hipError_t launch_cols(int n, const float2* scratch, const int2* rowinfo, int nframes, int is_int, float scale,
                       const float* win, const float2* tw, const float2* wx, float2* hist, int T, int first_slot,
                       hipStream_t st) {
  switch (n) {
#define CASE(N_) \
  case N_: return launch_cols_t<N_>(scratch, rowinfo, nframes, is_int, scale, win, tw, wx, hist, T, first_slot, st);
    CASE(16) CASE(32) CASE(64) CASE(128) CASE(256) CASE(512) CASE(1024) CASE(2048)
    CASE(48) CASE(96) CASE(192) CASE(384) CASE(768) CASE(1536)
#undef CASE
  }
  return hipErrorInvalidValue;
}

template <int LOGT>
static hipError_t launch_temporal_t(const float2* hist, int T, const int4* items, int nitems, const int* bin_modes,
                                    const uint32_t* order, const float* mask, int identity, float inv_n0,
                                    const float2* tw, float* S_part, float* p_part, float* P_part, int KT,
                                    float* G_part, hipStream_t st) {
  constexpr int n = 1 << LOGT;
  size_t lds = sizeof(float2) * 2 * n + sizeof(float) * 2 * (NT / 64);
  auto kern = k_temporal<LOGT>;
  hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(kern, dim3(nitems), dim3(NT), lds, st, hist, T, items, bin_modes, order, mask, identity, inv_n0,
                     tw, S_part, p_part, P_part, KT, G_part);
  return hipGetLastError();
}

hipError_t launch_temporal(int logt, const float2* hist, int T, const int4* items, int nitems, const int* bin_modes,
                           const uint32_t* order, const float* mask, int identity, float inv_n0, const float2* tw,
                           float* S_part, float* p_part, float* P_part, int KT, float* G_part, hipStream_t st) {
  if (KT < 1 || KT > NT / 4 || KT > T) return hipErrorInvalidValue;
  switch (logt) {
#define CASE(LG_)                                                                                                    \
  case LG_:                                                                                                        \
    return launch_temporal_t<LG_>(hist, T, items, nitems, bin_modes, order, mask, identity, inv_n0, tw, S_part, p_part, \
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

template <int LOGT, int LM>
static hipError_t launch_temporal_w_t(const float2* hist, int T, const int4* items, int nitems, const int* bin_modes,
                                      const uint32_t* order, const float* mask, int identity, float inv_n0,
                                      const float2* tw, const float2* tw_f, float* S_part, float* p_part,
                                      float* P_part, int KT, float* G_part, hipStream_t st) {
  const size_t lds = temporal_w_lds<LOGT, LM>();
  auto kern = k_temporal_w<LOGT, LM>;
  hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(kern, dim3(nitems), dim3(NT), lds, st, hist, T, items, bin_modes, order, mask, identity, inv_n0,
                     tw, tw_f, S_part, p_part, P_part, KT, G_part);
  return hipGetLastError();
}

__global__ void k_fill_int(int* __restrict__ p, int v) { p[threadIdx.x] = v; }

template <int LOGT, int LM>
static hipError_t launch_temporal_wq_t(const TemporalArgs* vids, int nvid, int nwork, int* counter, const float2* tw,
                                       const float2* tw_f, int grid, hipStream_t st) {
  const size_t lds = temporal_w_lds<LOGT, LM>();
  auto kern = k_temporal_wq<LOGT, LM>;
  hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  // the queue's counter is cleared by a kernel (a memset node would not do
  // in a captured graph)
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(1), 0, st, counter, 0);
  hipLaunchKernelGGL(kern, dim3(grid), dim3(NT), lds, st, vids, nvid, nwork, counter, tw, tw_f);
  return hipGetLastError();
}

template <int LOGT, int LM>
static int temporal_wq_grid_t() {
  const size_t lds = temporal_w_lds<LOGT, LM>();
  auto kern = k_temporal_wq<LOGT, LM>;
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
int temporal_wq_grid(int logt) {
  switch (logt) {
    case 8: return temporal_wq_grid_t<8, 0>();
    case 9: return temporal_wq_grid_t<9, 0>();
    case 10: return temporal_wq_grid_t<10, 1>();
    case 11: return temporal_wq_grid_t<11, 2>();
    case 12: return temporal_wq_grid_t<12, 3>();
  }
  return 0;
}

hipError_t launch_temporal_wq(int logt, const TemporalArgs* vids, int nvid, int nwork, int* counter, const float2* tw,
                              const float2* tw_f, int grid, hipStream_t st) {
  if (nvid < 1 || nwork < 1 || grid < 1) return hipErrorInvalidValue;
  switch (logt) {
#define CASE(LG_, LM_) \
  case LG_: return launch_temporal_wq_t<LG_, LM_>(vids, nvid, nwork, counter, tw, tw_f, grid, st);
    CASE(8, 0) CASE(9, 0) CASE(10, 1) CASE(11, 2) CASE(12, 3)
#undef CASE
  }
  return hipErrorInvalidValue;
}

// tw: exp(-2 pi i j/n) for the TPAD length n; tw_f: the same for the
// sub-transform length n/m (m = 1 for n <= 512, n/512 above).
hipError_t launch_temporal_w(int logt, const float2* hist, int T, const int4* items, int nitems, const int* bin_modes,
                             const uint32_t* order, const float* mask, int identity, float inv_n0, const float2* tw,
                             const float2* tw_f, float* S_part, float* p_part, float* P_part, int KT, float* G_part,
                             hipStream_t st) {
  // KT tail lags: four lanes each, and their samples must fit the exchange buffer
  if (KT < 1 || KT > 64 || KT > T || 2 * KT > xch_size(std::min(1 << logt, 512))) return hipErrorInvalidValue;
  switch (logt) {
#define CASE(LG_, LM_)                                                                                              \
  case LG_:                                                                                                         \
    return launch_temporal_w_t<LG_, LM_>(hist, T, items, nitems, bin_modes, order, mask, identity, inv_n0, tw, tw_f, \
                                         S_part, p_part, P_part, KT, G_part, st);
    CASE(8, 0) CASE(9, 0) CASE(10, 1) CASE(11, 2) CASE(12, 3)
#undef CASE
  }
  return hipErrorInvalidValue;
}

template <int LOGT>
static hipError_t launch_temporal_full_t(const float2* hist, int T, int L, long m0, int nm, const uint32_t* order,
                                         const float* mask, int identity, float inv_n0, const float* npairs,
                                         const float2* tw, int KT, const int* mode_list, float* outD, float2* outC,
                                         float* outP, hipStream_t st) {
  constexpr int n = 1 << LOGT;
  size_t lds = sizeof(float2) * (2 * n + 64) + sizeof(float) * (2 * 64 + 2 * (NT / 64));
  auto kern = k_temporal_full<LOGT>;
  hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(kern, dim3(nm), dim3(NT), lds, st, hist, T, L, m0, order, mask, identity, inv_n0, npairs, tw, KT,
                     mode_list, outD, outC, outP);
  return hipGetLastError();
}

hipError_t launch_temporal_full(int logt, const float2* hist, int T, int L, long m0, int nm, const uint32_t* order,
                                const float* mask, int identity, float inv_n0, const float* npairs, const float2* tw,
                                int KT, const int* mode_list, float* outD, float2* outC, float* outP, hipStream_t st) {
  if (KT < 1 || KT > 64 || KT > T || nm < 1) return hipErrorInvalidValue;
  switch (logt) {
#define CASE(LG_)                                                                                                   \
  case LG_:                                                                                                       \
    return launch_temporal_full_t<LG_>(hist, T, L, m0, nm, order, mask, identity, inv_n0, npairs, tw, KT,         \
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

template <typename Tin, int LOGN>
static hipError_t launch_rows_w_t(const void* frames, int nframes, float scale, const float* win, const float2* tw,
                                  float2* scratch, int2* rowinfo, hipStream_t st) {
  constexpr int N = 1 << LOGN, NH = N / 2 + 1;
  size_t lds = sizeof(float2) * (4 * xch_size(N) + NH * 17);
  auto kern = k_rows_w<LOGN, Tin>;
  hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(kern, dim3(nframes * (N / 16)), dim3(NT), lds, st, (const Tin*)frames, nframes, scale, win, tw,
                     scratch, rowinfo);
  return hipGetLastError();
}

template <int LOGN, int F>
static hipError_t launch_cols_w_t(const float2* scratch, const int2* rowinfo, int nframes, int is_int, float scale,
                                  const float* win, const float2* tw, const float2* wx, float2* hist, int T,
                                  int first_slot, hipStream_t st) {
  constexpr int N = 1 << LOGN, NH = N / 2 + 1;
  size_t tile = sizeof(float2) * (N / 2) * (F + 1), xch = sizeof(float2) * 4 * xch_size(N);
  size_t lds = (tile > xch ? tile : xch) + (LOGN <= 9 ? sizeof(float2) * twpad(N) : 0);
  auto kern = k_cols_w<LOGN, F>;
  hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(kern, dim3(NH, (nframes + F - 1) / F), dim3(NT), lds, st, scratch, rowinfo, nframes, is_int, scale,
                     win, tw, wx, hist, T, first_slot);
  return hipGetLastError();
}

// Wavefront spatial kernels: N = 512 and 1024 (registers hold N/64 points
// per lane); F = frames per column workgroup (8 or 16).
hipError_t launch_rows_w(int dtype, int logn, const void* frames, int nframes, float scale, const float* win,
                         const float2* tw, float2* scratch, int2* rowinfo, hipStream_t st) {
  if (logn == 9) {
    return dtype == 0 ? launch_rows_w_t<uint8_t, 9>(frames, nframes, scale, win, tw, scratch, rowinfo, st)
           : dtype == 1 ? launch_rows_w_t<uint16_t, 9>(frames, nframes, scale, win, tw, scratch, rowinfo, st)
                        : launch_rows_w_t<float, 9>(frames, nframes, scale, win, tw, scratch, rowinfo, st);
  }
  if (logn == 10) {
    return dtype == 0 ? launch_rows_w_t<uint8_t, 10>(frames, nframes, scale, win, tw, scratch, rowinfo, st)
           : dtype == 1 ? launch_rows_w_t<uint16_t, 10>(frames, nframes, scale, win, tw, scratch, rowinfo, st)
                        : launch_rows_w_t<float, 10>(frames, nframes, scale, win, tw, scratch, rowinfo, st);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_cols_w(int logn, int F, const float2* scratch, const int2* rowinfo, int nframes, int is_int,
                         float scale, const float* win, const float2* tw, const float2* wx, float2* hist, int T,
                         int first_slot, hipStream_t st) {
  if (logn == 9 && F == 16)
    return launch_cols_w_t<9, 16>(scratch, rowinfo, nframes, is_int, scale, win, tw, wx, hist, T, first_slot, st);
  if (logn == 9 && F == 8)
    return launch_cols_w_t<9, 8>(scratch, rowinfo, nframes, is_int, scale, win, tw, wx, hist, T, first_slot, st);
  if (logn == 10 && F == 8)
    return launch_cols_w_t<10, 8>(scratch, rowinfo, nframes, is_int, scale, win, tw, wx, hist, T, first_slot, st);
  if (logn == 10 && F == 16)
    return launch_cols_w_t<10, 16>(scratch, rowinfo, nframes, is_int, scale, win, tw, wx, hist, T, first_slot, st);
  return hipErrorInvalidValue;
}

}  // namespace ddm
