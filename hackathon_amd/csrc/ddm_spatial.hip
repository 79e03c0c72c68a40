// Spatial-stage HIP kernels of the DDM hot path for gfx950: the windowed R2C
// 2D FFT of each frame and the corner turn into the spectral history.
//
//  k_rows      frame stack -> row spectra.  Decode u8/u16/f32, exact per-row
//              mean removal, separable Blackman-Harris window, R2C FFT along x
//              (two real rows packed into one complex FFT), store kx-major so
//              the column pass reads contiguous columns.
//  k_cols      row spectra -> history.  Mean correction on the kx <= 3 columns
//              (turns the per-row means into the frame mean, see below), FFT
//              along y, x 1/N, corner turn into hist[m*hs + slot] (rows of hs =
//              T rounded up to 16 slots) with F frames per workgroup so every
//              store run is F*8 contiguous bytes, line-aligned from slot 16j.
//  k_rows_w / k_cols_w  the same passes with one wavefront per transform
//              (N = 512, 1024).
// Frame sides: powers of two, and 3 * 2^m through a radix-3 first pass.
//
// Reference: src/ddm.wgsl (reduce_mean/fft_rows/fft_cols :115-192) and its
// FP64 twin src/fft.js:149-183, src/ddm_ref.js:196-205.
#include "ddm_internal.h"
#include "ddm_fft.h"

namespace ddm {



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

// ----------------------------------------------------------------- k_rows
//
// Workgroup = RW = 2B consecutive rows of one frame.  Integer pixels: the row
// sum S_y is exact, and the value fed to the FFT is (I*N - S_y) * (s/N), i.e.
// the row mean is removed exactly.  f32 pixels: the row mean is removed as an
// unevaluated (hi, lo) pair, like reduce_mean does per frame (src/ddm.wgsl:129-150).
// rowinfo[y] keeps (S_y, 0) or (hi_y, lo_y) for the column pass.
template <int N, int B, typename Tin, int MAXLR>
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
  lds_fft_any<N, B, NT, false, MAXLR>(buf, tw, tid);
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
template <int N, int F, int MAXLR>
__global__ __launch_bounds__(NT) void k_cols(const float2* __restrict__ scratch, const int2* __restrict__ rowinfo,
                                             int nframes, int is_int, float scale, const float* __restrict__ win,
                                             const float2* __restrict__ tw_g, const float2* __restrict__ wx,
                                             float2* __restrict__ hist, int T, int hs, int first_slot) {
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
  lds_fft_any<N, F, NT, false, MAXLR>(buf, tw, tid);
  const float s = 1.0f / N;
  for (int e = tid; e < F * N; e += NT) {
    int f = e % F, ky = e / F;
    if (f0 + f >= nframes) continue;
    int slot = (first_slot + f0 + f) % T;
    float2 v = buf[f * N + ky];
    hist[((long)ky * NH + kx) * hs + slot] = make_float2(v.x * s, v.y * s);
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

// Row pairs per wave and frame in k_rows_w.  One at N = 512: 70 VGPRs and
// 36 KB of LDS per workgroup against 133 and 52 KB with two, so the other
// analysis's kernels find room beside it (two in flight: 1.129 -> 1.112 ms
// per 1024 x 512^2 step, the serial row pass unchanged); one at N = 2048,
// where the row-spectra tile shares the exchange buffers' LDS.
__host__ __device__ constexpr int rows_pairs_w(int logn) { return (logn == 9 || logn >= 11) ? 1 : 2; }

// XCD-aware block order for a kernel whose workgroups each store half of
// every 128-byte line they touch, the other half coming from a partner
// workgroup (logical blocks 2j and 2j + 1).  The dispatcher hands workgroup b
// to XCD b mod 8, so the partners go to physical blocks b and b + 8: the same
// XCD in consecutive rounds, where the two halves meet in that XCD's L2 and
// leave it as one full line instead of two partial-line writes.  Blocks at or
// past the last multiple of 16 keep their order.  A bijection of [0, G).
__device__ __forceinline__ int xcd_pair_block(int b, int G) {
  if (b >= (G & ~15)) return b;
  const int x = b & 7, r = b >> 3;
  return 2 * (x + 8 * (r >> 1)) + (r & 1);
}

// FPW frames per workgroup (the same rows of each): every wave transforms
// PPW row pairs per frame (rows_pairs_w: two at N = 256 and 1024, 16 rows per
// workgroup; one at N = 512 and 2048, 8 rows), and the next pair's pixels are
// loaded while the current pair is transformed (also across frames), so a
// wave keeps a load in flight through its whole run and its set-up
// (twiddles, LDS addresses) is shared by PPW FPW pairs.  At N = 2048 the kx-major tile shares its LDS with the
// exchange buffers: a wave keeps its untangled spectra in registers until
// every wave is done with its transform.
template <int LOGN, typename Tin, int MAXLR, int FPW>
__global__ __launch_bounds__(NT) void k_rows_w(const Tin* __restrict__ frames, int nframes, float scale,
                                               const float* __restrict__ win, const float2* __restrict__ tw,
                                               float2* __restrict__ scratch, int2* __restrict__ rowinfo) {
  constexpr int N = 1 << LOGN, NH = N / 2 + 1, P = N / 64;
  constexpr bool ALIAS = LOGN >= 11;           // tile and exchange buffers share the LDS
  constexpr int PPW = rows_pairs_w(LOGN);      // row pairs per wave and frame
  constexpr int RW = 8 * PPW;                  // rows per workgroup
  constexpr int XS = xch_size(N);
  constexpr int TS = RW + 1;                   // padded tile row
  constexpr int KQ = P / 2 + 1;                // untangled outputs per lane: k = lane + 64 q <= N/2
  constexpr bool IS_INT = !std::is_same<Tin, float>::value;
  extern __shared__ float2 lds[];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  float2* xch = lds + wave * XS;
  float2* tile = ALIAS ? lds : lds + 4 * XS;   // NH x TS
  const int groups = N / RW;
  // 8-row workgroups (N = 2048) store 64-byte runs of a 128-byte line: row
  // groups 2g and 2g + 1 of a frame run on one XCD
  const int blk = RW < 16 ? xcd_pair_block(blockIdx.x, gridDim.x) : blockIdx.x;
  const int f0 = (blk / groups) * FPW;
  const int y0 = (blk % groups) * RW;
  if (f0 >= nframes) return;
  const int nf = min(FPW, nframes - f0);       // workgroup-uniform
  const int npairs = PPW * nf;
  // pair p of this wave: frame f0 + p / PPW, rows ra = wave 2 PPW + 2 (p % PPW), ra + 1
  auto load_pair = [&](int p, Tin (&a)[P], Tin (&b)[P]) {
    const int ra = wave * 2 * PPW + 2 * (p % PPW);
    const Tin* rowa = frames + ((long)(f0 + p / PPW) * N + y0 + ra) * N;
    const Tin* rowb = rowa + N;
#pragma unroll
    for (int q = 0; q < P; q++) {
      // frames are read once: non-temporal loads
      a[q] = __builtin_nontemporal_load(&rowa[lane + 64 * q]);
      b[q] = __builtin_nontemporal_load(&rowb[lane + 64 * q]);
    }
  };
  float2 ua[ALIAS ? KQ : 1], ub[ALIAS ? KQ : 1];   // ALIAS: the pair's untangled spectra until the tile is free
  auto transform_pair = [&](int p, const Tin (&pa)[P], const Tin (&pb)[P]) {
    const int f = f0 + p / PPW;
    const int ra = wave * 2 * PPW + 2 * (p % PPW);   // rows ra, ra+1 of the tile
    float xa[P], xb[P];
#pragma unroll
    for (int q = 0; q < P; q++) {
      xa[q] = px_to_float(pa[q]);
      xb[q] = px_to_float(pb[q]);
    }
    float ha = 0.f, hb = 0.f, la = 0.f, lb = 0.f;
    int Sa = 0, Sb = 0;
    if constexpr (IS_INT) {
      // exact integer row sums
      int ia = 0, ib = 0;
#pragma unroll
      for (int q = 0; q < P; q++) {
        ia += (int)pa[q];
        ib += (int)pb[q];
      }
      Sa = wave_total(ia);
      Sb = wave_total(ib);
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
        da = (float)((int)pa[q] * N - Sa) * (scale / N);
        db = (float)((int)pb[q] * N - Sb) * (scale / N);
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
    wave_fft<LOGN, false, false, false, MAXLR>(v, xch, tw, lane);
    // untangle: X_a(k) = (Z(k) + conj Z(N-k))/2, X_b(k) = (Z(k) - conj Z(N-k))/(2i)
    wave_sync();
#pragma unroll
    for (int q = 0; q < P; q++) xch[xpad(lane + 64 * q)] = v[q];
    wave_sync();
#pragma unroll
    for (int q = 0; q < KQ; q++) {
      const int k = lane + 64 * q;
      if (k <= N / 2) {
        const float2 z = v[q];
        const float2 zc = cconj(xch[xpad((N - k) & (N - 1))]);
        const float2 u = csub(z, zc);
        const float2 xa2 = make_float2(0.5f * (z.x + zc.x), 0.5f * (z.y + zc.y));
        const float2 xb2 = make_float2(0.5f * u.y, -0.5f * u.x);
        if constexpr (ALIAS) {
          ua[q] = xa2;
          ub[q] = xb2;
        } else {
          tile[k * TS + ra] = xa2;
          tile[k * TS + ra + 1] = xb2;
        }
      }
    }
  };
  auto finish_frame = [&](int p) {   // after the frame's last pair p
    const int f = f0 + p / PPW;
    if constexpr (ALIAS) {
      __syncthreads();   // every wave is done with its exchange buffer (the tile's LDS)
      const int ra = wave * 2 * PPW + 2 * (p % PPW);
#pragma unroll
      for (int q = 0; q < KQ; q++) {
        const int k = lane + 64 * q;
        if (k <= N / 2) {
          tile[k * TS + ra] = ua[q];
          tile[k * TS + ra + 1] = ub[q];
        }
      }
    }
    __syncthreads();
    for (int e = tid; e < NH * RW; e += NT) {
      const int r = e % RW, kx = e / RW;
      scratch[((long)f * NH + kx) * N + y0 + r] = tile[kx * TS + r];
    }
    __syncthreads();   // the tile is stored before the next frame's rows overwrite it
  };
  Tin A0[P], A1[P], B0[P], B1[P];
  load_pair(0, A0, A1);
#pragma unroll 1
  for (int p = 0; p < npairs; p += 2) {
    if (p + 1 < npairs) load_pair(p + 1, B0, B1);
    transform_pair(p, A0, A1);
    if (p % PPW == PPW - 1) finish_frame(p);
    if (p + 2 < npairs) load_pair(p + 2, A0, A1);
    if (p + 1 < npairs) {
      transform_pair(p + 1, B0, B1);
      if ((p + 1) % PPW == PPW - 1) finish_frame(p + 1);
    }
  }
}

// Eight waves per column workgroup (F/8 frames per wave): at N = 2048 one
// frame per wave takes 216 VGPRs where two took 336 (two waves per SIMD
// instead of one), and at every N the extra waves hide the column loads
// (per frame, against four waves: 2048 10.7 -> 8.8 us, 1024 2.79 -> 2.63,
// 512 0.52 -> 0.47, 256 0.121 -> 0.110).
template <int LOGN> constexpr int cols_waves() { return 8; }

template <int LOGN, int F, int MAXLR>
__global__ __launch_bounds__(64 * cols_waves<LOGN>()) void k_cols_w(
    const float2* __restrict__ scratch, const int2* __restrict__ rowinfo, int nframes, int is_int, float scale,
    const float* __restrict__ win, const float2* __restrict__ tw, const float2* __restrict__ wx,
    float2* __restrict__ hist, int T, int hs, int first_slot) {
  constexpr int N = 1 << LOGN, NH = N / 2 + 1, P = N / 64;
  constexpr int W = cols_waves<LOGN>(), NTC = 64 * W;
  static_assert(F % W == 0, "whole frames per wave");
  constexpr int FW = F / W;                    // frames per wave
  constexpr int XS = xch_size(N);
  constexpr int TS = F + 1;
  extern __shared__ float2 lds[];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  float2* xch = lds + wave * XS;
  float2* tile = lds;                          // (N/2) x TS, reused after every wave's FFTs
  int kx, ft;   // column kx, frame tile ft
  if constexpr (F < 16) {
    // F-frame (64-byte) runs of the history: frame tiles 2i and 2i + 1 of a
    // column run on one XCD (a 1D grid of NH x tiles; an odd last tile alone)
    const int ntiles = gridDim.x / NH, Gp = NH * (ntiles & ~1);
    const int L = blockIdx.x < Gp ? xcd_pair_block(blockIdx.x, Gp) : blockIdx.x;
    if (L < Gp) {
      kx = (L >> 1) % NH;
      ft = 2 * ((L >> 1) / NH) + (L & 1);
    } else {
      kx = L - Gp;
      ft = ntiles - 1;
    }
  } else {
    kx = blockIdx.x;
    ft = blockIdx.y;
  }
  const int f0 = ft * F;
  // N <= 512: the column FFT's twiddle table in LDS (twpad slots, behind the
  // tile/exchange area), read there instead of waiting on global loads; at
  // N = 1024 the extra 8.5 KB would cost a resident workgroup
  constexpr bool TWL = LOGN <= 9;
  float2* twl = lds + ((N / 2) * TS > W * XS ? (N / 2) * TS : W * XS);
  if constexpr (TWL) {
    for (int i = tid; i < N; i += NTC) twl[twpad(i)] = tw[i];
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
      wave_fft<LOGN, false, false, TWL, MAXLR>(res[i], xch, TWL ? twl : tw, lane);
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
    for (int e = tid; e < (N / 2) * F; e += NTC) {
      const int fl = e % F, ky = h * (N / 2) + e / F;
      if (f0 + fl >= nframes) continue;
      const int slot = (first_slot + f0 + fl) % T;
      // non-temporal: the history is read back only by the temporal pass, after
      // the whole window; keeping it out of the caches leaves them to the row spectra
      const float2 z = tile[(e / F) * TS + fl];
      float2* dst = &hist[((long)ky * NH + kx) * hs + slot];
      __builtin_nontemporal_store(z.x, &dst->x);
      __builtin_nontemporal_store(z.y, &dst->y);
    }
  }
}

// ----------------------------------------------------------------- launchers

template <typename Tin, int N, int MAXLR>
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
  auto kern = k_rows<N, B, Tin, MAXLR>;
  hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  dim3 grid(nframes * (N / RW));
  hipLaunchKernelGGL(kern, grid, dim3(NT), lds, st, (const Tin*)frames, nframes, scale, win, tw, scratch, rowinfo);
  return hipGetLastError();
}

template <typename Tin, int MAXLR>
static hipError_t launch_rows_dt(int n, const void* frames, int nframes, float scale, const float* win,
                                 const float2* tw, float2* scratch, int2* rowinfo, hipStream_t st) {
  switch (n) {
#define CASE(N_) \
  case N_: return launch_rows_t<Tin, N_, MAXLR>(frames, nframes, scale, win, tw, scratch, rowinfo, st);
    CASE(16) CASE(32) CASE(64) CASE(128) CASE(256) CASE(512) CASE(1024) CASE(2048)
    CASE(48) CASE(96) CASE(192) CASE(384) CASE(768) CASE(1536)
#undef CASE
  }
  return hipErrorInvalidValue;
}

template <int MAXLR>
static hipError_t launch_rows_r(int dtype, int n, const void* frames, int nframes, float scale, const float* win,
                                const float2* tw, float2* scratch, int2* rowinfo, hipStream_t st) {
  if (dtype == 0) return launch_rows_dt<uint8_t, MAXLR>(n, frames, nframes, scale, win, tw, scratch, rowinfo, st);
  if (dtype == 1) return launch_rows_dt<uint16_t, MAXLR>(n, frames, nframes, scale, win, tw, scratch, rowinfo, st);
  return launch_rows_dt<float, MAXLR>(n, frames, nframes, scale, win, tw, scratch, rowinfo, st);
}

hipError_t launch_rows(int dtype, int n, int maxlr, const void* frames, int nframes, float scale, const float* win,
                       const float2* tw, float2* scratch, int2* rowinfo, hipStream_t st) {
  return maxlr == 2 ? launch_rows_r<2>(dtype, n, frames, nframes, scale, win, tw, scratch, rowinfo, st)
                    : launch_rows_r<3>(dtype, n, frames, nframes, scale, win, tw, scratch, rowinfo, st);
}

// Frames per column workgroup: the corner turn writes runs of this many slots
// per mode.  At N = 2048 eight (64 B runs, one 147 KB workgroup per CU)
// measured 20 % faster than four (32 B runs, two per CU).
int cols_frames_per_wg(int n) { return n >= 1024 ? 8 : 16; }

template <int N, int MAXLR>
static hipError_t launch_cols_t(const float2* scratch, const int2* rowinfo, int nframes, int is_int, float scale,
                                const float* win, const float2* tw, const float2* wx, float2* hist, int T,
                                int hs, int first_slot, hipStream_t st) {
  constexpr int NH = N / 2 + 1;
  constexpr int F = N >= 1024 ? 8 : 16;
  size_t lds = sizeof(float2) * (F * N + N) + sizeof(float) * 4 * F + sizeof(long long) * F;
  auto kern = k_cols<N, F, MAXLR>;
  hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  dim3 grid(NH, (nframes + F - 1) / F);
  hipLaunchKernelGGL(kern, grid, dim3(NT), lds, st, scratch, rowinfo, nframes, is_int, scale, win, tw, wx, hist, T, hs,
                     first_slot);
  return hipGetLastError();
}

template <int MAXLR>
static hipError_t launch_cols_r(int n, const float2* scratch, const int2* rowinfo, int nframes, int is_int,
                                float scale, const float* win, const float2* tw, const float2* wx, float2* hist, int T,
                                int hs, int first_slot, hipStream_t st) {
  switch (n) {
#define CASE(N_) \
  case N_:       \
    return launch_cols_t<N_, MAXLR>(scratch, rowinfo, nframes, is_int, scale, win, tw, wx, hist, T, hs, first_slot, st);
    CASE(16) CASE(32) CASE(64) CASE(128) CASE(256) CASE(512) CASE(1024) CASE(2048)
    CASE(48) CASE(96) CASE(192) CASE(384) CASE(768) CASE(1536)
#undef CASE
  }
  return hipErrorInvalidValue;
}

hipError_t launch_cols(int n, int maxlr, const float2* scratch, const int2* rowinfo, int nframes, int is_int,
                       float scale, const float* win, const float2* tw, const float2* wx, float2* hist, int T,
                       int hs, int first_slot, hipStream_t st) {
  return maxlr == 2 ? launch_cols_r<2>(n, scratch, rowinfo, nframes, is_int, scale, win, tw, wx, hist, T, hs, first_slot, st)
                    : launch_cols_r<3>(n, scratch, rowinfo, nframes, is_int, scale, win, tw, wx, hist, T, hs, first_slot, st);
}

template <typename Tin, int LOGN, int MAXLR, int FPW>
static hipError_t launch_rows_w_f(const void* frames, int nframes, float scale, const float* win, const float2* tw,
                                  float2* scratch, int2* rowinfo, hipStream_t st) {
  constexpr int N = 1 << LOGN, NH = N / 2 + 1;
  constexpr int RW = 8 * rows_pairs_w(LOGN);   // rows per workgroup (k_rows_w)
  const size_t xch = sizeof(float2) * 4 * xch_size(N), tile = sizeof(float2) * NH * (RW + 1);
  size_t lds = LOGN >= 11 ? (xch > tile ? xch : tile) : xch + tile;
  auto kern = k_rows_w<LOGN, Tin, MAXLR, FPW>;
  hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(kern, dim3((nframes + FPW - 1) / FPW * (N / RW)), dim3(NT), lds, st, (const Tin*)frames, nframes,
                     scale, win, tw, scratch, rowinfo);
  return hipGetLastError();
}

// Two frames per row workgroup: measured at 512^2 u16, 42.2 us per 112-frame
// launch against 44.9 (one frame), 49.6 (four) and 54.5 (eight).
template <typename Tin, int LOGN, int MAXLR>
static hipError_t launch_rows_w_t(const void* frames, int nframes, float scale, const float* win, const float2* tw,
                                  float2* scratch, int2* rowinfo, hipStream_t st) {
  return launch_rows_w_f<Tin, LOGN, MAXLR, 2>(frames, nframes, scale, win, tw, scratch, rowinfo, st);
}

template <int LOGN, int F, int MAXLR>
static hipError_t launch_cols_w_t(const float2* scratch, const int2* rowinfo, int nframes, int is_int, float scale,
                                  const float* win, const float2* tw, const float2* wx, float2* hist, int T,
                                  int hs, int first_slot, hipStream_t st) {
  constexpr int N = 1 << LOGN, NH = N / 2 + 1;
  size_t tile = sizeof(float2) * (N / 2) * (F + 1), xch = sizeof(float2) * cols_waves<LOGN>() * xch_size(N);
  size_t lds = (tile > xch ? tile : xch) + (LOGN <= 9 ? sizeof(float2) * twpad(N) : 0);
  auto kern = k_cols_w<LOGN, F, MAXLR>;
  hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  const int ntiles = (nframes + F - 1) / F;
  const dim3 grid = F < 16 ? dim3(NH * ntiles) : dim3(NH, ntiles);
  hipLaunchKernelGGL(kern, grid, dim3(64 * cols_waves<LOGN>()), lds, st, scratch, rowinfo, nframes, is_int, scale, win,
                     tw, wx, hist, T, hs, first_slot);
  return hipGetLastError();
}

template <int LOGN, int MAXLR>
static hipError_t launch_rows_w_l(int dtype, const void* frames, int nframes, float scale, const float* win,
                                  const float2* tw, float2* scratch, int2* rowinfo, hipStream_t st) {
  return dtype == 0 ? launch_rows_w_t<uint8_t, LOGN, MAXLR>(frames, nframes, scale, win, tw, scratch, rowinfo, st)
         : dtype == 1 ? launch_rows_w_t<uint16_t, LOGN, MAXLR>(frames, nframes, scale, win, tw, scratch, rowinfo, st)
                      : launch_rows_w_t<float, LOGN, MAXLR>(frames, nframes, scale, win, tw, scratch, rowinfo, st);
}

// Wavefront spatial kernels: N = 256 .. 2048 (registers hold N/64 points
// per lane); F = frames per column workgroup (8 or 16); maxlr: radix cap.
hipError_t launch_rows_w(int dtype, int logn, int maxlr, const void* frames, int nframes, float scale,
                         const float* win, const float2* tw, float2* scratch, int2* rowinfo, hipStream_t st) {
  const bool r4 = maxlr == 2;
  if (logn == 8)
    return r4 ? launch_rows_w_l<8, 2>(dtype, frames, nframes, scale, win, tw, scratch, rowinfo, st)
              : launch_rows_w_l<8, 3>(dtype, frames, nframes, scale, win, tw, scratch, rowinfo, st);
  if (logn == 9)
    return r4 ? launch_rows_w_l<9, 2>(dtype, frames, nframes, scale, win, tw, scratch, rowinfo, st)
              : launch_rows_w_l<9, 3>(dtype, frames, nframes, scale, win, tw, scratch, rowinfo, st);
  if (logn == 10)
    return r4 ? launch_rows_w_l<10, 2>(dtype, frames, nframes, scale, win, tw, scratch, rowinfo, st)
              : launch_rows_w_l<10, 3>(dtype, frames, nframes, scale, win, tw, scratch, rowinfo, st);
  if (logn == 11)
    return r4 ? launch_rows_w_l<11, 2>(dtype, frames, nframes, scale, win, tw, scratch, rowinfo, st)
              : launch_rows_w_l<11, 3>(dtype, frames, nframes, scale, win, tw, scratch, rowinfo, st);
  return hipErrorInvalidValue;
}

hipError_t launch_cols_w(int logn, int F, int maxlr, const float2* scratch, const int2* rowinfo, int nframes,
                         int is_int, float scale, const float* win, const float2* tw, const float2* wx, float2* hist,
                         int T, int hs, int first_slot, hipStream_t st) {
  switch ((logn * 32 + F) * 4 + (maxlr == 2 ? 2 : 3)) {
#define CASE(LG_, F_, R_)                  \
  case (LG_ * 32 + F_) * 4 + R_:           \
    return launch_cols_w_t<LG_, F_, R_>(scratch, rowinfo, nframes, is_int, scale, win, tw, wx, hist, T, hs, first_slot, st);
    CASE(8, 16, 3) CASE(8, 8, 3) CASE(9, 16, 3) CASE(9, 8, 3) CASE(10, 8, 3) CASE(10, 16, 3) CASE(11, 8, 3)
    CASE(8, 16, 2) CASE(8, 8, 2) CASE(9, 16, 2) CASE(9, 8, 2) CASE(10, 8, 2) CASE(10, 16, 2) CASE(11, 8, 2)
#undef CASE
  }
  return hipErrorInvalidValue;
}

}  // namespace ddm
