// Surrogate MLP for LM seeding: 1536 -> 256 -> 256 -> 256 -> 68, GELU, BF16
// weights and activations on MFMA with FP32 accumulation; forward, backward
// and Adam are hand-written kernels (no BLAS).
//
// GEMM: C[M x N] (f32) = A[M x K] (bf16) * B[K x N] (bf16), operands addressed
// through (row, col) strides so the same kernel serves X W, dY W^T and X^T dY.
// A workgroup of 4 waves computes a 64 x 64 tile (each wave 32 x 32 with
// gfx950's v_mfma_f32_32x32x16_bf16), K staged through LDS 32 at a time from
// 16-byte global loads.
#include <cstdint>
#include "ddm_internal.h"


namespace ddm {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float float16_t __attribute__((ext_vector_type(16)));
typedef unsigned short ushort8_t __attribute__((ext_vector_type(8)));

__device__ __forceinline__ unsigned short f2bf(float x) {
  // round to nearest even; NaN stays NaN
  unsigned int u = __float_as_uint(x);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float(((unsigned int)h) << 16); }

namespace {
constexpr int GT = 64;   // output tile
constexpr int GK = 32;   // K chunk
constexpr int KP = GK + 8;   // LDS row pitch (elements): 80 B rows keep the 16-byte k reads of a wave spread over the banks
}

// Operand staging of one K chunk: A(i, k) for the tile's 64 rows and B(k, j)
// for its 64 columns, both stored k-contiguous in LDS (As[i][k], Bt[j][k]) so
// that every lane reads its 8 consecutive k of an MFMA operand with one
// 16-byte LDS load.  Each thread fetches one 8-element run along the operand's
// contiguous global dimension (16-byte loads where the run is inside the
// matrix and aligned, `vec`), else element by element with bounds checks.
struct GemmOp {
  const unsigned short* p;
  long s_outer, s_k;   // A: (s_i, s_k); B: (s_j, s_k)
  int n_outer;         // M for A, N for B
  int kmajor;          // 1: k is the contiguous dimension (s_k == 1), 0: the outer one is (s_outer == 1)
  int vec;             // 16-byte loads allowed
};

__device__ __forceinline__ void gemm_fetch(const GemmOp& o, int t, int o0, int k0, int K, ushort8_t& r) {
  // run t of 256: kmajor -> outer = t / 4, k = 8 (t % 4); else k = t / 8, outer = 8 (t % 8)
  const int oo = o.kmajor ? t / 4 : 8 * (t % 8);
  const int kk = o.kmajor ? 8 * (t % 4) : t / 8;
  const int go = o0 + oo, gk = k0 + kk;
  if (o.kmajor) {
    if (o.vec && go < o.n_outer && gk + 8 <= K) {
      r = *reinterpret_cast<const ushort8_t*>(o.p + go * o.s_outer + gk);
    } else {
#pragma unroll
      for (int v = 0; v < 8; v++) r[v] = (go < o.n_outer && gk + v < K) ? o.p[go * o.s_outer + gk + v] : (unsigned short)0;
    }
  } else {
    if (o.vec && go + 8 <= o.n_outer && gk < K) {
      r = *reinterpret_cast<const ushort8_t*>(o.p + go + gk * o.s_k);
    } else {
#pragma unroll
      for (int v = 0; v < 8; v++) r[v] = (go + v < o.n_outer && gk < K) ? o.p[go + v + gk * o.s_k] : (unsigned short)0;
    }
  }
}

__device__ __forceinline__ void gemm_stash(const GemmOp& o, int t, const ushort8_t& r, unsigned short (*S)[KP]) {
  if (o.kmajor) {
    *reinterpret_cast<ushort8_t*>(&S[t / 4][8 * (t % 4)]) = r;
  } else {
#pragma unroll
    for (int v = 0; v < 8; v++) S[8 * (t % 8) + v][t / 8] = r[v];
  }
}

// C = acc ? C + A B : A B on v_mfma_f32_32x32x16_bf16 (gfx950): a workgroup of
// 4 waves computes a 64 x 64 tile, each wave 32 x 32, K in chunks of 32 (two
// MFMAs per wave per chunk); the next chunk's operands are fetched from global
// memory into registers while the current chunk is multiplied.
// MFMA operand layout: lane l holds A(i0 + l % 32, k0 + 8 (l / 32) + v) and
// B(k0 + 8 (l / 32) + v, j0 + l % 32), v = 0..7; C register r holds row
// 8 (r / 4) + 4 (l / 32) + r % 4, column l % 32.
__global__ __launch_bounds__(256) void k_gemm_bf16(GemmOp a, GemmOp b, float* __restrict__ C, long ldc, int M, int N,
                                                   int K, int acc) {
  __shared__ __attribute__((aligned(16))) unsigned short As[GT][KP];
  __shared__ __attribute__((aligned(16))) unsigned short Bt[GT][KP];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int i0 = blockIdx.y * GT, j0 = blockIdx.x * GT;
  const int wi = (wave >> 1) * 32, wj = (wave & 1) * 32;
  float16_t c = {0};
  ushort8_t ra, rb;
  gemm_fetch(a, tid, i0, 0, K, ra);
  gemm_fetch(b, tid, j0, 0, K, rb);
  for (int k0 = 0; k0 < K; k0 += GK) {
    __syncthreads();   // the previous chunk's LDS reads are done
    gemm_stash(a, tid, ra, As);
    gemm_stash(b, tid, rb, Bt);
    __syncthreads();
    if (k0 + GK < K) {
      gemm_fetch(a, tid, i0, k0 + GK, K, ra);
      gemm_fetch(b, tid, j0, k0 + GK, K, rb);
    }
#pragma unroll
    for (int kk = 0; kk < GK; kk += 16) {
      const int kb = kk + 8 * (lane >> 5);
      const bf16x8_t av = __builtin_bit_cast(bf16x8_t, *reinterpret_cast<const ushort8_t*>(&As[wi + (lane & 31)][kb]));
      const bf16x8_t bv = __builtin_bit_cast(bf16x8_t, *reinterpret_cast<const ushort8_t*>(&Bt[wj + (lane & 31)][kb]));
      c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, c, 0, 0, 0);
    }
  }
#pragma unroll
  for (int r = 0; r < 16; r++) {
    const int gi = i0 + wi + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    const int gj = j0 + wj + (lane & 31);
    if (gi < M && gj < N) {
      float* p = C + gi * ldc + gj;
      *p = acc ? *p + c[r] : c[r];
    }
  }
}

// Z (f32, pre-activation, in place + bias) -> H = GELU(Z) in bf16 (or Z itself
// when gelu == 0).  GELU(z) = z Phi(z) with the exact normal cdf.
__global__ void k_bias_act(float* __restrict__ Z, const float* __restrict__ bias, unsigned short* __restrict__ H,
                           int M, int N, int gelu) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)M * N) return;
  const float z = Z[i] + bias[i % N];
  Z[i] = z;
  const float h = gelu ? 0.5f * z * (1.f + erff(z * 0.70710678118654752f)) : z;
  H[i] = f2bf(h);
}

// dZ = dH * GELU'(Z) (bf16 out for the next GEMMs); dbias[j] += sum_i dZ[i][j]
// (one thread per column, fixed order).
__global__ void k_act_bwd(const float* __restrict__ dH, const float* __restrict__ Z, unsigned short* __restrict__ dZ,
                          float* __restrict__ dZf, int M, int N, int gelu) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)M * N) return;
  const float z = Z[i];
  float g = 1.f;
  if (gelu) {
    const float cdf = 0.5f * (1.f + erff(z * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * expf(-0.5f * z * z);
    g = cdf + z * pdf;
  }
  const float d = dH[i] * g;
  dZf[i] = d;
  dZ[i] = f2bf(d);
}

__global__ void k_colsum(const float* __restrict__ X, float* __restrict__ out, int M, int N) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  float s = 0.f;
  for (int i = 0; i < M; i++) s += X[(long)i * N + j];
  out[j] = s;
}

// Loss 0.5 mean over rows of sum_j wt[j] (y - t)^2; dY = wt (y - t) / M.
__global__ void k_mse(const float* __restrict__ Y, const float* __restrict__ Tg, const float* __restrict__ wt,
                      float* __restrict__ dY, float* __restrict__ rowloss, int M, int N) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  float s = 0.f;
  for (int j = 0; j < N; j++) {
    const float e = Y[(long)i * N + j] - Tg[(long)i * N + j];
    s += wt[j] * e * e;
    dY[(long)i * N + j] = wt[j] * e / M;
  }
  rowloss[i] = 0.5f * s;
}

// Adam on FP32 master weights, refreshing the BF16 compute copy.
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       unsigned short* __restrict__ pb, long n, float lr, float b1, float b2, float eps, float c1,
                       float c2) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  p[i] = pi;
  if (pb) pb[i] = f2bf(pi);
}

__global__ void k_to_bf16(const float* __restrict__ x, unsigned short* __restrict__ y, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = f2bf(x[i]);
}

// One operand's staging description: the contiguous global dimension (k, or
// the outer one) and whether its 8-element runs may be read as aligned
// 16-byte loads (base 16-byte aligned, every run start a multiple of 8).
static GemmOp gemm_op(const unsigned short* p, long s_outer, long s_k, int n_outer) {
  GemmOp o{p, s_outer, s_k, n_outer, 0, 0};
  const bool base16 = ((uintptr_t)p & 15) == 0;
  if (s_k == 1) {
    o.kmajor = 1;
    o.vec = base16 && s_outer % 8 == 0;
  } else if (s_outer == 1) {
    o.kmajor = 0;
    o.vec = base16 && s_k % 8 == 0;
  } else {
    o.kmajor = 1;   // neither dimension contiguous: element loads
  }
  return o;
}

hipError_t launch_gemm_bf16(const unsigned short* A, long sai, long sak, const unsigned short* B, long sbk, long sbj,
                            float* C, long ldc, int M, int N, int K, int acc, hipStream_t st) {
  dim3 grid((N + GT - 1) / GT, (M + GT - 1) / GT);
  hipLaunchKernelGGL(k_gemm_bf16, grid, dim3(256), 0, st, gemm_op(A, sai, sak, M), gemm_op(B, sbj, sbk, N), C, ldc,
                     M, N, K, acc);
  return hipGetLastError();
}

hipError_t launch_bias_act(float* Z, const float* bias, unsigned short* H, int M, int N, int gelu, hipStream_t st) {
  const long n = (long)M * N;
  hipLaunchKernelGGL(k_bias_act, dim3((n + 255) / 256), dim3(256), 0, st, Z, bias, H, M, N, gelu);
  return hipGetLastError();
}

hipError_t launch_act_bwd(const float* dH, const float* Z, unsigned short* dZ, float* dZf, float* dbias, int M, int N,
                          int gelu, hipStream_t st) {
  const long n = (long)M * N;
  hipLaunchKernelGGL(k_act_bwd, dim3((n + 255) / 256), dim3(256), 0, st, dH, Z, dZ, dZf, M, N, gelu);
  hipLaunchKernelGGL(k_colsum, dim3((N + 255) / 256), dim3(256), 0, st, dZf, dbias, M, N);
  return hipGetLastError();
}

hipError_t launch_mse(const float* Y, const float* Tg, const float* wt, float* dY, float* rowloss, int M, int N,
                      hipStream_t st) {
  hipLaunchKernelGGL(k_mse, dim3((M + 255) / 256), dim3(256), 0, st, Y, Tg, wt, dY, rowloss, M, N);
  return hipGetLastError();
}

hipError_t launch_adam(float* p, const float* g, float* m, float* v, unsigned short* pb, long n, float lr, float b1,
                       float b2, float eps, float c1, float c2, hipStream_t st) {
  hipLaunchKernelGGL(k_adam, dim3((n + 255) / 256), dim3(256), 0, st, p, g, m, v, pb, n, lr, b1, b2, eps, c1, c2);
  return hipGetLastError();
}

hipError_t launch_to_bf16(const float* x, unsigned short* y, long n, hipStream_t st) {
  hipLaunchKernelGGL(k_to_bf16, dim3((n + 255) / 256), dim3(256), 0, st, x, y, n);
  return hipGetLastError();
}

}  // namespace ddm
