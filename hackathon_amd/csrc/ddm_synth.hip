// Seeded synthetic DDM videos on the GPU: Gaussian spots on a constant
// background, a fraction alpha of straight swimmers (Schulz speeds, uniform
// heading), the rest Brownian, periodic wrap, Poisson photon noise, clip to the
// pixel type.  The model is the reference movie simulator's
// (motility-ddm test/sim.mjs:90-186) in pixel units with Poisson noise.
//
// Random numbers: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random
// numbers: as easy as 1, 2, 3", SC'11), counter = (index, frame, stream, draw),
// key = the 64-bit seed, so every draw is addressable and a CPU restatement
// reproduces the streams exactly.  Normals: Box-Muller in FP64.  Schulz speeds:
// Gamma(Z+1, vbar/(Z+1)) by Marsaglia & Tsang (ACM TOMS 26(3), 2000).
// Poisson: inversion by sequential products for lambda < 10, the transformed
// rejection method PTRS (Hoermann, Insurance: Mathematics and Economics 12,
// 1993) above.
//
// Kernels: k_synth_init (one thread per particle: start, velocity),
// k_synth_tracks (one thread per particle, all frames: wrapped positions),
// k_synth_render (one workgroup per 16x16 tile of one frame: the tile's
// particle list is compacted in particle order, each pixel sums its spots in
// that order, draws its count and stores the pixel).
#include "ddm_internal.h"

#include <stdint.h>

namespace ddm {

namespace {

struct Philox {
  uint32_t c[4];
  uint32_t k0, k1;
  uint32_t out[4];
  int used;

  __device__ Philox(uint32_t idx, uint32_t frame, uint32_t stream, uint64_t seed)
      : c{idx, frame, stream, 0u}, k0((uint32_t)seed), k1((uint32_t)(seed >> 32)), used(4) {}

  __device__ void refill() {
    uint32_t x0 = c[0], x1 = c[1], x2 = c[2], x3 = c[3];
    uint32_t a = k0, b = k1;
#pragma unroll
    for (int r = 0; r < 10; r++) {
      const uint32_t lo0 = 0xD2511F53u * x0, hi0 = __umulhi(0xD2511F53u, x0);
      const uint32_t lo1 = 0xCD9E8D57u * x2, hi1 = __umulhi(0xCD9E8D57u, x2);
      const uint32_t y0 = hi1 ^ x1 ^ a, y1 = lo1, y2 = hi0 ^ x3 ^ b, y3 = lo0;
      x0 = y0, x1 = y1, x2 = y2, x3 = y3;
      a += 0x9E3779B9u;
      b += 0xBB67AE85u;
    }
    out[0] = x0, out[1] = x1, out[2] = x2, out[3] = x3;
    c[3]++;
    used = 0;
  }
  __device__ uint32_t next() {
    if (used == 4) refill();
    return out[used++];
  }
  // (0, 1) with 53 random bits
  __device__ double uniform() {
    const uint32_t a = next() >> 5, b = next() >> 6;
    return ((double)a * 67108864.0 + (double)b + 0.5) * (1.0 / 9007199254740992.0);
  }
  __device__ void normal2(double& n1, double& n2) {
    const double u1 = uniform(), u2 = uniform();
    const double r = sqrt(-2.0 * log(u1)), th = 6.283185307179586 * u2;
    n1 = r * cos(th);
    n2 = r * sin(th);
  }
};

constexpr uint32_t S_INIT = 0, S_STEP = 1, S_NOISE = 2;

__device__ double gamma_mt(Philox& g, double shape, double scale) {
  // Marsaglia-Tsang, shape >= 1
  const double d = shape - 1.0 / 3.0, c = 1.0 / sqrt(9.0 * d);
  for (;;) {
    double x, unused;
    g.normal2(x, unused);
    double v = 1.0 + c * x;
    if (v <= 0.0) continue;
    v = v * v * v;
    const double u = g.uniform();
    if (u < 1.0 - 0.0331 * (x * x) * (x * x)) return d * v * scale;
    if (log(u) < 0.5 * x * x + d * (1.0 - v + log(v))) return d * v * scale;
  }
}

__device__ double poisson(Philox& g, double lam) {
  if (!(lam > 0.0)) return 0.0;
  if (lam < 10.0) {
    const double L = exp(-lam);
    double p = 1.0;
    int k = 0;
    for (;;) {
      p *= g.uniform();
      if (p <= L) return (double)k;
      k++;
    }
  }
  const double slam = sqrt(lam), loglam = log(lam);
  const double b = 0.931 + 2.53 * slam, a = -0.059 + 0.02483 * b;
  const double invalpha = 1.1239 + 1.1328 / (b - 3.4), vr = 0.9277 - 3.6224 / (b - 2.0);
  for (;;) {
    const double U = g.uniform() - 0.5, V = g.uniform();
    const double us = 0.5 - fabs(U);
    const double k = floor((2.0 * a / us + b) * U + lam + 0.43);
    if (us >= 0.07 && V <= vr) return k;
    if (k < 0.0 || (us < 0.013 && V > us)) continue;
    if (log(V) + log(invalpha) - log(a / (us * us) + b) <= -lam + k * loglam - lgamma(k + 1.0)) return k;
  }
}

}  // namespace

// particle p: start (x0, y0) uniform on the torus, velocity (px/frame) for swimmers
// (the reference simulator's draw order: position, heading, speed, then the
// polar direction when headings are 3D, test/sim.mjs:109-121)
__global__ void k_synth_init(int n, int n_sw, int N, double vbar, double Z, int heading3d, uint64_t seed,
                             double4* __restrict__ part) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  Philox g((uint32_t)p, 0u, S_INIT, seed);
  const double x0 = g.uniform() * N, y0 = g.uniform() * N;
  double vx = 0.0, vy = 0.0;
  if (p < n_sw) {
    const double phi = 6.283185307179586 * g.uniform();
    double s = gamma_mt(g, Z + 1.0, vbar / (Z + 1.0));
    if (heading3d) {   // isotropic direction in 3D, projected on the image plane
      const double cz = 2.0 * g.uniform() - 1.0;
      s *= sqrt(1.0 - cz * cz);
    }
    vx = s * cos(phi);
    vy = s * sin(phi);
  }
  part[p] = make_double4(x0, y0, vx, vy);
}

// positions pos[t][p] = wrap(start + t v + sum_{s=1..t} step_s), steps N(0, 2D) per axis
// for the non-swimmers, and for the swimmers too when swim_diffuse is set (the
// reference simulator moves every mobile particle by a diffusive step,
// test/sim.mjs:162-168)
__global__ void k_synth_tracks(int n, int n_sw, int N, int T, double D, int swim_diffuse, uint64_t seed,
                               const double4* __restrict__ part, float2* __restrict__ pos) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const double4 s0 = part[p];
  double bx = 0.0, by = 0.0;
  const double sd = sqrt(2.0 * D);
  const bool diff = (p >= n_sw || swim_diffuse) && D > 0.0;
  for (int t = 0; t < T; t++) {
    if (diff && t > 0) {
      Philox g((uint32_t)p, (uint32_t)t, S_STEP, seed);
      double n1, n2;
      g.normal2(n1, n2);
      bx += sd * n1;
      by += sd * n2;
    }
    double x = s0.x + t * s0.z + bx, y = s0.y + t * s0.w + by;
    x -= floor(x / N) * N;
    y -= floor(y / N) * N;
    if (x >= N) x -= N;   // floor rounding at the top edge
    if (y >= N) y -= N;
    pos[(long)t * n + p] = make_float2((float)x, (float)y);
  }
}

template <typename Tout>
__device__ __forceinline__ Tout store_px(double v);
template <> __device__ __forceinline__ uint8_t store_px<uint8_t>(double v) { return (uint8_t)fmin(fmax(v, 0.0), 255.0); }
template <> __device__ __forceinline__ uint16_t store_px<uint16_t>(double v) {
  return (uint16_t)fmin(fmax(v, 0.0), 65535.0);
}
template <> __device__ __forceinline__ float store_px<float>(double v) { return (float)v; }

// one workgroup (256 threads) per 16x16 tile of one frame; grid (tiles, T)
template <typename Tout>
__global__ __launch_bounds__(256) void k_synth_render(int n, int N, int W, float inv2s2, float amp, float bg, int noise,
                                                      uint64_t seed, const float2* __restrict__ pos,
                                                      Tout* __restrict__ out) {
  __shared__ float2 list[256];
  __shared__ int wcount[4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int tpr = N >> 4;
  const int tx0 = (blockIdx.x % tpr) * 16, ty0 = (blockIdx.x / tpr) * 16;
  const int t = blockIdx.y;
  const int x = tx0 + (tid & 15), y = ty0 + (tid >> 4);
  const float halfN = 0.5f * N;
  const float reach = (float)W + 8.f;
  const float2* pt = pos + (long)t * n;
  float acc = 0.f;
  for (int c0 = 0; c0 < n; c0 += 256) {
    const int p = c0 + tid;
    float2 q = make_float2(0.f, 0.f);
    bool hit = false;
    if (p < n) {
      q = pt[p];
      float dx = (tx0 + 7.5f) - q.x, dy = (ty0 + 7.5f) - q.y;
      dx = dx >= halfN ? dx - N : (dx < -halfN ? dx + N : dx);
      dy = dy >= halfN ? dy - N : (dy < -halfN ? dy + N : dy);
      hit = fabsf(dx) <= reach && fabsf(dy) <= reach;
    }
    const unsigned long long m = __ballot(hit);
    const int before = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) wcount[wave] = __popcll(m);
    __syncthreads();
    int base = 0, total = 0;
    for (int w = 0; w < 4; w++) {
      base += w < wave ? wcount[w] : 0;
      total += wcount[w];
    }
    if (hit) list[base + before] = q;
    __syncthreads();
    for (int j = 0; j < total; j++) {
      const float2 s = list[j];
      float dx = (float)x - s.x, dy = (float)y - s.y;
      dx = dx >= halfN ? dx - N : (dx < -halfN ? dx + N : dx);
      dy = dy >= halfN ? dy - N : (dy < -halfN ? dy + N : dy);
      if (fabsf(dx) <= (float)W && fabsf(dy) <= (float)W) acc += expf(-(dx * dx + dy * dy) * inv2s2);
    }
    __syncthreads();
  }
  const double lam = (double)bg + (double)amp * (double)acc;
  const long i = (long)y * N + x;
  double v = lam;
  if (noise) {
    Philox g((uint32_t)i, (uint32_t)t, S_NOISE, seed);
    v = poisson(g, lam);
  }
  out[(long)t * N * N + i] = store_px<Tout>(v);
}

hipError_t launch_synth(int N, int T, int n, int n_sw, double sigma, double amp, double bg, double vbar, double Z,
                        double D, int heading3d, int swim_diffuse, uint64_t seed, int dtype, int noise,
                        double4* part, float2* pos, void* out, hipStream_t st) {
  if (n > 0) {
    hipLaunchKernelGGL(k_synth_init, dim3((n + 255) / 256), dim3(256), 0, st, n, n_sw, N, vbar, Z, heading3d, seed,
                       part);
    hipLaunchKernelGGL(k_synth_tracks, dim3((n + 63) / 64), dim3(64), 0, st, n, n_sw, N, T, D, swim_diffuse, seed,
                       (const double4*)part, pos);
  }
  if (!out) return hipGetLastError();
  const int W = (int)ceil(4.0 * sigma) + 1;
  const float inv2s2 = (float)(1.0 / (2.0 * sigma * sigma));
  const dim3 grid((N / 16) * (N / 16), T);
  switch (dtype) {
    case 0:
      hipLaunchKernelGGL(k_synth_render<uint8_t>, grid, dim3(256), 0, st, n, N, W, inv2s2, (float)amp, (float)bg,
                         noise, seed, (const float2*)pos, (uint8_t*)out);
      break;
    case 1:
      hipLaunchKernelGGL(k_synth_render<uint16_t>, grid, dim3(256), 0, st, n, N, W, inv2s2, (float)amp, (float)bg,
                         noise, seed, (const float2*)pos, (uint16_t*)out);
      break;
    default:
      hipLaunchKernelGGL(k_synth_render<float>, grid, dim3(256), 0, st, n, N, W, inv2s2, (float)amp, (float)bg,
                         noise, seed, (const float2*)pos, (float*)out);
  }
  return hipGetLastError();
}

}  // namespace ddm
