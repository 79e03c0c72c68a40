// libddm engine: host side of the C ABI in include/ddm.h.
//
// Owns the device buffers of one geometry, the host FP64 tables (window,
// twiddles, bins, temporal metadata) and one HIP stream.  It is the MI355X
// counterpart of DdmEngine (motility-ddm src/ddm-engine.ts:62-426) and of the
// host precompute it pulls from src/ddm_ref.js (makeBins :84-170, temporalMeta
// :538-560, gpuTwiddles :566-574) and src/fft.js (blackmanHarris :126-133).
#include "../../include/ddm.h"
#include "ddm_internal.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

thread_local std::string g_create_error;

constexpr double kPi = 3.14159265358979323846;

int ilog2(long n) {
  int l = 0;
  while ((1L << l) < n) l++;
  return l;
}
bool is_pow2(long n) { return n >= 1 && (n & (n - 1)) == 0; }

// Tuning override from the environment (clamped), or the default.
long env_long(const char* name, long dflt, long lo, long hi) {
  const char* v = std::getenv(name);
  if (!v || !*v) return dflt;
  char* end = nullptr;
  long x = std::strtol(v, &end, 10);
  if (end == v) return dflt;
  return std::min(hi, std::max(lo, x));
}

// ------------------------------------------------------------ host tables

// 4-term periodic Blackman-Harris in FP64 (src/fft.js:126-133).
std::vector<double> blackman_harris(int n) {
  std::vector<double> w(n);
  for (int i = 0; i < n; i++) {
    double x = 2 * kPi * i / n;
    w[i] = 0.35875 - 0.48829 * std::cos(x) + 0.14128 * std::cos(2 * x) - 0.01168 * std::cos(3 * x);
  }
  return w;
}

// exp(-2 pi i m / n), m < n, FP64 rounded once to f32 (src/ddm_ref.js:566-574
// keeps a half table for radix 2; the radix-8 passes index up to n-1).
std::vector<float> twiddles(int n) {
  std::vector<float> t(2 * n);
  for (int m = 0; m < n; m++) {
    double a = -2 * kPi * m / n;
    t[2 * m] = (float)std::cos(a);
    t[2 * m + 1] = (float)std::sin(a);
  }
  return t;
}

// The same in FP64 (the bin-level inverse transforms).
std::vector<double> twiddles64(int n) {
  std::vector<double> t(2 * n);
  for (int m = 0; m < n; m++) {
    double a = -2 * kPi * m / n;
    t[2 * m] = std::cos(a);
    t[2 * m + 1] = std::sin(a);
  }
  return t;
}

// V8's two-argument Math.hypot: scale by the larger magnitude, Kahan-sum the
// squared ratios in argument order, sqrt times the scale.  The reference bins
// by it (src/ddm_ref.js:93,122,135); it is not correctly rounded, so a
// different hypot moves Pythagorean bin-edge modes.  This translation unit is
// compiled with -ffp-contract=off so no step is fused.
double v8_hypot(double a, double b) {
  double aa = std::fabs(a), ab = std::fabs(b);
  if (std::isinf(aa) || std::isinf(ab)) return INFINITY;
  if (std::isnan(aa) || std::isnan(ab)) return NAN;
  double mx = std::max(aa, ab);
  if (mx == 0) return 0;
  double sum = 0, comp = 0;
  const double v[2] = {aa, ab};
  for (double x : v) {
    double r = x / mx;
    double summand = r * r - comp;
    double prelim = sum + summand;
    comp = (prelim - sum) - summand;
    sum = prelim;
  }
  return std::sqrt(sum) * mx;
}

// Math.round: nearest integer, ties toward +infinity.
double js_round(double x) {
  double r = std::floor(x);
  return (x - r >= 0.5) ? r + 1 : r;
}

struct Bins {
  int NB = 0, K = 0;
  std::vector<uint32_t> offsets;   // NB+1
  std::vector<int32_t> modes4;     // 4 per binned mode: mode, kx, ky, 0
  std::vector<double> kcenter;     // NB
  std::vector<uint32_t> nmodes;    // NB
  std::vector<double> knodes;      // NB*K
};

// Radial bins linear in |k| on [kmin, kmax): half-plane modes, skipping the
// kx = 0, ky <= 0 half axis (conjugates of ky > 0) and the Nyquist lines;
// per-bin mean |k|; K equal-weight quantile nodes of the bin's |k| distribution
// convolved with the window's power leakage (src/ddm_ref.js:40-55, 84-170).
Bins make_bins(int N, int NB, double kmin, double kmax, int hypot_mode, int K) {
  auto hyp = [hypot_mode](double a, double b) { return hypot_mode ? std::hypot(a, b) : v8_hypot(a, b); };
  const int NH = N / 2 + 1;
  const double width = (kmax - kmin) / NB;
  std::vector<std::vector<int>> lists(NB);
  for (int kyi = 0; kyi < N; kyi++) {
    const int ky = kyi < N / 2 ? kyi : kyi - N;
    for (int kx = 0; kx < NH; kx++) {
      if (kx == 0 && ky <= 0) continue;
      if (kx == N / 2 || ky == -N / 2) continue;
      double b = std::floor((hyp(kx, ky) - kmin) / width);
      if (b >= 0 && b < NB) lists[(int)b].push_back(kyi * NH + kx);
    }
  }
  Bins out;
  out.NB = NB;
  out.K = K;
  out.offsets.resize(NB + 1);
  out.kcenter.resize(NB);
  out.nmodes.resize(NB);
  out.knodes.assign((size_t)NB * K, 0.0);
  uint32_t total = 0;
  for (int b = 0; b < NB; b++) {
    out.offsets[b] = total;
    total += (uint32_t)lists[b].size();
  }
  out.offsets[NB] = total;
  out.modes4.resize(4 * (size_t)total);
  // window power leakage at offsets -4..4 step 0.5
  std::vector<double> w = blackman_harris(N), ld, lp;
  for (double x = -4; x <= 4 + 1e-9; x += 0.5) {
    double re = 0, im = 0;
    for (int i = 0; i < N; i++) {
      double a = (-2 * kPi * x * i) / N;
      re += w[i] * std::cos(a);
      im += w[i] * std::sin(a);
    }
    ld.push_back(x);
    lp.push_back(re * re + im * im);
  }
  double lsum = 0;
  for (double v : lp) lsum += v;
  for (double& v : lp) v /= lsum;
  const double h0 = std::max(0.0, kmin - 6), hres = 0.01;
  const long hlen = (long)std::ceil((kmax + 6 - h0) / hres) + 1;
  std::vector<double> hist(hlen), hk(hlen);
  for (int b = 0; b < NB; b++) {
    const auto& lst = lists[b];
    double sk = 0;
    for (size_t i = 0; i < lst.size(); i++) {
      int m = lst[i], kx = m % NH, kyi = m / NH, ky = kyi < N / 2 ? kyi : kyi - N;
      size_t o = 4 * (out.offsets[b] + i);
      out.modes4[o] = m;
      out.modes4[o + 1] = kx;
      out.modes4[o + 2] = ky;
      out.modes4[o + 3] = 0;
      sk += hyp(kx, ky);
    }
    out.nmodes[b] = (uint32_t)lst.size();
    out.kcenter[b] = lst.empty() ? kmin + (b + 0.5) * width : sk / (double)lst.size();
    if (K <= 0) continue;
    std::fill(hist.begin(), hist.end(), 0.0);
    std::fill(hk.begin(), hk.end(), 0.0);
    for (int m : lst) {
      int kx = m % NH, kyi = m / NH, ky = kyi < N / 2 ? kyi : kyi - N;
      for (size_t i = 0; i < ld.size(); i++) {
        double ax = kx + ld[i];
        for (size_t j = 0; j < ld.size(); j++) {
          double k = hyp(ax, ky + ld[j]);
          double wt = lp[i] * lp[j];
          double c = std::min((double)(hlen - 1), std::max(0.0, js_round((k - h0) / hres)));
          hist[(long)c] += wt;
          hk[(long)c] += wt * k;
        }
      }
    }
    double wsum = 0;
    for (double v : hist) wsum += v;
    if (!(wsum > 0)) {
      for (int n = 0; n < K; n++) out.knodes[(size_t)b * K + n] = out.kcenter[b];
      continue;
    }
    double acc = 0, sw = 0, swk = 0;
    int n = 0;
    for (long c = 0; c < hlen && n < K; c++) {
      double rem = hist[c];
      double kc = hist[c] > 0 ? hk[c] / hist[c] : 0;
      while (rem > 0 && n < K) {
        double take = std::min(rem, ((double)(n + 1) / K) * wsum - acc);
        sw += take;
        swk += take * kc;
        acc += take;
        rem -= take;
        if (acc >= ((double)(n + 1) / K) * wsum * (1 - 1e-12)) {
          out.knodes[(size_t)b * K + n] = sw > 0 ? swk / sw : kc;
          n++;
          sw = 0;
          swk = 0;
        }
      }
    }
    for (; n < K; n++) out.knodes[(size_t)b * K + n] = kmax;
  }
  return out;
}

}  // namespace

struct ddm_engine {
  ddm_geometry g{};
  int NH = 0, TPAD = 0, logn = 0, logt = 0;
  int HS = 0;             // history row stride: T rounded up to 16 slots (128 B)
  long modes = 0;
  int chunk = 0;
  int cols_f = 16;        // frames per column workgroup (wavefront kernels)
  int maxlr = 3;          // FFT radix cap (log2): 3 radix 8, 2 radix 4
  int tkern = 0;          // temporal kernel: 0 one mode per wave (one per workgroup above TPAD 4096), 1 cooperative
  int twg_cu = 0;         // cooperative kernel: 0 one workgroup per item, k > 0 a persistent grid of k per CU
  int cus = 0;            // compute units of the engine's device
  int* tcounter = nullptr;  // the persistent grid's item counter
  bool wave_spatial = false;
  int mean_frame = -1;    // rowinfo frame of the last frame ingested (-1: none yet)
  int mean_dtype = 0;
  hipStream_t stream = nullptr;
  std::string err;
  Bins bins;
  int nitems = 0;
  std::vector<int> bin_items_h;
  // device buffers
  float2* hist = nullptr;
  float2* scratch = nullptr;
  int2* rowinfo = nullptr;
  void* stage = nullptr;
  size_t stage_bytes = 0;
  float* win = nullptr;
  float2* tw_n = nullptr;
  float2* tw_t = nullptr;
  float2* tw_f = nullptr;   // twiddles of the temporal sub-transform length (min(TPAD, 512))
  float2* wx = nullptr;
  int4* items = nullptr;
  int* bin_modes = nullptr;
  int* bin_items = nullptr;
  uint32_t* nmodes = nullptr;
  uint32_t* order = nullptr;
  float* mask = nullptr;
  float* npairs = nullptr;
  float* S_part = nullptr;
  float* p_part = nullptr;
  float* P_part = nullptr;
  float* radD = nullptr;
  float* radP = nullptr;
  double* Sbin = nullptr;  // NB*TPAD summed |Y|^2 per bin (frequency order)
  float* G_part = nullptr; // items*KT direct D N(tau) sums of the tail lags
  int KT = 1;              // tail lags T-KT..T-1 summed directly (ceil(T/64); computeTemporal may use more)
  int KT_max = 1;          // capacity of G_part per work item
  double* Gbin = nullptr;  // NB*L summed Re G(tau) - G(0) per bin, D N(tau) at the tail lags (FP64)
  double2* tw_t64 = nullptr; // FP64 twiddles of the TPAD transform (bin-level inverse transforms)
  double* pbin = nullptr;  // NB*T summed p per bin
  float* Pbin = nullptr;   // NB summed P per bin
  bool temporal_busy = false;
  // set while the engine's work is being recorded into a HIP graph (batch):
  // no host waits, the pinned metadata staging then belongs to the graph
  bool capturing = false;
  ddm_tuning tune{};
  // the last computeTemporal's metadata, for readTemporalFull
  bool meta_valid = false, last_identity = true;
  float last_n0 = 0.f;
  // the analysis begun by computeTemporal's first half (temporal_begin), for
  // its temporal launch and radial finalisation
  int cur_KT = 1;
  bool cur_full = true;
  float *fullD = nullptr, *fullP = nullptr;   // per-mode outputs of one chunk of modes
  float2* fullC = nullptr;
  double* drift_acc = nullptr;   // NB*L*8 phase-drift sums (FP64)
  int2* kxy = nullptr;           // (kx, ky) per binned mode
  int* bin_off = nullptr;        // NB+1 binned-mode offsets
  // pinned staging of the per-analysis metadata (order, 0/1 mask, N(tau));
  // meta_ev marks when the last upload from it finished
  uint32_t* meta_order = nullptr;
  float* meta_mask = nullptr;
  float* meta_np = nullptr;
  hipEvent_t meta_ev = nullptr;
  // profiling
  bool prof = false;
  struct Rec {
    int kind;
    hipEvent_t a, b;
  };
  std::vector<Rec> recs;
  std::vector<hipEvent_t> pool;
};

namespace {

int fail(ddm_engine* e, int code, const std::string& msg) {
  if (e)
    e->err = msg;
  else
    g_create_error = msg;
  return code;
}

int hip_fail(ddm_engine* e, hipError_t he, const char* what) {
  return fail(e, DDM_EHIP, std::string(what) + ": " + hipGetErrorString(he));
}

#define HIPCHK(e, call, what)                          \
  do {                                                 \
    hipError_t _he = (call);                           \
    if (_he != hipSuccess) return hip_fail(e, _he, what); \
  } while (0)

template <typename T>
hipError_t dalloc(T** p, size_t count) {
  return hipMalloc((void**)p, std::max<size_t>(count, 1) * sizeof(T));
}

hipEvent_t take_event(ddm_engine* e) {
  if (!e->pool.empty()) {
    hipEvent_t ev = e->pool.back();
    e->pool.pop_back();
    return ev;
  }
  hipEvent_t ev = nullptr;
  if (hipEventCreate(&ev) != hipSuccess) return nullptr;
  return ev;
}

// Wrap a launch with events when profiling is on.
template <typename F>
hipError_t timed(ddm_engine* e, int kind, F&& launch) {
  if (!e->prof || e->capturing) return launch();
  hipEvent_t a = take_event(e), b = take_event(e);
  if (!a || !b) return launch();
  hipEventRecord(a, e->stream);
  hipError_t r = launch();
  hipEventRecord(b, e->stream);
  e->recs.push_back({kind, a, b});
  return r;
}

void free_all(ddm_engine* e) {
  void* ptrs[] = {e->hist, e->scratch, e->rowinfo, e->stage, e->win, e->tw_n, e->tw_t, e->tw_t64, e->tw_f, e->wx,
                  e->items, e->bin_modes, e->bin_items, e->nmodes, e->order, e->mask, e->npairs, e->S_part,
                  e->p_part, e->P_part, e->radD, e->radP, e->Gbin, e->pbin, e->Pbin, e->Sbin, e->G_part,
                  e->fullD, e->fullC, e->fullP, e->drift_acc, e->kxy, e->bin_off, e->tcounter};
  for (void* p : ptrs)
    if (p) hipFree(p);
  void* hptrs[] = {e->meta_order, e->meta_mask, e->meta_np};
  for (void* p : hptrs)
    if (p) hipHostFree(p);
  if (e->meta_ev) hipEventDestroy(e->meta_ev);
  for (auto& r : e->recs) {
    hipEventDestroy(r.a);
    hipEventDestroy(r.b);
  }
  for (auto ev : e->pool) hipEventDestroy(ev);
  if (e->stream) hipStreamDestroy(e->stream);
}

}  // namespace

extern "C" {

const char* ddm_last_error(const ddm_engine* e) { return e ? e->err.c_str() : g_create_error.c_str(); }

int ddm_create(const ddm_geometry* gp, ddm_engine** out) { return ddm_create_tuned(gp, nullptr, out); }

int ddm_create_tuned(const ddm_geometry* gp, const ddm_tuning* tp, ddm_engine** out) {
  if (!gp || !out) return fail(nullptr, DDM_EINVAL, "ddm_create: null argument");
  ddm_tuning tn{};
  if (tp) tn = *tp;
  if (tn.temporal_items < 0 || tn.chunk_mb < 0 || tn.wave_spatial < 0 || tn.wave_spatial > 2 ||
      (tn.cols_f != 0 && tn.cols_f != 8 && tn.cols_f != 16) ||
      (tn.fft_radix != 0 && tn.fft_radix != 4 && tn.fft_radix != 8) || tn.temporal_kernel < 0 ||
      tn.temporal_kernel > 2 || tn.temporal_wg_per_cu < 0 || tn.temporal_wg_per_cu > 16)
    return fail(nullptr, DDM_EINVAL, "ddm_create: tuning out of range");
  *out = nullptr;
  const ddm_geometry g = *gp;
  const bool n3 = g.N % 3 == 0 && is_pow2(g.N / 3) && g.N >= 48 && g.N <= 1536;
  if (!((is_pow2(g.N) && g.N >= 16 && g.N <= 2048) || n3))
    return fail(nullptr, DDM_EINVAL,
                "ddm_create: N must be a power of two in [16, 2048] or 3 * 2^m in [48, 1536], got " +
                    std::to_string(g.N));
  if (g.T < 2 || g.T > 4096) return fail(nullptr, DDM_EINVAL, "ddm_create: T must be in [2, 4096]");
  if (g.L < 1 || g.L > g.T) return fail(nullptr, DDM_EINVAL, "ddm_create: L must be in [1, T]");
  if (g.NB < 1) return fail(nullptr, DDM_EINVAL, "ddm_create: NB must be >= 1");
  if (!(g.kmax > g.kmin) || !(g.kmin >= 0)) return fail(nullptr, DDM_EINVAL, "ddm_create: need 0 <= kmin < kmax");
  if (g.knodes < 0 || g.knodes > 64) return fail(nullptr, DDM_EINVAL, "ddm_create: knodes must be in [0, 64]");
  ddm_engine* e = new (std::nothrow) ddm_engine();
  if (!e) return fail(nullptr, DDM_ENOMEM, "ddm_create: out of host memory");
  e->g = g;
  e->NH = g.N / 2 + 1;
  e->modes = (long)g.N * e->NH;
  e->logn = ilog2(g.N);
  e->TPAD = std::max(256, 1 << ilog2(2L * g.T));
  // Each mode's series starts on a 128-byte line, so the column pass's
  // 16-frame corner-turn runs from slot 16j fill whole lines: with a stride
  // of T they straddle two lines for T % 16 != 0, and the pass ran 1.5-2.4x
  // slower (512^2 frames, T = 1025 .. 1100).  TPAD/2 >= 128 is a power of two
  // >= T, so HS <= TPAD/2.
  e->HS = (g.T + 15) & ~15;
  // Lags with few valid pairs take Re G from direct sums: the FFT route's
  // error is a few ulp x sum_t |y|^2 whatever the lag, i.e. ~ulp x T / N(tau)
  // relative to D(tau) there.  KT (the default, for readTemporalFull) covers
  // the last T/64 lags; computeTemporal extends the direct tail per mask up to
  // KT_max lags so that every transform lag keeps N(tau) >= T/32 (measured at
  // 1024 frames: the error at the last transform lag is ~2e-7 x T / N(tau)).
  e->KT = std::max(1, (g.T + 63) / 64);
  e->KT_max = std::min(64, g.T);
  e->logt = ilog2(e->TPAD);
  auto bail = [&](int code) {
    g_create_error = e->err;
    free_all(e);
    delete e;
    return code;
  };
  hipError_t he = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
  if (he != hipSuccess) return (hip_fail(e, he, "hipStreamCreate"), bail(DDM_EHIP));
  {
    int dev = 0;
    if ((he = hipGetDevice(&dev)) != hipSuccess ||
        (he = hipDeviceGetAttribute(&e->cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess)
      return (hip_fail(e, he, "ddm_create: device query"), bail(DDM_EHIP));
  }

  e->bins = make_bins(g.N, g.NB, g.kmin, g.kmax, g.hypot_mode, g.knodes);
  const long binned = e->bins.offsets[g.NB];

  // Temporal work items: (bin, first mode, count); up to M modes of one bin.
  const long target_items =
      tn.temporal_items > 0 ? tn.temporal_items : env_long("DDM_TEMPORAL_ITEMS", 2048, 1, 1L << 22);
  // temporal kernel: the cooperative one (every wave of a workgroup on one
  // mode) takes TPAD >= 2048; at TPAD 8192 its items hold at most 16 modes
  // (its per-wave sums are one-level there)
  // (0: the default, the cooperative kernel wherever it applies)
  const long tk = tn.temporal_kernel ? tn.temporal_kernel : env_long("DDM_TEMPORAL_KERNEL", 0, 0, 2);
  e->tkern = tk != 1 && e->logt >= 11 ? 1 : 0;
  e->twg_cu = e->tkern ? (int)(tn.temporal_wg_per_cu ? tn.temporal_wg_per_cu
                                                     : env_long("DDM_TEMPORAL_WG_PER_CU", 0, 0, 16)) : 0;
  const int Mcap = e->tkern && e->logt >= 13 ? 16 : 64;
  int M = (int)std::min<long>(Mcap, std::max<long>(1, (binned + target_items - 1) / target_items));
  std::vector<int> items_h, bin_modes_h(std::max<long>(binned, 1));
  e->bin_items_h.assign(g.NB + 1, 0);
  for (int b = 0; b < g.NB; b++) {
    e->bin_items_h[b] = (int)(items_h.size() / 4);
    for (uint32_t i = e->bins.offsets[b]; i < e->bins.offsets[b + 1]; i += M) {
      int cnt = (int)std::min<uint32_t>(M, e->bins.offsets[b + 1] - i);
      items_h.insert(items_h.end(), {b, (int)i, cnt, 0});
    }
  }
  e->nitems = (int)(items_h.size() / 4);
  e->bin_items_h[g.NB] = e->nitems;
  for (long i = 0; i < binned; i++) bin_modes_h[i] = e->bins.modes4[4 * i];

  // Spatial chunk: frames whose row spectra (NH*N*8 B each) fit the chunk
  // budget (default 128 MB: the scratch stays in the 256 MB last-level cache).
  // wave_spatial: 0 default (environment, else on where built), 1 off, 2 on
  // (N = 2048: 8-frame column tiles; 2.6x faster per frame than the workgroup-level kernels, measured)
  const long ws = tn.wave_spatial ? tn.wave_spatial - 1 : env_long("DDM_WAVE_SPATIAL", 1, 0, 1);
  e->wave_spatial = is_pow2(g.N) && e->logn >= 8 && e->logn <= 11 && ws;
  e->cols_f = e->logn == 11 ? 8 : (tn.cols_f ? tn.cols_f : env_long("DDM_COLS_F", 16, 8, 16)) >= 16 ? 16 : 8;
  const int F = 16;   // chunks are a multiple of every column-kernel frame tile
  long per_frame = (long)e->NH * g.N * 8;
  const long chunk_mb = tn.chunk_mb > 0 ? tn.chunk_mb : env_long("DDM_CHUNK_MB", 128, 1, 1 << 16);
  const long chunk_bytes = chunk_mb << 20;
  e->tune.temporal_items = (int32_t)target_items;
  e->tune.chunk_mb = (int32_t)chunk_mb;
  e->tune.wave_spatial = e->wave_spatial ? 2 : 1;
  e->tune.cols_f = e->cols_f;
  e->maxlr = (tn.fft_radix ? tn.fft_radix : env_long("DDM_FFT_RADIX", 8, 4, 8)) == 4 ? 2 : 3;
  e->tune.fft_radix = e->maxlr == 2 ? 4 : 8;
  e->tune.temporal_kernel = e->tkern ? 2 : 1;
  e->tune.temporal_wg_per_cu = e->twg_cu;
  int chunk = (int)std::max<long>(F, (chunk_bytes / per_frame) / F * F);
  chunk = std::min(chunk, ((g.T + F - 1) / F) * F);
  e->chunk = chunk;
  e->stage_bytes = (size_t)chunk * g.N * g.N * 4;

  const size_t nitems = std::max(e->nitems, 1);
  // Rows of HS slots (the HS - T behind each series stay zero), then TPAD/2 -
  // HS spare samples behind the last row: the temporal kernel reads every
  // series (and the frame order) as TPAD/2 entries without a guard; the
  // entries past T belong to the row's padding, the next mode or the spare
  // samples and are masked off, the order's pad stays 0
  const size_t hist_len = (size_t)e->modes * e->HS + (size_t)(e->TPAD / 2 - e->HS);
  if ((he = dalloc(&e->hist, hist_len)) != hipSuccess ||
      (he = dalloc(&e->scratch, (size_t)chunk * e->NH * g.N)) != hipSuccess ||
      (he = dalloc(&e->rowinfo, (size_t)chunk * g.N)) != hipSuccess ||
      (he = hipMalloc(&e->stage, e->stage_bytes)) != hipSuccess || (he = dalloc(&e->win, g.N)) != hipSuccess ||
      (he = dalloc(&e->tw_n, g.N)) != hipSuccess || (he = dalloc(&e->tw_t, e->TPAD)) != hipSuccess ||
      (he = dalloc(&e->tw_t64, e->TPAD)) != hipSuccess ||
      (he = dalloc(&e->tw_f, std::min(e->TPAD, 512))) != hipSuccess ||
      (he = dalloc(&e->wx, 4)) != hipSuccess || (he = dalloc(&e->items, nitems)) != hipSuccess ||
      (he = dalloc(&e->bin_modes, bin_modes_h.size())) != hipSuccess ||
      (he = dalloc(&e->bin_items, g.NB + 1)) != hipSuccess || (he = dalloc(&e->nmodes, g.NB)) != hipSuccess ||
      (he = dalloc(&e->order, (size_t)e->TPAD / 2)) != hipSuccess || (he = dalloc(&e->mask, g.T)) != hipSuccess ||
      (he = dalloc(&e->npairs, g.L)) != hipSuccess || (he = dalloc(&e->S_part, nitems * e->TPAD)) != hipSuccess ||
      (he = dalloc(&e->p_part, nitems * g.T)) != hipSuccess || (he = dalloc(&e->P_part, nitems)) != hipSuccess ||
      (he = dalloc(&e->radD, (size_t)g.NB * g.L)) != hipSuccess || (he = dalloc(&e->radP, g.NB)) != hipSuccess ||
      (he = dalloc(&e->Gbin, (size_t)g.NB * g.L)) != hipSuccess ||
      (he = dalloc(&e->Sbin, (size_t)g.NB * e->TPAD)) != hipSuccess ||
      (he = dalloc(&e->G_part, nitems * (size_t)e->KT_max)) != hipSuccess || (he = dalloc(&e->pbin, (size_t)g.NB * g.T)) != hipSuccess ||
      (he = dalloc(&e->Pbin, g.NB)) != hipSuccess || (he = dalloc(&e->tcounter, 1)) != hipSuccess) {
    hip_fail(e, he, "ddm_create: device allocation");
    return bail(he == hipErrorOutOfMemory ? DDM_ENOMEM : DDM_EHIP);
  }
  if ((he = hipHostMalloc((void**)&e->meta_order, g.T * 4)) != hipSuccess ||
      (he = hipHostMalloc((void**)&e->meta_mask, g.T * 4)) != hipSuccess ||
      (he = hipHostMalloc((void**)&e->meta_np, g.L * 4)) != hipSuccess ||
      (he = hipEventCreateWithFlags(&e->meta_ev, hipEventDisableTiming)) != hipSuccess) {
    hip_fail(e, he, "ddm_create: pinned host allocation");
    return bail(DDM_EHIP);
  }

  std::vector<double> w64 = blackman_harris(g.N);
  std::vector<float> w32(g.N);
  for (int i = 0; i < g.N; i++) w32[i] = (float)w64[i];
  // DFT of the f32 window at kx = 0..3, in FP64: the only columns the row-mean
  // correction touches (the 4-term window's DFT vanishes for 4 <= kx <= N-4).
  std::vector<float> wx_h(8);
  for (int kx = 0; kx < 4; kx++) {
    double re = 0, im = 0;
    for (int x = 0; x < g.N; x++) {
      double a = -2 * kPi * ((long)kx * x % g.N) / g.N;
      re += (double)w32[x] * std::cos(a);
      im += (double)w32[x] * std::sin(a);
    }
    wx_h[2 * kx] = (float)re;
    wx_h[2 * kx + 1] = (float)im;
  }
  std::vector<float> twn = twiddles(g.N), twt = twiddles(e->TPAD), twf = twiddles(std::min(e->TPAD, 512));
  std::vector<double> twt64 = twiddles64(e->TPAD);
  std::vector<int4> items4(nitems, int4{0, 0, 0, 0});
  std::memcpy(items4.data(), items_h.data(), items_h.size() * sizeof(int));
  if ((he = hipMemcpy(e->win, w32.data(), g.N * 4, hipMemcpyHostToDevice)) != hipSuccess ||
      (he = hipMemcpy(e->tw_n, twn.data(), twn.size() * 4, hipMemcpyHostToDevice)) != hipSuccess ||
      (he = hipMemcpy(e->tw_t, twt.data(), twt.size() * 4, hipMemcpyHostToDevice)) != hipSuccess ||
      (he = hipMemcpy(e->tw_t64, twt64.data(), twt64.size() * 8, hipMemcpyHostToDevice)) != hipSuccess ||
      (he = hipMemcpy(e->tw_f, twf.data(), twf.size() * 4, hipMemcpyHostToDevice)) != hipSuccess ||
      (he = hipMemcpy(e->wx, wx_h.data(), 32, hipMemcpyHostToDevice)) != hipSuccess ||
      (he = hipMemcpy(e->items, items4.data(), items4.size() * sizeof(int4), hipMemcpyHostToDevice)) != hipSuccess ||
      (he = hipMemcpy(e->bin_modes, bin_modes_h.data(), bin_modes_h.size() * 4, hipMemcpyHostToDevice)) !=
          hipSuccess ||
      (he = hipMemcpy(e->bin_items, e->bin_items_h.data(), (g.NB + 1) * 4, hipMemcpyHostToDevice)) != hipSuccess ||
      (he = hipMemcpy(e->nmodes, e->bins.nmodes.data(), g.NB * 4, hipMemcpyHostToDevice)) != hipSuccess ||
      // clears on the engine's own (non-blocking) stream, finished before
      // create returns: a null-stream memset is not ordered before the
      // engine's kernels, and a 69 GB history is still being cleared when
      // the first frames' spectra arrive
      (he = hipMemsetAsync(e->hist, 0, hist_len * 8, e->stream)) != hipSuccess ||
      (he = hipMemsetAsync(e->order, 0, (size_t)e->TPAD / 2 * 4, e->stream)) != hipSuccess ||
      (he = hipMemsetAsync(e->radD, 0, (size_t)g.NB * g.L * 4, e->stream)) != hipSuccess ||
      (he = hipMemsetAsync(e->radP, 0, (size_t)g.NB * 4, e->stream)) != hipSuccess ||
      (he = hipStreamSynchronize(e->stream)) != hipSuccess) {
    hip_fail(e, he, "ddm_create: table upload");
    return bail(DDM_EHIP);
  }
  *out = e;
  return DDM_OK;
}

void ddm_destroy(ddm_engine* e) {
  if (!e) return;
  hipStreamSynchronize(e->stream);
  free_all(e);
  delete e;
}

int ddm_get_tuning(const ddm_engine* e, ddm_tuning* t) {
  if (!e || !t) return DDM_EINVAL;
  *t = e->tune;
  return DDM_OK;
}

int ddm_get_sizes(const ddm_engine* e, ddm_sizes* s) {
  if (!e || !s) return DDM_EINVAL;
  s->N = e->g.N;
  s->NH = e->NH;
  s->T = e->g.T;
  s->TPAD = e->TPAD;
  s->L = e->g.L;
  s->NB = e->g.NB;
  s->knodes = e->g.knodes;
  s->modes = e->modes;
  s->binned_modes = e->bins.offsets[e->g.NB];
  s->hist_bytes = (int64_t)e->modes * e->g.T * 8;
  s->work_items = e->nitems;
  s->chunk_frames = e->chunk;
  s->hist_stride = e->HS;
  return DDM_OK;
}

int ddm_get_bins(const ddm_engine* e, uint32_t* off, int32_t* modes4, double* kc, uint32_t* nm, double* kn) {
  if (!e) return DDM_EINVAL;
  const Bins& b = e->bins;
  if (off) std::memcpy(off, b.offsets.data(), b.offsets.size() * 4);
  if (modes4) std::memcpy(modes4, b.modes4.data(), b.modes4.size() * 4);
  if (kc) std::memcpy(kc, b.kcenter.data(), b.kcenter.size() * 8);
  if (nm) std::memcpy(nm, b.nmodes.data(), b.nmodes.size() * 4);
  if (kn) std::memcpy(kn, b.knodes.data(), b.knodes.size() * 8);
  return DDM_OK;
}

static int check_ingest(ddm_engine* e, const void* frames, int32_t dtype, int32_t nframes, int32_t first_slot) {
  if (!e) return DDM_EINVAL;
  if (!frames) return fail(e, DDM_EINVAL, "ingest: null frame pointer");
  if (dtype < DDM_U8 || dtype > DDM_F32) return fail(e, DDM_EINVAL, "ingest: unknown dtype");
  if (nframes < 0) return fail(e, DDM_EINVAL, "ingest: negative frame count");
  if (first_slot < 0) return fail(e, DDM_EINVAL, "ingest: negative slot");
  if (nframes > e->g.T) return fail(e, DDM_EINVAL, "ingest: at most T frames per call (one ring turn)");
  return DDM_OK;
}

static int run_spatial(ddm_engine* e, const void* dframes, int dtype, int nframes, int slot) {
  const float scale = dtype == DDM_U8 ? 1.0f / 255.0f : dtype == DDM_U16 ? 1.0f / 65535.0f : 1.0f;
  hipError_t he = timed(e, DDM_K_ROWS, [&] {
    if (e->wave_spatial)
      return ddm::launch_rows_w(dtype, e->logn, e->maxlr, dframes, nframes, scale, e->win, e->tw_n, e->scratch, e->rowinfo,
                                e->stream);
    return ddm::launch_rows(dtype, e->g.N, e->maxlr, dframes, nframes, scale, e->win, e->tw_n, e->scratch, e->rowinfo,
                            e->stream);
  });
  if (he != hipSuccess) return hip_fail(e, he, "spatial rows launch");
  he = timed(e, DDM_K_COLS, [&] {
    if (e->wave_spatial)
      return ddm::launch_cols_w(e->logn, e->cols_f, e->maxlr, e->scratch, e->rowinfo, nframes, dtype != DDM_F32, scale, e->win,
                                e->tw_n, e->wx, e->hist, e->g.T, e->HS, slot % e->g.T, e->stream);
    return ddm::launch_cols(e->g.N, e->maxlr, e->scratch, e->rowinfo, nframes, dtype != DDM_F32, scale, e->win, e->tw_n, e->wx,
                            e->hist, e->g.T, e->HS, slot % e->g.T, e->stream);
  });
  if (he != hipSuccess) return hip_fail(e, he, "spatial cols launch");
  e->mean_frame = nframes - 1;   // its row sums stay in rowinfo until the next ingest
  e->mean_dtype = dtype;
  return DDM_OK;
}

int ddm_ingest(ddm_engine* e, const void* frames, int32_t dtype, int32_t nframes, int32_t first_slot) {
  int rc = check_ingest(e, frames, dtype, nframes, first_slot);
  if (rc) return rc;
  const size_t px = (size_t)e->g.N * e->g.N;
  const size_t bpp = dtype == DDM_U8 ? 1 : dtype == DDM_U16 ? 2 : 4;
  for (int f0 = 0; f0 < nframes; f0 += e->chunk) {
    int nf = std::min(e->chunk, nframes - f0);
    HIPCHK(e, hipMemcpyAsync(e->stage, (const char*)frames + f0 * px * bpp, nf * px * bpp, hipMemcpyHostToDevice,
                             e->stream),
           "ingest: host to device copy");
    if ((rc = run_spatial(e, e->stage, dtype, nf, first_slot + f0))) return rc;
  }
  // the caller may reuse its (pageable) buffer as soon as this returns
  HIPCHK(e, hipStreamSynchronize(e->stream), "ingest: stream");
  return DDM_OK;
}

int ddm_ingest_device(ddm_engine* e, const void* dframes, int32_t dtype, int32_t nframes, int32_t first_slot) {
  int rc = check_ingest(e, dframes, dtype, nframes, first_slot);
  if (rc) return rc;
  const size_t px = (size_t)e->g.N * e->g.N;
  const size_t bpp = dtype == DDM_U8 ? 1 : dtype == DDM_U16 ? 2 : 4;
  for (int f0 = 0; f0 < nframes; f0 += e->chunk) {
    int nf = std::min(e->chunk, nframes - f0);
    if ((rc = run_spatial(e, (const char*)dframes + f0 * px * bpp, dtype, nf, first_slot + f0))) return rc;
  }
  return DDM_OK;
}

// computeTemporal: metadata, the temporal kernel, the radial finalisation
static int compute_temporal(ddm_engine* e, const uint32_t* order, const float* mask) {
  int rc = ddm::engine_temporal_begin(e, order, mask);
  if (rc) return rc;
  hipError_t he;
  if (e->nitems > 0 &&
      (he = timed(e, DDM_K_TEMPORAL, [&] {
         const float inv_n0 = e->last_n0 > 0.5f ? 1.0f / e->last_n0 : 0.0f;
         const int identity = e->last_identity ? 1 : 0;
         if (e->tkern && e->twg_cu > 0) {   // cooperative kernel on a persistent grid
           const ddm::TemporalArgs a{e->hist, e->items, e->bin_modes, e->order, e->mask, e->S_part, e->p_part,
                                     e->P_part, e->G_part, e->g.T, e->HS, identity, e->cur_KT, 0, inv_n0};
           return ddm::launch_temporal_cp(e->logt, e->maxlr, a, e->nitems, e->tcounter, e->tw_t, e->tw_f,
                                          std::min(e->nitems, e->twg_cu * e->cus), e->stream);
         }
         if (e->tkern)        // cooperative kernel (512-point sub-transforms, r-major spectra)
           return ddm::launch_temporal_c(e->logt, e->maxlr, e->hist, e->g.T, e->HS, e->items, e->nitems, e->bin_modes,
                                         e->order, e->mask, identity, e->cur_full && 2 * e->g.T == e->TPAD, inv_n0,
                                         e->tw_t, e->tw_f, e->S_part, e->p_part, e->P_part, e->cur_KT, e->G_part,
                                         e->stream);
         if (e->logt <= 12)   // wavefront-per-mode kernel (sub-transforms of at most 512 points)
           return ddm::launch_temporal_w(e->logt, e->maxlr, e->hist, e->g.T, e->HS, e->items, e->nitems, e->bin_modes, e->order,
                                         e->mask, identity, inv_n0, e->tw_t, e->tw_f, e->S_part, e->p_part,
                                         e->P_part, e->cur_KT, e->G_part, e->stream);
         return ddm::launch_temporal(e->logt, e->hist, e->g.T, e->HS, e->items, e->nitems, e->bin_modes, e->order, e->mask,
                                     identity, inv_n0, e->tw_t, e->S_part, e->p_part, e->P_part, e->cur_KT,
                                     e->G_part, e->stream);
       })) != hipSuccess)
    return hip_fail(e, he, "temporal launch");
  return ddm::engine_temporal_end(e);
}

int ddm_compute_temporal(ddm_engine* e, const uint32_t* order, const float* mask) {
  if (!e) return DDM_EINVAL;
  if (e->temporal_busy) return fail(e, DDM_ESTATE, "computeTemporal is already running");
  e->temporal_busy = true;
  int rc = compute_temporal(e, order, mask);
  e->temporal_busy = false;
  return rc;
}


// Phase-drift sums (opt-in: only when the caller asks for them): C(tau) of
// every binned mode from the per-mode kernel, in chunks of the binned-mode
// list, reduced per (bin, lag) in FP64 in stream order.
static int compute_drift(ddm_engine* e, float* drift) {
  const int NB = e->g.NB, L = e->g.L, T = e->g.T;
  const size_t nout = (size_t)NB * L * 8;
  if (!e->meta_valid) {   // nothing analysed yet: the reference's record holds zeros
    std::memset(drift, 0, nout * sizeof(float));
    return DDM_OK;
  }
  const long MC = 8192;
  hipError_t he;
  if (!e->fullC && (he = dalloc(&e->fullC, (size_t)MC * L)) != hipSuccess) return hip_fail(e, he, "drift: scratch");
  if (!e->drift_acc) {
    const long binned = e->bins.offsets[NB];
    std::vector<int2> kxy(std::max<long>(binned, 1));
    for (long i = 0; i < binned; i++) kxy[i] = make_int2(e->bins.modes4[4 * i + 1], e->bins.modes4[4 * i + 2]);
    std::vector<int> off(e->bins.offsets.begin(), e->bins.offsets.end());
    if ((he = dalloc(&e->drift_acc, nout)) != hipSuccess || (he = dalloc(&e->kxy, kxy.size())) != hipSuccess ||
        (he = dalloc(&e->bin_off, off.size())) != hipSuccess ||
        (he = hipMemcpy(e->kxy, kxy.data(), kxy.size() * sizeof(int2), hipMemcpyHostToDevice)) != hipSuccess ||
        (he = hipMemcpy(e->bin_off, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice)) != hipSuccess)
      return hip_fail(e, he, "drift: tables");
  }
  HIPCHK(e, hipMemsetAsync(e->drift_acc, 0, nout * sizeof(double), e->stream), "drift: clear");
  const long binned = e->bins.offsets[NB];
  const float inv_n0 = e->last_n0 > 0.5f ? 1.0f / e->last_n0 : 0.0f;
  const auto& off = e->bins.offsets;
  for (long c0 = 0; c0 < binned; c0 += MC) {
    const int nc = (int)std::min<long>(MC, binned - c0);
    // bins whose mode ranges meet [c0, c0 + nc)
    const int b0 = (int)(std::upper_bound(off.begin(), off.end(), (uint32_t)c0) - off.begin()) - 1;
    const int b1 = (int)(std::upper_bound(off.begin(), off.end(), (uint32_t)(c0 + nc - 1)) - off.begin()) - 1;
    HIPCHK(e, ddm::launch_temporal_full(e->logt, e->hist, T, e->HS, L, c0, nc, e->order, e->mask, e->last_identity ? 1 : 0,
                                        inv_n0, e->npairs, e->tw_t, e->KT, e->bin_modes, nullptr, e->fullC, nullptr,
                                        e->stream),
           "drift: per-mode C");
    HIPCHK(e, ddm::launch_drift(e->fullC, L, (int)c0, nc, e->bin_off, e->kxy, b0, b1 - b0 + 1, e->drift_acc, e->stream),
           "drift: sums");
  }
  std::vector<double> acc(nout);
  HIPCHK(e, hipMemcpyAsync(acc.data(), e->drift_acc, nout * sizeof(double), hipMemcpyDeviceToHost, e->stream),
         "drift: copy");
  HIPCHK(e, hipStreamSynchronize(e->stream), "drift: stream");
  for (size_t i = 0; i < nout; i++) drift[i] = (float)acc[i];
  return DDM_OK;
}

int ddm_read_radial(ddm_engine* e, float* D, float* P, float* drift) {
  if (!e) return DDM_EINVAL;
  HIPCHK(e, hipStreamSynchronize(e->stream), "readRadial: stream");
  if (D) HIPCHK(e, hipMemcpy(D, e->radD, (size_t)e->g.NB * e->g.L * 4, hipMemcpyDeviceToHost), "readRadial: D");
  if (P) HIPCHK(e, hipMemcpy(P, e->radP, (size_t)e->g.NB * 4, hipMemcpyDeviceToHost), "readRadial: P");
  if (drift) return compute_drift(e, drift);
  return DDM_OK;
}

int ddm_read_mean(ddm_engine* e, double* mean) {
  if (!e) return DDM_EINVAL;
  if (!mean) return fail(e, DDM_EINVAL, "readMean: null output");
  HIPCHK(e, hipStreamSynchronize(e->stream), "readMean: stream");
  *mean = 0;
  if (e->mean_frame < 0) return DDM_OK;
  const int N = e->g.N;
  std::vector<int32_t> ri(2 * (size_t)N);
  HIPCHK(e, hipMemcpy(ri.data(), e->rowinfo + (size_t)e->mean_frame * N, (size_t)N * 8, hipMemcpyDeviceToHost),
         "readMean: copy");
  if (e->mean_dtype == DDM_F32) {
    // per row the mean as an unevaluated (hi, lo) pair
    double s = 0;
    for (int y = 0; y < N; y++) {
      float hi, lo;
      std::memcpy(&hi, &ri[2 * y], 4);
      std::memcpy(&lo, &ri[2 * y + 1], 4);
      s += (double)hi + (double)lo;
    }
    *mean = s / N;
  } else {
    // exact integer row sums, scaled to luma as the spectra are
    int64_t s = 0;
    for (int y = 0; y < N; y++) s += ri[2 * y];
    *mean = (double)s / ((double)N * N) / (e->mean_dtype == DDM_U8 ? 255.0 : 65535.0);
  }
  return DDM_OK;
}

int ddm_read_history(ddm_engine* e, float* out) {
  if (!e) return DDM_EINVAL;
  if (!out) return fail(e, DDM_EINVAL, "readHistory: null output");
  HIPCHK(e, hipStreamSynchronize(e->stream), "readHistory: stream");
  const size_t row = (size_t)e->g.T * 8, pitch = (size_t)e->HS * 8;
  if (row == pitch) {
    HIPCHK(e, hipMemcpy(out, e->hist, (size_t)e->modes * row, hipMemcpyDeviceToHost), "readHistory: copy");
    return DDM_OK;
  }
  // rows of HS slots on the device, of T in the output (strided copies of at
  // most 16384 rows each)
  for (long m0 = 0; m0 < e->modes; m0 += 16384) {
    const size_t rows = (size_t)std::min<long>(16384, e->modes - m0);
    HIPCHK(e, hipMemcpy2D(reinterpret_cast<char*>(out) + m0 * row, row, e->hist + m0 * e->HS, pitch, row, rows,
                          hipMemcpyDeviceToHost),
           "readHistory: copy");
  }
  return DDM_OK;
}

int ddm_read_temporal_full(ddm_engine* e, float* D, float* C, float* P) {
  if (!e) return DDM_EINVAL;
  if (!e->meta_valid) return fail(e, DDM_ESTATE, "readTemporalFull: computeTemporal has not run");
  const int T = e->g.T, L = e->g.L;
  const long MC = 8192;   // modes per launch: bounds the device scratch to MC * L * 12 bytes
  hipError_t he;
  if ((!e->fullD && (he = dalloc(&e->fullD, (size_t)MC * L)) != hipSuccess) ||
      (!e->fullC && (he = dalloc(&e->fullC, (size_t)MC * L)) != hipSuccess) ||
      (!e->fullP && (he = dalloc(&e->fullP, (size_t)MC)) != hipSuccess))
    return hip_fail(e, he, "readTemporalFull: device scratch");
  const float inv_n0 = e->last_n0 > 0.5f ? 1.0f / e->last_n0 : 0.0f;
  for (long m0 = 0; m0 < e->modes; m0 += MC) {
    const int nm = (int)std::min<long>(MC, e->modes - m0);
    HIPCHK(e, ddm::launch_temporal_full(e->logt, e->hist, T, e->HS, L, m0, nm, e->order, e->mask, e->last_identity ? 1 : 0,
                                        inv_n0, e->npairs, e->tw_t, e->KT, nullptr, e->fullD, e->fullC, e->fullP,
                                        e->stream),
           "readTemporalFull: launch");
    if (D)
      HIPCHK(e, hipMemcpyAsync(D + m0 * L, e->fullD, (size_t)nm * L * 4, hipMemcpyDeviceToHost, e->stream),
             "readTemporalFull: D");
    if (C)
      HIPCHK(e, hipMemcpyAsync(C + m0 * L * 2, e->fullC, (size_t)nm * L * 8, hipMemcpyDeviceToHost, e->stream),
             "readTemporalFull: C");
    if (P)
      HIPCHK(e, hipMemcpyAsync(P + m0, e->fullP, (size_t)nm * 4, hipMemcpyDeviceToHost, e->stream),
             "readTemporalFull: P");
    HIPCHK(e, hipStreamSynchronize(e->stream), "readTemporalFull: stream");
  }
  return DDM_OK;
}

int ddm_idle(ddm_engine* e) {
  if (!e) return DDM_EINVAL;
  HIPCHK(e, hipStreamSynchronize(e->stream), "idle");
  return DDM_OK;
}

int ddm_history_device(ddm_engine* e, void** dptr) {
  if (!e || !dptr) return DDM_EINVAL;
  *dptr = e->hist;
  return DDM_OK;
}

int ddm_set_profiling(ddm_engine* e, int32_t enable) {
  if (!e) return DDM_EINVAL;
  HIPCHK(e, hipStreamSynchronize(e->stream), "set_profiling: stream");
  for (auto& r : e->recs) {
    e->pool.push_back(r.a);
    e->pool.push_back(r.b);
  }
  e->recs.clear();
  e->prof = enable != 0;
  return DDM_OK;
}

int ddm_kernel_times(ddm_engine* e, double* ms, int32_t* count, int32_t n) {
  if (!e || !ms || !count) return DDM_EINVAL;
  HIPCHK(e, hipStreamSynchronize(e->stream), "kernel_times: stream");
  for (int k = 0; k < n; k++) {
    ms[k] = 0;
    count[k] = 0;
  }
  for (auto& r : e->recs) {
    float t = 0;
    HIPCHK(e, hipEventElapsedTime(&t, r.a, r.b), "kernel_times: elapsed");
    if (r.kind < n) {
      ms[r.kind] += t;
      count[r.kind] += 1;
    }
  }
  return DDM_OK;
}

}  // extern "C"

namespace ddm {
hipStream_t engine_stream(ddm_engine* e) { return e->stream; }
void engine_set_capturing(ddm_engine* e, bool on) { e->capturing = on; }

// A recorded graph keeps reading the pinned order / mask / N(tau) staging it
// was captured with: hand those buffers to the graph's owner and give the
// engine fresh ones, so a later direct computeTemporal on the engine cannot
// change what the graph uploads.
hipError_t engine_detach_meta(ddm_engine* e, void** owned3) {
  owned3[0] = e->meta_order;
  owned3[1] = e->meta_mask;
  owned3[2] = e->meta_np;
  e->meta_order = nullptr;
  e->meta_mask = nullptr;
  e->meta_np = nullptr;
  hipError_t he;
  if ((he = hipHostMalloc((void**)&e->meta_order, e->g.T * 4)) != hipSuccess ||
      (he = hipHostMalloc((void**)&e->meta_mask, e->g.T * 4)) != hipSuccess ||
      (he = hipHostMalloc((void**)&e->meta_np, e->g.L * 4)) != hipSuccess)
    return he;
  return hipSuccess;
}
void engine_get_meta(const ddm_engine* e, bool* identity, float* n0) {
  *identity = e->last_identity;
  *n0 = e->last_n0;
}
// after a graph run: the engine's device metadata is the graph's again
// computeTemporal's host planning (temporalMeta, src/ddm_ref.js:522-560):
// validation, identity/full flags, the exact valid-pair counts N(tau) into np
// (when not null), N(0) and the direct-tail length KT.  No GPU work.
static int plan_temporal(ddm_engine* e, const uint32_t* order, const float* mask, float* m01, uint32_t* ord_out,
                         float* np, bool* identity_out, bool* full_out, float* n0_out, int* KT_out) {
  const int T = e->g.T, L = e->g.L;
  for (int j = 0; j < T; j++)
    if (order[j] >= (uint32_t)T)
      return fail(e, DDM_EINVAL, "computeTemporal: order[" + std::to_string(j) + "] is not a slot < T");
  bool identity = true, full = true;
  for (int j = 0; j < T; j++) {
    identity = identity && order[j] == (uint32_t)j;
    full = full && mask[j] != 0.f;
    if (ord_out) ord_out[j] = order[j];
    if (m01) m01[j] = mask[j] != 0.f ? 1.f : 0.f;
  }
  // exact valid-pair counts N(tau) = sum_t m(t) m(t + tau) (temporalMeta,
  // src/ddm_ref.js:552-556): T - tau when every frame is valid, else popcounts
  // of the mask packed into 64-bit words (T/64 word operations per lag)
  const int nw = (T + 63) / 64;
  std::vector<uint64_t> bits(full ? 0 : nw, 0);
  if (!full)
    for (int j = 0; j < T; j++)
      if (mask[j] != 0.f) bits[j >> 6] |= 1ull << (j & 63);
  auto pairs = [&](int tau) -> int {
    if (full) return T - tau;
    const int ws = tau >> 6, bs = tau & 63;
    int c = 0;
    for (int w = 0; w + ws < nw; w++) {
      uint64_t sh = bits[w + ws] >> bs;   // bit t of sh = m(64 w + t + tau)
      if (bs && w + ws + 1 < nw) sh |= bits[w + ws + 1] << (64 - bs);
      c += __builtin_popcountll(bits[w] & sh);
    }
    return c;
  };
  if (np)
    for (int tau = 0; tau < L; tau++) np[tau] = (float)pairs(tau);
  // direct tail: the lags from the first one (within reach of KT_max) with
  // N(tau) < need on, so that every transform lag keeps N(tau) >= need =
  // max(T/32, min(32, T/4)) (with dropped frames N(tau) is not monotone: a look
  // at the boundary lag alone would leave single-pair lags to the transform)
  int KT = e->KT;
  const int need = std::max((T + 31) / 32, std::min(32, T / 4));
  for (int tau = std::max(1, T - e->KT_max); tau < T - KT; tau++)
    if (pairs(tau) < need) {
      KT = T - tau;
      break;
    }
  *identity_out = identity;
  *full_out = full;
  *n0_out = (float)pairs(0);
  *KT_out = KT;
  return DDM_OK;
}

// The first half of computeTemporal: planning and the metadata uploads on the
// engine's stream (order, 0/1 mask, N(tau)).
int engine_temporal_begin(ddm_engine* e, const uint32_t* order, const float* mask) {
  if (!order || !mask) return fail(e, DDM_EINVAL, "computeTemporal: order/mask must have length T");
  const int T = e->g.T, L = e->g.L;
  // the pinned staging buffers may still feed the previous upload
  hipError_t he = e->capturing ? hipSuccess : hipEventSynchronize(e->meta_ev);
  if (he != hipSuccess) return hip_fail(e, he, "computeTemporal: metadata staging");
  bool identity = true, full = true;
  float n0 = 0.f;
  int KT = e->KT;
  int rc = plan_temporal(e, order, mask, e->meta_mask, e->meta_order, e->meta_np, &identity, &full, &n0, &KT);
  if (rc) return rc;
  e->meta_valid = true;
  e->last_identity = identity;
  e->last_n0 = n0;
  e->cur_KT = KT;
  e->cur_full = full;
  if ((he = hipMemcpyAsync(e->order, e->meta_order, T * 4, hipMemcpyHostToDevice, e->stream)) != hipSuccess ||
      (he = hipMemcpyAsync(e->mask, e->meta_mask, T * 4, hipMemcpyHostToDevice, e->stream)) != hipSuccess ||
      (he = hipMemcpyAsync(e->npairs, e->meta_np, L * 4, hipMemcpyHostToDevice, e->stream)) != hipSuccess ||
      (!e->capturing && (he = hipEventRecord(e->meta_ev, e->stream)) != hipSuccess))
    return hip_fail(e, he, "computeTemporal: metadata upload");
  return DDM_OK;
}

// The radial finalisation of the analysis temporal_begin started, once its
// temporal partials are written (on the engine's stream).
int engine_temporal_end(ddm_engine* e) {
  hipError_t he = timed(e, DDM_K_FINALIZE, [&] {
    // the wavefront temporal kernel splits TPAD > 512 into 2^(logt-9)
    // sub-transforms and stores their spectra r-major
    const int lm = (e->tkern || e->logt <= 12) && e->logt > 9 ? e->logt - 9 : 0;
    return ddm::launch_bin_finalize(e->logt, lm, e->S_part, e->p_part, e->P_part, e->G_part, e->cur_KT,
                                    e->bin_items, e->g.NB, e->nmodes, e->mask, e->cur_full ? 1 : 0, e->npairs,
                                    e->g.T, e->g.L, e->last_n0, e->tw_t64, e->Sbin, e->Gbin, e->pbin, e->Pbin,
                                    e->radD, e->radP, e->stream);
  });
  if (he != hipSuccess) return hip_fail(e, he, "radial finalize launch");
  return DDM_OK;
}

// The wavefront temporal kernel's arguments for (order, mask) -- what
// computeTemporal would launch -- without any GPU work; logt 0 when the
// geometry takes the workgroup kernel instead (TPAD > 4096).
int engine_temporal_plan(ddm_engine* e, const uint32_t* order, const float* mask, TemporalArgs* a, int* logt) {
  bool identity = true, full = true;
  float n0 = 0.f;
  int KT = e->KT;
  int rc = plan_temporal(e, order, mask, nullptr, nullptr, nullptr, &identity, &full, &n0, &KT);
  if (rc) return rc;
  *a = TemporalArgs{e->hist, e->items, e->bin_modes, e->order, e->mask, e->S_part, e->p_part, e->P_part,
                    e->G_part, e->g.T, e->HS, identity ? 1 : 0, KT, 0, n0 > 0.5f ? 1.0f / n0 : 0.0f};
  *logt = e->logt <= 12 || e->tkern ? e->logt : 0;
  return DDM_OK;
}

int engine_tkern(const ddm_engine* e) { return e->tkern; }

int engine_nitems(const ddm_engine* e) { return e->nitems; }
int engine_maxlr(const ddm_engine* e) { return e->maxlr; }
const float2* engine_tw_t(const ddm_engine* e) { return e->tw_t; }
const float2* engine_tw_f(const ddm_engine* e) { return e->tw_f; }

void engine_set_meta(ddm_engine* e, bool identity, float n0) {
  e->meta_valid = true;
  e->last_identity = identity;
  e->last_n0 = n0;
}
}  // namespace ddm
