// Host side of the batched per-q LM fitter (ddm_fit in include/ddm.h): packs
// the batch into device arrays (f32 model data, one row of <= 128 lags per
// fit), launches k_lm_fit on the fitter's stream and times it with HIP events.
#include "../../include/ddm.h"
#include "ddm_internal.h"
#include "ddm_isf2d.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

struct ddm_fitter {
  hipStream_t stream = nullptr;
  int device = 0;
  hipError_t noalloc = hipErrorNotInitialized;   // set-up of the allocation-free path on `device`
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::string err;
  size_t cap = 0;  // fits the buffers hold
  float *tau = nullptr, *d = nullptr, *w = nullptr, *qn = nullptr, *bounds = nullptr, *seeds = nullptr,
        *vmag = nullptr;
  int* ntau = nullptr;
  float* theta = nullptr;
  double* chi = nullptr;
  int* steps = nullptr;
  void* lmws = nullptr;   // LM workspace (ranked starts and per-start results)
  // global-parameter fits: device arrays (grown on demand)
  size_t gcap_vid = 0, gcap_bins = 0;
  float *gtau = nullptr, *gd = nullptr, *gw = nullptr, *gqn = nullptr, *gbounds = nullptr, *gseeds = nullptr,
        *gvmag = nullptr;
  int *gboff = nullptr, *gntau = nullptr, *gsteps = nullptr;
  double *gtheta = nullptr, *gchi = nullptr, *gbrown = nullptr, *gab = nullptr;
  // in-plane ISF table (uploaded on the first in-plane fit)
  double *t2v = nullptr, *t2z = nullptr, *t2u0 = nullptr, *t2du = nullptr;
  int *t2off = nullptr, *t2n = nullptr;
  ddm::Isf2dDev t2d{};
};

namespace {
constexpr int kMaxT = 128;
constexpr int kMaxK = 8;

void release(ddm_fitter* f) {
  void* ptrs[] = {f->tau, f->d, f->w, f->qn, f->bounds, f->seeds, f->vmag, f->ntau, f->theta, f->chi, f->steps,
                  f->lmws};
  for (void* p : ptrs)
    if (p) hipFree(p);
  f->tau = f->d = f->w = f->qn = f->bounds = f->seeds = f->vmag = f->theta = nullptr;
  f->ntau = f->steps = nullptr;
  f->chi = nullptr;
  f->lmws = nullptr;
  f->cap = 0;
}

void release_global(ddm_fitter* f) {
  void* ptrs[] = {f->gtau, f->gd, f->gw, f->gqn, f->gbounds, f->gseeds, f->gvmag, f->gboff, f->gntau, f->gsteps,
                  f->gtheta, f->gchi, f->gbrown, f->gab};
  for (void* p : ptrs)
    if (p) hipFree(p);
  f->gtau = f->gd = f->gw = f->gqn = f->gbounds = f->gseeds = f->gvmag = nullptr;
  f->gboff = f->gntau = f->gsteps = nullptr;
  f->gtheta = f->gchi = f->gbrown = f->gab = nullptr;
  f->gcap_vid = f->gcap_bins = 0;
}

hipError_t grow_global(ddm_fitter* f, size_t nv, size_t nb) {
  if (nv <= f->gcap_vid && nb <= f->gcap_bins) return hipSuccess;
  release_global(f);
  nv = std::max(nv, (size_t)1);
  nb = std::max(nb, (size_t)1);
  hipError_t e;
  if ((e = hipMalloc((void**)&f->gtau, nv * kMaxT * 4)) != hipSuccess ||
      (e = hipMalloc((void**)&f->gd, nb * kMaxT * 4)) != hipSuccess ||
      (e = hipMalloc((void**)&f->gw, nb * 4)) != hipSuccess ||
      (e = hipMalloc((void**)&f->gqn, nb * kMaxK * 4)) != hipSuccess ||
      (e = hipMalloc((void**)&f->gbounds, nv * 16)) != hipSuccess ||
      (e = hipMalloc((void**)&f->gseeds, nv * 16)) != hipSuccess ||
      (e = hipMalloc((void**)&f->gvmag, nv * 4)) != hipSuccess ||
      (e = hipMalloc((void**)&f->gboff, (nv + 1) * 4)) != hipSuccess ||
      (e = hipMalloc((void**)&f->gntau, nv * 4)) != hipSuccess ||
      (e = hipMalloc((void**)&f->gsteps, nv * 4)) != hipSuccess ||
      (e = hipMalloc((void**)&f->gtheta, nv * 32)) != hipSuccess ||
      (e = hipMalloc((void**)&f->gchi, nv * 8)) != hipSuccess ||
      (e = hipMalloc((void**)&f->gbrown, nv * 16)) != hipSuccess ||
      (e = hipMalloc((void**)&f->gab, nb * 16)) != hipSuccess) {
    release_global(f);
    return e;
  }
  f->gcap_vid = nv;
  f->gcap_bins = nb;
  return hipSuccess;
}

void free_isf2d(ddm_fitter* f) {
  for (void** p : {(void**)&f->t2v, (void**)&f->t2z, (void**)&f->t2u0, (void**)&f->t2du, (void**)&f->t2off,
                   (void**)&f->t2n}) {
    if (*p) hipFree(*p);
    *p = nullptr;
  }
}

// The in-plane ISF table of ddm_isf2d.h (its row layout), built in device
// memory once by the FP64 quadrature on the GPU (a failed build leaves nothing
// behind, so the next in-plane fit retries).
hipError_t upload_isf2d(ddm_fitter* f) {
  const ddm::Isf2dTable& T = ddm::isf2d_table();
  const size_t nz = T.Z.size();
  hipError_t e;
  if ((e = hipMalloc((void**)&f->t2v, T.v.size() * 8)) != hipSuccess ||
      (e = hipMalloc((void**)&f->t2z, nz * 8)) != hipSuccess || (e = hipMalloc((void**)&f->t2u0, nz * 8)) != hipSuccess ||
      (e = hipMalloc((void**)&f->t2du, nz * 8)) != hipSuccess || (e = hipMalloc((void**)&f->t2off, nz * 4)) != hipSuccess ||
      (e = hipMalloc((void**)&f->t2n, nz * 4)) != hipSuccess)
    return free_isf2d(f), e;
  if ((e = ddm::build_isf2d_device(T, f->t2v, f->t2z, f->t2u0, f->t2du, f->t2off, f->t2n, f->stream)) != hipSuccess)
    return free_isf2d(f), e;
  f->t2d = ddm::Isf2dDev{f->t2v, f->t2z, f->t2u0, f->t2du, f->t2off, f->t2n, T.lz0, T.dlz};
  return hipSuccess;
}

int ffail(ddm_fitter* f, int code, const std::string& m) {
  f->err = m;
  return code;
}

hipError_t grow(ddm_fitter* f, size_t n) {
  if (n <= f->cap) return hipSuccess;
  release(f);
  hipError_t e;
  if ((e = hipMalloc((void**)&f->tau, n * kMaxT * 4)) != hipSuccess) return e;
  if ((e = hipMalloc((void**)&f->d, n * kMaxT * 4)) != hipSuccess) return e;
  if ((e = hipMalloc((void**)&f->w, n * 4)) != hipSuccess) return e;
  if ((e = hipMalloc((void**)&f->qn, n * kMaxK * 4)) != hipSuccess) return e;
  if ((e = hipMalloc((void**)&f->bounds, n * 16)) != hipSuccess) return e;
  if ((e = hipMalloc((void**)&f->seeds, n * 24)) != hipSuccess) return e;
  if ((e = hipMalloc((void**)&f->vmag, n * 4)) != hipSuccess) return e;
  if ((e = hipMalloc((void**)&f->ntau, n * 4)) != hipSuccess) return e;
  if ((e = hipMalloc((void**)&f->theta, n * 24)) != hipSuccess) return e;
  if ((e = hipMalloc((void**)&f->chi, n * 8)) != hipSuccess) return e;
  if ((e = hipMalloc((void**)&f->steps, n * 4)) != hipSuccess) return e;
  if ((e = hipMalloc(&f->lmws, ddm::lm_fit_workspace_bytes((int)n))) != hipSuccess) return e;
  f->cap = n;
  return hipSuccess;
}
}  // namespace

extern "C" {

int ddm_fitter_create(ddm_fitter** out) {
  if (!out) return DDM_EINVAL;
  ddm_fitter* f = new ddm_fitter();
  if (hipGetDevice(&f->device) != hipSuccess || hipStreamCreateWithFlags(&f->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&f->ev0) != hipSuccess || hipEventCreate(&f->ev1) != hipSuccess) {
    delete f;
    return DDM_EHIP;
  }
  // the allocation-free path's fixed scratch on this device (its failure is
  // reported by ddm_fit_device_noalloc; ddm_fit does not need it)
  f->noalloc = ddm::noalloc_setup();
  *out = f;
  return DDM_OK;
}

void ddm_fitter_destroy(ddm_fitter* f) {
  if (!f) return;
  hipStreamSynchronize(f->stream);
  release(f);
  release_global(f);
  free_isf2d(f);
  hipEventDestroy(f->ev0);
  hipEventDestroy(f->ev1);
  hipStreamDestroy(f->stream);
  delete f;
}

const char* ddm_fitter_last_error(const ddm_fitter* f) { return f ? f->err.c_str() : "null fitter"; }

int ddm_fit(ddm_fitter* f, const ddm_fit_batch* b, double* theta, double* chi2, int32_t* iters, double* kernel_ms) {
  if (!f || !b) return DDM_EINVAL;
  if (b->nfits < 0) return ffail(f, DDM_EINVAL, "fit: negative batch size");
  if (b->nfits == 0) return DDM_OK;
  if (b->ntau_max < 1 || b->ntau_max > kMaxT) return ffail(f, DDM_EINVAL, "fit: ntau_max must be in [1, 128]");
  if (b->knodes < 1 || b->knodes > kMaxK) return ffail(f, DDM_EINVAL, "fit: knodes must be in [1, 8]");
  if (b->model != DDM_MODEL_BROWNIAN && b->model != DDM_MODEL_SWIMMER_3D && b->model != DDM_MODEL_SWIMMER_2D)
    return ffail(f, DDM_EINVAL, "fit: model");
  if (b->seed_mode != DDM_SEED_GRID && b->seed_mode != DDM_SEED_GIVEN && b->seed_mode != DDM_SEED_NONLINEAR)
    return ffail(f, DDM_EINVAL, "fit: seed_mode");
  const bool seeded = b->seed_mode != DDM_SEED_GRID;
  if (seeded && !b->seeds) return ffail(f, DDM_EINVAL, "fit: seeds required");
  if (!b->ntau || !b->tau || !b->data || !b->weight || !b->qnodes || !b->bounds || !theta || !chi2 || !iters)
    return ffail(f, DDM_EINVAL, "fit: null array");
  const size_t n = (size_t)b->nfits;
  std::vector<float> tau(n * kMaxT, 0.f), d(n * kMaxT, 0.f), w(n), qn(n * b->knodes), bd(n * 4), sd(n * 6, 0.f),
      vm(b->vmag ? n : 0);
  std::vector<int> nt(n);
  for (size_t i = 0; i < n; i++) {
    const int m = b->ntau[i];
    if (m < 1 || m > b->ntau_max) return ffail(f, DDM_EINVAL, "fit: ntau[i] out of range");
    nt[i] = m;
    for (int j = 0; j < m; j++) {
      tau[i * kMaxT + j] = (float)b->tau[i * b->ntau_max + j];
      d[i * kMaxT + j] = (float)b->data[i * b->ntau_max + j];
    }
    w[i] = (float)b->weight[i];
    for (int k = 0; k < b->knodes; k++) qn[i * b->knodes + k] = (float)b->qnodes[i * b->knodes + k];
    for (int k = 0; k < 4; k++) bd[i * 4 + k] = (float)b->bounds[i * 4 + k];
    if (!(bd[i * 4] > 0 && bd[i * 4 + 1] > bd[i * 4] && bd[i * 4 + 2] > 0 && bd[i * 4 + 3] > bd[i * 4 + 2]))
      return ffail(f, DDM_EINVAL, "fit: bounds must satisfy 0 < vmin < vmax, 0 < Dmin < Dmax");
    if (seeded)
      for (int k = 0; k < 6; k++) sd[i * 6 + k] = (float)b->seeds[i * 6 + k];
    if (b->vmag) {
      if (!(b->vmag[i] >= 0 && std::isfinite(b->vmag[i]))) return ffail(f, DDM_EINVAL, "fit: vmag must be >= 0");
      vm[i] = (float)b->vmag[i];
    }
  }
  hipError_t e = grow(f, n);
  if (e != hipSuccess) return ffail(f, DDM_EHIP, std::string("fit: allocation: ") + hipGetErrorString(e));
  auto up = [&](void* dst, const void* src, size_t bytes) {
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, f->stream);
  };
  if ((e = up(f->tau, tau.data(), tau.size() * 4)) != hipSuccess || (e = up(f->d, d.data(), d.size() * 4)) != hipSuccess ||
      (e = up(f->w, w.data(), n * 4)) != hipSuccess || (e = up(f->qn, qn.data(), qn.size() * 4)) != hipSuccess ||
      (e = up(f->bounds, bd.data(), n * 16)) != hipSuccess || (e = up(f->seeds, sd.data(), n * 24)) != hipSuccess ||
      (e = up(f->ntau, nt.data(), n * 4)) != hipSuccess || (b->vmag && (e = up(f->vmag, vm.data(), n * 4)) != hipSuccess))
    return ffail(f, DDM_EHIP, std::string("fit: upload: ") + hipGetErrorString(e));
  if (b->model == DDM_MODEL_SWIMMER_2D && !f->t2v && (e = upload_isf2d(f)) != hipSuccess)
    return ffail(f, DDM_EHIP, std::string("fit: in-plane ISF table: ") + hipGetErrorString(e));
  hipEventRecord(f->ev0, f->stream);
  e = ddm::launch_lm_fit(f->tau, f->d, f->ntau, f->w, f->qn, b->knodes, f->bounds, f->seeds,
                         b->vmag ? f->vmag : nullptr, b->nfits, b->model, &f->t2d, b->seed_mode,
                         std::max(1, std::min(3, b->nstarts)), std::max(0, b->max_iter), f->theta, f->chi, f->steps,
                         f->lmws, f->stream);
  if (e != hipSuccess) return ffail(f, DDM_EHIP, std::string("fit: launch: ") + hipGetErrorString(e));
  hipEventRecord(f->ev1, f->stream);
  std::vector<float> th(n * 6);
  std::vector<int> st(n);
  if ((e = hipMemcpyAsync(th.data(), f->theta, n * 24, hipMemcpyDeviceToHost, f->stream)) != hipSuccess ||
      (e = hipMemcpyAsync(chi2, f->chi, n * 8, hipMemcpyDeviceToHost, f->stream)) != hipSuccess ||
      (e = hipMemcpyAsync(st.data(), f->steps, n * 4, hipMemcpyDeviceToHost, f->stream)) != hipSuccess ||
      (e = hipStreamSynchronize(f->stream)) != hipSuccess)
    return ffail(f, DDM_EHIP, std::string("fit: ") + hipGetErrorString(e));
  for (size_t i = 0; i < n * 6; i++) theta[i] = th[i];
  for (size_t i = 0; i < n; i++) iters[i] = st[i];
  if (kernel_ms) {
    float ms = 0;
    hipEventElapsedTime(&ms, f->ev0, f->ev1);
    *kernel_ms = ms;
  }
  return DDM_OK;
}

int ddm_fit_global(ddm_fitter* f, const ddm_global_fit_batch* b, double* theta, double* rss, int32_t* iters,
                   double* brown, double* amps, double* kernel_ms) {
  if (!f || !b) return DDM_EINVAL;
  if (b->nvideos < 0) return ffail(f, DDM_EINVAL, "global fit: negative batch size");
  if (b->nvideos == 0) return DDM_OK;
  if (b->ntau_max < 1 || b->ntau_max > kMaxT) return ffail(f, DDM_EINVAL, "global fit: ntau_max must be in [1, 128]");
  if (b->knodes < 1 || b->knodes > kMaxK) return ffail(f, DDM_EINVAL, "global fit: knodes must be in [1, 8]");
  if (b->model != DDM_MODEL_BROWNIAN && b->model != DDM_MODEL_SWIMMER_3D && b->model != DDM_MODEL_SWIMMER_2D)
    return ffail(f, DDM_EINVAL, "global fit: model");
  if (b->seed_mode != DDM_SEED_GRID && b->seed_mode != DDM_SEED_GIVEN)
    return ffail(f, DDM_EINVAL, "global fit: seed_mode must be DDM_SEED_GRID or DDM_SEED_GIVEN");
  if (b->seed_mode == DDM_SEED_GIVEN && !b->seeds) return ffail(f, DDM_EINVAL, "global fit: seeds required");
  if (!b->bin_offset || !b->ntau || !b->tau || !b->data || !b->weight || !b->qnodes || !b->bounds || !theta || !rss ||
      !iters)
    return ffail(f, DDM_EINVAL, "global fit: null array");
  const size_t nv = (size_t)b->nvideos;
  if (b->bin_offset[0] != 0) return ffail(f, DDM_EINVAL, "global fit: bin_offset[0] must be 0");
  for (size_t v = 0; v < nv; v++) {
    const int nbv = b->bin_offset[v + 1] - b->bin_offset[v];
    if (nbv < 1 || nbv > DDM_GLOBAL_MAX_BINS)
      return ffail(f, DDM_EINVAL, "global fit: every video needs 1.." + std::to_string(DDM_GLOBAL_MAX_BINS) + " bins");
    if (b->ntau[v] < 1 || b->ntau[v] > b->ntau_max) return ffail(f, DDM_EINVAL, "global fit: ntau[v] out of range");
    const double* bd = b->bounds + 4 * v;
    if (!(bd[0] > 0 && bd[1] > bd[0] && bd[2] > 0 && bd[3] > bd[2]))
      return ffail(f, DDM_EINVAL, "global fit: bounds must satisfy 0 < vmin < vmax, 0 < Dmin < Dmax");
    if (b->vmag && !(b->vmag[v] >= 0 && std::isfinite(b->vmag[v])))
      return ffail(f, DDM_EINVAL, "global fit: vmag must be >= 0");
  }
  const size_t nb = (size_t)b->bin_offset[nv];
  std::vector<float> tau(nv * kMaxT, 0.f), d(nb * kMaxT, 0.f), w(nb), qn(nb * b->knodes), bd(nv * 4), sd(nv * 4, 0.f),
      vm(nv, 0.f);
  std::vector<int> nt(nv), boff(nv + 1);
  for (size_t v = 0; v < nv; v++) {
    nt[v] = b->ntau[v];
    for (int j = 0; j < nt[v]; j++) tau[v * kMaxT + j] = (float)b->tau[v * b->ntau_max + j];
    for (int k = 0; k < 4; k++) bd[v * 4 + k] = (float)b->bounds[v * 4 + k];
    if (b->seeds)
      for (int k = 0; k < 4; k++) sd[v * 4 + k] = (float)b->seeds[v * 4 + k];
    if (b->vmag) vm[v] = (float)b->vmag[v];
    for (int i = b->bin_offset[v]; i < b->bin_offset[v + 1]; i++)
      for (int j = 0; j < nt[v]; j++) d[(size_t)i * kMaxT + j] = (float)b->data[(size_t)i * b->ntau_max + j];
  }
  for (size_t v = 0; v <= nv; v++) boff[v] = b->bin_offset[v];
  for (size_t i = 0; i < nb; i++) {
    w[i] = (float)b->weight[i];
    for (int k = 0; k < b->knodes; k++) qn[i * b->knodes + k] = (float)b->qnodes[i * b->knodes + k];
  }
  // small plain swimmer batches: one workgroup per (video, start), so a batch
  // of few videos is not held up by each video's starts in sequence
  const int ns = std::max(1, std::min(3, b->nstarts));
  const int split = (b->seed_mode == DDM_SEED_GRID && b->model != DDM_MODEL_BROWNIAN && ns > 1 && nv * ns <= 512)
                        ? ns : 0;
  const size_t slots = split ? nv * split : nv;
  hipError_t e = grow_global(f, slots, split ? nb * split : nb);
  if (e != hipSuccess) return ffail(f, DDM_EHIP, std::string("global fit: allocation: ") + hipGetErrorString(e));
  auto up = [&](void* dst, const void* src, size_t bytes) {
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, f->stream);
  };
  if ((e = up(f->gtau, tau.data(), tau.size() * 4)) != hipSuccess || (e = up(f->gd, d.data(), d.size() * 4)) != hipSuccess ||
      (e = up(f->gw, w.data(), nb * 4)) != hipSuccess || (e = up(f->gqn, qn.data(), qn.size() * 4)) != hipSuccess ||
      (e = up(f->gbounds, bd.data(), nv * 16)) != hipSuccess || (e = up(f->gseeds, sd.data(), nv * 16)) != hipSuccess ||
      (e = up(f->gvmag, vm.data(), nv * 4)) != hipSuccess || (e = up(f->gboff, boff.data(), (nv + 1) * 4)) != hipSuccess ||
      (e = up(f->gntau, nt.data(), nv * 4)) != hipSuccess)
    return ffail(f, DDM_EHIP, std::string("global fit: upload: ") + hipGetErrorString(e));
  if (b->model == DDM_MODEL_SWIMMER_2D && !f->t2v && (e = upload_isf2d(f)) != hipSuccess)
    return ffail(f, DDM_EHIP, std::string("global fit: in-plane ISF table: ") + hipGetErrorString(e));
  hipEventRecord(f->ev0, f->stream);
  e = ddm::launch_lm_global(f->gtau, f->gd, f->gw, f->gqn, b->knodes, f->gboff, f->gntau, f->gbounds,
                            b->seed_mode == DDM_SEED_GIVEN ? f->gseeds : nullptr, b->vmag ? f->gvmag : nullptr,
                            (int)nv, b->model, &f->t2d, b->seed_mode == DDM_SEED_GIVEN ? 0 : 1,
                            ns, std::max(0, b->max_iter), split, f->gtheta, f->gchi, f->gsteps, f->gbrown, f->gab,
                            f->stream);
  if (e != hipSuccess) return ffail(f, DDM_EHIP, std::string("global fit: launch: ") + hipGetErrorString(e));
  hipEventRecord(f->ev1, f->stream);
  const size_t nab = split ? nb * split : nb;
  std::vector<int> st(slots);
  std::vector<double> th(slots * 4), ch(slots), ab(amps ? nab * 2 : 0), br(brown ? nv * 2 : 0);
  if ((e = hipMemcpyAsync(th.data(), f->gtheta, slots * 32, hipMemcpyDeviceToHost, f->stream)) != hipSuccess ||
      (e = hipMemcpyAsync(ch.data(), f->gchi, slots * 8, hipMemcpyDeviceToHost, f->stream)) != hipSuccess ||
      (e = hipMemcpyAsync(st.data(), f->gsteps, slots * 4, hipMemcpyDeviceToHost, f->stream)) != hipSuccess ||
      (amps && (e = hipMemcpyAsync(ab.data(), f->gab, nab * 16, hipMemcpyDeviceToHost, f->stream)) != hipSuccess) ||
      (brown && (e = hipMemcpyAsync(br.data(), f->gbrown, nv * 16, hipMemcpyDeviceToHost, f->stream)) != hipSuccess) ||
      (e = hipStreamSynchronize(f->stream)) != hipSuccess)
    return ffail(f, DDM_EHIP, std::string("global fit: ") + hipGetErrorString(e));
  for (size_t v = 0; v < nv; v++) {
    // the best start (the first of equal ones, as the sequential loop keeps it)
    const int per = split ? split : 1;
    int best = 0, total = 0;
    for (int s2 = 0; s2 < per; s2++) {
      total += st[v * per + s2];
      if (ch[v * per + s2] < ch[v * per + best]) best = s2;
    }
    const size_t sl = v * per + best;
    for (int k = 0; k < 4; k++) theta[v * 4 + k] = th[sl * 4 + k];
    rss[v] = ch[sl];
    iters[v] = total;
    if (amps) {
      const size_t b0 = b->bin_offset[v], nbv = b->bin_offset[v + 1] - b0;
      const size_t src = split ? 2 * (b0 * split + best * nbv) : 2 * b0;
      std::copy(ab.begin() + src, ab.begin() + src + 2 * nbv, amps + 2 * b0);
    }
  }
  if (brown) {
    if (b->seed_mode == DDM_SEED_GRID) std::copy(br.begin(), br.end(), brown);
    else std::fill(brown, brown + nv * 2, NAN);
  }
  if (kernel_ms) {
    float ms = 0;
    hipEventElapsedTime(&ms, f->ev0, f->ev1);
    *kernel_ms = ms;
  }
  return DDM_OK;
}

int ddm_isf2d_eval(ddm_fitter* f, const double* x, const double* Z, int32_t n, int32_t exact, double* out) {
  if (!f) return DDM_EINVAL;
  if (n < 0 || (n > 0 && (!x || !Z || !out))) return ffail(f, DDM_EINVAL, "isf2d: bad arguments");
  if (n == 0) return DDM_OK;
  hipError_t e;
  if (!f->t2v && (e = upload_isf2d(f)) != hipSuccess)
    return ffail(f, DDM_EHIP, std::string("isf2d: table: ") + hipGetErrorString(e));
  double* d = nullptr;
  if ((e = hipMalloc((void**)&d, (size_t)n * 24)) != hipSuccess)
    return ffail(f, DDM_EHIP, std::string("isf2d: allocation: ") + hipGetErrorString(e));
  if ((e = hipMemcpyAsync(d, x, (size_t)n * 8, hipMemcpyHostToDevice, f->stream)) != hipSuccess ||
      (e = hipMemcpyAsync(d + n, Z, (size_t)n * 8, hipMemcpyHostToDevice, f->stream)) != hipSuccess ||
      (e = ddm::launch_isf2d_eval(d, d + n, n, exact ? 1 : 0, &f->t2d, d + 2 * (size_t)n, f->stream)) != hipSuccess ||
      (e = hipMemcpyAsync(out, d + 2 * (size_t)n, (size_t)n * 8, hipMemcpyDeviceToHost, f->stream)) != hipSuccess ||
      (e = hipStreamSynchronize(f->stream)) != hipSuccess) {
    hipFree(d);
    return ffail(f, DDM_EHIP, std::string("isf2d: ") + hipGetErrorString(e));
  }
  hipFree(d);
  return DDM_OK;
}

int ddm_fit_device_noalloc(ddm_fitter* f, const ddm_fit_batch* b, double* theta, double* chi2, int32_t* iters,
                           int32_t wait) {
  if (!f || !b) return DDM_EINVAL;
  if (b->nfits < 0 || b->nfits > DDM_FIT_NOALLOC_MAX)
    return ffail(f, DDM_EINVAL, "fit (no-alloc): nfits must be in [0, " + std::to_string(DDM_FIT_NOALLOC_MAX) + "]");
  if (b->nfits == 0) return DDM_OK;
  if (b->ntau_max < 1 || b->ntau_max > kMaxT) return ffail(f, DDM_EINVAL, "fit (no-alloc): ntau_max must be in [1, 128]");
  if (b->knodes < 1 || b->knodes > kMaxK) return ffail(f, DDM_EINVAL, "fit (no-alloc): knodes must be in [1, 8]");
  if (b->model != DDM_MODEL_BROWNIAN && b->model != DDM_MODEL_SWIMMER_3D && b->model != DDM_MODEL_SWIMMER_2D)
    return ffail(f, DDM_EINVAL, "fit (no-alloc): model");
  if (b->seed_mode != DDM_SEED_GRID && b->seed_mode != DDM_SEED_GIVEN && b->seed_mode != DDM_SEED_NONLINEAR)
    return ffail(f, DDM_EINVAL, "fit (no-alloc): seed_mode");
  if (b->seed_mode != DDM_SEED_GRID && !b->seeds) return ffail(f, DDM_EINVAL, "fit (no-alloc): seeds required");
  if (!b->ntau || !b->tau || !b->data || !b->weight || !b->qnodes || !b->bounds || !theta || !chi2 || !iters)
    return ffail(f, DDM_EINVAL, "fit (no-alloc): null array");
  if (f->noalloc != hipSuccess)
    return ffail(f, DDM_EHIP, std::string("fit (no-alloc): scratch set-up failed on device ") +
                                  std::to_string(f->device) + ": " + hipGetErrorString(f->noalloc));
  hipError_t e = ddm::launch_fit_noalloc(f->device, b, theta, chi2, iters, f->stream);
  if (e != hipSuccess) return ffail(f, DDM_EHIP, std::string("fit (no-alloc): launch: ") + hipGetErrorString(e));
  if (wait && (e = hipStreamSynchronize(f->stream)) != hipSuccess)
    return ffail(f, DDM_EHIP, std::string("fit (no-alloc): ") + hipGetErrorString(e));
  return DDM_OK;
}

}  // extern "C"
