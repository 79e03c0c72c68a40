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
  // in-plane ISF table (uploaded on the first in-plane fit)
  double *t2v = nullptr, *t2z = nullptr, *t2u0 = nullptr, *t2du = nullptr;
  int *t2off = nullptr, *t2n = nullptr;
  ddm::Isf2dDev t2d{};
};

namespace {
constexpr int kMaxT = 128;
constexpr int kMaxK = 8;

void release(ddm_fitter* f) {
  void* ptrs[] = {f->tau, f->d, f->w, f->qn, f->bounds, f->seeds, f->vmag, f->ntau, f->theta, f->chi, f->steps};
  for (void* p : ptrs)
    if (p) hipFree(p);
  f->tau = f->d = f->w = f->qn = f->bounds = f->seeds = f->vmag = f->theta = nullptr;
  f->ntau = f->steps = nullptr;
  f->chi = nullptr;
  f->cap = 0;
}

void free_isf2d(ddm_fitter* f) {
  for (void** p : {(void**)&f->t2v, (void**)&f->t2z, (void**)&f->t2u0, (void**)&f->t2du, (void**)&f->t2off,
                   (void**)&f->t2n}) {
    if (*p) hipFree(*p);
    *p = nullptr;
  }
}

// The in-plane ISF table of ddm_isf2d.h, copied to device memory once (a
// failed upload leaves nothing behind, so the next in-plane fit retries).
hipError_t upload_isf2d(ddm_fitter* f) {
  const ddm::Isf2dTable& T = ddm::isf2d_table();
  const size_t nz = T.Z.size();
  hipError_t e;
  if ((e = hipMalloc((void**)&f->t2v, T.v.size() * 8)) != hipSuccess ||
      (e = hipMalloc((void**)&f->t2z, nz * 8)) != hipSuccess || (e = hipMalloc((void**)&f->t2u0, nz * 8)) != hipSuccess ||
      (e = hipMalloc((void**)&f->t2du, nz * 8)) != hipSuccess || (e = hipMalloc((void**)&f->t2off, nz * 4)) != hipSuccess ||
      (e = hipMalloc((void**)&f->t2n, nz * 4)) != hipSuccess)
    return free_isf2d(f), e;
  if ((e = hipMemcpy(f->t2v, T.v.data(), T.v.size() * 8, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(f->t2z, T.Z.data(), nz * 8, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(f->t2u0, T.u0.data(), nz * 8, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(f->t2du, T.du.data(), nz * 8, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(f->t2off, T.off.data(), nz * 4, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(f->t2n, T.n.data(), nz * 4, hipMemcpyHostToDevice)) != hipSuccess)
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
                         f->stream);
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
