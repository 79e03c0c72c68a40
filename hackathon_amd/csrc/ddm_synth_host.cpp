// Host side of the synthetic video generator (ddm_synth_* in include/ddm.h).
#include "../../include/ddm.h"
#include "ddm_internal.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

struct ddm_synth {
  hipStream_t stream = nullptr;
  std::string err;
  double4* part = nullptr;
  float2* pos = nullptr;
  long cap_part = 0, cap_pos = 0;
};

namespace {
int sfail(ddm_synth* s, int code, const std::string& msg) {
  s->err = msg;
  return code;
}

int validate(ddm_synth* s, const ddm_video_params* p) {
  if (!p) return sfail(s, DDM_EINVAL, "synth: null parameters");
  if (p->N < 16 || p->N > 4096 || p->N % 16) return sfail(s, DDM_EINVAL, "synth: N must be a multiple of 16 in [16, 4096]");
  if (p->T < 1 || p->T > 65535 || p->n_spots < 0)
    return sfail(s, DDM_EINVAL, "synth: T in [1, 65535] and n_spots >= 0 required");
  if (p->dtype < DDM_U8 || p->dtype > DDM_F32) return sfail(s, DDM_EINVAL, "synth: unknown pixel type");
  if (!(p->sigma > 0) || !(p->alpha >= 0 && p->alpha <= 1) || !(p->vbar >= 0) || !(p->Z >= 0) || !(p->D >= 0) ||
      !(p->background >= 0) || !(p->amp >= 0))
    return sfail(s, DDM_EINVAL, "synth: parameter out of range");
  return DDM_OK;
}

int run(ddm_synth* s, const ddm_video_params* p, void* dframes, float2* dpos) {
  int rc = validate(s, p);
  if (rc) return rc;
  const long n = p->n_spots, npos = (long)p->T * n;
  hipError_t e;
  if (n > s->cap_part) {
    if (s->part) hipFree(s->part);
    s->part = nullptr;
    if ((e = hipMalloc((void**)&s->part, n * sizeof(double4))) != hipSuccess) {
      s->cap_part = 0;
      return sfail(s, DDM_ENOMEM, std::string("synth: ") + hipGetErrorString(e));
    }
    s->cap_part = n;
  }
  float2* pos = dpos;
  if (!pos) {
    if (npos > s->cap_pos) {
      if (s->pos) hipFree(s->pos);
      s->pos = nullptr;
      if ((e = hipMalloc((void**)&s->pos, npos * sizeof(float2))) != hipSuccess) {
        s->cap_pos = 0;
        return sfail(s, DDM_ENOMEM, std::string("synth: ") + hipGetErrorString(e));
      }
      s->cap_pos = npos;
    }
    pos = s->pos;
  }
  const int n_sw = (int)std::floor(p->alpha * n + 0.5);
  e = ddm::launch_synth(p->N, p->T, (int)n, n_sw, p->sigma, p->amp, p->background, p->vbar, p->Z, p->D, p->heading3d, p->swim_diffuse, p->seed,
                        p->dtype, p->noise, s->part, pos, dframes, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  if (e != hipSuccess) return sfail(s, DDM_EHIP, std::string("synth: ") + hipGetErrorString(e));
  return DDM_OK;
}
}  // namespace

extern "C" {

int ddm_synth_create(ddm_synth** out) {
  if (!out) return DDM_EINVAL;
  ddm_synth* s = new ddm_synth();
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
    delete s;
    return DDM_EHIP;
  }
  *out = s;
  return DDM_OK;
}

void ddm_synth_destroy(ddm_synth* s) {
  if (!s) return;
  if (s->stream) hipStreamSynchronize(s->stream);
  if (s->part) hipFree(s->part);
  if (s->pos) hipFree(s->pos);
  if (s->stream) hipStreamDestroy(s->stream);
  delete s;
}

const char* ddm_synth_last_error(const ddm_synth* s) { return s ? s->err.c_str() : "null synth"; }

int ddm_synth_video(ddm_synth* s, const ddm_video_params* p, void* dframes) {
  if (!s || !dframes) return DDM_EINVAL;
  return run(s, p, dframes, nullptr);
}

int ddm_synth_tracks(ddm_synth* s, const ddm_video_params* p, float* dpos) {
  if (!s || !dpos) return DDM_EINVAL;
  return run(s, p, nullptr, reinterpret_cast<float2*>(dpos));
}

}  // extern "C"
