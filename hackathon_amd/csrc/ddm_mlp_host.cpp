// Host side of the LM-seeding MLP (ddm_mlp_* in include/ddm.h): parameter and
// activation buffers, forward / backward launch sequences, Adam.
#include "../../include/ddm.h"
#include "ddm_internal.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

namespace {
constexpr int IN = DDM_MLP_IN, HID = DDM_MLP_HIDDEN, OUT = DDM_MLP_OUT;
constexpr int NL = 4;
constexpr int DIN[NL] = {IN, HID, HID, HID};
constexpr int DOUT[NL] = {HID, HID, HID, OUT};

long woff(int l) {
  long o = 0;
  for (int i = 0; i < l; i++) o += (long)DIN[i] * DOUT[i] + DOUT[i];
  return o;
}
long boff(int l) { return woff(l) + (long)DIN[l] * DOUT[l]; }
}  // namespace

struct ddm_mlp {
  int maxb = 0;
  hipStream_t stream = nullptr;
  std::string err;
  long step = 0;
  float *p = nullptr, *g = nullptr, *mo = nullptr, *vo = nullptr;  // FP32 master, grads, Adam moments
  unsigned short* pb = nullptr;                                    // BF16 compute copy
  float* xf = nullptr;        // input f32 [maxb][IN]
  unsigned short* xb = nullptr;
  float* Z[NL] = {};          // pre-activations f32
  unsigned short* H[NL] = {}; // activations bf16 (H[3] unused)
  float *tg = nullptr, *wt = nullptr, *dY = nullptr, *rowloss = nullptr;
  float *dH = nullptr, *dZf = nullptr;
  unsigned short* dZb = nullptr;
};

namespace {
int mfail(ddm_mlp* m, int code, const std::string& msg) {
  m->err = msg;
  return code;
}
#define MCHK(m, call)                                                                  \
  do {                                                                                 \
    hipError_t _e = (call);                                                            \
    if (_e != hipSuccess) return mfail(m, DDM_EHIP, std::string(#call ": ") + hipGetErrorString(_e)); \
  } while (0)

int forward(ddm_mlp* m, int M) {
  const unsigned short* in = m->xb;
  long ld = IN;
  for (int l = 0; l < NL; l++) {
    MCHK(m, ddm::launch_gemm_bf16(in, ld, 1, m->pb + woff(l), DOUT[l], 1, m->Z[l], DOUT[l], M, DOUT[l], DIN[l], 0,
                                  m->stream));
    MCHK(m, ddm::launch_bias_act(m->Z[l], m->p + boff(l), m->H[l], M, DOUT[l], l < NL - 1, m->stream));
    in = m->H[l];
    ld = DOUT[l];
  }
  return DDM_OK;
}

int backward(ddm_mlp* m, int M) {
  // dY (f32, M x OUT) is in m->dY; walk the layers down
  const float* dHl = m->dY;
  for (int l = NL - 1; l >= 0; l--) {
    const int N = DOUT[l], K = DIN[l];
    MCHK(m, ddm::launch_act_bwd(dHl, m->Z[l], m->dZb, m->dZf, m->g + boff(l), M, N, l < NL - 1, m->stream));
    const unsigned short* Hin = l ? m->H[l - 1] : m->xb;
    // dW = Hin^T dZ: A(i=k_in, k=row) = Hin[row * K + k_in]
    MCHK(m, ddm::launch_gemm_bf16(Hin, 1, K, m->dZb, N, 1, m->g + woff(l), N, K, N, M, 0, m->stream));
    if (l > 0) {
      // dH_{l-1} = dZ W^T: B(k=j_out, j=k_in) = W[k_in * N + j_out]
      MCHK(m, ddm::launch_gemm_bf16(m->dZb, N, 1, m->pb + woff(l), 1, N, m->dH, K, M, K, N, 0, m->stream));
      dHl = m->dH;
    }
  }
  return DDM_OK;
}

int upload_batch(ddm_mlp* m, const float* X, const float* Tg, const float* wt, int M) {
  if (M < 1 || M > m->maxb) return mfail(m, DDM_EINVAL, "mlp: batch size out of range");
  MCHK(m, hipMemcpyAsync(m->xf, X, (size_t)M * IN * 4, hipMemcpyHostToDevice, m->stream));
  MCHK(m, ddm::launch_to_bf16(m->xf, m->xb, (long)M * IN, m->stream));
  if (Tg) MCHK(m, hipMemcpyAsync(m->tg, Tg, (size_t)M * OUT * 4, hipMemcpyHostToDevice, m->stream));
  if (wt) MCHK(m, hipMemcpyAsync(m->wt, wt, OUT * 4, hipMemcpyHostToDevice, m->stream));
  return DDM_OK;
}

int loss_and_backward(ddm_mlp* m, int M, float* loss) {
  int rc;
  if ((rc = forward(m, M))) return rc;
  MCHK(m, ddm::launch_mse(m->Z[NL - 1], m->tg, m->wt, m->dY, m->rowloss, M, OUT, m->stream));
  if ((rc = backward(m, M))) return rc;
  if (loss) {
    std::vector<float> rl(M);
    MCHK(m, hipMemcpyAsync(rl.data(), m->rowloss, M * 4, hipMemcpyDeviceToHost, m->stream));
    MCHK(m, hipStreamSynchronize(m->stream));
    double s = 0;
    for (float v : rl) s += v;
    *loss = (float)(s / M);
  }
  return DDM_OK;
}
}  // namespace

extern "C" {

int ddm_mlp_create(int32_t max_batch, ddm_mlp** out) {
  if (!out || max_batch < 1) return DDM_EINVAL;
  ddm_mlp* m = new ddm_mlp();
  m->maxb = max_batch;
  const long P = DDM_MLP_PARAMS, B = max_batch;
  bool ok = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) == hipSuccess &&
            hipMalloc((void**)&m->p, P * 4) == hipSuccess && hipMalloc((void**)&m->g, P * 4) == hipSuccess &&
            hipMalloc((void**)&m->mo, P * 4) == hipSuccess && hipMalloc((void**)&m->vo, P * 4) == hipSuccess &&
            hipMalloc((void**)&m->pb, P * 2) == hipSuccess && hipMalloc((void**)&m->xf, B * IN * 4) == hipSuccess &&
            hipMalloc((void**)&m->xb, B * IN * 2) == hipSuccess && hipMalloc((void**)&m->tg, B * OUT * 4) == hipSuccess &&
            hipMalloc((void**)&m->wt, OUT * 4) == hipSuccess && hipMalloc((void**)&m->dY, B * OUT * 4) == hipSuccess &&
            hipMalloc((void**)&m->rowloss, B * 4) == hipSuccess && hipMalloc((void**)&m->dH, B * HID * 4) == hipSuccess &&
            hipMalloc((void**)&m->dZf, B * HID * 4) == hipSuccess && hipMalloc((void**)&m->dZb, B * HID * 2) == hipSuccess;
  for (int l = 0; l < NL && ok; l++)
    ok = hipMalloc((void**)&m->Z[l], B * DOUT[l] * 4) == hipSuccess &&
         hipMalloc((void**)&m->H[l], B * DOUT[l] * 2) == hipSuccess;
  // on the network's own (non-blocking) stream, which a null-stream memset would not be ordered before
  if (ok)
    ok = hipMemsetAsync(m->mo, 0, P * 4, m->stream) == hipSuccess &&
         hipMemsetAsync(m->vo, 0, P * 4, m->stream) == hipSuccess &&
         hipMemsetAsync(m->p, 0, P * 4, m->stream) == hipSuccess &&
         hipMemsetAsync(m->pb, 0, P * 2, m->stream) == hipSuccess && hipStreamSynchronize(m->stream) == hipSuccess;
  if (!ok) {
    ddm_mlp_destroy(m);
    return DDM_ENOMEM;
  }
  *out = m;
  return DDM_OK;
}

void ddm_mlp_destroy(ddm_mlp* m) {
  if (!m) return;
  if (m->stream) hipStreamSynchronize(m->stream);
  void* ptrs[] = {m->p, m->g, m->mo, m->vo, m->pb, m->xf, m->xb, m->tg, m->wt, m->dY, m->rowloss, m->dH, m->dZf, m->dZb};
  for (void* q : ptrs)
    if (q) hipFree(q);
  for (int l = 0; l < NL; l++) {
    if (m->Z[l]) hipFree(m->Z[l]);
    if (m->H[l]) hipFree(m->H[l]);
  }
  if (m->stream) hipStreamDestroy(m->stream);
  delete m;
}

const char* ddm_mlp_last_error(const ddm_mlp* m) { return m ? m->err.c_str() : "null mlp"; }

int ddm_mlp_set_params(ddm_mlp* m, const float* params) {
  if (!m || !params) return DDM_EINVAL;
  MCHK(m, hipMemcpyAsync(m->p, params, (size_t)DDM_MLP_PARAMS * 4, hipMemcpyHostToDevice, m->stream));
  MCHK(m, ddm::launch_to_bf16(m->p, m->pb, DDM_MLP_PARAMS, m->stream));
  MCHK(m, hipMemsetAsync(m->mo, 0, (size_t)DDM_MLP_PARAMS * 4, m->stream));
  MCHK(m, hipMemsetAsync(m->vo, 0, (size_t)DDM_MLP_PARAMS * 4, m->stream));
  m->step = 0;
  MCHK(m, hipStreamSynchronize(m->stream));
  return DDM_OK;
}

int ddm_mlp_get_params(ddm_mlp* m, float* params) {
  if (!m || !params) return DDM_EINVAL;
  MCHK(m, hipMemcpyAsync(params, m->p, (size_t)DDM_MLP_PARAMS * 4, hipMemcpyDeviceToHost, m->stream));
  MCHK(m, hipStreamSynchronize(m->stream));
  return DDM_OK;
}

int ddm_mlp_forward(ddm_mlp* m, const float* X, int32_t M, float* Y) {
  if (!m || !X || !Y) return DDM_EINVAL;
  int rc;
  if ((rc = upload_batch(m, X, nullptr, nullptr, M))) return rc;
  if ((rc = forward(m, M))) return rc;
  MCHK(m, hipMemcpyAsync(Y, m->Z[NL - 1], (size_t)M * OUT * 4, hipMemcpyDeviceToHost, m->stream));
  MCHK(m, hipStreamSynchronize(m->stream));
  return DDM_OK;
}

int ddm_mlp_loss_grad(ddm_mlp* m, const float* X, const float* Tg, const float* wt, int32_t M, float* loss,
                      float* grads) {
  if (!m || !X || !Tg || !wt) return DDM_EINVAL;
  int rc;
  if ((rc = upload_batch(m, X, Tg, wt, M))) return rc;
  if ((rc = loss_and_backward(m, M, loss))) return rc;
  if (grads) {
    MCHK(m, hipMemcpyAsync(grads, m->g, (size_t)DDM_MLP_PARAMS * 4, hipMemcpyDeviceToHost, m->stream));
    MCHK(m, hipStreamSynchronize(m->stream));
  }
  return DDM_OK;
}

int ddm_mlp_train_step(ddm_mlp* m, const float* X, const float* Tg, const float* wt, int32_t M, float lr,
                       float* loss) {
  if (!m || !X || !Tg || !wt) return DDM_EINVAL;
  int rc;
  if ((rc = upload_batch(m, X, Tg, wt, M))) return rc;
  if ((rc = loss_and_backward(m, M, loss))) return rc;
  m->step++;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - std::pow(b1, (float)m->step), c2 = 1.f - std::pow(b2, (float)m->step);
  MCHK(m, ddm::launch_adam(m->p, m->g, m->mo, m->vo, m->pb, DDM_MLP_PARAMS, lr, b1, b2, 1e-8f, c1, c2, m->stream));
  if (!loss) MCHK(m, hipStreamSynchronize(m->stream));
  return DDM_OK;
}

}  // extern "C"
