// Internal declarations shared by the kernel launchers and the engine.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

struct ddm_engine;
struct ddm_fit_batch;

namespace ddm {

// block spatial kernels: n = frame side, a power of two in [16, 2048] or 3 * 2^m in [48, 1536]
// maxlr (every FFT launcher): the radix cap, 3 (radix 8) or 2 (radix 4)
hipError_t launch_rows(int dtype, int n, int maxlr, const void* frames, int nframes, float scale, const float* win,
                       const float2* tw, float2* scratch, int2* rowinfo, hipStream_t st);
int cols_frames_per_wg(int n);
hipError_t launch_cols(int n, int maxlr, const float2* scratch, const int2* rowinfo, int nframes, int is_int, float scale,
                       const float* win, const float2* tw, const float2* wx, float2* hist, int T, int hs,
                       int first_slot, hipStream_t st);
hipError_t launch_temporal(int logt, const float2* hist, int T, int hs, const int4* items, int nitems, const int* bin_modes,
                           const uint32_t* order, const float* mask, int identity, float inv_n0, const float2* tw,
                           float* S_part, float* p_part, float* P_part, int KT, float* G_part, hipStream_t st);
hipError_t launch_temporal_w(int logt, int maxlr, const float2* hist, int T, int hs, const int4* items, int nitems, const int* bin_modes,
                             const uint32_t* order, const float* mask, int identity, float inv_n0, const float2* tw,
                             const float2* tw_f, float* S_part, float* p_part, float* P_part, int KT, float* G_part,
                             hipStream_t st);
// the workgroup-cooperative form for logt 11..13 (items of at most 64 modes,
// 16 at logt 13; S_part rows r-major with lm = logt - 9)
// dense: T = 2^logt / 2 and every frame valid (the mask tests compiled out)
hipError_t launch_temporal_c(int logt, int maxlr, const float2* hist, int T, int hs, const int4* items, int nitems,
                             const int* bin_modes, const uint32_t* order, const float* mask, int identity, int dense,
                             float inv_n0, const float2* tw, const float2* tw_f, float* S_part, float* p_part,
                             float* P_part, int KT, float* G_part, hipStream_t st);
// per-mode D, C, P of modes m0 .. m0 + nm - 1 (readTemporalFull)
// (mode_list: entries m0 .. m0 + nm - 1 of a mode list instead; outD/outP may be null)
hipError_t launch_temporal_full(int logt, const float2* hist, int T, int hs, int L, long m0, int nm, const uint32_t* order,
                                const float* mask, int identity, float inv_n0, const float* npairs, const float2* tw,
                                int KT, const int* mode_list, float* outD, float2* outC, float* outP, hipStream_t st);
// phase-drift sums of bins b0 .. b0 + nb - 1 over binned modes c0 .. c0 + nc - 1 (C rows of those modes)
hipError_t launch_drift(const float2* C, int L, int c0, int nc, const int* bin_offsets, const int2* kxy, int b0,
                        int nb, double* acc, hipStream_t st);
// KT: the temporal kernels sum D N(tau) of the last KT lags (T-KT..T-1)
// directly into G_part[item][KT]; lm: log2 of the r-major interleave of S_part rows
// (0 = frequency order); full != 0: every frame valid (prefix-sum form of the
// masked correlations); tw64: FP64 twiddles of the TPAD transform; Sbin,
// pbin: FP64 bin sums; Gbin: NB*L doubles, Re G(tau) - G(0) per bin (D N(tau)
// at the tail lags)
hipError_t launch_bin_finalize(int logt, int lm, const float* S_part, const float* p_part, const float* P_part,
                               const float* G_part, int KT, const int* bin_items, int NB, const uint32_t* nmodes,
                               const float* mask, int full, const float* npairs, int T, int L, float n0,
                               const double2* tw64, double* Sbin, double* Gbin, double* pbin, float* Pbin, float* radD,
                               float* radP, hipStream_t st);

hipError_t launch_rows_w(int dtype, int logn, int maxlr, const void* frames, int nframes, float scale, const float* win,
                         const float2* tw, float2* scratch, int2* rowinfo, hipStream_t st);
hipError_t launch_cols_w(int logn, int F, int maxlr, const float2* scratch, const int2* rowinfo, int nframes, int is_int,
                         float scale, const float* win, const float2* tw, const float2* wx, float2* hist, int T,
                         int hs, int first_slot, hipStream_t st);

// Device copy of the in-plane ISF table (ddm_isf2d.h): row i's values at
// v[off[i] .. off[i] + n[i] + 1], sampled at ln Lambda = u0[i] + k du[i].
// One video's arguments of the wavefront temporal stage (k_temporal_w); in a
// batch's work queue also item0, the index of its first work item there.
struct TemporalArgs {
  const float2* hist;
  const int4* items;
  const int* bin_modes;
  const uint32_t* order;
  const float* mask;
  float* S_part;
  float* p_part;
  float* P_part;
  float* G_part;
  int T, hs, identity, KT, item0;   // hs: history row stride (complex elements)
  float inv_n0;
};

struct Isf2dDev {
  const double *v, *Z, *u0, *du;
  const int *off, *n;
  double lz0, dlz;
};

struct Isf2dTable;
// the in-plane ISF table built on the device (layout from T, every value by
// the FP64 Legendre quadrature); synchronous
hipError_t build_isf2d_device(const Isf2dTable& T, double* v, double* z, double* u0, double* du, int* off, int* n,
                              hipStream_t st);
// out[i] = isf2d(x[i], Z[i]) by the quadrature (exact = 1) or the table (0); device arrays
hipError_t launch_isf2d_eval(const double* x, const double* Z, int n, int exact, const Isf2dDev* T, double* out,
                             hipStream_t st);

// model: 0 Brownian, 1 3D swimmer, 2 in-plane swimmer (t2d required); vmag:
// per-fit uniform drift speed or null; nstarts in [1, 3]; ws: device
// workspace of lm_fit_workspace_bytes(nfits) bytes (each fit's starts and
// their results)
size_t lm_fit_workspace_bytes(int nfits);
hipError_t launch_lm_fit(const float* tau, const float* d, const int* ntau, const float* w, const float* qn, int K,
                         const float* bounds, const float* seeds, const float* vmag, int nfits, int model,
                         const Isf2dDev* t2d, int seed_mode, int nstarts, int maxit, float* theta, double* chi,
                         int* steps, void* ws, hipStream_t st);

// global-parameter fits (one workgroup per video); model 0 Brownian, 1 3D
// swimmer, 2 in-plane swimmer (t2d required); mode 0 from seeds, 1 plain;
// split > 0: that many workgroups per video, one LM start each (outputs per
// workgroup: theta[4], chi, steps, and (A, B) of the video's bins at
// 2 (b0 split + s nbins))
hipError_t launch_lm_global(const float* tau, const float* d, const float* w, const float* qn, int K,
                            const int* bin_off, const int* ntau, const float* bounds, const float* seeds,
                            const float* vmag, int nvid, int model, const Isf2dDev* t2d, int mode, int nstarts,
                            int maxit, int split, double* theta, double* chi, int* steps, double* brown, double* ab,
                            hipStream_t st);

// allocation-free fits (ddm_fit.hip): per-device set-up of the fixed scratch
// array (current device), and pack -> LM -> unpack on device-resident arrays
hipError_t noalloc_setup();
hipError_t launch_fit_noalloc(int dev, const ::ddm_fit_batch* b, double* theta, double* chi2, int* iters,
                              hipStream_t st);

hipError_t launch_gemm_bf16(const unsigned short* A, long sai, long sak, const unsigned short* B, long sbk, long sbj,
                            float* C, long ldc, int M, int N, int K, int acc, hipStream_t st);
hipError_t launch_bias_act(float* Z, const float* bias, unsigned short* H, int M, int N, int gelu, hipStream_t st);
hipError_t launch_act_bwd(const float* dH, const float* Z, unsigned short* dZ, float* dZf, float* dbias, int M, int N,
                          int gelu, hipStream_t st);
hipError_t launch_mse(const float* Y, const float* Tg, const float* wt, float* dY, float* rowloss, int M, int N,
                      hipStream_t st);
hipError_t launch_adam(float* p, const float* g, float* m, float* v, unsigned short* pb, long n, float lr, float b1,
                       float b2, float eps, float c1, float c2, hipStream_t st);
hipError_t launch_to_bf16(const float* x, unsigned short* y, long n, hipStream_t st);

hipError_t launch_synth(int N, int T, int n, int n_sw, double sigma, double amp, double bg, double vbar, double Z,
                        double D, int heading3d, int swim_diffuse, uint64_t seed, int dtype, int noise,
                        double4* part, float2* pos, void* out, hipStream_t st);

// engine internals used by the batch runner (ddm_batch.cpp)
// computeTemporal in parts, for the batch's work queue: begin (planning and
// metadata uploads on the engine's stream), the temporal kernel (the caller's),
// end (radial finalisation); plan gives the kernel arguments with no GPU work
// (logt 0: the geometry takes the workgroup-per-mode kernel, TPAD 8192 without the cooperative one)
int engine_temporal_begin(::ddm_engine* e, const uint32_t* order, const float* mask);
int engine_temporal_end(::ddm_engine* e);
int engine_temporal_plan(::ddm_engine* e, const uint32_t* order, const float* mask, TemporalArgs* a, int* logt);
int engine_nitems(const ::ddm_engine* e);
const float2* engine_tw_t(const ::ddm_engine* e);
const float2* engine_tw_f(const ::ddm_engine* e);
// the persistent temporal kernel over the work items of nvid videos of one
// transform length 2^logt (vids: device table, item0 increasing; counter: one
// device int, zeroed here by a fill kernel); grid from temporal_wq_grid
hipError_t launch_temporal_wq(int logt, int maxlr, const TemporalArgs* vids, int nvid, int nwork, int* counter, const float2* tw,
                              const float2* tw_f, int grid, hipStream_t st);
int temporal_wq_grid(int logt, int maxlr);
// the same over the cooperative kernel (logt 11..13)
hipError_t launch_temporal_cq(int logt, int maxlr, const TemporalArgs* vids, int nvid, int nwork, int* counter,
                              const float2* tw, const float2* tw_f, int grid, hipStream_t st);
int temporal_cq_grid(int logt, int maxlr);
// one video's items through a persistent grid of `grid` workgroups (counter: one device int, zeroed here)
hipError_t launch_temporal_cp(int logt, int maxlr, const TemporalArgs& a, int nwork, int* counter, const float2* tw,
                              const float2* tw_f, int grid, hipStream_t st);
// 1 when the engine's temporal stage runs the cooperative kernel
int engine_tkern(const ::ddm_engine* e);
int engine_maxlr(const ::ddm_engine* e);
hipStream_t engine_stream(::ddm_engine* e);
void engine_set_capturing(::ddm_engine* e, bool on);
hipError_t engine_detach_meta(::ddm_engine* e, void** owned3);
void engine_get_meta(const ::ddm_engine* e, bool* identity, float* n0);
void engine_set_meta(::ddm_engine* e, bool identity, float n0);

}  // namespace ddm
