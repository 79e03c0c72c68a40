/*
 * ddm.h -- C ABI of the MI355X differential dynamic microscopy engine (libddm.so).
 *
 * This is the drop-in boundary for the motility-ddm hot path: a frame stack goes
 * in; the half-plane spectral history, the bin-averaged structure function
 * D(q, tau) and static power P(q) come out.  It replaces the WebGPU engine class
 * `DdmEngine` (motility-ddm src/ddm-engine.ts:62-426) and the host-side tables it
 * takes from src/ddm_ref.js.  The N-API addon (hackathon_amd/napi/ddm_napi.c)
 * binds each entry point 1:1 for the TypeScript facade; Python binds it with
 * ctypes (hackathon_amd/engine.py).
 *
 * Conventions
 *   - Every call returns an int status: DDM_OK (0), DDM_DROPPED (1, never
 *     returned by this engine: ingest is stream-ordered and does not drop), or
 *     a negative error whose text ddm_last_error() returns.  The reference
 *     throws `Error(message)` in the same situations (src/ddm-engine.ts:126-163,
 *     :313, :332; src/ddm_ref.js:540).
 *   - No C++ exception crosses the ABI.
 *   - Host buffers are owned by the caller; device memory is owned by the
 *     engine, except the device frame stack passed to ddm_ingest_device.
 *   - One engine = one HIP stream.  An engine is not thread-safe (as the
 *     reference engine, single JS thread).
 *   - Layouts are the reference's: mode m = ky*NH + kx with NH = N/2+1,
 *     history hist[(m*T + slot)*2 + {re,im}] (src/ddm_ref.js:175-205),
 *     radial D[b*L + tau] and P[b] (src/ddm_ref.js:336-348).
 */
#ifndef DDM_H
#define DDM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ddm_engine ddm_engine;

enum {
  DDM_OK = 0,
  DDM_DROPPED = 1,
  DDM_EINVAL = -1,   /* bad argument or geometry */
  DDM_EHIP = -2,     /* HIP runtime error (message has the HIP error string) */
  DDM_ESTATE = -3,   /* call not valid in the engine's state (e.g. re-entry) */
  DDM_ENOMEM = -4    /* device allocation failed */
};

/* Pixel types of a frame stack.  Integer frames are scaled to [0, 1]:
 * u8 by 1/255 (what the reference's luma path does, test/sim.mjs:188-193),
 * u16 by 1/65535; f32 frames are taken as luma already (ingestLuma). */
enum { DDM_U8 = 0, DDM_U16 = 1, DDM_F32 = 2 };

/* Geometry; the reference fixes it (DDM, src/ddm_ref.js:18-31: N=256, T=256,
 * L=128, NB=32, kmin=4, kmax=100, KNODES=4).  Here every field is a runtime
 * parameter.  N: power of two in [16, 2048] or 3 * 2^m in [48, 1536].  T: ring slots, 2..4096 (the
 * temporal FFT length is the next power of two >= 2T).  L: lags 1..T.
 * NB >= 1 bins linear in |k| on [kmin, kmax).  hypot_mode 0 bins by V8's
 * Math.hypot (bit-exact with the reference), 1 by the correctly rounded
 * hypot.  knodes: effective-|k| quadrature nodes per bin (reference: 4). */
typedef struct ddm_geometry {
  int32_t N, T, L, NB;
  double kmin, kmax;
  int32_t hypot_mode;
  int32_t knodes;
} ddm_geometry;

/* Derived sizes of an engine (bytes where named so). */
typedef struct ddm_sizes {
  int32_t N, NH, T, TPAD, L, NB, knodes;
  int64_t modes;          /* N * NH */
  int64_t binned_modes;   /* modes that fall in a bin */
  int64_t hist_bytes;     /* modes * T * 8 (ddm_read_history's output) */
  int32_t work_items;     /* temporal work items (bin, mode range) */
  int32_t chunk_frames;   /* frames per spatial chunk */
  int32_t hist_stride;    /* slots per mode row of the device history: T rounded up to 16 */
} ddm_sizes;

/* Launch-shape tuning of an engine (0 in any field = the default, which the
 * environment variables DDM_TEMPORAL_ITEMS, DDM_CHUNK_MB, DDM_WAVE_SPATIAL,
 * DDM_COLS_F may override).  Every choice computes the same D(q, tau) up to
 * FP32 summation order; the autotuner (hackathon_amd/autotune.py) checks that
 * before it times a candidate. */
typedef struct ddm_tuning {
  int32_t temporal_items;  /* target count of temporal work items (default 2048; <= 64 modes each) */
  int32_t chunk_mb;        /* spatial chunk budget in MiB of row spectra (default 128) */
  int32_t wave_spatial;    /* 1: workgroup-level spatial FFTs, 2: wavefront-level (N = 256 .. 2048 only) */
  int32_t cols_f;          /* frames per column workgroup of the wavefront kernel: 8 or 16 */
  int32_t fft_radix;       /* largest FFT radix of the spatial and temporal passes: 8 (default) or 4 */
  int32_t temporal_kernel; /* 1: one mode per wavefront (per workgroup above TPAD 4096), 2: one mode per
                              workgroup of four cooperating wavefronts (TPAD >= 2048) */
  int32_t temporal_wg_per_cu; /* cooperative kernel: 0 one workgroup per work item; k in 1..16 a persistent
                                 grid of k workgroups per CU taking the items from a counter */
} ddm_tuning;

/* DdmEngine.create (src/ddm-engine.ts:125-149): allocate device buffers for the
 * geometry, build the bin tables (makeBins, src/ddm_ref.js:84-170) and the
 * twiddle / window tables on the host in FP64 (src/ddm-engine.ts:193-197). */
int ddm_create(const ddm_geometry *g, ddm_engine **out);
/* Same with explicit tuning (t may be NULL); ddm_get_tuning reports the
 * values in effect. */
int ddm_create_tuned(const ddm_geometry *g, const ddm_tuning *t, ddm_engine **out);
int ddm_get_tuning(const ddm_engine *e, ddm_tuning *t);

/* DdmEngine.destroy (src/ddm-engine.ts:421-425).  NULL is ignored. */
void ddm_destroy(ddm_engine *e);

/* Message of the last failed call on `e` (or of the last failed ddm_create
 * when e is NULL).  Valid until the next call on the same engine. */
const char *ddm_last_error(const ddm_engine *e);

int ddm_get_sizes(const ddm_engine *e, ddm_sizes *s);

/* Bin tables of makeBins (src/ddm_ref.js:84-170): binOffsets[NB+1],
 * binModes[4*binned_modes] = (mode, kx, ky, 0), kCenter[NB], nModes[NB],
 * kNodes[NB*knodes].  Any pointer may be NULL. */
int ddm_get_bins(const ddm_engine *e, uint32_t *binOffsets, int32_t *binModes, double *kCenter, uint32_t *nModes,
                 double *kNodes);

/* Frame-stack ingest from HOST memory: nframes frames of N*N pixels (row-major,
 * y*N + x) of `dtype`, transformed (mean removal, separable Blackman-Harris
 * window, R2C 2D FFT, x 1/N) into ring slots (first_slot + i) mod T.
 * Batched form of ingestLuma / History.ingest (src/ddm-engine.ts:308-324,
 * src/ddm_ref.js:196-205); ingestLuma is nframes = 1, dtype = DDM_F32. */
int ddm_ingest(ddm_engine *e, const void *frames, int32_t dtype, int32_t nframes, int32_t first_slot);

/* Same, with the frame stack already resident in device memory (HBM).  The
 * call is asynchronous on the engine's stream; `dframes` must stay valid until
 * the next ddm_idle / read. */
int ddm_ingest_device(ddm_engine *e, const void *dframes, int32_t dtype, int32_t nframes, int32_t first_slot);

/* computeTemporal (src/ddm-engine.ts:331-365): masked Wiener-Khinchin
 * structure function of every binned mode and its radial reduction.
 * order[j] = ring slot holding time index j, mask[j] != 0 if valid (length T).
 * Asynchronous; read results with ddm_read_radial. */
int ddm_compute_temporal(ddm_engine *e, const uint32_t *order, const float *mask);

/* readRadial (src/ddm-engine.ts:391-395, radialFromGpu src/ddm_ref.js:584-590):
 * D[NB*L], P[NB]; `drift` (NB*L*8, may be NULL): the phase-drift least-squares
 * sums drift[(b*L + tau)*8 + c] (c = Sw, Sxx, Sxy, Syy, Rx, Ry, Rpp, count) of
 * accumulateMode (src/ddm_ref.js:360-381), computed on demand from per-mode
 * C(tau) only when the pointer is non-NULL (a per-mode pass over the binned
 * modes; zeros before the first ddm_compute_temporal). */
int ddm_read_radial(ddm_engine *e, float *D, float *P, float *drift);

/* readHistory (src/ddm-engine.ts:397-400): hist[(m*T + slot)*2 + {re,im}],
 * modes*T*2 floats. */
int ddm_read_history(ddm_engine *e, float *out);

/* readMean (src/ddm-engine.ts:410-414): the mean luma of the last frame
 * ingested (the reference's hi + lo of reduce_mean, src/ddm.wgsl:135-150), in
 * the units the spectra use (u8 / 255, u16 / 65535, f32 as given), from the
 * row sums the row transform keeps (exact integers for u8/u16); 0 before the
 * first frame. */
int ddm_read_mean(ddm_engine *e, double *mean);

/* readTemporalFull (src/ddm-engine.ts:403-408): per-mode results for EVERY
 * mode m (binned or not) of the current history under the last
 * ddm_compute_temporal's order and mask: D[m*L + tau], C[(m*L + tau)*2 +
 * {re,im}] = G(tau)/N(tau) (masked autocorrelation, no mean removal) and
 * P[m] = sum |x|^2 / N(0).  Opt-in: computed on demand in chunks of 8192
 * modes (modes*L*12 bytes of host output).  Any pointer may be NULL. */
int ddm_read_temporal_full(ddm_engine *e, float *D, float *C, float *P);

/* idle (src/ddm-engine.ts:417-419): wait for all work on the engine's stream. */
int ddm_idle(ddm_engine *e);

/* Device pointer of the spectral history, for callers that keep data on the
 * GPU: modes rows of hist_stride complex f32 (ddm_sizes), mode-major; slots
 * 0 .. T-1 of a row hold the mode's series, the rest stay zero. */
int ddm_history_device(ddm_engine *e, void **dptr);

/* Per-kernel timing with HIP events on the engine's stream.  enable != 0 starts
 * recording (and clears previous records).  ddm_kernel_times waits for the
 * stream and returns, for each kernel kind k (0 spatial rows, 1 spatial cols /
 * corner turn, 2 temporal, 3 radial finalize), the summed milliseconds and the
 * launch count. */
enum { DDM_K_ROWS = 0, DDM_K_COLS = 1, DDM_K_TEMPORAL = 2, DDM_K_FINALIZE = 3, DDM_K_COUNT = 4 };
int ddm_set_profiling(ddm_engine *e, int32_t enable);
int ddm_kernel_times(ddm_engine *e, double *ms, int32_t *count, int32_t n);

/* ------------------------------------------------------------------ per-q motility fits
 *
 * Batched Levenberg-Marquardt fits of the reference's per-bin model
 * (modelColumn, motility-ddm src/motility_fit.js:417-432; prepare :383-401
 * defines the data, lags, weights and q nodes; bounds :480-491):
 *   D_b(tau) = A c(tau) + B,  c = 1 - <exp(-D q^2 tau) (1 - alpha + alpha g(q vbar tau; Z))>_nodes
 * one fit per q-bin, all bins of all videos in one launch.  The reference fits
 * alpha, vbar, Z, D globally with Nelder-Mead over VarPro
 * (fitSwimmer/fitBrownian :500-611); this entry point is the per-q LM of the
 * MI355X design (no reference counterpart for the LM itself).
 * theta = (A, B, alpha, ln vbar, ln Z, ln D); chi^2 = w^2 sum (d - f)^2.
 * g: DDM_MODEL_SWIMMER_3D the Schulz swimmer ISF sin(Z atan L)/(Z L (1+L^2)^(Z/2))
 * (isf3d :140-144), DDM_MODEL_SWIMMER_2D the in-plane one from the reference's
 * tabulated kernel (isf2dFast :168-219, the same table as the verdict layer). */
typedef struct ddm_fitter ddm_fitter;

enum { DDM_MODEL_BROWNIAN = 0, DDM_MODEL_SWIMMER_3D = 1, DDM_MODEL_SWIMMER_2D = 2 };
enum { DDM_SEED_GRID = 0, DDM_SEED_GIVEN = 1, DDM_SEED_NONLINEAR = 2 };

typedef struct ddm_fit_batch {
  int32_t nfits;
  int32_t ntau_max;      /* row stride of tau/data, <= 128 */
  int32_t knodes;        /* q nodes per bin, <= 8 */
  int32_t model;         /* DDM_MODEL_* */
  int32_t seed_mode;     /* DDM_SEED_GRID: the reference's start grid, LM from the best nstarts;
                            DDM_SEED_GIVEN: LM from seeds[];
                            DDM_SEED_NONLINEAR: alpha, ln vbar, ln Z, ln D from seeds[],
                            A and B by the exact 2-column NNLS at those values */
  int32_t nstarts;       /* 1..3 */
  int32_t max_iter;      /* LM steps per start; 0 evaluates chi^2 at the (clamped) start */
  const int32_t *ntau;   /* [nfits] lags used by each fit */
  const double *tau;     /* [nfits * ntau_max] lag times (s) */
  const double *data;    /* [nfits * ntau_max] D_b(tau) */
  const double *weight;  /* [nfits] w = 1 / max_tau D_b */
  const double *qnodes;  /* [nfits * knodes] effective |q| nodes (rad/um) */
  const double *bounds;  /* [nfits * 4] vmin, vmax, Dmin, Dmax */
  const double *seeds;   /* [nfits * 6] theta seeds (DDM_SEED_GIVEN / _NONLINEAR), else NULL */
  const double *vmag;    /* [nfits] a fixed uniform drift speed |V| >= 0 (um/s) per fit, or NULL for none:
                            E = exp(-D q^2 tau) J0(q |V| tau) in the column, as modelColumn takes the
                            drift estimate (src/motility_fit.js:417-432, J0 :54-68) */
} ddm_fit_batch;

int ddm_fitter_create(ddm_fitter **out);
void ddm_fitter_destroy(ddm_fitter *f);
const char *ddm_fitter_last_error(const ddm_fitter *f);
/* Runs the batch; blocks until done.  theta (nfits*6), chi2 (nfits), iters
 * (nfits, LM steps summed over starts) are host outputs; *kernel_ms (may be
 * NULL) is the fit kernel's duration from HIP events. */
int ddm_fit(ddm_fitter *f, const ddm_fit_batch *b, double *theta, double *chi2, int32_t *iters, double *kernel_ms);

/* Global-parameter fits: the reference's model with alpha, vbar, Z and D SHARED by all bins of a video and
 * per-bin amplitudes A_b, B_b >= 0 (fitSwimmer / fitBrownian, src/motility_fit.js:500-536, :577-611, over
 * evalModel's variable projection :450-474), one workgroup per video, all videos in one launch.  The
 * objective is the reference's RSS = sum_b w_b^2 min_{A_b,B_b >= 0} |d_b - A_b c_b - B_b|^2.  The reference
 * minimises it by Nelder-Mead; here LM on (alpha, ln vbar, ln Z, ln D) with the amplitudes eliminated
 * exactly at every step.  seed_mode DDM_SEED_GRID: the reference's seeding (49-point ln D grid refined in
 * ln D = the Brownian fit, then the 16 vbar x 2 Z x 2 alpha start grid at that D, LM from the best
 * nstarts); DDM_SEED_GIVEN: LM from seeds[4 v ..] = (alpha, ln vbar, ln Z, ln D).  model
 * DDM_MODEL_BROWNIAN fits D alone. */
enum { DDM_GLOBAL_MAX_BINS = 1024 };
typedef struct ddm_global_fit_batch {
  int32_t nvideos;
  int32_t ntau_max;         /* row stride of tau/data, <= 128 */
  int32_t knodes;           /* q nodes per bin, <= 8 */
  int32_t model;            /* DDM_MODEL_* */
  int32_t seed_mode;        /* DDM_SEED_GRID or DDM_SEED_GIVEN */
  int32_t nstarts;          /* 1..3 */
  int32_t max_iter;         /* LM steps per start */
  const int32_t *bin_offset; /* [nvideos + 1]: video v owns bins bin_offset[v] .. bin_offset[v+1]-1
                               (1..DDM_GLOBAL_MAX_BINS of them) */
  const int32_t *ntau;      /* [nvideos] lags of the video's bins (the same lag grid for all of them) */
  const double *tau;        /* [nvideos * ntau_max] lag times (s) */
  const double *data;       /* [nbins * ntau_max] D_b(tau) */
  const double *weight;     /* [nbins] w_b = 1 / max_tau D_b */
  const double *qnodes;     /* [nbins * knodes] */
  const double *bounds;     /* [nvideos * 4] vmin, vmax, Dmin, Dmax (bounds, :480-491) */
  const double *seeds;      /* [nvideos * 4] (DDM_SEED_GIVEN), else NULL */
  const double *vmag;       /* [nvideos] uniform drift speed in the model, or NULL */
} ddm_global_fit_batch;
/* theta[nvideos * 4] = (alpha, ln vbar, ln Z, ln D); rss[nvideos]; iters[nvideos] (LM steps); brown
 * [nvideos * 2] = (D, RSS) of the Brownian fit (DDM_SEED_GRID only; may be NULL); amps[nbins * 2] = (A_b,
 * B_b) at theta (may be NULL); *kernel_ms (may be NULL) from HIP events.  Blocks until done. */
int ddm_fit_global(ddm_fitter *f, const ddm_global_fit_batch *b, double *theta, double *rss, int32_t *iters,
                   double *brown, double *amps, double *kernel_ms);

/* The in-plane swimmer ISF on the GPU, FP64: exact != 0 evaluates isf2d (src/motility_fit.js:152-159) with
 * legendreP's trapezoid quadrature of Laplace's integral (:120-132); exact == 0 evaluates isf2dFast
 * (:161-220), the table the in-plane fits interpolate, which the fitter builds on the device from the same
 * quadrature.  x, Z, out: host arrays of n values.  Blocks until done. */
int ddm_isf2d_eval(ddm_fitter *f, const double *x, const double *Z, int32_t n, int32_t exact, double *out);

/* ddm_fit on device-resident arrays, allocating nothing.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Every array of *b (ntau, tau, data, weight, qnodes, bounds,
 * seeds, vmag) and theta (nfits*6), chi2 (nfits), iters (nfits) are DEVICE
 * pointers; *b itself is read on the host.  The work is enqueued on the
 * fitter's stream; wait != 0 blocks until it is done (leave it 0 under graph
 * capture).  Results are bit-identical to ddm_fit's.  A fit whose ntau,
 * bounds or vmag fails ddm_fit's checks gets chi2 = NaN, iters = -1 and NaN
 * theta instead of an error status.  Buffers beyond the inputs and outputs
 * live in one fixed device array of about 23 MiB per GPU (set up by
 * ddm_fitter_create), so nfits <= DDM_FIT_NOALLOC_MAX; every call on a GPU
 * shares that array, so calls that overlap on one GPU (several fitters'
 * streams at once) race. */
enum { DDM_FIT_NOALLOC_MAX = 16384 };
int ddm_fit_device_noalloc(ddm_fitter *f, const ddm_fit_batch *b, double *theta, double *chi2, int32_t *iters,
                           int32_t wait);

/* ------------------------------------------------------------------ LM-seeding surrogate (MLP)
 *
 * 1536 -> 256 -> 256 -> 256 -> 68 with GELU after the hidden layers, BF16
 * weights and activations on MFMA with FP32 accumulation, FP32 master weights
 * and Adam.  No reference counterpart.  Parameters are one flat FP32 array of
 * DDM_MLP_PARAMS values: W0[1536][256], b0[256], W1[256][256], b1, W2[256][256],
 * b2, W3[256][68], b3[68] (row-major, y = x W + b). */
typedef struct ddm_mlp ddm_mlp;
enum { DDM_MLP_IN = 1536, DDM_MLP_HIDDEN = 256, DDM_MLP_OUT = 68, DDM_MLP_PARAMS = 542532 };

int ddm_mlp_create(int32_t max_batch, ddm_mlp **out);
void ddm_mlp_destroy(ddm_mlp *m);
const char *ddm_mlp_last_error(const ddm_mlp *m);
int ddm_mlp_set_params(ddm_mlp *m, const float *params);
int ddm_mlp_get_params(ddm_mlp *m, float *params);
/* inference: X[M][1536] -> Y[M][68] (host arrays) */
int ddm_mlp_forward(ddm_mlp *m, const float *X, int32_t M, float *Y);
/* weighted MSE loss 0.5/M sum_i sum_j wt[j] (y_ij - t_ij)^2 and its gradient
 * w.r.t. every parameter (grads[DDM_MLP_PARAMS], may be NULL), no update */
int ddm_mlp_loss_grad(ddm_mlp *m, const float *X, const float *Tg, const float *wt, int32_t M, float *loss,
                      float *grads);
/* one Adam step (t counts steps from 1) on a batch; *loss before the update */
int ddm_mlp_train_step(ddm_mlp *m, const float *X, const float *Tg, const float *wt, int32_t M, float lr,
                       float *loss);

/* ------------------------------------------------------------------ batches of videos
 *
 * Many videos (or ROIs) of uneven frame size and length through the spatial,
 * temporal and radial stages in one submission: one engine per video, every
 * video's kernels recorded once into a HIP graph whose branches run
 * concurrently (DDM_BATCH_GRAPH), or enqueued on the videos' streams on every
 * run (DDM_BATCH_EAGER).  Each video's results are bit-identical to running
 * its engine alone.  No reference counterpart (the reference analyses one
 * video at a time). */
typedef struct ddm_batch ddm_batch;
typedef struct ddm_batch_video {
  ddm_geometry g;
  ddm_tuning tuning;        /* zeros: defaults */
  int32_t dtype;            /* DDM_U8 / DDM_U16 / DDM_F32 */
  int32_t nframes;          /* frames ingested into slots 0..nframes-1 (<= T) */
  const void *dframes;      /* device frame stack; must stay valid while the batch exists */
  const uint32_t *order;    /* T entries (host), NULL: linear order */
  const float *mask;        /* T entries (host), NULL: every frame valid */
} ddm_batch_video;
/* mode: DDM_BATCH_EAGER or DDM_BATCH_GRAPH, optionally | DDM_BATCH_QUEUE: the
 * videos' temporal stages then run as one persistent work-queue kernel per
 * temporal transform length (every video's work items in one queue, taken by
 * a grid that fills the device), between the videos' spatial stages and
 * their radial finalisations; the results stay bit-identical. */
enum { DDM_BATCH_EAGER = 0, DDM_BATCH_GRAPH = 1, DDM_BATCH_QUEUE = 2 };

/* Creates the engines (and records the graph).  `chains` (>= 1) splits the
 * videos into that many independent chains that may run concurrently, balanced
 * by estimated work (each video, longest first, onto the least loaded chain);
 * videos within a chain run one after the other. */
int ddm_batch_create(const ddm_batch_video *v, int32_t n, int32_t mode, int32_t chains, ddm_batch **out);
/* One pass over every video; blocks until done; *gpu_ms (may be NULL) is the
 * pass's duration from HIP events. */
int ddm_batch_run(ddm_batch *b, double *gpu_ms);
/* The engine of video i (owned by the batch): read results with
 * ddm_read_radial etc. after ddm_batch_run. */
int ddm_batch_engine(ddm_batch *b, int32_t i, ddm_engine **e);
const char *ddm_batch_last_error(const ddm_batch *b);
void ddm_batch_destroy(ddm_batch *b);

/* ------------------------------------------------------------------ frame clock (streaming ring)
 *
 * The live producer's side of the ring (FrameClock, src/camera.ts:306-425).
 * Semantics:
 *   - frame index n = round((t - t0) / dt) (halves up), t0 = the epoch's first
 *     time stamp, dt = 1000 / fps ms; frame n lives in history slot n mod T;
 *   - a frame whose n is not above the newest one is a duplicate (counted,
 *     not placed); a gap of more than T frames starts a new epoch at that
 *     frame (n = 0); switching between the media timeline and the performance
 *     clock (capture / watchdog stamps) starts a new epoch too; missed counts
 *     the skipped indices between consecutive placed frames;
 *   - a snapshot lists the window n = newest - T + 1 .. newest oldest first:
 *     order[j] = slot of that frame, mask[j] = 1 iff n >= 0, j >= guard and the
 *     slot still holds frame n (not dropped, not overwritten, this epoch);
 *   - the median interval is over the last 120 intervals between placed
 *     frames of an epoch (the upper median for an even count; NaN if none;
 *     kept across epochs).
 * Host code only; one clock per producer (not thread-safe). */
typedef struct ddm_clock ddm_clock;
enum { DDM_TIME_CAPTURE = 0, DDM_TIME_WATCHDOG = 1, DDM_TIME_MEDIA = 2 };
typedef struct ddm_clock_state {
  int64_t newest_n;       /* newest frame index of the epoch, -1 before its first frame */
  int64_t frames, missed, duplicates, resets;
  int32_t valid;          /* frames of the window still in their slots */
  int32_t time_source;    /* DDM_TIME_* of the last placed frame, -1 in a fresh epoch */
  double median_dt_ms;
} ddm_clock_state;
/* new FrameClock(fps) with a ring of T slots (the reference's T = 256) */
int ddm_clock_create(double fps, int32_t T, ddm_clock **out);
void ddm_clock_destroy(ddm_clock *c);
/* place (camera.ts:345-376): DDM_OK with *slot and *n, or DDM_DROPPED for a duplicate */
int ddm_clock_place(ddm_clock *c, double t_ms, int32_t source, int32_t *slot, int64_t *n);
/* drop (camera.ts:379-382): frame n was not ingested; its slot no longer counts */
int ddm_clock_drop(ddm_clock *c, int64_t n);
/* reset (camera.ts:334-341): a new epoch */
int ddm_clock_reset(ddm_clock *c);
/* counters, newestN, timeSource, validCount, medianDtMs (camera.ts:326-398) */
int ddm_clock_state_get(const ddm_clock *c, ddm_clock_state *s);
/* snapshot (camera.ts:405-424) without the shuffle: order[T], mask[T], *used = valid entries */
int ddm_clock_snapshot(const ddm_clock *c, int32_t guard, uint32_t *order, float *mask, int32_t *used);
/* shuffleOrder (src/ddm_ref.js:510-519), the frame-order null test: a Fisher-Yates permutation of
 * the (order, mask) pairs in place; u holds the n-1 uniform draws in [0, 1) the caller's generator
 * gives, consumed for positions n-1 down to 1 (position i trades with floor(u (i + 1))). */
int ddm_shuffle_order(uint32_t *order, float *mask, int32_t n, const double *u);

/* ------------------------------------------------------------------ synthetic videos
 *
 * Seeded video generator on the GPU (the benchmark and training workloads):
 * n_spots Gaussian spots (width sigma px, peak amp) on a constant background,
 * the first round(alpha * n_spots) swim in straight lines (uniform heading,
 * speed ~ Gamma(Z+1, vbar/(Z+1)) px/frame; with heading3d the direction is
 * isotropic in 3D and the velocity its projection on the image plane, as in
 * the reference simulator's default), the rest take N(0, 2D) steps per
 * axis per frame (D in px^2/frame; with swim_diffuse the swimmers too, as in
 * the reference simulator), periodic wrap, Poisson photon noise
 * (noise = 0 stores the expected intensity instead), clipped to the pixel type.
 * The model follows the reference's movie simulator (motility-ddm
 * test/sim.mjs:90-186) in pixel units; random draws are Philox4x32-10 streams
 * keyed by `seed`, so a CPU restatement reproduces them. */
typedef struct ddm_synth ddm_synth;
typedef struct ddm_video_params {
  int32_t N;        /* frame side, multiple of 16, 16..4096 */
  int32_t T;        /* frames */
  int32_t n_spots;
  int32_t dtype;    /* DDM_U8 / DDM_U16 / DDM_F32 */
  int32_t noise;    /* 1: Poisson counts, 0: expected intensity */
  int32_t heading3d; /* 1: isotropic 3D headings projected on the image plane */
  int32_t swim_diffuse; /* 1: swimmers also take the diffusive steps */
  int32_t pad_;
  double sigma, amp, background;
  double alpha, vbar, Z, D;
  uint64_t seed;
} ddm_video_params;

int ddm_synth_create(ddm_synth **out);
void ddm_synth_destroy(ddm_synth *s);
const char *ddm_synth_last_error(const ddm_synth *s);
/* Renders T frames of N*N pixels into device memory `dframes` (frame-major,
 * row-major); returns when the frames are written. */
int ddm_synth_video(ddm_synth *s, const ddm_video_params *p, void *dframes);
/* Particle positions only: dpos[(t * n_spots + i) * 2 + {x, y}] (device, f32). */
int ddm_synth_tracks(ddm_synth *s, const ddm_video_params *p, float *dpos);

/* ---------------------------------------------------------------- verdict layer
 * analyzeMotility (src/motility_fit.js:808-990) on the host, FP64: global
 * VarPro fits of the Brownian and Schulz-swimmer models (grid + golden
 * section, multi-start Nelder-Mead), the drift estimate and a drift-aware
 * second pass, BIC on the effective sample size, TSI, motile share, tau_1/2
 * scaling, gates G1-G5 and the verdict.  Host code only: no GPU work. */
enum { DDM_GEOMETRY_3D = 0, DDM_GEOMETRY_2D = 1 };
enum {
  DDM_VERDICT_MOTILE = 0,
  DDM_VERDICT_NOT_MOTILE = 1,
  DDM_VERDICT_DIRECTED_TRANSPORT = 2,
  DDM_VERDICT_NO_DYNAMICS = 3,
  DDM_VERDICT_UNLOCKED_EXPOSURE = 4
};

/* The Radial record (src/ddm_ref.js:336-348) as readRadial returns it, FP64:
 * kNodes[NB * knodes] (knodes = 0: q-nodes are the bin centres), D[NB * L],
 * drift[NB * L * 8] (the phase sums, ddm_read_radial with drift). */
typedef struct ddm_radial_record {
  int32_t N, NB, L, knodes;
  const double *kCenter, *kNodes, *D, *drift;
  const uint32_t *nModes;
} ddm_radial_record;

typedef struct ddm_motility_opts {
  double dx_um, fps;        /* AnalysisOptions.dxUm, .fps */
  int32_t geometry;         /* DDM_GEOMETRY_3D (default of the reference) or _2D */
  int32_t exposure_locked;  /* AnalysisOptions.exposureLocked */
} ddm_motility_opts;

/* estimateDrift's result (src/motility_fit.js:671-725) */
typedef struct ddm_drift_estimate {
  double Vx, Vy, mag, sigma, angle_deg, samples;
  int32_t significant, pairs;
} ddm_drift_estimate;

/* Gate: id 0 (usable q-bins), 1 (TSI), 2 (uniform flow), 3 (delta BIC),
 * 4 (motile share), 5 (exposure locked), in the reference's order. */
typedef struct ddm_gate {
  int32_t id, pass;
  double value, sigma;
  int32_t resolved_bins, pad_;
} ddm_gate;

typedef struct ddm_motility_result {
  int32_t verdict, physics_verdict, insufficient, ngates;
  ddm_gate gates[6];
  double TSI, alpha, vbar, sigma_v, Z, Ddiff, Ddiff_brownian, dBIC, mu, nData, nEff, residual_rho;
  double swim_rss, swim_Vmag, brown_rss, brown_Vmag;
  int32_t alpha_resolved, resolved_bins, nbins, nscaling;
  ddm_drift_estimate drift;
} ddm_motility_result;

/* Optional per-bin detail (any pointer may be NULL; arrays hold cap entries):
 * usable bins' index and q, both models' amplitudes (A, B) per bin, and the
 * (q, tau_1/2) points of the scaling fit. */
typedef struct ddm_motility_detail {
  int32_t cap, pad_;
  int32_t *bin;
  double *q, *swim_amps, *brown_amps, *scaling_q, *scaling_tau_half;
} ddm_motility_detail;

int ddm_analyze_motility(const ddm_radial_record *r, const ddm_motility_opts *opt, ddm_motility_result *res,
                         ddm_motility_detail *detail);
/* n records at once on host threads (nthreads <= 0: one per hardware thread):
 * r, opt, res (and detail, when not NULL) are arrays of n.  Returns the first
 * failing record's status; the message names the record. */
int ddm_analyze_motility_batch(const ddm_radial_record *r, const ddm_motility_opts *opt, int32_t n,
                               ddm_motility_result *res, ddm_motility_detail *detail, int32_t nthreads);
const char *ddm_verdict_last_error(void);

/* Building blocks of the fits, for callers that prepare per-q problems themselves (host code):
 * lagGrid (src/motility_fit.js:358-369): the lags 1..L-1 the fits sample; writes at most cap of them
 * and returns their count (negative status on L < 2). */
int ddm_lag_grid(int32_t L, int32_t *lags, int32_t cap);
/* prepare's bin selection (src/motility_fit.js:383-401): the bins with nModes >= 3 and a finite positive
 * maximum of D over the lag grid, in order; writes at most cap and returns their count.  The drift sums
 * may be NULL. */
int ddm_usable_bins(const ddm_radial_record *r, int32_t *bins, int32_t cap);
/* The swimmer self-ISF g(x; Z) of the fits: isf3d (src/motility_fit.js:135-144) for DDM_GEOMETRY_3D,
 * the tabulated in-plane kernel isf2dFast (:168-219) for DDM_GEOMETRY_2D; n points. */
int ddm_swimmer_isf(int32_t geometry, const double *x, int32_t n, double Z, double *g);
/* estimateDrift (src/motility_fit.js:671-725) over the record's phase sums (drift required): swimmer =
 * NULL, or {vbar (um/s), Z} of a swimmer component whose ISF gates the lags (|g| < 0.2). */
int ddm_estimate_drift(const ddm_radial_record *r, const ddm_motility_opts *opt, const double *swimmer,
                       ddm_drift_estimate *out);

#ifdef __cplusplus
}
#endif
#endif
