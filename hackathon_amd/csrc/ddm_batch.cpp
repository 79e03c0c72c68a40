// Batches of uneven videos (ddm_batch_* in include/ddm.h): one engine per
// video; the videos' ingest + temporal + radial work is recorded into one HIP
// graph with `chains` concurrent branches (fork/join through events on the
// batch stream), or enqueued eagerly with the same dependencies.  In
// work-queue mode the temporal stages of all videos of one transform length
// run as one persistent kernel fed from a queue of their work items.
#include "../../include/ddm.h"
#include "ddm_internal.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

struct ddm_batch {
  std::vector<ddm_engine*> eng;
  std::vector<ddm_batch_video> vids;
  std::vector<std::vector<uint32_t>> order;
  std::vector<std::vector<float>> mask;
  std::vector<hipEvent_t> evs;
  // graph mode: the pinned metadata staging the graph reads (detached from
  // the engines) and each video's order/mask flags at capture
  std::vector<void*> meta_owned;
  std::vector<char> meta_identity;
  std::vector<float> meta_n0;
  int mode = DDM_BATCH_EAGER, chains = 1;
  // videos in enqueue order (longest first) and the chain each one runs on
  std::vector<int> seq, chain_of;
  // work-queue mode: one persistent temporal launch per transform length
  struct Queue {
    int logt = 0, maxlr = 3, nwork = 0, grid = 0;
    int coop = 0;                         // 1: the cooperative kernel's queue form
    std::vector<int> vids;                // videos in the queue, in item order
    ddm::TemporalArgs* dargs = nullptr;   // device table, one entry per video
    const float2 *tw = nullptr, *tw_f = nullptr;
  };
  std::vector<Queue> queues;
  std::vector<char> queued;               // video i's temporal stage runs from a queue
  int* counters = nullptr;                // one device int per queue
  hipStream_t s0 = nullptr;
  hipEvent_t t0 = nullptr, t1 = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::string err;
};

namespace {
int bfail(ddm_batch* b, int code, const std::string& msg) {
  b->err = msg;
  return code;
}

hipEvent_t new_event(ddm_batch* b) {
  hipEvent_t ev = nullptr;
  if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) return nullptr;
  b->evs.push_back(ev);
  return ev;
}

// Enqueue every video on its engine's stream, chained per `chains`, forked from
// and joined back into b->s0 (identical when recorded into a graph).  In
// work-queue mode the queued videos stop after their metadata uploads; the
// persistent temporal kernels follow on b->s0 once every chain has joined,
// then the queued videos' radial finalisations, chained again.
int enqueue_all(ddm_batch* b, bool capturing) {
  const int n = (int)b->eng.size();
  hipEvent_t fork = new_event(b);
  if (!fork || hipEventRecord(fork, b->s0) != hipSuccess) return bfail(b, DDM_EHIP, "batch: fork event");
  auto run_chains = [&](hipEvent_t start, bool second) -> int {
    std::vector<hipEvent_t> tail(b->chains, start);
    for (int k = 0; k < n; k++) {
      const int i = b->seq[k];
      if (second && !b->queued[i]) continue;
      ddm_engine* e = b->eng[i];
      hipStream_t es = ddm::engine_stream(e);
      const int c = b->chain_of[i];
      if (hipStreamWaitEvent(es, tail[c], 0) != hipSuccess) return bfail(b, DDM_EHIP, "batch: chain wait");
      ddm::engine_set_capturing(e, capturing);
      const ddm_batch_video& v = b->vids[i];
      int rc = DDM_OK;
      if (second) {
        rc = ddm::engine_temporal_end(e);
      } else {
        rc = ddm_ingest_device(e, v.dframes, v.dtype, v.nframes, 0);
        if (!rc) rc = b->queued[i] ? ddm::engine_temporal_begin(e, b->order[i].data(), b->mask[i].data())
                                   : ddm_compute_temporal(e, b->order[i].data(), b->mask[i].data());
      }
      ddm::engine_set_capturing(e, false);
      if (rc) return bfail(b, rc, "batch: video " + std::to_string(i) + ": " + ddm_last_error(e));
      hipEvent_t done = new_event(b);
      if (!done || hipEventRecord(done, es) != hipSuccess) return bfail(b, DDM_EHIP, "batch: chain event");
      tail[c] = done;
    }
    for (int c = 0; c < b->chains; c++)
      if (tail[c] != start && hipStreamWaitEvent(b->s0, tail[c], 0) != hipSuccess)
        return bfail(b, DDM_EHIP, "batch: join");
    return DDM_OK;
  };
  int rc = run_chains(fork, false);
  if (rc || b->queues.empty()) return rc;
  for (size_t q = 0; q < b->queues.size(); q++) {
    const ddm_batch::Queue& Q = b->queues[q];
    hipError_t he =
        Q.coop ? ddm::launch_temporal_cq(Q.logt, Q.maxlr, Q.dargs, (int)Q.vids.size(), Q.nwork, b->counters + q, Q.tw,
                                         Q.tw_f, Q.grid, b->s0)
               : ddm::launch_temporal_wq(Q.logt, Q.maxlr, Q.dargs, (int)Q.vids.size(), Q.nwork, b->counters + q, Q.tw,
                                         Q.tw_f, Q.grid, b->s0);
    if (he != hipSuccess) return bfail(b, DDM_EHIP, std::string("batch: work-queue launch: ") + hipGetErrorString(he));
  }
  hipEvent_t mid = new_event(b);
  if (!mid || hipEventRecord(mid, b->s0) != hipSuccess) return bfail(b, DDM_EHIP, "batch: queue event");
  return run_chains(mid, true);
}

// Work-queue set-up: each video's temporal arguments planned on the host (no
// GPU work), grouped by transform length into device tables.
int build_queues(ddm_batch* b) {
  const int n = (int)b->eng.size();
  b->queued.assign(n, 0);
  // one queue per (kernel, transform length, radix cap): slot 32 coop + 2 logt + (cap == 2)
  std::vector<std::vector<ddm::TemporalArgs>> args(64);
  std::vector<ddm_batch::Queue> qs(64);
  for (int i = 0; i < n; i++) {
    ddm::TemporalArgs a{};
    int logt = 0;
    int rc = ddm::engine_temporal_plan(b->eng[i], b->order[i].data(), b->mask[i].data(), &a, &logt);
    if (rc) return bfail(b, rc, "batch: video " + std::to_string(i) + ": " + ddm_last_error(b->eng[i]));
    const int ni = ddm::engine_nitems(b->eng[i]);
    if (logt < 8 || logt > 13 || ni < 1) continue;   // the video's own computeTemporal
    const int maxlr = ddm::engine_maxlr(b->eng[i]);
    const int coop = ddm::engine_tkern(b->eng[i]);
    const int slot = 32 * coop + 2 * logt + (maxlr == 2 ? 1 : 0);
    ddm_batch::Queue& Q = qs[slot];
    a.item0 = Q.nwork;
    Q.nwork += ni;
    Q.logt = logt;
    Q.maxlr = maxlr;
    Q.coop = coop;
    Q.vids.push_back(i);
    if (!Q.tw) {
      Q.tw = ddm::engine_tw_t(b->eng[i]);
      Q.tw_f = ddm::engine_tw_f(b->eng[i]);
    }
    args[slot].push_back(a);
    b->queued[i] = 1;
  }
  for (int l = 0; l < 64; l++) {
    if (qs[l].vids.empty()) continue;
    ddm_batch::Queue& Q = qs[l];
    const int g = Q.coop ? ddm::temporal_cq_grid(Q.logt, Q.maxlr) : ddm::temporal_wq_grid(Q.logt, Q.maxlr);
    if (g < 1) return bfail(b, DDM_EHIP, "batch: work-queue occupancy query");
    Q.grid = std::min(g, Q.nwork);
    hipError_t he = hipMalloc((void**)&Q.dargs, args[l].size() * sizeof(ddm::TemporalArgs));
    if (he == hipSuccess)
      he = hipMemcpy(Q.dargs, args[l].data(), args[l].size() * sizeof(ddm::TemporalArgs), hipMemcpyHostToDevice);
    b->queues.push_back(Q);
    if (he != hipSuccess) return bfail(b, DDM_EHIP, std::string("batch: work-queue table: ") + hipGetErrorString(he));
  }
  if (!b->queues.empty() && hipMalloc((void**)&b->counters, b->queues.size() * sizeof(int)) != hipSuccess)
    return bfail(b, DDM_EHIP, "batch: work-queue counters");
  return DDM_OK;
}

// Chains balanced by estimated work, longest video first onto the least
// loaded chain (LPT).  Work model: the spatial stage ~ 44 N^2 per frame, the
// temporal stage ~ N^2 TPAD log2 TPAD per analysis (their ratio measured at
// 1024 frames of 512^2: 0.9 ms spatial, 0.45 ms temporal at TPAD 2048).
void balance_chains(ddm_batch* b) {
  const int n = (int)b->vids.size();
  std::vector<double> w(n);
  for (int i = 0; i < n; i++) {
    const ddm_batch_video& v = b->vids[i];
    const double px = (double)v.g.N * v.g.N;
    int lt = 1;
    while ((1L << lt) < 2L * v.g.T) lt++;
    w[i] = px * (44.0 * v.nframes + (double)(1L << lt) * lt);
  }
  b->seq.resize(n);
  for (int i = 0; i < n; i++) b->seq[i] = i;
  std::stable_sort(b->seq.begin(), b->seq.end(), [&](int x, int y) { return w[x] > w[y]; });
  std::vector<double> load(b->chains, 0.0);
  b->chain_of.assign(n, 0);
  for (int i : b->seq) {
    const int c = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    b->chain_of[i] = c;
    load[c] += w[i];
  }
}

void release(ddm_batch* b) {
  if (b->s0) hipStreamSynchronize(b->s0);
  for (auto* e : b->eng) ddm_destroy(e);
  if (b->exec) hipGraphExecDestroy(b->exec);
  if (b->graph) hipGraphDestroy(b->graph);
  for (auto ev : b->evs) hipEventDestroy(ev);
  if (b->t0) hipEventDestroy(b->t0);
  if (b->t1) hipEventDestroy(b->t1);
  if (b->s0) hipStreamDestroy(b->s0);
  for (void* p : b->meta_owned)
    if (p) hipHostFree(p);
  for (auto& q : b->queues)
    if (q.dargs) hipFree(q.dargs);
  if (b->counters) hipFree(b->counters);
}
}  // namespace

extern "C" {

static std::string g_batch_create_error;

int ddm_batch_create(const ddm_batch_video* v, int32_t n, int32_t mode, int32_t chains, ddm_batch** out) {
  if (!out) return DDM_EINVAL;
  *out = nullptr;
  if (!v || n < 1 || mode < 0 || mode > (DDM_BATCH_GRAPH | DDM_BATCH_QUEUE) || chains < 1) {
    g_batch_create_error = "batch: bad arguments";
    return DDM_EINVAL;
  }
  ddm_batch* b = new ddm_batch();
  b->mode = mode & DDM_BATCH_GRAPH;
  b->chains = std::min<int>(chains, n);
  auto bail = [&](int rc) {
    g_batch_create_error = b->err;
    release(b);
    delete b;
    return rc;
  };
  if (hipStreamCreateWithFlags(&b->s0, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&b->t0) != hipSuccess || hipEventCreate(&b->t1) != hipSuccess)
    return (bfail(b, DDM_EHIP, "batch: stream/event creation"), bail(DDM_EHIP));
  for (int i = 0; i < n; i++) {
    const ddm_batch_video& vi = v[i];
    if (!vi.dframes || vi.nframes < 1 || vi.nframes > vi.g.T)
      return (bfail(b, DDM_EINVAL, "batch: video " + std::to_string(i) + ": need 1 <= nframes <= T and frames"),
              bail(DDM_EINVAL));
    ddm_engine* e = nullptr;
    int rc = ddm_create_tuned(&vi.g, &vi.tuning, &e);
    if (rc) return (bfail(b, rc, "batch: video " + std::to_string(i) + ": " + ddm_last_error(nullptr)), bail(rc));
    b->eng.push_back(e);
    b->vids.push_back(vi);
    std::vector<uint32_t> o(vi.g.T);
    std::vector<float> m(vi.g.T, 1.f);
    for (int j = 0; j < vi.g.T; j++) {
      o[j] = vi.order ? vi.order[j] : (uint32_t)j;
      if (vi.mask) m[j] = vi.mask[j];
      else if (j >= vi.nframes) m[j] = 0.f;   // slots never written are invalid
    }
    b->order.push_back(std::move(o));
    b->mask.push_back(std::move(m));
  }
  b->queued.assign(n, 0);
  balance_chains(b);
  if (mode & DDM_BATCH_QUEUE) {
    const int qrc = build_queues(b);
    if (qrc) return bail(qrc);
  }
  if (b->mode == DDM_BATCH_GRAPH) {
    if (hipStreamBeginCapture(b->s0, hipStreamCaptureModeRelaxed) != hipSuccess)
      return (bfail(b, DDM_EHIP, "batch: begin capture"), bail(DDM_EHIP));
    int rc = enqueue_all(b, true);
    hipGraph_t g = nullptr;
    hipError_t he = hipStreamEndCapture(b->s0, &g);
    if (rc) {
      if (g) hipGraphDestroy(g);
      return bail(rc);
    }
    if (he != hipSuccess) return (bfail(b, DDM_EHIP, std::string("batch: end capture: ") + hipGetErrorString(he)),
                                  bail(DDM_EHIP));
    b->graph = g;
    if ((he = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0)) != hipSuccess)
      return (bfail(b, DDM_EHIP, std::string("batch: instantiate: ") + hipGetErrorString(he)), bail(DDM_EHIP));
    for (ddm_engine* e : b->eng) {
      bool id = true;
      float n0 = 0.f;
      ddm::engine_get_meta(e, &id, &n0);
      b->meta_identity.push_back(id ? 1 : 0);
      b->meta_n0.push_back(n0);
      void* owned[3] = {nullptr, nullptr, nullptr};
      he = ddm::engine_detach_meta(e, owned);
      b->meta_owned.insert(b->meta_owned.end(), owned, owned + 3);
      if (he != hipSuccess)
        return (bfail(b, DDM_EHIP, std::string("batch: pinned metadata: ") + hipGetErrorString(he)),
                bail(DDM_EHIP));
    }
  }
  *out = b;
  return DDM_OK;
}

int ddm_batch_run(ddm_batch* b, double* gpu_ms) {
  if (!b) return DDM_EINVAL;
  if (hipEventRecord(b->t0, b->s0) != hipSuccess) return bfail(b, DDM_EHIP, "batch: timing event");
  int rc = DDM_OK;
  if (b->mode == DDM_BATCH_GRAPH) {
    hipError_t he = hipGraphLaunch(b->exec, b->s0);
    if (he != hipSuccess) return bfail(b, DDM_EHIP, std::string("batch: graph launch: ") + hipGetErrorString(he));
  } else {
    // events of earlier eager runs are no longer waited on: recycle them
    hipStreamSynchronize(b->s0);
    for (auto ev : b->evs) hipEventDestroy(ev);
    b->evs.clear();
    if ((rc = enqueue_all(b, false))) return rc;
  }
  if (hipEventRecord(b->t1, b->s0) != hipSuccess) return bfail(b, DDM_EHIP, "batch: timing event");
  hipError_t he = hipEventSynchronize(b->t1);
  if (he != hipSuccess) return bfail(b, DDM_EHIP, std::string("batch: run: ") + hipGetErrorString(he));
  if (b->mode == DDM_BATCH_GRAPH)   // readers of the engines see the graph's order and mask again
    for (size_t i = 0; i < b->eng.size(); i++) ddm::engine_set_meta(b->eng[i], b->meta_identity[i] != 0, b->meta_n0[i]);
  if (gpu_ms) {
    float ms = 0;
    hipEventElapsedTime(&ms, b->t0, b->t1);
    *gpu_ms = ms;
  }
  return DDM_OK;
}

int ddm_batch_engine(ddm_batch* b, int32_t i, ddm_engine** e) {
  if (!b || !e || i < 0 || i >= (int)b->eng.size()) return DDM_EINVAL;
  *e = b->eng[i];
  return DDM_OK;
}

const char* ddm_batch_last_error(const ddm_batch* b) { return b ? b->err.c_str() : g_batch_create_error.c_str(); }

void ddm_batch_destroy(ddm_batch* b) {
  if (!b) return;
  release(b);
  delete b;
}

}  // extern "C"
