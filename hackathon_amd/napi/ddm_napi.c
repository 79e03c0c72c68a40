/*
 * ddm_napi.c -- thin N-API addon over libddm.so for the TypeScript facade.
 *
 * One JS function per C-ABI entry point (include/ddm.h); no logic of its own.
 * Errors: a negative status throws (or rejects with) Error(ddm_last_error()),
 * the way the reference engine throws (motility-ddm src/ddm-engine.ts:126-163,
 * :313, :332).  The engine handle is a napi_external whose finalizer calls
 * ddm_destroy.
 *
 * The engine calls that wait on the GPU or on large copies (frame-stack
 * ingest, computeTemporal through to D, readRadial, readHistory,
 * readTemporalFull, idle) run as napi_async_work on libuv's thread pool and
 * return Promises, so the event loop keeps running while they work, as the
 * reference's async engine methods let it (src/ddm-engine.ts:331-352).  An
 * engine is not thread-safe: every call on it holds the handle's mutex; the
 * GPU wait at the end of computeTemporal holds a waiter count instead, so a
 * single-frame ingestLuma (synchronous, as in the reference) can be issued
 * while an analysis finishes on the GPU; destroy waits for both.
 *
 * Built with gcc against node_api.h (N-API >= 6), linked to libddm.so with an
 * $ORIGIN rpath.
 */
#define NAPI_VERSION 6
#include <node_api.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/ddm.h"

#define NAPI_CALL(env, call)                                          \
  do {                                                                \
    if ((call) != napi_ok) {                                          \
      napi_throw_error((env), NULL, "N-API call failed: " #call);     \
      return NULL;                                                    \
    }                                                                 \
  } while (0)

typedef struct {
  ddm_engine *e;
  pthread_mutex_t mu;     /* held by every call on e */
  pthread_cond_t cv;      /* signalled when waiters drops to 0 */
  int waiters;            /* GPU waits in flight outside the mutex */
} handle_t;

/* destroy the engine once no call is using it (main thread) */
static void handle_close(handle_t *h) {
  pthread_mutex_lock(&h->mu);
  while (h->waiters > 0) pthread_cond_wait(&h->cv, &h->mu);
  if (h->e) ddm_destroy(h->e);
  h->e = NULL;
  pthread_mutex_unlock(&h->mu);
}

static void finalize_handle(napi_env env, void *data, void *hint) {
  (void)env;
  (void)hint;
  handle_t *h = (handle_t *)data;
  handle_close(h);
  pthread_cond_destroy(&h->cv);
  pthread_mutex_destroy(&h->mu);
  free(h);
}

static napi_value throw_msg(napi_env env, const char *msg) {
  napi_throw_error(env, NULL, msg);
  return NULL;
}

static int get_handle(napi_env env, napi_value v, handle_t **out) {
  void *p = NULL;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
    napi_throw_type_error(env, NULL, "expected a DDM engine handle");
    return 0;
  }
  *out = (handle_t *)p;
  if (!(*out)->e) {
    napi_throw_error(env, NULL, "engine destroyed");
    return 0;
  }
  return 1;
}

static double get_prop_num(napi_env env, napi_value obj, const char *name, double dflt) {
  bool has = false;
  napi_value v;
  double d = dflt;
  if (napi_has_named_property(env, obj, name, &has) == napi_ok && has &&
      napi_get_named_property(env, obj, name, &v) == napi_ok) {
    napi_valuetype t;
    napi_typeof(env, v, &t);
    if (t == napi_number) napi_get_value_double(env, v, &d);
  }
  return d;
}

static napi_value make_typed(napi_env env, napi_typedarray_type t, size_t count, size_t esize, void **data) {
  napi_value ab, ta;
  if (napi_create_arraybuffer(env, count * esize, data, &ab) != napi_ok) return NULL;
  if (napi_create_typedarray(env, t, count, ab, 0, &ta) != napi_ok) return NULL;
  return ta;
}

static void set_prop(napi_env env, napi_value obj, const char *name, napi_value v) {
  if (v) napi_set_named_property(env, obj, name, v);
}

static napi_value num(napi_env env, double d) {
  napi_value v;
  napi_create_double(env, d, &v);
  return v;
}

/* create({N, T, L, NB, kmin, kmax, hypotMode, knodes}) -> handle */
static napi_value js_create(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  if (argc < 1) {
    napi_throw_type_error(env, NULL, "create expects a geometry object");
    return NULL;
  }
  ddm_geometry g;
  g.N = (int32_t)get_prop_num(env, argv[0], "N", 256);
  g.T = (int32_t)get_prop_num(env, argv[0], "T", 256);
  g.L = (int32_t)get_prop_num(env, argv[0], "L", 128);
  g.NB = (int32_t)get_prop_num(env, argv[0], "NB", 32);
  g.kmin = get_prop_num(env, argv[0], "kmin", 4);
  g.kmax = get_prop_num(env, argv[0], "kmax", 100);
  g.hypot_mode = (int32_t)get_prop_num(env, argv[0], "hypotMode", 0);
  g.knodes = (int32_t)get_prop_num(env, argv[0], "knodes", 4);
  ddm_engine *e = NULL;
  int rc = ddm_create(&g, &e);
  if (rc != DDM_OK) return throw_msg(env, ddm_last_error(NULL));
  handle_t *h = (handle_t *)calloc(1, sizeof(handle_t));
  if (!h) {
    ddm_destroy(e);
    return throw_msg(env, "create: out of memory");
  }
  h->e = e;
  pthread_mutex_init(&h->mu, NULL);
  pthread_cond_init(&h->cv, NULL);
  napi_value out;
  NAPI_CALL(env, napi_create_external(env, h, finalize_handle, NULL, &out));
  return out;
}

static napi_value js_destroy(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  void *p = NULL;
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  if (argc >= 1 && napi_get_value_external(env, argv[0], &p) == napi_ok && p) handle_close((handle_t *)p);
  return NULL;
}

/* sizes(h) -> {N, NH, T, TPAD, L, NB, knodes, modes, binnedModes, histBytes, workItems, chunkFrames,
 * histStride} */
static napi_value js_sizes(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1], out;
  handle_t *h;
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  if (argc < 1 || !get_handle(env, argv[0], &h)) return NULL;
  ddm_sizes s;
  int rc = ddm_get_sizes(h->e, &s);   /* fixed at create: no lock needed */
  if (rc != DDM_OK) return throw_msg(env, "sizes: invalid engine");
  NAPI_CALL(env, napi_create_object(env, &out));
  set_prop(env, out, "N", num(env, s.N));
  set_prop(env, out, "NH", num(env, s.NH));
  set_prop(env, out, "T", num(env, s.T));
  set_prop(env, out, "TPAD", num(env, s.TPAD));
  set_prop(env, out, "L", num(env, s.L));
  set_prop(env, out, "NB", num(env, s.NB));
  set_prop(env, out, "knodes", num(env, s.knodes));
  set_prop(env, out, "modes", num(env, (double)s.modes));
  set_prop(env, out, "binnedModes", num(env, (double)s.binned_modes));
  set_prop(env, out, "histBytes", num(env, (double)s.hist_bytes));
  set_prop(env, out, "workItems", num(env, s.work_items));
  set_prop(env, out, "chunkFrames", num(env, s.chunk_frames));
  set_prop(env, out, "histStride", num(env, s.hist_stride));
  return out;
}

/* bins(h) -> {binOffsets, binModes, kCenter, nModes, kNodes} (makeBins layout) */
static napi_value js_bins(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1], out;
  handle_t *h;
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  if (argc < 1 || !get_handle(env, argv[0], &h)) return NULL;
  ddm_sizes s;
  ddm_get_sizes(h->e, &s);
  void *off, *modes, *kc, *nm, *kn;
  napi_value voff = make_typed(env, napi_uint32_array, (size_t)s.NB + 1, 4, &off);
  napi_value vmodes = make_typed(env, napi_int32_array, (size_t)(4 * s.binned_modes), 4, &modes);
  napi_value vkc = make_typed(env, napi_float64_array, (size_t)s.NB, 8, &kc);
  napi_value vnm = make_typed(env, napi_uint32_array, (size_t)s.NB, 4, &nm);
  napi_value vkn = make_typed(env, napi_float64_array, (size_t)s.NB * s.knodes, 8, &kn);
  if (!voff || !vmodes || !vkc || !vnm || !vkn) return throw_msg(env, "bins: allocation failed");
  int rc = ddm_get_bins(h->e, (uint32_t *)off, (int32_t *)modes, (double *)kc, (uint32_t *)nm, (double *)kn);
  if (rc != DDM_OK) return throw_msg(env, "bins: invalid engine");
  NAPI_CALL(env, napi_create_object(env, &out));
  set_prop(env, out, "binOffsets", voff);
  set_prop(env, out, "binModes", vmodes);
  set_prop(env, out, "kCenter", vkc);
  set_prop(env, out, "nModes", vnm);
  set_prop(env, out, "kNodes", vkn);
  return out;
}

/* ------------------------------------------------------------------ async engine calls */

typedef enum { OP_INGEST, OP_TEMPORAL, OP_RADIAL, OP_HISTORY, OP_FULL, OP_IDLE } op_kind;

typedef struct {
  napi_async_work work;
  napi_deferred deferred;
  napi_ref href;            /* the engine handle, alive until the op completes */
  napi_ref keep[4];         /* typed arrays the op reads or fills */
  const char *names[4];     /* result properties of the filled arrays */
  int nkeep;
  handle_t *h;
  op_kind kind;
  const void *in;
  int32_t dtype, nframes, slot;
  const uint32_t *order;
  const float *mask;
  void *out[3];
  int rc;
  char err[512];
} op_t;

static void op_execute(napi_env env, void *data) {
  (void)env;
  op_t *op = (op_t *)data;
  handle_t *h = op->h;
  pthread_mutex_lock(&h->mu);
  ddm_engine *e = h->e;
  int wait = 0;
  if (!e) {
    op->rc = DDM_ESTATE;
    snprintf(op->err, sizeof op->err, "engine destroyed");
  } else {
    switch (op->kind) {
      case OP_INGEST: op->rc = ddm_ingest(e, op->in, op->dtype, op->nframes, op->slot); break;
      case OP_TEMPORAL:
        op->rc = ddm_compute_temporal(e, op->order, op->mask);
        wait = op->rc == DDM_OK;
        break;
      case OP_RADIAL: op->rc = ddm_read_radial(e, (float *)op->out[0], (float *)op->out[1], (float *)op->out[2]); break;
      case OP_HISTORY: op->rc = ddm_read_history(e, (float *)op->out[0]); break;
      case OP_FULL:
        op->rc = ddm_read_temporal_full(e, (float *)op->out[0], (float *)op->out[1], (float *)op->out[2]);
        break;
      case OP_IDLE: op->rc = ddm_idle(e); break;
    }
    if (op->rc < 0) snprintf(op->err, sizeof op->err, "%s", ddm_last_error(e));
    if (wait) h->waiters++;
  }
  pthread_mutex_unlock(&h->mu);
  if (wait) {
    /* the analysis is enqueued on the engine's stream: wait for D outside the
       mutex (the stream handle is fixed for the engine's life, and destroy
       waits for this waiter) */
    const int rc = ddm_idle(e);
    pthread_mutex_lock(&h->mu);
    if (rc < 0) {
      op->rc = rc;
      snprintf(op->err, sizeof op->err, "%s", ddm_last_error(e));
    }
    if (--h->waiters == 0) pthread_cond_broadcast(&h->cv);
    pthread_mutex_unlock(&h->mu);
  }
}

static void op_complete(napi_env env, napi_status status, void *data) {
  op_t *op = (op_t *)data;
  if (status != napi_ok && op->rc >= 0) {
    op->rc = DDM_ESTATE;
    snprintf(op->err, sizeof op->err, "engine call cancelled");
  }
  if (op->rc < 0) {
    napi_value msg, err;
    napi_create_string_utf8(env, op->err, NAPI_AUTO_LENGTH, &msg);
    napi_create_error(env, NULL, msg, &err);
    napi_reject_deferred(env, op->deferred, err);
  } else {
    napi_value result = NULL;
    if (op->kind == OP_INGEST) {
      napi_get_boolean(env, op->rc == DDM_OK, &result);
    } else if (op->kind == OP_HISTORY) {
      napi_get_reference_value(env, op->keep[0], &result);
    } else if (op->kind == OP_RADIAL || op->kind == OP_FULL) {
      napi_create_object(env, &result);
      for (int i = 0; i < op->nkeep; i++) {
        napi_value v;
        if (napi_get_reference_value(env, op->keep[i], &v) == napi_ok) set_prop(env, result, op->names[i], v);
      }
    } else {
      napi_get_undefined(env, &result);
    }
    napi_resolve_deferred(env, op->deferred, result);
  }
  for (int i = 0; i < op->nkeep; i++) napi_delete_reference(env, op->keep[i]);
  napi_delete_reference(env, op->href);
  napi_delete_async_work(env, op->work);
  free(op);
}

static op_t *op_new(napi_env env, handle_t *h, op_kind kind) {
  op_t *op = (op_t *)calloc(1, sizeof(op_t));
  if (!op) napi_throw_error(env, NULL, "engine call: out of memory");
  else {
    op->h = h;
    op->kind = kind;
  }
  return op;
}

static void op_keep(napi_env env, op_t *op, napi_value v, const char *name) {
  if (op->nkeep < 4 && napi_create_reference(env, v, 1, &op->keep[op->nkeep]) == napi_ok) op->names[op->nkeep++] = name;
}

static napi_value op_queue(napi_env env, op_t *op, napi_value handle_value, const char *name) {
  napi_value promise, rname;
  if (napi_create_promise(env, &op->deferred, &promise) != napi_ok ||
      napi_create_reference(env, handle_value, 1, &op->href) != napi_ok ||
      napi_create_string_utf8(env, name, NAPI_AUTO_LENGTH, &rname) != napi_ok ||
      napi_create_async_work(env, NULL, rname, op_execute, op_complete, op, &op->work) != napi_ok ||
      napi_queue_async_work(env, op->work) != napi_ok) {
    for (int i = 0; i < op->nkeep; i++) napi_delete_reference(env, op->keep[i]);
    free(op);
    return throw_msg(env, "engine call: could not queue the work");
  }
  return promise;
}

/* the frame stack of an ingest call: typed array -> (data, dtype), length checked */
static int frames_arg(napi_env env, handle_t *h, napi_value arr, int32_t nframes, void **data, int32_t *dtype) {
  napi_typedarray_type t;
  size_t len, off;
  napi_value ab;
  if (napi_get_typedarray_info(env, arr, &t, &len, data, &ab, &off) != napi_ok ||
      (t != napi_uint8_array && t != napi_uint16_array && t != napi_float32_array)) {
    napi_throw_type_error(env, NULL, "ingest expects a Uint8Array, Uint16Array or Float32Array");
    return 0;
  }
  *dtype = t == napi_uint8_array ? DDM_U8 : t == napi_uint16_array ? DDM_U16 : DDM_F32;
  ddm_sizes s;
  ddm_get_sizes(h->e, &s);
  if ((int64_t)len != (int64_t)nframes * s.N * s.N) {
    char msg[128];
    snprintf(msg, sizeof msg, "ingest expects %lld values", (long long)nframes * s.N * s.N);
    napi_throw_error(env, NULL, msg);
    return 0;
  }
  return 1;
}

/* ingest(h, frames, nframes, firstSlot) -> Promise<boolean>: a frame stack, off the event loop */
static napi_value js_ingest(napi_env env, napi_callback_info info) {
  size_t argc = 4;
  napi_value argv[4];
  handle_t *h;
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  if (argc < 4 || !get_handle(env, argv[0], &h)) return NULL;
  int32_t nframes, slot, dtype;
  void *data;
  NAPI_CALL(env, napi_get_value_int32(env, argv[2], &nframes));
  NAPI_CALL(env, napi_get_value_int32(env, argv[3], &slot));
  if (!frames_arg(env, h, argv[1], nframes, &data, &dtype)) return NULL;
  op_t *op = op_new(env, h, OP_INGEST);
  if (!op) return NULL;
  op->in = data;
  op->dtype = dtype;
  op->nframes = nframes;
  op->slot = slot;
  op_keep(env, op, argv[1], NULL);
  return op_queue(env, op, argv[0], "ddm.ingest");
}

/* ingestSync(h, frames, nframes, slot) -> boolean: ingestLuma's single frame, on the caller's thread */
static napi_value js_ingest_sync(napi_env env, napi_callback_info info) {
  size_t argc = 4;
  napi_value argv[4], out;
  handle_t *h;
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  if (argc < 4 || !get_handle(env, argv[0], &h)) return NULL;
  int32_t nframes, slot, dtype;
  void *data;
  NAPI_CALL(env, napi_get_value_int32(env, argv[2], &nframes));
  NAPI_CALL(env, napi_get_value_int32(env, argv[3], &slot));
  if (!frames_arg(env, h, argv[1], nframes, &data, &dtype)) return NULL;
  pthread_mutex_lock(&h->mu);
  int rc = h->e ? ddm_ingest(h->e, data, dtype, nframes, slot) : DDM_ESTATE;
  char err[512] = "engine destroyed";
  if (rc < 0 && h->e) snprintf(err, sizeof err, "%s", ddm_last_error(h->e));
  pthread_mutex_unlock(&h->mu);
  if (rc < 0) return throw_msg(env, err);
  NAPI_CALL(env, napi_get_boolean(env, rc == DDM_OK, &out));
  return out;
}

/* computeTemporal(h, Uint32Array order, Float32Array mask) -> Promise: resolves when D is ready */
static napi_value js_compute_temporal(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  handle_t *h;
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  if (argc < 3 || !get_handle(env, argv[0], &h)) return NULL;
  napi_typedarray_type t1, t2;
  size_t n1, n2, o;
  void *d1, *d2;
  napi_value ab;
  if (napi_get_typedarray_info(env, argv[1], &t1, &n1, &d1, &ab, &o) != napi_ok || t1 != napi_uint32_array ||
      napi_get_typedarray_info(env, argv[2], &t2, &n2, &d2, &ab, &o) != napi_ok || t2 != napi_float32_array) {
    napi_throw_type_error(env, NULL, "computeTemporal expects (Uint32Array order, Float32Array mask)");
    return NULL;
  }
  ddm_sizes s;
  ddm_get_sizes(h->e, &s);
  if ((int)n1 != s.T || (int)n2 != s.T) {
    char msg[96];
    snprintf(msg, sizeof msg, "temporalMeta: order/mask must have length %d", s.T);
    return throw_msg(env, msg);
  }
  op_t *op = op_new(env, h, OP_TEMPORAL);
  if (!op) return NULL;
  op->order = (const uint32_t *)d1;
  op->mask = (const float *)d2;
  op_keep(env, op, argv[1], NULL);
  op_keep(env, op, argv[2], NULL);
  return op_queue(env, op, argv[0], "ddm.computeTemporal");
}

/* readRadial(h, withDrift) -> Promise<{D: Float32Array, P: Float32Array, drift: Float32Array}> */
static napi_value js_read_radial(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  handle_t *h;
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  if (argc < 1 || !get_handle(env, argv[0], &h)) return NULL;
  bool with_drift = true;
  if (argc >= 2) napi_get_value_bool(env, argv[1], &with_drift);
  ddm_sizes s;
  ddm_get_sizes(h->e, &s);
  void *D, *P, *dr;
  napi_value vD = make_typed(env, napi_float32_array, (size_t)s.NB * s.L, 4, &D);
  napi_value vP = make_typed(env, napi_float32_array, (size_t)s.NB, 4, &P);
  napi_value vdr = make_typed(env, napi_float32_array, with_drift ? (size_t)s.NB * s.L * 8 : 0, 4, &dr);
  if (!vD || !vP || !vdr) return throw_msg(env, "readRadial: allocation failed");
  op_t *op = op_new(env, h, OP_RADIAL);
  if (!op) return NULL;
  op->out[0] = D;
  op->out[1] = P;
  op->out[2] = with_drift ? dr : NULL;
  op_keep(env, op, vD, "D");
  op_keep(env, op, vP, "P");
  op_keep(env, op, vdr, "drift");
  return op_queue(env, op, argv[0], "ddm.readRadial");
}

/* readHistory(h) -> Promise<Float32Array> hist[(m*T + slot)*2 + {0,1}] */
static napi_value js_read_history(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  handle_t *h;
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  if (argc < 1 || !get_handle(env, argv[0], &h)) return NULL;
  ddm_sizes s;
  ddm_get_sizes(h->e, &s);
  void *d;
  napi_value v = make_typed(env, napi_float32_array, (size_t)s.hist_bytes / 4, 4, &d);
  if (!v) return throw_msg(env, "readHistory: allocation failed");
  op_t *op = op_new(env, h, OP_HISTORY);
  if (!op) return NULL;
  op->out[0] = d;
  op_keep(env, op, v, NULL);
  return op_queue(env, op, argv[0], "ddm.readHistory");
}

/* readMean(h) -> number (a scalar copy: synchronous) */
static napi_value js_read_mean(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1], out;
  handle_t *h;
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  if (argc < 1 || !get_handle(env, argv[0], &h)) return NULL;
  double m = 0;
  pthread_mutex_lock(&h->mu);
  int rc = h->e ? ddm_read_mean(h->e, &m) : DDM_ESTATE;
  char err[512] = "engine destroyed";
  if (rc < 0 && h->e) snprintf(err, sizeof err, "%s", ddm_last_error(h->e));
  pthread_mutex_unlock(&h->mu);
  if (rc != DDM_OK) return throw_msg(env, err);
  NAPI_CALL(env, napi_create_double(env, m, &out));
  return out;
}

/* idle(h) -> Promise: every call queued before it has finished on the GPU */
static napi_value js_idle(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  handle_t *h;
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  if (argc < 1 || !get_handle(env, argv[0], &h)) return NULL;
  op_t *op = op_new(env, h, OP_IDLE);
  if (!op) return NULL;
  return op_queue(env, op, argv[0], "ddm.idle");
}

/* readTemporalFull(h) -> Promise<{D: Float32Array(modes*L), C: Float32Array(modes*L*2), P: Float32Array(modes)}> */
static napi_value js_read_temporal_full(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  handle_t *h;
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  if (argc < 1 || !get_handle(env, argv[0], &h)) return NULL;
  ddm_sizes s;
  ddm_get_sizes(h->e, &s);
  void *D, *C, *P;
  napi_value vD = make_typed(env, napi_float32_array, (size_t)s.modes * s.L, 4, &D);
  napi_value vC = make_typed(env, napi_float32_array, (size_t)s.modes * s.L * 2, 4, &C);
  napi_value vP = make_typed(env, napi_float32_array, (size_t)s.modes, 4, &P);
  if (!vD || !vC || !vP) return throw_msg(env, "readTemporalFull: allocation failed");
  op_t *op = op_new(env, h, OP_FULL);
  if (!op) return NULL;
  op->out[0] = D;
  op->out[1] = C;
  op->out[2] = P;
  op_keep(env, op, vD, "D");
  op_keep(env, op, vC, "C");
  op_keep(env, op, vP, "P");
  return op_queue(env, op, argv[0], "ddm.readTemporalFull");
}

/* ------------------------------------------------------------------ per-q fits */

static void finalize_fitter(napi_env env, void *data, void *hint) {
  (void)env;
  (void)hint;
  ddm_fitter_destroy((ddm_fitter *)data);
}

/* fitterCreate() -> external handle of a ddm_fitter */
static napi_value js_fitter_create(napi_env env, napi_callback_info info) {
  (void)info;
  ddm_fitter *f = NULL;
  if (ddm_fitter_create(&f) != DDM_OK) {
    napi_throw_error(env, NULL, "ddm_fitter_create failed");
    return NULL;
  }
  napi_value out;
  NAPI_CALL(env, napi_create_external(env, f, finalize_fitter, NULL, &out));
  return out;
}

static int get_typed(napi_env env, napi_value obj, const char *name, napi_typedarray_type want, void **data,
                     size_t *len, int required) {
  napi_value v;
  bool has = false;
  *data = NULL;
  *len = 0;
  if (napi_has_named_property(env, obj, name, &has) != napi_ok || !has) return !required;
  if (napi_get_named_property(env, obj, name, &v) != napi_ok) return 0;
  napi_valuetype vt;
  napi_typeof(env, v, &vt);
  if (vt == napi_undefined || vt == napi_null) return !required;
  napi_typedarray_type t;
  napi_value ab;
  size_t off;
  if (napi_get_typedarray_info(env, v, &t, len, data, &ab, &off) != napi_ok || t != want) return 0;
  return 1;
}

/* fit(fitter, {nfits, ntauMax, knodes, model, seedMode, nstarts, maxIter, ntau: Int32Array,
 *              tau, data, weight, qnodes, bounds: Float64Array, seeds?, vmag?: Float64Array})
 *   -> {theta: Float64Array(6 nfits), chi2: Float64Array, iters: Int32Array, kernelMs} */
static napi_value js_fit(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2], out;
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  void *fp = NULL;
  if (argc < 2 || napi_get_value_external(env, argv[0], &fp) != napi_ok || !fp) {
    napi_throw_type_error(env, NULL, "fit(fitter, batch)");
    return NULL;
  }
  ddm_fitter *f = (ddm_fitter *)fp;
  ddm_fit_batch b;
  memset(&b, 0, sizeof b);
  b.nfits = (int32_t)get_prop_num(env, argv[1], "nfits", 0);
  b.ntau_max = (int32_t)get_prop_num(env, argv[1], "ntauMax", 0);
  b.knodes = (int32_t)get_prop_num(env, argv[1], "knodes", 1);
  b.model = (int32_t)get_prop_num(env, argv[1], "model", DDM_MODEL_SWIMMER_3D);
  b.seed_mode = (int32_t)get_prop_num(env, argv[1], "seedMode", DDM_SEED_GRID);
  b.nstarts = (int32_t)get_prop_num(env, argv[1], "nstarts", 3);
  b.max_iter = (int32_t)get_prop_num(env, argv[1], "maxIter", 400);
  void *ntau, *tau, *data, *w, *qn, *bd, *sd, *vm;
  size_t n0, n1, n2, n3, n4, n5, n6, n7;
  if (!get_typed(env, argv[1], "ntau", napi_int32_array, &ntau, &n0, 1) ||
      !get_typed(env, argv[1], "tau", napi_float64_array, &tau, &n1, 1) ||
      !get_typed(env, argv[1], "data", napi_float64_array, &data, &n2, 1) ||
      !get_typed(env, argv[1], "weight", napi_float64_array, &w, &n3, 1) ||
      !get_typed(env, argv[1], "qnodes", napi_float64_array, &qn, &n4, 1) ||
      !get_typed(env, argv[1], "bounds", napi_float64_array, &bd, &n5, 1) ||
      !get_typed(env, argv[1], "seeds", napi_float64_array, &sd, &n6, 0) ||
      !get_typed(env, argv[1], "vmag", napi_float64_array, &vm, &n7, 0)) {
    napi_throw_type_error(env, NULL, "fit: batch arrays must be typed (Int32Array ntau, Float64Array others)");
    return NULL;
  }
  size_t n = (size_t)(b.nfits > 0 ? b.nfits : 0);
  if (n0 < n || n1 < n * b.ntau_max || n2 < n * b.ntau_max || n3 < n || n4 < n * b.knodes || n5 < 4 * n ||
      (sd && n6 < 6 * n) || (vm && n7 < n)) {
    napi_throw_range_error(env, NULL, "fit: array shorter than nfits x its row size");
    return NULL;
  }
  b.ntau = (const int32_t *)ntau;
  b.tau = (const double *)tau;
  b.data = (const double *)data;
  b.weight = (const double *)w;
  b.qnodes = (const double *)qn;
  b.bounds = (const double *)bd;
  b.seeds = (const double *)sd;
  b.vmag = (const double *)vm;
  void *th, *chi, *it;
  napi_value vth = make_typed(env, napi_float64_array, 6 * (n ? n : 1), 8, &th);
  napi_value vchi = make_typed(env, napi_float64_array, n ? n : 1, 8, &chi);
  napi_value vit = make_typed(env, napi_int32_array, n ? n : 1, 4, &it);
  if (!vth || !vchi || !vit) {
    napi_throw_error(env, NULL, "fit: allocation failed");
    return NULL;
  }
  double ms = 0;
  int rc = ddm_fit(f, &b, (double *)th, (double *)chi, (int32_t *)it, &ms);
  if (rc != DDM_OK) {
    napi_throw_error(env, NULL, ddm_fitter_last_error(f));
    return NULL;
  }
  NAPI_CALL(env, napi_create_object(env, &out));
  set_prop(env, out, "theta", vth);
  set_prop(env, out, "chi2", vchi);
  set_prop(env, out, "iters", vit);
  set_prop(env, out, "kernelMs", num(env, ms));
  return out;
}

/* analyzeMotility(radial, opt): radial {N, NB, L, kCenter, kNodes, nModes, D, drift} as
 * Float64Array / Uint32Array, opt {dxUm, fps, geometry (0 = 3d, 1 = 2d), exposureLocked (0/1)}.
 * Returns the flat result; motility-fit.mjs shapes it like the reference's MotilityResult. */
static napi_value js_analyze_motility(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2], out;
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  if (argc < 2) {
    napi_throw_type_error(env, NULL, "analyzeMotility(radial, opt)");
    return NULL;
  }
  ddm_radial_record r;
  memset(&r, 0, sizeof r);
  r.N = (int32_t)get_prop_num(env, argv[0], "N", 0);
  r.NB = (int32_t)get_prop_num(env, argv[0], "NB", 0);
  r.L = (int32_t)get_prop_num(env, argv[0], "L", 0);
  void *kc, *kn, *nm, *D, *dr;
  size_t nkc, nkn = 0, nnm, nD, ndr;
  if (!get_typed(env, argv[0], "kCenter", napi_float64_array, &kc, &nkc, 1) ||
      !get_typed(env, argv[0], "kNodes", napi_float64_array, &kn, &nkn, 0) ||
      !get_typed(env, argv[0], "nModes", napi_uint32_array, &nm, &nnm, 1) ||
      !get_typed(env, argv[0], "D", napi_float64_array, &D, &nD, 1) ||
      !get_typed(env, argv[0], "drift", napi_float64_array, &dr, &ndr, 1)) {
    napi_throw_type_error(env, NULL, "analyzeMotility: radial arrays must be Float64Array (nModes Uint32Array)");
    return NULL;
  }
  const size_t NB = r.NB > 0 ? (size_t)r.NB : 0, L = r.L > 0 ? (size_t)r.L : 0;
  if (nkc < NB || nnm < NB || nD < NB * L || ndr < NB * L * 8) {
    napi_throw_range_error(env, NULL, "analyzeMotility: radial arrays shorter than NB, NB x L, NB x L x 8");
    return NULL;
  }
  r.knodes = (kn && NB) ? (int32_t)(nkn / NB) : 0;
  r.kCenter = (const double *)kc;
  r.kNodes = r.knodes ? (const double *)kn : NULL;
  r.nModes = (const uint32_t *)nm;
  r.D = (const double *)D;
  r.drift = (const double *)dr;
  ddm_motility_opts o;
  o.dx_um = get_prop_num(env, argv[1], "dxUm", 1);
  o.fps = get_prop_num(env, argv[1], "fps", 30);
  o.geometry = (int32_t)get_prop_num(env, argv[1], "geometry", DDM_GEOMETRY_3D);
  o.exposure_locked = (int32_t)get_prop_num(env, argv[1], "exposureLocked", 1);
  void *bins, *q, *sa, *ba, *sq, *st, *gt;
  const size_t cap = NB ? NB : 1;
  napi_value vb = make_typed(env, napi_int32_array, cap, 4, &bins), vq = make_typed(env, napi_float64_array, cap, 8, &q);
  napi_value vsa = make_typed(env, napi_float64_array, 2 * cap, 8, &sa), vba = make_typed(env, napi_float64_array, 2 * cap, 8, &ba);
  napi_value vsq = make_typed(env, napi_float64_array, cap, 8, &sq), vst = make_typed(env, napi_float64_array, cap, 8, &st);
  napi_value vg = make_typed(env, napi_float64_array, 6 * 5, 8, &gt);
  if (!vb || !vq || !vsa || !vba || !vsq || !vst || !vg) {
    napi_throw_error(env, NULL, "analyzeMotility: allocation failed");
    return NULL;
  }
  ddm_motility_detail det = {(int32_t)cap, 0, (int32_t *)bins, (double *)q, (double *)sa, (double *)ba, (double *)sq,
                             (double *)st};
  ddm_motility_result res;
  if (ddm_analyze_motility(&r, &o, &res, &det) != DDM_OK) {
    napi_throw_error(env, NULL, ddm_verdict_last_error());
    return NULL;
  }
  double *g = (double *)gt;
  for (int i = 0; i < res.ngates; i++) {
    g[5 * i] = res.gates[i].id;
    g[5 * i + 1] = res.gates[i].pass;
    g[5 * i + 2] = res.gates[i].value;
    g[5 * i + 3] = res.gates[i].sigma;
    g[5 * i + 4] = res.gates[i].resolved_bins;
  }
  NAPI_CALL(env, napi_create_object(env, &out));
  const struct { const char *k; double v; } f[] = {
      {"verdict", res.verdict}, {"physicsVerdict", res.physics_verdict}, {"insufficient", res.insufficient},
      {"ngates", res.ngates}, {"TSI", res.TSI}, {"alpha", res.alpha}, {"vbar", res.vbar}, {"sigmaV", res.sigma_v},
      {"Z", res.Z}, {"Ddiff", res.Ddiff}, {"DdiffBrownian", res.Ddiff_brownian}, {"dBIC", res.dBIC}, {"mu", res.mu},
      {"nData", res.nData}, {"nEff", res.nEff}, {"residualRho", res.residual_rho}, {"swimRss", res.swim_rss},
      {"swimVmag", res.swim_Vmag}, {"brownRss", res.brown_rss}, {"brownVmag", res.brown_Vmag},
      {"alphaResolved", res.alpha_resolved}, {"resolvedBins", res.resolved_bins}, {"nbins", res.nbins},
      {"nscaling", res.nscaling}, {"driftVx", res.drift.Vx}, {"driftVy", res.drift.Vy}, {"driftMag", res.drift.mag},
      {"driftSigma", res.drift.sigma}, {"driftAngleDeg", res.drift.angle_deg}, {"driftSamples", res.drift.samples},
      {"driftSignificant", res.drift.significant}, {"driftPairs", res.drift.pairs}};
  for (size_t i = 0; i < sizeof f / sizeof f[0]; i++) set_prop(env, out, f[i].k, num(env, f[i].v));
  set_prop(env, out, "gates", vg);
  set_prop(env, out, "bins", vb);
  set_prop(env, out, "q", vq);
  set_prop(env, out, "swimAmps", vsa);
  set_prop(env, out, "brownAmps", vba);
  set_prop(env, out, "scalingQ", vsq);
  set_prop(env, out, "scalingTauHalf", vst);
  return out;
}

/* ------------------------------------------------------------------ fit building blocks */

/* a Radial record object -> ddm_radial_record over its typed arrays (drift optional unless need_drift) */
static int radial_arg(napi_env env, napi_value obj, ddm_radial_record *r, int need_drift) {
  memset(r, 0, sizeof *r);
  r->N = (int32_t)get_prop_num(env, obj, "N", 0);
  r->NB = (int32_t)get_prop_num(env, obj, "NB", 0);
  r->L = (int32_t)get_prop_num(env, obj, "L", 0);
  void *kc, *kn, *nm, *D, *dr;
  size_t nkc, nkn = 0, nnm, nD, ndr = 0;
  if (!get_typed(env, obj, "kCenter", napi_float64_array, &kc, &nkc, 1) ||
      !get_typed(env, obj, "kNodes", napi_float64_array, &kn, &nkn, 0) ||
      !get_typed(env, obj, "nModes", napi_uint32_array, &nm, &nnm, 1) ||
      !get_typed(env, obj, "D", napi_float64_array, &D, &nD, 1) ||
      !get_typed(env, obj, "drift", napi_float64_array, &dr, &ndr, need_drift)) {
    napi_throw_type_error(env, NULL, "radial arrays must be Float64Array (nModes Uint32Array)");
    return 0;
  }
  const size_t NB = r->NB > 0 ? (size_t)r->NB : 0, L = r->L > 0 ? (size_t)r->L : 0;
  if (nkc < NB || nnm < NB || nD < NB * L || (need_drift && ndr < NB * L * 8)) {
    napi_throw_range_error(env, NULL, "radial arrays shorter than NB, NB x L (drift NB x L x 8)");
    return 0;
  }
  r->knodes = (kn && NB) ? (int32_t)(nkn / NB) : 0;
  r->kCenter = (const double *)kc;
  r->kNodes = r->knodes ? (const double *)kn : NULL;
  r->nModes = (const uint32_t *)nm;
  r->D = (const double *)D;
  r->drift = (dr && ndr >= NB * L * 8) ? (const double *)dr : NULL;
  return 1;
}

/* lagGrid(L) -> Int32Array */
static napi_value js_lag_grid(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  int32_t L = 0;
  if (argc < 1 || napi_get_value_int32(env, argv[0], &L) != napi_ok) return throw_msg(env, "lagGrid(L)");
  const int n = ddm_lag_grid(L, NULL, 0);
  if (n < 0) return throw_msg(env, ddm_verdict_last_error());
  void *d;
  napi_value v = make_typed(env, napi_int32_array, (size_t)n, 4, &d);
  if (!v) return throw_msg(env, "lagGrid: allocation failed");
  ddm_lag_grid(L, (int32_t *)d, n);
  return v;
}

/* usableBins(radial) -> Int32Array of prepare's bins */
static napi_value js_usable_bins(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  ddm_radial_record r;
  if (argc < 1 || !radial_arg(env, argv[0], &r, 0)) return NULL;
  const int n = ddm_usable_bins(&r, NULL, 0);
  if (n < 0) return throw_msg(env, ddm_verdict_last_error());
  void *d;
  napi_value v = make_typed(env, napi_int32_array, (size_t)n, 4, &d);
  if (!v) return throw_msg(env, "usableBins: allocation failed");
  ddm_usable_bins(&r, (int32_t *)d, n);
  return v;
}

/* swimmerIsf(geometry 0|1, Float64Array x, Z) -> Float64Array g */
static napi_value js_swimmer_isf(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  int32_t geo = 0;
  double Z = 0;
  napi_typedarray_type t;
  size_t n, off;
  void *x;
  napi_value ab;
  if (argc < 3 || napi_get_value_int32(env, argv[0], &geo) != napi_ok ||
      napi_get_typedarray_info(env, argv[1], &t, &n, &x, &ab, &off) != napi_ok || t != napi_float64_array ||
      napi_get_value_double(env, argv[2], &Z) != napi_ok)
    return throw_msg(env, "swimmerIsf(geometry, Float64Array x, Z)");
  void *g;
  napi_value v = make_typed(env, napi_float64_array, n, 8, &g);
  if (!v) return throw_msg(env, "swimmerIsf: allocation failed");
  if (ddm_swimmer_isf(geo, (const double *)x, (int32_t)n, Z, (double *)g) != DDM_OK)
    return throw_msg(env, ddm_verdict_last_error());
  return v;
}

/* estimateDrift(radial, {dxUm, fps, geometry 0|1}, swimmer {vbar, Z} | null) -> the estimate */
static napi_value js_estimate_drift(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3], out;
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  ddm_radial_record r;
  if (argc < 2 || !radial_arg(env, argv[0], &r, 1)) return NULL;
  ddm_motility_opts o;
  memset(&o, 0, sizeof o);
  o.dx_um = get_prop_num(env, argv[1], "dxUm", 1);
  o.fps = get_prop_num(env, argv[1], "fps", 30);
  o.geometry = (int32_t)get_prop_num(env, argv[1], "geometry", DDM_GEOMETRY_3D);
  double sw[2];
  const double *swp = NULL;
  if (argc >= 3) {
    napi_valuetype vt;
    napi_typeof(env, argv[2], &vt);
    if (vt == napi_object) {
      sw[0] = get_prop_num(env, argv[2], "vbar", 0);
      sw[1] = get_prop_num(env, argv[2], "Z", 1);
      swp = sw;
    }
  }
  ddm_drift_estimate d;
  if (ddm_estimate_drift(&r, &o, swp, &d) != DDM_OK) return throw_msg(env, ddm_verdict_last_error());
  NAPI_CALL(env, napi_create_object(env, &out));
  set_prop(env, out, "Vx", num(env, d.Vx));
  set_prop(env, out, "Vy", num(env, d.Vy));
  set_prop(env, out, "mag", num(env, d.mag));
  set_prop(env, out, "sigma", num(env, d.sigma));
  set_prop(env, out, "angleDeg", num(env, d.angle_deg));
  napi_value b;
  napi_get_boolean(env, d.significant != 0, &b);
  set_prop(env, out, "significant", b);
  set_prop(env, out, "pairs", num(env, d.pairs));
  set_prop(env, out, "samples", num(env, d.samples));
  return out;
}

/* ------------------------------------------------------------------ frame clock */

static void finalize_clock(napi_env env, void *data, void *hint) {
  (void)env;
  (void)hint;
  ddm_clock_destroy((ddm_clock *)data);
}

static ddm_clock *get_clock(napi_env env, napi_value v) {
  void *p = NULL;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
    napi_throw_type_error(env, NULL, "expected a frame clock handle");
    return NULL;
  }
  return (ddm_clock *)p;
}

/* clockCreate(fps, T) -> handle */
static napi_value js_clock_create(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2], out;
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  double fps = 0;
  int32_t T = 0;
  if (argc < 2 || napi_get_value_double(env, argv[0], &fps) != napi_ok ||
      napi_get_value_int32(env, argv[1], &T) != napi_ok)
    return throw_msg(env, "clockCreate(fps, T)");
  ddm_clock *c = NULL;
  if (ddm_clock_create(fps, T, &c) != DDM_OK) return throw_msg(env, "FrameClock: fps must be positive and T >= 1");
  NAPI_CALL(env, napi_create_external(env, c, finalize_clock, NULL, &out));
  return out;
}

/* clockPlace(c, tMs, source 0|1|2) -> [slot, n] or null (duplicate) */
static napi_value js_clock_place(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3], out;
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  ddm_clock *c;
  double t = 0;
  int32_t src = 0;
  if (argc < 3 || !(c = get_clock(env, argv[0])) || napi_get_value_double(env, argv[1], &t) != napi_ok ||
      napi_get_value_int32(env, argv[2], &src) != napi_ok)
    return NULL;
  int32_t slot = 0;
  int64_t n = 0;
  const int rc = ddm_clock_place(c, t, src, &slot, &n);
  if (rc < 0) return throw_msg(env, "FrameClock.place: time stamps must be finite numbers");
  if (rc == DDM_DROPPED) {
    napi_get_null(env, &out);
    return out;
  }
  NAPI_CALL(env, napi_create_array_with_length(env, 2, &out));
  napi_set_element(env, out, 0, num(env, slot));
  napi_set_element(env, out, 1, num(env, (double)n));
  return out;
}

/* clockDrop(c, n), clockReset(c) */
static napi_value js_clock_drop(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  ddm_clock *c;
  int64_t n = 0;
  if (argc < 2 || !(c = get_clock(env, argv[0])) || napi_get_value_int64(env, argv[1], &n) != napi_ok) return NULL;
  ddm_clock_drop(c, n);
  return NULL;
}

static napi_value js_clock_reset(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  ddm_clock *c;
  if (argc < 1 || !(c = get_clock(env, argv[0]))) return NULL;
  ddm_clock_reset(c);
  return NULL;
}

/* clockState(c) -> {newestN, frames, missed, duplicates, resets, valid, timeSource, medianDtMs} */
static napi_value js_clock_state(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1], out;
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  ddm_clock *c;
  if (argc < 1 || !(c = get_clock(env, argv[0]))) return NULL;
  ddm_clock_state s;
  ddm_clock_state_get(c, &s);
  NAPI_CALL(env, napi_create_object(env, &out));
  set_prop(env, out, "newestN", num(env, (double)s.newest_n));
  set_prop(env, out, "frames", num(env, (double)s.frames));
  set_prop(env, out, "missed", num(env, (double)s.missed));
  set_prop(env, out, "duplicates", num(env, (double)s.duplicates));
  set_prop(env, out, "resets", num(env, (double)s.resets));
  set_prop(env, out, "valid", num(env, s.valid));
  set_prop(env, out, "timeSource", num(env, s.time_source));
  set_prop(env, out, "medianDtMs", num(env, s.median_dt_ms));
  return out;
}

/* clockSnapshot(c, guard) -> {order: Uint32Array, mask: Float32Array, used} */
static napi_value js_clock_snapshot(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3], out;
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  ddm_clock *c;
  int32_t guard = 0, T = 0;
  if (argc < 3 || !(c = get_clock(env, argv[0])) || napi_get_value_int32(env, argv[1], &guard) != napi_ok ||
      napi_get_value_int32(env, argv[2], &T) != napi_ok || T < 1)
    return throw_msg(env, "clockSnapshot(c, guard, T)");
  void *o, *m;
  napi_value vo = make_typed(env, napi_uint32_array, (size_t)T, 4, &o);
  napi_value vm = make_typed(env, napi_float32_array, (size_t)T, 4, &m);
  if (!vo || !vm) return throw_msg(env, "snapshot: allocation failed");
  int32_t used = 0;
  if (ddm_clock_snapshot(c, guard, (uint32_t *)o, (float *)m, &used) != DDM_OK)
    return throw_msg(env, "FrameClock.snapshot: guard must be >= 0");
  NAPI_CALL(env, napi_create_object(env, &out));
  set_prop(env, out, "order", vo);
  set_prop(env, out, "mask", vm);
  set_prop(env, out, "used", num(env, used));
  return out;
}

/* shuffleOrder(Uint32Array order, Float32Array mask, Float64Array draws): permutes both in place */
static napi_value js_shuffle_order(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3], ab;
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  napi_typedarray_type t0, t1, t2;
  size_t n0, n1, n2, off;
  void *o, *m, *u;
  if (argc < 3 || napi_get_typedarray_info(env, argv[0], &t0, &n0, &o, &ab, &off) != napi_ok ||
      t0 != napi_uint32_array || napi_get_typedarray_info(env, argv[1], &t1, &n1, &m, &ab, &off) != napi_ok ||
      t1 != napi_float32_array || napi_get_typedarray_info(env, argv[2], &t2, &n2, &u, &ab, &off) != napi_ok ||
      t2 != napi_float64_array || n1 != n0 || (n0 > 1 && n2 < n0 - 1))
    return throw_msg(env, "shuffleOrder(Uint32Array order, Float32Array mask, Float64Array draws)");
  if (ddm_shuffle_order((uint32_t *)o, (float *)m, (int32_t)n0, (const double *)u) != DDM_OK)
    return throw_msg(env, "shuffleOrder: rand() must return values in [0, 1)");
  return NULL;
}

static napi_value init(napi_env env, napi_value exports) {
  napi_property_descriptor d[] = {
      {"create", NULL, js_create, NULL, NULL, NULL, napi_default, NULL},
      {"destroy", NULL, js_destroy, NULL, NULL, NULL, napi_default, NULL},
      {"sizes", NULL, js_sizes, NULL, NULL, NULL, napi_default, NULL},
      {"bins", NULL, js_bins, NULL, NULL, NULL, napi_default, NULL},
      {"ingest", NULL, js_ingest, NULL, NULL, NULL, napi_default, NULL},
      {"ingestSync", NULL, js_ingest_sync, NULL, NULL, NULL, napi_default, NULL},
      {"computeTemporal", NULL, js_compute_temporal, NULL, NULL, NULL, napi_default, NULL},
      {"readRadial", NULL, js_read_radial, NULL, NULL, NULL, napi_default, NULL},
      {"readHistory", NULL, js_read_history, NULL, NULL, NULL, napi_default, NULL},
      {"readMean", NULL, js_read_mean, NULL, NULL, NULL, napi_default, NULL},
      {"idle", NULL, js_idle, NULL, NULL, NULL, napi_default, NULL},
      {"readTemporalFull", NULL, js_read_temporal_full, NULL, NULL, NULL, napi_default, NULL},
      {"fitterCreate", NULL, js_fitter_create, NULL, NULL, NULL, napi_default, NULL},
      {"fit", NULL, js_fit, NULL, NULL, NULL, napi_default, NULL},
      {"analyzeMotility", NULL, js_analyze_motility, NULL, NULL, NULL, napi_default, NULL},
      {"lagGrid", NULL, js_lag_grid, NULL, NULL, NULL, napi_default, NULL},
      {"usableBins", NULL, js_usable_bins, NULL, NULL, NULL, napi_default, NULL},
      {"swimmerIsf", NULL, js_swimmer_isf, NULL, NULL, NULL, napi_default, NULL},
      {"estimateDrift", NULL, js_estimate_drift, NULL, NULL, NULL, napi_default, NULL},
      {"clockCreate", NULL, js_clock_create, NULL, NULL, NULL, napi_default, NULL},
      {"clockPlace", NULL, js_clock_place, NULL, NULL, NULL, napi_default, NULL},
      {"clockDrop", NULL, js_clock_drop, NULL, NULL, NULL, napi_default, NULL},
      {"clockReset", NULL, js_clock_reset, NULL, NULL, NULL, napi_default, NULL},
      {"clockState", NULL, js_clock_state, NULL, NULL, NULL, napi_default, NULL},
      {"clockSnapshot", NULL, js_clock_snapshot, NULL, NULL, NULL, napi_default, NULL},
      {"shuffleOrder", NULL, js_shuffle_order, NULL, NULL, NULL, napi_default, NULL},
  };
  if (napi_define_properties(env, exports, sizeof d / sizeof d[0], d) != napi_ok) return NULL;
  return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, init)
