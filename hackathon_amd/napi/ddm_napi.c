/*
 * ddm_napi.c -- thin N-API addon over libddm.so for the TypeScript facade.
 *
 * One JS function per C-ABI entry point (include/ddm.h); no logic of its own.
 * Errors: a negative status throws Error(ddm_last_error()), the way the
 * reference engine throws (motility-ddm src/ddm-engine.ts:126-163, :313, :332).
 * The engine handle is a napi_external whose finalizer calls ddm_destroy.
 * Built with gcc against node_api.h (N-API >= 6), linked to libddm.so with an
 * $ORIGIN rpath.
 */
#define NAPI_VERSION 6
#include <node_api.h>
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
} handle_t;

static void finalize_handle(napi_env env, void *data, void *hint) {
  (void)env;
  (void)hint;
  handle_t *h = (handle_t *)data;
  if (h->e) ddm_destroy(h->e);
  free(h);
}

static napi_value throw_status(napi_env env, ddm_engine *e, int rc) {
  (void)rc;
  napi_throw_error(env, NULL, ddm_last_error(e));
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
  if (rc != DDM_OK) return throw_status(env, NULL, rc);
  handle_t *h = (handle_t *)calloc(1, sizeof(handle_t));
  h->e = e;
  napi_value out;
  NAPI_CALL(env, napi_create_external(env, h, finalize_handle, NULL, &out));
  return out;
}

static napi_value js_destroy(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  void *p = NULL;
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  if (argc >= 1 && napi_get_value_external(env, argv[0], &p) == napi_ok && p) {
    handle_t *h = (handle_t *)p;
    if (h->e) ddm_destroy(h->e);
    h->e = NULL;
  }
  return NULL;
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

/* sizes(h) -> {N, NH, T, TPAD, L, NB, knodes, modes, binnedModes, histBytes, workItems, chunkFrames} */
static napi_value js_sizes(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1], out;
  handle_t *h;
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  if (argc < 1 || !get_handle(env, argv[0], &h)) return NULL;
  ddm_sizes s;
  int rc = ddm_get_sizes(h->e, &s);
  if (rc != DDM_OK) return throw_status(env, h->e, rc);
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
  if (!voff || !vmodes || !vkc || !vnm || !vkn) {
    napi_throw_error(env, NULL, "bins: allocation failed");
    return NULL;
  }
  int rc = ddm_get_bins(h->e, (uint32_t *)off, (int32_t *)modes, (double *)kc, (uint32_t *)nm, (double *)kn);
  if (rc != DDM_OK) return throw_status(env, h->e, rc);
  NAPI_CALL(env, napi_create_object(env, &out));
  set_prop(env, out, "binOffsets", voff);
  set_prop(env, out, "binModes", vmodes);
  set_prop(env, out, "kCenter", vkc);
  set_prop(env, out, "nModes", vnm);
  set_prop(env, out, "kNodes", vkn);
  return out;
}

/* ingest(h, Uint8Array | Uint16Array | Float32Array, nframes, firstSlot) -> boolean */
static napi_value js_ingest(napi_env env, napi_callback_info info) {
  size_t argc = 4;
  napi_value argv[4];
  handle_t *h;
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  if (argc < 4 || !get_handle(env, argv[0], &h)) return NULL;
  napi_typedarray_type t;
  size_t len;
  void *data;
  napi_value ab;
  size_t off;
  if (napi_get_typedarray_info(env, argv[1], &t, &len, &data, &ab, &off) != napi_ok) {
    napi_throw_type_error(env, NULL, "ingest expects a Uint8Array, Uint16Array or Float32Array");
    return NULL;
  }
  int32_t dtype = t == napi_uint8_array ? DDM_U8 : t == napi_uint16_array ? DDM_U16 : t == napi_float32_array ? DDM_F32 : -1;
  if (dtype < 0) {
    napi_throw_type_error(env, NULL, "ingest expects a Uint8Array, Uint16Array or Float32Array");
    return NULL;
  }
  int32_t nframes, slot;
  NAPI_CALL(env, napi_get_value_int32(env, argv[2], &nframes));
  NAPI_CALL(env, napi_get_value_int32(env, argv[3], &slot));
  ddm_sizes s;
  ddm_get_sizes(h->e, &s);
  if ((int64_t)len != (int64_t)nframes * s.N * s.N) {
    char msg[128];
    snprintf(msg, sizeof msg, "ingest expects %lld values", (long long)nframes * s.N * s.N);
    napi_throw_error(env, NULL, msg);
    return NULL;
  }
  int rc = ddm_ingest(h->e, data, dtype, nframes, slot);
  if (rc < 0) return throw_status(env, h->e, rc);
  napi_value out;
  NAPI_CALL(env, napi_get_boolean(env, rc == DDM_OK, &out));
  return out;
}

/* computeTemporal(h, Uint32Array order, Float32Array mask) */
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
    napi_throw_error(env, NULL, msg);
    return NULL;
  }
  int rc = ddm_compute_temporal(h->e, (const uint32_t *)d1, (const float *)d2);
  if (rc != DDM_OK) return throw_status(env, h->e, rc);
  return NULL;
}

/* readRadial(h, withDrift) -> {D: Float32Array, P: Float32Array, drift: Float32Array} */
static napi_value js_read_radial(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2], out;
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
  napi_value vdr = make_typed(env, napi_float32_array, (size_t)s.NB * s.L * 8, 4, &dr);
  if (!vD || !vP || !vdr) {
    napi_throw_error(env, NULL, "readRadial: allocation failed");
    return NULL;
  }
  int rc = ddm_read_radial(h->e, (float *)D, (float *)P, with_drift ? (float *)dr : NULL);
  if (rc != DDM_OK) return throw_status(env, h->e, rc);
  NAPI_CALL(env, napi_create_object(env, &out));
  set_prop(env, out, "D", vD);
  set_prop(env, out, "P", vP);
  set_prop(env, out, "drift", vdr);
  return out;
}

/* readHistory(h) -> Float32Array hist[(m*T + slot)*2 + {0,1}] */
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
  if (!v) {
    napi_throw_error(env, NULL, "readHistory: allocation failed");
    return NULL;
  }
  int rc = ddm_read_history(h->e, (float *)d);
  if (rc != DDM_OK) return throw_status(env, h->e, rc);
  return v;
}

static napi_value js_read_mean(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1], out;
  handle_t *h;
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  if (argc < 1 || !get_handle(env, argv[0], &h)) return NULL;
  double m = 0;
  int rc = ddm_read_mean(h->e, &m);
  if (rc != DDM_OK) return throw_status(env, h->e, rc);
  NAPI_CALL(env, napi_create_double(env, m, &out));
  return out;
}

static napi_value js_idle(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  handle_t *h;
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  if (argc < 1 || !get_handle(env, argv[0], &h)) return NULL;
  int rc = ddm_idle(h->e);
  if (rc != DDM_OK) return throw_status(env, h->e, rc);
  return NULL;
}

/* readTemporalFull(h) -> {D: Float32Array(modes*L), C: Float32Array(modes*L*2), P: Float32Array(modes)} */
static napi_value js_read_temporal_full(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1], out;
  handle_t *h;
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  if (argc < 1 || !get_handle(env, argv[0], &h)) return NULL;
  ddm_sizes s;
  ddm_get_sizes(h->e, &s);
  void *D, *C, *P;
  napi_value vD = make_typed(env, napi_float32_array, (size_t)s.modes * s.L, 4, &D);
  napi_value vC = make_typed(env, napi_float32_array, (size_t)s.modes * s.L * 2, 4, &C);
  napi_value vP = make_typed(env, napi_float32_array, (size_t)s.modes, 4, &P);
  if (!vD || !vC || !vP) {
    napi_throw_error(env, NULL, "readTemporalFull: allocation failed");
    return NULL;
  }
  int rc = ddm_read_temporal_full(h->e, (float *)D, (float *)C, (float *)P);
  if (rc != DDM_OK) return throw_status(env, h->e, rc);
  NAPI_CALL(env, napi_create_object(env, &out));
  set_prop(env, out, "D", vD);
  set_prop(env, out, "C", vC);
  set_prop(env, out, "P", vP);
  return out;
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
  b.max_iter = (int32_t)get_prop_num(env, argv[1], "maxIter", 100);
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

static napi_value init(napi_env env, napi_value exports) {
  napi_property_descriptor d[] = {
      {"create", NULL, js_create, NULL, NULL, NULL, napi_default, NULL},
      {"destroy", NULL, js_destroy, NULL, NULL, NULL, napi_default, NULL},
      {"sizes", NULL, js_sizes, NULL, NULL, NULL, napi_default, NULL},
      {"bins", NULL, js_bins, NULL, NULL, NULL, napi_default, NULL},
      {"ingest", NULL, js_ingest, NULL, NULL, NULL, napi_default, NULL},
      {"computeTemporal", NULL, js_compute_temporal, NULL, NULL, NULL, napi_default, NULL},
      {"readRadial", NULL, js_read_radial, NULL, NULL, NULL, napi_default, NULL},
      {"readHistory", NULL, js_read_history, NULL, NULL, NULL, napi_default, NULL},
      {"readMean", NULL, js_read_mean, NULL, NULL, NULL, napi_default, NULL},
      {"idle", NULL, js_idle, NULL, NULL, NULL, napi_default, NULL},
      {"readTemporalFull", NULL, js_read_temporal_full, NULL, NULL, NULL, napi_default, NULL},
      {"fitterCreate", NULL, js_fitter_create, NULL, NULL, NULL, napi_default, NULL},
      {"fit", NULL, js_fit, NULL, NULL, NULL, napi_default, NULL},
      {"analyzeMotility", NULL, js_analyze_motility, NULL, NULL, NULL, napi_default, NULL},
  };
  if (napi_define_properties(env, exports, sizeof d / sizeof d[0], d) != napi_ok) return NULL;
  return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, init)
