/*
 * Node-API binding of libpyr3hip (include/pyr3hip.h) for a TypeScript host.
 *
 * It exposes the stage interface the reference's Renderer drives
 * (src/renderer.ts:163-262): create / reset / iterate / present /
 * readHistogram / readImage / destroy, taking the same packed blobs the
 * reference's packers produce (packXforms, packXformDistrib, the 128-byte
 * chaos uniforms, the DE and visualize uniform blocks, the baked palette) as
 * TypedArrays. Errors from the C ABI become thrown JS Errors carrying
 * p3_last_error(). As in the reference, reset/iterate/present return after
 * queueing GPU work, and the readbacks return Promises
 * (src/renderer.ts:196-209 readIndexMap / readHistogram): the copy runs on a
 * libuv worker thread (napi_async_work) that waits for the handle's stream
 * while the JS thread goes on; the promise resolves on the JS thread. A
 * per-handle mutex keeps the C ABI's one-thread-at-a-time contract: a call
 * made while a readback is in flight waits for it.
 */
#include <node_api.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "pyr3hip.h"

#define CHECK_NAPI(call)                                              \
  do {                                                                \
    if ((call) != napi_ok) {                                          \
      napi_throw_error(env, NULL, "pyr3hip: N-API call failed");      \
      return NULL;                                                    \
    }                                                                 \
  } while (0)

static napi_value throw_p3(napi_env env, p3_renderer* r, int rc) {
  char msg[600];
  snprintf(msg, sizeof msg, "pyr3hip error %d: %s", rc, p3_last_error(r));
  napi_throw_error(env, NULL, msg);
  return NULL;
}

/* A renderer handle and the lock that serialises calls on it. */
typedef struct {
  p3_renderer* r;
  pthread_mutex_t mu;
} handle_t;

static void finalize_handle(napi_env env, void* data, void* hint) {
  (void)env;
  (void)hint;
  handle_t* h = (handle_t*)data;
  pthread_mutex_lock(&h->mu);
  p3_destroy(h->r);
  pthread_mutex_unlock(&h->mu);
  pthread_mutex_destroy(&h->mu);
  free(h);
}

/* Typed-array view: pointer and element count (NULL for undefined/null). */
static int typed(napi_env env, napi_value v, void** data, size_t* n) {
  napi_valuetype t;
  *data = NULL;
  *n = 0;
  if (napi_typeof(env, v, &t) != napi_ok) return -1;
  if (t == napi_undefined || t == napi_null) return 0;
  bool is_ta = false;
  if (napi_is_typedarray(env, v, &is_ta) != napi_ok || !is_ta) return -1;
  napi_typedarray_type tt;
  napi_value buf;
  size_t off;
  if (napi_get_typedarray_info(env, v, &tt, n, data, &buf, &off) != napi_ok) return -1;
  return 0;
}

static handle_t* handle_of(napi_env env, napi_value v) {
  void* p = NULL;
  if (napi_get_value_external(env, v, &p) != napi_ok) return NULL;
  return (handle_t*)p;
}

/* Runs `call` on the handle under its lock and throws on failure. */
#define LOCKED(h, call)                                  \
  do {                                                   \
    pthread_mutex_lock(&(h)->mu);                        \
    int rc_ = (call);                                    \
    pthread_mutex_unlock(&(h)->mu);                      \
    if (rc_) return throw_p3(env, (h)->r, rc_);          \
  } while (0)

/* create(width, height, oversample, taps: Float32Array) -> handle */
static napi_value js_create(napi_env env, napi_callback_info info) {
  size_t argc = 4;
  napi_value a[4];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  int32_t w, h, os;
  CHECK_NAPI(napi_get_value_int32(env, a[0], &w));
  CHECK_NAPI(napi_get_value_int32(env, a[1], &h));
  CHECK_NAPI(napi_get_value_int32(env, a[2], &os));
  void* taps;
  size_t nt;
  if (typed(env, a[3], &taps, &nt)) return throw_p3(env, NULL, P3_E_ARG);
  p3_renderer* r = NULL;
  int rc = p3_create(w, h, os, (const float*)taps, (int)nt, &r);
  if (rc) return throw_p3(env, NULL, rc);
  handle_t* hd = (handle_t*)calloc(1, sizeof(handle_t));
  if (!hd) {
    p3_destroy(r);
    napi_throw_error(env, NULL, "pyr3hip: out of host memory");
    return NULL;
  }
  hd->r = r;
  pthread_mutex_init(&hd->mu, NULL);
  napi_value ext;
  CHECK_NAPI(napi_create_external(env, hd, finalize_handle, NULL, &ext));
  return ext;
}

/* reset(handle, palette: Float32Array(1024)) */
static napi_value js_reset(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value a[2];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  handle_t* h = handle_of(env, a[0]);
  void* pal;
  size_t n;
  if (!h || typed(env, a[1], &pal, &n) || n != 1024) return throw_p3(env, h ? h->r : NULL, P3_E_ARG);
  LOCKED(h, p3_reset(h->r, (const float*)pal));
  return NULL;
}

/* iterate(handle, xforms: Float32Array, distrib: Uint32Array, uniforms: Uint32Array(32)) */
static napi_value js_iterate(napi_env env, napi_callback_info info) {
  size_t argc = 4;
  napi_value a[4];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  handle_t* h = handle_of(env, a[0]);
  void *xf, *ds, *u;
  size_t nxf, nds, nu;
  if (!h || typed(env, a[1], &xf, &nxf) || typed(env, a[2], &ds, &nds) ||
      typed(env, a[3], &u, &nu) || nu != P3_CHAOS_UNIFORM_WORDS || nxf % P3_XFORM_FLOATS ||
      nds % P3_GRAIN)
    return throw_p3(env, h ? h->r : NULL, P3_E_ARG);
  LOCKED(h, p3_iterate(h->r, (const float*)xf, (int)(nxf / P3_XFORM_FLOATS),
                       (const uint32_t*)ds, (int)(nds / P3_GRAIN), (const uint32_t*)u));
  return NULL;
}

/* present(handle, deUniforms: Uint32Array(8) | null, vizUniforms: Uint32Array(24)) */
static napi_value js_present(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value a[3];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  handle_t* h = handle_of(env, a[0]);
  void *de, *viz;
  size_t nde, nviz;
  if (!h || typed(env, a[1], &de, &nde) || typed(env, a[2], &viz, &nviz) ||
      (de && nde != P3_DE_UNIFORM_WORDS) || nviz != P3_VIZ_UNIFORM_WORDS)
    return throw_p3(env, h ? h->r : NULL, P3_E_ARG);
  LOCKED(h, p3_present(h->r, (const uint32_t*)de, (const uint32_t*)viz));
  return NULL;
}

/* ---- asynchronous readbacks ---------------------------------------- */

enum { RD_HIST = 0, RD_IMAGE = 1, RD_INDEX = 2 };

typedef struct {
  handle_t* h;
  int kind;
  int dims[4];                 /* width, height, superW, superH */
  size_t n;                    /* elements of the main buffer */
  void* data;                  /* main buffer (JS-owned, kept alive by ref) */
  void* data2;                 /* readIndexMap: the count buffer */
  napi_ref ref, ref2, handle_ref;
  napi_deferred deferred;
  napi_async_work work;
  int rc;
  char err[600];
} read_job;

static void read_execute(napi_env env, void* p) {
  (void)env;
  read_job* j = (read_job*)p;
  pthread_mutex_lock(&j->h->mu);
  if (j->kind == RD_HIST) {
    j->rc = p3_read_histogram(j->h->r, (uint32_t*)j->data, j->n);
  } else if (j->kind == RD_IMAGE) {
    j->rc = p3_read_image(j->h->r, (uint8_t*)j->data, j->n);
  } else {
    /* idx_sum, then the count channel of the histogram */
    j->rc = p3_read_index(j->h->r, (uint32_t*)j->data, j->n);
    if (!j->rc) {
      uint32_t* hist = (uint32_t*)malloc(j->n * 4 * sizeof(uint32_t));
      if (!hist) {
        j->rc = P3_E_ARG;
      } else {
        j->rc = p3_read_histogram(j->h->r, hist, j->n * 4);
        uint32_t* cnt = (uint32_t*)j->data2;
        for (size_t i = 0; !j->rc && i < j->n; ++i) cnt[i] = hist[4 * i + 3];
        free(hist);
      }
    }
  }
  if (j->rc) snprintf(j->err, sizeof j->err, "pyr3hip error %d: %s", j->rc, p3_last_error(j->h->r));
  pthread_mutex_unlock(&j->h->mu);
}

static napi_value typed_from_ref(napi_env env, napi_ref ref, napi_typedarray_type t, size_t n) {
  napi_value buf, arr;
  if (napi_get_reference_value(env, ref, &buf) != napi_ok) return NULL;
  if (napi_create_typedarray(env, t, n, buf, 0, &arr) != napi_ok) return NULL;
  return arr;
}

static void set_int(napi_env env, napi_value obj, const char* key, int v) {
  napi_value x;
  if (napi_create_int32(env, v, &x) == napi_ok) napi_set_named_property(env, obj, key, x);
}

static void read_complete(napi_env env, napi_status status, void* p) {
  read_job* j = (read_job*)p;
  napi_value result = NULL;
  if (status != napi_ok || j->rc) {
    napi_value msg;
    napi_create_string_utf8(env, j->rc ? j->err : "pyr3hip: readback cancelled",
                            NAPI_AUTO_LENGTH, &msg);
    napi_create_error(env, NULL, msg, &result);
    napi_reject_deferred(env, j->deferred, result);
  } else {
    if (j->kind == RD_IMAGE) {
      result = typed_from_ref(env, j->ref, napi_uint8_array, j->n);
    } else if (j->kind == RD_HIST) {
      napi_create_object(env, &result);
      napi_set_named_property(env, result, "rgba",
                              typed_from_ref(env, j->ref, napi_uint32_array, j->n));
      set_int(env, result, "superW", j->dims[2]);
      set_int(env, result, "superH", j->dims[3]);
    } else {
      napi_create_object(env, &result);
      napi_set_named_property(env, result, "idxSum",
                              typed_from_ref(env, j->ref, napi_uint32_array, j->n));
      napi_set_named_property(env, result, "count",
                              typed_from_ref(env, j->ref2, napi_uint32_array, j->n));
      set_int(env, result, "width", j->dims[2]);
      set_int(env, result, "height", j->dims[3]);
    }
    napi_resolve_deferred(env, j->deferred, result);
  }
  napi_delete_reference(env, j->ref);
  if (j->ref2) napi_delete_reference(env, j->ref2);
  napi_delete_reference(env, j->handle_ref);
  napi_delete_async_work(env, j->work);
  free(j);
}

/* Queues a readback of `kind` and returns its promise. */
static napi_value start_read(napi_env env, napi_callback_info info, int kind) {
  size_t argc = 1;
  napi_value a[1];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  handle_t* h = handle_of(env, a[0]);
  if (!h) return throw_p3(env, NULL, P3_E_ARG);
  read_job* j = (read_job*)calloc(1, sizeof(read_job));
  if (!j) {
    napi_throw_error(env, NULL, "pyr3hip: out of host memory");
    return NULL;
  }
  j->h = h;
  j->kind = kind;
  pthread_mutex_lock(&h->mu);
  int rc = p3_dims(h->r, j->dims);
  pthread_mutex_unlock(&h->mu);
  if (rc) {
    free(j);
    return throw_p3(env, h->r, rc);
  }
  const size_t super_bins = (size_t)j->dims[2] * j->dims[3];
  j->n = kind == RD_HIST ? super_bins * 4 : kind == RD_IMAGE ? (size_t)j->dims[0] * j->dims[1] * 4
                                                             : super_bins;
  const size_t bytes = kind == RD_IMAGE ? j->n : j->n * 4;
  napi_value buf, buf2, promise, name;
  if (napi_create_arraybuffer(env, bytes, &j->data, &buf) != napi_ok ||
      napi_create_reference(env, buf, 1, &j->ref) != napi_ok ||
      napi_create_reference(env, a[0], 1, &j->handle_ref) != napi_ok) {
    free(j);
    napi_throw_error(env, NULL, "pyr3hip: N-API call failed");
    return NULL;
  }
  if (kind == RD_INDEX) {
    CHECK_NAPI(napi_create_arraybuffer(env, bytes, &j->data2, &buf2));
    CHECK_NAPI(napi_create_reference(env, buf2, 1, &j->ref2));
  }
  CHECK_NAPI(napi_create_promise(env, &j->deferred, &promise));
  CHECK_NAPI(napi_create_string_utf8(env, "pyr3hip.read", NAPI_AUTO_LENGTH, &name));
  CHECK_NAPI(napi_create_async_work(env, NULL, name, read_execute, read_complete, j, &j->work));
  CHECK_NAPI(napi_queue_async_work(env, j->work));
  return promise;
}

/* readHistogram(handle) -> Promise<{rgba: Uint32Array(superW * superH * 4), superW, superH}> */
static napi_value js_read_histogram(napi_env env, napi_callback_info info) {
  return start_read(env, info, RD_HIST);
}

/* readImage(handle) -> Promise<Uint8Array(width * height * 4)>, RGBA8 row-major */
static napi_value js_read_image(napi_env env, napi_callback_info info) {
  return start_read(env, info, RD_IMAGE);
}

/* readIndexMap(handle) -> Promise<{idxSum, count: Uint32Array(superW * superH), width, height}>
 * (super-res dims, as src/renderer.ts:196-204). */
static napi_value js_read_index_map(napi_env env, napi_callback_info info) {
  return start_read(env, info, RD_INDEX);
}

/* sync(handle) */
static napi_value js_sync(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value a[1];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  handle_t* h = handle_of(env, a[0]);
  if (!h) return throw_p3(env, NULL, P3_E_ARG);
  LOCKED(h, p3_sync(h->r));
  return NULL;
}

static napi_value init(napi_env env, napi_value exports) {
  napi_property_descriptor props[] = {
      {"create", NULL, js_create, NULL, NULL, NULL, napi_enumerable, NULL},
      {"reset", NULL, js_reset, NULL, NULL, NULL, napi_enumerable, NULL},
      {"iterate", NULL, js_iterate, NULL, NULL, NULL, napi_enumerable, NULL},
      {"present", NULL, js_present, NULL, NULL, NULL, napi_enumerable, NULL},
      {"readHistogram", NULL, js_read_histogram, NULL, NULL, NULL, napi_enumerable, NULL},
      {"readImage", NULL, js_read_image, NULL, NULL, NULL, napi_enumerable, NULL},
      {"readIndexMap", NULL, js_read_index_map, NULL, NULL, NULL, napi_enumerable, NULL},
      {"sync", NULL, js_sync, NULL, NULL, NULL, napi_enumerable, NULL},
  };
  CHECK_NAPI(napi_define_properties(env, exports, sizeof props / sizeof props[0], props));
  return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, init)
