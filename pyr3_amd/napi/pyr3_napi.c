/*
 * Node-API binding of libpyr3hip (include/pyr3hip.h) for a TypeScript host.
 *
 * It exposes the stage interface the reference's Renderer drives
 * (src/renderer.ts:163-262): create / reset / iterate / present /
 * readHistogram / readImage / destroy, taking the same packed blobs the
 * reference's packers produce (packXforms, packXformDistrib, the 128-byte
 * chaos uniforms, the DE and visualize uniform blocks, the baked palette) as
 * TypedArrays. Errors from the C ABI become thrown JS Errors carrying
 * p3_last_error(). Calls are synchronous; iterate/present return after
 * queueing GPU work, reads block until the stream drains.
 */
#include <node_api.h>
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

static void finalize_handle(napi_env env, void* data, void* hint) {
  (void)env;
  (void)hint;
  p3_destroy((p3_renderer*)data);
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

static p3_renderer* handle_of(napi_env env, napi_value v) {
  void* p = NULL;
  if (napi_get_value_external(env, v, &p) != napi_ok) return NULL;
  return (p3_renderer*)p;
}

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
  napi_value ext;
  CHECK_NAPI(napi_create_external(env, r, finalize_handle, NULL, &ext));
  return ext;
}

/* reset(handle, palette: Float32Array(1024)) */
static napi_value js_reset(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value a[2];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  p3_renderer* r = handle_of(env, a[0]);
  void* pal;
  size_t n;
  if (!r || typed(env, a[1], &pal, &n) || n != 1024) return throw_p3(env, r, P3_E_ARG);
  int rc = p3_reset(r, (const float*)pal);
  return rc ? throw_p3(env, r, rc) : NULL;
}

/* iterate(handle, xforms: Float32Array, distrib: Uint32Array, uniforms: Uint32Array(32)) */
static napi_value js_iterate(napi_env env, napi_callback_info info) {
  size_t argc = 4;
  napi_value a[4];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  p3_renderer* r = handle_of(env, a[0]);
  void *xf, *ds, *u;
  size_t nxf, nds, nu;
  if (!r || typed(env, a[1], &xf, &nxf) || typed(env, a[2], &ds, &nds) ||
      typed(env, a[3], &u, &nu) || nu != P3_CHAOS_UNIFORM_WORDS || nxf % P3_XFORM_FLOATS ||
      nds % P3_GRAIN)
    return throw_p3(env, r, P3_E_ARG);
  int rc = p3_iterate(r, (const float*)xf, (int)(nxf / P3_XFORM_FLOATS), (const uint32_t*)ds,
                      (int)(nds / P3_GRAIN), (const uint32_t*)u);
  return rc ? throw_p3(env, r, rc) : NULL;
}

/* present(handle, deUniforms: Uint32Array(8) | null, vizUniforms: Uint32Array(24)) */
static napi_value js_present(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value a[3];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  p3_renderer* r = handle_of(env, a[0]);
  void *de, *viz;
  size_t nde, nviz;
  if (!r || typed(env, a[1], &de, &nde) || typed(env, a[2], &viz, &nviz) ||
      (de && nde != P3_DE_UNIFORM_WORDS) || nviz != P3_VIZ_UNIFORM_WORDS)
    return throw_p3(env, r, P3_E_ARG);
  int rc = p3_present(r, (const uint32_t*)de, (const uint32_t*)viz);
  return rc ? throw_p3(env, r, rc) : NULL;
}

/* readHistogram(handle) -> Uint32Array(superW * superH * 4) */
static napi_value js_read_histogram(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value a[1];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  p3_renderer* r = handle_of(env, a[0]);
  int d[4];
  if (!r || p3_dims(r, d)) return throw_p3(env, r, P3_E_ARG);
  size_t n = (size_t)d[2] * d[3] * 4;
  void* data;
  napi_value buf, arr;
  CHECK_NAPI(napi_create_arraybuffer(env, n * 4, &data, &buf));
  int rc = p3_read_histogram(r, (uint32_t*)data, n);
  if (rc) return throw_p3(env, r, rc);
  CHECK_NAPI(napi_create_typedarray(env, napi_uint32_array, n, buf, 0, &arr));
  return arr;
}

/* readImage(handle) -> Uint8Array(width * height * 4), RGBA8 row-major */
static napi_value js_read_image(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value a[1];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  p3_renderer* r = handle_of(env, a[0]);
  int d[4];
  if (!r || p3_dims(r, d)) return throw_p3(env, r, P3_E_ARG);
  size_t n = (size_t)d[0] * d[1] * 4;
  void* data;
  napi_value buf, arr;
  CHECK_NAPI(napi_create_arraybuffer(env, n, &data, &buf));
  int rc = p3_read_image(r, (uint8_t*)data, n);
  if (rc) return throw_p3(env, r, rc);
  CHECK_NAPI(napi_create_typedarray(env, napi_uint8_array, n, buf, 0, &arr));
  return arr;
}

/* sync(handle) */
static napi_value js_sync(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value a[1];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  p3_renderer* r = handle_of(env, a[0]);
  if (!r) return throw_p3(env, r, P3_E_ARG);
  int rc = p3_sync(r);
  return rc ? throw_p3(env, r, rc) : NULL;
}

static napi_value init(napi_env env, napi_value exports) {
  napi_property_descriptor props[] = {
      {"create", NULL, js_create, NULL, NULL, NULL, napi_enumerable, NULL},
      {"reset", NULL, js_reset, NULL, NULL, NULL, napi_enumerable, NULL},
      {"iterate", NULL, js_iterate, NULL, NULL, NULL, napi_enumerable, NULL},
      {"present", NULL, js_present, NULL, NULL, NULL, napi_enumerable, NULL},
      {"readHistogram", NULL, js_read_histogram, NULL, NULL, NULL, napi_enumerable, NULL},
      {"readImage", NULL, js_read_image, NULL, NULL, NULL, napi_enumerable, NULL},
      {"sync", NULL, js_sync, NULL, NULL, NULL, napi_enumerable, NULL},
  };
  CHECK_NAPI(napi_define_properties(env, exports, sizeof props / sizeof props[0], props));
  return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, init)
