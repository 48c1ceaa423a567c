/*
 * mbvh_napi.c — Node-API addon over libmbvh.so. It only marshals typed arrays and numbers into
 * the C ABI of include/mbvh.h and turns status codes into JavaScript exceptions; all logic lives
 * in the library (kernels, upload) and in the JavaScript package (the MeshBVH API).
 *
 * Device contexts are process-wide, one per device, created on first use and destroyed by an
 * environment cleanup hook. Node runs that hook before the finalizers of externals, so the hook first
 * releases every scene, batch and mesh still alive (the handles are kept in lists) and the
 * finalizers that run after it find nothing left to free on a destroyed context.
 */
#define NAPI_VERSION 4
#include <math.h>
#include <node_api.h>
#include <pthread.h>

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mbvh.h"

#define MAX_DEVICES 64
#define MAX_ARGS 16

static mbvh_ctx* g_ctx[MAX_DEVICES];

typedef struct MeshHandle {
  mbvh_mesh* mesh;
  int device;
  int batch_refs;  /* batches using this mesh: it cannot be freed explicitly meanwhile */
  int async_refs;  /* asynchronous queries in flight on this mesh: likewise */
  struct MeshHandle *prev, *next;  /* the live mesh handles */
} MeshHandle;

typedef struct BatchHandle {
  mbvh_batch* batch;
  mbvh_scene* scene;    /* scenes use the same handle (one of batch / scene is set) */
  int device;
  uint32_t n_meshes;
  MeshHandle** meshes;  /* kept alive (and unfreeable) while the batch exists */
  napi_ref* refs;
  struct BatchHandle *prev, *next;  /* the live batch / scene handles */
} BatchHandle;

/* Handles whose externals exist, for the environment cleanup hook (main thread only). */
static MeshHandle* g_meshes;
static BatchHandle* g_batches;

static void link_mesh(MeshHandle* h) {
  h->prev = NULL;
  h->next = g_meshes;
  if (g_meshes) g_meshes->prev = h;
  g_meshes = h;
}

static void unlink_mesh(MeshHandle* h) {
  if (h->prev) h->prev->next = h->next;
  else if (g_meshes == h) g_meshes = h->next;
  if (h->next) h->next->prev = h->prev;
  h->prev = h->next = NULL;
}

static void link_batch(BatchHandle* b) {
  b->prev = NULL;
  b->next = g_batches;
  if (g_batches) g_batches->prev = b;
  g_batches = b;
}

static void unlink_batch(BatchHandle* b) {
  if (b->prev) b->prev->next = b->next;
  else if (g_batches == b) g_batches = b->next;
  if (b->next) b->next->prev = b->prev;
  b->prev = b->next = NULL;
}

/* ------------------------------------------------------------------------------------------ */

static napi_value throw_status(napi_env env, int rc, const char* where) {
  char msg[768];
  snprintf(msg, sizeof(msg), "%s failed (mbvh status %d): %s", where, rc, mbvh_last_error());
  char code[32];
  snprintf(code, sizeof(code), "MBVH_%d", -rc);
  napi_throw_error(env, code, msg);
  return NULL;
}

static napi_value throw_msg(napi_env env, const char* msg) {
  napi_throw_type_error(env, NULL, msg);
  return NULL;
}

#define CHECK_NAPI(call)                                   \
  do {                                                     \
    if ((call) != napi_ok) return throw_msg(env, #call);   \
  } while (0)

static int get_args(napi_env env, napi_callback_info info, size_t want, napi_value* argv) {
  size_t argc = MAX_ARGS;
  napi_value tmp[MAX_ARGS];
  if (napi_get_cb_info(env, info, &argc, tmp, NULL, NULL) != napi_ok) return 0;
  for (size_t i = 0; i < want; i++) {
    if (i < argc) {
      argv[i] = tmp[i];
    } else {
      napi_get_undefined(env, &argv[i]);
    }
  }
  return 1;
}

static int is_nullish(napi_env env, napi_value v) {
  napi_valuetype t;
  if (napi_typeof(env, v, &t) != napi_ok) return 1;
  return t == napi_undefined || t == napi_null;
}

/* Typed array view: data pointer, element count and element size. */
typedef struct {
  void* data;
  size_t length;
  int elt;
  napi_typedarray_type type;
} View;

static int get_view(napi_env env, napi_value v, View* out) {
  memset(out, 0, sizeof(*out));
  bool is_ta = false;
  if (napi_is_typedarray(env, v, &is_ta) != napi_ok || !is_ta) return 0;
  napi_value ab;
  size_t off;
  if (napi_get_typedarray_info(env, v, &out->type, &out->length, &out->data, &ab, &off) != napi_ok) return 0;
  switch (out->type) {
    case napi_int8_array: case napi_uint8_array: case napi_uint8_clamped_array: out->elt = 1; break;
    case napi_int16_array: case napi_uint16_array: out->elt = 2; break;
    case napi_int32_array: case napi_uint32_array: case napi_float32_array: out->elt = 4; break;
    default: out->elt = 8; break;
  }
  return 1;
}

static int out_view(napi_env env, napi_value v, napi_typedarray_type want, size_t n, void** ptr);

static int get_f64(napi_env env, napi_value v, double* out) {
  return napi_get_value_double(env, v, out) == napi_ok;
}

static int get_i32(napi_env env, napi_value v, int32_t* out) {
  return napi_get_value_int32(env, v, out) == napi_ok;
}

static int get_named_i32(napi_env env, napi_value obj, const char* name, int32_t dflt, int32_t* out) {
  bool has = false;
  *out = dflt;
  if (napi_has_named_property(env, obj, name, &has) != napi_ok || !has) return 1;
  napi_value v;
  if (napi_get_named_property(env, obj, name, &v) != napi_ok) return 0;
  if (is_nullish(env, v)) return 1;
  napi_valuetype t;
  napi_typeof(env, v, &t);
  if (t == napi_boolean) {
    bool b;
    napi_get_value_bool(env, v, &b);
    *out = b ? 1 : 0;
    return 1;
  }
  return get_i32(env, v, out);
}

static float to_float(double v) { return (float)v; }

static int context_for(napi_env env, int32_t device, mbvh_ctx** out) {
  (void)env;
  if (device < 0 || device >= MAX_DEVICES) return MBVH_ERR_NO_DEVICE;
  if (!g_ctx[device]) {
    int rc = mbvh_ctx_create(device, &g_ctx[device]);
    if (rc) return rc;
  }
  *out = g_ctx[device];
  return MBVH_OK;
}

static void stop_query_thread(void);

static napi_env g_env;

static void cleanup_contexts(void* arg) {
  (void)arg;
  g_env = NULL;
  stop_query_thread();
  /* scenes and batches before the meshes they hold, every library object before its context */
  for (BatchHandle* b = g_batches; b; b = b->next) {
    if (b->batch) mbvh_batch_free(b->batch);
    if (b->scene) mbvh_scene_free(b->scene);
    b->batch = NULL;
    b->scene = NULL;
  }
  for (MeshHandle* h = g_meshes; h; h = h->next) {
    if (h->mesh) mbvh_mesh_free(h->mesh);
    h->mesh = NULL;
  }
  for (int i = 0; i < MAX_DEVICES; i++) {
    if (g_ctx[i]) mbvh_ctx_destroy(g_ctx[i]);
    g_ctx[i] = NULL;
  }
}

static void finalize_mesh(napi_env env, void* data, void* hint) {
  (void)env;
  (void)hint;
  MeshHandle* h = (MeshHandle*)data;
  if (h) {
    unlink_mesh(h);
    if (h->mesh) mbvh_mesh_free(h->mesh);
    free(h);
  }
}

static MeshHandle* get_mesh(napi_env env, napi_value v) {
  void* p = NULL;
  if (napi_get_value_external(env, v, &p) != napi_ok) return NULL;
  MeshHandle* h = (MeshHandle*)p;
  return (h && h->mesh) ? h : NULL;
}

/* ------------------------------------------------------------------------------------------ */
/* build(positions, stride, offset, index|null, ranges, opts) -> { roots, indirect }          */
/* ------------------------------------------------------------------------------------------ */
static napi_value js_build(napi_env env, napi_callback_info info) {
  napi_value argv[6];
  if (!get_args(env, info, 6, argv)) return throw_msg(env, "build: bad arguments");
  View pos, idx, rng;
  if (!get_view(env, argv[0], &pos) || pos.type != napi_float32_array)
    return throw_msg(env, "build: positions must be a Float32Array");
  int32_t stride, offset;
  if (!get_i32(env, argv[1], &stride) || !get_i32(env, argv[2], &offset) || stride <= 0 || offset < 0)
    return throw_msg(env, "build: stride / offset must be integers");
  int has_index = !is_nullish(env, argv[3]);
  if (has_index && (!get_view(env, argv[3], &idx) || (idx.type != napi_uint16_array && idx.type != napi_uint32_array)))
    return throw_msg(env, "build: index must be a Uint16Array or Uint32Array");
  if (!get_view(env, argv[4], &rng) || rng.type != napi_uint32_array || rng.length < 2 || rng.length % 2)
    return throw_msg(env, "build: ranges must be a Uint32Array of (offset, count) pairs");
  mbvh_build_opts opts;
  if (!get_named_i32(env, argv[5], "strategy", MBVH_CENTER, &opts.strategy) ||
      !get_named_i32(env, argv[5], "maxDepth", 40, &opts.max_depth) ||
      !get_named_i32(env, argv[5], "targetLeafSize", 10, &opts.target_leaf_size) ||
      !get_named_i32(env, argv[5], "indirect", 0, &opts.indirect) ||
      !get_named_i32(env, argv[5], "strictLeafSize", 0, &opts.strict_leaf_size))
    return throw_msg(env, "build: bad options");
  int32_t device = -1; /* >= 0: build on that device (mbvh_build_gpu), same result */
  if (!get_named_i32(env, argv[5], "device", -1, &device)) return throw_msg(env, "build: bad options");

  mbvh_tree* tree = NULL;
  int rc;
  if (device >= 0) {
    mbvh_ctx* ctx = NULL;
    rc = context_for(env, device, &ctx);
    if (rc) return throw_status(env, rc, "mbvh_ctx_create");
    rc = mbvh_build_gpu(ctx, (const float*)pos.data, pos.length / (size_t)stride, (uint32_t)stride, (uint32_t)offset,
                        has_index ? idx.data : NULL, has_index ? idx.elt : 0, has_index ? idx.length : 0,
                        (const uint32_t*)rng.data, (uint32_t)(rng.length / 2), &opts, &tree);
    if (rc) return throw_status(env, rc, "mbvh_build_gpu");
  } else {
    rc = mbvh_build((const float*)pos.data, pos.length / (size_t)stride, (uint32_t)stride, (uint32_t)offset,
                    has_index ? idx.data : NULL, has_index ? idx.elt : 0, has_index ? idx.length : 0,
                    (const uint32_t*)rng.data, (uint32_t)(rng.length / 2), &opts, &tree);
    if (rc) return throw_status(env, rc, "mbvh_build");
  }

  uint32_t n_roots = 0;
  int32_t ieb = 0;
  uint64_t icount = 0;
  mbvh_tree_info(tree, &n_roots, &ieb, &icount);
  napi_value result, roots, indirect;
  napi_create_object(env, &result);
  napi_create_array_with_length(env, n_roots, &roots);
  for (uint32_t r = 0; r < n_roots; r++) {
    uint64_t bytes = 0;
    mbvh_tree_root_bytes(tree, r, &bytes);
    void* data = NULL;
    napi_value ab;
    if (napi_create_arraybuffer(env, (size_t)bytes, &data, &ab) != napi_ok) {
      mbvh_tree_free(tree);
      return throw_msg(env, "build: out of memory");
    }
    mbvh_tree_copy_root(tree, r, data, bytes);
    napi_set_element(env, roots, r, ab);
  }
  if (ieb) {
    void* data = NULL;
    napi_value ab;
    napi_create_arraybuffer(env, (size_t)(icount * (uint64_t)ieb), &data, &ab);
    mbvh_tree_copy_indirect(tree, data, icount * (uint64_t)ieb);
    napi_create_typedarray(env, ieb == 2 ? napi_uint16_array : napi_uint32_array, (size_t)icount, ab, 0, &indirect);
  } else {
    napi_get_null(env, &indirect);
  }
  mbvh_tree_free(tree);
  napi_set_named_property(env, result, "roots", roots);
  napi_set_named_property(env, result, "indirect", indirect);
  return result;
}

/* deviceCount() */
static napi_value js_device_count(napi_env env, napi_callback_info info) {
  (void)info;
  napi_value out;
  napi_create_int32(env, mbvh_device_count(), &out);
  return out;
}

/* setQueryLanes(device, pointLanes, rayLanes): lanes per query of the device's context
 * (mbvh_ctx_set_point_lanes / mbvh_ctx_set_ray_lanes); a null or undefined count is left as it is */
static napi_value js_set_query_lanes(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  get_args(env, info, 3, argv);
  int32_t device, lanes;
  if (!get_i32(env, argv[0], &device)) return throw_msg(env, "setQueryLanes: device must be an integer");
  mbvh_ctx* ctx = NULL;
  int rc = context_for(env, device, &ctx);
  if (rc) return throw_status(env, rc, "mbvh_ctx_create");
  if (!is_nullish(env, argv[1])) {
    if (!get_i32(env, argv[1], &lanes)) return throw_msg(env, "setQueryLanes: pointLanes must be an integer");
    if ((rc = mbvh_ctx_set_point_lanes(ctx, lanes))) return throw_status(env, rc, "mbvh_ctx_set_point_lanes");
  }
  if (!is_nullish(env, argv[2])) {
    if (!get_i32(env, argv[2], &lanes)) return throw_msg(env, "setQueryLanes: rayLanes must be an integer");
    if ((rc = mbvh_ctx_set_ray_lanes(ctx, lanes))) return throw_status(env, rc, "mbvh_ctx_set_ray_lanes");
  }
  return NULL;
}

/* uploadMesh(device, roots[], index|null, indirect|null, positions, stride, offset) -> external */
static napi_value js_upload(napi_env env, napi_callback_info info) {
  napi_value argv[7];
  if (!get_args(env, info, 7, argv)) return throw_msg(env, "uploadMesh: bad arguments");
  int32_t device;
  if (!get_i32(env, argv[0], &device)) return throw_msg(env, "uploadMesh: device must be an integer");
  bool is_arr = false;
  napi_is_array(env, argv[1], &is_arr);
  if (!is_arr) return throw_msg(env, "uploadMesh: roots must be an array of ArrayBuffers");
  uint32_t n_roots = 0;
  napi_get_array_length(env, argv[1], &n_roots);
  if (n_roots == 0 || n_roots > 1u << 20) return throw_msg(env, "uploadMesh: roots must not be empty");
  const void** root_ptrs = (const void**)calloc(n_roots, sizeof(void*));
  uint64_t* root_bytes = (uint64_t*)calloc(n_roots, sizeof(uint64_t));
  if (!root_ptrs || !root_bytes) {
    free(root_ptrs);
    free(root_bytes);
    return throw_msg(env, "uploadMesh: out of memory");
  }
  for (uint32_t r = 0; r < n_roots; r++) {
    napi_value e;
    napi_get_element(env, argv[1], r, &e);
    void* data = NULL;
    size_t len = 0;
    bool is_ab = false;
    napi_is_arraybuffer(env, e, &is_ab);
    if (is_ab) {
      napi_get_arraybuffer_info(env, e, &data, &len);
    } else {
      View v;
      if (!get_view(env, e, &v)) {
        free(root_ptrs);
        free(root_bytes);
        return throw_msg(env, "uploadMesh: each root must be an ArrayBuffer");
      }
      data = v.data;
      len = v.length * (size_t)v.elt;
    }
    root_ptrs[r] = data;
    root_bytes[r] = len;
  }
  View idx, ind, pos;
  int has_index = !is_nullish(env, argv[2]);
  int has_ind = !is_nullish(env, argv[3]);
  if ((has_index && !get_view(env, argv[2], &idx)) || (has_ind && !get_view(env, argv[3], &ind)) ||
      !get_view(env, argv[4], &pos) || pos.type != napi_float32_array) {
    free(root_ptrs);
    free(root_bytes);
    return throw_msg(env, "uploadMesh: index / indirect must be integer typed arrays, positions a Float32Array");
  }
  int32_t stride, offset;
  if (!get_i32(env, argv[5], &stride) || !get_i32(env, argv[6], &offset) || stride <= 0 || offset < 0) {
    free(root_ptrs);
    free(root_bytes);
    return throw_msg(env, "uploadMesh: stride / offset must be integers");
  }
  mbvh_ctx* ctx = NULL;
  int rc = context_for(env, device, &ctx);
  if (rc) {
    free(root_ptrs);
    free(root_bytes);
    return throw_status(env, rc, "mbvh_ctx_create");
  }
  mbvh_mesh* mesh = NULL;
  rc = mbvh_mesh_upload(ctx, root_ptrs, root_bytes, n_roots, has_index ? idx.data : NULL, has_index ? idx.elt : 0,
                        has_index ? idx.length : 0, has_ind ? ind.data : NULL, has_ind ? ind.elt : 0,
                        has_ind ? ind.length : 0, (const float*)pos.data, pos.length / (size_t)stride,
                        (uint32_t)stride, (uint32_t)offset, &mesh);
  free(root_ptrs);
  free(root_bytes);
  if (rc) return throw_status(env, rc, "mbvh_mesh_upload");
  MeshHandle* h = (MeshHandle*)calloc(1, sizeof(MeshHandle));
  if (!h) {
    mbvh_mesh_free(mesh);
    return throw_msg(env, "uploadMesh: out of memory");
  }
  h->mesh = mesh;
  h->device = device;
  napi_value ext;
  if (napi_create_external(env, h, finalize_mesh, NULL, &ext) != napi_ok) {
    mbvh_mesh_free(mesh);
    free(h);
    return throw_msg(env, "uploadMesh: cannot wrap the mesh");
  }
  link_mesh(h);
  return ext;
}

/* freeMesh(mesh): releases device memory now instead of at garbage collection */
static napi_value js_free_mesh(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  get_args(env, info, 1, argv);
  void* p = NULL;
  if (napi_get_value_external(env, argv[0], &p) == napi_ok && p) {
    MeshHandle* h = (MeshHandle*)p;
    if (h->batch_refs > 0) return throw_msg(env, "freeMesh: the mesh is used by a batch; free the batch first");
    if (h->async_refs > 0) return throw_msg(env, "freeMesh: an asynchronous query on the mesh is in flight; await it first");
    if (h->mesh) mbvh_mesh_free(h->mesh);
    h->mesh = NULL;
  }
  return NULL;
}

static void release_batch(napi_env env, BatchHandle* b) {
  if (!b) return;
  if (b->batch) mbvh_batch_free(b->batch);
  if (b->scene) mbvh_scene_free(b->scene);
  b->batch = NULL;
  b->scene = NULL;
  for (uint32_t i = 0; i < b->n_meshes; i++) {
    if (b->meshes && b->meshes[i]) b->meshes[i]->batch_refs--;
    if (b->refs && b->refs[i]) napi_delete_reference(env, b->refs[i]);
  }
  free(b->meshes);
  free(b->refs);
  b->meshes = NULL;
  b->refs = NULL;
  b->n_meshes = 0;
}

static void finalize_batch(napi_env env, void* data, void* hint) {
  (void)hint;
  BatchHandle* b = (BatchHandle*)data;
  unlink_batch(b);
  release_batch(env, b);
  free(b);
}

/* A batch / scene handle holding n uploaded meshes of `device` (kept alive and unfreeable meanwhile);
 * raw receives their C handles. Returns an error message, or NULL. */
static const char* hold_meshes(napi_env env, napi_value arr, uint32_t n, int32_t device, BatchHandle** out,
                               mbvh_mesh*** raw_out) {
  BatchHandle* b = (BatchHandle*)calloc(1, sizeof(BatchHandle));
  mbvh_mesh** raw = (mbvh_mesh**)calloc(n, sizeof(mbvh_mesh*));
  if (!b || !raw) {
    free(b);
    free(raw);
    return "out of memory";
  }
  b->device = device;
  b->meshes = (MeshHandle**)calloc(n, sizeof(MeshHandle*));
  b->refs = (napi_ref*)calloc(n, sizeof(napi_ref));
  if (!b->meshes || !b->refs) {
    free(raw);
    release_batch(env, b);
    free(b);
    return "out of memory";
  }
  for (uint32_t i = 0; i < n; i++) {
    napi_value v;
    napi_get_element(env, arr, i, &v);
    MeshHandle* h = get_mesh(env, v);
    if (!h || h->device != device) {
      free(raw);
      release_batch(env, b);
      free(b);
      return "every mesh must be an uploaded mesh of this device";
    }
    b->n_meshes = i + 1;
    b->meshes[i] = h;
    h->batch_refs++;
    napi_create_reference(env, v, 1, &b->refs[i]);
    raw[i] = h->mesh;
  }
  *out = b;
  *raw_out = raw;
  return NULL;
}

/* batchCreate(device, [mesh, ...]) -> batch (external) */
static napi_value js_batch_create(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  get_args(env, info, 2, argv);
  int32_t device;
  if (!get_i32(env, argv[0], &device)) return throw_msg(env, "batchCreate: device must be a number");
  bool is_array = false;
  if (napi_is_array(env, argv[1], &is_array) != napi_ok || !is_array)
    return throw_msg(env, "batchCreate: meshes must be an array");
  uint32_t n = 0;
  napi_get_array_length(env, argv[1], &n);
  if (n == 0) return throw_msg(env, "batchCreate: no meshes");
  BatchHandle* b = NULL;
  mbvh_mesh** raw = NULL;
  const char* err = hold_meshes(env, argv[1], n, device, &b, &raw);
  if (err) return throw_msg(env, err);
  mbvh_ctx* ctx = NULL;
  int rc = context_for(env, device, &ctx);
  if (!rc) rc = mbvh_batch_create(ctx, (const mbvh_mesh* const*)raw, n, &b->batch);
  free(raw);
  if (rc) {
    release_batch(env, b);
    free(b);
    return throw_status(env, rc, "mbvh_batch_create");
  }
  napi_value out;
  if (napi_create_external(env, b, finalize_batch, NULL, &out) != napi_ok) {
    release_batch(env, b);
    free(b);
    return throw_msg(env, "batchCreate: cannot wrap the handle");
  }
  link_batch(b);
  return out;
}

/* freeBatch(batch) */
static napi_value js_free_batch(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  get_args(env, info, 1, argv);
  void* p = NULL;
  if (napi_get_value_external(env, argv[0], &p) == napi_ok && p) release_batch(env, (BatchHandle*)p);
  return NULL;
}

/* batchRaycastFirst(batch, matrixWorld, matrixWorldInverse, instanceGeometry, origins, dirs, rayInstance,
 *                   side, near, far, outFace, outDist, outPoint, outBary) */
static napi_value js_batch_raycast_first(napi_env env, napi_callback_info info) {
  napi_value argv[14];
  get_args(env, info, 14, argv);
  void* p = NULL;
  if (napi_get_value_external(env, argv[0], &p) != napi_ok || !p || !((BatchHandle*)p)->batch)
    return throw_msg(env, "batchRaycastFirst: invalid batch");
  BatchHandle* b = (BatchHandle*)p;
  View mw, mi, ig, o, d, ri;
  if (!get_view(env, argv[1], &mw) || !get_view(env, argv[2], &mi) || !get_view(env, argv[3], &ig) ||
      mw.type != napi_float32_array || mi.type != napi_float32_array || ig.type != napi_uint32_array ||
      mw.length != 16 * ig.length || mi.length != mw.length)
    return throw_msg(env, "batchRaycastFirst: matrices must be Float32Arrays of 16 per instance, instanceGeometry a Uint32Array");
  if (!get_view(env, argv[4], &o) || !get_view(env, argv[5], &d) || !get_view(env, argv[6], &ri) ||
      o.type != napi_float32_array || d.type != napi_float32_array || ri.type != napi_uint32_array ||
      o.length != d.length || o.length % 3 || ri.length * 3 != o.length)
    return throw_msg(env, "batchRaycastFirst: origins / directions (3n floats) and rayInstance (n) do not match");
  const size_t n = ri.length;
  int32_t side;
  double nr, fr;
  if (!get_i32(env, argv[7], &side) || !get_f64(env, argv[8], &nr) || !get_f64(env, argv[9], &fr))
    return throw_msg(env, "batchRaycastFirst: side / near / far must be numbers");
  void *face, *dist, *pt, *bc;
  if (!out_view(env, argv[10], napi_uint32_array, n, &face) || !face ||
      !out_view(env, argv[11], napi_float32_array, n, &dist) || !out_view(env, argv[12], napi_float32_array, 3 * n, &pt) ||
      !out_view(env, argv[13], napi_float32_array, 3 * n, &bc))
    return throw_msg(env, "batchRaycastFirst: output arrays have the wrong type or length");
  mbvh_ctx* ctx = NULL;
  int rc = context_for(env, b->device, &ctx);
  if (!rc)
    rc = mbvh_batch_raycast_first(ctx, b->batch, (const float*)mw.data, (const float*)mi.data, (const uint32_t*)ig.data,
                                  (uint32_t)ig.length, (const float*)o.data, (const float*)d.data,
                                  (const uint32_t*)ri.data, n, side, to_float(nr), to_float(fr), (uint32_t*)face,
                                  (float*)dist, (float*)pt, (float*)bc);
  if (rc) return throw_status(env, rc, "mbvh_batch_raycast_first");
  return NULL;
}

/* batchRaycastFirstLocal(batch, origins, dirs, windows|null, rayMesh, side, outFace, outDist, outPoint, outBary):
 * double-precision rays already in their mesh's frame (Float64Arrays), windows 2 per ray (near, far) */
static napi_value js_batch_raycast_first_local(napi_env env, napi_callback_info info) {
  napi_value argv[10];
  get_args(env, info, 10, argv);
  void* p = NULL;
  if (napi_get_value_external(env, argv[0], &p) != napi_ok || !p || !((BatchHandle*)p)->batch)
    return throw_msg(env, "batchRaycastFirstLocal: invalid batch");
  BatchHandle* b = (BatchHandle*)p;
  View o, d, w, rm;
  if (!get_view(env, argv[1], &o) || !get_view(env, argv[2], &d) || !get_view(env, argv[4], &rm) ||
      o.type != napi_float64_array || d.type != napi_float64_array || rm.type != napi_uint32_array ||
      o.length != d.length || o.length != 3 * rm.length)
    return throw_msg(env, "batchRaycastFirstLocal: origins / directions (Float64Array, 3n) and rayMesh (Uint32Array, n) do not match");
  const size_t n = rm.length;
  const double* win = NULL;
  if (!is_nullish(env, argv[3])) {
    if (!get_view(env, argv[3], &w) || w.type != napi_float64_array || w.length != 2 * n)
      return throw_msg(env, "batchRaycastFirstLocal: windows must be a Float64Array of 2n (near, far)");
    win = (const double*)w.data;
  }
  int32_t side;
  if (!get_i32(env, argv[5], &side)) return throw_msg(env, "batchRaycastFirstLocal: side must be a number");
  void *face, *dist, *pt, *bc;
  if (!out_view(env, argv[6], napi_uint32_array, n, &face) || !face ||
      !out_view(env, argv[7], napi_float32_array, n, &dist) || !out_view(env, argv[8], napi_float32_array, 3 * n, &pt) ||
      !out_view(env, argv[9], napi_float32_array, 3 * n, &bc))
    return throw_msg(env, "batchRaycastFirstLocal: output arrays have the wrong type or length");
  mbvh_ctx* ctx = NULL;
  int rc = context_for(env, b->device, &ctx);
  if (!rc)
    rc = mbvh_batch_raycast_first_local_f64(ctx, b->batch, (const double*)o.data, (const double*)d.data, win,
                                            (const uint32_t*)rm.data, n, side, (uint32_t*)face, (float*)dist,
                                            (float*)pt, (float*)bc);
  if (rc) return throw_status(env, rc, "mbvh_batch_raycast_first_local_f64");
  return NULL;
}

/* sceneCreate(device, [mesh, ...], matrixWorld, matrixWorldInverse, instanceGeometry, visible|null) -> scene */
static napi_value js_scene_create(napi_env env, napi_callback_info info) {
  napi_value argv[6];
  get_args(env, info, 6, argv);
  int32_t device;
  if (!get_i32(env, argv[0], &device)) return throw_msg(env, "sceneCreate: device must be a number");
  bool is_array = false;
  if (napi_is_array(env, argv[1], &is_array) != napi_ok || !is_array)
    return throw_msg(env, "sceneCreate: meshes must be an array");
  uint32_t n = 0;
  napi_get_array_length(env, argv[1], &n);
  if (n == 0) return throw_msg(env, "sceneCreate: no meshes");
  View mw, mi, ig, vis;
  if (!get_view(env, argv[2], &mw) || !get_view(env, argv[3], &mi) || !get_view(env, argv[4], &ig) ||
      mw.type != napi_float32_array || mi.type != napi_float32_array || ig.type != napi_uint32_array ||
      mw.length != 16 * ig.length || mi.length != mw.length)
    return throw_msg(env, "sceneCreate: matrices must be Float32Arrays of 16 per instance, instanceGeometry a Uint32Array");
  const int has_vis = !is_nullish(env, argv[5]);
  if (has_vis && (!get_view(env, argv[5], &vis) || vis.elt != 1 || vis.length != ig.length))
    return throw_msg(env, "sceneCreate: visible must be a Uint8Array of one entry per instance");
  BatchHandle* b = NULL;
  mbvh_mesh** raw = NULL;
  const char* err = hold_meshes(env, argv[1], n, device, &b, &raw);
  if (err) return throw_msg(env, err);
  mbvh_ctx* ctx = NULL;
  int rc = context_for(env, device, &ctx);
  if (!rc)
    rc = mbvh_scene_create(ctx, (const mbvh_mesh* const*)raw, n, (const float*)mw.data, (const float*)mi.data,
                           (const uint32_t*)ig.data, has_vis ? (const uint8_t*)vis.data : NULL, (uint32_t)ig.length,
                           &b->scene);
  free(raw);
  if (rc) {
    release_batch(env, b);
    free(b);
    return throw_status(env, rc, "mbvh_scene_create");
  }
  napi_value out;
  if (napi_create_external(env, b, finalize_batch, NULL, &out) != napi_ok) {
    release_batch(env, b);
    free(b);
    return throw_msg(env, "sceneCreate: cannot wrap the handle");
  }
  link_batch(b);
  return out;
}

/* sceneRaycastFirst(scene, origins, dirs, side, near, far, outFace, outInstance, outDist, outPoint, outBary) */
static napi_value js_scene_raycast_first(napi_env env, napi_callback_info info) {
  napi_value argv[11];
  get_args(env, info, 11, argv);
  void* p = NULL;
  if (napi_get_value_external(env, argv[0], &p) != napi_ok || !p || !((BatchHandle*)p)->scene)
    return throw_msg(env, "sceneRaycastFirst: invalid scene");
  BatchHandle* b = (BatchHandle*)p;
  View o, d;
  if (!get_view(env, argv[1], &o) || !get_view(env, argv[2], &d) || o.type != napi_float32_array ||
      d.type != napi_float32_array || o.length != d.length || o.length % 3)
    return throw_msg(env, "sceneRaycastFirst: origins / directions must be Float32Arrays of 3n floats");
  const size_t n = o.length / 3;
  int32_t side;
  double nr, fr;
  if (!get_i32(env, argv[3], &side) || !get_f64(env, argv[4], &nr) || !get_f64(env, argv[5], &fr))
    return throw_msg(env, "sceneRaycastFirst: side / near / far must be numbers");
  void *face, *inst, *dist, *pt, *bc;
  if (!out_view(env, argv[6], napi_uint32_array, n, &face) || !face ||
      !out_view(env, argv[7], napi_uint32_array, n, &inst) || !out_view(env, argv[8], napi_float32_array, n, &dist) ||
      !out_view(env, argv[9], napi_float32_array, 3 * n, &pt) || !out_view(env, argv[10], napi_float32_array, 3 * n, &bc))
    return throw_msg(env, "sceneRaycastFirst: output arrays have the wrong type or length");
  mbvh_ctx* ctx = NULL;
  int rc = context_for(env, b->device, &ctx);
  if (!rc)
    rc = mbvh_scene_raycast_first(ctx, b->scene, (const float*)o.data, (const float*)d.data, n, side, to_float(nr),
                                  to_float(fr), (uint32_t*)face, (uint32_t*)inst, (float*)dist, (float*)pt, (float*)bc);
  if (rc) return throw_status(env, rc, "mbvh_scene_raycast_first");
  return NULL;
}

/* sceneClosestPoint(scene, points F32 (3n), face U32 (n), instance U32 (n), dist F32 (n), point F32 (3n),
   bary F32 (3n)) */
static napi_value js_scene_closest_point(napi_env env, napi_callback_info info) {
  napi_value argv[7];
  get_args(env, info, 7, argv);
  void* p = NULL;
  if (napi_get_value_external(env, argv[0], &p) != napi_ok || !p || !((BatchHandle*)p)->scene)
    return throw_msg(env, "sceneClosestPoint: invalid scene");
  BatchHandle* b = (BatchHandle*)p;
  View q;
  if (!get_view(env, argv[1], &q) || q.type != napi_float32_array || q.length % 3)
    return throw_msg(env, "sceneClosestPoint: points must be a Float32Array of 3n");
  const size_t n = q.length / 3;
  void *face, *inst, *dist, *pt, *bc;
  if (!out_view(env, argv[2], napi_uint32_array, n, &face) || !out_view(env, argv[3], napi_uint32_array, n, &inst) ||
      !out_view(env, argv[4], napi_float32_array, n, &dist) || !out_view(env, argv[5], napi_float32_array, 3 * n, &pt) ||
      !out_view(env, argv[6], napi_float32_array, 3 * n, &bc))
    return throw_msg(env, "sceneClosestPoint: output arrays have the wrong type or length");
  mbvh_ctx* ctx = NULL;
  int rc = context_for(env, b->device, &ctx);
  if (!rc)
    rc = mbvh_scene_closest_point(ctx, b->scene, (const float*)q.data, n, (uint32_t*)face, (uint32_t*)inst,
                                  (float*)dist, (float*)pt, (float*)bc);
  if (rc) return throw_status(env, rc, "mbvh_scene_closest_point");
  return NULL;
}

/* intersectsBox(mesh, boxMin F64 (3n), boxMax F64 (3n), boxToMesh F64 (16n), out U8 (n)) */
static napi_value js_intersects_box(napi_env env, napi_callback_info info) {
  napi_value argv[5];
  get_args(env, info, 5, argv);
  MeshHandle* h = get_mesh(env, argv[0]);
  if (!h) return throw_msg(env, "intersectsBox: invalid mesh");
  View mn, mx, m, out;
  if (!get_view(env, argv[1], &mn) || !get_view(env, argv[2], &mx) || !get_view(env, argv[3], &m) ||
      !get_view(env, argv[4], &out) || mn.type != napi_float64_array || mx.type != napi_float64_array ||
      m.type != napi_float64_array || out.elt != 1 || mn.length != 3 * out.length || mx.length != mn.length ||
      m.length != 16 * out.length)
    return throw_msg(env, "intersectsBox: boxes must be Float64Arrays (3n, 3n, 16n) and out a Uint8Array (n)");
  mbvh_ctx* ctx = NULL;
  int rc = context_for(env, h->device, &ctx);
  if (!rc)
    rc = mbvh_intersects_box(ctx, h->mesh, (const double*)mn.data, (const double*)mx.data, (const double*)m.data,
                             out.length, (uint8_t*)out.data);
  if (rc) return throw_status(env, rc, "mbvh_intersects_box");
  return NULL;
}

/* closestPointToGeometry(mesh, positions F32, stride, offset, index U16/U32|null, triCount, geometryToBvh F64 (16),
   minThreshold, maxThreshold, out F64 (10): distance, point xyz, otherPoint xyz, otherDistance, faceIndex,
   otherFaceIndex) -> true when a pair was found */
static napi_value js_closest_to_geometry(napi_env env, napi_callback_info info) {
  napi_value argv[10];
  get_args(env, info, 10, argv);
  MeshHandle* h = get_mesh(env, argv[0]);
  if (!h) return throw_msg(env, "closestPointToGeometry: invalid mesh");
  View pos, idx, mat, out;
  int32_t stride = 3, offset = 0;
  double tris = 0, lo = 0, hi = 0;
  const int has_index = !is_nullish(env, argv[4]);
  if (!get_view(env, argv[1], &pos) || pos.type != napi_float32_array || !get_i32(env, argv[2], &stride) ||
      !get_i32(env, argv[3], &offset) || (has_index && (!get_view(env, argv[4], &idx) ||
      (idx.type != napi_uint16_array && idx.type != napi_uint32_array))) || !get_f64(env, argv[5], &tris) ||
      !get_view(env, argv[6], &mat) || mat.type != napi_float64_array || mat.length != 16 ||
      !get_f64(env, argv[7], &lo) || !get_f64(env, argv[8], &hi) || !get_view(env, argv[9], &out) ||
      out.type != napi_float64_array || out.length != 10 || stride < 3 || offset < 0 || tris < 0 ||
      (has_index && 3 * (size_t)tris > idx.length))
    return throw_msg(env, "closestPointToGeometry: bad arguments");
  const uint64_t nv = pos.length >= (size_t)offset + 3 ? (pos.length - (size_t)offset - 3) / (size_t)stride + 1 : 0;
  mbvh_ctx* ctx = NULL;
  mbvh_geometry_hit r;
  int rc = context_for(env, h->device, &ctx);
  if (!rc)
    rc = mbvh_closest_point_to_geometry(ctx, h->mesh, (const float*)pos.data, nv, (uint32_t)stride, (uint32_t)offset,
                                        has_index ? idx.data : NULL, has_index ? (uint32_t)idx.elt : 0, (uint64_t)tris,
                                        (const double*)mat.data, lo, hi, &r);
  if (rc) return throw_status(env, rc, "mbvh_closest_point_to_geometry");
  double* o = (double*)out.data;
  o[0] = r.distance;
  for (int k = 0; k < 3; k++) {
    o[1 + k] = r.point[k];
    o[4 + k] = r.other_point[k];
  }
  o[7] = r.other_distance;
  o[8] = r.face_index;
  o[9] = r.other_face_index;
  napi_value res;
  napi_get_boolean(env, r.face_index != 0xFFFFFFFFu, &res);
  return res;
}

/* textureSizes(mesh) -> [bounds, contents, index, position] texture sides */
static napi_value js_texture_sizes(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  get_args(env, info, 1, argv);
  MeshHandle* h = get_mesh(env, argv[0]);
  if (!h) return throw_msg(env, "textureSizes: invalid mesh");
  uint32_t sides[4];
  int rc = mbvh_mesh_texture_sizes(h->mesh, sides);
  if (rc) return throw_status(env, rc, "mbvh_mesh_texture_sizes");
  napi_value out;
  napi_create_array_with_length(env, 4, &out);
  for (uint32_t i = 0; i < 4; i++) {
    napi_value v;
    napi_create_uint32(env, sides[i], &v);
    napi_set_element(env, out, i, v);
  }
  return out;
}

/* exportTextures(mesh, positions|null, stride, offset, bounds F32, contents U32, index U32|null, position F32|null) */
static napi_value js_export_textures(napi_env env, napi_callback_info info) {
  napi_value argv[8];
  get_args(env, info, 8, argv);
  MeshHandle* h = get_mesh(env, argv[0]);
  if (!h) return throw_msg(env, "exportTextures: invalid mesh");
  uint32_t sides[4];
  int rc = mbvh_mesh_texture_sizes(h->mesh, sides);
  if (rc) return throw_status(env, rc, "mbvh_mesh_texture_sizes");
  View pos;
  const int has_pos = !is_nullish(env, argv[1]);
  if (has_pos && (!get_view(env, argv[1], &pos) || pos.type != napi_float32_array))
    return throw_msg(env, "exportTextures: positions must be a Float32Array");
  int32_t stride = 3, offset = 0;
  if (!get_i32(env, argv[2], &stride) || !get_i32(env, argv[3], &offset) || stride <= 0 || offset < 0)
    return throw_msg(env, "exportTextures: stride / offset must be integers");
  void *bounds, *contents, *index, *position;
  if (!out_view(env, argv[4], napi_float32_array, 4ull * sides[0] * sides[0], &bounds) || !bounds ||
      !out_view(env, argv[5], napi_uint32_array, 2ull * sides[1] * sides[1], &contents) || !contents ||
      !out_view(env, argv[6], napi_uint32_array, 4ull * sides[2] * sides[2], &index) ||
      !out_view(env, argv[7], napi_float32_array, 4ull * sides[3] * sides[3], &position))
    return throw_msg(env, "exportTextures: output arrays have the wrong type or length");
  mbvh_ctx* ctx = NULL;
  rc = context_for(env, h->device, &ctx);
  if (!rc)
    rc = mbvh_mesh_export_textures(ctx, h->mesh, has_pos ? (const float*)pos.data : NULL, (uint32_t)stride,
                                   (uint32_t)offset, (float*)bounds, (uint32_t*)contents, (uint32_t*)index,
                                   (float*)position);
  if (rc) return throw_status(env, rc, "mbvh_mesh_export_textures");
  return NULL;
}

/* exportNodes(mesh, out Uint32Array of 8 words per node) */
static napi_value js_export_nodes(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  get_args(env, info, 2, argv);
  MeshHandle* h = get_mesh(env, argv[0]);
  if (!h) return throw_msg(env, "exportNodes: invalid mesh");
  View out;
  if (!get_view(env, argv[1], &out) || out.type != napi_uint32_array)
    return throw_msg(env, "exportNodes: out must be a Uint32Array");
  mbvh_ctx* ctx = NULL;
  int rc = context_for(env, h->device, &ctx);
  if (!rc) rc = mbvh_mesh_export_nodes(ctx, h->mesh, (uint32_t*)out.data, 4ull * out.length);
  if (rc) return throw_status(env, rc, "mbvh_mesh_export_nodes");
  return NULL;
}

/* setSideMask(mesh, Uint8Array) */
static napi_value js_set_side_mask(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  get_args(env, info, 2, argv);
  MeshHandle* h = get_mesh(env, argv[0]);
  if (!h) return throw_msg(env, "setSideMask: invalid mesh");
  View m;
  if (!get_view(env, argv[1], &m) || m.elt != 1) return throw_msg(env, "setSideMask: mask must be a Uint8Array");
  mbvh_ctx* ctx = NULL;
  int rc = context_for(env, h->device, &ctx);
  if (!rc) rc = mbvh_mesh_set_side_mask(ctx, h->mesh, (const uint8_t*)m.data, m.length);
  if (rc) return throw_status(env, rc, "mbvh_mesh_set_side_mask");
  return NULL;
}

static int out_view(napi_env env, napi_value v, napi_typedarray_type want, size_t n, void** ptr) {
  *ptr = NULL;
  if (is_nullish(env, v)) return 1;
  View w;
  if (!get_view(env, v, &w) || w.type != want || w.length < n) return 0;
  *ptr = w.data;
  return 1;
}

/* Float32Array or Float64Array rays of equal type and length 3n; *f64 tells which. */
static int get_rays(napi_env env, napi_value vo, napi_value vd, View* o, View* d, int* f64) {
  if (!get_view(env, vo, o) || !get_view(env, vd, d) || o->type != d->type || o->length != d->length ||
      o->length % 3)
    return 0;
  *f64 = o->type == napi_float64_array;
  return o->type == napi_float32_array || *f64;
}

/* One first-hit or closest-point query, parsed from its arguments; run on the calling thread
 * (raycastFirst, closestPoint) or on the addon's query thread (raycastFirstAsync, closestPointAsync). */
typedef struct Query {
  int kind;  /* 0: first hit, 1: closest point */
  MeshHandle* h;
  mbvh_ctx* ctx;
  int f64;
  const void* a;  /* origins, or points */
  const void* b;  /* directions */
  size_t n;
  int32_t side;
  double p0, p1;  /* near / far, or min / max threshold */
  void *face, *dist, *pt, *bc;
  int rc;
  char msg[512];  /* mbvh_last_error() of the thread that ran it */
  napi_deferred deferred;
  napi_ref refs[8];  /* the mesh and the typed arrays, alive until an asynchronous query settles */
  int n_refs;
  struct Query* next;  /* the query thread's FIFO */
} Query;

/* raycastFirst(mesh, origins, dirs, side, near, far, face, dist, point, bary): origins / dirs either
 * Float32Arrays (FP32 rays) or Float64Arrays (double rays, traced as the reference traces them) */
static const char* parse_raycast_first(napi_env env, napi_value* argv, Query* q) {
  memset(q, 0, sizeof(*q));
  q->kind = 0;
  q->h = get_mesh(env, argv[0]);
  if (!q->h) return "raycastFirst: invalid mesh";
  View o, d;
  if (!get_rays(env, argv[1], argv[2], &o, &d, &q->f64))
    return "raycastFirst: origins / directions must be Float32Arrays or Float64Arrays of equal length 3n";
  q->a = o.data;
  q->b = d.data;
  q->n = o.length / 3;
  if (!get_i32(env, argv[3], &q->side) || !get_f64(env, argv[4], &q->p0) || !get_f64(env, argv[5], &q->p1))
    return "raycastFirst: side / near / far must be numbers";
  if (!out_view(env, argv[6], napi_uint32_array, q->n, &q->face) || !q->face ||
      !out_view(env, argv[7], napi_float32_array, q->n, &q->dist) ||
      !out_view(env, argv[8], napi_float32_array, 3 * q->n, &q->pt) ||
      !out_view(env, argv[9], napi_float32_array, 3 * q->n, &q->bc))
    return "raycastFirst: output arrays have the wrong type or length";
  return NULL;
}

/* closestPoint(mesh, points, minThreshold, maxThreshold, face, dist, point, bary): points a
 * Float32Array or a Float64Array (double points) */
static const char* parse_closest_point(napi_env env, napi_value* argv, Query* q) {
  memset(q, 0, sizeof(*q));
  q->kind = 1;
  q->h = get_mesh(env, argv[0]);
  if (!q->h) return "closestPoint: invalid mesh";
  View p;
  if (!get_view(env, argv[1], &p) || (p.type != napi_float32_array && p.type != napi_float64_array) || p.length % 3)
    return "closestPoint: points must be a Float32Array or Float64Array of length 3n";
  q->f64 = p.type == napi_float64_array;
  q->a = p.data;
  q->n = p.length / 3;
  if (!get_f64(env, argv[2], &q->p0) || !get_f64(env, argv[3], &q->p1)) return "closestPoint: thresholds must be numbers";
  if (!out_view(env, argv[4], napi_uint32_array, q->n, &q->face) || !q->face ||
      !out_view(env, argv[5], napi_float32_array, q->n, &q->dist) ||
      !out_view(env, argv[6], napi_float32_array, 3 * q->n, &q->pt) ||
      !out_view(env, argv[7], napi_float32_array, 3 * q->n, &q->bc))
    return "closestPoint: output arrays have the wrong type or length";
  return NULL;
}

static const char* query_where(const Query* q) {
  if (q->kind == 0) return q->f64 ? "mbvh_raycast_first_f64" : "mbvh_raycast_first";
  return q->f64 ? "mbvh_closest_point_f64" : "mbvh_closest_point";
}

/* Runs the query through the C ABI (no JavaScript values touched: safe on the query thread). */
static void run_query(Query* q) {
  if (q->kind == 0 && q->f64)
    q->rc = mbvh_raycast_first_f64(q->ctx, q->h->mesh, (const double*)q->a, (const double*)q->b, q->n, q->side, q->p0,
                                   q->p1, (uint32_t*)q->face, (float*)q->dist, (float*)q->pt, (float*)q->bc);
  else if (q->kind == 0)
    q->rc = mbvh_raycast_first(q->ctx, q->h->mesh, (const float*)q->a, (const float*)q->b, q->n, q->side,
                               to_float(q->p0), to_float(q->p1), (uint32_t*)q->face, (float*)q->dist, (float*)q->pt,
                               (float*)q->bc);
  else if (q->f64)
    q->rc = mbvh_closest_point_f64(q->ctx, q->h->mesh, (const double*)q->a, q->n, q->p0, q->p1, (uint32_t*)q->face,
                                   (float*)q->dist, (float*)q->pt, (float*)q->bc);
  else
    q->rc = mbvh_closest_point(q->ctx, q->h->mesh, (const float*)q->a, q->n, to_float(q->p0), to_float(q->p1),
                               (uint32_t*)q->face, (float*)q->dist, (float*)q->pt, (float*)q->bc);
  if (q->rc) snprintf(q->msg, sizeof(q->msg), "%s", mbvh_last_error());
}

static napi_value run_sync(napi_env env, napi_callback_info info, size_t argc,
                           const char* (*parse)(napi_env, napi_value*, Query*)) {
  napi_value argv[10];
  get_args(env, info, argc, argv);
  Query q;
  const char* bad = parse(env, argv, &q);
  if (bad) return throw_msg(env, bad);
  int rc = context_for(env, q.h->device, &q.ctx);
  if (rc) return throw_status(env, rc, query_where(&q));
  run_query(&q);
  if (q.rc) return throw_status(env, q.rc, query_where(&q));
  return NULL;
}

static napi_value js_raycast_first(napi_env env, napi_callback_info info) {
  return run_sync(env, info, 10, parse_raycast_first);
}

static napi_value js_closest_point(napi_env env, napi_callback_info info) {
  return run_sync(env, info, 8, parse_closest_point);
}

/* ---- asynchronous queries: ONE query thread owned by the addon runs the C calls in submission
 * order (calls on one device context are serialised by the library anyway), and a thread-safe
 * function settles each Promise on the main thread. Not libuv's worker pool: libuv joins its pool
 * threads at process exit after the shared libraries' destructors have run, so the per-thread state
 * the HIP runtime keeps for a thread that called it would be torn down after the runtime itself. The
 * environment cleanup hook stops and joins the query thread while the runtime is still alive. ---- */
typedef struct {
  pthread_mutex_t mu;
  pthread_cond_t cv;
  Query *head, *tail;  /* submitted, not yet run */
  int stop;
  int started;
  pthread_t thread;
  napi_threadsafe_function done;  /* settles a finished query's Promise */
  uint32_t in_flight;             /* main thread: submitted and not yet settled (keeps the loop alive) */
} QueryThread;

static QueryThread g_qt = {PTHREAD_MUTEX_INITIALIZER, PTHREAD_COND_INITIALIZER, NULL, NULL, 0, 0, 0, NULL, 0};

static void* query_main(void* arg) {
  (void)arg;
  for (;;) {
    pthread_mutex_lock(&g_qt.mu);
    while (!g_qt.head && !g_qt.stop) pthread_cond_wait(&g_qt.cv, &g_qt.mu);
    Query* q = NULL;
    if (!g_qt.stop) {
      q = g_qt.head;
      g_qt.head = q->next;
      if (!g_qt.head) g_qt.tail = NULL;
    }
    pthread_mutex_unlock(&g_qt.mu);
    if (!q) return NULL;  /* stopping: queries not yet run are dropped with the environment */
    run_query(q);
    napi_call_threadsafe_function(g_qt.done, q, napi_tsfn_blocking);
  }
}

/* Main thread: a finished query's Promise (env NULL: the environment is going away). */
static void settle_query(napi_env env, napi_value js_cb, void* context, void* data) {
  (void)js_cb;
  (void)context;
  Query* q = (Query*)data;
  if (!env) {
    free(q);
    return;
  }
  for (int i = 0; i < q->n_refs; i++) napi_delete_reference(env, q->refs[i]);
  q->h->async_refs--;
  if (q->rc) {
    char msg[768], code[32];
    snprintf(msg, sizeof(msg), "%s failed (mbvh status %d): %s", query_where(q), q->rc, q->msg);
    snprintf(code, sizeof(code), "MBVH_%d", -q->rc);
    napi_value m, c, err;
    napi_create_string_utf8(env, msg, NAPI_AUTO_LENGTH, &m);
    napi_create_string_utf8(env, code, NAPI_AUTO_LENGTH, &c);
    napi_create_error(env, c, m, &err);
    napi_reject_deferred(env, q->deferred, err);
  } else {
    napi_value undef;
    napi_get_undefined(env, &undef);
    napi_resolve_deferred(env, q->deferred, undef);
  }
  free(q);
  if (g_qt.in_flight > 0 && --g_qt.in_flight == 0) napi_unref_threadsafe_function(env, g_qt.done);
}

static napi_value js_noop(napi_env env, napi_callback_info info) {
  (void)env;
  (void)info;
  return NULL;
}

/* Main thread: starts the query thread on first use. */
static int start_query_thread(napi_env env) {
  if (g_qt.started) return 1;
  napi_value name, fn;
  if (napi_create_string_utf8(env, "mbvh.query", NAPI_AUTO_LENGTH, &name) != napi_ok ||
      napi_create_function(env, "mbvhQueryDone", NAPI_AUTO_LENGTH, js_noop, NULL, &fn) != napi_ok ||
      napi_create_threadsafe_function(env, fn, NULL, name, 0, 1, NULL, NULL, NULL, settle_query, &g_qt.done) != napi_ok)
    return 0;
  napi_unref_threadsafe_function(env, g_qt.done);  /* referenced only while queries are in flight */
  g_qt.stop = 0;
  if (pthread_create(&g_qt.thread, NULL, query_main, NULL) != 0) {
    napi_release_threadsafe_function(g_qt.done, napi_tsfn_abort);
    g_qt.done = NULL;
    return 0;
  }
  g_qt.started = 1;
  return 1;
}

/* Environment cleanup: stop and join the query thread (before any library object is freed). */
static void stop_query_thread(void) {
  if (!g_qt.started) return;
  pthread_mutex_lock(&g_qt.mu);
  g_qt.stop = 1;
  pthread_cond_signal(&g_qt.cv);
  pthread_mutex_unlock(&g_qt.mu);
  pthread_join(g_qt.thread, NULL);
  g_qt.started = 0;
}

static napi_value run_async(napi_env env, napi_callback_info info, size_t argc,
                            const char* (*parse)(napi_env, napi_value*, Query*), const char* name) {
  (void)name;
  napi_value argv[10];
  get_args(env, info, argc, argv);
  Query* q = (Query*)malloc(sizeof(Query));
  if (!q) return throw_msg(env, "out of memory");
  const char* bad = parse(env, argv, q);
  if (bad) {
    free(q);
    return throw_msg(env, bad);
  }
  int rc = context_for(env, q->h->device, &q->ctx);  /* created here, on the main thread */
  if (rc) {
    const char* where = query_where(q);
    free(q);
    return throw_status(env, rc, where);
  }
  if (!start_query_thread(env)) {
    free(q);
    return throw_msg(env, "could not start the query thread");
  }
  /* the mesh external and every typed-array argument stay referenced until the promise settles */
  for (size_t i = 0; i < argc; i++) {
    napi_valuetype t;
    napi_typeof(env, argv[i], &t);
    if ((t == napi_object || t == napi_external) && q->n_refs < 8) napi_create_reference(env, argv[i], 1, &q->refs[q->n_refs++]);
  }
  napi_value promise;
  napi_create_promise(env, &q->deferred, &promise);
  q->h->async_refs++;
  if (g_qt.in_flight++ == 0) napi_ref_threadsafe_function(env, g_qt.done);
  q->next = NULL;
  pthread_mutex_lock(&g_qt.mu);
  if (g_qt.tail)
    g_qt.tail->next = q;
  else
    g_qt.head = q;
  g_qt.tail = q;
  pthread_cond_signal(&g_qt.cv);
  pthread_mutex_unlock(&g_qt.mu);
  return promise;
}

/* raycastFirstAsync(same arguments as raycastFirst) -> Promise (resolved once the outputs are written) */
static napi_value js_raycast_first_async(napi_env env, napi_callback_info info) {
  return run_async(env, info, 10, parse_raycast_first, "mbvh.raycastFirstAsync");
}

/* closestPointAsync(same arguments as closestPoint) -> Promise */
static napi_value js_closest_point_async(napi_env env, napi_callback_info info) {
  return run_async(env, info, 8, parse_closest_point, "mbvh.closestPointAsync");
}

/* raycast(mesh, origins, dirs, side, near, far) -> { offsets, faceIndex, distance, point, barycoord } */
static napi_value js_raycast(napi_env env, napi_callback_info info) {
  napi_value argv[6];
  get_args(env, info, 6, argv);
  MeshHandle* h = get_mesh(env, argv[0]);
  if (!h) return throw_msg(env, "raycast: invalid mesh");
  View o, d;
  int f64 = 0;
  if (!get_rays(env, argv[1], argv[2], &o, &d, &f64))
    return throw_msg(env, "raycast: origins / directions must be Float32Arrays or Float64Arrays of equal length 3n");
  const size_t n = o.length / 3;
  int32_t side;
  double nr, fr;
  if (!get_i32(env, argv[3], &side) || !get_f64(env, argv[4], &nr) || !get_f64(env, argv[5], &fr))
    return throw_msg(env, "raycast: side / near / far must be numbers");
  mbvh_ctx* ctx = NULL;
  int rc = context_for(env, h->device, &ctx);
  uint64_t* offs = (uint64_t*)calloc(n + 1, sizeof(uint64_t));
  if (!offs) return throw_msg(env, "raycast: out of memory");
  if (!rc && f64)
    rc = mbvh_raycast_f64(ctx, h->mesh, (const double*)o.data, (const double*)d.data, n, side, nr, fr, offs);
  else if (!rc)
    rc = mbvh_raycast(ctx, h->mesh, (const float*)o.data, (const float*)d.data, n, side, to_float(nr), to_float(fr), offs);
  if (rc) {
    free(offs);
    return throw_status(env, rc, "mbvh_raycast");
  }
  const size_t total = (size_t)offs[n];
  napi_value result, offsets, face, dist, pt, bc, ab;
  void *p_off, *p_face, *p_dist, *p_pt, *p_bc;
  napi_create_arraybuffer(env, (n + 1) * sizeof(double), &p_off, &ab);
  napi_create_typedarray(env, napi_float64_array, n + 1, ab, 0, &offsets);
  for (size_t i = 0; i <= n; i++) ((double*)p_off)[i] = (double)offs[i];
  free(offs);
  napi_create_arraybuffer(env, total * 4, &p_face, &ab);
  napi_create_typedarray(env, napi_uint32_array, total, ab, 0, &face);
  napi_create_arraybuffer(env, total * 4, &p_dist, &ab);
  napi_create_typedarray(env, napi_float32_array, total, ab, 0, &dist);
  napi_create_arraybuffer(env, total * 12, &p_pt, &ab);
  napi_create_typedarray(env, napi_float32_array, 3 * total, ab, 0, &pt);
  napi_create_arraybuffer(env, total * 12, &p_bc, &ab);
  napi_create_typedarray(env, napi_float32_array, 3 * total, ab, 0, &bc);
  rc = mbvh_raycast_fetch(ctx, (uint32_t*)p_face, (float*)p_dist, (float*)p_pt, (float*)p_bc, total);
  if (rc) return throw_status(env, rc, "mbvh_raycast_fetch");
  napi_create_object(env, &result);
  napi_set_named_property(env, result, "offsets", offsets);
  napi_set_named_property(env, result, "faceIndex", face);
  napi_set_named_property(env, result, "distance", dist);
  napi_set_named_property(env, result, "point", pt);
  napi_set_named_property(env, result, "barycoord", bc);
  return result;
}

/* meshInfo(mesh) -> { triCount, vertexCount, maxDepth } */
static napi_value js_mesh_info(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  get_args(env, info, 1, argv);
  MeshHandle* h = get_mesh(env, argv[0]);
  if (!h) return throw_msg(env, "meshInfo: invalid mesh");
  uint64_t t = 0, v = 0;
  uint32_t d = 0;
  int rc = mbvh_mesh_info(h->mesh, &t, &v, &d);
  if (rc) return throw_status(env, rc, "mbvh_mesh_info");
  napi_value out, a, b, c;
  napi_create_object(env, &out);
  napi_create_double(env, (double)t, &a);
  napi_create_double(env, (double)v, &b);
  napi_create_uint32(env, d, &c);
  napi_set_named_property(env, out, "triCount", a);
  napi_set_named_property(env, out, "vertexCount", b);
  napi_set_named_property(env, out, "maxDepth", c);
  return out;
}

/* raycastFirstBackward(mesh, dirs, face, dist, bary, gDist|null, gPoint|null, gBary|null,
 *                      outGradOrigins|null, outGradDirs|null, outGradPositions|null, mode) */
static napi_value js_raycast_first_bwd(napi_env env, napi_callback_info info) {
  napi_value argv[12];
  get_args(env, info, 12, argv);
  MeshHandle* h = get_mesh(env, argv[0]);
  if (!h) return throw_msg(env, "raycastFirstBackward: invalid mesh");
  View d;
  if (!get_view(env, argv[1], &d) || d.type != napi_float32_array || d.length % 3)
    return throw_msg(env, "raycastFirstBackward: directions must be a Float32Array of length 3n");
  const size_t n = d.length / 3;
  uint64_t t = 0, v = 0;
  uint32_t depth = 0;
  int rc = mbvh_mesh_info(h->mesh, &t, &v, &depth);
  if (rc) return throw_status(env, rc, "mbvh_mesh_info");
  void *face, *dist, *bary, *gd, *gp, *gb, *oo, *od, *ov;
  if (!out_view(env, argv[2], napi_uint32_array, n, &face) || !face ||
      !out_view(env, argv[3], napi_float32_array, n, &dist) || !dist ||
      !out_view(env, argv[4], napi_float32_array, 3 * n, &bary) || !bary)
    return throw_msg(env, "raycastFirstBackward: saved faceIndex / distance / barycoord have the wrong type or length");
  if (!out_view(env, argv[5], napi_float32_array, n, &gd) || !out_view(env, argv[6], napi_float32_array, 3 * n, &gp) ||
      !out_view(env, argv[7], napi_float32_array, 3 * n, &gb))
    return throw_msg(env, "raycastFirstBackward: upstream gradients have the wrong type or length");
  if (!out_view(env, argv[8], napi_float32_array, 3 * n, &oo) || !out_view(env, argv[9], napi_float32_array, 3 * n, &od) ||
      !out_view(env, argv[10], napi_float32_array, 3 * (size_t)v, &ov))
    return throw_msg(env, "raycastFirstBackward: output gradients have the wrong type or length");
  int32_t mode;
  if (!get_i32(env, argv[11], &mode)) return throw_msg(env, "raycastFirstBackward: mode must be a number");
  mbvh_ctx* ctx = NULL;
  rc = context_for(env, h->device, &ctx);
  if (!rc)
    rc = mbvh_raycast_first_bwd(ctx, h->mesh, (const float*)d.data, n, (const uint32_t*)face, (const float*)dist,
                                (const float*)bary, (const float*)gd, (const float*)gp, (const float*)gb, (float*)oo,
                                (float*)od, (float*)ov, mode);
  if (rc) return throw_status(env, rc, "mbvh_raycast_first_bwd");
  return NULL;
}

/* closestPointBackward(mesh, points, face, dist, closest, bary, gDist, outGradPoints|null,
 *                      outGradPositions|null, mode) */
static napi_value js_closest_point_bwd(napi_env env, napi_callback_info info) {
  napi_value argv[10];
  get_args(env, info, 10, argv);
  MeshHandle* h = get_mesh(env, argv[0]);
  if (!h) return throw_msg(env, "closestPointBackward: invalid mesh");
  View q;
  if (!get_view(env, argv[1], &q) || q.type != napi_float32_array || q.length % 3)
    return throw_msg(env, "closestPointBackward: points must be a Float32Array of length 3n");
  const size_t n = q.length / 3;
  uint64_t t = 0, v = 0;
  uint32_t depth = 0;
  int rc = mbvh_mesh_info(h->mesh, &t, &v, &depth);
  if (rc) return throw_status(env, rc, "mbvh_mesh_info");
  void *face, *dist, *cp, *bary, *gd, *oq, *ov;
  if (!out_view(env, argv[2], napi_uint32_array, n, &face) || !face ||
      !out_view(env, argv[3], napi_float32_array, n, &dist) || !dist ||
      !out_view(env, argv[4], napi_float32_array, 3 * n, &cp) || !cp ||
      !out_view(env, argv[5], napi_float32_array, 3 * n, &bary) || !bary)
    return throw_msg(env, "closestPointBackward: saved faceIndex / distance / point / barycoord have the wrong type or length");
  if (!out_view(env, argv[6], napi_float32_array, n, &gd) || !gd)
    return throw_msg(env, "closestPointBackward: the upstream distance gradient must be a Float32Array of length n");
  if (!out_view(env, argv[7], napi_float32_array, 3 * n, &oq) ||
      !out_view(env, argv[8], napi_float32_array, 3 * (size_t)v, &ov))
    return throw_msg(env, "closestPointBackward: output gradients have the wrong type or length");
  int32_t mode;
  if (!get_i32(env, argv[9], &mode)) return throw_msg(env, "closestPointBackward: mode must be a number");
  mbvh_ctx* ctx = NULL;
  rc = context_for(env, h->device, &ctx);
  if (!rc)
    rc = mbvh_closest_point_bwd(ctx, h->mesh, (const float*)q.data, n, (const uint32_t*)face, (const float*)dist,
                                (const float*)cp, (const float*)bary, (const float*)gd, (float*)oq, (float*)ov, mode);
  if (rc) return throw_status(env, rc, "mbvh_closest_point_bwd");
  return NULL;
}

static napi_value js_abi_version(napi_env env, napi_callback_info info) {
  (void)info;
  napi_value out;
  napi_create_uint32(env, mbvh_abi_version(), &out);
  return out;
}

#define EXPORT(name, fn)                                                        \
  do {                                                                          \
    napi_value f;                                                               \
    napi_create_function(env, name, NAPI_AUTO_LENGTH, fn, NULL, &f);            \
    napi_set_named_property(env, exports, name, f);                             \
  } while (0)

/* intersectsSphere(mesh, centers Float32Array(3n), radii Float32Array(n), out Uint8Array(n)) */
static napi_value js_intersects_sphere(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  get_args(env, info, 4, argv);
  MeshHandle* h = get_mesh(env, argv[0]);
  if (!h) return throw_msg(env, "intersectsSphere: invalid mesh");
  View c, r;
  if (!get_view(env, argv[1], &c) || c.type != napi_float32_array || c.length % 3 || !get_view(env, argv[2], &r) ||
      r.type != napi_float32_array || r.length != c.length / 3)
    return throw_msg(env, "intersectsSphere: centers must be a Float32Array of length 3n, radii one of length n");
  const size_t n = r.length;
  void* out;
  if (!out_view(env, argv[3], napi_uint8_array, n, &out) || !out)
    return throw_msg(env, "intersectsSphere: out must be a Uint8Array of length n");
  mbvh_ctx* ctx = NULL;
  int rc = context_for(env, h->device, &ctx);
  if (!rc) rc = mbvh_intersects_sphere(ctx, h->mesh, (const float*)c.data, (const float*)r.data, n, (uint8_t*)out);
  if (rc) return throw_status(env, rc, "mbvh_intersects_sphere");
  return NULL;
}

/* sceneExportWebGPU(scene, primitiveLimit = Infinity) -> { nodes, transforms, index } (Uint32Arrays): the
 * scene in the reference's WebGPU packing (mbvh_scene_export_webgpu); roots are cut into clusters of
 * fewer than primitiveLimit triangles, Infinity keeps them whole */
static napi_value new_u32(napi_env env, uint64_t n, uint32_t** data) {
  napi_value ab, out;
  void* p = NULL;
  if (napi_create_arraybuffer(env, (size_t)n * 4, &p, &ab) != napi_ok) return NULL;
  if (napi_create_typedarray(env, napi_uint32_array, (size_t)n, ab, 0, &out) != napi_ok) return NULL;
  *data = (uint32_t*)p;
  return out;
}

static napi_value js_scene_export_webgpu(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  get_args(env, info, 2, argv);
  uint32_t limit = 0; /* 0: whole roots */
  if (!is_nullish(env, argv[1])) {
    double v;
    if (!get_f64(env, argv[1], &v)) return throw_msg(env, "sceneExportWebGPU: primitiveLimit must be a number");
    /* a node becomes a cluster when it spans fewer than v triangles: count < v <=> count < ceil(v) */
    if (v < 4294967296.0) limit = v < 1.0 ? 1u : (uint32_t)ceil(v);
  }
  void* p = NULL;
  if (napi_get_value_external(env, argv[0], &p) != napi_ok || !p || !((BatchHandle*)p)->scene)
    return throw_msg(env, "sceneExportWebGPU: invalid scene");
  BatchHandle* b = (BatchHandle*)p;
  mbvh_ctx* ctx = NULL;
  uint64_t nn = 0, nt = 0, ni = 0;
  int rc = context_for(env, b->device, &ctx);
  if (!rc) rc = mbvh_scene_webgpu_sizes(ctx, b->scene, limit, &nn, &nt, &ni);
  if (rc) return throw_status(env, rc, "mbvh_scene_webgpu_sizes");
  uint32_t *pn = NULL, *pt = NULL, *pi = NULL;
  napi_value vn = new_u32(env, 8 * nn, &pn), vt = new_u32(env, 36 * nt, &pt), vi = new_u32(env, ni, &pi);
  if (!vn || !vt || !vi) return throw_msg(env, "sceneExportWebGPU: out of memory");
  rc = mbvh_scene_export_webgpu(ctx, b->scene, limit, pn, 8 * nn, pt, 36 * nt, pi, ni);
  if (rc) return throw_status(env, rc, "mbvh_scene_export_webgpu");
  napi_value out;
  napi_create_object(env, &out);
  napi_set_named_property(env, out, "nodes", vn);
  napi_set_named_property(env, out, "transforms", vt);
  napi_set_named_property(env, out, "index", vi);
  return out;
}

/* refitMesh(mesh, positions, stride, offset, roots[], nodeIds = null): device refit from new positions;
 * the refit reference-format nodes are written into the roots' ArrayBuffers. nodeIds (a Uint32Array
 * of node indices over all roots) makes it a partial refit. */
static napi_value js_refit(napi_env env, napi_callback_info info) {
  napi_value argv[6];
  if (!get_args(env, info, 6, argv)) return throw_msg(env, "refitMesh: bad arguments");
  View ids;
  memset(&ids, 0, sizeof(ids));
  const int partial = !is_nullish(env, argv[5]);
  if (partial && (!get_view(env, argv[5], &ids) || ids.type != napi_uint32_array))
    return throw_msg(env, "refitMesh: nodeIds must be a Uint32Array or null");
  MeshHandle* h = get_mesh(env, argv[0]);
  if (!h) return throw_msg(env, "refitMesh: invalid mesh");
  View pos;
  if (!get_view(env, argv[1], &pos) || pos.type != napi_float32_array)
    return throw_msg(env, "refitMesh: positions must be a Float32Array");
  int32_t stride, offset;
  if (!get_i32(env, argv[2], &stride) || !get_i32(env, argv[3], &offset) || stride <= 0 || offset < 0)
    return throw_msg(env, "refitMesh: stride / offset must be integers");
  bool is_arr = false;
  napi_is_array(env, argv[4], &is_arr);
  if (!is_arr) return throw_msg(env, "refitMesh: roots must be an array of ArrayBuffers");
  uint32_t n_roots = 0;
  napi_get_array_length(env, argv[4], &n_roots);
  if (n_roots == 0 || n_roots > 1u << 20) return throw_msg(env, "refitMesh: roots must not be empty");
  void** root_ptrs = (void**)calloc(n_roots, sizeof(void*));
  uint64_t* root_bytes = (uint64_t*)calloc(n_roots, sizeof(uint64_t));
  if (!root_ptrs || !root_bytes) {
    free(root_ptrs);
    free(root_bytes);
    return throw_msg(env, "refitMesh: out of memory");
  }
  for (uint32_t r = 0; r < n_roots; r++) {
    napi_value e;
    napi_get_element(env, argv[4], r, &e);
    bool is_ab = false;
    napi_is_arraybuffer(env, e, &is_ab);
    size_t len = 0;
    if (!is_ab || napi_get_arraybuffer_info(env, e, &root_ptrs[r], &len) != napi_ok) {
      free(root_ptrs);
      free(root_bytes);
      return throw_msg(env, "refitMesh: each root must be an ArrayBuffer");
    }
    root_bytes[r] = len;
  }
  mbvh_ctx* ctx = NULL;
  int rc = context_for(env, h->device, &ctx);
  static const uint32_t no_ids = 0;
  if (!rc && !partial)
    rc = mbvh_mesh_refit(ctx, h->mesh, (const float*)pos.data, pos.length / (size_t)stride, (uint32_t)stride,
                         (uint32_t)offset, root_ptrs, root_bytes, n_roots);
  else if (!rc)
    rc = mbvh_mesh_refit_nodes(ctx, h->mesh, (const float*)pos.data, pos.length / (size_t)stride, (uint32_t)stride,
                               (uint32_t)offset, ids.length ? (const uint32_t*)ids.data : &no_ids, ids.length,
                               root_ptrs, root_bytes, n_roots);
  free(root_ptrs);
  free(root_bytes);
  if (rc) return throw_status(env, rc, partial ? "mbvh_mesh_refit_nodes" : "mbvh_mesh_refit");
  return NULL;
}

/* The environment the process-wide state (contexts, handle lists, query thread) belongs to. */
static napi_env g_env = NULL;

static napi_value init(napi_env env, napi_value exports) {
  if (g_env && g_env != env) {
    napi_throw_error(env, NULL,
                     "three-mesh-bvh-amd: the addon serves one Node environment per process (its device contexts, "
                     "handles and query thread are process-wide); load it from one thread only");
    return NULL;
  }
  g_env = env;
  EXPORT("abiVersion", js_abi_version);
  EXPORT("deviceCount", js_device_count);
  EXPORT("setQueryLanes", js_set_query_lanes);
  EXPORT("build", js_build);
  EXPORT("uploadMesh", js_upload);
  EXPORT("freeMesh", js_free_mesh);
  EXPORT("meshInfo", js_mesh_info);
  EXPORT("setSideMask", js_set_side_mask);
  EXPORT("raycastFirst", js_raycast_first);
  EXPORT("closestPoint", js_closest_point);
  EXPORT("raycastFirstAsync", js_raycast_first_async);
  EXPORT("closestPointAsync", js_closest_point_async);
  EXPORT("raycast", js_raycast);
  EXPORT("raycastFirstBackward", js_raycast_first_bwd);
  EXPORT("batchCreate", js_batch_create);
  EXPORT("freeBatch", js_free_batch);
  EXPORT("batchRaycastFirst", js_batch_raycast_first);
  EXPORT("batchRaycastFirstLocal", js_batch_raycast_first_local);
  EXPORT("sceneCreate", js_scene_create);
  EXPORT("textureSizes", js_texture_sizes);
  EXPORT("intersectsBox", js_intersects_box);
  EXPORT("sceneClosestPoint", js_scene_closest_point);
  EXPORT("closestPointToGeometry", js_closest_to_geometry);
  EXPORT("exportTextures", js_export_textures);
  EXPORT("exportNodes", js_export_nodes);
  EXPORT("freeScene", js_free_batch);
  EXPORT("sceneRaycastFirst", js_scene_raycast_first);
  EXPORT("sceneExportWebGPU", js_scene_export_webgpu);
  EXPORT("closestPointBackward", js_closest_point_bwd);
  EXPORT("refitMesh", js_refit);
  EXPORT("intersectsSphere", js_intersects_sphere);
  napi_add_env_cleanup_hook(env, cleanup_contexts, NULL);
  return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, init)
