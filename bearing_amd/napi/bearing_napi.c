/* N-API addon: the whole-world C ABI (bm_world_*) for a JavaScript host.
 *
 * It binds what the reference's World does per step (src/world.js:314 constructor,
 * :563 addBody, :880 step, :547 transforms) and nothing more: a world is an external
 * value, bodies arrive as addBody has resolved them (typed-array columns named as
 * bm_bodies names them), and each call's status becomes a thrown Error worded
 * "caller: ..." as src/core/check.js:15-22 words the reference's. No host pointer is
 * kept past a call. Built by gcc against Node's own node_api.h (N-API 8: Node 12.22
 * and later), linked to libbearing_mi355x.so beside it. */
#define NAPI_VERSION 8
#include <node_api.h>
#include <stdbool.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "bearing_mi355x.h"

typedef struct holder {
  bm_world* world;
} holder;

#define NAPI(env, call)                                                      \
  do {                                                                       \
    if ((call) != napi_ok) {                                                 \
      napi_throw_error((env), NULL, "bearing: N-API call failed: " #call); \
      return NULL;                                                           \
    }                                                                        \
  } while (0)

static void release(napi_env env, void* data, void* hint) {
  (void)env;
  (void)hint;
  holder* h = (holder*)data;
  if (h->world) bm_world_destroy(h->world);
  free(h);
}

static napi_value fail(napi_env env, const char* caller, const char* message) {
  char text[512];
  snprintf(text, sizeof text, "%s: %s", caller, message);
  napi_throw_error(env, NULL, text);
  return NULL;
}

/* The holder behind argument 0, or NULL with an Error thrown. */
static holder* world_of(napi_env env, napi_value v, const char* caller) {
  napi_valuetype t;
  void* p = NULL;
  if (napi_typeof(env, v, &t) != napi_ok || t != napi_external || napi_get_value_external(env, v, &p) != napi_ok) {
    fail(env, caller, "not a world");
    return NULL;
  }
  holder* h = (holder*)p;
  if (!h->world) {
    fail(env, caller, "the world was destroyed");
    return NULL;
  }
  return h;
}

/* Column `name` of obj: a typed array of type `want` holding at least `count` entries,
 * or absent (undefined / null: *out = NULL). Returns false with an Error thrown. */
static bool column(napi_env env, napi_value obj, const char* name, napi_typedarray_type want, size_t count,
                   const void** out) {
  napi_value v;
  napi_valuetype t;
  bool is = false;
  *out = NULL;
  if (napi_get_named_property(env, obj, name, &v) != napi_ok || napi_typeof(env, v, &t) != napi_ok) {
    fail(env, "addBody", "bodies must be an object of typed arrays");
    return false;
  }
  if (t == napi_undefined || t == napi_null) return true;
  napi_typedarray_type got;
  size_t length = 0, offset = 0;
  void* data = NULL;
  napi_value buffer;
  if (napi_is_typedarray(env, v, &is) != napi_ok || !is ||
      napi_get_typedarray_info(env, v, &got, &length, &data, &buffer, &offset) != napi_ok || got != want ||
      length < count) {
    char text[160];
    snprintf(text, sizeof text, "column %s must be a typed array of the column's type with %zu entries", name, count);
    fail(env, "addBody", text);
    return false;
  }
  *out = data; /* already advanced by the view's byte offset */
  return true;
}

/* create(gravity: Float64Array(3) | null, substeps: int, sleeping: bool) -> world */
static napi_value js_create(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  if (argc < 3) return fail(env, "World", "create(gravity, substeps, sleeping)");
  const void* gravity = NULL;
  napi_valuetype t;
  NAPI(env, napi_typeof(env, argv[0], &t));
  if (t != napi_undefined && t != napi_null) {
    bool is = false;
    napi_typedarray_type got;
    size_t length = 0, offset = 0;
    void* data = NULL;
    napi_value buffer;
    if (napi_is_typedarray(env, argv[0], &is) != napi_ok || !is ||
        napi_get_typedarray_info(env, argv[0], &got, &length, &data, &buffer, &offset) != napi_ok ||
        got != napi_float64_array || length != 3)
      return fail(env, "World", "gravity must be a Float64Array of 3");
    gravity = data;
  }
  int32_t substeps = 0;
  bool sleeping = true;
  NAPI(env, napi_get_value_int32(env, argv[1], &substeps));
  NAPI(env, napi_get_value_bool(env, argv[2], &sleeping));
  bm_world* world = NULL;
  int rc = bm_world_create(&world, 0, (const double*)gravity, substeps, sleeping ? 1 : 0);
  if (rc != BM_OK || !world) {
    char text[96];
    snprintf(text, sizeof text, "no HIP device 0 (status %d)", rc);
    return fail(env, "World", text);
  }
  holder* h = (holder*)malloc(sizeof *h);
  if (!h) {
    bm_world_destroy(world);
    return fail(env, "World", "out of host memory");
  }
  h->world = world;
  napi_value out;
  if (napi_create_external(env, h, release, NULL, &out) != napi_ok) {
    release(env, h, NULL);
    return fail(env, "World", "could not wrap the world");
  }
  return out;
}

/* add(world, bodies): bodies = {n, type, shape, lock (Uint8Array), position ... restitution
 * (Float64Array, 3 or 4 a body where bm_bodies says so), hull, field (Int32Array), framed
 * (Uint8Array), offset, frame (Float64Array)}; absent columns are NULL. */
static napi_value js_add(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  if (argc < 2) return fail(env, "addBody", "add(world, bodies)");
  holder* h = world_of(env, argv[0], "addBody");
  if (!h) return NULL;
  napi_value nv;
  int32_t n = 0;
  NAPI(env, napi_get_named_property(env, argv[1], "n", &nv));
  if (napi_get_value_int32(env, nv, &n) != napi_ok || n < 0) return fail(env, "addBody", "n must be a count");
  bm_bodies b = {0};
  b.n = n;
  const size_t k = (size_t)n;
  struct {
    const char* name;
    napi_typedarray_type type;
    size_t per;
    const void** at;
  } cols[] = {
      {"type", napi_uint8_array, 1, (const void**)&b.type},
      {"shape", napi_uint8_array, 1, (const void**)&b.shape},
      {"lock", napi_uint8_array, 1, (const void**)&b.lock},
      {"position", napi_float64_array, 3, (const void**)&b.position},
      {"rotation", napi_float64_array, 4, (const void**)&b.rotation},
      {"velocity", napi_float64_array, 3, (const void**)&b.velocity},
      {"angular", napi_float64_array, 3, (const void**)&b.angular},
      {"dims", napi_float64_array, 3, (const void**)&b.dims},
      {"inertia", napi_float64_array, 3, (const void**)&b.inertia},
      {"inv_inertia", napi_float64_array, 3, (const void**)&b.inv_inertia},
      {"mass", napi_float64_array, 1, (const void**)&b.mass},
      {"inv_mass", napi_float64_array, 1, (const void**)&b.inv_mass},
      {"gravity_scale", napi_float64_array, 1, (const void**)&b.gravity_scale},
      {"damping", napi_float64_array, 1, (const void**)&b.damping},
      {"angular_damping", napi_float64_array, 1, (const void**)&b.angular_damping},
      {"size", napi_float64_array, 1, (const void**)&b.size},
      {"reach", napi_float64_array, 1, (const void**)&b.reach},
      {"friction", napi_float64_array, 1, (const void**)&b.friction},
      {"restitution", napi_float64_array, 1, (const void**)&b.restitution},
      {"hull", napi_int32_array, 1, (const void**)&b.hull},
      {"framed", napi_uint8_array, 1, (const void**)&b.framed},
      {"offset", napi_float64_array, 3, (const void**)&b.offset},
      {"frame", napi_float64_array, 4, (const void**)&b.frame},
      {"field", napi_int32_array, 1, (const void**)&b.field},
  };
  for (size_t i = 0; i < sizeof cols / sizeof cols[0]; i++)
    if (!column(env, argv[1], cols[i].name, cols[i].type, cols[i].per * k, cols[i].at)) return NULL;
  if (bm_world_add(h->world, &b) != BM_OK) return fail(env, "addBody", bm_world_error(h->world));
  return NULL;
}

/* Typed array `name` of obj, of type `want`: its data and length; false with an Error thrown. */
static bool array_of(napi_env env, napi_value obj, const char* caller, const char* name, napi_typedarray_type want,
                     void** out, size_t* length) {
  napi_value v, buffer;
  bool is = false;
  napi_typedarray_type got;
  size_t offset = 0;
  if (napi_get_named_property(env, obj, name, &v) != napi_ok || napi_is_typedarray(env, v, &is) != napi_ok || !is ||
      napi_get_typedarray_info(env, v, &got, length, out, &buffer, &offset) != napi_ok || got != want) {
    char text[160];
    snprintf(text, sizeof text, "%s must be a typed array of its column's type", name);
    fail(env, caller, text);
    return false;
  }
  return true;
}

/* setHulls(world, table): the hull bodies' hulls in their frames, table = {hullV, hullF, hullL,
 * start, loops (Int32Array), vertices, planes (Float64Array)}, n = hullV.length - 1 hulls, laid
 * out as bm_world_set_hulls takes them. */
static napi_value js_set_hulls(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  if (argc < 2) return fail(env, "addBody", "setHulls(world, table)");
  holder* h = world_of(env, argv[0], "addBody");
  if (!h) return NULL;
  void *hv, *hf, *hl, *st, *lp, *vx, *pl;
  size_t nv, nf, nl, n1, n2, n3, n4;
  if (!array_of(env, argv[1], "addBody", "hullV", napi_int32_array, &hv, &nv) ||
      !array_of(env, argv[1], "addBody", "hullF", napi_int32_array, &hf, &nf) ||
      !array_of(env, argv[1], "addBody", "hullL", napi_int32_array, &hl, &nl) ||
      !array_of(env, argv[1], "addBody", "start", napi_int32_array, &st, &n1) ||
      !array_of(env, argv[1], "addBody", "loops", napi_int32_array, &lp, &n2) ||
      !array_of(env, argv[1], "addBody", "vertices", napi_float64_array, &vx, &n3) ||
      !array_of(env, argv[1], "addBody", "planes", napi_float64_array, &pl, &n4))
    return NULL;
  if (nv < 1 || nf != nv || nl != nv) return fail(env, "addBody", "hullV, hullF and hullL must each hold n + 1 offsets");
  const int32_t n = (int32_t)nv - 1;
  const int32_t* V = (const int32_t*)hv;
  const int32_t* F = (const int32_t*)hf;
  const int32_t* L = (const int32_t*)hl;
  if ((size_t)V[n] * 3 > n3 || (size_t)F[n] * 4 > n4 || (size_t)F[n] > n1 || (size_t)L[n] > n2)
    return fail(env, "addBody", "the hull offsets run past the tables");
  if (bm_world_set_hulls(h->world, n, V, F, L, (const double*)vx, (const int32_t*)st, (const int32_t*)lp,
                         (const double*)pl) != BM_OK)
    return fail(env, "addBody", bm_world_error(h->world));
  return NULL;
}

/* setFields(world, table): the heightfields, table = {columns, rows (Int32Array), spacing
 * (Float64Array), offsets (Float64Array of whole numbers, n + 1: field f's heights at
 * offsets[f] ..), heights (Float64Array)}, as bm_world_set_fields takes them. */
static napi_value js_set_fields(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  if (argc < 2) return fail(env, "addBody", "setFields(world, table)");
  holder* h = world_of(env, argv[0], "addBody");
  if (!h) return NULL;
  void *cl, *rw, *sp, *of, *ht;
  size_t nc, nr, ns, no, nh;
  if (!array_of(env, argv[1], "addBody", "columns", napi_int32_array, &cl, &nc) ||
      !array_of(env, argv[1], "addBody", "rows", napi_int32_array, &rw, &nr) ||
      !array_of(env, argv[1], "addBody", "spacing", napi_float64_array, &sp, &ns) ||
      !array_of(env, argv[1], "addBody", "offsets", napi_float64_array, &of, &no) ||
      !array_of(env, argv[1], "addBody", "heights", napi_float64_array, &ht, &nh))
    return NULL;
  if (nr != nc || ns != nc || no != nc + 1) return fail(env, "addBody", "a heightfield table must hold n columns, rows and spacings and n + 1 offsets");
  int64_t* offsets = (int64_t*)malloc(sizeof(int64_t) * no);
  if (!offsets) return fail(env, "addBody", "out of memory");
  const double* o = (const double*)of;
  for (size_t i = 0; i < no; i++) {
    if (!(o[i] >= 0 && o[i] <= (double)nh && o[i] == (double)(int64_t)o[i])) {
      free(offsets);
      return fail(env, "addBody", "a heightfield's offsets must be whole numbers within its heights");
    }
    offsets[i] = (int64_t)o[i];
  }
  const int rc = bm_world_set_fields(h->world, (int32_t)nc, (const int32_t*)cl, (const int32_t*)rw, (const double*)sp,
                                     offsets, (const double*)ht);
  free(offsets);
  if (rc != BM_OK) return fail(env, "addBody", bm_world_error(h->world));
  return NULL;
}

/* step(world, dt) */
static napi_value js_step(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  if (argc < 2) return fail(env, "step", "step(world, dt)");
  holder* h = world_of(env, argv[0], "step");
  if (!h) return NULL;
  double dt = 0;
  if (napi_get_value_double(env, argv[1], &dt) != napi_ok) return fail(env, "step", "dt must be a number");
  if (bm_world_step(h->world, dt) != BM_OK) return fail(env, "step", bm_world_error(h->world));
  return NULL;
}

static napi_value set_int(napi_env env, napi_value obj, const char* name, double value) {
  napi_value v;
  NAPI(env, napi_create_double(env, value, &v));
  NAPI(env, napi_set_named_property(env, obj, name, v));
  return obj;
}

/* A Float32Array argument of `per` floats a slot, or NULL when it is null or absent; *bad is
 * set when it is given but wrong (the exception is then pending). */
static float* floats_arg(napi_env env, napi_value v, const char* name, size_t n, int* bad) {
  napi_valuetype t;
  *bad = 0;
  if (napi_typeof(env, v, &t) != napi_ok) return (*bad = 1), NULL;
  if (t == napi_undefined || t == napi_null) return NULL;
  const void* data = NULL;
  napi_value holder_obj;
  if (napi_create_object(env, &holder_obj) != napi_ok || napi_set_named_property(env, holder_obj, name, v) != napi_ok ||
      !column(env, holder_obj, name, napi_float32_array, n, &data))
    return (*bad = 1), NULL;
  return (float*)data;
}

/* read(world, transforms: Float32Array | null, previous: Float32Array | null) -> {top, awake,
 * contacts, pairs, frozen, impacts, catches}; transforms and previous, when given, receive 7
 * floats a slot for `top` slots. */
static napi_value js_read(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  if (argc < 1) return fail(env, "transforms", "read(world, transforms, previous)");
  holder* h = world_of(env, argv[0], "transforms");
  if (!h) return NULL;
  bm_world_state s = {0};
  if (bm_world_read(h->world, &s) != BM_OK) return fail(env, "transforms", bm_world_error(h->world));
  int bad = 0;
  if (argc >= 2) {
    s.transforms = floats_arg(env, argv[1], "transforms", (size_t)s.top * 7, &bad);
    if (bad) return NULL;
  }
  if (argc >= 3) {
    s.previous = floats_arg(env, argv[2], "previous", (size_t)s.top * 7, &bad);
    if (bad) return NULL;
  }
  if ((s.transforms || s.previous) && bm_world_read(h->world, &s) != BM_OK)
    return fail(env, "transforms", bm_world_error(h->world));
  napi_value out;
  NAPI(env, napi_create_object(env, &out));
  if (!set_int(env, out, "top", s.top) || !set_int(env, out, "awake", s.awake) ||
      !set_int(env, out, "contacts", s.contacts) || !set_int(env, out, "pairs", s.pairs) ||
      !set_int(env, out, "frozen", s.frozen) || !set_int(env, out, "impacts", (double)s.impacts) ||
      !set_int(env, out, "catches", (double)s.catches))
    return NULL;
  return out;
}

/* moved(world, since, out: Int32Array with room for every slot) -> how many slots it wrote:
 * world.moved(since), src/world.js:771-803. */
static napi_value js_moved(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  if (argc < 3) return fail(env, "moved", "moved(world, since, out)");
  holder* h = world_of(env, argv[0], "moved");
  if (!h) return NULL;
  int32_t since = 0;
  if (napi_get_value_int32(env, argv[1], &since) != napi_ok) return fail(env, "moved", "since must be a number");
  bm_world_state s = {0};
  if (bm_world_read(h->world, &s) != BM_OK) return fail(env, "moved", bm_world_error(h->world));
  napi_value holder_obj;
  const void* data = NULL;
  NAPI(env, napi_create_object(env, &holder_obj));
  NAPI(env, napi_set_named_property(env, holder_obj, "out", argv[2]));
  if (!column(env, holder_obj, "out", napi_int32_array, (size_t)s.top, &data)) return NULL;
  int32_t count = 0;
  if (bm_world_moved(h->world, since, (int32_t*)data, &count) != BM_OK) return fail(env, "moved", bm_world_error(h->world));
  napi_value out;
  NAPI(env, napi_create_int32(env, count, &out));
  return out;
}

/* contacts(world) -> Float64Array(10 n): the last step's contact events (bm_world_events_of),
 * [type, slot a, slot b, point 3, normal 3, speed] an event, in the order the step wrote them. */
static napi_value js_contacts(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  if (argc < 1) return fail(env, "contacts", "contacts(world)");
  holder* h = world_of(env, argv[0], "contacts");
  if (!h) return NULL;
  bm_world_events e = {0};
  if (bm_world_events_of(h->world, &e) != BM_OK) return fail(env, "contacts", bm_world_error(h->world));
  napi_value buffer, out;
  void* data = NULL;
  const size_t n = (size_t)e.count;
  NAPI(env, napi_create_arraybuffer(env, sizeof(double) * 10 * (n > 0 ? n : 1), &data, &buffer));
  double* d = (double*)data;
  for (size_t i = 0; i < n; i++) {
    double* r = d + 10 * i;
    r[0] = e.type[i];
    r[1] = e.a[i];
    r[2] = e.b[i];
    for (int k = 0; k < 3; k++) {
      r[3 + k] = e.point[3 * i + k];
      r[6 + k] = e.normal[3 * i + k];
    }
    r[9] = e.speed[i];
  }
  NAPI(env, napi_create_typedarray(env, napi_float64_array, 10 * n, buffer, 0, &out));
  return out;
}

/* times(world) -> Float64Array(BM_WORLD_TIMES): milliseconds spent in each part so far. */
static napi_value js_times(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  if (argc < 1) return fail(env, "times", "times(world)");
  holder* h = world_of(env, argv[0], "times");
  if (!h) return NULL;
  napi_value buffer, out;
  void* data = NULL;
  NAPI(env, napi_create_arraybuffer(env, sizeof(double) * BM_WORLD_TIMES, &data, &buffer));
  if (bm_world_times(h->world, (double*)data) != BM_OK) return fail(env, "times", bm_world_error(h->world));
  NAPI(env, napi_create_typedarray(env, napi_float64_array, BM_WORLD_TIMES, buffer, 0, &out));
  return out;
}

/* destroy(world): frees the world now rather than at collection; later calls on it throw. */
static napi_value js_destroy(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  if (argc < 1) return fail(env, "destroy", "destroy(world)");
  holder* h = world_of(env, argv[0], "destroy");
  if (!h) return NULL;
  bm_world_destroy(h->world);
  h->world = NULL;
  return NULL;
}

NAPI_MODULE_INIT() {
  static const struct {
    const char* name;
    napi_callback fn;
  } fns[] = {{"create", js_create}, {"add", js_add}, {"setHulls", js_set_hulls}, {"setFields", js_set_fields}, {"step", js_step}, {"read", js_read}, {"moved", js_moved}, {"contacts", js_contacts}, {"times", js_times}, {"destroy", js_destroy}};
  for (size_t i = 0; i < sizeof fns / sizeof fns[0]; i++) {
    napi_value f;
    if (napi_create_function(env, fns[i].name, NAPI_AUTO_LENGTH, fns[i].fn, NULL, &f) != napi_ok ||
        napi_set_named_property(env, exports, fns[i].name, f) != napi_ok)
      return NULL;
  }
  return exports;
}
