/*
 * Node N-API addon over liboutbreak_hip.so (include/outbreak_hip.h).
 *
 * A thin, 1:1 layer: each export calls one C-ABI entry point with the data of a
 * TypedArray; errors become JS exceptions carrying the library's message.
 * obl_sync runs as napi_async_work and resolves a Promise, the counterpart of
 * the reference's queue.onSubmittedWorkDone() (app.js:1704). The JavaScript
 * classes RegionalModel / PoissonDirect (outbreak.js) are built on these.
 *
 * Plain C against node_api.h (NAPI_VERSION 8); the library is loaded with
 * dlopen from the package directory, so the addon itself links no HIP code.
 */
#define NAPI_VERSION 8
#include <dlfcn.h>
#include <node_api.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "outbreak_hip.h"

typedef struct {
  void* so;
  int (*create)(const obl_grid*, obl_model**);
  void (*destroy)(obl_model*);
  const char* (*last_error)(const obl_model*);
  const char* (*global_error)(void);
  int (*set_base_state)(obl_model*, const float*, int);
  int (*set_boundaries)(obl_model*, const float*, int);
  int (*set_ci)(obl_model*, const float*, int);
  int (*set_surface)(obl_model*, const float*, size_t);
  int (*set_prep_field)(obl_model*, int);
  int (*set_vortex)(obl_model*, const float*, int, size_t);
  int (*vortex_table)(int, double, double, double, const float*, const double*, int, float*, size_t, double*);
  double (*grid_bytes)(const obl_grid*);
  int (*reset)(obl_model*, const float*, int);
  int (*step)(obl_model*, const float*, int, unsigned);
  int (*maps_request)(obl_model*, const float*, int, int*);
  int (*maps_collect)(obl_model*, int, float*, size_t);
  int (*r_create)(obl_model*, size_t, size_t, obl_radar**);
  void (*r_destroy)(obl_radar*);
  int (*r_set_tornadoes)(obl_model*, const float*, int);
  int (*r_cut)(obl_radar*, const float*, int);
  int (*r_readback)(obl_radar*, int*);
  int (*r_collect)(obl_radar*, int, void*, size_t);
  int (*add_bubble)(obl_model*, double, double, double, double, double, double);
  int (*upload)(obl_model*, int, int, const void*, size_t, size_t);
  int (*download)(obl_model*, int, int, void*, size_t, size_t);
  int (*field_ptr)(obl_model*, int, int, void**, size_t*);
  int (*sync)(obl_model*);
  int (*set_solver_mode)(obl_model*, int, int, double);
  int (*profile)(obl_model*, int);
  int (*profile_read)(obl_model*, char*, size_t, double*, int*, int);
  int (*p_create)(int, int, int, double, double, obl_poisson**);
  void (*p_destroy)(obl_poisson*);
  int (*p_set_density)(obl_poisson*, const double*, const double*);
  int (*p_solve_host)(obl_poisson*, const float*, float*, int, int, double);
  const char* (*p_last_error)(const obl_poisson*);
  /* microphysics surrogate (no reference counterpart) */
  int (*dump_enable)(obl_model*, unsigned, float, float);
  int (*dump_read)(obl_model*, float*, size_t, size_t*, void**);
  int (*set_surrogate)(obl_model*, obl_mlp*, unsigned);
  int (*surrogate_stats)(obl_model*, obl_surr_stats*, int);
  int (*m_create)(obl_mlp**);
  void (*m_destroy)(obl_mlp*);
  const char* (*m_last_error)(const obl_mlp*);
  int (*m_nparams)(void);
  int (*m_set_params)(obl_mlp*, const float*, int);
  int (*m_get_params)(obl_mlp*, float*, int);
  int (*m_set_scales)(obl_mlp*, const float*, const float*, const float*);
  int (*m_set_dead_zone)(obl_mlp*, float);
  int (*m_set_hand_min)(obl_mlp*, float);
  int (*m_load_rows)(obl_mlp*, const float*, size_t, int);
  int (*m_train)(obl_mlp*, const float*, size_t, int, int, int, float, unsigned, int, float*);
  int (*m_infer)(obl_mlp*, const float*, int, int, float*);
} Api;

static Api A;

#define CHECK_NAPI(call)                                              \
  do {                                                                \
    if ((call) != napi_ok) {                                          \
      napi_throw_error(env, NULL, "N-API call failed: " #call);       \
      return NULL;                                                    \
    }                                                                 \
  } while (0)

static napi_value throw_msg(napi_env env, const char* msg) {
  napi_throw_error(env, NULL, msg && *msg ? msg : "outbreak_hip: unknown error");
  return NULL;
}

static int load_api(const char* path, char* err, size_t n) {
  if (A.so) return 0;
  void* so = dlopen(path, RTLD_NOW | RTLD_LOCAL);
  if (!so) {
    snprintf(err, n, "cannot load %s: %s", path, dlerror());
    return -1;
  }
#define SYM(field, name)                                         \
  *(void**)(&A.field) = dlsym(so, name);                         \
  if (!A.field) {                                                \
    snprintf(err, n, "%s does not export %s", path, name);       \
    return -1;                                                   \
  }
  SYM(create, "obl_create");
  SYM(destroy, "obl_destroy");
  SYM(last_error, "obl_last_error");
  SYM(global_error, "obl_global_error");
  SYM(set_base_state, "obl_set_base_state");
  SYM(set_boundaries, "obl_set_boundaries");
  SYM(set_ci, "obl_set_ci");
  SYM(set_surface, "obl_set_surface");
  SYM(set_prep_field, "obl_set_prep_field");
  SYM(set_vortex, "obl_set_vortex");
  SYM(vortex_table, "obl_vortex_table");
  SYM(grid_bytes, "obl_grid_bytes");
  SYM(reset, "obl_reset");
  SYM(step, "obl_step");
  SYM(maps_request, "obl_maps_request");
  SYM(maps_collect, "obl_maps_collect");
  SYM(r_create, "obl_radar_create");
  SYM(r_destroy, "obl_radar_destroy");
  SYM(r_set_tornadoes, "obl_radar_set_tornadoes");
  SYM(r_cut, "obl_radar_cut");
  SYM(r_readback, "obl_radar_readback");
  SYM(r_collect, "obl_radar_collect");
  SYM(add_bubble, "obl_add_bubble");
  SYM(upload, "obl_upload");
  SYM(download, "obl_download");
  SYM(field_ptr, "obl_field_ptr");
  SYM(sync, "obl_sync");
  SYM(set_solver_mode, "obl_set_solver_mode");
  SYM(profile, "obl_profile");
  SYM(profile_read, "obl_profile_read");
  SYM(p_create, "obl_poisson_create");
  SYM(p_destroy, "obl_poisson_destroy");
  SYM(p_set_density, "obl_poisson_set_density");
  SYM(p_solve_host, "obl_poisson_solve_host");
  SYM(p_last_error, "obl_poisson_last_error");
  SYM(dump_enable, "obl_dump_enable");
  SYM(dump_read, "obl_dump_read");
  SYM(set_surrogate, "obl_set_surrogate");
  SYM(surrogate_stats, "obl_surrogate_stats");
  SYM(m_create, "obl_mlp_create");
  SYM(m_destroy, "obl_mlp_destroy");
  SYM(m_last_error, "obl_mlp_last_error");
  SYM(m_nparams, "obl_mlp_nparams");
  SYM(m_set_params, "obl_mlp_set_params");
  SYM(m_get_params, "obl_mlp_get_params");
  SYM(m_set_scales, "obl_mlp_set_scales");
  SYM(m_set_dead_zone, "obl_mlp_set_dead_zone");
  SYM(m_set_hand_min, "obl_mlp_set_hand_min");
  SYM(m_load_rows, "obl_mlp_load_rows");
  SYM(m_train, "obl_mlp_train");
  SYM(m_infer, "obl_mlp_infer");
#undef SYM
  A.so = so;
  return 0;
}

/* ---------------- argument helpers ---------------- */
static int get_args(napi_env env, napi_callback_info info, size_t want, napi_value* argv) {
  size_t argc = want;
  if (napi_get_cb_info(env, info, &argc, argv, NULL, NULL) != napi_ok) return -1;
  for (size_t i = argc; i < want; i++) napi_get_undefined(env, &argv[i]);
  return 0;
}
static double num(napi_env env, napi_value v) {
  double d = 0;
  napi_get_value_double(env, v, &d);
  return d;
}
static void* ext(napi_env env, napi_value v) {
  void* p = NULL;
  napi_get_value_external(env, v, &p);
  return p;
}
/* data pointer and byte length of a TypedArray (or NULL) */
static void* tarr(napi_env env, napi_value v, size_t* bytes) {
  bool is = false;
  *bytes = 0;
  if (napi_is_typedarray(env, v, &is) != napi_ok || !is) return NULL;
  napi_typedarray_type t;
  size_t len = 0, off = 0;
  void* data = NULL;
  napi_value ab;
  napi_get_typedarray_info(env, v, &t, &len, &data, &ab, &off);
  size_t el = 1;
  switch (t) {
    case napi_int16_array: case napi_uint16_array: el = 2; break;
    case napi_int32_array: case napi_uint32_array: case napi_float32_array: el = 4; break;
    case napi_float64_array: case napi_bigint64_array: case napi_biguint64_array: el = 8; break;
    default: el = 1;
  }
  *bytes = len * el;
  return data;
}
static napi_value ok_or_throw(napi_env env, int rc, obl_model* h) {
  if (rc) return throw_msg(env, h ? A.last_error(h) : A.global_error());
  napi_value u;
  napi_get_undefined(env, &u);
  return u;
}

/* ---------------- exports ---------------- */
static napi_value js_load(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  get_args(env, info, 1, argv);
  char path[4096], err[4200];
  size_t n = 0;
  napi_get_value_string_utf8(env, argv[0], path, sizeof path, &n);
  if (load_api(path, err, sizeof err)) return throw_msg(env, err);
  napi_value t;
  napi_get_boolean(env, true, &t);
  return t;
}

static void model_finalize(napi_env env, void* data, void* hint) {
  (void)env; (void)hint;
  if (data && A.destroy) A.destroy((obl_model*)data);
}

static napi_value js_create(napi_env env, napi_callback_info info) {
  napi_value argv[6];
  get_args(env, info, 6, argv);
  if (!A.so) return throw_msg(env, "outbreak_hip library not loaded");
  obl_grid g = {(int)num(env, argv[0]), (int)num(env, argv[1]), (int)num(env, argv[2]), num(env, argv[3]),
                num(env, argv[4]), (int)num(env, argv[5])};
  obl_model* h = NULL;
  if (A.create(&g, &h)) return throw_msg(env, A.global_error());
  napi_value out;
  CHECK_NAPI(napi_create_external(env, h, model_finalize, NULL, &out));
  return out;
}

static napi_value js_destroy(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  get_args(env, info, 1, argv);
  /* freed by the finalizer; an explicit destroy only synchronises */
  obl_model* h = (obl_model*)ext(env, argv[0]);
  if (h) A.sync(h);
  napi_value u;
  napi_get_undefined(env, &u);
  return u;
}

static napi_value js_set_base_state(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  get_args(env, info, 3, argv);
  obl_model* h = ext(env, argv[0]);
  size_t b;
  float* B = tarr(env, argv[1], &b);
  int nz = (int)num(env, argv[2]);
  if (!B || b < (size_t)(nz + 1) * 32) return throw_msg(env, "setBaseState: Float32Array of (nz+1)*8 expected");
  return ok_or_throw(env, A.set_base_state(h, B, nz), h);
}

static napi_value js_set_boundaries(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  get_args(env, info, 3, argv);
  obl_model* h = ext(env, argv[0]);
  size_t b;
  float* f = tarr(env, argv[1], &b);
  int count = (int)num(env, argv[2]);
  if (count && (!f || b < (size_t)count * OBL_BOUND_FLOATS * 4)) return throw_msg(env, "setBoundaries: short array");
  return ok_or_throw(env, A.set_boundaries(h, f, count), h);
}

static napi_value js_set_ci(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  get_args(env, info, 3, argv);
  obl_model* h = ext(env, argv[0]);
  size_t b;
  float* f = tarr(env, argv[1], &b);
  int count = (int)num(env, argv[2]);
  if (count && (!f || b < (size_t)count * OBL_CI_FLOATS * 4)) return throw_msg(env, "setCI: short array");
  return ok_or_throw(env, A.set_ci(h, f, count), h);
}

/* setSurface(h, Float32Array(4 nx ny) | null) */
static napi_value js_set_surface(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  get_args(env, info, 2, argv);
  obl_model* h = ext(env, argv[0]);
  size_t b;
  float* f = tarr(env, argv[1], &b);
  return ok_or_throw(env, A.set_surface(h, f, f ? b : 0), h);
}
/* setPrepField(h, on) */
static napi_value js_set_prep_field(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  get_args(env, info, 2, argv);
  obl_model* h = ext(env, argv[0]);
  return ok_or_throw(env, A.set_prep_field(h, num(env, argv[1]) != 0.0 ? 1 : 0), h);
}
/* setVortex(h, Float32Array(4 nz nr) | null, nr) */
static napi_value js_set_vortex(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  get_args(env, info, 3, argv);
  obl_model* h = ext(env, argv[0]);
  size_t b;
  float* f = tarr(env, argv[1], &b);
  int nr = (int)num(env, argv[2]);
  return ok_or_throw(env, A.set_vortex(h, f, f ? nr : 0, f ? b : 0), h);
}

/* vortexTable(nz, dz, dx, lat, Float32Array B, Float64Array(4) spec, mid) -> {tab: Float32Array(nz nr 4),
 * info: [dr, nr, pc, minStab, warm]} (host code, no model handle) */
static napi_value js_vortex_table(napi_env env, napi_callback_info info) {
  napi_value argv[7];
  get_args(env, info, 7, argv);
  if (!A.so) return throw_msg(env, "outbreak_hip library not loaded");
  int nz = (int)num(env, argv[0]);
  size_t bb, sb;
  float* B = tarr(env, argv[4], &bb);
  double* spec = tarr(env, argv[5], &sb);
  if (nz < 2 || !B || bb < (size_t)(nz + 1) * 32 || !spec || sb < 32)
    return throw_msg(env, "vortexTable: Float32Array((nz + 1) * 8) base state and Float64Array(4) spec expected");
  size_t cap = (size_t)nz * OBL_VORTEX_MAX_RADII * 4;
  float* tab = malloc(cap * 4);
  double out[5];
  if (!tab) return throw_msg(env, "vortexTable: out of host memory");
  int rc = A.vortex_table(nz, num(env, argv[1]), num(env, argv[2]), num(env, argv[3]), B, spec, (int)num(env, argv[6]), tab,
                          cap, out);
  if (rc) {
    free(tab);
    return throw_msg(env, "setVortex: bad vortex");
  }
  size_t nf = (size_t)nz * (size_t)out[1] * 4;
  napi_value ab, arr, inf, res, v;
  void* data = NULL;
  napi_status st = napi_create_arraybuffer(env, nf * 4, &data, &ab);
  if (st == napi_ok) memcpy(data, tab, nf * 4);
  free(tab);
  if (st != napi_ok) return throw_msg(env, "vortexTable: N-API allocation failed");
  CHECK_NAPI(napi_create_typedarray(env, napi_float32_array, nf, ab, 0, &arr));
  CHECK_NAPI(napi_create_array_with_length(env, 5, &inf));
  for (uint32_t i = 0; i < 5; i++) {
    CHECK_NAPI(napi_create_double(env, out[i], &v));
    CHECK_NAPI(napi_set_element(env, inf, i, v));
  }
  CHECK_NAPI(napi_create_object(env, &res));
  CHECK_NAPI(napi_set_named_property(env, res, "tab", arr));
  CHECK_NAPI(napi_set_named_property(env, res, "info", inf));
  return res;
}
/* gridBytes(nx, ny, nz, dx, dz, pad): the device bytes obl_create allocates (host code) */
static napi_value js_grid_bytes(napi_env env, napi_callback_info info) {
  napi_value argv[6], out;
  get_args(env, info, 6, argv);
  if (!A.so) return throw_msg(env, "outbreak_hip library not loaded");
  obl_grid g = {(int)num(env, argv[0]), (int)num(env, argv[1]), (int)num(env, argv[2]), num(env, argv[3]),
                num(env, argv[4]), (int)num(env, argv[5])};
  CHECK_NAPI(napi_create_double(env, A.grid_bytes(&g), &out));
  return out;
}

static napi_value js_reset(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  get_args(env, info, 3, argv);
  obl_model* h = ext(env, argv[0]);
  size_t b;
  float* u = tarr(env, argv[1], &b);
  if (!u || b < OBL_UNIFORM_FLOATS * 4) return throw_msg(env, "reset: Float32Array(76) uniform block expected");
  return ok_or_throw(env, A.reset(h, u, (int)num(env, argv[2])), h);
}

static napi_value js_step(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  get_args(env, info, 4, argv);
  obl_model* h = ext(env, argv[0]);
  size_t b;
  float* u = tarr(env, argv[1], &b);
  if (!u || b < OBL_UNIFORM_FLOATS * 4) return throw_msg(env, "step: Float32Array(76) uniform block expected");
  return ok_or_throw(env, A.step(h, u, (int)num(env, argv[2]), (unsigned)num(env, argv[3])), h);
}

/* mapsRequest(h, uniforms, parity) -> slot (0 / 1), or -1 while both slots await collection */
static napi_value js_maps_request(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  get_args(env, info, 3, argv);
  obl_model* h = ext(env, argv[0]);
  size_t b;
  float* u = tarr(env, argv[1], &b);
  if (!u || b < OBL_UNIFORM_FLOATS * 4) return throw_msg(env, "requestMaps: Float32Array(76) uniform block expected");
  int slot = -1;
  if (A.maps_request(h, u, (int)num(env, argv[2]), &slot)) return throw_msg(env, A.last_error(h));
  napi_value r;
  CHECK_NAPI(napi_create_int32(env, slot, &r));
  return r;
}

/* mapsCollect(h, slot, Float32Array(16 nx ny)) */
static napi_value js_maps_collect(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  get_args(env, info, 3, argv);
  obl_model* h = ext(env, argv[0]);
  size_t b;
  float* dst = tarr(env, argv[2], &b);
  if (!dst) return throw_msg(env, "collectMaps: Float32Array expected");
  return ok_or_throw(env, A.maps_collect(h, (int)num(env, argv[1]), dst, b / 4), h);
}

/* ---- virtual radars (radar.js:271-413); the model owns them, so the external has no finalizer ---- */
/* radarCreate(h, bytes, kdpFloats) -> radar */
static napi_value js_radar_create(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  get_args(env, info, 3, argv);
  obl_model* h = ext(env, argv[0]);
  obl_radar* r = NULL;
  if (A.r_create(h, (size_t)num(env, argv[1]), (size_t)num(env, argv[2]), &r))
    return throw_msg(env, h ? A.last_error(h) : A.global_error());
  napi_value out;
  CHECK_NAPI(napi_create_external(env, r, NULL, NULL, &out));
  return out;
}
/* radarDestroy(radar) */
static napi_value js_radar_destroy(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  get_args(env, info, 1, argv);
  A.r_destroy((obl_radar*)ext(env, argv[0]));
  napi_value u;
  napi_get_undefined(env, &u);
  return u;
}
/* radarSetTornadoes(h, Float32Array(8 n), n) */
static napi_value js_radar_set_tornadoes(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  get_args(env, info, 3, argv);
  obl_model* h = ext(env, argv[0]);
  size_t b;
  float* f = tarr(env, argv[1], &b);
  int n = (int)num(env, argv[2]);
  if (n > 0 && (!f || b < (size_t)n * 32)) return throw_msg(env, "setTornadoes: Float32Array(8 n) expected");
  return ok_or_throw(env, A.r_set_tornadoes(h, f, n), h);
}
/* radarCut(h, radar, Float32Array(44), parity) */
static napi_value js_radar_cut(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  get_args(env, info, 4, argv);
  obl_model* h = ext(env, argv[0]);
  size_t b;
  float* u = tarr(env, argv[2], &b);
  if (!u || b < 44 * 4) return throw_msg(env, "encCut: Float32Array(44) RADAR_FIELDS block expected");
  return ok_or_throw(env, A.r_cut((obl_radar*)ext(env, argv[1]), u, (int)num(env, argv[3])), h);
}
/* radarReadback(h, radar) -> slot (0 / 1), or -1 while both slots await collection */
static napi_value js_radar_readback(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  get_args(env, info, 2, argv);
  obl_model* h = ext(env, argv[0]);
  int slot = -1;
  if (A.r_readback((obl_radar*)ext(env, argv[1]), &slot)) return throw_msg(env, A.last_error(h));
  napi_value r;
  CHECK_NAPI(napi_create_int32(env, slot, &r));
  return r;
}
/* radarCollect(h, radar, slot, Uint8Array(bytes)) */
static napi_value js_radar_collect(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  get_args(env, info, 4, argv);
  obl_model* h = ext(env, argv[0]);
  size_t b;
  void* dst = tarr(env, argv[3], &b);
  if (!dst) return throw_msg(env, "collect: Uint8Array expected");
  return ok_or_throw(env, A.r_collect((obl_radar*)ext(env, argv[1]), (int)num(env, argv[2]), dst, b), h);
}

static napi_value js_add_bubble(napi_env env, napi_callback_info info) {
  napi_value argv[7];
  get_args(env, info, 7, argv);
  obl_model* h = ext(env, argv[0]);
  return ok_or_throw(env, A.add_bubble(h, num(env, argv[1]), num(env, argv[2]), num(env, argv[3]), num(env, argv[4]),
                                       num(env, argv[5]), num(env, argv[6])), h);
}

static napi_value js_download(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  get_args(env, info, 4, argv);
  obl_model* h = ext(env, argv[0]);
  size_t b;
  void* dst = tarr(env, argv[3], &b);
  if (!dst) return throw_msg(env, "download: TypedArray expected");
  return ok_or_throw(env, A.download(h, (int)num(env, argv[1]), (int)num(env, argv[2]), dst, b, 0), h);
}

static napi_value js_upload(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  get_args(env, info, 4, argv);
  obl_model* h = ext(env, argv[0]);
  size_t b;
  void* src = tarr(env, argv[3], &b);
  if (!src) return throw_msg(env, "upload: TypedArray expected");
  return ok_or_throw(env, A.upload(h, (int)num(env, argv[1]), (int)num(env, argv[2]), src, b, 0), h);
}

static napi_value js_field_ptr(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  get_args(env, info, 3, argv);
  obl_model* h = ext(env, argv[0]);
  void* p = NULL;
  size_t n = 0;
  if (A.field_ptr(h, (int)num(env, argv[1]), (int)num(env, argv[2]), &p, &n)) return throw_msg(env, A.last_error(h));
  napi_value o, vp, vn;
  CHECK_NAPI(napi_create_object(env, &o));
  CHECK_NAPI(napi_create_bigint_uint64(env, (uint64_t)(uintptr_t)p, &vp));
  CHECK_NAPI(napi_create_double(env, (double)n, &vn));
  napi_set_named_property(env, o, "ptr", vp);
  napi_set_named_property(env, o, "bytes", vn);
  return o;
}

static napi_value js_set_solver_mode(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  get_args(env, info, 4, argv);
  obl_model* h = ext(env, argv[0]);
  return ok_or_throw(env, A.set_solver_mode(h, (int)num(env, argv[1]), (int)num(env, argv[2]), num(env, argv[3])), h);
}

static napi_value js_profile(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  get_args(env, info, 2, argv);
  obl_model* h = ext(env, argv[0]);
  return ok_or_throw(env, A.profile(h, (int)num(env, argv[1])), h);
}

static napi_value js_profile_read(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  get_args(env, info, 1, argv);
  obl_model* h = ext(env, argv[0]);
  char names[4096];
  double ms[64];
  int cnt[64];
  int n = A.profile_read(h, names, sizeof names, ms, cnt, 64);
  if (n < 0) return throw_msg(env, A.last_error(h));
  napi_value o;
  CHECK_NAPI(napi_create_object(env, &o));
  char* s = names;
  for (int i = 0; i < n; i++) {
    char* e = strchr(s, '\n');
    if (e) *e = 0;
    napi_value v;
    napi_create_double(env, ms[i], &v);
    napi_set_named_property(env, o, s, v);
    if (!e) break;
    s = e + 1;
  }
  return o;
}

/* sync as async work -> Promise (onSubmittedWorkDone) */
typedef struct {
  obl_model* h;
  int rc;
  napi_deferred d;
  napi_async_work w;
} SyncJob;
static void sync_exec(napi_env env, void* data) {
  (void)env;
  SyncJob* j = data;
  j->rc = A.sync(j->h);
}
static void sync_done(napi_env env, napi_status st, void* data) {
  SyncJob* j = data;
  napi_value v;
  if (st == napi_ok && j->rc == 0) {
    napi_get_undefined(env, &v);
    napi_resolve_deferred(env, j->d, v);
  } else {
    napi_value msg;
    const char* m = A.last_error(j->h);
    napi_create_string_utf8(env, m && *m ? m : "sync failed", NAPI_AUTO_LENGTH, &msg);
    napi_create_error(env, NULL, msg, &v);
    napi_reject_deferred(env, j->d, v);
  }
  napi_delete_async_work(env, j->w);
  free(j);
}
static napi_value js_sync(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  get_args(env, info, 1, argv);
  SyncJob* j = calloc(1, sizeof *j);
  j->h = ext(env, argv[0]);
  napi_value promise, name;
  CHECK_NAPI(napi_create_promise(env, &j->d, &promise));
  napi_create_string_utf8(env, "obl_sync", NAPI_AUTO_LENGTH, &name);
  CHECK_NAPI(napi_create_async_work(env, NULL, name, sync_exec, sync_done, j, &j->w));
  CHECK_NAPI(napi_queue_async_work(env, j->w));
  return promise;
}

/* ---------------- PoissonDirect ---------------- */
static void poisson_finalize(napi_env env, void* data, void* hint) {
  (void)env; (void)hint;
  if (data && A.p_destroy) A.p_destroy((obl_poisson*)data);
}
static napi_value js_p_create(napi_env env, napi_callback_info info) {
  napi_value argv[5];
  get_args(env, info, 5, argv);
  if (!A.so) return throw_msg(env, "outbreak_hip library not loaded");
  obl_poisson* p = NULL;
  if (A.p_create((int)num(env, argv[0]), (int)num(env, argv[1]), (int)num(env, argv[2]), num(env, argv[3]),
                 num(env, argv[4]), &p))
    return throw_msg(env, A.global_error());
  napi_value out;
  CHECK_NAPI(napi_create_external(env, p, poisson_finalize, NULL, &out));
  return out;
}
static napi_value js_p_set_density(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  get_args(env, info, 3, argv);
  obl_poisson* p = ext(env, argv[0]);
  size_t b1, b2;
  double* rc = tarr(env, argv[1], &b1);
  double* rw = tarr(env, argv[2], &b2);
  if (!rc || !rw) return throw_msg(env, "setDensity: Float64Arrays expected");
  if (A.p_set_density(p, rc, rw)) return throw_msg(env, A.p_last_error(p));
  napi_value u;
  napi_get_undefined(env, &u);
  return u;
}
static napi_value js_p_solve(napi_env env, napi_callback_info info) {
  napi_value argv[6];
  get_args(env, info, 6, argv);
  obl_poisson* p = ext(env, argv[0]);
  size_t b1, b2;
  float* rhs = tarr(env, argv[1], &b1);
  float* phi = tarr(env, argv[2], &b2);
  if (!rhs || !phi || b1 != b2) return throw_msg(env, "solve: two Float32Arrays of the grid size expected");
  if (A.p_solve_host(p, rhs, phi, (int)num(env, argv[3]), (int)num(env, argv[4]), num(env, argv[5])))
    return throw_msg(env, A.p_last_error(p));
  napi_value u;
  napi_get_undefined(env, &u);
  return u;
}

/* ---------------- microphysics surrogate (no reference counterpart) ---------------- */
static napi_value undef(napi_env env) {
  napi_value u;
  napi_get_undefined(env, &u);
  return u;
}
static napi_value mlp_ok(napi_env env, int rc, obl_mlp* m) {
  if (rc) return throw_msg(env, A.m_last_error(m));
  return undef(env);
}
static napi_value js_sample_micro(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  get_args(env, info, 4, argv);
  obl_model* h = ext(env, argv[0]);
  return ok_or_throw(env, A.dump_enable(h, (unsigned)num(env, argv[1]), (float)num(env, argv[2]), (float)num(env, argv[3])),
                     h);
}
/* the sampled teacher rows as a Float32Array (rows x 24) */
static napi_value js_micro_samples(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  get_args(env, info, 1, argv);
  obl_model* h = ext(env, argv[0]);
  size_t rows = 0;
  if (A.dump_read(h, NULL, 0, &rows, NULL)) return throw_msg(env, A.last_error(h));
  napi_value ab, arr;
  void* data = NULL;
  CHECK_NAPI(napi_create_arraybuffer(env, rows * 24 * 4, &data, &ab));
  if (rows && A.dump_read(h, (float*)data, rows, &rows, NULL)) return throw_msg(env, A.last_error(h));
  CHECK_NAPI(napi_create_typedarray(env, napi_float32_array, rows * 24, ab, 0, &arr));
  return arr;
}
static void mlp_finalize(napi_env env, void* data, void* hint) {
  (void)env; (void)hint;
  if (data && A.m_destroy) A.m_destroy((obl_mlp*)data);
}
static napi_value js_mlp_create(napi_env env, napi_callback_info info) {
  (void)info;
  if (!A.so) return throw_msg(env, "outbreak_hip library not loaded");
  obl_mlp* m = NULL;
  if (A.m_create(&m)) return throw_msg(env, A.m_last_error(NULL));
  napi_value out;
  CHECK_NAPI(napi_create_external(env, m, mlp_finalize, NULL, &out));
  return out;
}
static napi_value js_mlp_nparams(napi_env env, napi_callback_info info) {
  (void)info;
  napi_value v;
  napi_create_double(env, (double)A.m_nparams(), &v);
  return v;
}
static napi_value js_mlp_params(napi_env env, napi_callback_info info) {
  /* (m, Float32Array): set when given with data, else get into it */
  napi_value argv[3];
  get_args(env, info, 3, argv);
  obl_mlp* m = ext(env, argv[0]);
  size_t b;
  float* p = tarr(env, argv[1], &b);
  if (!p) return throw_msg(env, "params: Float32Array expected");
  bool set = false;
  napi_get_value_bool(env, argv[2], &set);
  return mlp_ok(env, set ? A.m_set_params(m, p, (int)(b / 4)) : A.m_get_params(m, p, (int)(b / 4)), m);
}
static napi_value js_mlp_set_scales(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  get_args(env, info, 4, argv);
  obl_mlp* m = ext(env, argv[0]);
  size_t b0, b1, b2;
  float* os = tarr(env, argv[1], &b0);
  float* lo = tarr(env, argv[2], &b1);
  float* hi = tarr(env, argv[3], &b2);
  if (!os || b0 < 36 || (lo && b1 < 56) || (hi && b2 < 56)) return throw_msg(env, "setScales: Float32Array(9), (14), (14)");
  return mlp_ok(env, A.m_set_scales(m, os, lo, hi), m);
}
static napi_value js_mlp_set_dead_zone(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  get_args(env, info, 2, argv);
  obl_mlp* m = ext(env, argv[0]);
  return mlp_ok(env, A.m_set_dead_zone(m, (float)num(env, argv[1])), m);
}
static napi_value js_mlp_set_hand_min(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  get_args(env, info, 2, argv);
  obl_mlp* m = ext(env, argv[0]);
  return mlp_ok(env, A.m_set_hand_min(m, (float)num(env, argv[1])), m);
}
static napi_value js_mlp_load_rows(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  get_args(env, info, 3, argv);
  obl_mlp* m = ext(env, argv[0]);
  size_t b;
  float* r = tarr(env, argv[1], &b);
  int stride = (int)num(env, argv[2]);
  if (!r || stride < 23 || b < (size_t)stride * 4) return throw_msg(env, "loadRows: Float32Array of rows x stride expected");
  return mlp_ok(env, A.m_load_rows(m, r, b / 4 / (size_t)stride, stride), m);
}
/* (m, steps, batch, lr, seed, quant) on the loaded rows -> mean loss */
static napi_value js_mlp_train(napi_env env, napi_callback_info info) {
  napi_value argv[6];
  get_args(env, info, 6, argv);
  obl_mlp* m = ext(env, argv[0]);
  float loss = 0.0f;
  if (A.m_train(m, NULL, 0, 0, (int)num(env, argv[1]), (int)num(env, argv[2]), (float)num(env, argv[3]),
                (unsigned)num(env, argv[4]), (int)num(env, argv[5]), &loss))
    return throw_msg(env, A.m_last_error(m));
  napi_value v;
  napi_create_double(env, (double)loss, &v);
  return v;
}
/* (m, rows Float32Array, stride) -> Float32Array(n x 16) of raw outputs */
static napi_value js_mlp_infer(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  get_args(env, info, 3, argv);
  obl_mlp* m = ext(env, argv[0]);
  size_t b;
  float* r = tarr(env, argv[1], &b);
  int stride = (int)num(env, argv[2]);
  if (!r || stride < 14 || b < (size_t)stride * 4) return throw_msg(env, "infer: Float32Array of rows x stride expected");
  int n = (int)(b / 4 / (size_t)stride);
  napi_value ab, arr;
  void* data = NULL;
  CHECK_NAPI(napi_create_arraybuffer(env, (size_t)n * 16 * 4, &data, &ab));
  if (A.m_infer(m, r, n, stride, (float*)data)) return throw_msg(env, A.m_last_error(m));
  CHECK_NAPI(napi_create_typedarray(env, napi_float32_array, (size_t)n * 16, ab, 0, &arr));
  return arr;
}
static napi_value js_set_surrogate(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  get_args(env, info, 3, argv);
  obl_model* h = ext(env, argv[0]);
  napi_valuetype t;
  napi_typeof(env, argv[1], &t);
  obl_mlp* m = t == napi_external ? (obl_mlp*)ext(env, argv[1]) : NULL;
  return ok_or_throw(env, A.set_surrogate(h, m, (unsigned)num(env, argv[2])), h);
}
static void put(napi_env env, napi_value o, const char* k, double v) {
  napi_value x;
  napi_create_double(env, v, &x);
  napi_set_named_property(env, o, k, x);
}
static napi_value js_surrogate_stats(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  get_args(env, info, 2, argv);
  obl_model* h = ext(env, argv[0]);
  bool reset = false;
  napi_get_value_bool(env, argv[1], &reset);
  obl_surr_stats st;
  if (A.surrogate_stats(h, &st, reset ? 1 : 0)) return throw_msg(env, A.last_error(h));
  napi_value o, d2, r2;
  CHECK_NAPI(napi_create_object(env, &o));
  put(env, o, "steps", st.steps); put(env, o, "points", st.points); put(env, o, "fallbacks", st.fallbacks);
  put(env, o, "overflow", st.overflow); put(env, o, "judged", st.judged);
  CHECK_NAPI(napi_create_array_with_length(env, 8, &d2));
  CHECK_NAPI(napi_create_array_with_length(env, 8, &r2));
  for (uint32_t i = 0; i < 8; i++) {
    napi_value a, b;
    napi_create_double(env, st.d2[i], &a);
    napi_create_double(env, st.r2[i], &b);
    napi_set_element(env, d2, i, a);
    napi_set_element(env, r2, i, b);
  }
  napi_set_named_property(env, o, "d2", d2);
  napi_set_named_property(env, o, "r2", r2);
  put(env, o, "ccD2", st.cc_d2); put(env, o, "ccR2", st.cc_r2); put(env, o, "ccMaxDiff", st.cc_maxdiff);
  put(env, o, "ccMax", st.cc_max); put(env, o, "dbzD2", st.dbz_d2); put(env, o, "dbzCols", st.dbz_cols);
  put(env, o, "dbzMaxDiff", st.dbz_maxdiff); put(env, o, "prD2", st.pr_d2); put(env, o, "prR2", st.pr_r2);
  put(env, o, "prSumS", st.pr_sum_s); put(env, o, "prSumP", st.pr_sum_p);
  return o;
}

#define EXPORT(name, fn)                                                      \
  do {                                                                        \
    napi_value f;                                                             \
    napi_create_function(env, name, NAPI_AUTO_LENGTH, fn, NULL, &f);          \
    napi_set_named_property(env, exports, name, f);                          \
  } while (0)

static napi_value init(napi_env env, napi_value exports) {
  EXPORT("load", js_load);
  EXPORT("create", js_create);
  EXPORT("destroy", js_destroy);
  EXPORT("setBaseState", js_set_base_state);
  EXPORT("setBoundaries", js_set_boundaries);
  EXPORT("setCI", js_set_ci);
  EXPORT("reset", js_reset);
  EXPORT("setSurface", js_set_surface);
  EXPORT("setPrepField", js_set_prep_field);
  EXPORT("setVortex", js_set_vortex);
  EXPORT("vortexTable", js_vortex_table);
  EXPORT("gridBytes", js_grid_bytes);
  EXPORT("step", js_step);
  EXPORT("mapsRequest", js_maps_request);
  EXPORT("mapsCollect", js_maps_collect);
  EXPORT("radarCreate", js_radar_create);
  EXPORT("radarDestroy", js_radar_destroy);
  EXPORT("radarSetTornadoes", js_radar_set_tornadoes);
  EXPORT("radarCut", js_radar_cut);
  EXPORT("radarReadback", js_radar_readback);
  EXPORT("radarCollect", js_radar_collect);
  EXPORT("addBubble", js_add_bubble);
  EXPORT("download", js_download);
  EXPORT("upload", js_upload);
  EXPORT("fieldPtr", js_field_ptr);
  EXPORT("setSolverMode", js_set_solver_mode);
  EXPORT("profile", js_profile);
  EXPORT("profileRead", js_profile_read);
  EXPORT("sync", js_sync);
  EXPORT("poissonCreate", js_p_create);
  EXPORT("poissonSetDensity", js_p_set_density);
  EXPORT("poissonSolve", js_p_solve);
  EXPORT("sampleMicro", js_sample_micro);
  EXPORT("microSamples", js_micro_samples);
  EXPORT("setSurrogate", js_set_surrogate);
  EXPORT("surrogateStats", js_surrogate_stats);
  EXPORT("mlpCreate", js_mlp_create);
  EXPORT("mlpNparams", js_mlp_nparams);
  EXPORT("mlpParams", js_mlp_params);
  EXPORT("mlpSetScales", js_mlp_set_scales);
  EXPORT("mlpSetDeadZone", js_mlp_set_dead_zone);
  EXPORT("mlpSetHandMin", js_mlp_set_hand_min);
  EXPORT("mlpLoadRows", js_mlp_load_rows);
  EXPORT("mlpTrain", js_mlp_train);
  EXPORT("mlpInfer", js_mlp_infer);
  return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, init)
