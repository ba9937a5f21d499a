/*
 * firex_napi.c — Node-API (N-API 8) addon over the C ABI of libfirex_hip.so.
 *
 * This is the binding the reference's TypeScript host loads in place of its
 * WebGPU compute path (FireXEngine.encodeSimulationStep,
 * src/sim/firex/FireXEngine.ts:2115-2276).  It is a thin marshalling layer:
 * every function forwards to one firex_* call and turns a FIREX_ERR_* status
 * into a thrown JS Error carrying firex_last_error() (an aborted offline
 * render throws an Error named "AbortError", as FireXEngine.ts:2571-2574 does).
 *
 * Solver handles are napi_external values whose finalizer destroys the solver.
 */
#include <node_api.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/firex.h"

#define NAPI_CALL(env, call)                                            \
  do {                                                                  \
    if ((call) != napi_ok) {                                            \
      napi_throw_error((env), NULL, "N-API call failed: " #call);       \
      return NULL;                                                      \
    }                                                                   \
  } while (0)

typedef struct {
  firex_solver* solver;
} handle_t;

static napi_value throw_status(napi_env env, int rc) {
  const char* msg = firex_last_error();
  char code[32];
  snprintf(code, sizeof code, "FIREX_ERR_%d", rc);
  if (rc == FIREX_ERR_ABORTED) {
    napi_value err, m, name;
    napi_create_string_utf8(env, msg && *msg ? msg : "aborted", NAPI_AUTO_LENGTH, &m);
    napi_create_error(env, NULL, m, &err);
    napi_create_string_utf8(env, "AbortError", NAPI_AUTO_LENGTH, &name);
    napi_set_named_property(env, err, "name", name);
    napi_throw(env, err);
    return NULL;
  }
  napi_throw_error(env, code, msg && *msg ? msg : "Fire-X native call failed");
  return NULL;
}

#define FX_CHECK(env, rc)                      \
  do {                                         \
    int rc_ = (rc);                            \
    if (rc_ != FIREX_OK) return throw_status((env), rc_); \
  } while (0)

static void finalize_handle(napi_env env, void* data, void* hint) {
  (void)env;
  (void)hint;
  handle_t* h = (handle_t*)data;
  if (h->solver) firex_destroy(h->solver);
  free(h);
}

static firex_solver* get_solver(napi_env env, napi_value v) {
  handle_t* h = NULL;
  if (napi_get_value_external(env, v, (void**)&h) != napi_ok || !h || !h->solver) {
    napi_throw_type_error(env, NULL, "expected a live Fire-X solver handle");
    return NULL;
  }
  return h->solver;
}

static int get_args(napi_env env, napi_callback_info info, size_t want, napi_value* argv) {
  size_t argc = want;
  if (napi_get_cb_info(env, info, &argc, argv, NULL, NULL) != napi_ok) return 0;
  for (size_t i = argc; i < want; ++i) napi_get_undefined(env, &argv[i]);
  return 1;
}

static double num_or(napi_env env, napi_value obj, const char* key, double fallback) {
  bool has = false;
  if (napi_has_named_property(env, obj, key, &has) != napi_ok || !has) return fallback;
  napi_value v;
  napi_valuetype t;
  napi_get_named_property(env, obj, key, &v);
  napi_typeof(env, v, &t);
  if (t == napi_boolean) {
    bool b;
    napi_get_value_bool(env, v, &b);
    return b ? 1 : 0;
  }
  if (t != napi_number) return fallback;
  double d;
  napi_get_value_double(env, v, &d);
  return d;
}

static int str_arg(napi_env env, napi_value v, char* buf, size_t cap) {
  size_t len = 0;
  return napi_get_value_string_utf8(env, v, buf, cap, &len) == napi_ok;
}

static const char* storage_names[] = {"f32", "f16", "bf16"};
static const char* heavy_names[] = {"auto", "off", "f16", "f32"};
static const char* plan_names[] = {"reference", "fused"};

static int lookup(const char* s, const char** names, int n) {
  for (int i = 0; i < n; ++i)
    if (strcmp(s, names[i]) == 0) return i;
  return -1;
}

static int str_prop(napi_env env, napi_value obj, const char* key, char* buf, size_t cap) {
  bool has = false;
  if (napi_has_named_property(env, obj, key, &has) != napi_ok || !has) return 0;
  napi_value v;
  napi_valuetype t;
  napi_get_named_property(env, obj, key, &v);
  napi_typeof(env, v, &t);
  if (t != napi_string) return 0;
  return str_arg(env, v, buf, cap);
}

/* Options object -> firex_config: {tier, dims:[x,y,z], pressureIterations,
 * correctionIterations, timeStep, maximumSubsteps, speciesStorage,
 * heavyMode, plan, cinematic, device, memoryBudgetBytes,
 * jacobiSweepsPerLaunch, captureGraphs, jacobiAutotune, jacobiTileHeight,
 * jacobiPrefetch, particleCount, raySteps}.  Returns 0 after throwing. */
static int parse_config(napi_env env, napi_value obj, firex_config* cfg) {
  napi_valuetype t;
  napi_typeof(env, obj, &t);
  firex_config_init(cfg);
  char buf[64];
  if (t != napi_object) return 1;
  if (str_prop(env, obj, "tier", buf, sizeof buf)) {
    int rc = firex_config_from_tier(cfg, buf);
    if (rc != FIREX_OK) {
      throw_status(env, rc);
      return 0;
    }
  }
  bool has = false;
  napi_has_named_property(env, obj, "dims", &has);
  if (has) {
    napi_value dims;
    napi_get_named_property(env, obj, "dims", &dims);
    napi_valuetype dt;
    napi_typeof(env, dims, &dt);
    if (dt != napi_null && dt != napi_undefined) {
      for (uint32_t i = 0; i < 3; ++i) {
        napi_value e;
        double d = 0;
        if (napi_get_element(env, dims, i, &e) != napi_ok || napi_get_value_double(env, e, &d) != napi_ok) {
          napi_throw_type_error(env, NULL, "dims must be an array of three numbers");
          return 0;
        }
        cfg->dims[i] = (int32_t)d;
      }
    }
  }
  cfg->pressure_iterations = (int32_t)num_or(env, obj, "pressureIterations", cfg->pressure_iterations);
  cfg->correction_iterations = (int32_t)num_or(env, obj, "correctionIterations", cfg->correction_iterations);
  cfg->time_step = num_or(env, obj, "timeStep", cfg->time_step);
  cfg->maximum_substeps = (int32_t)num_or(env, obj, "maximumSubsteps", cfg->maximum_substeps);
  cfg->cinematic = (int32_t)num_or(env, obj, "cinematic", cfg->cinematic);
  cfg->device = (int32_t)num_or(env, obj, "device", cfg->device);
  cfg->memory_budget_bytes = (uint64_t)num_or(env, obj, "memoryBudgetBytes", 0);
  cfg->jacobi_sweeps_per_launch = (int32_t)num_or(env, obj, "jacobiSweepsPerLaunch", 0);
  cfg->capture_graphs = (int32_t)num_or(env, obj, "captureGraphs", 0);
  cfg->jacobi_autotune = (int32_t)num_or(env, obj, "jacobiAutotune", 0);
  cfg->jacobi_tile_height = (int32_t)num_or(env, obj, "jacobiTileHeight", 0);
  cfg->jacobi_prefetch = (int32_t)num_or(env, obj, "jacobiPrefetch", 0);
  cfg->jacobi_ring_waves = (int32_t)num_or(env, obj, "jacobiRingWaves", 0);
  cfg->particle_count = (int32_t)num_or(env, obj, "particleCount", cfg->particle_count);
  cfg->ray_steps = (int32_t)num_or(env, obj, "raySteps", cfg->ray_steps);
  if (str_prop(env, obj, "speciesStorage", buf, sizeof buf)) cfg->species_storage = lookup(buf, storage_names, 3);
  if (str_prop(env, obj, "heavyMode", buf, sizeof buf)) cfg->heavy_mode = lookup(buf, heavy_names, 4);
  if (str_prop(env, obj, "plan", buf, sizeof buf)) cfg->plan = lookup(buf, plan_names, 2);
  return 1;
}

/* create(options): see parse_config. */
static napi_value js_create(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  firex_config cfg;
  if (!parse_config(env, argv[0], &cfg)) return NULL;
  firex_solver* s = NULL;
  FX_CHECK(env, firex_create(&cfg, &s));
  handle_t* h = (handle_t*)malloc(sizeof *h);
  h->solver = s;
  napi_value ext;
  NAPI_CALL(env, napi_create_external(env, h, finalize_handle, NULL, &ext));
  return ext;
}

/* reconfigure(h, options, guardBytes) -> true when the buffer set was rebuilt */
static napi_value js_reconfigure(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return NULL;
  firex_solver* s = get_solver(env, argv[0]);
  if (!s) return NULL;
  firex_config cfg;
  if (!parse_config(env, argv[1], &cfg)) return NULL;
  double guard = 0;
  napi_valuetype t;
  napi_typeof(env, argv[2], &t);
  if (t == napi_number) napi_get_value_double(env, argv[2], &guard);
  int32_t rebuilt = 0;
  FX_CHECK(env, firex_reconfigure(s, &cfg, guard > 0 ? (uint64_t)guard : 0, &rebuilt));
  napi_value out;
  NAPI_CALL(env, napi_get_boolean(env, rebuilt != 0, &out));
  return out;
}

static napi_value dims_array(napi_env env, const int32_t* d);

/* configFromTier(name) -> the tier's solver configuration (FIREX_QUALITY /
 * FIREX_OFFLINE_QUALITY, FireXEngine.ts:187-277) */
static napi_value js_config_from_tier(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  char tier[32];
  if (!str_arg(env, argv[0], tier, sizeof tier)) {
    napi_throw_type_error(env, NULL, "tier must be a string");
    return NULL;
  }
  firex_config cfg;
  firex_config_init(&cfg);
  FX_CHECK(env, firex_config_from_tier(&cfg, tier));
  napi_value out, v;
  NAPI_CALL(env, napi_create_object(env, &out));
  NAPI_CALL(env, napi_set_named_property(env, out, "dimensions", dims_array(env, cfg.dims)));
  const struct { const char* key; double value; } fields[] = {
      {"pressureIterations", cfg.pressure_iterations}, {"correctionIterations", cfg.correction_iterations},
      {"timeStep", cfg.time_step}, {"maximumSubsteps", cfg.maximum_substeps},
      {"particleCount", cfg.particle_count}, {"cinematic", cfg.cinematic}, {"raySteps", cfg.ray_steps},
  };
  for (size_t i = 0; i < sizeof fields / sizeof fields[0]; ++i) {
    NAPI_CALL(env, napi_create_double(env, fields[i].value, &v));
    NAPI_CALL(env, napi_set_named_property(env, out, fields[i].key, v));
  }
  return out;
}

static napi_value js_destroy(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  handle_t* h = NULL;
  if (napi_get_value_external(env, argv[0], (void**)&h) == napi_ok && h && h->solver) {
    firex_destroy(h->solver);
    h->solver = NULL;
  }
  return NULL;
}

static napi_value js_reset(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  firex_solver* s = get_solver(env, argv[0]);
  if (!s) return NULL;
  FX_CHECK(env, firex_reset(s));
  return NULL;
}

/* setParameter(h, name, number | boolean): FireXEngine.setParameter. */
static napi_value js_set_parameter(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return NULL;
  firex_solver* s = get_solver(env, argv[0]);
  if (!s) return NULL;
  char name[64];
  if (!str_arg(env, argv[1], name, sizeof name)) {
    napi_throw_type_error(env, NULL, "parameter name must be a string");
    return NULL;
  }
  napi_valuetype t;
  napi_typeof(env, argv[2], &t);
  if (t == napi_boolean) {
    bool b;
    napi_get_value_bool(env, argv[2], &b);
    FX_CHECK(env, firex_set_parameter_bool(s, name, b ? 1 : 0));
  } else if (t == napi_number) {
    double d;
    napi_get_value_double(env, argv[2], &d);
    FX_CHECK(env, firex_set_parameter(s, name, d));
  }
  /* other value types are ignored, as finiteNumber() ignores them */
  return NULL;
}

static napi_value js_get_parameter(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  firex_solver* s = get_solver(env, argv[0]);
  if (!s) return NULL;
  char name[64];
  if (!str_arg(env, argv[1], name, sizeof name)) return NULL;
  double v = 0;
  FX_CHECK(env, firex_get_parameter(s, name, &v));
  napi_value out;
  napi_create_double(env, v, &out);
  return out;
}

/* setEmitters(h, [[cx, cy, cz, radius, amplitude], ...], jitter, noise, seed) */
static napi_value js_set_emitters(napi_env env, napi_callback_info info) {
  napi_value argv[5];
  if (!get_args(env, info, 5, argv)) return NULL;
  firex_solver* s = get_solver(env, argv[0]);
  if (!s) return NULL;
  uint32_t n = 0;
  napi_get_array_length(env, argv[1], &n);
  if (n > 8) n = 9; /* let the library reject it */
  firex_emitter em[9];
  memset(em, 0, sizeof em);
  for (uint32_t i = 0; i < n && i < 9; ++i) {
    napi_value row;
    napi_get_element(env, argv[1], i, &row);
    double v[5] = {0, 0, 0, 0, 0};
    for (uint32_t k = 0; k < 5; ++k) {
      napi_value e;
      napi_get_element(env, row, k, &e);
      napi_get_value_double(env, e, &v[k]);
    }
    em[i].center[0] = (float)v[0];
    em[i].center[1] = (float)v[1];
    em[i].center[2] = (float)v[2];
    em[i].radius = (float)v[3];
    em[i].amplitude = (float)v[4];
  }
  double jitter = 0, noise = 0, seed = 0;
  napi_get_value_double(env, argv[2], &jitter);
  napi_get_value_double(env, argv[3], &noise);
  napi_get_value_double(env, argv[4], &seed);
  FX_CHECK(env, firex_set_emitters(s, em, (int32_t)n, (float)jitter, (float)noise, (uint64_t)seed));
  return NULL;
}

static napi_value js_step(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return NULL;
  firex_solver* s = get_solver(env, argv[0]);
  if (!s) return NULL;
  double dt = 0, count = 1;
  napi_get_value_double(env, argv[1], &dt);
  napi_valuetype t;
  napi_typeof(env, argv[2], &t);
  if (t == napi_number) napi_get_value_double(env, argv[2], &count);
  FX_CHECK(env, firex_step(s, dt, (int32_t)count));
  return NULL;
}

/* renderOfflineFrame(h, deltaSeconds, abortFlag?: Int32Array) — waits. */
static napi_value js_render_offline_frame(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return NULL;
  firex_solver* s = get_solver(env, argv[0]);
  if (!s) return NULL;
  double delta = 0;
  if (napi_get_value_double(env, argv[1], &delta) != napi_ok) {
    napi_throw_type_error(env, NULL, "Offline frame delta must be a number.");
    return NULL;
  }
  const volatile int32_t* flag = NULL;
  bool is_ta = false;
  napi_is_typedarray(env, argv[2], &is_ta);
  if (is_ta) {
    napi_typedarray_type type;
    size_t length;
    void* data;
    napi_get_typedarray_info(env, argv[2], &type, &length, &data, NULL, NULL);
    if (type == napi_int32_array && length > 0) flag = (const volatile int32_t*)data;
  }
  FX_CHECK(env, firex_render_offline_frame(s, delta, flag));
  return NULL;
}

static napi_value js_frame(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  firex_solver* s = get_solver(env, argv[0]);
  if (!s) return NULL;
  double delta = 0;
  napi_get_value_double(env, argv[1], &delta);
  int32_t steps = 0;
  FX_CHECK(env, firex_frame(s, delta, &steps));
  napi_value out;
  napi_create_int32(env, steps, &out);
  return out;
}

static napi_value js_sync(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  firex_solver* s = get_solver(env, argv[0]);
  if (!s) return NULL;
  FX_CHECK(env, firex_sync(s));
  return NULL;
}

static int field_id(const char* name) {
  static const char* names[] = {"flow",          "species",    "pressure",        "correction",
                                "divergence",    "reaction",   "liquid",          "flowPredicted",
                                "packedFlow",    "packedChemistry", "packedMedia",
                                "particles",     "particleDensity", "scene",         "frame"};
  for (int i = 0; i < (int)(sizeof names / sizeof names[0]); ++i)
    if (strcmp(name, names[i]) == 0) return i;
  return -1;
}

static int typed_dtype(napi_typedarray_type t) {
  if (t == napi_float32_array) return FIREX_DTYPE_F32;
  if (t == napi_float64_array) return FIREX_DTYPE_F64;
  if (t == napi_uint32_array) return FIREX_DTYPE_U32;
  return -1;
}

/* readField(h, name, out: Float32Array | Float64Array | Uint32Array) */
static napi_value js_field(napi_env env, napi_callback_info info, int write) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return NULL;
  firex_solver* s = get_solver(env, argv[0]);
  if (!s) return NULL;
  char name[32];
  if (!str_arg(env, argv[1], name, sizeof name) || field_id(name) < 0) {
    napi_throw_range_error(env, NULL, "unknown Fire-X field");
    return NULL;
  }
  bool is_ta = false;
  napi_is_typedarray(env, argv[2], &is_ta);
  if (!is_ta) {
    napi_throw_type_error(env, NULL, "expected a Float32Array, Float64Array or Uint32Array");
    return NULL;
  }
  napi_typedarray_type type;
  size_t length;
  void* data;
  NAPI_CALL(env, napi_get_typedarray_info(env, argv[2], &type, &length, &data, NULL, NULL));
  int dtype = typed_dtype(type);
  if (dtype < 0) {
    napi_throw_type_error(env, NULL, "expected a Float32Array, Float64Array or Uint32Array");
    return NULL;
  }
  size_t bytes = length * (dtype == FIREX_DTYPE_F64 ? 8 : 4);
  if (write)
    FX_CHECK(env, firex_write_field(s, field_id(name), data, bytes, dtype));
  else
    FX_CHECK(env, firex_read_field(s, field_id(name), data, bytes, dtype));
  return argv[2];
}

static napi_value js_read_field(napi_env env, napi_callback_info info) { return js_field(env, info, 0); }

static void set_num(napi_env env, napi_value obj, const char* key, double v);
static napi_value dims_array(napi_env env, const int32_t* d);

static napi_value hex64(napi_env env, uint64_t v) {
  char buf[24];
  snprintf(buf, sizeof buf, "%016llx", (unsigned long long)v);
  napi_value out;
  napi_create_string_utf8(env, buf, NAPI_AUTO_LENGTH, &out);
  return out;
}

/* renderVolume(h, {width, height, raySteps}): the volume ray marcher
 * (qualityShaders.ts:374-575) into the rgba16f scene image ("scene"). */
static napi_value js_render_volume(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  firex_solver* s = get_solver(env, argv[0]);
  if (!s) return NULL;
  firex_render_config rc;
  memset(&rc, 0, sizeof rc);
  rc.width = (int32_t)num_or(env, argv[1], "width", 0);
  rc.height = (int32_t)num_or(env, argv[1], "height", 0);
  rc.ray_steps = (int32_t)num_or(env, argv[1], "raySteps", 0);
  FX_CHECK(env, firex_render_volume(s, &rc));
  return NULL;
}

/* renderFrame(h, {width, height, raySteps}): a presented rgba8 frame ("frame"). */
static napi_value js_render_frame(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  firex_solver* s = get_solver(env, argv[0]);
  if (!s) return NULL;
  firex_render_config rc;
  memset(&rc, 0, sizeof rc);
  rc.width = (int32_t)num_or(env, argv[1], "width", 0);
  rc.height = (int32_t)num_or(env, argv[1], "height", 0);
  rc.ray_steps = (int32_t)num_or(env, argv[1], "raySteps", 0);
  FX_CHECK(env, firex_render_frame(s, &rc));
  return NULL;
}

/* resolveOfflineDims(tier, override: [x, y, z] | null) -> [x, y, z]
 * (resolveFireXOfflineDimensions, FireXEngine.ts:279-290) */
static napi_value js_resolve_offline_dims(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  char tier[16];
  if (!str_arg(env, argv[0], tier, sizeof tier)) {
    napi_throw_type_error(env, NULL, "tier must be a string");
    return NULL;
  }
  napi_valuetype t;
  napi_typeof(env, argv[1], &t);
  int32_t over[3], out[3];
  const int32_t* op = NULL;
  if (t == napi_object) {
    for (uint32_t i = 0; i < 3; ++i) {
      napi_value e;
      double d = 0;
      if (napi_get_element(env, argv[1], i, &e) != napi_ok || napi_get_value_double(env, e, &d) != napi_ok) {
        napi_throw_type_error(env, NULL, "override must be an array of three numbers");
        return NULL;
      }
      over[i] = (int32_t)d;
    }
    op = over;
  }
  FX_CHECK(env, firex_resolve_offline_dims(tier, op, out));
  return dims_array(env, out);
}

/* fieldLayout(h, name) -> {kind, components, records, storedBytes} */
static napi_value js_field_layout(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  firex_solver* s = get_solver(env, argv[0]);
  if (!s) return NULL;
  char name[32];
  if (!str_arg(env, argv[1], name, sizeof name) || field_id(name) < 0) {
    napi_throw_range_error(env, NULL, "unknown Fire-X field");
    return NULL;
  }
  firex_field_layout l;
  FX_CHECK(env, firex_get_field_layout(s, field_id(name), &l));
  static const char* kinds[] = {"f32", "f16", "bf16", "u32", "u8"};
  napi_value o, k;
  napi_create_object(env, &o);
  napi_create_string_utf8(env, l.kind >= 0 && l.kind <= 4 ? kinds[l.kind] : "?", NAPI_AUTO_LENGTH, &k);
  napi_set_named_property(env, o, "kind", k);
  set_num(env, o, "kindCode", l.kind);
  set_num(env, o, "components", l.components);
  set_num(env, o, "records", (double)l.records);
  set_num(env, o, "storedBytes", (double)l.stored_bytes);
  return o;
}

/* readFieldRaw(h, name, out: Uint8Array of storedBytes) -> out */
static napi_value js_read_field_raw(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return NULL;
  firex_solver* s = get_solver(env, argv[0]);
  if (!s) return NULL;
  char name[32];
  if (!str_arg(env, argv[1], name, sizeof name) || field_id(name) < 0) {
    napi_throw_range_error(env, NULL, "unknown Fire-X field");
    return NULL;
  }
  napi_typedarray_type type;
  size_t length;
  void* data;
  bool is_ta = false;
  napi_is_typedarray(env, argv[2], &is_ta);
  if (!is_ta || napi_get_typedarray_info(env, argv[2], &type, &length, &data, NULL, NULL) != napi_ok ||
      type != napi_uint8_array) {
    napi_throw_type_error(env, NULL, "expected a Uint8Array");
    return NULL;
  }
  FX_CHECK(env, firex_read_field(s, field_id(name), data, length, FIREX_DTYPE_RAW));
  return argv[2];
}

/* fieldChecksum(h, name) -> 16 hex digits (firex_field_checksum) */
static napi_value js_field_checksum(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  firex_solver* s = get_solver(env, argv[0]);
  if (!s) return NULL;
  char name[32];
  if (!str_arg(env, argv[1], name, sizeof name) || field_id(name) < 0) {
    napi_throw_range_error(env, NULL, "unknown Fire-X field");
    return NULL;
  }
  uint64_t h = 0;
  FX_CHECK(env, firex_field_checksum(s, field_id(name), &h));
  return hex64(env, h);
}

/* checksumBytes(data: Uint8Array) -> 16 hex digits (firex_checksum_bytes) */
static napi_value js_checksum_bytes(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  napi_typedarray_type type;
  size_t length;
  void* data;
  bool is_ta = false;
  napi_is_typedarray(env, argv[0], &is_ta);
  if (!is_ta || napi_get_typedarray_info(env, argv[0], &type, &length, &data, NULL, NULL) != napi_ok ||
      type != napi_uint8_array) {
    napi_throw_type_error(env, NULL, "expected a Uint8Array");
    return NULL;
  }
  uint64_t h = 0;
  FX_CHECK(env, firex_checksum_bytes(data, length, &h));
  return hex64(env, h);
}
static napi_value js_write_field(napi_env env, napi_callback_info info) { return js_field(env, info, 1); }

/* packFields(h): the three render volumes (qualityShaders.ts:37-73). */
static napi_value js_pack_fields(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  firex_solver* s = get_solver(env, argv[0]);
  if (!s) return NULL;
  FX_CHECK(env, firex_pack_fields(s));
  return NULL;
}

static napi_value js_reduce(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  firex_solver* s = get_solver(env, argv[0]);
  if (!s) return NULL;
  double which = 0, out = 0;
  napi_get_value_double(env, argv[1], &which);
  FX_CHECK(env, firex_reduce(s, (int32_t)which, &out));
  napi_value v;
  napi_create_double(env, out, &v);
  return v;
}

static void set_num(napi_env env, napi_value obj, const char* key, double v) {
  napi_value n;
  napi_create_double(env, v, &n);
  napi_set_named_property(env, obj, key, n);
}

static napi_value dims_array(napi_env env, const int32_t* d) {
  napi_value arr;
  napi_create_array_with_length(env, 3, &arr);
  for (uint32_t i = 0; i < 3; ++i) {
    napi_value e;
    napi_create_int32(env, d[i], &e);
    napi_set_element(env, arr, i, e);
  }
  return arr;
}

static napi_value js_get_info(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  firex_solver* s = get_solver(env, argv[0]);
  if (!s) return NULL;
  firex_info fi;
  FX_CHECK(env, firex_get_info(s, &fi));
  napi_value o;
  napi_create_object(env, &o);
  napi_set_named_property(env, o, "dimensions", dims_array(env, fi.dims));
  set_num(env, o, "cellCount", (double)fi.cell_count);
  set_num(env, o, "simulationTime", fi.simulation_time);
  set_num(env, o, "frameIndex", (double)fi.frame_index);
  set_num(env, o, "speciesIndex", fi.species_index);
  set_num(env, o, "pressureIndex", fi.pressure_index);
  set_num(env, o, "heavy", fi.heavy);
  set_num(env, o, "f16Pressure", fi.f16_pressure);
  set_num(env, o, "deviceBytes", (double)fi.device_bytes);
  set_num(env, o, "jacobiSweepsPerLaunch", fi.jacobi_sweeps_per_launch);
  set_num(env, o, "jacobiChunkZ", fi.jacobi_chunk_z);
  set_num(env, o, "jacobiTileHeight", fi.jacobi_tile_height);
  set_num(env, o, "particleCount", fi.particle_count);
  return o;
}

/* estimateMemory([x, y, z], particleCount) — estimateFireXGridMemory. */
static napi_value js_estimate_memory(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  int32_t d[3];
  for (uint32_t i = 0; i < 3; ++i) {
    napi_value e;
    double v = 0;
    napi_get_element(env, argv[0], i, &e);
    napi_get_value_double(env, e, &v);
    d[i] = (int32_t)v;
  }
  double particles = 0;
  napi_get_value_double(env, argv[1], &particles);
  firex_memory_estimate m;
  FX_CHECK(env, firex_estimate_memory(d, (int64_t)particles, &m));
  napi_value o;
  napi_create_object(env, &o);
  set_num(env, o, "cellCount", (double)m.cell_count);
  set_num(env, o, "gridBytes", (double)m.grid_bytes);
  set_num(env, o, "particleBytes", (double)m.particle_bytes);
  set_num(env, o, "fixedBytes", (double)m.fixed_bytes);
  set_num(env, o, "totalBytes", (double)m.total_bytes);
  set_num(env, o, "largestStorageBufferBytes", (double)m.largest_storage_buffer_bytes);
  return o;
}

/* splitOfflineDelta(delta, maximumStep) — splitOfflineSimulationDelta. */
static napi_value js_split_offline_delta(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  double delta = 0, maxstep = 0;
  napi_get_value_double(env, argv[0], &delta);
  napi_get_value_double(env, argv[1], &maxstep);
  int32_t count = 0;
  int rc = firex_split_offline_delta(delta, maxstep, NULL, 0, &count);
  if (rc != FIREX_OK) {
    napi_throw_range_error(env, NULL, firex_last_error());
    return NULL;
  }
  napi_value arr;
  napi_create_array_with_length(env, (size_t)count, &arr);
  double step = count > 0 ? delta / count : 0;
  for (int32_t i = 0; i < count; ++i) {
    napi_value e;
    napi_create_double(env, step, &e);
    napi_set_element(env, arr, (uint32_t)i, e);
  }
  return arr;
}

static napi_value js_device_count(napi_env env, napi_callback_info info) {
  (void)info;
  napi_value v;
  napi_create_int32(env, firex_device_count(), &v);
  return v;
}

static napi_value js_abi_version(napi_env env, napi_callback_info info) {
  (void)info;
  napi_value v;
  napi_create_int32(env, firex_abi_version(), &v);
  return v;
}

static napi_value init(napi_env env, napi_value exports) {
  napi_property_descriptor props[] = {
      {"create", NULL, js_create, NULL, NULL, NULL, napi_default, NULL},
      {"destroy", NULL, js_destroy, NULL, NULL, NULL, napi_default, NULL},
      {"reconfigure", NULL, js_reconfigure, NULL, NULL, NULL, napi_default, NULL},
      {"configFromTier", NULL, js_config_from_tier, NULL, NULL, NULL, napi_default, NULL},
      {"reset", NULL, js_reset, NULL, NULL, NULL, napi_default, NULL},
      {"setParameter", NULL, js_set_parameter, NULL, NULL, NULL, napi_default, NULL},
      {"getParameter", NULL, js_get_parameter, NULL, NULL, NULL, napi_default, NULL},
      {"setEmitters", NULL, js_set_emitters, NULL, NULL, NULL, napi_default, NULL},
      {"step", NULL, js_step, NULL, NULL, NULL, napi_default, NULL},
      {"renderOfflineFrame", NULL, js_render_offline_frame, NULL, NULL, NULL, napi_default, NULL},
      {"frame", NULL, js_frame, NULL, NULL, NULL, napi_default, NULL},
      {"sync", NULL, js_sync, NULL, NULL, NULL, napi_default, NULL},
      {"readField", NULL, js_read_field, NULL, NULL, NULL, napi_default, NULL},
      {"writeField", NULL, js_write_field, NULL, NULL, NULL, napi_default, NULL},
      {"reduce", NULL, js_reduce, NULL, NULL, NULL, napi_default, NULL},
      {"packFields", NULL, js_pack_fields, NULL, NULL, NULL, napi_default, NULL},
      {"fieldLayout", NULL, js_field_layout, NULL, NULL, NULL, napi_default, NULL},
      {"resolveOfflineDims", NULL, js_resolve_offline_dims, NULL, NULL, NULL, napi_default, NULL},
      {"renderVolume", NULL, js_render_volume, NULL, NULL, NULL, napi_default, NULL},
      {"renderFrame", NULL, js_render_frame, NULL, NULL, NULL, napi_default, NULL},
      {"readFieldRaw", NULL, js_read_field_raw, NULL, NULL, NULL, napi_default, NULL},
      {"fieldChecksum", NULL, js_field_checksum, NULL, NULL, NULL, napi_default, NULL},
      {"checksumBytes", NULL, js_checksum_bytes, NULL, NULL, NULL, napi_default, NULL},
      {"getInfo", NULL, js_get_info, NULL, NULL, NULL, napi_default, NULL},
      {"estimateMemory", NULL, js_estimate_memory, NULL, NULL, NULL, napi_default, NULL},
      {"splitOfflineDelta", NULL, js_split_offline_delta, NULL, NULL, NULL, napi_default, NULL},
      {"deviceCount", NULL, js_device_count, NULL, NULL, NULL, napi_default, NULL},
      {"abiVersion", NULL, js_abi_version, NULL, NULL, NULL, napi_default, NULL},
  };
  napi_define_properties(env, exports, sizeof props / sizeof props[0], props);
  return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, init)
