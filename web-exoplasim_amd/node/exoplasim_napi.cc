// Node-API binding of libexoplasim_mi355x.so: the engine calls the
// reference's worker makes (runner/webgpu/gcm.worker.ts) as synchronous
// functions over the C-ABI. A failed call throws an Error carrying
// exo_last_error()'s text.
#include <node_api.h>

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/exoplasim_mi355x.h"

namespace {

#define NAPI_OK(call)                                           \
  do {                                                          \
    if ((call) != napi_ok) {                                    \
      napi_throw_error(env, nullptr, "Node-API call failed");   \
      return nullptr;                                           \
    }                                                           \
  } while (0)

#define EXO_OK_OR_THROW(call)                                   \
  do {                                                          \
    if ((call) != EXO_OK) {                                     \
      napi_throw_error(env, nullptr, exo_last_error());         \
      return nullptr;                                           \
    }                                                           \
  } while (0)

struct Args {
  size_t n = 8;
  napi_value v[8];
};

bool get_args(napi_env env, napi_callback_info info, Args& a, size_t need) {
  if (napi_get_cb_info(env, info, &a.n, a.v, nullptr, nullptr) != napi_ok) return false;
  if (a.n < need) {
    napi_throw_type_error(env, nullptr, "missing arguments");
    return false;
  }
  return true;
}

exo_engine* engine_of(napi_env env, napi_value v) {
  void* p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
    napi_throw_type_error(env, nullptr, "not an engine handle (or destroyed)");
    return nullptr;
  }
  return static_cast<exo_engine*>(p);
}

int32_t int_of(napi_env env, napi_value v) {
  int32_t x = 0;
  napi_get_value_int32(env, v, &x);
  return x;
}

napi_value number(napi_env env, double x) {
  napi_value r;
  napi_create_double(env, x, &r);
  return r;
}

napi_value double_array(napi_env env, const std::vector<double>& data) {
  napi_value buf, arr;
  void* p = nullptr;
  if (napi_create_arraybuffer(env, data.size() * sizeof(double), &p, &buf) != napi_ok) return nullptr;
  if (!data.empty()) std::memcpy(p, data.data(), data.size() * sizeof(double));
  if (napi_create_typedarray(env, napi_float64_array, data.size(), buf, 0, &arr) != napi_ok) return nullptr;
  return arr;
}

// The data of a typed array of `type` (Float32Array / Float64Array), or
// nullptr with a TypeError thrown.
void* typed_data(napi_env env, napi_value v, napi_typedarray_type type, size_t& length) {
  bool is_typed = false;
  napi_typedarray_type t;
  void* p = nullptr;
  napi_value buf;
  size_t offset = 0;
  if (napi_is_typedarray(env, v, &is_typed) != napi_ok || !is_typed ||
      napi_get_typedarray_info(env, v, &t, &length, &p, &buf, &offset) != napi_ok || t != type) {
    napi_throw_type_error(env, nullptr, type == napi_float32_array ? "expected a Float32Array" : "expected a Float64Array");
    return nullptr;
  }
  return p;
}

double number_of(napi_env env, napi_value obj, const char* key) {
  napi_value v;
  double x = 0;
  if (napi_get_named_property(env, obj, key, &v) == napi_ok) napi_get_value_double(env, v, &x);
  return x;
}

bool is_nullish(napi_env env, napi_value v) {
  napi_valuetype t;
  return napi_typeof(env, v, &t) != napi_ok || t == napi_null || t == napi_undefined;
}

napi_value float_array(napi_env env, const std::vector<float>& data) {
  napi_value buf, arr;
  void* p = nullptr;
  if (napi_create_arraybuffer(env, data.size() * sizeof(float), &p, &buf) != napi_ok) return nullptr;
  if (!data.empty()) std::memcpy(p, data.data(), data.size() * sizeof(float));
  if (napi_create_typedarray(env, napi_float32_array, data.size(), buf, 0, &arr) != napi_ok) return nullptr;
  return arr;
}

// exo_config fields by their C names; the JS side maps GcmConfig onto them.
struct Field {
  const char* name;
  char kind;  // i: int, d: double, u: uint64, l: int64, s: int32[8] (an array)
  size_t offset;
};
#define EXO_FIELD(k, n) {#n, k, offsetof(exo_config, n)}
const Field kConfigFields[] = {
    EXO_FIELD('i', truncation), EXO_FIELD('i', levels), EXO_FIELD('d', timestep_minutes),
    EXO_FIELD('d', radius_earth), EXO_FIELD('d', gravity), EXO_FIELD('d', rotation_period_days),
    EXO_FIELD('d', pressure_bar), EXO_FIELD('d', gas_constant), EXO_FIELD('d', co2_ppm),
    EXO_FIELD('d', flux_wm2), EXO_FIELD('d', obliquity_deg), EXO_FIELD('d', eccentricity),
    EXO_FIELD('d', year_days), EXO_FIELD('d', ocean_depth_m), EXO_FIELD('i', ozone),
    EXO_FIELD('d', land_fraction), EXO_FIELD('d', terrain_max_m), EXO_FIELD('d', initial_temperature),
    EXO_FIELD('d', initial_humidity), EXO_FIELD('d', noise_kelvin), EXO_FIELD('d', sea_surface_k),
    EXO_FIELD('u', terrain_seed), EXO_FIELD('u', noise_seed), EXO_FIELD('i', storm_diagnostics),
    EXO_FIELD('i', vegetation), EXO_FIELD('d', initial_growth), EXO_FIELD('i', glaciers),
    EXO_FIELD('d', glacier_min_depth_m), EXO_FIELD('d', glacier_initial_height_m),
    EXO_FIELD('d', star_temperature_k), EXO_FIELD('i', synchronous), EXO_FIELD('d', substellar_longitude_deg),
    EXO_FIELD('i', fixed_orbit), EXO_FIELD('i', keplerian), EXO_FIELD('d', periapse_longitude_deg),
    EXO_FIELD('d', initial_mean_anomaly_deg), EXO_FIELD('d', soil_depth_scale),
    EXO_FIELD('d', soil_water_capacity_m), EXO_FIELD('d', soil_heat_capacity), EXO_FIELD('d', land_albedo),
    EXO_FIELD('d', ocean_albedo), EXO_FIELD('d', snow_ice_albedo), EXO_FIELD('i', wet_soil),
    EXO_FIELD('i', filter), EXO_FIELD('i', initial_state), EXO_FIELD('i', kick),
    EXO_FIELD('s', seed), EXO_FIELD('l', clock_ms), EXO_FIELD('i', nspinit),
};
#undef EXO_FIELD

napi_value config_object(napi_env env, const exo_config& c) {
  napi_value obj;
  NAPI_OK(napi_create_object(env, &obj));
  for (const Field& f : kConfigFields) {
    const char* at = reinterpret_cast<const char*>(&c) + f.offset;
    if (f.kind == 's') {
      napi_value arr;
      NAPI_OK(napi_create_array_with_length(env, 8, &arr));
      for (uint32_t j = 0; j < 8; j++)
        NAPI_OK(napi_set_element(env, arr, j, number(env, reinterpret_cast<const int32_t*>(at)[j])));
      NAPI_OK(napi_set_named_property(env, obj, f.name, arr));
      continue;
    }
    double x = f.kind == 'i' ? *reinterpret_cast<const int*>(at)
               : f.kind == 'd' ? *reinterpret_cast<const double*>(at)
               : f.kind == 'l' ? (double)*reinterpret_cast<const int64_t*>(at)
                               : (double)*reinterpret_cast<const uint64_t*>(at);
    NAPI_OK(napi_set_named_property(env, obj, f.name, number(env, x)));
  }
  return obj;
}

napi_value config_default(napi_env env, napi_callback_info) {
  exo_config c;
  exo_config_default(&c);
  return config_object(env, c);
}

std::string string_of(napi_env env, napi_value v) {
  size_t len = 0;
  if (napi_get_value_string_utf8(env, v, nullptr, 0, &len) != napi_ok) return std::string();
  std::string s(len, '\0');
  napi_get_value_string_utf8(env, v, &s[0], len + 1, &len);
  return s;
}

bool config_from(napi_env env, napi_value obj, exo_config& c);

// configFromNamelists(truncation, levels, {name: text}, base?) -> config:
// exo_config_from_namelists; the fields no namelist carries come from
// `base` (an exo_config object over the defaults) or the defaults.
napi_value config_from_namelists(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 3)) return nullptr;
  exo_config c;
  napi_valuetype base_type = napi_undefined;
  if (a.n > 3) napi_typeof(env, a.v[3], &base_type);
  if (base_type == napi_object) {
    if (!config_from(env, a.v[3], c)) return nullptr;
  } else {
    exo_config_default(&c);
  }
  napi_value keys;
  uint32_t n = 0;
  NAPI_OK(napi_get_property_names(env, a.v[2], &keys));
  NAPI_OK(napi_get_array_length(env, keys, &n));
  std::vector<std::string> names(n), texts(n);
  std::vector<const char*> np(n), tp(n);
  for (uint32_t i = 0; i < n; i++) {
    napi_value k, v;
    NAPI_OK(napi_get_element(env, keys, i, &k));
    NAPI_OK(napi_get_property(env, a.v[2], k, &v));
    napi_valuetype t;
    NAPI_OK(napi_typeof(env, v, &t));
    names[i] = string_of(env, k);
    if (t != napi_string) {
      napi_throw_type_error(env, nullptr, (names[i] + " must be the text of a namelist file").c_str());
      return nullptr;
    }
    texts[i] = string_of(env, v);
    np[i] = names[i].c_str();
    tp[i] = texts[i].c_str();
  }
  EXO_OK_OR_THROW(exo_config_from_namelists(int_of(env, a.v[0]), int_of(env, a.v[1]), np.data(), tp.data(),
                                            (int)n, &c));
  return config_object(env, c);
}

// Defaults overlaid with the object's numeric properties; unknown keys throw.
bool config_from(napi_env env, napi_value obj, exo_config& c) {
  exo_config_default(&c);
  napi_value keys;
  uint32_t n = 0;
  if (napi_get_property_names(env, obj, &keys) != napi_ok || napi_get_array_length(env, keys, &n) != napi_ok)
    return false;
  for (uint32_t i = 0; i < n; i++) {
    napi_value k, v;
    napi_get_element(env, keys, i, &k);
    char name[64] = {0};
    size_t len = 0;
    napi_get_value_string_utf8(env, k, name, sizeof(name), &len);
    const Field* f = nullptr;
    for (const Field& x : kConfigFields)
      if (std::strcmp(x.name, name) == 0) f = &x;
    if (!f) {
      napi_throw_error(env, nullptr, (std::string("unknown config field ") + name).c_str());
      return false;
    }
    napi_get_property(env, obj, k, &v);
    char* at = reinterpret_cast<char*>(&c) + f->offset;
    if (f->kind == 's') {
      bool is_array = false;
      uint32_t count = 0;
      if (napi_is_array(env, v, &is_array) != napi_ok || !is_array || napi_get_array_length(env, v, &count) != napi_ok ||
          count > 8) {
        napi_throw_type_error(env, nullptr, (std::string(name) + " must be an array of at most 8 numbers").c_str());
        return false;
      }
      int32_t* words = reinterpret_cast<int32_t*>(at);
      for (uint32_t j = 0; j < 8; j++) {
        double d = 0;
        if (j < count) {
          napi_value e;
          napi_get_element(env, v, j, &e);
          if (napi_get_value_double(env, e, &d) != napi_ok) {
            napi_throw_type_error(env, nullptr, (std::string(name) + " must be an array of at most 8 numbers").c_str());
            return false;
          }
        }
        // Fortran reads the list from the front and leaves the rest zero
        words[j] = (int32_t)(uint32_t)(int64_t)d;
      }
    } else if (f->kind == 'l') {
      double d = 0;
      if (napi_get_value_double(env, v, &d) != napi_ok) {
        napi_throw_type_error(env, nullptr, (std::string(name) + " must be a number").c_str());
        return false;
      }
      *reinterpret_cast<int64_t*>(at) = (int64_t)d;
    } else if (f->kind == 'u') {
      bool lossless = false;
      uint64_t x = 0;
      if (napi_get_value_bigint_uint64(env, v, &x, &lossless) != napi_ok) {
        double d = 0;
        napi_get_value_double(env, v, &d);
        x = (uint64_t)d;
      }
      *reinterpret_cast<uint64_t*>(at) = x;
    } else {
      double d = 0;
      if (napi_get_value_double(env, v, &d) != napi_ok) {
        napi_throw_type_error(env, nullptr, (std::string(name) + " must be a number").c_str());
        return false;
      }
      if (f->kind == 'i')
        *reinterpret_cast<int*>(at) = (int)d;
      else
        *reinterpret_cast<double*>(at) = d;
    }
  }
  return true;
}

napi_value create(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 2)) return nullptr;
  exo_engine* e = nullptr;
  EXO_OK_OR_THROW(exo_create(int_of(env, a.v[0]), int_of(env, a.v[1]), &e));
  napi_value r;
  NAPI_OK(napi_create_external(env, e, nullptr, nullptr, &r));
  return r;
}

napi_value destroy(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 1)) return nullptr;
  exo_engine* e = engine_of(env, a.v[0]);
  if (!e) return nullptr;
  exo_destroy(e);
  return nullptr;
}

napi_value add_member(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 3)) return nullptr;
  exo_engine* e = engine_of(env, a.v[0]);
  if (!e) return nullptr;
  exo_config c;
  if (!config_from(env, a.v[1], c)) return nullptr;
  int id = -1;
  EXO_OK_OR_THROW(exo_add_member(e, &c, int_of(env, a.v[2]), &id));
  return number(env, id);
}

// Gaussian latitudes of a truncation (0 for one that is not built).
int exo_latitudes_of(int truncation) { return truncation == 21 ? 32 : truncation == 42 ? 64 : truncation == 63 ? 96 : 0; }

// addMemberSurface(engine, config, precision, [{code, months, values:
// Float64Array}]) -> id: exo_add_member_surface (the planet's .sra fields).
napi_value add_member_surface(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 4)) return nullptr;
  exo_engine* e = engine_of(env, a.v[0]);
  if (!e) return nullptr;
  exo_config c;
  if (!config_from(env, a.v[1], c)) return nullptr;
  uint32_t n = 0;
  NAPI_OK(napi_get_array_length(env, a.v[3], &n));
  std::vector<exo_surface_field> fields(n);
  for (uint32_t i = 0; i < n; i++) {
    napi_value f, v;
    NAPI_OK(napi_get_element(env, a.v[3], i, &f));
    fields[i].code = (int)number_of(env, f, "code");
    fields[i].months = (int)number_of(env, f, "months");
    NAPI_OK(napi_get_named_property(env, f, "values", &v));
    size_t count = 0;
    fields[i].values = static_cast<const double*>(typed_data(env, v, napi_float64_array, count));
    if (!fields[i].values) return nullptr;
    if (count != (size_t)fields[i].months * 2 * exo_latitudes_of(c.truncation) * exo_latitudes_of(c.truncation)) {
      napi_throw_range_error(env, nullptr, "a surface field does not match the grid");
      return nullptr;
    }
  }
  int id = -1;
  EXO_OK_OR_THROW(exo_add_member_surface(e, &c, int_of(env, a.v[2]), fields.data(), (int)n, &id));
  return number(env, id);
}

// atmosphereResolve([9 partial pressures, NaN for null], pressureAndGasConstant,
// pressureBar, gasConstant) -> {pressureBar, gasConstant, co2Ppm}:
// GCM_ATMOSPHERE.resolve.
napi_value atmosphere_resolve(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 4)) return nullptr;
  double partial[9];
  for (uint32_t i = 0; i < 9; i++) {
    napi_value v;
    NAPI_OK(napi_get_element(env, a.v[0], i, &v));
    if (napi_get_value_double(env, v, &partial[i]) != napi_ok) {
      napi_throw_type_error(env, nullptr, "partial pressures must be numbers");
      return nullptr;
    }
  }
  bool mode = false;
  napi_get_value_bool(env, a.v[1], &mode);
  double bar = 0, gas = 0, out[3];
  napi_get_value_double(env, a.v[2], &bar);
  napi_get_value_double(env, a.v[3], &gas);
  EXO_OK_OR_THROW(exo_atmosphere_resolve(partial, mode ? 1 : 0, bar, gas, &out[0], &out[1], &out[2]));
  napi_value obj;
  NAPI_OK(napi_create_object(env, &obj));
  NAPI_OK(napi_set_named_property(env, obj, "pressureBar", number(env, out[0])));
  NAPI_OK(napi_set_named_property(env, obj, "gasConstant", number(env, out[1])));
  NAPI_OK(napi_set_named_property(env, obj, "co2Ppm", number(env, out[2])));
  return obj;
}

napi_value remove_member(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 2)) return nullptr;
  exo_engine* e = engine_of(env, a.v[0]);
  if (!e) return nullptr;
  EXO_OK_OR_THROW(exo_remove_member(e, int_of(env, a.v[1])));
  return nullptr;
}

napi_value step(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 2)) return nullptr;
  exo_engine* e = engine_of(env, a.v[0]);
  if (!e) return nullptr;
  EXO_OK_OR_THROW(exo_step(e, int_of(env, a.v[1])));
  return nullptr;
}

napi_value sync(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 1)) return nullptr;
  exo_engine* e = engine_of(env, a.v[0]);
  if (!e) return nullptr;
  EXO_OK_OR_THROW(exo_sync(e));
  return nullptr;
}

napi_value set_publish_every(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 2)) return nullptr;
  exo_engine* e = engine_of(env, a.v[0]);
  if (!e) return nullptr;
  EXO_OK_OR_THROW(exo_set_publish_every(e, int_of(env, a.v[1])));
  return nullptr;
}

napi_value member_step(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 2)) return nullptr;
  exo_engine* e = engine_of(env, a.v[0]);
  if (!e) return nullptr;
  int64_t s = 0;
  EXO_OK_OR_THROW(exo_member_step(e, int_of(env, a.v[1]), &s));
  return number(env, (double)s);
}

napi_value member_date(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 2)) return nullptr;
  exo_engine* e = engine_of(env, a.v[0]);
  if (!e) return nullptr;
  int d[5];
  EXO_OK_OR_THROW(exo_member_date(e, int_of(env, a.v[1]), d));
  napi_value obj;
  NAPI_OK(napi_create_object(env, &obj));
  const char* names[5] = {"year", "month", "day", "hour", "minute"};
  for (int i = 0; i < 5; i++) NAPI_OK(napi_set_named_property(env, obj, names[i], number(env, d[i])));
  return obj;
}

napi_value dims(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 2)) return nullptr;
  exo_engine* e = engine_of(env, a.v[0]);
  if (!e) return nullptr;
  int d[4];
  EXO_OK_OR_THROW(exo_dims(e, int_of(env, a.v[1]), d));
  napi_value arr;
  NAPI_OK(napi_create_array_with_length(env, 4, &arr));
  for (uint32_t i = 0; i < 4; i++) NAPI_OK(napi_set_element(env, arr, i, number(env, d[i])));
  return arr;
}

napi_value read_field(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 3)) return nullptr;
  exo_engine* e = engine_of(env, a.v[0]);
  if (!e) return nullptr;
  char name[64] = {0};
  size_t len = 0;
  napi_get_value_string_utf8(env, a.v[2], name, sizeof(name), &len);
  size_t off = 0, n = 0;
  const int id = int_of(env, a.v[1]);
  EXO_OK_OR_THROW(exo_slot(e, id, name, &off, &n));
  std::vector<float> out(n);
  EXO_OK_OR_THROW(exo_read_field(e, id, name, out.data(), n));
  return float_array(env, out);
}

napi_value read_snapshot(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 2)) return nullptr;
  exo_engine* e = engine_of(env, a.v[0]);
  if (!e) return nullptr;
  const int id = int_of(env, a.v[1]);
  size_t off = 0, n = 0;
  EXO_OK_OR_THROW(exo_slot(e, id, "snapshot", &off, &n));
  std::vector<float> out(n);
  EXO_OK_OR_THROW(exo_read_snapshot(e, id, out.data(), n));
  return float_array(env, out);
}

napi_value global_means(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 2)) return nullptr;
  exo_engine* e = engine_of(env, a.v[0]);
  if (!e) return nullptr;
  double ts = 0, toa = 0;
  EXO_OK_OR_THROW(exo_global_means(e, int_of(env, a.v[1]), &ts, &toa));
  napi_value obj;
  NAPI_OK(napi_create_object(env, &obj));
  NAPI_OK(napi_set_named_property(env, obj, "surfaceTemperatureK", number(env, ts)));
  NAPI_OK(napi_set_named_property(env, obj, "toaNetWm2", number(env, toa)));
  return obj;
}

napi_value checkpoint(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 2)) return nullptr;
  exo_engine* e = engine_of(env, a.v[0]);
  if (!e) return nullptr;
  const int id = int_of(env, a.v[1]);
  size_t bytes = 0;
  EXO_OK_OR_THROW(exo_checkpoint_size(e, id, &bytes));
  napi_value buf;
  void* p = nullptr;
  NAPI_OK(napi_create_arraybuffer(env, bytes, &p, &buf));
  EXO_OK_OR_THROW(exo_checkpoint_save(e, id, p, bytes));
  return buf;
}

napi_value restore(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 3)) return nullptr;
  exo_engine* e = engine_of(env, a.v[0]);
  if (!e) return nullptr;
  void* p = nullptr;
  size_t bytes = 0;
  if (napi_get_arraybuffer_info(env, a.v[2], &p, &bytes) != napi_ok) {
    napi_throw_type_error(env, nullptr, "restart must be an ArrayBuffer");
    return nullptr;
  }
  // restore(engine, id, restart, previousConfig): with the configuration
  // the restart was saved under, GCM_CONTINUATION.fields' resets apply
  if (a.n > 3 && !is_nullish(env, a.v[3])) {
    exo_config previous;
    if (!config_from(env, a.v[3], previous)) return nullptr;
    EXO_OK_OR_THROW(exo_checkpoint_continue(e, int_of(env, a.v[1]), p, bytes, &previous));
    return nullptr;
  }
  EXO_OK_OR_THROW(exo_checkpoint_restore(e, int_of(env, a.v[1]), p, bytes));
  return nullptr;
}

// ---- averages and balance (GCM_AVERAGES / GCM_BALANCE) ----------------

napi_value set_averaging(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 3)) return nullptr;
  exo_engine* e = engine_of(env, a.v[0]);
  if (!e) return nullptr;
  bool on = false;
  napi_get_value_bool(env, a.v[2], &on);
  EXO_OK_OR_THROW(exo_set_averaging(e, int_of(env, a.v[1]), on ? 1 : 0));
  return nullptr;
}

napi_value month_count(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 2)) return nullptr;
  exo_engine* e = engine_of(env, a.v[0]);
  if (!e) return nullptr;
  int n = 0;
  EXO_OK_OR_THROW(exo_month_count(e, int_of(env, a.v[1]), &n));
  return number(env, n);
}

bool set_month_info(napi_env env, napi_value obj, const exo_month_info& m) {
  return napi_set_named_property(env, obj, "year", number(env, m.year)) == napi_ok &&
         napi_set_named_property(env, obj, "month", number(env, m.month)) == napi_ok &&
         napi_set_named_property(env, obj, "samples", number(env, m.samples)) == napi_ok &&
         napi_set_named_property(env, obj, "step", number(env, (double)m.step)) == napi_ok;
}

exo_month_info month_info_of(napi_env env, napi_value obj) {
  return exo_month_info{(int)number_of(env, obj, "year"), (int)number_of(env, obj, "month"),
                        (int)number_of(env, obj, "samples"), (int64_t)number_of(env, obj, "step")};
}

size_t snapshot_floats(napi_env env, exo_engine* e, int id) {
  size_t off = 0, n = 0;
  if (exo_slot(e, id, "snapshot", &off, &n) != EXO_OK) return 0;
  return n;
}

// {year, month, samples, step, values: Float32Array} of closed month index;
// readMonth(engine, id, index, false) leaves the values out.
napi_value read_month(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 3)) return nullptr;
  exo_engine* e = engine_of(env, a.v[0]);
  if (!e) return nullptr;
  const int id = int_of(env, a.v[1]);
  bool with_values = true;
  if (a.n > 3) napi_get_value_bool(env, a.v[3], &with_values);
  std::vector<float> out(with_values ? snapshot_floats(env, e, id) : 0);
  exo_month_info m{};
  EXO_OK_OR_THROW(exo_read_month(e, id, int_of(env, a.v[2]), &m, with_values ? out.data() : nullptr, out.size()));
  napi_value obj;
  NAPI_OK(napi_create_object(env, &obj));
  if (!set_month_info(env, obj, m)) return nullptr;
  if (with_values) NAPI_OK(napi_set_named_property(env, obj, "values", float_array(env, out)));
  return obj;
}

napi_value read_annual(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 2)) return nullptr;
  exo_engine* e = engine_of(env, a.v[0]);
  if (!e) return nullptr;
  const int id = int_of(env, a.v[1]);
  std::vector<float> out(snapshot_floats(env, e, id));
  EXO_OK_OR_THROW(exo_read_annual(e, id, out.data(), out.size()));
  return float_array(env, out);
}

napi_value balance_object(napi_env env, const exo_yearly_balance& y) {
  napi_value obj;
  if (napi_create_object(env, &obj) != napi_ok) return nullptr;
  napi_set_named_property(env, obj, "year", number(env, y.year));
  napi_set_named_property(env, obj, "toaNetRadiation", number(env, y.toa_net_radiation));
  napi_set_named_property(env, obj, "surfaceNetHeat", number(env, y.surface_net_heat));
  napi_set_named_property(env, obj, "surfaceTemperature", number(env, y.surface_temperature));
  return obj;
}

// [GcmYearlyBalance] oldest first.
napi_value yearly_balances(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 2)) return nullptr;
  exo_engine* e = engine_of(env, a.v[0]);
  if (!e) return nullptr;
  const int id = int_of(env, a.v[1]);
  int n = 0;
  EXO_OK_OR_THROW(exo_yearly_balances(e, id, nullptr, 0, &n));
  std::vector<exo_yearly_balance> years(n);
  if (n > 0) EXO_OK_OR_THROW(exo_yearly_balances(e, id, years.data(), n, &n));
  napi_value arr;
  NAPI_OK(napi_create_array_with_length(env, years.size(), &arr));
  for (uint32_t i = 0; i < years.size(); i++) NAPI_OK(napi_set_element(env, arr, i, balance_object(env, years[i])));
  return arr;
}

// null, or {year, month, samples, step, sum: Float64Array} of the open month.
napi_value read_accumulator(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 2)) return nullptr;
  exo_engine* e = engine_of(env, a.v[0]);
  if (!e) return nullptr;
  const int id = int_of(env, a.v[1]);
  exo_month_info m{};
  EXO_OK_OR_THROW(exo_read_accumulator(e, id, &m, nullptr, 0));
  napi_value obj;
  if (m.samples == 0) {
    NAPI_OK(napi_get_null(env, &obj));
    return obj;
  }
  std::vector<double> sum(snapshot_floats(env, e, id));
  EXO_OK_OR_THROW(exo_read_accumulator(e, id, &m, sum.data(), sum.size()));
  NAPI_OK(napi_create_object(env, &obj));
  if (!set_month_info(env, obj, m)) return nullptr;
  NAPI_OK(napi_set_named_property(env, obj, "sum", double_array(env, sum)));
  return obj;
}

// loadAverager(engine, id, months, years, current): months
// [{year, month, samples, step, values: Float32Array}], years
// [GcmYearlyBalance], current null or {year, month, samples, step,
// sum: Float64Array}.
napi_value load_averager(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 5)) return nullptr;
  exo_engine* e = engine_of(env, a.v[0]);
  if (!e) return nullptr;
  const int id = int_of(env, a.v[1]);
  const size_t floats = snapshot_floats(env, e, id);
  uint32_t n_months = 0, n_years = 0;
  NAPI_OK(napi_get_array_length(env, a.v[2], &n_months));
  NAPI_OK(napi_get_array_length(env, a.v[3], &n_years));
  std::vector<exo_month_info> months(n_months);
  std::vector<float> values((size_t)n_months * floats);
  for (uint32_t i = 0; i < n_months; i++) {
    napi_value m, v;
    NAPI_OK(napi_get_element(env, a.v[2], i, &m));
    months[i] = month_info_of(env, m);
    NAPI_OK(napi_get_named_property(env, m, "values", &v));
    size_t length = 0;
    const void* p = typed_data(env, v, napi_float32_array, length);
    if (!p) return nullptr;
    if (length != floats) {
      napi_throw_range_error(env, nullptr, ("a month holds " + std::to_string(floats) + " values").c_str());
      return nullptr;
    }
    std::memcpy(values.data() + (size_t)i * floats, p, floats * sizeof(float));
  }
  std::vector<exo_yearly_balance> years(n_years);
  for (uint32_t i = 0; i < n_years; i++) {
    napi_value y;
    NAPI_OK(napi_get_element(env, a.v[3], i, &y));
    years[i] = exo_yearly_balance{(int)number_of(env, y, "year"), number_of(env, y, "toaNetRadiation"),
                                  number_of(env, y, "surfaceNetHeat"), number_of(env, y, "surfaceTemperature")};
  }
  exo_month_info current{};
  const double* sum = nullptr;
  if (!is_nullish(env, a.v[4])) {
    current = month_info_of(env, a.v[4]);
    napi_value v;
    NAPI_OK(napi_get_named_property(env, a.v[4], "sum", &v));
    size_t length = 0;
    sum = static_cast<const double*>(typed_data(env, v, napi_float64_array, length));
    if (!sum) return nullptr;
    if (length != floats) {
      napi_throw_range_error(env, nullptr, ("a month sum holds " + std::to_string(floats) + " values").c_str());
      return nullptr;
    }
  }
  EXO_OK_OR_THROW(exo_load_averager(e, id, months.data(), values.data(), (int)n_months, years.data(), (int)n_years,
                                    &current, sum, floats));
  return nullptr;
}

// null, or {year, month, samples, step, years, values: Float32Array}.
napi_value read_climatology(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 3)) return nullptr;
  exo_engine* e = engine_of(env, a.v[0]);
  if (!e) return nullptr;
  const int id = int_of(env, a.v[1]);
  std::vector<float> out(snapshot_floats(env, e, id));
  exo_month_info m{};
  int years = 0;
  EXO_OK_OR_THROW(exo_read_climatology(e, id, int_of(env, a.v[2]), &m, out.data(), out.size(), &years));
  napi_value obj;
  if (years == 0) {
    NAPI_OK(napi_get_null(env, &obj));
    return obj;
  }
  NAPI_OK(napi_create_object(env, &obj));
  if (!set_month_info(env, obj, m)) return nullptr;
  NAPI_OK(napi_set_named_property(env, obj, "years", number(env, years)));
  NAPI_OK(napi_set_named_property(env, obj, "values", float_array(env, out)));
  return obj;
}

// GCM_BALANCE.assess(years) -> GcmBalanceAssessment.
napi_value balance_assess(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 1)) return nullptr;
  uint32_t n = 0;
  NAPI_OK(napi_get_array_length(env, a.v[0], &n));
  std::vector<exo_yearly_balance> years(n);
  for (uint32_t i = 0; i < n; i++) {
    napi_value y;
    NAPI_OK(napi_get_element(env, a.v[0], i, &y));
    years[i] = exo_yearly_balance{(int)number_of(env, y, "year"), number_of(env, y, "toaNetRadiation"),
                                  number_of(env, y, "surfaceNetHeat"), number_of(env, y, "surfaceTemperature")};
  }
  exo_balance_assessment out{};
  EXO_OK_OR_THROW(exo_balance_assess(years.data(), (int)n, &out));
  const char* status = out.status == EXO_BALANCE_BALANCED   ? "balanced"
                       : out.status == EXO_BALANCE_DRIFTING ? "drifting"
                                                            : "spinning-up";
  napi_value obj, s;
  NAPI_OK(napi_create_object(env, &obj));
  NAPI_OK(napi_create_string_utf8(env, status, NAPI_AUTO_LENGTH, &s));
  NAPI_OK(napi_set_named_property(env, obj, "status", s));
  NAPI_OK(napi_set_named_property(env, obj, "yearsNeeded", number(env, out.years_needed)));
  NAPI_OK(napi_set_named_property(env, obj, "toaDriftPerYear", number(env, out.toa_drift_per_year)));
  NAPI_OK(napi_set_named_property(env, obj, "surfaceDriftPerYear", number(env, out.surface_drift_per_year)));
  return obj;
}

// ---- boundary data (GCM_BOUNDARY) ----------------------------------------

// boundaryFromTerrain(engine, {width, height, elevationM: Float32Array,
// land: Float32Array}, latitudes, gravityMs2) -> {landFraction,
// geopotential} (Float64Array each): GCM_BOUNDARY.fromTerrain.
napi_value boundary_from_terrain(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 4)) return nullptr;
  exo_engine* e = engine_of(env, a.v[0]);
  if (!e) return nullptr;
  const int width = (int)number_of(env, a.v[1], "width"), height = (int)number_of(env, a.v[1], "height");
  napi_value ev, lv;
  NAPI_OK(napi_get_named_property(env, a.v[1], "elevationM", &ev));
  NAPI_OK(napi_get_named_property(env, a.v[1], "land", &lv));
  size_t ne = 0, nl = 0;
  const float* elev = static_cast<const float*>(typed_data(env, ev, napi_float32_array, ne));
  if (!elev) return nullptr;
  const float* land = static_cast<const float*>(typed_data(env, lv, napi_float32_array, nl));
  if (!land) return nullptr;
  if (ne != (size_t)width * height || nl != ne) {
    napi_throw_error(env, nullptr, "Terrain arrays do not match the image dimensions");
    return nullptr;
  }
  const int nlat = int_of(env, a.v[2]);
  double gravity = 0;
  napi_get_value_double(env, a.v[3], &gravity);
  std::vector<double> fraction(2 * (size_t)nlat * nlat), geo(fraction.size());
  EXO_OK_OR_THROW(exo_boundary_from_terrain(e, elev, land, width, height, nlat, gravity, fraction.data(), geo.data()));
  napi_value obj;
  NAPI_OK(napi_create_object(env, &obj));
  NAPI_OK(napi_set_named_property(env, obj, "landFraction", double_array(env, fraction)));
  NAPI_OK(napi_set_named_property(env, obj, "geopotential", double_array(env, geo)));
  return obj;
}

// boundarySra(code, latitudes, values: Float64Array) -> string (toSra).
napi_value boundary_sra(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 3)) return nullptr;
  size_t n = 0;
  const double* values = static_cast<const double*>(typed_data(env, a.v[2], napi_float64_array, n));
  if (!values) return nullptr;
  const int nlat = int_of(env, a.v[1]);
  if (n != 2 * (size_t)nlat * nlat) {
    napi_throw_range_error(env, nullptr, "the field does not match the grid");
    return nullptr;
  }
  size_t len = 0;
  EXO_OK_OR_THROW(exo_boundary_sra(int_of(env, a.v[0]), nlat, values, nullptr, 0, &len));
  std::string text(len, '\0');
  EXO_OK_OR_THROW(exo_boundary_sra(int_of(env, a.v[0]), nlat, values, text.data(), len, &len));
  napi_value out;
  NAPI_OK(napi_create_string_utf8(env, text.data(), len, &out));
  return out;
}

napi_value init(napi_env env, napi_value exports) {
  const std::pair<const char*, napi_callback> fns[] = {
      {"configDefault", config_default}, {"configFromNamelists", config_from_namelists}, {"create", create},
      {"destroy", destroy},              {"addMember", add_member},
      {"removeMember", remove_member},   {"step", step},
      {"sync", sync},                    {"setPublishEvery", set_publish_every},
      {"memberStep", member_step},       {"memberDate", member_date},
      {"dims", dims},                    {"readField", read_field},
      {"readSnapshot", read_snapshot},   {"globalMeans", global_means},
      {"checkpoint", checkpoint},        {"restore", restore},
      {"setAveraging", set_averaging},   {"monthCount", month_count},
      {"readMonth", read_month},         {"readAnnual", read_annual},
      {"yearlyBalances", yearly_balances}, {"readAccumulator", read_accumulator},
      {"loadAverager", load_averager},   {"readClimatology", read_climatology},
      {"balanceAssess", balance_assess},
      {"boundaryFromTerrain", boundary_from_terrain}, {"boundarySra", boundary_sra},
      {"addMemberSurface", add_member_surface}, {"atmosphereResolve", atmosphere_resolve},
  };
  for (const auto& f : fns) {
    napi_value fn;
    if (napi_create_function(env, f.first, NAPI_AUTO_LENGTH, f.second, nullptr, &fn) != napi_ok) return nullptr;
    if (napi_set_named_property(env, exports, f.first, fn) != napi_ok) return nullptr;
  }
  napi_value v;
  napi_create_int32(env, EXO_PRECISION_FP32, &v);
  napi_set_named_property(env, exports, "PRECISION_FP32", v);
  napi_create_int32(env, EXO_PRECISION_SPLIT, &v);
  napi_set_named_property(env, exports, "PRECISION_SPLIT", v);
  return exports;
}

}  // namespace

NAPI_MODULE(NODE_GYP_MODULE_NAME, init)
