// soup_napi.cc -- the Node binding of libsoup_hip (N-API): what the reference's TypeScript host
// (soup/src/sim.ts, soup/src/soup-types.ts:181-442) calls in place of its WebGPU backend.
//
// The host keeps createSoup's resolution (sim.ts:96-278) and hands this module the packed
// SoupConfig bytes and the initial state's typed arrays; every entry point is a thin call into the
// C-ABI of include/soup_abi.h. A nonzero status becomes a JS Error carrying soup_last_error, whose
// messages keep the reference's own stems ("createSoup: the starting composition is empty",
// "neighbour grid: ...", "Verlet list overflow", ...).
//
// Exports:
//   version() -> string                      soup_version
//   configBytes() -> number                  sizeof(SoupConfig): the packer's layout check
//   create(config: ArrayBuffer, init: {positions, velocities, bondSlots, centerLink, centerHeldSteps,
//          frozen, bondRng, thermoRng, charges, events, desorbEvents (typed arrays), globalStep,
//          nextSweepAt, protonationRng (numbers)}) -> handle (external; destroy() or the GC frees it)
//   destroy(h)
//   step(h, n) -> Promise<void>              soup_step on a worker thread (the event loop stays free);
//                                            until it settles, every other call on that handle throws
//                                            "soup_napi: a step is still running on this system"
//   stepSync(h, n)
//   read(h, field, typedArray) -> typedArray soup_read into the caller's array (exact byte size)
//   write(h, field, typedArray)              soup_write
//   fieldBytes(h, field) -> number
//   forces(h, mode)                          soup_forces (0 rebuild, 1 brute force, 2 current lists)
//   relax(h, iterations, d0)                 soup_relax (relaxColdStart)
//   globalStep(h) -> number
//   protonationSweep(h) -> {attempts, accepted, heads, deprotonated, alpha, dEsMeanKT}
#define NAPI_VERSION 6
#define NODE_GYP_MODULE_NAME soup_napi
#include <node_api.h>

#include <stdint.h>
#include <string.h>

#include <string>

#include "soup_abi.h"

namespace {

struct Handle {
  SoupHandle* h = nullptr;
  // set while an asynchronous step runs on a worker thread; read and written on the JS thread only
  // (Step sets it, step_complete clears it), so no lock is needed
  bool busy = false;
};

napi_value throw_error(napi_env env, const std::string& msg) {
  napi_throw_error(env, nullptr, msg.c_str());
  return nullptr;
}

napi_value throw_soup(napi_env env, const SoupHandle* h, int code) {
  const char* m = soup_last_error(h);
  std::string msg = (m && *m) ? std::string(m) : std::string("libsoup_hip call failed");
  napi_value err, msgv, codev;
  napi_create_string_utf8(env, msg.c_str(), NAPI_AUTO_LENGTH, &msgv);
  napi_create_error(env, nullptr, msgv, &err);
  napi_create_int32(env, code, &codev);
  napi_set_named_property(env, err, "code", codev);
  napi_throw(env, err);
  return nullptr;
}

bool get_args(napi_env env, napi_callback_info info, size_t want, napi_value* argv) {
  size_t argc = want;
  if (napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr) != napi_ok || argc < want) {
    throw_error(env, "soup_napi: expected " + std::to_string(want) + " arguments");
    return false;
  }
  return true;
}

// the live, idle handle behind v (nullptr with a thrown error otherwise): the library's calls on one
// system do not overlap, as the reference's host awaits each call before the next
Handle* handle_of(napi_env env, napi_value v) {
  void* p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p || !static_cast<Handle*>(p)->h) {
    throw_error(env, "soup_napi: not a live soup handle");
    return nullptr;
  }
  if (static_cast<Handle*>(p)->busy) {
    throw_error(env, "soup_napi: a step is still running on this system; await it before the next call");
    return nullptr;
  }
  return static_cast<Handle*>(p);
}

bool get_u32(napi_env env, napi_value v, uint32_t* out, const char* what) {
  if (napi_get_value_uint32(env, v, out) != napi_ok) {
    throw_error(env, std::string("soup_napi: ") + what + " must be a number");
    return false;
  }
  return true;
}

// the data of typed-array property `name` of obj (nullptr with a thrown error when absent)
const void* typed_prop(napi_env env, napi_value obj, const char* name, size_t* bytes) {
  napi_value v;
  bool is_ta = false;
  if (napi_get_named_property(env, obj, name, &v) != napi_ok || napi_is_typedarray(env, v, &is_ta) != napi_ok || !is_ta) {
    throw_error(env, std::string("soup_napi: init.") + name + " must be a typed array");
    return nullptr;
  }
  napi_typedarray_type t;
  size_t len, off;
  void* data;
  napi_value ab;
  napi_get_typedarray_info(env, v, &t, &len, &data, &ab, &off);
  size_t el = (t == napi_float64_array || t == napi_bigint64_array || t == napi_biguint64_array) ? 8
              : (t == napi_int16_array || t == napi_uint16_array)                               ? 2
              : (t == napi_int8_array || t == napi_uint8_array || t == napi_uint8_clamped_array) ? 1
                                                                                                 : 4;
  if (bytes) *bytes = len * el;
  return data;
}

double num_prop(napi_env env, napi_value obj, const char* name, double dflt) {
  bool has = false;
  napi_has_named_property(env, obj, name, &has);
  if (!has) return dflt;
  napi_value v;
  double d = dflt;
  napi_get_named_property(env, obj, name, &v);
  if (napi_get_value_double(env, v, &d) != napi_ok) return dflt;
  return d;
}

void finalize_handle(napi_env, void* p, void*) {
  Handle* hd = static_cast<Handle*>(p);
  if (hd->h) soup_destroy(hd->h);
  delete hd;
}

napi_value Version(napi_env env, napi_callback_info) {
  napi_value v;
  napi_create_string_utf8(env, soup_version(), NAPI_AUTO_LENGTH, &v);
  return v;
}

napi_value ConfigBytes(napi_env env, napi_callback_info) {
  napi_value v;
  napi_create_uint32(env, (uint32_t)sizeof(SoupConfig), &v);
  return v;
}

napi_value Create(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  bool is_ab = false;
  napi_is_arraybuffer(env, argv[0], &is_ab);
  if (!is_ab) return throw_error(env, "soup_napi: create(config, init): config must be an ArrayBuffer");
  void* cfg = nullptr;
  size_t cfg_len = 0;
  napi_get_arraybuffer_info(env, argv[0], &cfg, &cfg_len);
  if (cfg_len != sizeof(SoupConfig))
    return throw_error(env, "soup_napi: the config holds " + std::to_string(cfg_len) + " bytes, SoupConfig " +
                                std::to_string(sizeof(SoupConfig)));
  const SoupConfig* c = static_cast<const SoupConfig*>(cfg);
  SoupInitState init{};
  const char* names[11] = {"positions", "velocities", "bondSlots", "centerLink", "centerHeldSteps", "frozen",
                           "bondRng",   "thermoRng",  "charges",   "events",     "desorbEvents"};
  // bytes each array must hold for the config's n particles (the C-ABI takes bare pointers)
  const size_t n = c->n;
  const size_t need[11] = {16 * n, 16 * n, 12 * n, 4 * n, 4 * n, 4 * n, 4 * n, 4 * n, 4 * n, 8 * (size_t)c->n_rules, 8};
  const void* ptr[11];
  for (int k = 0; k < 11; ++k) {
    size_t bytes = 0;
    ptr[k] = typed_prop(env, argv[1], names[k], &bytes);
    if (!ptr[k] && need[k]) return nullptr;
    if (bytes < need[k])
      return throw_error(env, std::string("soup_napi: init.") + names[k] + " holds " + std::to_string(bytes) + " bytes, " +
                                  std::to_string(n) + " particles need " + std::to_string(need[k]));
  }
  init.positions = static_cast<const float*>(ptr[0]);
  init.velocities = static_cast<const float*>(ptr[1]);
  init.bond_slots = static_cast<const uint32_t*>(ptr[2]);
  init.center_link = static_cast<const uint32_t*>(ptr[3]);
  init.center_held = static_cast<const uint32_t*>(ptr[4]);
  init.frozen = static_cast<const uint32_t*>(ptr[5]);
  init.bond_rng = static_cast<const uint32_t*>(ptr[6]);
  init.thermo_rng = static_cast<const uint32_t*>(ptr[7]);
  init.charges = static_cast<const float*>(ptr[8]);
  init.events = static_cast<const uint32_t*>(ptr[9]);
  init.desorb_events = static_cast<const uint32_t*>(ptr[10]);
  init.global_step = (uint64_t)num_prop(env, argv[1], "globalStep", 0.0);
  const double nsa = num_prop(env, argv[1], "nextSweepAt", -1.0);
  init.next_sweep_at = nsa < 0.0 ? UINT64_MAX : (uint64_t)nsa;
  init.protonation_rng = (uint32_t)num_prop(env, argv[1], "protonationRng", 0.0);
  SoupHandle* h = nullptr;
  const int r = soup_create(c, &init, &h);
  if (r) return throw_soup(env, nullptr, r);
  Handle* hd = new Handle();
  hd->h = h;
  napi_value ext;
  napi_create_external(env, hd, finalize_handle, nullptr, &ext);
  return ext;
}

napi_value Destroy(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  Handle* hd = handle_of(env, argv[0]);
  if (!hd) return nullptr;
  soup_destroy(hd->h);
  hd->h = nullptr;
  return nullptr;
}

napi_value StepSync(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  Handle* hd = handle_of(env, argv[0]);
  uint32_t n;
  if (!hd || !get_u32(env, argv[1], &n, "step count")) return nullptr;
  const int r = soup_step(hd->h, n);
  if (r) return throw_soup(env, hd->h, r);
  return nullptr;
}

struct StepWork {
  Handle* hd;
  uint32_t n;
  int status = 0;
  napi_deferred deferred;
  napi_async_work work;
  napi_ref keep;  // keeps the handle's external alive while the step runs
};

void step_execute(napi_env, void* data) {
  StepWork* w = static_cast<StepWork*>(data);
  w->status = soup_step(w->hd->h, w->n);
}

void step_complete(napi_env env, napi_status, void* data) {
  StepWork* w = static_cast<StepWork*>(data);
  w->hd->busy = false;
  if (w->status == 0) {
    napi_value undef;
    napi_get_undefined(env, &undef);
    napi_resolve_deferred(env, w->deferred, undef);
  } else {
    const char* m = soup_last_error(w->hd->h);
    napi_value msgv, err, codev;
    napi_create_string_utf8(env, (m && *m) ? m : "libsoup_hip call failed", NAPI_AUTO_LENGTH, &msgv);
    napi_create_error(env, nullptr, msgv, &err);
    napi_create_int32(env, w->status, &codev);
    napi_set_named_property(env, err, "code", codev);
    napi_reject_deferred(env, w->deferred, err);
  }
  napi_delete_reference(env, w->keep);
  napi_delete_async_work(env, w->work);
  delete w;
}

napi_value Step(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  Handle* hd = handle_of(env, argv[0]);
  uint32_t n;
  if (!hd || !get_u32(env, argv[1], &n, "step count")) return nullptr;
  StepWork* w = new StepWork();
  w->hd = hd;
  w->n = n;
  hd->busy = true;
  napi_value promise, name;
  napi_create_promise(env, &w->deferred, &promise);
  napi_create_reference(env, argv[0], 1, &w->keep);
  napi_create_string_utf8(env, "soup_step", NAPI_AUTO_LENGTH, &name);
  napi_create_async_work(env, nullptr, name, step_execute, step_complete, w, &w->work);
  napi_queue_async_work(env, w->work);
  return promise;
}

napi_value Read(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  Handle* hd = handle_of(env, argv[0]);
  uint32_t field;
  if (!hd || !get_u32(env, argv[1], &field, "field")) return nullptr;
  bool is_ta = false;
  napi_is_typedarray(env, argv[2], &is_ta);
  if (!is_ta) return throw_error(env, "soup_napi: read(h, field, out): out must be a typed array");
  napi_typedarray_type t;
  size_t len, off;
  void* data;
  napi_value ab;
  napi_get_typedarray_info(env, argv[2], &t, &len, &data, &ab, &off);
  size_t bytes = 0;
  napi_get_arraybuffer_info(env, ab, nullptr, &bytes);
  const size_t el = (t == napi_float64_array || t == napi_bigint64_array || t == napi_biguint64_array) ? 8
                    : (t == napi_uint8_array || t == napi_int8_array || t == napi_uint8_clamped_array)  ? 1
                    : (t == napi_uint16_array || t == napi_int16_array)                                 ? 2
                                                                                                        : 4;
  const int r = soup_read(hd->h, (int)field, data, len * el);
  if (r) return throw_soup(env, hd->h, r);
  return argv[2];
}

napi_value Write(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  Handle* hd = handle_of(env, argv[0]);
  uint32_t field;
  if (!hd || !get_u32(env, argv[1], &field, "field")) return nullptr;
  bool is_ta = false;
  napi_is_typedarray(env, argv[2], &is_ta);
  if (!is_ta) return throw_error(env, "soup_napi: write(h, field, src): src must be a typed array");
  napi_typedarray_type t;
  size_t len, off;
  void* data;
  napi_value ab;
  napi_get_typedarray_info(env, argv[2], &t, &len, &data, &ab, &off);
  const size_t el = (t == napi_float64_array) ? 8 : (t == napi_uint8_array || t == napi_int8_array) ? 1 : 4;
  const int r = soup_write(hd->h, (int)field, data, len * el);
  if (r) return throw_soup(env, hd->h, r);
  return nullptr;
}

napi_value FieldBytes(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  Handle* hd = handle_of(env, argv[0]);
  uint32_t field;
  if (!hd || !get_u32(env, argv[1], &field, "field")) return nullptr;
  napi_value v;
  napi_create_double(env, (double)soup_field_bytes(hd->h, (int)field), &v);
  return v;
}

napi_value Forces(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  Handle* hd = handle_of(env, argv[0]);
  uint32_t mode;
  if (!hd || !get_u32(env, argv[1], &mode, "mode")) return nullptr;
  const int r = soup_forces(hd->h, (int)mode);
  if (r) return throw_soup(env, hd->h, r);
  return nullptr;
}

napi_value Relax(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  Handle* hd = handle_of(env, argv[0]);
  uint32_t it;
  double d0 = 0.0;
  if (!hd || !get_u32(env, argv[1], &it, "iterations")) return nullptr;
  napi_get_value_double(env, argv[2], &d0);
  const int r = soup_relax(hd->h, it, d0);
  if (r) return throw_soup(env, hd->h, r);
  return nullptr;
}

napi_value GlobalStep(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  Handle* hd = handle_of(env, argv[0]);
  if (!hd) return nullptr;
  uint64_t g = 0;
  const int r = soup_global_step(hd->h, &g);
  if (r) return throw_soup(env, hd->h, r);
  napi_value v;
  napi_create_double(env, (double)g, &v);
  return v;
}

napi_value ProtonationSweep(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  Handle* hd = handle_of(env, argv[0]);
  if (!hd) return nullptr;
  SoupSweepResult s{};
  const int r = soup_protonation_sweep(hd->h, &s);
  if (r) return throw_soup(env, hd->h, r);
  napi_value o, v;
  napi_create_object(env, &o);
  napi_create_uint32(env, s.attempts, &v);
  napi_set_named_property(env, o, "attempts", v);
  napi_create_uint32(env, s.accepted, &v);
  napi_set_named_property(env, o, "accepted", v);
  napi_create_uint32(env, s.heads, &v);
  napi_set_named_property(env, o, "heads", v);
  napi_create_uint32(env, s.deprotonated, &v);
  napi_set_named_property(env, o, "deprotonated", v);
  napi_create_double(env, s.alpha, &v);
  napi_set_named_property(env, o, "alpha", v);
  napi_create_double(env, s.dEsMeanKT, &v);
  napi_set_named_property(env, o, "dEsMeanKT", v);
  return o;
}

napi_value Init(napi_env env, napi_value exports) {
  const napi_property_descriptor props[] = {
      {"version", nullptr, Version, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"configBytes", nullptr, ConfigBytes, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"create", nullptr, Create, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"destroy", nullptr, Destroy, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"step", nullptr, Step, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"stepSync", nullptr, StepSync, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"read", nullptr, Read, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"write", nullptr, Write, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"fieldBytes", nullptr, FieldBytes, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"forces", nullptr, Forces, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"relax", nullptr, Relax, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"globalStep", nullptr, GlobalStep, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"protonationSweep", nullptr, ProtonationSweep, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
  };
  napi_define_properties(env, exports, sizeof(props) / sizeof(props[0]), props);
  return exports;
}

}  // namespace

NAPI_MODULE(NODE_GYP_MODULE_NAME, Init)
