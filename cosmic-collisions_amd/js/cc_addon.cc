// Node N-API addon over the C ABI of libcc_hip.so (include/cc_hip.h).
// Each export is one C-ABI call: typed arrays in and out, a thrown Error with
// the library's message on a negative return code.  The engine object
// (nbody.js) is the only intended caller.
#include <node_api.h>

#include <cstring>
#include <string>
#include <vector>

#include "cc_hip.h"

namespace {

#define NAPI_OK(env, call)                                         \
  do {                                                             \
    if ((call) != napi_ok) {                                       \
      napi_throw_error(env, nullptr, "N-API call failed: " #call); \
      return nullptr;                                              \
    }                                                              \
  } while (0)

napi_value throw_cc(napi_env env, cc_ctx* c, const char* what, int code) {
  std::string msg = std::string(what) + " failed (" + std::to_string(code) + "): " + cc_last_error(c);
  napi_throw_error(env, nullptr, msg.c_str());
  return nullptr;
}

struct Args {
  napi_env env;
  size_t argc = 12;
  napi_value argv[12];
  explicit Args(napi_env e, napi_callback_info info) : env(e) {
    napi_get_cb_info(e, info, &argc, argv, nullptr, nullptr);
  }
};

// A context handle is a JS external holding this box: destroy() frees the
// context at once and empties the box, the finalizer frees what is left.
struct Handle {
  cc_ctx* c;
};

cc_ctx* ctx_of(napi_env env, napi_value v) {
  void* p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p || !static_cast<Handle*>(p)->c) {
    napi_throw_type_error(env, nullptr, "expected a live cc context handle");
    return nullptr;
  }
  return static_cast<Handle*>(p)->c;
}

bool is_nullish(napi_env env, napi_value v) {
  napi_valuetype t;
  napi_typeof(env, v, &t);
  return t == napi_null || t == napi_undefined;
}

// typed array data pointer (nullptr for null/undefined); checks element type and minimum length
void* typed(napi_env env, napi_value v, napi_typedarray_type want, size_t minLen, bool* ok) {
  *ok = true;
  if (is_nullish(env, v)) return nullptr;
  bool isTA = false;
  napi_is_typedarray(env, v, &isTA);
  if (!isTA) { napi_throw_type_error(env, nullptr, "expected a typed array"); *ok = false; return nullptr; }
  napi_typedarray_type t;
  size_t len;
  void* data;
  napi_value ab;
  size_t off;
  napi_get_typedarray_info(env, v, &t, &len, &data, &ab, &off);
  if (t != want) { napi_throw_type_error(env, nullptr, "typed array of the wrong element type"); *ok = false; return nullptr; }
  if (len < minLen) { napi_throw_range_error(env, nullptr, "typed array too short"); *ok = false; return nullptr; }
  return data;
}

double num(napi_env env, napi_value v, double dflt = 0.0) {
  double d = dflt;
  if (!is_nullish(env, v)) napi_get_value_double(env, v, &d);
  return d;
}

// a JS array or typed array of numbers into doubles
bool nums(napi_env env, napi_value v, double* out, size_t n) {
  if (is_nullish(env, v)) { for (size_t i = 0; i < n; i++) out[i] = 0.0; return true; }
  for (uint32_t i = 0; i < n; i++) {
    napi_value e;
    if (napi_get_element(env, v, i, &e) != napi_ok) return false;
    if (napi_get_value_double(env, e, &out[i]) != napi_ok) return false;
  }
  return true;
}

napi_value prop(napi_env env, napi_value obj, const char* name) {
  napi_value v;
  napi_get_named_property(env, obj, name, &v);
  return v;
}

bool read_params(napi_env env, napi_value o, cc_params* p) {
  memset(p, 0, sizeof *p);
  p->N = (int)num(env, prop(env, o, "N"));
  p->W = (int)num(env, prop(env, o, "W"));
  p->H = (int)num(env, prop(env, o, "H"));
  p->G = (int)num(env, prop(env, o, "G"));
  p->m = num(env, prop(env, o, "m"));
  p->a = num(env, prop(env, o, "a"));
  p->cell = num(env, prop(env, o, "cell"));
  p->dt = num(env, prop(env, o, "dt"));
  p->k = num(env, prop(env, o, "k"));
  p->c = num(env, prop(env, o, "c"));
  p->creep = num(env, prop(env, o, "creep"));
  p->damp = num(env, prop(env, o, "damp"));
  p->tK = num(env, prop(env, o, "tK"));
  p->kappa = num(env, prop(env, o, "kappa"));
  p->cpIron = num(env, prop(env, o, "cpIron"));
  p->seatSlack = num(env, prop(env, o, "seatSlack"), 0.6);
  p->field = (int)num(env, prop(env, o, "field"));
  p->frozenOn = (int)num(env, prop(env, o, "frozenOn"));
  p->frozen = num(env, prop(env, o, "frozen"));
  p->meshEvery = (int)num(env, prop(env, o, "meshEvery"), 8);
  return nums(env, prop(env, o, "spin0"), p->spin0, 3) && nums(env, prop(env, o, "spin1"), p->spin1, 3) &&
         nums(env, prop(env, o, "gField"), p->gField, 3);
}

bool read_sim_args(napi_env env, napi_value o, cc_sim_args* a) {
  memset(a, 0, sizeof *a);
  bool b = false;
  napi_value v = prop(env, o, "settle");
  if (!is_nullish(env, v)) { napi_coerce_to_bool(env, v, &v); napi_get_value_bool(env, v, &b); a->settle = b; }
  v = prop(env, o, "kick");
  if (!is_nullish(env, v)) { napi_coerce_to_bool(env, v, &v); napi_get_value_bool(env, v, &b); a->kick = b; }
  a->gdt = num(env, prop(env, o, "gdt"));
  return nums(env, prop(env, o, "bulk0"), a->bulk0, 3) && nums(env, prop(env, o, "bulk1"), a->bulk1, 3) &&
         nums(env, prop(env, o, "cen0"), a->cen0, 3) && nums(env, prop(env, o, "cen1"), a->cen1, 3);
}

napi_value undefined(napi_env env) {
  napi_value u;
  napi_get_undefined(env, &u);
  return u;
}

void finalize_ctx(napi_env, void* data, void*) {
  Handle* h = static_cast<Handle*>(data);
  if (h->c) cc_destroy(h->c);
  delete h;
}

// ---------------------------------------------------------------- exports --
napi_value Create(napi_env env, napi_callback_info info) {
  Args a(env, info);
  cc_params p;
  if (!read_params(env, a.argv[0], &p)) { napi_throw_type_error(env, nullptr, "bad params"); return nullptr; }
  cc_ctx* c = nullptr;
  int r = cc_create(&p, &c);
  if (r) return throw_cc(env, nullptr, "cc_create", r);
  napi_value out;
  NAPI_OK(env, napi_create_external(env, new Handle{ c }, finalize_ctx, nullptr, &out));
  return out;
}

#define CTX0()                                  \
  Args a(env, info);                            \
  cc_ctx* c = ctx_of(env, a.argv[0]);           \
  if (!c) return nullptr;
#define RET(call, name)                         \
  do {                                          \
    int _r = (call);                            \
    if (_r) return throw_cc(env, c, name, _r);  \
    return undefined(env);                      \
  } while (0)
#define TA(var, T, idx, type, len)                                  \
  bool ok_##var;                                                    \
  T* var = (T*)typed(env, a.argv[idx], type, (len), &ok_##var);     \
  if (!ok_##var) return nullptr;

napi_value SetParams(napi_env env, napi_callback_info info) {
  CTX0();
  cc_params p;
  if (!read_params(env, a.argv[1], &p)) { napi_throw_type_error(env, nullptr, "bad params"); return nullptr; }
  RET(cc_set_params(c, &p), "cc_set_params");
}

napi_value Upload(napi_env env, napi_callback_info info) {
  CTX0();
  int N = 0;
  napi_get_value_int32(env, a.argv[5], &N);
  TA(pos, float, 1, napi_float32_array, (size_t)N * 4);
  TA(vel, float, 2, napi_float32_array, (size_t)N * 4);
  TA(mat, float, 3, napi_float32_array, (size_t)N);
  TA(q, float, 4, napi_float32_array, (size_t)N);
  RET(cc_upload_state(c, pos, vel, mat, q), "cc_upload_state");
}

napi_value Download(napi_env env, napi_callback_info info) {
  CTX0();
  int N = 0;
  napi_get_value_int32(env, a.argv[6], &N);
  TA(pos, float, 1, napi_float32_array, (size_t)N * 4);
  TA(vel, float, 2, napi_float32_array, (size_t)N * 4);
  TA(aux, float, 3, napi_float32_array, (size_t)N * 4);
  TA(q, float, 4, napi_float32_array, (size_t)N);
  TA(mat, float, 5, napi_float32_array, (size_t)N);
  RET(cc_download_state(c, pos, vel, aux, q, mat), "cc_download_state");
}

napi_value SetBoxes(napi_env env, napi_callback_info info) {
  CTX0();
  double LA[3], LB[3];
  if (!nums(env, a.argv[1], LA, 3) || !nums(env, a.argv[2], LB, 3)) { napi_throw_type_error(env, nullptr, "bad box"); return nullptr; }
  bool on = false;
  napi_value v;
  napi_coerce_to_bool(env, a.argv[3], &v);
  napi_get_value_bool(env, v, &on);
  RET(cc_set_boxes(c, LA, LB, on), "cc_set_boxes");
}

napi_value SetMono(napi_env env, napi_callback_info info) {
  CTX0();
  int b = 0;
  napi_get_value_int32(env, a.argv[1], &b);
  TA(m, float, 2, napi_float32_array, 4);
  RET(cc_set_mono(c, b, m), "cc_set_mono");
}

napi_value Slots(napi_env env, napi_callback_info info) { CTX0(); RET(cc_slots(c), "cc_slots"); }
napi_value Mesh(napi_env env, napi_callback_info info) { CTX0(); RET(cc_mesh(c), "cc_mesh"); }
napi_value PP(napi_env env, napi_callback_info info) { CTX0(); RET(cc_pp(c), "cc_pp"); }
napi_value Sync(napi_env env, napi_callback_info info) { CTX0(); RET(cc_sync(c), "cc_sync"); }

napi_value BoxPass(napi_env env, napi_callback_info info, int (*f)(cc_ctx*, int), const char* name) {
  CTX0();
  int b = 0;
  napi_get_value_int32(env, a.argv[1], &b);
  RET(f(c, b), name);
}
napi_value Deposit(napi_env env, napi_callback_info info) { return BoxPass(env, info, cc_deposit, "cc_deposit"); }
napi_value Coarse(napi_env env, napi_callback_info info) { return BoxPass(env, info, cc_coarse, "cc_coarse"); }
napi_value Cell(napi_env env, napi_callback_info info) { return BoxPass(env, info, cc_cell, "cc_cell"); }

napi_value ReadSeats(napi_env env, napi_callback_info info) {
  CTX0();
  TA(sa, uint32_t, 1, napi_uint32_array, 4);
  TA(sb, uint32_t, 2, napi_uint32_array, 4);
  RET(cc_read_seats(c, sa, sb), "cc_read_seats");
}

static size_t mesh_len(int which) { return which == 2 ? 4 : (which == 0 || which == 6) ? 513 * 512 * 4 : 64 * 64 * 4; }

napi_value ReadMesh(napi_env env, napi_callback_info info) {
  CTX0();
  int b = 0, w = 0;
  napi_get_value_int32(env, a.argv[1], &b);
  napi_get_value_int32(env, a.argv[2], &w);
  TA(out, float, 3, napi_float32_array, mesh_len(w));
  RET(cc_read_mesh(c, b, w, out), "cc_read_mesh");
}
napi_value WriteMesh(napi_env env, napi_callback_info info) {
  CTX0();
  int b = 0, w = 0;
  napi_get_value_int32(env, a.argv[1], &b);
  napi_get_value_int32(env, a.argv[2], &w);
  TA(in, float, 3, napi_float32_array, mesh_len(w));
  RET(cc_write_mesh(c, b, w, in), "cc_write_mesh");
}

napi_value SetPPTable(napi_env env, napi_callback_info info) {
  CTX0();
  TA(t, float, 1, napi_float32_array, 128);
  RET(cc_set_pp_table(c, t), "cc_set_pp_table");
}
napi_value SetPP(napi_env env, napi_callback_info info) {
  CTX0();
  TA(l, int32_t, 1, napi_int32_array, 0);
  int n = 0;
  napi_get_value_int32(env, a.argv[2], &n);
  RET(cc_set_pp(c, l, n), "cc_set_pp");
}
napi_value ReadPP(napi_env env, napi_callback_info info) {
  CTX0();
  TA(l, float, 1, napi_float32_array, 8192 * 4);
  TA(b, float, 2, napi_float32_array, 8192);
  TA(f, float, 3, napi_float32_array, 4 * 8192 * 4);
  RET(cc_read_pp(c, l, b, f), "cc_read_pp");
}

napi_value Sim(napi_env env, napi_callback_info info) {
  CTX0();
  cc_sim_args s;
  if (!read_sim_args(env, a.argv[1], &s)) { napi_throw_type_error(env, nullptr, "bad sim args"); return nullptr; }
  RET(cc_sim(c, &s), "cc_sim");
}
napi_value Step(napi_env env, napi_callback_info info) {
  CTX0();
  cc_sim_args s;
  if (!read_sim_args(env, a.argv[1], &s)) { napi_throw_type_error(env, nullptr, "bad sim args"); return nullptr; }
  RET(cc_step(c, s.settle, s.kick, &s), "cc_step");
}
napi_value StepN(napi_env env, napi_callback_info info) {
  CTX0();
  int n = 0;
  napi_get_value_int32(env, a.argv[1], &n);
  cc_sim_args s;
  if (!read_sim_args(env, a.argv[2], &s)) { napi_throw_type_error(env, nullptr, "bad sim args"); return nullptr; }
  bool okb;
  size_t len = 0;
  double* bl = (double*)typed(env, a.argv[3], napi_float64_array, 0, &okb);
  if (!okb) return nullptr;
  if (bl) { napi_typedarray_type t; void* d; napi_value ab; size_t off; napi_get_typedarray_info(env, a.argv[3], &t, &len, &d, &ab, &off); }
  int r = cc_step_n(c, n, &s, bl, (int)(len / 7));
  if (r) return throw_cc(env, c, "cc_step_n", r);
  double ms = 0;
  cc_last_step_ms(c, &ms);
  napi_value out;
  NAPI_OK(env, napi_create_double(env, ms, &out));
  return out;
}
napi_value GetStepNo(napi_env env, napi_callback_info info) {
  CTX0();
  int n = 0;
  int r = cc_get_step_no(c, &n);
  if (r) return throw_cc(env, c, "cc_get_step_no", r);
  napi_value out;
  NAPI_OK(env, napi_create_int32(env, n, &out));
  return out;
}
napi_value SetStepNo(napi_env env, napi_callback_info info) {
  CTX0();
  int n = 0;
  napi_get_value_int32(env, a.argv[1], &n);
  RET(cc_set_step_no(c, n), "cc_set_step_no");
}

napi_value Destroy(napi_env env, napi_callback_info info) {
  Args a(env, info);
  void* p = nullptr;
  if (napi_get_value_external(env, a.argv[0], &p) != napi_ok || !p) {
    napi_throw_type_error(env, nullptr, "expected a cc context handle");
    return nullptr;
  }
  Handle* h = static_cast<Handle*>(p);
  if (h->c) { cc_destroy(h->c); h->c = nullptr; }
  return undefined(env);
}

napi_value SetGraphs(napi_env env, napi_callback_info info) {
  CTX0();
  bool on = true;
  napi_value v = a.argv[1];
  napi_coerce_to_bool(env, v, &v);
  napi_get_value_bool(env, v, &on);
  RET(cc_set_graphs(c, on ? 1 : 0), "cc_set_graphs");
}

napi_value Rigid(napi_env env, napi_callback_info info) {
  CTX0();
  double b0[3], b1[3], c0[3], c1[3], r0[9], r1[9];
  if (!nums(env, a.argv[1], b0, 3) || !nums(env, a.argv[2], b1, 3) || !nums(env, a.argv[3], c0, 3) ||
      !nums(env, a.argv[4], c1, 3) || !nums(env, a.argv[5], r0, 9) || !nums(env, a.argv[6], r1, 9)) {
    napi_throw_type_error(env, nullptr, "bad rigid arguments");
    return nullptr;
  }
  RET(cc_rigid(c, b0, b1, c0, c1, r0, r1, num(env, a.argv[7])), "cc_rigid");
}

napi_value Couple(napi_env env, napi_callback_info info) {
  CTX0();
  cc_couple_args u;
  memset(&u, 0, sizeof u);
  napi_value o = a.argv[1];
  if (!nums(env, prop(env, o, "centre"), u.centre, 3)) { napi_throw_type_error(env, nullptr, "bad centre"); return nullptr; }
  u.vimp = num(env, prop(env, o, "vimp")); u.aimp = num(env, prop(env, o, "aimp")); u.c1 = num(env, prop(env, o, "c1"));
  u.invMu = num(env, prop(env, o, "invMu")); u.pexp = num(env, prop(env, o, "pexp"));
  u.stall = num(env, prop(env, o, "stall")); u.z = num(env, prop(env, o, "z")); u.heat = num(env, prop(env, o, "heat"));
  u.tK = num(env, prop(env, o, "tK")); u.frozen = num(env, prop(env, o, "frozen"));
  u.ambient = num(env, prop(env, o, "ambient")); u.melt = num(env, prop(env, o, "melt"));
  RET(cc_couple(c, &u), "cc_couple");
}

napi_value Diag(napi_env env, napi_callback_info info) {
  CTX0();
  TA(out, float, 1, napi_float32_array, 4);
  RET(cc_diag(c, out), "cc_diag");
}

napi_value Morton(napi_env env, napi_callback_info info) {
  CTX0();
  double core[3];
  if (!nums(env, a.argv[1], core, 3)) { napi_throw_type_error(env, nullptr, "bad core"); return nullptr; }
  TA(perm, int32_t, 2, napi_int32_array, 0);
  TA(into, int32_t, 3, napi_int32_array, 0);
  RET(cc_morton(c, core, perm, into), "cc_morton");
}

napi_value Books(napi_env env, napi_callback_info info) {
  CTX0();
  napi_value ab, out;
  void* data;
  NAPI_OK(env, napi_create_arraybuffer(env, 13 * sizeof(double), &data, &ab));
  int r = cc_books(c, static_cast<double*>(data));
  if (r) return throw_cc(env, c, "cc_books", r);
  NAPI_OK(env, napi_create_typedarray(env, napi_float64_array, 13, ab, 0, &out));
  return out;
}

napi_value set_num(napi_env env, napi_value o, const char* k, double v) {
  napi_value x;
  napi_create_double(env, v, &x);
  napi_set_named_property(env, o, k, x);
  return o;
}
napi_value set_vec(napi_env env, napi_value o, const char* k, const double* v, int n) {
  napi_value arr;
  napi_create_array_with_length(env, n, &arr);
  for (int i = 0; i < n; i++) {
    napi_value x;
    napi_create_double(env, v[i], &x);
    napi_set_element(env, arr, i, x);
  }
  napi_set_named_property(env, o, k, arr);
  return o;
}

// analyse(ctx, {m, a, spring, touch, cap, flags}) -> fof()'s report (src/nbody.js:888-892)
// fof()'s report (src/nbody.js:888-892) as the reference's object, pp.list over `ab`
static napi_value report_object(napi_env env, const cc_report& r, napi_value ab) {
  napi_value list;
  if (napi_create_typedarray(env, napi_int32_array, (size_t)r.ppCount, ab, 0, &list) != napi_ok) return nullptr;
  napi_value out, second, pp, nat;
  napi_create_object(env, &out);
  set_num(env, out, "largest", r.largest); set_num(env, out, "largestImp", r.largestImp);
  set_num(env, out, "largestIron", r.largestIron); set_num(env, out, "day", r.day); set_num(env, out, "tilt", r.tilt);
  set_vec(env, out, "axis", r.axis, 3); set_vec(env, out, "bodyVel", r.bodyVel, 3);
  set_num(env, out, "orbit", r.orbit); set_num(env, out, "escape", r.escape);
  if (r.hasSecond) {
    const cc_second& q = r.second;
    napi_create_object(env, &second);
    set_num(env, second, "mass", q.mass); set_num(env, second, "dist", q.dist);
    napi_value b;
    napi_get_boolean(env, q.bound != 0, &b);
    napi_set_named_property(env, second, "bound", b);
    set_num(env, second, "imp", q.imp); set_num(env, second, "iron", q.iron);
    set_vec(env, second, "com", q.com, 3); set_vec(env, second, "core", q.core, 3); set_vec(env, second, "vel", q.vel, 3);
    set_num(env, second, "R", q.R); set_num(env, second, "Redge", q.Redge); set_num(env, second, "Tsurf", q.Tsurf);
    set_vec(env, second, "hotCom", q.hotCom, 3); set_num(env, second, "hotR", q.hotR);
  } else {
    napi_get_null(env, &second);
  }
  napi_set_named_property(env, out, "second", second);
  set_vec(env, out, "com", r.com, 3); set_vec(env, out, "core", r.core, 3); set_vec(env, out, "vel", r.vel, 3);
  set_num(env, out, "R", r.R); set_num(env, out, "Redge", r.Redge); set_num(env, out, "Tsurf", r.Tsurf);
  set_vec(env, out, "hotCom", r.hotCom, 3); set_num(env, out, "hotR", r.hotR);
  set_num(env, out, "impMass", r.impMass); set_vec(env, out, "impCom", r.impCom, 3); set_num(env, out, "impRms", r.impRms);
  set_num(env, out, "M", r.M); set_num(env, out, "iron", r.iron); set_num(env, out, "drift", r.drift);
  set_num(env, out, "L", r.L); set_num(env, out, "KE", r.KE); set_num(env, out, "PE", r.PE); set_num(env, out, "EL", r.EL);
  set_num(env, out, "Q", r.Q); set_num(env, out, "unseated", r.unseated); set_num(env, out, "full", r.full);
  napi_create_object(env, &pp);
  napi_set_named_property(env, pp, "list", list);
  set_num(env, pp, "loose", r.ppLoose);
  napi_set_named_property(env, out, "pp", pp);
  set_num(env, out, "jD", r.jD); set_num(env, out, "moon", r.moon);
  // what the reference's report does not carry: group count and sizes, whether the unions were replayed, timing
  napi_create_object(env, &nat);
  set_num(env, nat, "groups", r.groups); set_num(env, nat, "largestN", r.largestN);
  set_num(env, nat, "secondN", r.secondN); set_num(env, nat, "replayed", r.replayed);
  set_vec(env, nat, "ms", r.ms, 3);
  napi_set_named_property(env, out, "analysis", nat);
  return out;
}

static bool analyse_args(napi_env env, napi_value o, cc_analyse_args* in) {
  memset(in, 0, sizeof *in);
  in->m = num(env, prop(env, o, "m"));
  in->a = num(env, prop(env, o, "a"));
  in->spring = num(env, prop(env, o, "spring"));
  in->touch = num(env, prop(env, o, "touch"));
  in->cap = (int32_t)num(env, prop(env, o, "cap"), 8192);
  in->flags = (int32_t)num(env, prop(env, o, "flags"));
  in->diag = nullptr;   // the DIAG read-out cc_diag left on the device
  if (in->cap < 0 || in->cap > 8192) { napi_throw_range_error(env, nullptr, "cap outside 0..8192"); return false; }
  return true;
}

napi_value Analyse(napi_env env, napi_callback_info info) {
  CTX0();
  cc_analyse_args in;
  if (!analyse_args(env, a.argv[1], &in)) return nullptr;
  cc_report r;
  napi_value ab;
  void* data;
  NAPI_OK(env, napi_create_arraybuffer(env, (size_t)(in.cap > 0 ? in.cap : 1) * 4, &data, &ab));
  int code = cc_analyse(c, &in, &r, static_cast<int32_t*>(data));
  if (code) return throw_cc(env, c, "cc_analyse", code);
  return report_object(env, r, ab);
}

// analyseBegin(ctx, args): the device's half now, the host's sums on a thread (src/nbody.js:967-1008)
napi_value AnalyseBegin(napi_env env, napi_callback_info info) {
  CTX0();
  cc_analyse_args in;
  if (!analyse_args(env, a.argv[1], &in)) return nullptr;
  RET(cc_analyse_begin(c, &in), "cc_analyse_begin");
}

// analyseEnd(ctx, wait): the begun analysis' report, or null while its sums still run (collect, src/nbody.js:1011-1025)
napi_value AnalyseEnd(napi_env env, napi_callback_info info) {
  CTX0();
  bool wait = false;
  napi_get_value_bool(env, a.argv[1], &wait);
  cc_report r;
  napi_value ab;
  void* data;
  NAPI_OK(env, napi_create_arraybuffer(env, (size_t)8192 * 4, &data, &ab));
  int code = cc_analyse_end(c, wait ? 1 : 0, &r, static_cast<int32_t*>(data));
  if (code < 0) return throw_cc(env, c, "cc_analyse_end", code);
  if (code == 1) { napi_value nul; napi_get_null(env, &nul); return nul; }
  return report_object(env, r, ab);
}

// relaxPiles([{p: Float32Array (n x 3), n, a, A, C, confine, sweeps}, ...]) -> [{sweeps, worst}, ...]:
// relax() of src/bodies.js:97-146 on each pile, the piles side by side (cc_relax_piles)
napi_value RelaxPiles(napi_env env, napi_callback_info info) {
  Args a(env, info);
  bool isArr = false;
  if (a.argc < 1 || napi_is_array(env, a.argv[0], &isArr) != napi_ok || !isArr) {
    napi_throw_type_error(env, nullptr, "relaxPiles takes an array of piles");
    return nullptr;
  }
  uint32_t count = 0;
  napi_get_array_length(env, a.argv[0], &count);
  std::vector<cc_pile> piles(count);
  std::vector<cc_pile_report> reps(count);
  for (uint32_t k = 0; k < count; k++) {
    napi_value o;
    napi_get_element(env, a.argv[0], k, &o);
    cc_pile& pl = piles[k];
    memset(&pl, 0, sizeof pl);
    pl.n = (int32_t)num(env, prop(env, o, "n"));
    bool ok = true;
    pl.p = static_cast<float*>(typed(env, prop(env, o, "p"), napi_float32_array, (size_t)(pl.n > 0 ? pl.n : 0) * 3, &ok));
    if (!ok) return nullptr;
    pl.a = num(env, prop(env, o, "a"));
    pl.A = num(env, prop(env, o, "A"));
    pl.C = num(env, prop(env, o, "C"));
    pl.confine = (int32_t)num(env, prop(env, o, "confine"));
    pl.sweeps = (int32_t)num(env, prop(env, o, "sweeps"));
  }
  int code = cc_relax_piles(piles.data(), (int)count, reps.data());
  if (code) {
    std::string msg = "cc_relax_piles failed (" + std::to_string(code) + "): " + cc_bodies_last_error();
    napi_throw_error(env, nullptr, msg.c_str());
    return nullptr;
  }
  napi_value out;
  napi_create_array_with_length(env, count, &out);
  for (uint32_t k = 0; k < count; k++) {
    napi_value r;
    napi_create_object(env, &r);
    set_num(env, r, "sweeps", reps[k].sweeps);
    set_num(env, r, "worst", reps[k].worst);
    napi_set_element(env, out, k, r);
  }
  return out;
}

// craterStats(pos, vel, R, N, mref, {Rbox, lSI, vSI, gSI, rhoSI, vOrb, vEsc}) -> craterStats()'s report
// (chicxulub.html:848-919), the same names (cc_crater_stats)
napi_value CraterStats(napi_env env, napi_callback_info info) {
  Args a(env, info);
  if (a.argc < 6) { napi_throw_type_error(env, nullptr, "craterStats(pos, vel, R, N, mref, args)"); return nullptr; }
  int N = (int)num(env, a.argv[3]);
  if (N < 0) { napi_throw_range_error(env, nullptr, "N must be >= 0"); return nullptr; }
  bool ok = true;
  const float* P = static_cast<const float*>(typed(env, a.argv[0], napi_float32_array, (size_t)N * 4, &ok));
  if (!ok) return nullptr;
  const float* V = static_cast<const float*>(typed(env, a.argv[1], napi_float32_array, (size_t)N * 4, &ok));
  if (!ok) return nullptr;
  const float* R = static_cast<const float*>(typed(env, a.argv[2], napi_float32_array, (size_t)N, &ok));
  if (!ok) return nullptr;
  napi_value o = a.argv[5];
  cc_crater_args ca;
  ca.Rbox = num(env, prop(env, o, "Rbox")); ca.lSI = num(env, prop(env, o, "lSI")); ca.vSI = num(env, prop(env, o, "vSI"));
  ca.gSI = num(env, prop(env, o, "gSI")); ca.rhoSI = num(env, prop(env, o, "rhoSI"));
  ca.vOrb = num(env, prop(env, o, "vOrb")); ca.vEsc = num(env, prop(env, o, "vEsc"));
  cc_crater_report r;
  int code = cc_crater_stats(N, num(env, a.argv[4]), P, V, R, &ca, &r);
  if (code) {
    napi_throw_error(env, nullptr, ("cc_crater_stats failed (" + std::to_string(code) + "): bad arguments").c_str());
    return nullptr;
  }
  napi_value out;
  napi_create_object(env, &out);
  set_num(env, out, "craterD", r.craterD); set_num(env, out, "rimR", r.rimR); set_num(env, out, "rimH", r.rimH);
  set_num(env, out, "depth", r.depth); set_num(env, out, "excavated", r.excavated); set_num(env, out, "thrown", r.thrown);
  set_num(env, out, "platform", r.platform); set_num(env, out, "airborne", r.airborne);
  set_num(env, out, "fastOrb", r.fastOrb); set_num(env, out, "fastEsc", r.fastEsc); set_num(env, out, "melt", r.melt);
  set_num(env, out, "vapour", r.vapour); set_num(env, out, "Tmax", r.Tmax); set_num(env, out, "mass", r.mass);
  set_num(env, out, "KE", r.KE); set_num(env, out, "PE", r.PE); set_num(env, out, "drift", r.drift);
  set_num(env, out, "Tsurf", r.Tsurf); set_num(env, out, "R", r.R);
  const double zero[3] = { 0, 0, 0 };
  set_vec(env, out, "vel", zero, 3);
  set_vec(env, out, "com", r.com, 3);
  return out;
}

napi_value Init(napi_env env, napi_value exports) {
  struct { const char* name; napi_callback cb; } fns[] = {
    { "create", Create }, { "destroy", Destroy }, { "setParams", SetParams }, { "uploadState", Upload }, { "downloadState", Download },
    { "setBoxes", SetBoxes }, { "setMono", SetMono }, { "slots", Slots }, { "readSeats", ReadSeats },
    { "deposit", Deposit }, { "coarse", Coarse }, { "cell", Cell }, { "mesh", Mesh }, { "readMesh", ReadMesh },
    { "writeMesh", WriteMesh }, { "setPPTable", SetPPTable }, { "setPP", SetPP }, { "pp", PP }, { "readPP", ReadPP },
    { "sim", Sim }, { "step", Step }, { "stepN", StepN }, { "getStepNo", GetStepNo }, { "setStepNo", SetStepNo },
    { "setGraphs", SetGraphs },
    { "rigid", Rigid }, { "couple", Couple }, { "diag", Diag }, { "morton", Morton }, { "analyse", Analyse }, { "analyseBegin", AnalyseBegin }, { "analyseEnd", AnalyseEnd }, { "books", Books }, { "sync", Sync },
    { "relaxPiles", RelaxPiles }, { "craterStats", CraterStats },
  };
  for (auto& f : fns) {
    napi_value fn;
    NAPI_OK(env, napi_create_function(env, f.name, NAPI_AUTO_LENGTH, f.cb, nullptr, &fn));
    NAPI_OK(env, napi_set_named_property(env, exports, f.name, fn));
  }
  return exports;
}

}  // namespace

NAPI_MODULE(cc_addon, Init)
