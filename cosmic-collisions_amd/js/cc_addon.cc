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

napi_value Init(napi_env env, napi_value exports) {
  struct { const char* name; napi_callback cb; } fns[] = {
    { "create", Create }, { "destroy", Destroy }, { "setParams", SetParams }, { "uploadState", Upload }, { "downloadState", Download },
    { "setBoxes", SetBoxes }, { "setMono", SetMono }, { "slots", Slots }, { "readSeats", ReadSeats },
    { "deposit", Deposit }, { "coarse", Coarse }, { "cell", Cell }, { "mesh", Mesh }, { "readMesh", ReadMesh },
    { "writeMesh", WriteMesh }, { "setPPTable", SetPPTable }, { "setPP", SetPP }, { "pp", PP }, { "readPP", ReadPP },
    { "sim", Sim }, { "step", Step }, { "stepN", StepN }, { "getStepNo", GetStepNo }, { "setStepNo", SetStepNo },
    { "setGraphs", SetGraphs },
    { "rigid", Rigid }, { "couple", Couple }, { "diag", Diag }, { "morton", Morton }, { "books", Books }, { "sync", Sync },
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
