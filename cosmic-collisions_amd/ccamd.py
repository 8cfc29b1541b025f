"""Python binding of the C ABI in include/cc_hip.h (libcc_hip.so), over ctypes.

A thin layer: numpy arrays in, numpy arrays out, a Python exception for every
negative return code with the library's own message.  It is what the tests
and the benchmark drive; the reference-shaped engine object is the
JavaScript one in js/nbody.js.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcc_hip.so")
_LIB = None

FL = 6.0
FH = 2 * FL / 64
PP_CAP = 8192


class Params(C.Structure):
    _fields_ = [("N", C.c_int), ("W", C.c_int), ("H", C.c_int), ("G", C.c_int),
                ("m", C.c_double), ("a", C.c_double), ("cell", C.c_double), ("dt", C.c_double), ("k", C.c_double),
                ("c", C.c_double), ("creep", C.c_double), ("damp", C.c_double), ("tK", C.c_double),
                ("kappa", C.c_double), ("cpIron", C.c_double), ("seatSlack", C.c_double),
                ("spin0", C.c_double * 3), ("spin1", C.c_double * 3), ("field", C.c_int), ("gField", C.c_double * 3),
                ("frozenOn", C.c_int), ("frozen", C.c_double), ("meshEvery", C.c_int)]


class SimArgs(C.Structure):
    _fields_ = [("settle", C.c_int), ("kick", C.c_int), ("gdt", C.c_double), ("bulk0", C.c_double * 3),
                ("bulk1", C.c_double * 3), ("cen0", C.c_double * 3), ("cen1", C.c_double * 3)]


EXPORTS = ["cc_create", "cc_destroy", "cc_last_error", "cc_set_params", "cc_upload_state", "cc_download_state",
           "cc_set_boxes", "cc_set_mono", "cc_slots", "cc_read_seats", "cc_deposit", "cc_coarse", "cc_cell", "cc_mesh",
           "cc_read_mesh", "cc_write_mesh", "cc_set_pp_table", "cc_set_pp", "cc_pp", "cc_read_pp", "cc_sim", "cc_step",
           "cc_step_n", "cc_set_graphs", "cc_time_pass", "cc_get_step_no", "cc_set_step_no", "cc_rigid", "cc_couple",
           "cc_diag", "cc_morton", "cc_analyse", "cc_books", "cc_sync", "cc_last_step_ms", "cc_batch_create", "cc_batch_destroy", "cc_batch_last_error",
           "cc_batch_step_n", "cc_batch_last_ms", "cc_ff_create", "cc_ff_destroy", "cc_ff_last_error", "cc_ff_set_slope",
           "cc_ff_weight_count", "cc_ff_get_weights", "cc_ff_set_weights", "cc_ff_record", "cc_ff_push_sample",
           "cc_ff_read_sample", "cc_ff_samples", "cc_ff_clear", "cc_ff_train", "cc_ff_eval", "cc_ff_gradient",
           "cc_ff_passes", "cc_ff_read_pass", "cc_ff_clear_passes", "cc_ff_pair",
           "cc_set_far_field", "cc_ff_stats", "cc_relax_piles", "cc_bodies_last_error", "cc_autotune", "cc_set_tuning", "cc_set_visit", "cc_crater_stats", "cc_pp_noalloc", "cc_analyse_begin", "cc_analyse_end"]


class CraterArgs(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("Rbox", "lSI", "vSI", "gSI", "rhoSI", "vOrb", "vEsc")]


class CraterReport(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("craterD", "rimR", "rimH", "depth", "excavated", "thrown", "platform",
                                          "airborne", "fastOrb", "fastEsc", "melt", "vapour", "Tmax", "mass", "KE",
                                          "PE", "drift", "Tsurf", "R")] + [("com", C.c_double * 3)]


class TuneReport(C.Structure):
    _fields_ = [("cellBlocks", C.c_int32), ("simOrder", C.c_int32), ("cellMs", C.c_double * 4),
                ("simMs", C.c_double * 4)]


class Pile(C.Structure):
    """cc_pile: n x 3 FP32 positions relaxed in place (relax(), src/bodies.js:97-146)."""
    _fields_ = [("p", C.POINTER(C.c_float)), ("n", C.c_int32), ("sweeps", C.c_int32), ("a", C.c_double),
                ("A", C.c_double), ("C", C.c_double), ("confine", C.c_int32), ("pad_", C.c_int32)]


class PileReport(C.Structure):
    _fields_ = [("sweeps", C.c_int32), ("pad_", C.c_int32), ("worst", C.c_double)]


PILE_SPHEROID, PILE_GROUND = 0, 1


class AnalyseArgs(C.Structure):
    _fields_ = [("m", C.c_double), ("a", C.c_double), ("spring", C.c_double), ("touch", C.c_double),
                ("cap", C.c_int32), ("flags", C.c_int32), ("diag", C.POINTER(C.c_float))]


REPLAY_UNIONS = 1   # CC_ANALYSE_REPLAY_UNIONS

_V3 = C.c_double * 3


class Second(C.Structure):
    _fields_ = [("mass", C.c_double), ("dist", C.c_double), ("bound", C.c_int32), ("pad_", C.c_int32),
                ("imp", C.c_double), ("iron", C.c_double), ("com", _V3), ("core", _V3), ("vel", _V3),
                ("R", C.c_double), ("Redge", C.c_double), ("Tsurf", C.c_double), ("hotCom", _V3), ("hotR", C.c_double)]


class Report(C.Structure):
    _fields_ = ([(n, C.c_double) for n in ("largest", "largestImp", "largestIron", "day", "tilt")] +
                [("axis", _V3), ("bodyVel", _V3), ("orbit", C.c_double), ("escape", C.c_double),
                 ("com", _V3), ("core", _V3), ("vel", _V3), ("R", C.c_double), ("Redge", C.c_double),
                 ("Tsurf", C.c_double), ("hotCom", _V3), ("hotR", C.c_double), ("impMass", C.c_double),
                 ("impCom", _V3), ("impRms", C.c_double)] +
                [(n, C.c_double) for n in ("M", "iron", "drift", "L", "KE", "PE", "EL", "Q", "unseated", "full",
                                           "jD", "moon")] +
                [("hasSecond", C.c_int32), ("pad_", C.c_int32), ("second", Second), ("ppCount", C.c_int32),
                 ("ppLoose", C.c_int32), ("groups", C.c_int32), ("largestN", C.c_int32), ("secondN", C.c_int32),
                 ("replayed", C.c_int32), ("ms", C.c_double * 3)])


def _plain(st):
    """A ctypes structure as a dict of Python numbers and lists."""
    out = {}
    for name, _ in st._fields_:
        if name.startswith("pad_"):
            continue
        v = getattr(st, name)
        if isinstance(v, C.Structure):
            v = _plain(v)
        elif isinstance(v, C.Array):
            v = list(v)
        out[name] = v
    return out


def report_dict(rep, lst):
    """fof()'s report (src/nbody.js:888-892) from a Report and the loose list."""
    d = _plain(rep)
    second = d.pop("second")
    has = d.pop("hasSecond")
    if has:
        second["bound"] = bool(second["bound"])
    d["second"] = second if has else None
    d["pp"] = {"list": np.array(lst[:rep.ppCount], np.int32), "loose": rep.ppLoose}
    return d


class CoupleArgs(C.Structure):
    _fields_ = [("centre", C.c_double * 3)] + [(n, C.c_double) for n in (
        "vimp", "aimp", "c1", "invMu", "pexp", "stall", "z", "heat", "tK", "frozen", "ambient", "melt")]


class CCError(RuntimeError):
    pass


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise CCError("libcc_hip.so is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        dp, fp, ip, up = (C.POINTER(C.c_double), C.POINTER(C.c_float), C.POINTER(C.c_int32),
                          C.POINTER(C.c_uint32))
        sig = {
            "cc_create": ([C.POINTER(Params), C.POINTER(P)], C.c_int), "cc_destroy": ([P], C.c_int),
            "cc_last_error": ([P], C.c_char_p), "cc_set_params": ([P, C.POINTER(Params)], C.c_int),
            "cc_upload_state": ([P, fp, fp, fp, fp], C.c_int), "cc_download_state": ([P, fp, fp, fp, fp, fp], C.c_int),
            "cc_set_boxes": ([P, dp, dp, C.c_int], C.c_int), "cc_set_mono": ([P, C.c_int, fp], C.c_int),
            "cc_slots": ([P], C.c_int), "cc_read_seats": ([P, up, up], C.c_int),
            "cc_deposit": ([P, C.c_int], C.c_int), "cc_coarse": ([P, C.c_int], C.c_int),
            "cc_cell": ([P, C.c_int], C.c_int), "cc_mesh": ([P], C.c_int),
            "cc_read_mesh": ([P, C.c_int, C.c_int, fp], C.c_int), "cc_write_mesh": ([P, C.c_int, C.c_int, fp], C.c_int),
            "cc_set_pp_table": ([P, fp], C.c_int), "cc_set_pp": ([P, ip, C.c_int], C.c_int), "cc_pp": ([P], C.c_int),
            "cc_read_pp": ([P, fp, fp, fp], C.c_int), "cc_sim": ([P, C.POINTER(SimArgs)], C.c_int),
            "cc_step": ([P, C.c_int, C.c_int, C.POINTER(SimArgs)], C.c_int),
            "cc_step_n": ([P, C.c_int, C.POINTER(SimArgs), dp, C.c_int], C.c_int),
            "cc_set_graphs": ([P, C.c_int], C.c_int), "cc_time_pass": ([P, C.c_int, C.c_int, dp], C.c_int),
            "cc_get_step_no": ([P, ip], C.c_int), "cc_set_step_no": ([P, C.c_int], C.c_int),
            "cc_rigid": ([P, dp, dp, dp, dp, dp, dp, C.c_double], C.c_int), "cc_diag": ([P, fp], C.c_int),
            "cc_couple": ([P, C.POINTER(CoupleArgs)], C.c_int),
            "cc_morton": ([P, dp, ip, ip], C.c_int), "cc_books": ([P, dp], C.c_int), "cc_sync": ([P], C.c_int),
            "cc_analyse": ([P, C.POINTER(AnalyseArgs), C.POINTER(Report), ip], C.c_int),
            "cc_last_step_ms": ([P, dp], C.c_int),
            "cc_analyse_begin": ([P, C.POINTER(AnalyseArgs)], C.c_int),
            "cc_analyse_end": ([P, C.c_int, C.POINTER(Report), ip], C.c_int),
            "cc_batch_create": ([C.POINTER(P), C.c_int, C.POINTER(P)], C.c_int), "cc_batch_destroy": ([P], C.c_int),
            "cc_batch_last_error": ([P], C.c_char_p),
            "cc_batch_step_n": ([P, C.c_int, C.POINTER(SimArgs), dp, ip], C.c_int),
            "cc_batch_last_ms": ([P, dp], C.c_int),
            "cc_ff_create": ([C.c_int, C.POINTER(P)], C.c_int), "cc_ff_destroy": ([P], C.c_int),
            "cc_ff_last_error": ([P], C.c_char_p), "cc_ff_weight_count": ([C.POINTER(C.c_int64)], C.c_int),
            "cc_ff_get_weights": ([P, fp, C.c_int64], C.c_int), "cc_ff_set_weights": ([P, fp, C.c_int64], C.c_int),
            "cc_ff_record": ([P, P, C.c_int], C.c_int), "cc_ff_push_sample": ([P, fp, fp, fp], C.c_int),
            "cc_ff_read_sample": ([P, C.c_int, fp, fp, fp], C.c_int), "cc_ff_samples": ([P, ip], C.c_int),
            "cc_ff_clear": ([P], C.c_int), "cc_ff_set_slope": ([P, C.c_double], C.c_int),
            "cc_ff_passes": ([P, ip], C.c_int), "cc_ff_read_pass": ([P, C.c_int, fp, fp, fp], C.c_int),
            "cc_ff_clear_passes": ([P], C.c_int), "cc_ff_pair": ([P, C.c_int, C.c_int, ip], C.c_int),
            "cc_ff_train": ([P, C.c_int, C.c_int, C.c_double, C.c_uint32, dp], C.c_int),
            "cc_ff_eval": ([P, C.c_int, C.c_int, fp, dp, dp], C.c_int),
            "cc_ff_gradient": ([P, C.c_int, C.c_int, fp, dp], C.c_int),
            "cc_set_far_field": ([P, P, C.c_int, C.c_double], C.c_int), "cc_ff_stats": ([P, dp], C.c_int),
            "cc_relax_piles": ([C.POINTER(Pile), C.c_int, C.POINTER(PileReport)], C.c_int),
            "cc_autotune": ([P, C.c_int, C.POINTER(TuneReport)], C.c_int),
            "cc_set_tuning": ([P, C.c_int, C.c_int, C.c_int], C.c_int),
            "cc_set_visit": ([P, C.c_int], C.c_int),
            "cc_crater_stats": ([C.c_int, C.c_double, fp, fp, fp, C.POINTER(CraterArgs), C.POINTER(CraterReport)],
                                C.c_int),
            "cc_bodies_last_error": ([], C.c_char_p),
            "cc_pp_noalloc": ([P, P, C.c_int, dp, dp, C.c_int, P, C.c_int, C.c_double, C.c_double, P, P, P], C.c_int),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _LIB = L
    return _LIB


def _f(a):
    return np.ascontiguousarray(a, np.float32)


def _ptr(a, t=C.c_float):
    return a.ctypes.data_as(C.POINTER(t)) if a is not None else None


def _d3(v):
    return (C.c_double * len(v))(*[float(x) for x in v])


def params(N, W, G, m, a, cell, dt, k, c, creep=0.12, damp=None, tK=None, kappa=0.3, spin0=(0, 0, 0),
           spin1=(0, 0, 0), field=False, gField=(0, 0, 0), frozen=None, meshEvery=8, cpIron=450.0 / 1200.0,
           seatSlack=0.6):
    p = Params()
    p.N, p.W, p.H, p.G = N, W, N // W, G
    p.m, p.a, p.cell, p.dt, p.k, p.c, p.creep = m, a, cell, dt, k, c, creep
    p.damp = float(np.exp(-dt * 12)) if damp is None else damp
    p.tK = 0.5 * 6.256e7 / 1200.0 / m if tK is None else tK
    p.kappa, p.cpIron, p.seatSlack = kappa, cpIron, seatSlack
    p.spin0[:] = list(spin0)
    p.spin1[:] = list(spin1)
    p.field = 1 if field else 0
    p.gField[:] = list(gField)
    p.frozenOn = 0 if frozen is None else 1
    p.frozen = 0.0 if frozen is None else float(frozen)
    p.meshEvery = meshEvery
    return p


def sim_args(settle=False, kick=False, gdt=0.0, bulk0=(0, 0, 0), bulk1=(0, 0, 0), cen0=(0, 0, 0), cen1=(0, 0, 0)):
    s = SimArgs()
    s.settle, s.kick, s.gdt = int(bool(settle)), int(bool(kick)), float(gdt)
    s.bulk0[:], s.bulk1[:], s.cen0[:], s.cen1[:] = list(bulk0), list(bulk1), list(cen0), list(cen1)
    return s


class Context:
    """One simulation instance on the GPU (cc_ctx)."""

    def __init__(self, p):
        self.p = p
        self.N = p.N
        self.G = p.G
        sx = 16 if p.G == 128 else 8          # src/gl.js:73-83: SX slices across, G/SX down
        self.AW = sx * p.G
        self.AH = (p.G // sx) * p.G
        h = C.c_void_p()
        r = lib().cc_create(C.byref(p), C.byref(h))
        if r != 0:
            raise CCError("cc_create: %s" % lib().cc_last_error(None).decode())
        self.h = h

    def _chk(self, r, what):
        if r != 0:
            raise CCError("%s: %s" % (what, lib().cc_last_error(self.h).decode()))

    def close(self):
        if self.h:
            lib().cc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, p):
        self._chk(lib().cc_set_params(self.h, C.byref(p)), "cc_set_params")
        self.p = p

    def upload(self, pos=None, vel=None, mat=None, q=None):
        arrs = [None if x is None else _f(x) for x in (pos, vel, mat, q)]
        self._chk(lib().cc_upload_state(self.h, *[_ptr(a) for a in arrs]), "cc_upload_state")

    def download(self):
        N = self.N
        pos, vel, aux = (np.empty((N, 4), np.float32) for _ in range(3))
        q, mat = np.empty(N, np.float32), np.empty(N, np.float32)
        self._chk(lib().cc_download_state(self.h, _ptr(pos), _ptr(vel), _ptr(aux), _ptr(q), _ptr(mat)),
                  "cc_download_state")
        return pos, vel, aux, q, mat

    def set_boxes(self, LA, LB, boxB):
        self._chk(lib().cc_set_boxes(self.h, _d3(LA), _d3(LB), int(bool(boxB))), "cc_set_boxes")

    def set_mono(self, box, mono):
        self._chk(lib().cc_set_mono(self.h, box, _ptr(_f(mono))), "cc_set_mono")

    def slots(self):
        self._chk(lib().cc_slots(self.h), "cc_slots")

    def read_seats(self):
        a = np.empty((self.AH, self.AW, 4), np.uint32)
        b = np.empty((self.AH, self.AW, 4), np.uint32)
        self._chk(lib().cc_read_seats(self.h, _ptr(a, C.c_uint32), _ptr(b, C.c_uint32)), "cc_read_seats")
        return a, b

    def deposit(self, box=0):
        self._chk(lib().cc_deposit(self.h, box), "cc_deposit")

    def coarse(self, box=0):
        self._chk(lib().cc_coarse(self.h, box), "cc_coarse")

    def cell(self, box=0):
        self._chk(lib().cc_cell(self.h, box), "cc_cell")

    def mesh(self):
        self._chk(lib().cc_mesh(self.h), "cc_mesh")

    SHAPES = {0: (513, 512, 4), 1: (64, 64, 4), 2: (4,), 3: (64, 64, 4), 4: (64, 64, 4), 5: (64, 64, 4),
              6: (513, 512, 4)}
    NAMES = {"fine": 0, "coarse": 1, "mono": 2, "farA": 3, "farT0": 4, "farT1": 5, "force": 6}

    def read_mesh(self, box, which):
        w = self.NAMES.get(which, which)
        out = np.empty(self.SHAPES[w], np.float32)
        self._chk(lib().cc_read_mesh(self.h, box, w, _ptr(out)), "cc_read_mesh")
        return out

    def write_mesh(self, box, which, arr):
        w = self.NAMES.get(which, which)
        a = _f(arr).reshape(self.SHAPES[w])
        self._chk(lib().cc_write_mesh(self.h, box, w, _ptr(a)), "cc_write_mesh")

    def set_pp_table(self, tab):
        self._chk(lib().cc_set_pp_table(self.h, _ptr(_f(tab).reshape(64, 2))), "cc_set_pp_table")

    def set_pp(self, lst):
        a = np.ascontiguousarray(lst, np.int32)
        self._chk(lib().cc_set_pp(self.h, _ptr(a, C.c_int32), len(a)), "cc_set_pp")

    def pp(self):
        self._chk(lib().cc_pp(self.h), "cc_pp")

    def read_pp(self):
        lst = np.empty((PP_CAP, 4), np.float32)
        box = np.empty(PP_CAP, np.float32)
        force = np.empty((4, PP_CAP, 4), np.float32)
        self._chk(lib().cc_read_pp(self.h, _ptr(lst), _ptr(box), _ptr(force)), "cc_read_pp")
        return lst, box, force

    def sim(self, args):
        self._chk(lib().cc_sim(self.h, C.byref(args)), "cc_sim")

    def step(self, settle, kick, args):
        self._chk(lib().cc_step(self.h, int(settle), int(kick), C.byref(args)), "cc_step")

    def step_n(self, n, args, boxL=None):
        if boxL is None:
            self._chk(lib().cc_step_n(self.h, n, C.byref(args), None, 0), "cc_step_n")
        else:
            b = np.ascontiguousarray(boxL, np.float64).reshape(-1, 7)
            self._chk(lib().cc_step_n(self.h, n, C.byref(args), _ptr(b, C.c_double), len(b)), "cc_step_n")
        ms = C.c_double()
        lib().cc_last_step_ms(self.h, C.byref(ms))
        return ms.value

    def time_pass(self, which, reps):
        """Device ms per pass (cc_time_pass): which = 'sim' | 'slots' | 'mesh' | 'pp' | 'mesh_between' |
        'mesh_judge' | 'copy'."""
        ms = C.c_double()
        k = {"sim": 0, "slots": 1, "mesh": 2, "pp": 3, "mesh_between": 4, "mesh_judge": 5, "copy": 6}[which]
        self._chk(lib().cc_time_pass(self.h, k, int(reps), C.byref(ms)), "cc_time_pass")
        return ms.value

    def set_far_field(self, ff, judge_every=16, tol=1e-3):
        """Attach a FarField network (None: detach) to this instance's step."""
        self._chk(lib().cc_set_far_field(self.h, ff.h if ff is not None else None, int(judge_every), float(tol)),
                  "cc_set_far_field")

    FF_STATS = ("last_err", "judged", "fallbacks", "learned", "max_err", "sum_err", "usable", "spare")

    def ff_stats(self):
        """Per box (2 x 8, FF_STATS): last relative force error, judge passes, fallbacks, mesh steps on the
        network, the largest and the summed error and the count of judges measured from a usable base."""
        out = np.zeros(16, np.float64)
        self._chk(lib().cc_ff_stats(self.h, _ptr(out, C.c_double)), "cc_ff_stats")
        return out.reshape(2, 8)

    def set_graphs(self, on):
        self._chk(lib().cc_set_graphs(self.h, int(bool(on))), "cc_set_graphs")

    @property
    def step_no(self):
        v = C.c_int32()
        self._chk(lib().cc_get_step_no(self.h, C.byref(v)), "cc_get_step_no")
        return v.value

    @step_no.setter
    def step_no(self, n):
        self._chk(lib().cc_set_step_no(self.h, int(n)), "cc_set_step_no")

    def rigid(self, bulk0, bulk1, cen0, cen1, rot0, rot1, dtF):
        self._chk(lib().cc_rigid(self.h, _d3(bulk0), _d3(bulk1), _d3(cen0), _d3(cen1), _d3(rot0), _d3(rot1),
                                 float(dtF)), "cc_rigid")

    def couple(self, centre, **kw):
        a = CoupleArgs()
        a.centre[:] = [float(x) for x in centre]
        for k, v in kw.items():
            setattr(a, k, float(v))
        self._chk(lib().cc_couple(self.h, C.byref(a)), "cc_couple")

    def diag(self):
        out = np.empty((self.N, 4), np.float32)
        self._chk(lib().cc_diag(self.h, _ptr(out)), "cc_diag")
        return out

    def morton(self, core):
        perm = np.empty(self.N, np.int32)
        into = np.empty(self.N, np.int32)
        self._chk(lib().cc_morton(self.h, _d3(core), _ptr(perm, C.c_int32), _ptr(into, C.c_int32)), "cc_morton")
        return perm, into

    def analyse(self, m, a, spring, touch, cap=PP_CAP, diag=None, flags=0):
        """fof() of src/nbody.js:646-870 on the current state (cc_analyse): the
        report as a dict with the reference's names, pp.list an int32 array.
        diag: an N x 4 DIAG read-out to use instead of the one on the device."""
        d = None if diag is None else _f(diag).reshape(-1)
        if d is not None and d.size != self.N * 4:
            raise CCError("diag must hold N x 4 floats")
        args = AnalyseArgs(float(m), float(a), float(spring), float(touch), int(cap), int(flags), _ptr(d))
        rep = Report()
        lst = np.zeros(max(int(cap), 1), np.int32)
        self._chk(lib().cc_analyse(self.h, C.byref(args), C.byref(rep), _ptr(lst, C.c_int32)), "cc_analyse")
        return report_dict(rep, lst)

    def analyse_begin(self, m, a, spring, touch, cap=PP_CAP, diag=None, flags=0):
        """cc_analyse_begin: the analysis' device half now, its host sums on a thread
        (diag: an N x 4 DIAG read-out to use instead of the one on the device)."""
        d = None if diag is None else _f(diag).reshape(-1)
        if d is not None and d.size != self.N * 4:
            raise CCError("diag must hold N x 4 floats")
        args = AnalyseArgs(float(m), float(a), float(spring), float(touch), int(cap), int(flags), _ptr(d))
        self._chk(lib().cc_analyse_begin(self.h, C.byref(args)), "cc_analyse_begin")

    def analyse_end(self, wait=True):
        """cc_analyse_end: the begun analysis' report (as analyse()), or None while its sums still run."""
        rep = Report()
        lst = np.zeros(PP_CAP, np.int32)
        r = lib().cc_analyse_end(self.h, 1 if wait else 0, C.byref(rep), _ptr(lst, C.c_int32))
        if r == 1:
            return None
        self._chk(r, "cc_analyse_end")
        return report_dict(rep, lst)

    def autotune(self, reps=20):
        """cc_autotune at a mesh step: the chosen launch shapes and each candidate's time (ms)."""
        rep = TuneReport()
        self._chk(lib().cc_autotune(self.h, int(reps), C.byref(rep)), "cc_autotune")
        return {"cellBlocks": rep.cellBlocks, "simOrder": rep.simOrder, "cellMs": list(rep.cellMs),
                "simMs": list(rep.simMs)}

    def set_visit(self, every):
        """The step pass's visit order re-sorted every `every` steps (0: index order); results unchanged."""
        self._chk(lib().cc_set_visit(self.h, int(every)), "cc_set_visit")

    def set_tuning(self, cell_blocks, sim_order, fuse_count=1):
        self._chk(lib().cc_set_tuning(self.h, int(cell_blocks), int(sim_order), int(fuse_count)), "cc_set_tuning")

    BOOKS = ("M", "px", "py", "pz", "Jx", "Jy", "Jz", "KE", "PE", "Q", "iron", "unseated", "full")

    def books(self):
        out = np.empty(13, np.float64)
        self._chk(lib().cc_books(self.h, _ptr(out, C.c_double)), "cc_books")
        return dict(zip(self.BOOKS, out.tolist()))

    def sync(self):
        self._chk(lib().cc_sync(self.h), "cc_sync")


class Batch:
    """Instances stepped together, one graph launch per mesh cycle for the
    whole batch (cc_batch_*).  The Contexts must outlive the Batch."""

    def __init__(self, ctxs):
        self.ctxs = list(ctxs)
        arr = (C.c_void_p * len(self.ctxs))(*[c.h for c in self.ctxs])
        h = C.c_void_p()
        r = lib().cc_batch_create(arr, len(self.ctxs), C.byref(h))
        if r:
            raise CCError("cc_batch_create failed (%d): %s" % (r, lib().cc_batch_last_error(None).decode()))
        self.h = h

    def step_n(self, n, args=None, boxL=None):
        """args: one SimArgs for all or a list; boxL: a list (per instance) of (k, 7) arrays or None."""
        m = len(self.ctxs)
        if args is None:
            args = sim_args()
        if not isinstance(args, (list, tuple)):
            args = [args] * m
        arr = (SimArgs * m)(*args)
        if boxL is None:
            r = lib().cc_batch_step_n(self.h, n, arr, None, None)
        else:
            parts = [np.ascontiguousarray(b, np.float64).reshape(-1, 7) for b in boxL]
            flat = np.ascontiguousarray(np.concatenate(parts) if parts else np.zeros((0, 7)), np.float64)
            counts = np.array([len(p) for p in parts], np.int32)
            r = lib().cc_batch_step_n(self.h, n, arr, _ptr(flat, C.c_double), _ptr(counts, C.c_int32))
        if r:
            raise CCError("cc_batch_step_n failed (%d): %s" % (r, lib().cc_batch_last_error(self.h).decode()))
        ms = C.c_double()
        lib().cc_batch_last_ms(self.h, C.byref(ms))
        return ms.value

    def close(self):
        if self.h:
            lib().cc_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FarField:
    """The learned far field network (cc_ff_*)."""

    def __init__(self, capacity=0):
        h = C.c_void_p()
        r = lib().cc_ff_create(int(capacity), C.byref(h))
        if r:
            raise CCError("cc_ff_create failed (%d): %s" % (r, lib().cc_ff_last_error(None).decode()))
        self.h = h
        n = C.c_int64()
        lib().cc_ff_weight_count(C.byref(n))
        self.nw = n.value

    def _chk(self, r, what):
        if r:
            raise CCError("%s failed (%d): %s" % (what, r, lib().cc_ff_last_error(self.h).decode()))

    def set_slope(self, slope):
        """The leaky ReLU's negative slope (1: a linear operator)."""
        self._chk(lib().cc_ff_set_slope(self.h, float(slope)), "cc_ff_set_slope")

    def weights(self):
        w = np.empty(self.nw, np.float32)
        self._chk(lib().cc_ff_get_weights(self.h, _ptr(w), self.nw), "cc_ff_get_weights")
        return w

    def set_weights(self, w):
        w = np.ascontiguousarray(w, np.float32)
        self._chk(lib().cc_ff_set_weights(self.h, _ptr(w), w.size), "cc_ff_set_weights")

    def record(self, ctx, box=0):
        """Record box `box` of ctx's last (exact) mesh pass; returns the pass index."""
        self._chk(lib().cc_ff_record(self.h, ctx.h, int(box)), "cc_ff_record")
        return self.passes - 1

    @property
    def passes(self):
        n = C.c_int()
        self._chk(lib().cc_ff_passes(self.h, C.byref(n)), "cc_ff_passes")
        return n.value

    def read_pass(self, k):
        """(feat[4096,4], meta[4], record[4096,12]) of recorded pass k."""
        f, m, r = np.empty((4096, 4), np.float32), np.empty(4, np.float32), np.empty((4096, 12), np.float32)
        self._chk(lib().cc_ff_read_pass(self.h, int(k), _ptr(f), _ptr(m), _ptr(r)), "cc_ff_read_pass")
        return f, m, r

    def clear_passes(self):
        self._chk(lib().cc_ff_clear_passes(self.h), "cc_ff_clear_passes")

    def pair(self, base, cur):
        """A training sample: the prediction from recorded pass `base` to pass `cur`; returns its index."""
        k = C.c_int()
        self._chk(lib().cc_ff_pair(self.h, int(base), int(cur), C.byref(k)), "cc_ff_pair")
        return k.value

    def push(self, x, target, mask):
        self._chk(lib().cc_ff_push_sample(self.h, _ptr(_f(x)), _ptr(_f(target)), _ptr(_f(mask))), "cc_ff_push_sample")

    def sample(self, k):
        x, t, m = np.empty((4096, 16), np.float32), np.empty((4096, 16), np.float32), np.empty(4096, np.float32)
        self._chk(lib().cc_ff_read_sample(self.h, int(k), _ptr(x), _ptr(t), _ptr(m)), "cc_ff_read_sample")
        return x, t, m

    @property
    def count(self):
        n = C.c_int()
        self._chk(lib().cc_ff_samples(self.h, C.byref(n)), "cc_ff_samples")
        return n.value

    def clear(self):
        self._chk(lib().cc_ff_clear(self.h), "cc_ff_clear")

    def train(self, iters, batch=16, lr=1e-3, seed=1):
        losses = np.empty(int(iters), np.float64)
        self._chk(lib().cc_ff_train(self.h, int(iters), int(batch), float(lr), int(seed), _ptr(losses, C.c_double)),
                  "cc_ff_train")
        return losses

    def eval(self, first=0, count=None, outputs=False):
        count = self.count - first if count is None else count
        out = np.empty((count, 4096, 16), np.float32) if outputs else None
        ra, rl = C.c_double(), C.c_double()
        self._chk(lib().cc_ff_eval(self.h, int(first), int(count), _ptr(out), C.byref(ra), C.byref(rl)), "cc_ff_eval")
        return (ra.value, rl.value, out) if outputs else (ra.value, rl.value)

    def gradient(self, first, count):
        g = np.empty(self.nw, np.float32)
        loss = C.c_double()
        self._chk(lib().cc_ff_gradient(self.h, int(first), int(count), _ptr(g), C.byref(loss)), "cc_ff_gradient")
        return g, loss.value

    def close(self):
        if self.h:
            lib().cc_ff_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def pp_noalloc(pos, mat, N, LA, LB, boxB, index, count, m, a, tab, force, stream=None):
    """cc_pp_noalloc: the pairwise pass on device arrays (addresses as integers:
    pos float[N*4], mat float[N], index int32[count], tab float[128], force
    float[4*8192*4] out), queued on `stream` (a hipStream_t address, None: the
    null stream); no allocation, no wait."""
    r = lib().cc_pp_noalloc(C.c_void_p(pos), C.c_void_p(mat), int(N), _d3(LA), _d3(LB), int(bool(boxB)),
                            C.c_void_p(index), int(count), float(m), float(a), C.c_void_p(tab), C.c_void_p(force),
                            C.c_void_p(stream) if stream else None)
    if r:
        raise CCError("cc_pp_noalloc failed (%d): %s" % (r, lib().cc_last_error(None).decode()))


def relax_piles(piles):
    """Relaxes each (xyz float32 array of shape (n, 3), a, A, C, confine, sweeps) in place, the
    piles side by side (cc_relax_piles); returns [(sweeps, worst)]."""
    for q in piles:
        a = q[0]
        if not (isinstance(a, np.ndarray) and a.dtype == np.float32 and a.flags.c_contiguous and a.ndim == 2
                and a.shape[1] == 3):
            raise CCError("each pile must be a C-contiguous float32 (n, 3) array, relaxed in place")
    ps = (Pile * max(len(piles), 1))()
    for k, q in enumerate(piles):
        a = q[0]
        ps[k] = Pile(_ptr(a), len(a), int(q[5]), float(q[1]), float(q[2]), float(q[3]), int(q[4]), 0)
    reps = (PileReport * max(len(piles), 1))()
    r = lib().cc_relax_piles(ps, len(piles), reps)
    if r:
        raise CCError("cc_relax_piles failed (%d): %s" % (r, lib().cc_bodies_last_error().decode()))
    return [(reps[k].sweeps, reps[k].worst) for k in range(len(piles))]


def crater_stats(pos, vel, R, mref, **args):
    """craterStats() of chicxulub.html:848-919 (cc_crater_stats) on host arrays: the report as a dict."""
    pos, vel = (np.ascontiguousarray(x, np.float32).reshape(-1, 4) for x in (pos, vel))
    R = np.ascontiguousarray(R, np.float32).ravel()
    n = len(R)
    if len(pos) != n or len(vel) != n:
        raise CCError("pos, vel and R must hold the same particles")
    a = CraterArgs(*[float(args[k]) for k in ("Rbox", "lSI", "vSI", "gSI", "rhoSI", "vOrb", "vEsc")])
    rep = CraterReport()
    r = lib().cc_crater_stats(n, float(mref), _ptr(pos), _ptr(vel), _ptr(R), C.byref(a), C.byref(rep))
    if r:
        raise CCError("cc_crater_stats failed (%d)" % r)
    d = _plain(rep)
    d["vel"] = [0.0, 0.0, 0.0]
    return d
