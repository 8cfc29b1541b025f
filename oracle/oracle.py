"""ctypes front end of the CPU oracle (libcc_oracle.so) -- TEST INFRASTRUCTURE.

Only tests, the smoke check and the benchmark's CPU-baseline leg import this
module, and only as the checker.  Arrays are numpy, in the reference's texture
layouts (see cc_oracle.h).  Uniform dictionaries carry the float values the
reference's shaders receive (names as in src/glsl-sim.js).
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")


class Grid(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("G", "GShift", "GMask", "SX", "SXShift", "SXMask", "AW", "AH", "IB")]


class Box(C.Structure):
    _fields_ = [("invH", C.c_float), ("LA", C.c_float * 3), ("LB", C.c_float * 3), ("boxB", C.c_float)]


class SimU(C.Structure):
    _fields_ = ([("W", C.c_int), ("H", C.c_int)]
                + [(n, C.c_float) for n in ("invCell", "cell", "dt", "gdt", "touch", "link2", "reach2", "k", "c",
                                            "invM", "damp", "creep", "tK", "kappa", "cpIron", "settle", "kick")]
                + [(n, C.c_float * 3) for n in ("bulk0", "bulk1", "cen0", "cen1", "spin0", "spin1")]
                + [("grid", Grid), ("box", Box), ("gm", C.c_float), ("eps2", C.c_float), ("field", C.c_int),
                   ("gField", C.c_float * 3), ("frozenOn", C.c_int), ("frozen", C.c_float)])


def lib():
    global _LIB
    if _LIB is None:
        _LIB = _load(os.path.join(_HERE, "libcc_oracle.so"))
    return _LIB


class using:
    """Within the block, the oracle functions call the library at `path`
    (a build with a planted mistake, `make -C oracle mutants`) instead."""

    def __init__(self, path):
        self.path = path

    def __enter__(self):
        global _LIB
        lib()
        self.saved, _LIB = _LIB, _load(self.path)
        return self

    def __exit__(self, *exc):
        global _LIB
        _LIB = self.saved
        return False


def _load(path):
    if not os.path.exists(path):
        raise RuntimeError("oracle library not built: run `make -C oracle`")
    L = C.CDLL(path)
    L.cco_grid_make.argtypes = [C.c_int, C.POINTER(Grid)]
    L.cco_hash_at.argtypes = [C.POINTER(Grid), C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.cco_slots.argtypes = [f32p, C.c_int, C.c_float, C.POINTER(Grid), u32p, u32p]
    L.cco_box_of.argtypes = [C.POINTER(Box), f32p, f32p]
    L.cco_box_of.restype = C.c_int
    L.cco_deposit.argtypes = [f32p, f32p, C.c_int, C.POINTER(Box), C.c_int, f32p, C.c_float, f32p]
    L.cco_coarse.argtypes = [f32p, f32p]
    L.cco_mono.argtypes = [f32p, f32p]
    L.cco_far.argtypes = [f32p, f32p, f32p, f32p]
    L.cco_cell.argtypes = [f32p, f32p, f32p, f32p, f32p, C.c_float, C.c_float, f32p]
    L.cco_cell_f64.argtypes = [f32p, f32p, f32p, f32p, f32p, C.c_double, C.c_double, f64p]
    L.cco_pp_gather.argtypes = [f32p, f32p, f32p, C.POINTER(Box), f32p, f32p]
    L.cco_pp_force.argtypes = [f32p, f32p, C.c_int, C.c_float, C.c_float, C.c_float, f32p, f32p]
    L.cco_sim.argtypes = [C.POINTER(SimU)] + [f32p] * 4 + [u32p, u32p] + [f32p] * 6 + [f32p] * 4
    L.cco_rigid.argtypes = [C.c_int] + [f32p] * 4 + [f32p] * 8 + [C.c_float] + [f32p] * 4
    L.cco_perm.argtypes = [C.c_int] + [f32p] * 5 + [f32p] * 4
    L.cco_diag.argtypes = [C.POINTER(SimU), f32p, f32p, u32p, u32p] + [f32p] * 6 + [f32p]
    L.cco_couple.argtypes = [C.c_int, f32p] + [f32p] * 4 + [f32p] * 4
    L.cco_morton.argtypes = [f32p, C.c_int, f64p, i32p, f32p, i32p]
    L.cco_books.argtypes = [C.c_int, C.c_double, f32p, f32p, f32p, f32p, f64p]
    L.cco_mulberry_next.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_double)]
    L.cco_mulberry_next.restype = C.c_uint32
    L.cco_direct_f64.argtypes = [C.c_int, f32p, f32p, C.c_double, C.c_double, f64p]
    L.cco_mesh_pair_table.argtypes = [C.c_int, C.c_double, f32p, f32p]
    L.cco_fof_link.argtypes = [C.c_int, f32p, C.c_double, C.c_double, C.c_double, i32p, i32p, C.c_void_p, C.c_long,
                               C.POINTER(C.c_double), C.c_void_p, C.c_void_p, C.c_long]
    L.cco_fof_link.restype = C.c_long
    return L


def c32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def grid(G, ib=18):
    g = Grid()
    lib().cco_grid_make(G, C.byref(g))
    g.IB = ib
    return g


def box(invH, LA, LB, boxB):
    b = Box()
    b.invH = invH
    b.LA[:] = [float(x) for x in LA]
    b.LB[:] = [float(x) for x in LB]
    b.boxB = boxB
    return b


def hash_at(g, cx, cy, cz):
    ax, ay = C.c_int(), C.c_int()
    lib().cco_hash_at(C.byref(g), int(cx), int(cy), int(cz), C.byref(ax), C.byref(ay))
    return ax.value, ay.value


def slots(pos, inv_cell, g):
    n = g.AW * g.AH * 4
    a = np.empty(n, np.uint32)
    b = np.empty(n, np.uint32)
    lib().cco_slots(c32(pos).ravel(), pos.size // 4, inv_cell, C.byref(g), a, b)
    return a.reshape(g.AH, g.AW, 4), b.reshape(g.AH, g.AW, 4)


def deposit(pos, mat, bx, which, mono, eps2):
    fine = np.empty((513, 512, 4), np.float32)
    lib().cco_deposit(c32(pos).ravel(), c32(mat).ravel(), pos.size // 4, C.byref(bx), which, c32(mono).ravel(), eps2,
                      fine)
    return fine


def coarse(fine):
    out = np.empty((64, 64, 4), np.float32)
    lib().cco_coarse(c32(fine), out)
    return out


def mono(coarse_):
    out = np.empty(4, np.float32)
    lib().cco_mono(c32(coarse_), out)
    return out


def far(coarse_):
    A = np.empty((64, 64, 4), np.float32)
    T0 = np.empty((64, 64, 4), np.float32)
    T1 = np.empty((64, 64, 4), np.float32)
    lib().cco_far(c32(coarse_), A, T0, T1)
    return A, T0, T1


def cell(fine, coarse_, farA, farT0, farT1, eps2, gm):
    out = np.empty((513, 512, 4), np.float32)
    lib().cco_cell(c32(fine), c32(coarse_), c32(farA), c32(farT0), c32(farT1), eps2, gm, out)
    return out


def cell_f64(fine, coarse_, farA, farT0, farT1, eps2, gm):
    out = np.empty((513, 512, 4), np.float64)
    lib().cco_cell_f64(c32(fine), c32(coarse_), c32(farA), c32(farT0), c32(farT1), eps2, gm, out)
    return out


def pp_gather(pos, mat, index, bx):
    lst = np.empty((64, 128, 4), np.float32)
    bo = np.empty((64, 128), np.float32)
    lib().cco_pp_gather(c32(pos).ravel(), c32(mat).ravel(), c32(index).ravel(), C.byref(bx), lst, bo)
    return lst, bo


def pp_force(lst, bo, count, gm, cut, a2, tab):
    out = np.empty((256, 128, 4), np.float32)
    lib().cco_pp_force(c32(lst), c32(bo), count, gm, cut, a2, c32(tab).ravel(), out)
    return out


def sim(u, pos, vel, mat, q, seatA, seatB, force0, mono0, force1, mono1, ppSlot, ppForce):
    N = u.W * u.H
    oP = np.empty((N, 4), np.float32)
    oV = np.empty((N, 4), np.float32)
    oA = np.empty((N, 4), np.float32)
    oQ = np.empty(N, np.float32)
    lib().cco_sim(C.byref(u), c32(pos).ravel(), c32(vel).ravel(), c32(mat).ravel(), c32(q).ravel(),
                  np.ascontiguousarray(seatA, np.uint32).ravel(), np.ascontiguousarray(seatB, np.uint32).ravel(),
                  c32(force0).ravel(), c32(mono0).ravel(), c32(force1).ravel(), c32(mono1).ravel(),
                  c32(ppSlot).ravel(), c32(ppForce).ravel(), oP, oV, oA, oQ)
    return oP, oV, oA, oQ


def diag(u, pos, q, seatA, seatB, force0, mono0, force1, mono1, ppSlot, ppForce):
    N = u.W * u.H
    out = np.empty((N, 4), np.float32)
    lib().cco_diag(C.byref(u), c32(pos).ravel(), c32(q).ravel(), np.ascontiguousarray(seatA, np.uint32).ravel(),
                   np.ascontiguousarray(seatB, np.uint32).ravel(), c32(force0).ravel(), c32(mono0).ravel(),
                   c32(force1).ravel(), c32(mono1).ravel(), c32(ppSlot).ravel(), c32(ppForce).ravel(), out)
    return out


def rigid(pos, vel, aux, q, bulk0, bulk1, cen0, cen1, spin0, spin1, rot0, rot1, dt):
    N = pos.size // 4
    outs = [np.empty((N, 4), np.float32) for _ in range(3)] + [np.empty(N, np.float32)]
    lib().cco_rigid(N, c32(pos).ravel(), c32(vel).ravel(), c32(aux).ravel(), c32(q).ravel(),
                    *[c32(x) for x in (bulk0, bulk1, cen0, cen1, spin0, spin1, rot0, rot1)], dt, *outs)
    return outs


def perm(perm_, pos, vel, aux, q):
    N = pos.size // 4
    outs = [np.empty((N, 4), np.float32) for _ in range(3)] + [np.empty(N, np.float32)]
    lib().cco_perm(N, c32(perm_), c32(pos).ravel(), c32(vel).ravel(), c32(aux).ravel(), c32(q).ravel(), *outs)
    return outs


def couple(args, pos, vel, aux, mat):
    """args: centre[3], vimp, aimp, c1, invMu, pexp, stall, z, heat, tK, frozen, ambient, melt"""
    N = pos.size // 4
    outs = [np.empty((N, 4), np.float32) for _ in range(3)] + [np.empty(N, np.float32)]
    lib().cco_couple(N, c32(args), c32(pos).ravel(), c32(vel).ravel(), c32(aux).ravel(), c32(mat).ravel(), *outs)
    return outs


def morton(pos, core):
    N = pos.size // 4
    keys = np.empty(N, np.int32)
    pm = np.empty(N, np.float32)
    into = np.empty(N, np.int32)
    lib().cco_morton(c32(pos).ravel(), N, np.ascontiguousarray(core, np.float64), keys, pm, into)
    return keys, pm, into


BOOK_NAMES = ("M", "px", "py", "pz", "Jx", "Jy", "Jz", "KE", "PE", "Q", "unseated", "full", "iron", "drift")


def books(mref, pos, vel, mat, diag_):
    out = np.empty(14, np.float64)
    lib().cco_books(pos.size // 4, mref, c32(pos).ravel(), c32(vel).ravel(), c32(mat).ravel(), c32(diag_).ravel(), out)
    return dict(zip(BOOK_NAMES, out.tolist()))


def direct_f64(pos, mat, gm, a2):
    """FP64 direct-sum gravity softened at a: (N, 4) acc xyz, potential."""
    N = pos.size // 4
    out = np.empty((N, 4), np.float64)
    lib().cco_direct_f64(N, c32(pos).ravel(), c32(mat).ravel(), gm, a2, out)
    return out


def mulberry(seed, n):
    st = C.c_uint32(seed)
    d = C.c_double()
    out = []
    for _ in range(n):
        lib().cco_mulberry_next(C.byref(st), C.byref(d))
        out.append(d.value)
    return out


def mesh_pair_table(bins=64, cut=5 * 12.0 / 64):
    f = np.empty(bins, np.float32)
    p = np.empty(bins, np.float32)
    lib().cco_mesh_pair_table(bins, cut, f, p)
    return np.stack([f, p], axis=1)


def fof_link(pos, a, touch, spring, pairs=False):
    """Friends-of-friends linking as the reference runs it (src/nbody.js:647-684):
    {root, label (smallest index of the group), e2 (touching pairs' squared
    distances in pair order), elastic, and with pairs linkOff/partner}."""
    P = c32(pos).ravel()
    N = P.size // 4
    root = np.empty(N, np.int32)
    label = np.empty(N, np.int32)
    el = C.c_double()
    nt = lib().cco_fof_link(N, P, a, touch, spring, root, label, None, 0, C.byref(el), None, None, 0)
    e2 = np.empty(max(nt, 1), np.float64)
    off = np.empty(N + 1, np.int32) if pairs else None
    lib().cco_fof_link(N, P, a, touch, spring, root, label, e2.ctypes.data, nt, C.byref(el),
                       off.ctypes.data if pairs else None, None, 0)
    out = {"root": root, "label": label, "e2": e2[:nt], "elastic": el.value}
    if pairs:
        npairs = int(off[N])
        partner = np.empty(max(npairs, 1), np.int32)
        lib().cco_fof_link(N, P, a, touch, spring, root, label, None, 0, C.byref(el), off.ctypes.data,
                           partner.ctypes.data, npairs)
        out["linkOff"], out["partner"] = off, partner[:npairs]
    return out
