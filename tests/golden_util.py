"""Loading the golden fixtures (tests/golden/fixtures/*.npz) and turning the
uniform values the reference's passes received into oracle / library
arguments.  Test infrastructure."""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
FIX = os.path.join(HERE, "golden", "fixtures")
sys.path.insert(0, os.path.join(REPO, "oracle"))

FL = 6.0
FH = 2 * FL / 64
EMPTY = np.uint32(0xBF800000)


def fixture(name):
    return dict(np.load(os.path.join(FIX, name + ".npz"), allow_pickle=False))


def dense(fx, name, shape, empty=0):
    """Expand an (index, value) sparse texture back to its full atlas."""
    idx, val = fx[name + "_idx"], fx[name + "_val"]
    out = np.full((shape[0] * shape[1], 4), empty, dtype=val.dtype)
    out[idx] = val
    return out.reshape(shape[0], shape[1], 4)


def seats(fx, AW, AH):
    return (dense(fx, "seatA", (AH, AW), EMPTY), dense(fx, "seatB", (AH, AW), EMPTY))


def mesh(fx, b):
    return {"fine": dense(fx, "fine%d" % b, (513, 512)), "coarse": dense(fx, "coarse%d" % b, (64, 64)),
            "mono": dense(fx, "mono%d" % b, (1, 1)).reshape(4), "farA": dense(fx, "farA%d" % b, (64, 64)),
            "farT0": dense(fx, "farT0%d" % b, (64, 64)), "farT1": dense(fx, "farT1%d" % b, (64, 64))}


def draws(fx, kind):
    return [d["u"] for d in json.loads(str(fx["draws"])) if d["kind"] == kind]


def f(u, name, i=0):
    return np.float32(u[name][i])


def v3(u, name):
    return [float(np.float32(x)) for x in u[name][:3]]


def sim_u(u, mode="mesh"):
    """oracle.SimU from the uniform values the step pass (or DIAG) received."""
    import oracle
    s = oracle.SimU()
    s.W, s.H = int(u["uSize"][0]), int(u["uSize"][1])
    for nm, key in (("invCell", "uInvCell"), ("cell", "uCell"), ("dt", "uDt"), ("gdt", "uGDt"), ("touch", "uTouch"),
                    ("link2", "uLink2"), ("reach2", "uReach2"), ("k", "uK"), ("c", "uC"), ("invM", "uInvM"),
                    ("damp", "uDamp"), ("creep", "uCreep"), ("tK", "uTk"), ("kappa", "uKappa"),
                    ("cpIron", "uCpIron"), ("settle", "uSettle"), ("kick", "uKick")):
        if key in u:
            setattr(s, nm, float(f(u, key)))
    for nm, key in (("bulk0", "uBulk0"), ("bulk1", "uBulk1"), ("cen0", "uCen0"), ("cen1", "uCen1"),
                    ("spin0", "uSpin0"), ("spin1", "uSpin1")):
        if key in u:
            getattr(s, nm)[:] = v3(u, key)
    G = int(u["uGMask"][0]) + 1
    s.grid = oracle.grid(G)
    if mode == "mesh":
        s.box = oracle.box(float(f(u, "uInvH")), v3(u, "uLA"), v3(u, "uLB"), float(f(u, "uBoxB")))
        s.gm = float(f(u, "uGm"))
        s.eps2 = float(f(u, "uEps2"))
        s.field = 0
    else:
        s.field = 1
        s.gField[:] = v3(u, "uGField")
    if "uFrozen" in u:
        s.frozenOn = 1
        s.frozen = float(f(u, "uFrozen"))
    return s


CRATER_STEPS = (("b0", "s1", 1), ("s1", "s2", 1), ("post", "f1", 0), ("f1", "f2", 0))


def crater_state(fx, tag):
    """(pos, vel, q) of a crater dump; the build's q is zero."""
    pos = fx[tag + "_pos"].reshape(-1, 4)
    vel = fx[tag + "_vel"].reshape(-1, 4)
    q = fx[tag + "_q"].reshape(-1, 4)[:, 0] if tag + "_q" in fx else np.zeros(pos.shape[0], np.float32)
    return pos, vel, q


def crater_host(fx, op):
    return [o for o in json.loads(str(fx["host"])) if o["op"] == op]


def box_from(u):
    import oracle
    return oracle.box(float(f(u, "uInvH")), v3(u, "uLA"), v3(u, "uLB"), float(f(u, "uBoxB")))


def pp_slot(N, index):
    s = np.full(N, -1.0, np.float32)
    s[index] = np.arange(len(index), dtype=np.float32)
    return s


def rel_err(x, ref, floor_frac=1e-3):
    """|x - ref| / max(|ref|, floor_frac * rms(ref)) per component."""
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    rms = np.sqrt(np.mean(ref ** 2)) if ref.size else 0.0
    den = np.maximum(np.abs(ref), floor_frac * rms + 1e-300)
    return np.abs(x - ref) / den


def rel_err_vec(x, ref, floor_frac=1e-3):
    """Per-vector relative error |x - ref| / max(|ref|, floor_frac * rms|ref|)
    over the last axis (3-vectors: accelerations, velocities)."""
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    n = np.linalg.norm(ref, axis=-1)
    rms = np.sqrt(np.mean(n ** 2)) if n.size else 0.0
    return np.linalg.norm(x - ref, axis=-1) / np.maximum(n, floor_frac * rms + 1e-300)
