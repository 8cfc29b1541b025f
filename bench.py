"""Benchmark: particle-steps/s of the plain simulation step at 262,144 particles
(the 'twins' giant impact) on one MI355X.

A step is the reference's step() (src/nbody.js:551-558): contact seats every
second step, the two-box mesh and the pairwise correction every eighth, the
contact / heat / gravity / symplectic-Euler pass every step.  The state is
resident in HBM before the timed region; K steps are timed between two
device synchronisations, and the step kernel's own launches are timed with
HIP events on the library's stream.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import math
import os
import subprocess
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(REPO, "cosmic-collisions_amd")
sys.path.insert(0, PKG)

FL = 6.0
FH = 2 * FL / 64
HBM_PEAK_GBS = 8000.0


def twins_scalars(N=262144, q=1.0, dens=0.75, spin0=0.18, spin1=0.20, vfac=1.05, rest=0.3):
    """The sim scalars build() derives for the twins config (src/nbody.js:114-206)."""
    V1, V2 = 1.0, q / dens
    vp = (V1 + V2) / N
    s = ((16 * math.pi * vp / 3) ** (1.0 / 3.0))
    a = 0.96 * s / (2 * math.sqrt(2))
    n1 = round(N * V1 / (V1 + V2))
    M = 2.0                      # mass ratio 1: both bodies one Earth mass in the reference's units
    mref = 1.0 / (n1 * (0.325 / (2.5 - 0.325 * 1.5) * 2.5 + (1 - 0.325 / (2.5 - 0.325 * 1.5))))
    Rsum = 1.0 + ((V2) ** (1.0 / 3.0))
    vesc = math.sqrt(2 * M / Rsum)
    vc = vfac * vesc
    vref = max(1.5, vc)
    k = 4 * mref * vref * vref / (a * a)
    dt = 0.35 / math.sqrt(2 * k / (mref * min(1, dens)))
    lnE = math.log(rest)
    zeta = -lnE / math.sqrt(math.pi ** 2 + lnE ** 2)
    c = 2 * zeta * math.sqrt(k * mref / 2)
    return dict(N=N, W=512, G=128, s=s, a=a, n1=n1, m=mref, cell=2.6 * a, dt=dt, k=k, c=c, vc=vc,
                R1=1.0, R2=((V2) ** (1.0 / 3.0)), spin0=(0, spin0, 0), spin1=(0, spin1 * math.sqrt(dens), 0))


def lattice_ball(n, s, rng):
    """n points of a jittered FCC lattice of cube edge s, nearest to the centre."""
    r = (3 * n * s ** 3 / (16 * math.pi)) ** (1 / 3) * 1.1
    m = int(math.ceil(r / s)) + 1
    g = np.arange(-m, m + 1) * s
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    base = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    offs = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]]) * s
    pts = (base[:, None, :] + offs[None]).reshape(-1, 3)
    d = np.linalg.norm(pts, axis=1)
    pts = pts[np.argsort(d, kind="stable")[:n]]
    return pts + rng.uniform(-0.03, 0.03, pts.shape) * s


def synthetic_twins(sc, seed=0xC0FFEE):
    """Two lattice bodies of the twins' sizes, touching at 30 degrees, closing at
    the contact speed -- the configuration a step sees just after the kick."""
    rng = np.random.default_rng(seed)
    N, n1 = sc["N"], sc["n1"]
    b1 = lattice_ball(n1, sc["s"], rng)
    b2 = lattice_ball(N - n1, sc["s"], rng)
    ang = math.radians(30)
    dist = sc["R1"] + sc["R2"] - 0.05
    off = np.array([-dist * math.cos(ang), dist * math.sin(ang), 0.0])
    c0 = -off * 0.5
    c1 = off * 0.5
    pos = np.zeros((N, 4), np.float32)
    pos[:n1, :3] = b1 + c0
    pos[n1:, :3] = b2 + c1
    r1 = np.linalg.norm(b1, axis=1)
    r2 = np.linalg.norm(b2, axis=1)
    pos[:n1, 3] = np.where(r1 < 0.66 * sc["R1"], 0, np.where(r1 > 0.97 * sc["R1"], 2, 1))
    pos[n1:, 3] = 4 + np.where(r2 < 0.66 * sc["R2"], 0, np.where(r2 > 0.97 * sc["R2"], 2, 1))
    mat = np.where(pos[:, 3] % 4 < 0.5, 2.5, 1.0).astype(np.float32)
    mat[n1:] *= 0.75
    vel = np.zeros((N, 4), np.float32)
    vrel = np.array([sc["vc"], 0, 0]) * 0.5
    vel[:n1, :3] = vrel
    vel[n1:, :3] = -vrel
    w0, w1 = np.array(sc["spin0"]), np.array(sc["spin1"])
    vel[:n1, :3] += np.cross(w0, pos[:n1, :3] - c0)
    vel[n1:, :3] += np.cross(w1, pos[n1:, :3] - c1)
    return pos, vel, mat, np.zeros(N, np.float32), c0, c1


def morton_order(ctx, core):
    ctx.morton(core)


def box_uniform(cen):
    L = [FL - float(x) for x in cen]
    return L + [FL, FL, FL] + [0.0]


def cpu_baseline(sc, pos, vel, mat, q, budget_s=15.0):
    """The CPU oracle's step pass (one core) on the same state: particle-steps/s."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle
    subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle")], check=True)
    g = oracle.grid(sc["G"])
    inv = np.float32(1.0 / sc["cell"])
    seatA, seatB = oracle.slots(pos, inv, g)
    u = oracle.SimU()
    u.W, u.H = sc["W"], sc["N"] // sc["W"]
    for nm, v in (("invCell", 1 / sc["cell"]), ("cell", sc["cell"]), ("dt", sc["dt"]), ("gdt", 0.0),
                  ("touch", 2 * sc["a"]), ("link2", sc["cell"] ** 2), ("reach2", (sc["cell"] + 0.6 * sc["a"]) ** 2),
                  ("k", sc["k"]), ("c", sc["c"]), ("invM", 1 / sc["m"]), ("damp", math.exp(-sc["dt"] * 12)),
                  ("creep", 0.12), ("tK", 0.5 * 6.256e7 / 1200 / sc["m"]), ("kappa", 0.3), ("cpIron", 0.375)):
        setattr(u, nm, float(np.float32(v)))
    u.grid = g
    u.box = oracle.box(np.float32(1 / FH), [FL] * 3, [FL] * 3, 0.0)
    u.gm = float(np.float32(sc["m"]))
    u.eps2 = float(np.float32(FH * FH))
    zero_mesh = np.zeros((513, 512, 4), np.float32)
    mono = np.zeros(4, np.float32)
    slot = np.full(sc["N"], -1, np.float32)
    ppf = np.zeros((256, 128, 4), np.float32)
    n, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < budget_s or n == 0:
        oracle.sim(u, pos, vel, mat, q, seatA, seatB, zero_mesh, mono, zero_mesh, mono, slot, ppf)
        n += 1
    el = time.perf_counter() - t0
    return {"value": n * sc["N"] / el, "unit": "particle-steps/s", "cores": 1, "kind": "port",
            "sample": "%d step passes (contacts, heat, gravity gather, integration) of the oracle on the same "
                      "262,144-particle state, one core, %.1f s" % (n, el)}


TWINS_OPT = {"nIdx": 4, "q": 1.0, "dens": 0.75, "core": 0.325, "spin0": 0.18, "spin1": 0.20, "angle": 30, "incl": 0,
             "vfac": 1.05, "rest": 0.3}


def prepare_twins(out_dir, extra_steps=0, reuse=False):
    """The twins run through the JavaScript engine's own phases (build, settle,
    approach, kick, a look with the Morton sort) on this GPU, written to out_dir
    (or read from there when reuse is set and a run for these options is there)."""
    have = os.path.exists(os.path.join(out_dir, "run.json")) and \
        json.load(open(os.path.join(out_dir, "run.json"))).get("opt") == TWINS_OPT
    if not (reuse and have):
        cmd = ["node", os.path.join(PKG, "js", "prep_state.js"), json.dumps(TWINS_OPT), out_dir, str(extra_steps)]
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
        sys.stderr.write(r.stdout[-3000:])
        if r.returncode != 0:
            raise RuntimeError("state preparation failed:\n" + r.stderr[-4000:])
    run = json.load(open(os.path.join(out_dir, "run.json")))
    N = run["sim"]["N"]
    st = {k: np.fromfile(os.path.join(out_dir, "state.%s.f32" % k), np.float32) for k in ("pos", "vel", "q", "mat")}
    st["pos"] = st["pos"].reshape(N, 4)
    st["vel"] = st["vel"].reshape(N, 4)
    st["pptab"] = np.fromfile(os.path.join(out_dir, "pptab.f32"), np.float32).reshape(64, 2)
    st["pplist"] = np.fromfile(os.path.join(out_dir, "pplist.i32"), np.int32)
    return run, st


class Runner:
    """The engine's full-phase step loop over the C ABI: boxes() for every
    mesh step, the softening an hour and a half after first touch
    (src/nbody.js:538-558), sim.t advanced per step as the caller does."""

    def __init__(self, ccamd, run, st):
        s = run["sim"]
        self.ccamd, self.s = ccamd, dict(s)
        self.p = ccamd.params(N=s["N"], W=s["W"], G=s["G"], m=s["m"], a=s["a"], cell=s["cell"], dt=s["dt"], k=s["k"],
                              c=s["c"], creep=s["creep"], damp=s["damp"], tK=s["tK"], kappa=s["kappa"],
                              spin0=s["spin0"], spin1=s["spin1"])
        self.ctx = ccamd.Context(self.p)
        self.ctx.upload(st["pos"], st["vel"], st["mat"], st["q"])
        self.ctx.set_pp_table(st["pptab"])
        self.ctx.set_pp(st["pplist"])
        self.args = ccamd.sim_args(bulk0=s["bulk0"], bulk1=s["bulk1"], cen0=s["cen0"], cen1=s["cen1"])
        self.t, self.step_no = s["t"], s["stepNo"]

    def soften_if_due(self):
        s = self.s
        if not s["soft"] and self.t - s["impactT"] > 1.5:
            s["soft"] = True
            s["k"] *= 0.5
            s["c"] *= math.sqrt(0.5)
            s["dt"] *= math.sqrt(2.0)
            s["damp"] = math.exp(-s["dt"] * 12)
            self.p.k, self.p.c, self.p.dt, self.p.damp = s["k"], s["c"], s["dt"], s["damp"]
            self.ctx.set_params(self.p)

    def steps(self, n):
        """n steps; returns the device milliseconds (HIP events around the launches)."""
        ms, done = 0.0, 0
        while done < n:
            self.soften_if_due()
            s, t, chunk, L = self.s, self.t, 0, []
            # run until the softening would fire (it is checked at each step's start)
            while done + chunk < n and (chunk == 0 or s["soft"] or t - s["impactT"] <= 1.5):
                if (self.step_no + chunk) % 8 == 0:
                    for b in s["boxes"]:
                        dtb = t - b["t"]
                        L += [FL - b["cen"][k] - b["vel"][k] * dtb for k in range(3)]
                    L.append(1.0 if s["boxes"][1]["on"] else 0.0)
                t += s["dt"]
                chunk += 1
            self.ctx.step_no = self.step_no
            ms += self.ctx.step_n(chunk, self.args, np.array(L, np.float64).reshape(-1, 7) if L else None)
            self.t, self.step_no, done = t, self.step_no + chunk, done + chunk
        return ms


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--state-dir", default=os.path.join(REPO, "gpurun_out", "twins_state"))
    ap.add_argument("--reuse-state", action="store_true", help="use a prepared state in --state-dir if there is one")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (each box has one MI355X)", file=sys.stderr)
        return 2
    import ccamd

    run, st = prepare_twins(args.state_dir, reuse=args.reuse_state)
    import torch
    R = Runner(ccamd, run, st)
    N = run["sim"]["N"]
    R.steps(args.warmup)
    torch.cuda.synchronize()
    R.ctx.sync()
    t0 = time.perf_counter()
    dev_ms = R.steps(args.steps)
    R.ctx.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    value = N * args.steps / wall
    # the step pass alone (an odd step: no seats, no mesh), timed with HIP events on the library's stream
    ev = []
    saved = R.step_no
    for _ in range(50):
        R.ctx.step_no = 1
        ev.append(R.ctx.step_n(1, R.args))
    R.ctx.step_no = saved
    sim_ms = float(np.median(ev))
    alg_bytes = 92.0 * N     # pos+vel+mat+q read, pos+vel+aux+q written: each particle's own data once
    achieved = alg_bytes / (sim_ms * 1e-3) / 1e9
    out = {
        "metric": "particle-steps/s at 262k, plain vs batched/learned far field @<=1e-3 force err",
        "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "plain twins giant impact: 262,144 particles, two 0.5 M_earth differentiated bodies "
                               "at 30 deg and 1.05 v_esc; P3M 64^3 two-box mesh with pairwise correction, "
                               "spring-dashpot contacts with heat, contact grid G=128",
                   "particles": N, "path": "plain", "state": "the engine's own build/settle/approach/kick, "
                                                              "Morton-sorted, timed from first touch"},
        "device_ms_per_step": dev_ms / args.steps,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_sim (contacts + heat + gravity gather + symplectic Euler)", "kernel_ms": sim_ms},
    }
    if not args.no_cpu:
        P, V, _, Q, M = R.ctx.download()
        out["cpu_baseline"] = cpu_baseline(twins_like(run), P, V, M, Q)
    R.ctx.close()
    print(json.dumps(out))
    return 0


def twins_like(run):
    s = run["sim"]
    return dict(N=s["N"], W=s["W"], G=s["G"], a=s["a"], m=s["m"], cell=s["cell"], dt=s["dt"], k=s["k"], c=s["c"])


if __name__ == "__main__":
    sys.exit(main())
