"""Prepared runs over the C ABI: a configuration taken through the
JavaScript engine's own phases (build, settle, approach on rails, the kick
at first touch, a look with the Morton sort: js/prep_state.js) and then
stepped in the full phase as the engine steps it (src/nbody.js:538-558):
the boxes of every mesh step, the softening an hour and a half after first
touch, sim.t advanced per step.

Named configurations (build() options, src/nbody.js:114-250):
  THEIA    the plain theia-like impact, 16,384 particles (0.89 + 0.11 M_earth, 45 deg, 1.1 v_esc)
  TWINS    the plain twins, 262,144 particles (two 0.5 M_earth, 30 deg, 1.05 v_esc)
  SHATTER  262,144 particles, a 0.5 M_earth impactor head-on at 2.6 v_esc
"""
import json
import math
import os
import subprocess
import sys

import numpy as np

import ccamd

HERE = os.path.dirname(os.path.abspath(__file__))
FL = 6.0

THEIA = {"nIdx": 0, "q": 0.11 / 0.89, "dens": 0.75, "core": 0.325, "spin0": 0.18, "spin1": 0.20, "angle": 45,
         "incl": 50, "vfac": 1.10, "rest": 0.3}
TWINS = {"nIdx": 4, "q": 1.0, "dens": 0.75, "core": 0.325, "spin0": 0.18, "spin1": 0.20, "angle": 30, "incl": 0,
         "vfac": 1.05, "rest": 0.3}
SHATTER = {"nIdx": 4, "q": 0.5, "dens": 0.75, "core": 0.325, "spin0": 0.18, "spin1": 0.20, "angle": 0, "incl": 0,
           "vfac": 2.6, "rest": 0.3}


def prepare(opt, out_dir, extra_steps=0, reuse=False):
    """The configuration through the engine's phases on this GPU, written to
    out_dir (or read from there when reuse is set and a run of these options
    is there): (run description, state arrays)."""
    meta = os.path.join(out_dir, "run.json")
    have = os.path.exists(meta) and json.load(open(meta)).get("opt") == opt
    if not (reuse and have):
        cmd = ["node", os.path.join(HERE, "js", "prep_state.js"), json.dumps(opt), out_dir, str(extra_steps)]
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=900)
        sys.stderr.write(r.stdout[-3000:])
        if r.returncode != 0:
            raise RuntimeError("state preparation failed (exit %d):\n%s" % (r.returncode, r.stderr[-4000:]))
    return load(out_dir)


def load(d):
    run = json.load(open(os.path.join(d, "run.json")))
    N = run["sim"]["N"]
    st = {k: np.fromfile(os.path.join(d, "state.%s.f32" % k), np.float32) for k in ("pos", "vel", "q", "mat")}
    st["pos"] = st["pos"].reshape(N, 4)
    st["vel"] = st["vel"].reshape(N, 4)
    st["pptab"] = np.fromfile(os.path.join(d, "pptab.f32"), np.float32).reshape(64, 2)
    st["pplist"] = np.fromfile(os.path.join(d, "pplist.i32"), np.int32)
    return run, st


BATCH_SEED = 20261203


def prepare_batch(out_dir, count, seed=BATCH_SEED, reuse=False):
    """The batched sweep's instances (js/prep_batch.js: seeded uneven sizes and
    parameters, each through the engine's own build, settle, approach, kick
    and look), written to out_dir or read from there: [(run, state)]."""
    meta = os.path.join(out_dir, "batch.json")
    have = os.path.exists(meta) and json.load(open(meta)).get("count") == count
    if not (reuse and have):
        cmd = ["node", os.path.join(HERE, "js", "prep_batch.js"), out_dir, str(count), str(seed)]
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=1500)
        sys.stderr.write(r.stdout[-2000:])
        if r.returncode != 0:
            raise RuntimeError("batch preparation failed (exit %d):\n%s" % (r.returncode, r.stderr[-4000:]))
    return [load(os.path.join(out_dir, inst["dir"])) for inst in json.load(open(meta))["instances"]]


class Runner:
    """One instance in its full phase over the C ABI."""

    def __init__(self, run, st):
        s = run["sim"]
        self.s = dict(s)
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

    def steps_until_soften(self):
        """Steps that can run before the softening is due (a large number once softened)."""
        s = self.s
        if s["soft"]:
            return 1 << 30
        return max(1, int(math.floor((1.5 - (self.t - s["impactT"])) / s["dt"])) + 1)

    def plan(self, n):
        """The boxes of the mesh steps among the next n steps ((k, 7) array)
        and the time after them."""
        s, t, L = self.s, self.t, []
        for i in range(n):
            if (self.step_no + i) % 8 == 0:
                for b in s["boxes"]:
                    dtb = t - b["t"]
                    L += [FL - b["cen"][k] - b["vel"][k] * dtb for k in range(3)]
                L.append(1.0 if s["boxes"][1]["on"] else 0.0)
            t += s["dt"]
        return np.array(L, np.float64).reshape(-1, 7), t

    def steps(self, n):
        """n steps; returns the device milliseconds (HIP events around the launches)."""
        ms, done = 0.0, 0
        while done < n:
            self.soften_if_due()
            chunk = min(n - done, self.steps_until_soften())
            L, t = self.plan(chunk)
            self.ctx.step_no = self.step_no
            ms += self.ctx.step_n(chunk, self.args, L if len(L) else None)
            self.t, self.step_no, done = t, self.step_no + chunk, done + chunk
        return ms

    def close(self):
        self.ctx.close()


class BatchRunner:
    """Instances stepped as one batch: every instance's boxes and softening
    as its own Runner would have them, the batch advanced in chunks that end
    where some instance softens."""

    def __init__(self, runs):
        self.rs = [Runner(run, st) for run, st in runs]
        self.batch = ccamd.Batch([r.ctx for r in self.rs])
        self.N = sum(r.s["N"] for r in self.rs)

    def steps(self, n):
        ms, done = 0.0, 0
        while done < n:
            for r in self.rs:
                r.soften_if_due()
            chunk = min([n - done] + [r.steps_until_soften() for r in self.rs])
            boxL, args = [], []
            for r in self.rs:
                L, t = r.plan(chunk)
                r.ctx.step_no = r.step_no
                boxL.append(L)
                args.append(r.args)
                r.t, r.step_no = t, r.step_no + chunk
            ms += self.batch.step_n(chunk, args, boxL)
            done += chunk
        return ms

    def close(self):
        self.batch.close()
        for r in self.rs:
            r.close()
