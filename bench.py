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
import tempfile
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(REPO, "cosmic-collisions_amd")
sys.path.insert(0, PKG)
import ccamd  # noqa: E402
import runs  # noqa: E402  (the library's prepared runs; loads the HIP library lazily)

FL = 6.0
FH = 2 * FL / 64
HBM_PEAK_GBS = 8000.0


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


def reference_cpu():
    """The reference's own step() timed on SwiftShader in the development
    container (profiles/reference_swiftshader_cpu.md, written by
    tests/golden/time_reference.py): its twins figure, the workload of this
    benchmark, or None."""
    path = os.path.join(REPO, "profiles", "reference_swiftshader_cpu.md")
    if not os.path.exists(path):
        return None
    rows = {}
    for line in open(path):
        if line.startswith("cpu_reference "):
            name, rest = line[len("cpu_reference "):].split(":", 1)
            f = [x.strip() for x in rest.split(",")]
            rows[name] = {"value": float(f[0].split()[0]), "N": int(f[1].split()[1]), "steps": int(f[2].split()[1]),
                          "wall_s": float(f[3].split()[1]), "cores": int(f[4].split()[1])}
    tw = rows.get("twins")
    if tw is None:
        return None
    return {"value": tw["value"], "unit": "particle-steps/s", "cores": tw["cores"], "kind": "reference-swiftshader",
            "sample": "%d steps of the reference's own step() (src/nbody.js:551-558, its GLSL passes) on this "
                      "benchmark's twins state (262,144 particles, 8192-seat loose list) under SwiftShader 4.1.0.7 on "
                      "the development container's %d logical CPUs (Intel Xeon @ 2.70GHz), %.1f s; the per-cell pass "
                      "CELL_FS, which SwiftShader cannot run, replaced by a copy of an injected force mesh (its cost "
                      "left out); the reference cannot run on the GPU host" % (tw["steps"], tw["cores"], tw["wall_s"]),
            "source": "profiles/reference_swiftshader_cpu.md", "other": {k: v for k, v in rows.items() if k != "twins"}}


def measured_traffic():
    """k_sim's HBM bytes per launch as the committed counter profile states it
    (FETCH_SIZE + WRITE_SIZE scaled by the same counters on a plain copy of
    known size), or None."""
    path = os.path.join(REPO, "profiles", "twins262k_plain_step.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("k_sim HBM traffic per launch:"):
            return float(line.split(":")[1].split()[0])
    return None


def batch_leg(out_dir, count, steps, warmup, reuse):
    B = runs.BatchRunner(runs.prepare_batch(out_dir, count, reuse=reuse))
    B.steps(warmup)
    t0 = time.perf_counter()
    dev_ms = B.steps(steps)
    wall = time.perf_counter() - t0
    sizes = [r.s["N"] for r in B.rs]
    out = {"value": B.N * steps / wall, "unit": "particle-steps/s", "instances": len(sizes), "particles": B.N,
           "n_min": min(sizes), "n_max": max(sizes), "steps": steps, "warmup": warmup,
           "ms_per_step": wall * 1e3 / steps, "device_ms_per_step": dev_ms / steps,
           "workload": "batched what-if sweep: %d giant impacts, 16k-262k particles (log-uniform, multiples of 512), "
                       "mass ratio 0.02-1, angle 0-75 deg, 1.0-2.6 v_esc, core 0.2-0.4, seeded per instance; "
                       "one graph launch per 8-step cycle for the whole batch" % len(sizes)}
    B.close()
    return out


def _books(ctx):
    """The FP64 books after a DIAG pass: angular momentum and energy (KE + PE + heat)."""
    ctx.diag()
    b = ctx.books()
    return np.array([b["Jx"], b["Jy"], b["Jz"]]), b["KE"] + b["PE"] + b["Q"]


def learned_leg(state_root, batch_dir, steps, warmup, reuse, cycles, sweep, sweep_cycles, iters, batch_steps):
    """The learned far field, trained on the GPU from the exact far field of the
    theia-like and twins runs and a sweep of the generator's next seed, on what it
    was not trained on: the shatter impact, judged every 16 steps with fallback
    above a 1e-3 relative force error, beside the same run with the exact far
    field and with a zero network (each mesh step between judges keeps the base's
    far field: what the network's prediction adds to); and the batched sweep
    (C4) with the network on every instance, one inference per mesh step for
    the whole batch, beside the same batch with the exact far field."""
    import ff_train
    ff, rep = ff_train.build(["THEIA", "TWINS"], state_root, cycles, iters, reuse=reuse, sweep=sweep,
                             sweep_cycles=sweep_cycles, log=lambda m: print(m, file=sys.stderr, flush=True))
    zero = ccamd.FarField(capacity=0)
    zero.set_weights(np.zeros(zero.nw, np.float32))
    run, st = runs.prepare(runs.SHATTER, os.path.join(state_root, "cc_shatter_state"), reuse=True)
    N = run["sim"]["N"]
    out = {"workload": "shatter: 262,144 particles, 0.5 M_earth impactor head-on at 2.6 v_esc, %d steps from first "
                       "touch; between exact judges every 16 steps the far field's change since the last judge from "
                       "a 4-level 32-channel 3D conv operator (BF16 on the matrix cores) trained here on the GPU on "
                       "other impacts; fallback above 1e-3 relative force error" % (warmup + steps),
           "particles": N, "steps": steps, "training": rep}
    for mode, net in (("exact", None), ("learned", ff), ("zero_network", zero)):
        R = runs.Runner(run, st)
        J0, E0 = _books(R.ctx)
        if net is not None:
            R.ctx.set_far_field(net, 16, 1e-3)
        R.steps(warmup)
        R.ctx.sync()
        t0 = time.perf_counter()
        R.steps(steps)
        R.ctx.sync()
        wall = time.perf_counter() - t0
        J, E = _books(R.ctx)
        o = {"value": N * steps / wall, "ms_per_step": wall * 1e3 / steps,
             "dJ_rel": float(np.linalg.norm(J - J0) / np.linalg.norm(J0)), "E_end_rel": E / abs(E0)}
        if net is not None:
            st_ = dict(zip(ccamd.Context.FF_STATS, R.ctx.ff_stats()[0].tolist()))
            o.update({"judged": int(st_["judged"]), "fallbacks": int(st_["fallbacks"]),
                      "learned_mesh_steps": int(st_["learned"]), "last_force_err": st_["last_err"],
                      "max_force_err": st_["max_err"],
                      "mean_force_err": st_["sum_err"] / max(st_["usable"], 1.0)})
        out[mode] = o
        R.close()
    # the batched form: the C4 sweep (none of its instances in the training data), 64 instances stepped as one
    # batch with the exact far field and with the network attached to every instance
    insts = runs.prepare_batch(batch_dir, 64, reuse=True) if batch_steps > 0 else []
    bo = {"workload": "C4: the batched sweep's 64 giant impacts (none in the training data), one batch, %d steps "
                      "after 24; one inference over every instance's boxes per mesh step, one judge of the judged "
                      "ones after each 8-step cycle" % batch_steps}
    for mode in ("exact", "learned") if insts else ():
        B = runs.BatchRunner(insts)
        if mode == "learned":
            for r in B.rs:
                r.ctx.set_far_field(ff, 16, 1e-3)
        B.steps(24)
        t0 = time.perf_counter()
        B.steps(batch_steps)
        wall = time.perf_counter() - t0
        o = {"value": B.N * batch_steps / wall, "ms_per_step": wall * 1e3 / batch_steps}
        if mode == "learned":
            st_ = np.array([r.ctx.ff_stats() for r in B.rs])   # instance, box, FF_STATS
            tot = st_.sum(axis=(0, 1))
            o.update({"judged": int(tot[1]), "fallbacks": int(tot[2]), "learned_mesh_steps": int(tot[3]),
                      "mean_force_err": float(tot[5] / max(tot[6], 1.0)), "max_force_err": float(st_[:, :, 4].max()),
                      "instances_above_1e-3": int((st_[:, :, 4] > 1e-3).any(axis=1).sum())})
        bo[mode] = o
        bo["particles"] = B.N
        B.close()
    if insts:
        out["batched"] = bo
    ff.close()
    zero.close()
    for mode in ("learned", "zero_network"):   # energy (KE + PE + heat, no contact springs) against the exact run's
        out[mode]["dE_vs_exact"] = out[mode]["E_end_rel"] - out["exact"]["E_end_rel"]
    # the learned far field's throughput only where the network stood in on some mesh step; otherwise every
    # mesh step ran the exact far field (and the judge beside it), which is no learned figure
    stood_in = out["learned"]["learned_mesh_steps"] > 0
    out["value"] = out["learned"]["value"] if stood_in else None
    if not stood_in:
        out["learned"]["label"] = "exact far field with the judge: the network fell back on every judged mesh step"
    out["unit"] = "particle-steps/s"
    return out


def runto_leg(steps, every=360):
    """The twins run as configured, through the JavaScript engine as the
    reference's page drives it: from first touch, full steps with an analysis
    (groups, books, boxes, loose list, a Morton sort every eighth) every 360
    steps -- synchronous as runTo calls it, and overlapped (the host's sums
    running while the next 64 steps are taken, the worker's one-report lag)."""
    out = {"workload": "plain twins through the JavaScript engine (js/nbody.js over the N-API addon), %d steps from "
                       "first touch, an analysis every %d steps as index.html's runTo drives it; window_ms_per_step: "
                       "the step time in each window between looks" % (steps, every)}
    for mode in ("sync", "overlap"):
        cmd = ["node", os.path.join(PKG, "js", "run_timed.js"), json.dumps(runs.TWINS), str(steps), str(every), mode]
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=900)
        if r.returncode != 0:
            raise RuntimeError("run_timed.js failed (exit %d):\n%s" % (r.returncode, r.stderr[-3000:]))
        out[mode] = json.loads(r.stdout.strip().splitlines()[-1])
    return out


def configs_leg(theia_steps, crater_steps):
    """The other plain configurations on the GPU, each as configured: the
    theia-like impact (16,384 particles) through the JavaScript engine from
    first touch with an analysis every 360 steps (js/run_timed.js), and the
    Chicxulub crater patch at 524,288 grains through the crater engine --
    build, settle, approach, coupling, then the full steps (js/crater_timed.js)."""
    out = {}
    if theia_steps > 0:
        cmd = ["node", os.path.join(PKG, "js", "run_timed.js"), json.dumps(runs.THEIA), str(theia_steps), "360", "sync"]
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=900)
        if r.returncode != 0:
            raise RuntimeError("run_timed.js failed (exit %d):\n%s" % (r.returncode, r.stderr[-3000:]))
        t = json.loads(r.stdout.strip().splitlines()[-1])
        out["theia"] = {k: t[k] for k in ("particles", "steps", "every", "looks", "with_analysis", "steps_only",
                                          "ms_per_step", "ms_per_look", "books")}
        out["theia"]["workload"] = "plain theia-like impact: 16,384 particles (0.89 + 0.11 M_earth, 45 deg, 1.1 v_esc) " \
                                   "from first touch, an analysis every 360 steps"
    if crater_steps > 0:
        opt = {"imp": 14, "vimp": 20, "angle": 60, "sed": 3, "dens": 2500, "rest": 0.25, "nIdx": 4, "N": 524288}
        cmd = ["node", os.path.join(PKG, "js", "crater_timed.js"), json.dumps(opt), str(crater_steps)]
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=900)
        if r.returncode != 0:
            raise RuntimeError("crater_timed.js failed (exit %d):\n%s" % (r.returncode, r.stderr[-3000:]))
        out["crater"] = json.loads(r.stdout.strip().splitlines()[-1])
        out["crater"]["workload"] = "Chicxulub crater patch, constant g: 524,288 grains (carbonate over basement, a 14 " \
                                    "km impactor at 20 km/s and 60 deg, the Z-model coupling), full steps after touchdown"
    return out


def build_leg():
    """Body setup on the host: the twins build with its bodies relaxed in
    JavaScript and by the library, and the batched sweep's 64 instances'
    bodies relaxed in one library call (js/time_build.js)."""
    cmd = ["node", os.path.join(PKG, "js", "time_build.js"), json.dumps(runs.TWINS)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        raise RuntimeError("time_build.js failed (exit %d):\n%s" % (r.returncode, r.stderr[-3000:]))
    out = json.loads(r.stdout.strip().splitlines()[-1])
    out["workload"] = "body setup (onion shells laid in JavaScript, the overlap relaxation of src/bodies.js:97-146 in " \
                      "JavaScript or by cc_relax_piles, bit-identical): the twins build, and every body of the batched " \
                      "sweep's 64 instances in one call"
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--state-dir", default=os.path.join(tempfile.gettempdir(), "cc_twins_state"))
    ap.add_argument("--reuse-state", action="store_true", help="use a prepared state in --state-dir if there is one")
    ap.add_argument("--batch", type=int, default=64,
                    help="instances in the batched-sweep leg (64, the sweep's size; 0: skip it)")
    ap.add_argument("--batch-steps", type=int, default=240)
    ap.add_argument("--batch-dir", default=os.path.join(tempfile.gettempdir(), "cc_batch_state"))
    ap.add_argument("--runto", type=int, default=5000, help="steps of the engine-driven leg with analyses (0: skip)")
    ap.add_argument("--learned", type=int, default=1, help="the learned far field leg (0: skip it)")
    ap.add_argument("--build", type=int, default=1, help="the body setup leg on the host (0: skip it)")
    ap.add_argument("--tune", type=int, default=1, help="cc_autotune the plain leg's launch shapes first (0: defaults)")
    ap.add_argument("--theia-steps", type=int, default=1000, help="the theia-like leg's steps (0: skip it)")
    ap.add_argument("--crater-steps", type=int, default=20000, help="the 524,288-grain crater leg's steps (0: skip it)")
    ap.add_argument("--learned-steps", type=int, default=1000)
    ap.add_argument("--ff-cycles", type=int, default=250, help="8-step cycles recorded per training run")
    ap.add_argument("--ff-sweep", type=int, default=64, help="sweep instances recorded for training (0: none)")
    ap.add_argument("--ff-sweep-cycles", type=int, default=40)
    ap.add_argument("--ff-iters", type=int, default=6000)
    ap.add_argument("--ff-batch-steps", type=int, default=240, help="steps of the C4 sweep with the network (0: skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (each box has one MI355X)", file=sys.stderr)
        return 2
    run, st = runs.prepare(runs.TWINS, args.state_dir, reuse=args.reuse_state)
    R = runs.Runner(run, st)
    N = run["sim"]["N"]
    # Warm-up: steps to a mesh step, the launch shapes timed there (cc_autotune, results unchanged), then the
    # rest of the warm-up but at least one whole 8-step cycle, so that the cycle graph the timed steps replay is
    # captured and instantiated before the clock starts (its key includes the tuning).
    tune = None
    pre = (-R.step_no) % 8
    R.steps(pre)
    if args.tune:
        R.ctx.step_no = R.step_no
        tune = R.ctx.autotune(20)
    rest = max(args.warmup - pre, 8)
    R.steps(rest)
    warmup_run = pre + rest
    # the library's stream (the pairwise pass's second stream joins back into it) drained on both sides
    R.ctx.sync()
    t0 = time.perf_counter()
    dev_ms = R.steps(args.steps)
    R.ctx.sync()
    wall = time.perf_counter() - t0
    value = N * args.steps / wall
    # the step pass alone, 200 back to back between HIP events on the library's stream
    sim_ms = R.ctx.time_pass("sim", 200)
    R.ctx.time_pass("copy", 20)   # a copy of known size, the traffic counters' calibration in a profiled run
    alg_bytes = 92.0 * N     # pos+vel+mat+q read, pos+vel+aux+q written: each particle's own data once
    achieved = alg_bytes / (sim_ms * 1e-3) / 1e9
    # the whole step: B(N) = 117 + 12.6e6 / (8 N) compulsory bytes per particle-step with one mesh box
    # (contact grid 22, step pass 92, deposit 2.5, the rest of the mesh and its gather 12.6 MB per mesh step)
    step_bytes = 117.0 + 12.6e6 / (8.0 * N)
    step_gbs = value * step_bytes / 1e9
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
        "warmup_run": warmup_run,
        "autotune": tune,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(),
                     "kernel": "k_sim (contacts + heat + gravity gather + symplectic Euler)", "kernel_ms": sim_ms},
        "roofline_step": {"bound": "hbm", "achieved": step_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": step_gbs / HBM_PEAK_GBS, "bytes_per_particle_step": step_bytes,
                          "what": "the whole step at the metric's value: compulsory bytes per particle-step "
                                  "117 + 12.6e6 / (8 N)"},
    }
    out["cpu_baseline"] = reference_cpu()
    if not args.no_cpu:
        P, V, _, Q, M = R.ctx.download()
        out["cpu_baseline_port"] = cpu_baseline(twins_like(run), P, V, M, Q)
        if out["cpu_baseline"] is None:
            out["cpu_baseline"] = out["cpu_baseline_port"]
    R.close()
    if args.runto > 0:
        out["engine_runto"] = runto_leg(args.runto)
    if args.theia_steps > 0 or args.crater_steps > 0:
        out["configs"] = configs_leg(args.theia_steps, args.crater_steps)
    if args.batch > 0:
        out["batched"] = batch_leg(args.batch_dir, args.batch, args.batch_steps, 24, args.reuse_state)
    if args.build > 0:
        out["body_build"] = build_leg()
    if args.learned > 0:
        out["learned"] = learned_leg(tempfile.gettempdir(), args.batch_dir, args.learned_steps, 200,
                                     True, args.ff_cycles, args.ff_sweep,
                                     args.ff_sweep_cycles, args.ff_iters, args.ff_batch_steps)
    out["hip_runtime"] = hip_runtime()
    print(json.dumps(out))
    return 0


def hip_runtime():
    """The HIP runtime copies this process has mapped (the library's, and no other)."""
    return sorted({ln.split()[-1] for ln in open("/proc/self/maps") if "libamdhip64" in ln})


def twins_like(run):
    s = run["sim"]
    return dict(N=s["N"], W=s["W"], G=s["G"], a=s["a"], m=s["m"], cell=s["cell"], dt=s["dt"], k=s["k"], c=s["c"])


if __name__ == "__main__":
    sys.exit(main())
