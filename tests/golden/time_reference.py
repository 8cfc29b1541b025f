"""The reference itself timed on the CPU: its own step() (src/nbody.js:551-558)
and the crater page's step (chicxulub.html:779-789) run headless on
SwiftShader (the CPU OpenGL ES 3.0 of the kaleido package, run, never
inspected) through the golden tooling -- ref_trace.js records the reference's
GL calls, gl_replay.py replays them, and `mark` records around the timed
steps note the wall clock after a glFinish.  Development container only: it
reads the reference from /root/reference.

The per-cell gravity pass (CELL_FS) does not finish on SwiftShader in useful
time; as in the goldens it is replaced, through the reference's setProg hook,
by a copy of an injected force mesh, so the figures leave out its cost and
overstate the reference's rate on the mesh step.  Everything else -- the
eight depth-peeled seat passes, the deposit, coarse, monopole and far-field
passes, the pairwise pass, the step pass -- runs as written.

Workloads (the states the GPU benchmark and tests use):
  twins   262,144 particles at first touch, the GPU benchmark's state (written
          by runs.prepare on the GPU host: cosmic-collisions_amd/js/prep_state.js),
          its 8192-seat loose list, 16 steps from a mesh step
  theia   16,384 particles, the golden impact state with its 600-seat list,
          64 steps from a mesh step
  crater  the crater page at 262,144 grains (constant gravity: no mesh, no
          per-cell pass), 8 settle steps after its build

Usage: python tests/golden/time_reference.py --twins DIR [--out profiles/reference_swiftshader_cpu.md]
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "tests"))
import golden_util as gu  # noqa: E402


def replay(work, tag, ops):
    d = os.path.join(work, tag)
    os.makedirs(d, exist_ok=True)
    json.dump({"ops": ops}, open(os.path.join(d, "job.json"), "w"))
    t0 = time.time()
    subprocess.run(["node", os.path.join(HERE, "ref_trace.js"), os.path.join(d, "job.json"), d], check=True,
                   capture_output=True)
    record_s = time.time() - t0
    r = subprocess.run([sys.executable, os.path.join(HERE, "gl_replay.py"), d], check=True, capture_output=True,
                       text=True, timeout=7200)
    info = json.loads(r.stdout.strip().splitlines()[-1])
    marks = dict(info["marks"])
    return marks["end"] - marks["start"], info, record_s


def steps(n):
    return [{"op": "step", "settle": 0, "kick": 0, "advance": 1} for _ in range(n)]


def cpu_model():
    for ln in open("/proc/cpuinfo"):
        if ln.startswith("model name"):
            return ln.split(":", 1)[1].strip()
    return platform.processor()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--twins", required=True, help="the twins state directory written by runs.prepare")
    ap.add_argument("--work", default=os.path.join(REPO, "_work", "store", "timing"))
    ap.add_argument("--out", default=os.path.join(REPO, "profiles", "reference_swiftshader_cpu.md"))
    ap.add_argument("--twins-steps", type=int, default=16)
    ap.add_argument("--theia-steps", type=int, default=64)
    ap.add_argument("--crater-steps", type=int, default=8)
    args = ap.parse_args()
    rows = []

    # twins: the GPU benchmark's state at first touch
    run = json.load(open(os.path.join(args.twins, "run.json")))
    s = run["sim"]
    fields = {k: s[k] for k in ("t", "boxes", "bulk0", "bulk1", "cen0", "cen1", "impactT", "soft")}
    fields["stepNo"] = 0
    ops = [{"op": "build", "opt": run["opt"], "tag": "tw"},
           {"op": "set_state", "dir": args.twins, "name": "state"},
           {"op": "set_sim", "fields": fields},
           {"op": "set_pp", "list": os.path.join(args.twins, "pplist.i32")},
           {"op": "install_injector"},
           {"op": "mark", "tag": "start"}] + steps(args.twins_steps) + [{"op": "mark", "tag": "end"}]
    wall, info, _ = replay(args.work, "twins", ops)
    n_pp = os.path.getsize(os.path.join(args.twins, "pplist.i32")) // 4
    rows.append(("twins, first touch (the GPU benchmark's state)", s["N"], args.twins_steps, wall, n_pp))
    print(rows[-1], flush=True)

    # theia-like: the golden impact state (16,384) with its loose list
    fx = gu.fixture("c1_impact")
    d = os.path.join(args.work, "theia_state")
    os.makedirs(d, exist_ok=True)
    for nm in ("pos", "vel", "q", "mat"):
        np.ascontiguousarray(fx[nm], np.float32).tofile(os.path.join(d, "s.%s.f32" % nm))
    np.ascontiguousarray(fx["pp_index"], np.int32).tofile(os.path.join(d, "pp.i32"))
    f = json.loads(str(fx["fields"]))
    f.update({"stepNo": 0, "soft": False, "impactT": 0.0})
    ops = [{"op": "build", "opt": json.loads(str(gu.fixture("c1_build")["opt"])), "tag": "c1"},
           {"op": "set_state", "dir": d, "name": "s"}, {"op": "set_sim", "fields": f},
           {"op": "set_pp", "list": os.path.join(d, "pp.i32")}, {"op": "install_injector"},
           {"op": "mark", "tag": "start"}] + steps(args.theia_steps) + [{"op": "mark", "tag": "end"}]
    wall, info, _ = replay(args.work, "theia", ops)
    rows.append(("theia-like impact state (C1 size)", int(fx["pos"].shape[0]), args.theia_steps, wall,
                 len(fx["pp_index"])))
    print(rows[-1], flush=True)

    # the crater page at 262,144 grains: settle steps after its build
    copt = {"nIdx": 4, "imp": 14, "vimp": 20, "angle": 60, "sed": 3, "dens": 2500, "rest": 0.25}
    ops = [{"op": "crater_build", "opt": copt, "tag": "crater"}, {"op": "mark", "tag": "start"}]
    ops += [{"op": "crater_step", "settle": 1} for _ in range(args.crater_steps)] + [{"op": "mark", "tag": "end"}]
    wall, info, _ = replay(args.work, "crater", ops)
    rows.append(("crater page, constant g, settle steps", 262144, args.crater_steps, wall, 0))
    print(rows[-1], flush=True)

    cores = os.cpu_count()
    lines = ["# The reference on the CPU: SwiftShader in the development container", "",
             "Made by `python tests/golden/time_reference.py --twins <the twins state from runs.prepare>` (reads the "
             "reference from `/root/reference`; the twins state is the one `runs.prepare(runs.TWINS, ...)` writes on "
             "the GPU host, copied back).", "",
             "- CPU: %s, %d logical CPUs (SwiftShader's threads), %s" % (cpu_model(), cores, info["renderer"]),
             "- GL: %s" % info["version"],
             "- The per-cell gravity pass (CELL_FS) is a copy of an injected force mesh (SwiftShader cannot run it in "
             "useful time): the rates below leave out its cost.", "",
             "| Workload | N | loose list | steps | wall s | particle-steps/s |", "|---|---|---|---|---|---|"]
    for name, N, k, wall, npp in rows:
        lines.append("| %s | %d | %d | %d | %.2f | %.4g |" % (name, N, npp, k, wall, N * k / wall))
    lines += ["", "Machine-readable (read by bench.py):", ""]
    for name, N, k, wall, npp in rows:
        key = name.split(",")[0].split(" ")[0]
        lines.append("cpu_reference %s: %.6g particle-steps/s, N %d, steps %d, wall %.3f s, cores %d" %
                     (key, N * k / wall, N, k, wall, cores))
    open(args.out, "w").write("\n".join(lines) + "\n")
    print("\n".join(lines))
    return 0


if __name__ == "__main__":
    sys.exit(main())
