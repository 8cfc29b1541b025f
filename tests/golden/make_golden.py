"""Writes the golden fixtures under tests/golden/fixtures/ (development
container only: it reads the reference from REF_DIR, default /root/reference/src,
and needs node and SwiftShader).

For each case a seeded input state is built from the reference's own
``build()`` output (C1: 16,384 particles) and run through the reference's own
passes, recorded by ref_trace.js and replayed on SwiftShader by gl_replay.py:
contact seats, deposit, coarse, monopole, far field, pairwise correction,
the step pass in several variants, the per-particle read-out (DIAG), the
rigid approach pass.  The per-cell gravity pass cannot run on SwiftShader in
useful time; the force mesh the step reads is the oracle's (cco_cell applied
to SwiftShader's own fine/coarse/far meshes), injected through the
reference's setProg hook.  The reference's host analysis ``fof`` (books,
loose list, Morton permutation) is run by ref_fof.js on each case.

Fixture contents (npz): inputs, the uniform values each pass received (read
from the trace: the reference's own host arithmetic), and outputs.
Usage: python tests/golden/make_golden.py [--out DIR]
"""
import argparse
import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "oracle"))
import oracle  # noqa: E402

FL = 6.0
FH = 2 * FL / 64

C1_OPT = {"nIdx": 0, "q": 0.11 / 0.89, "dens": 0.75, "core": 0.325, "spin0": 0.18, "spin1": 0.20,
          "angle": 45, "incl": 50, "vfac": 1.10, "rest": 0.3}


def run(cmd, **kw):
    r = subprocess.run(cmd, capture_output=True, text=True, **kw)
    if r.returncode != 0:
        raise RuntimeError("%s failed:\n%s\n%s" % (cmd, r.stdout[-2000:], r.stderr[-4000:]))
    return r.stdout


def trace_and_replay(work, ops, tag):
    d = os.path.join(work, tag)
    os.makedirs(d, exist_ok=True)
    job = os.path.join(d, "job.json")
    json.dump({"ops": ops}, open(job, "w"))
    run(["node", os.path.join(HERE, "ref_trace.js"), job, d])
    run([sys.executable, os.path.join(HERE, "gl_replay.py"), d], timeout=1200)
    return d


def kind_of(fs):
    if "uInject" in fs:
        return "cell"
    for key, k in (("uVimp", "couple"), ("flat in float vSeat", "slot"), ("in vec4 vW", "dep"), ("oT1", "far"), ("uKick", "sim"),
                   ("uRot0", "rigid"), ("uPerm", "perm"), ("seated", "diag"), ("oBox", "ppgather"),
                   ("uTab", "ppforce"), ("c * 4 + ivec3(x, y, z)", "coarse"),
                   ("for (int y = 0; y < 64; y++)", "mono")):
        if key in fs:
            return k
    return "other"


def uniform_snapshots(d):
    """Uniform state of each program at each draw, read from the trace."""
    tr = json.load(open(os.path.join(d, "trace.json")))
    shaders = {int(k): v for k, v in tr["shaders"].items()}
    prog_fs, ustate, cur = {}, {}, None
    draws = []
    for rec in tr["trace"]:
        op, a = rec["op"], rec["args"]
        if op == "attachShader":
            src = shaders[a[1]["h"]]
            if "gl_FragCoord" in src or "out vec4 o" in src or "oPos" in src:
                prog_fs[a[0]["h"]] = src
        elif op == "useProgram":
            cur = a[0]["h"]
        elif op.startswith("uniform") and a[0] is not None:
            val = a[1:] if op not in ("uniform3fv", "uniformMatrix3fv") else (a[1] if op == "uniform3fv" else a[2])
            ustate.setdefault(a[0]["prog"], {})[a[0]["loc"]] = val
        elif op == "drawArrays":
            draws.append({"kind": kind_of(prog_fs.get(cur, "")), "u": json.loads(json.dumps(ustate.get(cur, {})))})
    return draws


def load(d, tag, name):
    return np.load(os.path.join(d, "%s.%s.npy" % (tag, name)))


def f32file(path, arr):
    np.ascontiguousarray(arr, np.float32).tofile(path)
    return path


def make_state(base_pos, base_vel, mat, sim, kind, rng):
    """A seeded input state from the built bodies."""
    pos = base_pos.copy().reshape(-1, 4)
    vel = base_vel.copy().reshape(-1, 4)
    N = pos.shape[0]
    imp = pos[:, 3] >= 3.5
    c0 = np.array(sim["cen0"])
    c1 = np.array(sim["cen1"])
    u = (c1 - c0) / np.linalg.norm(c1 - c0)
    if kind == "impact":
        # impactor moved into contact: overlapping the target by a tenth of a radius
        shift = c0 + u * (sim["R1"] + sim["R2"] - 0.1) - c1
        pos[imp, :3] += shift
        vel[:, :3] = rng.normal(0, 0.15, (N, 3))
        vel[imp, :3] += -u * 1.2
        strays = rng.choice(N, 40, replace=False)
        dirs = rng.normal(size=(40, 3))
        dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
        pos[strays, :3] = c0 + dirs * rng.uniform(6.5, 9.0, (40, 1))
    else:   # twobox: the impactor far off with its own box, debris between
        shift = c0 + u * 7.5 - c1
        pos[imp, :3] += shift
        vel[:, :3] = rng.normal(0, 0.05, (N, 3))
        deb = rng.choice(np.flatnonzero(~imp), 300, replace=False)
        pos[deb, :3] = c0 + u * rng.uniform(2.0, 5.5, (300, 1)) + rng.normal(0, 0.25, (300, 3))
    vel[:, 3] = rng.uniform(0, 6000, N).astype(np.float32)
    q = rng.uniform(0, 2e-5, N).astype(np.float32)
    return pos.astype(np.float32), vel.astype(np.float32), q


CRATER_OPT = {"nIdx": 0, "imp": 14, "vimp": 20, "angle": 60, "sed": 3, "dens": 2500, "rest": 0.25}


def make_crater(args):
    """The crater page (C3 preset at 16,384 particles): its build, two settle
    steps, the approach to touchdown, the coupling pass and two full steps,
    each state dumped so that every pass compares from identical input."""
    st = ["pos", "vel", "aux", "q"]
    ops = [{"op": "crater_build", "opt": CRATER_OPT, "tag": "crater"},
           {"op": "dump", "what": ["pos", "vel"], "tag": "b0"},
           {"op": "crater_step", "settle": 1}, {"op": "dump", "what": st + ["seatA", "seatB"], "tag": "s1"},
           {"op": "crater_step", "settle": 1}, {"op": "dump", "what": st, "tag": "s2"},
           {"op": "crater_approach", "what": st, "tagBefore": "pre"},
           {"op": "dump", "what": st, "tag": "post"},
           {"op": "crater_step", "settle": 0}, {"op": "dump", "what": st + ["seatA", "seatB"], "tag": "f1"},
           {"op": "crater_step", "settle": 0}, {"op": "dump", "what": st, "tag": "f2"}]
    d = trace_and_replay(args.work, ops, "crater")
    host = json.load(open(os.path.join(d, "host.json")))
    fx = {"mat": np.fromfile(os.path.join(d, "crater.mat.f32"), np.float32), "opt": json.dumps(CRATER_OPT),
          "host": json.dumps(host["ops"]), "draws": json.dumps(uniform_snapshots(d))}
    for tag, names in (("b0", ["pos", "vel"]), ("s1", st), ("s2", st), ("pre", st), ("post", st), ("f1", st),
                       ("f2", st)):
        for nm in names:
            fx["%s_%s" % (tag, nm)] = load(d, tag, nm)
    for tag in ("s1", "f1"):
        for nm in ("seatA", "seatB"):
            u32 = load(d, tag, nm).view(np.uint32).reshape(-1, 4)
            occ = np.flatnonzero((u32 != 0xBF800000).any(1))
            fx["%s_%s_idx" % (tag, nm)] = occ.astype(np.int32)
            fx["%s_%s_val" % (tag, nm)] = u32[occ]
    np.savez_compressed(os.path.join(args.out, "c3_crater.npz"), **fx)
    print("crater fixture: N=%d, touchdown %s" % (len(fx["mat"]), [o for o in host["ops"] if o["op"] == "crater_touchdown"]))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "fixtures"))
    ap.add_argument("--work", default=os.path.join(REPO, "_work", "golden"))
    ap.add_argument("--cases", default="c1,crater", help="comma list of: c1 (build, impact, twobox), crater")
    args = ap.parse_args()
    os.makedirs(args.out, exist_ok=True)
    os.makedirs(args.work, exist_ok=True)
    cases = args.cases.split(",")
    if "crater" in cases:
        make_crater(args)
    if "c1" not in cases:
        return 0

    # 1. the reference's body setup for C1, and its pair table
    d0 = trace_and_replay(args.work, [{"op": "build", "opt": C1_OPT, "tag": "c1"},
                                      {"op": "dump", "what": ["pos", "vel"], "tag": "b"}], "build")
    host = json.load(open(os.path.join(d0, "host.json")))
    sim = host["ops"][0]["sim"]
    pos0 = load(d0, "b", "pos").reshape(-1, 4)
    vel0 = load(d0, "b", "vel").reshape(-1, 4)
    mat = np.fromfile(os.path.join(d0, "c1.mat.f32"), np.float32)
    home = np.fromfile(os.path.join(d0, "c1.home.f32"), np.float32)
    blobs = np.fromfile(os.path.join(d0, "blobs.bin"), np.uint8)
    tb = host["ppTab"]
    pptab = blobs[tb["blob"]:tb["blob"] + tb["bytes"]].view(np.float32).reshape(64, 2).copy()
    np.savez_compressed(os.path.join(args.out, "c1_build.npz"), pos=pos0, vel=vel0, mat=mat, home=home,
                        pptab=pptab, sim=json.dumps(sim), opt=json.dumps(C1_OPT), consts=json.dumps(host["consts"]))
    print("build fixture: N=%d n1=%d" % (sim["N"], sim["n1"]))

    rng = np.random.default_rng(20261203)
    for kind in ("impact", "twobox"):
        d = os.path.join(args.work, kind)
        os.makedirs(d, exist_ok=True)
        pos, vel, q = make_state(pos0, vel0, mat, sim, kind, rng)
        for nm, arr in (("pos", pos), ("vel", vel), ("q", q), ("mat", mat)):
            f32file(os.path.join(d, "s.%s.f32" % nm), arr)
        c0 = list(map(float, sim["cen0"]))
        imp = pos[:, 3] >= 3.5
        impc = pos[imp, :3].mean(0).astype(float).tolist()
        boxes = [{"on": True, "cen": c0, "vel": [0.01, -0.02, 0.005], "t": -0.5},
                 {"on": kind == "twobox", "cen": impc, "vel": [0.0, 0.01, 0.0], "t": -0.25}]
        fields = {"t": 0.0, "boxes": boxes, "bulk0": [0.05, 0.0, -0.01], "bulk1": [-0.4, 0.02, 0.1],
                  "cen0": c0, "cen1": impc}
        # the loose list: the particles farthest from the target's centre
        r2 = ((pos[:, :3] - np.array(c0)) ** 2).sum(1)
        loose = np.argsort(-r2, kind="stable")[:600].astype(np.int32)
        loose.sort()
        loose.tofile(os.path.join(d, "pp.i32"))
        m1 = float(mat[imp].sum())
        mono1 = [impc[0] * m1, impc[1] * m1, impc[2] * m1, m1]
        base = [{"op": "build", "opt": C1_OPT, "tag": "c1"},
                {"op": "set_state", "dir": d, "name": "s"},
                {"op": "set_sim", "fields": fields}]
        if kind == "twobox":
            base.append({"op": "set_mono", "box": 1, "value": mono1})
        mesh_names = ["fine0", "coarse0", "mono0", "farA0", "farT00", "farT10"]
        if kind == "twobox":
            mesh_names += ["fine1", "coarse1", "mono1", "farA1", "farT01", "farT11"]
        # phase 1: seats, mesh, pairwise (CELL injected as zeros: its output unused here)
        p1 = trace_and_replay(args.work, base + [
            {"op": "install_injector"},
            {"op": "set_pp", "list": os.path.join(d, "pp.i32")},
            {"op": "diag"},
            {"op": "dump", "what": ["seatA", "seatB"] + mesh_names + ["ppList", "ppBox", "ppForce"], "tag": "p1"}],
            kind + "_p1")
        out = {}
        for nm in ["seatA", "seatB"] + mesh_names + ["ppList", "ppBox", "ppForce"]:
            out[nm] = load(p1, "p1", nm)
        draws = uniform_snapshots(p1)
        ucell = [x["u"] for x in draws if x["kind"] == "dep"]
        # the oracle's per-cell gravity from SwiftShader's own meshes (eps2 = 0.5 FH^2, G m = sim.m)
        eps2c, gm = np.float32(0.5 * FH * FH), np.float32(sim["m"])
        force = {}
        for b in ([0, 1] if kind == "twobox" else [0]):
            force[b] = oracle.cell(out["fine%d" % b], out["coarse%d" % b], out["farA%d" % b], out["farT0%d" % b],
                                   out["farT1%d" % b], eps2c, gm)
            f32file(os.path.join(d, "force%d.f32" % b), force[b])
        # phase 2: DIAG and the step pass on the injected force meshes (each box its own)
        inj = {"op": "install_injector", "file": os.path.join(d, "force0.f32")}
        if kind == "twobox":
            inj["file1"] = os.path.join(d, "force1.f32")
        ops = base + [inj,
                      {"op": "set_pp", "list": os.path.join(d, "pp.i32")},
                      {"op": "diag"},
                      {"op": "dump", "what": ["diag", "force0", "mono0"] + (["force1"] if kind == "twobox" else []),
                       "tag": "p2"}]
        variants = [("full", 0, 0, 8 * sim["dt"]), ("settle", 1, 0, 8 * sim["dt"]), ("kick", 0, 1, 0.0)]
        for vn, st, kk, gdt in variants:
            ops += [{"op": "set_state", "dir": d, "name": "s"},
                    {"op": "pass_sim", "settle": st, "kick": kk, "gdt": gdt},
                    {"op": "dump", "what": ["pos", "vel", "aux", "q"], "tag": "sim_" + vn}]
        # the rigid approach pass from the same state
        ops += [{"op": "set_state", "dir": d, "name": "s"},
                {"op": "set_sim", "fields": {"rel": {"r": [-6.0, 1.2, 1.0], "v": [0.8, 0.0, 0.0]}}},
                {"op": "rigid", "dtF": 0.01},
                {"op": "dump", "what": ["pos", "vel", "aux", "q"], "tag": "rigid"}]
        # the reference's own step() from the same state, eight of them from a mesh step (src/nbody.js:551-558):
        # seats every second step, the mesh, the pairwise pass and the 8 dt impulse on the first, the
        # softening (src/nbody.js:545-550) due on the fifth; sim.t advanced by the caller after each step
        steps = []
        if kind == "impact":
            soft_at = 4
            sfields = dict(fields, stepNo=0, soft=False, impactT=fields["t"] - 1.5 + (soft_at - 0.5) * sim["dt"])
            ops += [{"op": "set_state", "dir": d, "name": "s"},
                    {"op": "set_sim", "fields": sfields},
                    {"op": "set_pp", "list": os.path.join(d, "pp.i32")}]
            for n in range(8):
                ops.append({"op": "step", "settle": 0, "kick": 0, "advance": 1})
                if n in (0, 1, soft_at, 7):
                    ops.append({"op": "dump", "what": ["pos", "vel", "aux", "q"], "tag": "step%d" % n})
                    steps.append(n)
            ops.append({"op": "record_sim", "tag": "steps", "fields": ["t", "stepNo", "dt", "k", "c", "damp", "soft"]})
        p2 = trace_and_replay(args.work, ops, kind + "_p2")
        draws2 = uniform_snapshots(p2)
        out["diag"] = load(p2, "p2", "diag")
        assert np.array_equal(load(p2, "p2", "force0"), force[0].reshape(513, 512, 4)), "injection not exact"
        if kind == "twobox":
            assert np.array_equal(load(p2, "p2", "force1"), force[1].reshape(513, 512, 4)), "box B injection not exact"
        if steps:
            for n in steps:
                for nm in ("pos", "vel", "aux", "q"):
                    out["step%d_%s" % (n, nm)] = load(p2, "step%d" % n, nm)
            rec = [o for o in json.load(open(os.path.join(p2, "host.json")))["ops"] if o["op"] == "record_sim"][-1]
            out["steps_meta"] = json.dumps({"fields": sfields, "after": rec["sim"], "dumped": steps})
        for vn, *_ in variants:
            for nm in ("pos", "vel", "aux", "q"):
                out["sim_%s_%s" % (vn, nm)] = load(p2, "sim_" + vn, nm)
        for nm in ("pos", "vel", "aux", "q"):
            out["rigid_" + nm] = load(p2, "rigid", nm)
        # the reference's analysis on this state (positions, velocities, DIAG) with the Morton sort
        job = {"N": sim["N"], "m": sim["m"], "a": sim["a"], "k": sim["k"], "touch": 2 * sim["a"], "cap": 8192,
               "sort": True, "P": f32file(os.path.join(d, "P.f32"), pos), "V": f32file(os.path.join(d, "V.f32"), vel),
               "R": f32file(os.path.join(d, "R.f32"), mat), "D": f32file(os.path.join(d, "D.f32"), out["diag"]),
               "out": os.path.join(d, "fof")}
        json.dump(job, open(os.path.join(d, "fofjob.json"), "w"))
        run(["node", os.path.join(HERE, "ref_fof.js"), os.path.join(d, "fofjob.json")])
        report = json.load(open(os.path.join(d, "fof.report.json")))
        perm = np.fromfile(os.path.join(d, "fof.perm.f32"), np.float32)
        pplist = np.fromfile(os.path.join(d, "fof.pplist.f32"), np.float32)

        # keep fixtures small: seat atlases and meshes as (occupied texel index, value)
        fx = {"pos": pos, "vel": vel, "q": q, "mat": mat, "pp_index": loose, "pptab": pptab,
              "sim": json.dumps(sim), "fields": json.dumps(fields), "draws": json.dumps(draws + draws2),
              "report": json.dumps(report), "perm": perm, "pplist_fof": pplist}
        for nm in ("seatA", "seatB"):
            u32 = out[nm].view(np.uint32).reshape(-1, 4)
            occ = np.flatnonzero((u32 != 0xBF800000).any(1))
            fx[nm + "_idx"] = occ.astype(np.int32)
            fx[nm + "_val"] = u32[occ]
        for nm in mesh_names:
            arr = out[nm].reshape(-1, 4)
            occ = np.flatnonzero((arr != 0).any(1))
            fx[nm + "_idx"] = occ.astype(np.int32)
            fx[nm + "_val"] = arr[occ]
        for b in force:
            arr = force[b].reshape(-1, 4)
            occ = np.flatnonzero((arr != 0).any(1))
            fx["force%d_idx" % b] = occ.astype(np.int32)
            fx["force%d_val" % b] = arr[occ]
        for nm in ("ppList", "ppBox", "ppForce", "diag"):
            fx[nm] = out[nm]
        for k, v in out.items():
            if k.startswith("sim_") or k.startswith("rigid_") or k.startswith("step"):
                fx[k] = v
        np.savez_compressed(os.path.join(args.out, "c1_%s.npz" % kind), **fx)
        print("case %s: seats %d cells, fine %d texels" % (kind, len(fx["seatA_idx"]), len(fx["fine0_idx"])))
    return 0


if __name__ == "__main__":
    sys.exit(main())
