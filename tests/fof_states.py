"""States and checks shared by the analysis tests (test_fof_host.py on the
CPU, test_gpu_fof.py on the GPU): states built so that groups tie, the
oracle's whole analysis (oracle/fof.js) in node, and the report comparison."""
import json
import os
import subprocess

import numpy as np

import golden_util as gu
import oracle

ORACLE_DIR = os.path.join(gu.REPO, "oracle")
KEYS = ("largest", "largestImp", "largestIron", "day", "tilt", "axis", "bodyVel", "orbit", "escape", "com", "core",
        "vel", "R", "Redge", "Tsurf", "hotCom", "hotR", "impMass", "impCom", "impRms", "M", "iron", "drift", "L", "KE",
        "PE", "EL", "Q", "unseated", "full", "jD", "moon")


def run_oracle_js(tmp_path, P, V, R, D, m, a, spring, touch, cap=8192):
    """oracle/fof.js on the same state: (report dict, list)."""
    for k, x in (("P", P), ("V", V), ("R", R), ("D", D)):
        np.ascontiguousarray(x, np.float32).tofile(tmp_path / k)
    script = tmp_path / "run.js"
    script.write_text("""
const { analyse } = require('%s/fof.js'); const fs = require('fs');
const rd = (k) => { const b = fs.readFileSync('%s/' + k); return new Float32Array(b.buffer, b.byteOffset, b.byteLength / 4); };
const r = analyse(%d, %r, %r, rd('P'), rd('V'), rd('R'), rd('D'), %r, %r, %d);
fs.writeFileSync('%s/list', Buffer.from(r.pp.list.buffer));
const loose = r.pp.loose; delete r.pp; r.ppLoose = loose; console.log(JSON.stringify(r));
""" % (ORACLE_DIR, tmp_path, len(R), m, a, spring, touch, cap, tmp_path))
    out = subprocess.run(["node", str(script)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    return json.loads(out.stdout), np.fromfile(tmp_path / "list", np.int32)


def same_report(mine, ref):
    for k in KEYS:
        assert mine[k] == ref[k], k
    assert (mine["second"] is None) == (ref["second"] is None)
    if ref["second"] is not None:
        for k in ref["second"]:
            assert mine["second"][k] == ref["second"][k], k


def blob(n_side, spacing, centre):
    g = (np.arange(n_side) - (n_side - 1) / 2) * spacing
    x, y, z = np.meshgrid(g, g, g, indexing="ij")
    p = np.stack([x.ravel() + centre[0], y.ravel() + centre[1], z.ravel() + centre[2]], axis=1)
    return p.astype(np.float32)


def tie_state(kind):
    """The golden impact state with a copy or lattice clumps added so that two
    groups tie, their indices arranged so that the reference's union-find
    roots order the tied groups otherwise than their smallest indices do:
    'largest' -- the state and a shifted copy interleaved, the copy in
    reverse; 'second' -- two 5x5x5 lattice clumps (125 members each, more than
    N/1000) far off, the first clump's first member ahead of the second clump, and six
    lone particles."""
    fx = gu.fixture("c1_impact")
    s = json.loads(str(fx["sim"]))
    P, V, R, D = (np.asarray(fx[k], np.float32).reshape(len(fx["mat"]), -1) for k in ("pos", "vel", "mat", "diag"))
    R = R.ravel()
    a = s["a"]
    if kind == "largest":
        P2 = P.copy()
        P2[:, 0] += 40.0

        def weave(x, y):
            out = np.empty((2 * len(x),) + x.shape[1:], x.dtype)
            out[0::2], out[1::2] = x, y[::-1]
            return out
        P, V, R, D = weave(P, P2), weave(V, V), weave(R, R), weave(D, D)
    else:
        c1, c2 = (np.zeros((125, 4), np.float32) for _ in range(2))
        c1[:, :3] = blob(5, 1.5 * a, (30.0, 0, 0))
        c2[:, :3] = blob(5, 1.5 * a, (0, 30.0, 0))
        c1[:, 3] = c2[:, 3] = 1.0
        lone = np.zeros((6, 4), np.float32)   # single particles far off: N = 16,640 = 128 x 130
        lone[:, 2] = -30.0 - 10.0 * np.arange(6)
        P = np.vstack([P, c1[:1], c2, c1[1:], lone])
        extra = len(P) - len(R)
        V = np.vstack([V, np.zeros((extra, 4), np.float32)])
        R = np.concatenate([R, np.ones(extra, np.float32)])
        D = np.vstack([D, np.zeros((extra, 4), np.float32)])
    return P, V, R, D, s


def label_choice_differs(P, s, kind):
    """Whether the reference's choice (first root by size, src/nbody.js:680-684)
    differs from picking the tied group with the smaller smallest index."""
    lk = oracle.fof_link(P, s["a"], 2 * s["a"], s["k"])
    root, lab = lk["root"], lk["label"]
    N = len(root)
    size = np.bincount(root, minlength=N)
    g1 = g2 = -1
    for i in np.nonzero(size)[0]:
        if g1 < 0 or size[i] > size[g1]:
            g1, g2 = i, g1
        elif g2 < 0 or size[i] > size[g2]:
            g2 = i
    by_label = sorted(np.nonzero(np.bincount(lab, minlength=N))[0], key=lambda i: (-np.sum(lab == i), i))[:2]
    ref = [lab[g1], lab[g2]]
    return ref[0] != by_label[0] if kind == "largest" else ref[1] != by_label[1]


