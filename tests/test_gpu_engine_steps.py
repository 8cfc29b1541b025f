"""A13 on the GPU: the JavaScript engine (js/nbody.js over the N-API addon, the
drop-in for CC.nbody) takes the same eight step() calls from the same state as
the reference's golden (tests/golden/fixtures/c1_impact.npz: a mesh step, then
seven more, the softening due on the fifth).  Its sim scalars afterwards (t,
step number, dt, k, c, damping, soft) equal the reference's, and its states
after steps 0, 1, 4 and 7 equal those of the reference's schedule run on the
oracle's passes (bit for bit against the reference's own states on the CPU,
tests/test_oracle_golden.py) fed with the engine's own mesh-step force mesh,
monopole and pairwise sums -- so a seat rebuild, impulse or softening on
another step than the reference's shows at once.  Needs the GPU."""
import json
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

import golden_util as gu
from test_oracle_golden import step_chain

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(shutil.which("node") is None, reason="node is not installed")]


def test_engine_steps_follow_the_reference_schedule(tmp_path):
    fx = gu.fixture("c1_impact")
    meta = json.loads(str(fx["steps_meta"]))
    d = str(tmp_path)
    for nm in ("pos", "vel", "mat", "q"):
        np.ascontiguousarray(fx[nm], np.float32).tofile(os.path.join(d, nm + ".f32"))
    np.ascontiguousarray(fx["pp_index"], np.int32).tofile(os.path.join(d, "pp.i32"))
    job = {"opt": json.loads(str(gu.fixture("c1_build")["opt"])), "dir": d, "fields": meta["fields"], "steps": 8,
           "dump": meta["dumped"]}
    json.dump(job, open(os.path.join(d, "job.json"), "w"))
    print("engine steps from the impact state", flush=True)
    r = subprocess.run(["node", os.path.join(gu.HERE, "engine_steps.js"), os.path.join(d, "job.json"), d],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    after = json.load(open(os.path.join(d, "after.json")))
    for k in ("stepNo", "soft", "dt", "k", "c", "damp", "t"):
        assert after[k] == meta["after"][k], (k, after[k], meta["after"][k])
    N = fx["pos"].shape[0]
    rd = lambda nm, shape: np.fromfile(os.path.join(d, nm), np.float32).reshape(shape)
    out = step_chain(fx, rd("force0.f32", (513, 512, 4)), rd("mono0.f32", (4,)), rd("ppforce.f32", (256, 128, 4)))
    exact = True
    for n, (P, V, A, Q) in out.items():
        got = {nm: rd("step%d.%s.f32" % (n, nm), (N, 4)) for nm in ("pos", "vel", "aux")}
        gq = rd("step%d.q.f32" % n, (N,))
        assert np.array_equal(got["aux"][:, 3], A[:, 3]), (n, "neighbour counts")
        for mine, ref, what in ((got["pos"], P, "pos"), (got["vel"], V, "vel"), (got["aux"][:, :3], A[:, :3], "grav"),
                                (gq, Q, "heat")):
            exact = exact and np.array_equal(mine, ref)
            assert gu.rel_err(mine, ref).max() < 1e-5, (n, what, gu.rel_err(mine, ref).max())
    print("engine states bit-identical to the oracle's schedule:", exact, flush=True)
