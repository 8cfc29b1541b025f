"""The analysis on the GPU (cc_analyse: the hashed cells, the linked pairs and
the groups on the device, every FP64 sum on the host in the reference's
order).  Bit for bit against the reference's own reports on the golden
states; against the oracle's whole analysis (oracle/fof.js) on states where
groups tie (the union order replayed) and on a 262,144-particle state (16
copies of the golden state, each with a different number of its members
moved off as single particles, so that no two groups tie) whose loose list
overflows the 8192-entry cap.  Needs the GPU."""
import json
import os
import shutil
import sys

import numpy as np
import pytest

import golden_util as gu
from fof_states import run_oracle_js, same_report, tie_state

sys.path.insert(0, os.path.join(gu.REPO, "cosmic-collisions_amd"))
import ccamd  # noqa: E402

pytestmark = pytest.mark.gpu


def context(P, V, R, s, W):
    N = len(R)
    p = ccamd.params(N=N, W=W, G=s["G"], m=s["m"], a=s["a"], cell=s["cell"], dt=s["dt"], k=s["k"], c=s["c"],
                     creep=s["creep"], damp=s["damp"], tK=s["tK"], kappa=s["kappa"], spin0=s["spin0"], spin1=s["spin1"])
    ctx = ccamd.Context(p)
    ctx.upload(np.ascontiguousarray(P, np.float32), np.ascontiguousarray(V, np.float32),
               np.ascontiguousarray(R, np.float32), np.zeros(N, np.float32))
    return ctx


def analyse(ctx, s, D, flags=0):
    return ctx.analyse(s["m"], s["a"], s["k"], 2 * s["a"], diag=D, flags=flags)


@pytest.mark.parametrize("flags", [0, ccamd.REPLAY_UNIONS])
@pytest.mark.parametrize("name", ["impact", "twobox"])
def test_analysis_matches_reference(name, flags):
    fx = gu.fixture("c1_" + name)
    s = json.loads(str(fx["sim"]))
    print("analysis: golden %s, flags %d" % (name, flags), flush=True)
    ctx = context(fx["pos"], fx["vel"], fx["mat"], s, s["W"])
    try:
        mine = analyse(ctx, s, fx["diag"], flags)
    finally:
        ctx.close()
    ref = json.loads(str(fx["report"]))
    same_report(mine, ref)
    assert mine["pp"]["loose"] == ref["ppLoose"]
    into = np.empty(len(fx["perm"]), np.int64)
    into[fx["perm"].astype(np.int64)] = np.arange(len(fx["perm"]))
    assert np.array_equal(into[mine["pp"]["list"]].astype(np.float32), fx["pplist_fof"])
    if flags:
        assert mine["replayed"] == 1


@pytest.mark.skipif(shutil.which("node") is None, reason="node is not installed")
@pytest.mark.parametrize("kind,W", [("largest", 256), ("second", 130)])
def test_analysis_ties(tmp_path, kind, W):
    P, V, R, D, s = tie_state(kind)
    print("analysis: %s tie, N=%d" % (kind, len(R)), flush=True)
    ctx = context(P, V, R, s, W)
    try:
        mine = analyse(ctx, s, D)
    finally:
        ctx.close()
    ref, lst = run_oracle_js(tmp_path, P, V, R, D, s["m"], s["a"], s["k"], 2 * s["a"])
    assert mine["replayed"] == 1
    same_report(mine, ref)
    assert mine["pp"]["loose"] == ref["ppLoose"]
    assert np.array_equal(mine["pp"]["list"], lst)


def tiled_state():
    """16 copies of the golden impact state 40 apart on a 4 x 4 grid; from copy
    k, 37 k members of the largest group are set out as single particles in a
    row 1.0 apart, so every copy's largest group has its own size."""
    fx = gu.fixture("c1_impact")
    s = json.loads(str(fx["sim"]))
    n = len(fx["mat"])
    P0 = np.asarray(fx["pos"], np.float32).reshape(n, 4)
    V0 = np.asarray(fx["vel"], np.float32).reshape(n, 4)
    D0 = np.asarray(fx["diag"], np.float32).reshape(n, 4)
    R0 = np.asarray(fx["mat"], np.float32).ravel()
    import oracle
    lab = oracle.fof_link(P0, s["a"], 2 * s["a"], s["k"])["label"]
    big = np.nonzero(lab == np.bincount(lab).argmax())[0]
    Ps = []
    for k in range(16):
        P = P0.copy()
        P[:, 0] += 40.0 * (k % 4)
        P[:, 1] += 40.0 * (k // 4)
        out = big[:37 * k]
        P[out, 0] = 200.0 + np.arange(len(out), dtype=np.float32)
        P[out, 1] = 10.0 * k
        P[out, 2] = 0.0
        Ps.append(P)
    return np.vstack(Ps), np.tile(V0, (16, 1)), np.tile(R0, 16), np.tile(D0, (16, 1)), s


@pytest.mark.skipif(shutil.which("node") is None, reason="node is not installed")
def test_analysis_262k_against_oracle(tmp_path):
    P, V, R, D, s = tiled_state()
    assert len(R) == 262144
    print("analysis: 16 tiled copies, N=262144", flush=True)
    ctx = context(P, V, R, s, 512)
    try:
        mine = analyse(ctx, s, D)
        again = analyse(ctx, s, D)
    finally:
        ctx.close()
    print("cc_analyse at 262,144: groups %d, largest %d, second %d, loose %d, ms cells %.2f copy %.2f host %.2f"
          % (mine["groups"], mine["largestN"], mine["secondN"], mine["pp"]["loose"], *again["ms"]))
    ref, lst = run_oracle_js(tmp_path, P, V, R, D, s["m"], s["a"], s["k"], 2 * s["a"])
    assert mine["replayed"] == 0
    assert mine["pp"]["loose"] > 8192
    same_report(mine, ref)
    same_report(again, ref)
    assert mine["pp"]["loose"] == ref["ppLoose"]
    assert np.array_equal(mine["pp"]["list"], lst)


def test_overlapped_analysis_equals_synchronous():
    """cc_analyse_begin / cc_analyse_end (the reference's analyze(false) and
    collect(), src/nbody.js:967-1025): the device's half and the read-back at
    the call, the host's sums on a thread while the context steps on.  The
    report is the synchronous report of the state it was begun on, bit for
    bit; a second begin before the end, and an end with none begun, are
    refused."""
    fx = gu.fixture("c1_impact")
    s = json.loads(str(fx["sim"]))
    ctx = context(fx["pos"], fx["vel"], fx["mat"], s, s["W"])
    try:
        ref = analyse(ctx, s, fx["diag"])
        ctx.analyse_begin(s["m"], s["a"], s["k"], 2 * s["a"], diag=fx["diag"])
        with pytest.raises(ccamd.CCError):
            ctx.analyse_begin(s["m"], s["a"], s["k"], 2 * s["a"], diag=fx["diag"])
        print("overlapped analysis: stepping while the host sums run", flush=True)
        ctx.step_n(16, ccamd.sim_args())
        got = ctx.analyse_end(wait=True)
        assert not np.array_equal(ctx.download()[0], fx["pos"]), "the state moved meanwhile"
        for k, v in ref.items():
            if k == "ms":
                continue
            if k == "pp":
                assert np.array_equal(got["pp"]["list"], v["list"]) and got["pp"]["loose"] == v["loose"]
            else:
                assert got[k] == v, k
        with pytest.raises(ccamd.CCError):
            ctx.analyse_end(wait=True)
    finally:
        ctx.close()
