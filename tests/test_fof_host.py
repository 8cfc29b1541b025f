"""The analysis' host half (cosmic-collisions_amd/csrc/cc_fof_host.cpp: the
choice of the largest and second group, and every FP64 sum of the report in
the reference's order) on the CPU, fed by the oracle's restatement of the
reference's linking (oracle/cc_oracle.c, cco_fof_link) where the library
uses its device half.  Checked bit for bit against the reference's own
reports on the golden states, and against the oracle's whole analysis
(oracle/fof.js) on states built so that the largest or the second group ties
in size, where only a replay of the reference's union order settles which
group the reference reports."""
import ctypes as C
import json
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

import golden_util as gu
import oracle
from fof_states import label_choice_differs, run_oracle_js, same_report, tie_state

sys.path.insert(0, os.path.join(gu.REPO, "cosmic-collisions_amd"))
import ccamd  # noqa: E402

NODE = shutil.which("node")
GXX = shutil.which("g++")


@pytest.fixture(scope="module")
def host(tmp_path_factory):
    if GXX is None:
        pytest.skip("g++ is not installed")
    d = tmp_path_factory.mktemp("fofh")
    so = str(d / "libfofh.so")
    src = [os.path.join(gu.REPO, "tests", "fof_host_shim.cpp"),
           os.path.join(gu.REPO, "cosmic-collisions_amd", "csrc", "cc_fof_host.cpp")]
    subprocess.run([GXX, "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-I" + os.path.join(gu.REPO, "include"),
                    "-o", so] + src, check=True)
    L = C.CDLL(so)
    vp = C.c_void_p
    L.fofh_run.argtypes = [C.c_int, vp, vp, vp, vp, vp, vp, C.c_long, C.c_double, C.c_double, C.c_double, C.c_double,
                           C.c_int, C.c_int, vp, vp, C.POINTER(ccamd.Report), vp]
    L.fofh_run.restype = C.c_int
    return L


def run_host(L, P, V, R, D, m, a, spring, touch, cap=8192, flags=0):
    """The host half on the oracle's linking: (report dict, list)."""
    P, V, D = (np.ascontiguousarray(x, np.float32).reshape(-1, 4) for x in (P, V, D))
    R = np.ascontiguousarray(R, np.float32)
    N = len(R)
    lk = oracle.fof_link(P, a, touch, spring, pairs=True)
    rep = ccamd.Report()
    lst = np.zeros(max(cap, 1), np.int32)
    r = L.fofh_run(N, P.ctypes.data, V.ctypes.data, D.ctypes.data, R.ctypes.data, lk["label"].ctypes.data,
                   lk["e2"].ctypes.data, len(lk["e2"]), m, a, spring, touch, cap, flags, lk["linkOff"].ctypes.data,
                   lk["partner"].ctypes.data, C.byref(rep), lst.ctypes.data)
    assert r == 0
    d = ccamd.report_dict(rep, lst)
    assert d["EL"] == lk["elastic"]
    return d


@pytest.mark.parametrize("flags", [0, ccamd.REPLAY_UNIONS])
@pytest.mark.parametrize("name", ["impact", "twobox"])
def test_host_half_matches_reference(host, name, flags):
    """Every field of the reference's report on the golden state, bit for bit,
    with and without the replay of the union order; the loose list likewise."""
    fx = gu.fixture("c1_" + name)
    s = json.loads(str(fx["sim"]))
    mine = run_host(host, fx["pos"], fx["vel"], fx["mat"], fx["diag"], s["m"], s["a"], s["k"], 2 * s["a"], flags=flags)
    ref = json.loads(str(fx["report"]))
    same_report(mine, ref)
    assert mine["pp"]["loose"] == ref["ppLoose"]
    into = np.empty(len(fx["perm"]), np.int64)
    into[fx["perm"].astype(np.int64)] = np.arange(len(fx["perm"]))
    assert np.array_equal(into[mine["pp"]["list"]].astype(np.float32), fx["pplist_fof"])
    assert mine["replayed"] == (1 if flags else mine["replayed"])


@pytest.mark.skipif(NODE is None, reason="node is not installed")
@pytest.mark.parametrize("kind", ["largest", "second"])
def test_host_half_ties_replay(host, tmp_path, kind):
    """Equal groups: the labels cannot say which the reference reports, the
    host half replays the reference's union order and agrees with the oracle."""
    P, V, R, D, s = tie_state(kind)
    assert label_choice_differs(P, s, kind)
    mine = run_host(host, P, V, R, D, s["m"], s["a"], s["k"], 2 * s["a"])
    ref, lst = run_oracle_js(tmp_path, P, V, R, D, s["m"], s["a"], s["k"], 2 * s["a"])
    print("%s tie: groups %d, largest %d, second %d, replayed %d" % (kind, mine["groups"], mine["largestN"],
                                                                     mine["secondN"], mine["replayed"]))
    assert mine["replayed"] == 1
    same_report(mine, ref)
    assert mine["pp"]["loose"] == ref["ppLoose"]
    assert np.array_equal(mine["pp"]["list"], lst)


def test_oracle_link_brute_force():
    """The oracle's linking on a small random cloud: its groups are the
    connected components of the pairs closer than 2.4a (brute force), each
    labelled by its smallest index, and the touching pairs' count is that of
    the pairs closer than the contact distance."""
    rng = np.random.default_rng(3)
    a = 0.05
    P = np.zeros((600, 4), np.float32)
    P[:, :3] = rng.uniform(-1.0, 1.0, (600, 3))
    lk = oracle.fof_link(P, a, 2 * a, 1.0)
    x = P[:, :3].astype(np.float64)
    d2 = ((x[:, None, :] - x[None, :, :]) ** 2).sum(-1)
    adj = d2 < (2.4 * a) ** 2
    lab = np.arange(600)
    for _ in range(600):
        nb = np.where(adj, lab[None, :], 600).min(1)
        new = np.minimum(lab, nb)
        if np.array_equal(new, lab):
            break
        lab = new
    assert np.array_equal(lk["label"], lab)
    iu = np.triu_indices(600, 1)
    assert len(lk["e2"]) == int((d2[iu] < (2 * a) ** 2).sum())
