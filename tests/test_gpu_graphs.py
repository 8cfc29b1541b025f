"""Replays of the captured step-cycle graphs (cc_step_n) against the same
steps launched one by one (cc_set_graphs(ctx, 0)): bit-identical state after
every call, with eager steps, a read-out pass and host synchronisations
between the replays, starting on and off a mesh step; with an odd mesh
cadence (cycles starting on odd and even steps) across a Morton sort.  Needs
the GPU."""
import json
import os
import sys

import numpy as np
import pytest

import golden_util as gu

sys.path.insert(0, os.path.join(gu.REPO, "cosmic-collisions_amd"))
import ccamd  # noqa: E402

pytestmark = pytest.mark.gpu


def make(fx, graphs, mesh_every=8):
    s = json.loads(str(fx["sim"]))
    ctx = ccamd.Context(ccamd.params(N=s["N"], W=s["W"], G=s["G"], m=s["m"], a=s["a"], cell=s["cell"], dt=s["dt"],
                                     k=s["k"], c=s["c"], creep=s["creep"], damp=s["damp"], tK=s["tK"], kappa=s["kappa"],
                                     spin0=s["spin0"], spin1=s["spin1"], meshEvery=mesh_every))
    ctx.set_graphs(graphs)
    ctx.upload(fx["pos"], fx["vel"], fx["mat"], fx["q"])
    u = gu.draws(fx, "dep")[0]
    ctx.set_boxes([float(x) for x in u["uLA"]], [float(x) for x in u["uLB"]], float(u["uBoxB"][0]) > 0.5)
    if "mono1_idx" in fx:
        ctx.set_mono(1, gu.mesh(fx, 1)["mono"])
    ctx.set_pp_table(fx["pptab"])
    ctx.set_pp(fx["pp_index"])
    return ctx


@pytest.mark.parametrize("name", ["impact", "twobox"])
def test_graph_replays_equal_eager(name):
    fx = gu.fixture("c1_" + name)
    g, e = make(fx, True), make(fx, False)
    try:
        u = gu.draws(fx, "dep")[0]
        LA = np.array([float(x) for x in u["uLA"]])
        LB = np.array([float(x) for x in u["uLB"]])
        boxB = float(u["uBoxB"][0])
        args = ccamd.sim_args()
        for call, (start, n) in enumerate([(0, 24), (None, 19), (None, 32), (3, 29)]):
            print(name, "call", call, "start", start, "n", n, flush=True)
            for ctx in (g, e):
                if start is not None:
                    ctx.step_no = start
                n_mesh = sum(1 for i in range(n) if (ctx.step_no + i) % 8 == 0)
                boxL = np.array([np.concatenate([LA + 1e-3 * k, LB - 1e-3 * k, [boxB]]) for k in range(n_mesh)])
                ctx.step_n(n, args, boxL)
            P1, V1, A1, Q1, _ = g.download()
            P2, V2, A2, Q2, _ = e.download()
            assert g.step_no == e.step_no
            for a, b, what in ((P1, P2, "pos"), (V1, V2, "vel"), (A1, A2, "aux"), (Q1, Q2, "q")):
                assert np.array_equal(a, b), (call, what)
            for ctx in (g, e):   # eager work between replays
                ctx.step(0, 0, args)
                ctx.diag()
                ctx.sync()
    finally:
        g.close()
        e.close()


@pytest.mark.parametrize("name", ["impact", "twobox"])
def test_graph_replays_follow_the_loose_list(name):
    """A captured cycle's pairwise pass reads the loose list's length on the device, so a cycle captured with
    one list replays with the next (a look's new list, longer or shorter): bit-identical to eager steps."""
    fx = gu.fixture("c1_" + name)
    g, e = make(fx, True), make(fx, False)
    try:
        u = gu.draws(fx, "dep")[0]
        LA = np.array([float(x) for x in u["uLA"]])
        LB = np.array([float(x) for x in u["uLB"]])
        boxB = float(u["uBoxB"][0])
        pp = np.asarray(fx["pp_index"])
        args = ccamd.sim_args()
        for call, keep in enumerate((len(pp), len(pp) // 2, len(pp), 7, len(pp) // 3)):
            print(name, "call", call, "loose list", keep, flush=True)
            for ctx in (g, e):
                ctx.set_pp(pp[:keep])
                n_mesh = sum(1 for i in range(16) if (ctx.step_no + i) % 8 == 0)
                boxL = np.array([np.concatenate([LA + 1e-3 * k, LB - 1e-3 * k, [boxB]]) for k in range(n_mesh)])
                ctx.step_n(16, args, boxL)
            for a, b, what in zip(g.download(), e.download(), ("pos", "vel", "aux", "q", "mat")):
                assert np.array_equal(a, b), (call, keep, what)
            for a, b in zip(g.read_pp(), e.read_pp()):
                assert np.array_equal(a, b), (call, keep, "pairwise")
    finally:
        g.close()
        e.close()


def test_odd_mesh_cadence_across_a_morton_sort():
    """meshEvery = 3: a cycle starts on an odd step as often as on an even one,
    and which of its steps rebuild the seats differs; a Morton sort between
    calls moves the state, mass and seat buffers.  Replays equal eager steps."""
    fx = gu.fixture("c1_impact")
    g, e = make(fx, True, 3), make(fx, False, 3)
    try:
        u = gu.draws(fx, "dep")[0]
        LA = np.array([float(x) for x in u["uLA"]])
        LB = np.array([float(x) for x in u["uLB"]])
        boxB = float(u["uBoxB"][0])
        core = json.loads(str(fx["report"]))["core"]
        args = ccamd.sim_args()
        for call, n in enumerate([7, 9, 6, 11, 12]):
            print("meshEvery 3, call", call, "n", n, flush=True)
            for ctx in (g, e):
                n_mesh = sum(1 for i in range(n) if (ctx.step_no + i) % 3 == 0)
                boxL = np.array([np.concatenate([LA + 1e-3 * k, LB - 1e-3 * k, [boxB]]) for k in range(n_mesh)])
                ctx.step_n(n, args, boxL)
            P1, V1, A1, Q1, M1 = g.download()
            P2, V2, A2, Q2, M2 = e.download()
            assert g.step_no == e.step_no
            for a, b, what in ((P1, P2, "pos"), (V1, V2, "vel"), (A1, A2, "aux"), (Q1, Q2, "q"), (M1, M2, "mat")):
                assert np.array_equal(a, b), (call, what)
            assert all(np.array_equal(a, b) for a, b in zip(g.read_seats(), e.read_seats())), (call, "seats")
            if call in (1, 3):
                for ctx in (g, e):
                    ctx.morton(core)
    finally:
        g.close()
        e.close()
