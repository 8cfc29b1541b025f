"""The launch-shape tuning (cc_set_tuning, cc_autotune): every cell-pass grid
and step-pass work-group order, and the rebuild's count taken by the step
pass before it or by a pass of its own, give bit-identical meshes and
states (and the same seats), an
autotuned context runs on bit for bit with an untuned one from the same
state, and autotune refuses where its timing would disturb the run.  Needs
the GPU."""
import os
import sys

import numpy as np
import pytest

import golden_util as gu
from test_gpu_graphs import make

sys.path.insert(0, os.path.join(gu.REPO, "cosmic-collisions_amd"))
import ccamd  # noqa: E402

pytestmark = pytest.mark.gpu

CELL = (512, 1024, 2048, 4096)


def boxes(fx, n, start=0):
    u = gu.draws(fx, "dep")[0]
    LA = np.array([float(x) for x in u["uLA"]])
    LB = np.array([float(x) for x in u["uLB"]])
    k = sum(1 for i in range(n) if (start + i) % 8 == 0)
    return np.array([np.concatenate([LA + 1e-3 * j, LB - 1e-3 * j, [float(u["uBoxB"][0])]]) for j in range(k)])


@pytest.mark.parametrize("name", ["impact", "twobox"])
def test_every_tuning_gives_the_same_bits(name):
    fx = gu.fixture("c1_" + name)
    ref = None
    tunings = [(cb, order, 1) for cb in CELL for order in (0, 1)] + [(2048, 2, 1), (512, 2, 1), (2048, 0, 0),
                                                                    (2048, 2, 0), (2048, 3, 1), (2048, 3, 0)]
    for cb, order, fuse in tunings:
        print(name, "tuning", cb, order, fuse, flush=True)
        ctx = make(fx, True)
        try:
            ctx.set_tuning(cb, order, fuse)
            ctx.mesh()
            force = [ctx.read_mesh(b, "force") for b in (0, 1)]
            ctx.step_n(20, ccamd.sim_args(), boxes(fx, 20))
            got = force + list(ctx.download()[:4]) + list(ctx.read_seats())
        finally:
            ctx.close()
        if ref is None:
            ref = got
        for a, b, what in zip(ref, got, ("force0", "force1", "pos", "vel", "aux", "q", "seatA", "seatB")):
            assert np.array_equal(a, b), (cb, order, fuse, what)


@pytest.mark.parametrize("name", ["impact", "twobox"])
def test_visit_order_gives_the_same_bits(name):
    """The step pass's visit order (cc_set_visit): index order, a sort every
    8 steps and every 64, with cycle graphs and eagerly, through a Morton sort
    of the indices and an upload in mid-run: the same states and seats."""
    fx = gu.fixture("c1_" + name)
    N = fx["pos"].shape[0]
    perm_core = np.array([0.1, -0.2, 0.05])
    ref = None
    for every, graphs in ((0, True), (8, True), (64, True), (8, False)):
        print(name, "visit every", every, "graphs", graphs, flush=True)
        ctx = make(fx, graphs)
        try:
            ctx.set_visit(every)
            ctx.step_n(21, ccamd.sim_args(), boxes(fx, 21))
            ctx.morton(perm_core)
            ctx.step_n(19, ccamd.sim_args(), boxes(fx, 19, 21))
            P, V, A, Q, M = ctx.download()
            ctx.upload(P, V, M, Q)
            ctx.step_n(16, ccamd.sim_args(), boxes(fx, 16, 40))
            got = list(ctx.download()[:5]) + list(ctx.read_seats())
        finally:
            ctx.close()
        assert got[0].shape[0] == N
        if ref is None:
            ref = got
        for a, b, what in zip(ref, got, ("pos", "vel", "aux", "q", "mat", "seatA", "seatB")):
            assert np.array_equal(a, b), (every, graphs, what)


def test_count_in_step_pass_through_interruptions():
    """The rebuild's count taken by the step pass before it, against a count of
    its own, through what can leave a taken count stale: eager steps, a change
    of step number, an explicit step pass and rebuild, a Morton sort, an
    upload, and graph cycles starting on odd and even steps."""
    fx = gu.fixture("c1_impact")
    fused, alone = make(fx, True), make(fx, True)
    alone.set_tuning(2048, 0, 0)
    args = ccamd.sim_args()
    core = np.zeros(3)
    try:
        def both(what, f):
            print("interruption:", what, flush=True)
            for ctx in (fused, alone):
                f(ctx)
            a, b = fused.download(), alone.download()
            for x, y, k in zip(a, b, ("pos", "vel", "aux", "q", "mat")):
                assert np.array_equal(x, y), (what, k)
            for x, y in zip(fused.read_seats(), alone.read_seats()):
                assert np.array_equal(x, y), (what, "seats")
        both("13 steps", lambda c: c.step_n(13, args, boxes(fx, 13)))
        both("eager step", lambda c: c.step(0, 0, args))
        both("eager step", lambda c: c.step(0, 0, args))
        both("19 steps", lambda c: c.step_n(19, args, boxes(fx, 19, 15)))
        both("step number moved", lambda c: setattr(c, "step_no", 41))
        both("9 steps", lambda c: c.step_n(9, args, boxes(fx, 9, 41)))
        both("explicit step pass", lambda c: c.sim(args))
        both("5 steps", lambda c: c.step_n(5, args, boxes(fx, 5, 50)))
        both("explicit rebuild", lambda c: c.slots())
        both("7 steps", lambda c: c.step_n(7, args, boxes(fx, 7, 55)))
        both("Morton sort", lambda c: c.morton(core))
        both("11 steps", lambda c: c.step_n(11, args, boxes(fx, 11, 62)))
        P, V, A, Q, M = fused.download()
        both("upload", lambda c: c.upload(P, V, M, Q))
        both("16 steps", lambda c: c.step_n(16, args, boxes(fx, 16, 73)))
    finally:
        fused.close()
        alone.close()


def test_autotune_leaves_the_run_unchanged():
    fx = gu.fixture("c1_impact")
    tuned, plain = make(fx, True), make(fx, True)
    try:
        print("autotune", flush=True)
        rep = tuned.autotune(reps=5)
        assert rep["cellBlocks"] in CELL and rep["simOrder"] in (0, 1, 2, 3)
        assert all(t > 0 for t in rep["cellMs"] + rep["simMs"])
        args = ccamd.sim_args()
        for ctx in (tuned, plain):
            ctx.step_n(24, args, boxes(fx, 24))
        for a, b, what in zip(tuned.download(), plain.download(), ("pos", "vel", "aux", "q", "mat")):
            assert np.array_equal(a, b), what
        # off a mesh step the mesh it would time is still in use: refused
        tuned.step(0, 0, args)
        with pytest.raises(ccamd.CCError):
            tuned.autotune(reps=2)
        with pytest.raises(ccamd.CCError):
            tuned.set_tuning(3000, 0, 1)
    finally:
        tuned.close()
        plain.close()
