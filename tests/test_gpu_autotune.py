"""The launch-shape tuning (cc_set_tuning, cc_autotune): every cell-pass grid
and step-pass work-group order gives bit-identical meshes and states, an
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
    for cb in CELL:
        for order in (0, 1):
            print(name, "tuning", cb, order, flush=True)
            ctx = make(fx, True)
            try:
                ctx.set_tuning(cb, order)
                ctx.mesh()
                force = [ctx.read_mesh(b, "force") for b in (0, 1)]
                ctx.step_n(20, ccamd.sim_args(), boxes(fx, 20))
                got = force + list(ctx.download()[:4])
            finally:
                ctx.close()
            if ref is None:
                ref = got
            for a, b, what in zip(ref, got, ("force0", "force1", "pos", "vel", "aux", "q")):
                assert np.array_equal(a, b), (cb, order, what)


def test_autotune_leaves_the_run_unchanged():
    fx = gu.fixture("c1_impact")
    tuned, plain = make(fx, True), make(fx, True)
    try:
        print("autotune", flush=True)
        rep = tuned.autotune(reps=5)
        assert rep["cellBlocks"] in CELL and rep["simOrder"] in (0, 1)
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
            tuned.set_tuning(3000, 0)
    finally:
        tuned.close()
        plain.close()
