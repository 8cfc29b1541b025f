"""The batched what-if sweep at its stated size: 64 giant impacts of uneven
size (16,384-262,144 particles, log-uniform, multiples of 512, so most are
not powers of two), each with its own seeded mass ratio, angle, speed and
core fraction, every one through the engine's build, settle, approach on
rails, kick and look (js/prep_batch.js).  Stepped as one batch -- one graph
launch per 8-step cycle for all 64 -- every member's state is bit-identical
to the same instance stepped alone, after calls that start on and off a mesh
step.  Needs the GPU."""
import os
import sys

import numpy as np
import pytest

import golden_util as gu

sys.path.insert(0, os.path.join(gu.REPO, "cosmic-collisions_amd"))
import runs  # noqa: E402

pytestmark = pytest.mark.gpu


def test_sweep_of_64_members_each_equals_itself_alone(tmp_path):
    print("preparing the 64-instance sweep", flush=True)
    inst = runs.prepare_batch(str(tmp_path), 64)
    sizes = [run["sim"]["N"] for run, _ in inst]
    assert len(sizes) == 64 and min(sizes) >= 16384 and max(sizes) <= 262144
    assert sum(1 for n in sizes if n & (n - 1)) > 32, "mostly sizes that are not powers of two"
    assert len({run["sim"]["G"] for run, _ in inst}) == 2, "both contact-grid sizes"
    B = runs.BatchRunner(inst)
    alone = [runs.Runner(run, st) for run, st in inst]
    try:
        for call, n in enumerate((24, 21, 19)):
            print("sweep call", call, "steps", n, flush=True)
            B.steps(n)
            for r in alone:
                r.steps(n)
            for k, (rb, ra) in enumerate(zip(B.rs, alone)):
                assert rb.step_no == ra.step_no
                for x, y, what in zip(rb.ctx.download(), ra.ctx.download(), ("pos", "vel", "aux", "q", "mat")):
                    assert np.array_equal(x, y), (call, k, sizes[k], what)
    finally:
        B.close()
        for r in alone:
            r.close()
