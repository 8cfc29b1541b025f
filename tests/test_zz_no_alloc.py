"""cc_pp_noalloc -- the pairwise pass on the caller's device arrays, with no
context and no allocation (its workspace a static device array) -- against
the context's own pass (cc_set_pp + cc_pp) on the same state and boxes: the
forces bit for bit, on two calls in a row, the second with a smaller loose
list; and the size limit refused.  The device arrays are made through the
HIP runtime the library itself runs on.  Needs the GPU."""
import ctypes as C
import json
import os
import sys

import numpy as np
import pytest

import golden_util as gu

sys.path.insert(0, os.path.join(gu.REPO, "cosmic-collisions_amd"))
import ccamd  # noqa: E402

pytestmark = pytest.mark.gpu


class Dev:
    """A device copy of a host array, through the library's own HIP runtime."""

    def __init__(self, hip, host):
        self.hip, self.host = hip, np.ascontiguousarray(host)
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), C.c_size_t(self.host.nbytes)) == 0
        assert hip.hipMemcpy(self.p, self.host.ctypes.data_as(C.c_void_p), C.c_size_t(self.host.nbytes), 1) == 0

    def read(self):
        out = np.empty_like(self.host)
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, C.c_size_t(out.nbytes), 2) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def test_pp_noalloc_equals_the_context_pass():
    fx = gu.fixture("c1_impact")
    s = json.loads(str(fx["sim"]))
    ctx = ccamd.Context(ccamd.params(N=s["N"], W=s["W"], G=s["G"], m=s["m"], a=s["a"], cell=s["cell"], dt=s["dt"],
                                     k=s["k"], c=s["c"], creep=s["creep"], damp=s["damp"], tK=s["tK"],
                                     kappa=s["kappa"], spin0=s["spin0"], spin1=s["spin1"]))
    hip = C.CDLL("libamdhip64.so.7")   # the runtime the library has loaded
    bufs = []
    try:
        N = ctx.N
        ctx.upload(fx["pos"], fx["vel"], fx["mat"], fx["q"])
        u = gu.draws(fx, "ppgather")[0]
        LA, LB, boxB = [float(x) for x in u["uLA"]], [float(x) for x in u["uLB"]], float(u["uBoxB"][0]) > 0.5
        ctx.set_boxes(LA, LB, boxB)
        ctx.set_pp_table(fx["pptab"])
        pos, mat, tab = (Dev(hip, np.asarray(x, np.float32)) for x in (fx["pos"], fx["mat"], fx["pptab"]))
        force = Dev(hip, np.full((4, 8192, 4), np.nan, np.float32))   # every element must be written
        bufs += [pos, mat, tab, force]
        rng = np.random.default_rng(11)
        lists = [rng.choice(N, 3000, replace=False).astype(np.int32), np.asarray(fx["pp_index"], np.int32)]
        assert len(lists[1]) < len(lists[0])
        for k, lst in enumerate(lists):
            print("cc_pp_noalloc call", k, "count", len(lst), flush=True)
            idx = Dev(hip, lst)
            bufs.append(idx)
            ccamd.pp_noalloc(pos.p.value, mat.p.value, N, LA, LB, boxB, idx.p.value, len(lst), s["m"], s["a"],
                             tab.p.value, force.p.value)
            got = force.read()
            ctx.set_pp(lst)
            _, _, want = ctx.read_pp()
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), k
            assert got[:, :len(lst)].any() and not got[:, len(lst):].any()
        with pytest.raises(ccamd.CCError, match="at most 8192"):
            ccamd.pp_noalloc(pos.p.value, mat.p.value, N, LA, LB, boxB, idx.p.value, 8193, s["m"], s["a"],
                             tab.p.value, force.p.value)
    finally:
        for b in bufs:
            b.free()
        ctx.close()
