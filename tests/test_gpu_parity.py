"""The HIP library (through its C ABI) against the reference's golden outputs
and the CPU oracle.  Needs the GPU.

Bars: bit-exact for the seat atlases (hashed cells, contact lists, seat
words), neighbour counts, the loose-list gather and the Morton permutation;
1e-5 relative (|x - ref| <= 1e-5 max(|ref|, 1e-3 rms(ref))) for FP32 values --
component-wise for the step pass's outputs, per vector for the mesh and
pairwise accelerations, whose sums run in another order than the reference's.
"""
import json
import os
import sys

import numpy as np
import pytest

import golden_util as gu
import oracle

sys.path.insert(0, os.path.join(gu.REPO, "cosmic-collisions_amd"))
import ccamd  # noqa: E402

pytestmark = pytest.mark.gpu
CASES = ["impact", "twobox"]


def params_from(fx):
    s = json.loads(str(fx["sim"]))
    return ccamd.params(N=s["N"], W=s["W"], G=s["G"], m=s["m"], a=s["a"], cell=s["cell"], dt=s["dt"], k=s["k"],
                        c=s["c"], creep=s["creep"], damp=s["damp"], tK=s["tK"], kappa=s["kappa"], spin0=s["spin0"],
                        spin1=s["spin1"])


def boxes_from(u):
    return [float(x) for x in u["uLA"]], [float(x) for x in u["uLB"]], float(u["uBoxB"][0]) > 0.5


@pytest.fixture(scope="module", params=CASES)
def case(request):
    fx = gu.fixture("c1_" + request.param)
    fx["name"] = request.param
    return fx


@pytest.fixture(scope="module")
def ctxs():
    made = {}

    def get(fx):
        if fx["name"] not in made:
            made[fx["name"]] = ccamd.Context(params_from(fx))
        return made[fx["name"]]
    yield get
    for c in made.values():
        c.close()


def load_state(ctx, fx):
    ctx.upload(fx["pos"], fx["vel"], fx["mat"], fx["q"])


def test_seats_bit_exact(case, ctxs):
    ctx = ctxs(case)
    load_state(ctx, case)
    print("case", case["name"], "N", ctx.N, flush=True)
    ctx.slots()
    a, b = ctx.read_seats()
    ra, rb = gu.seats(case, ctx.AW, ctx.AH)
    assert np.array_equal(a, ra)
    assert np.array_equal(b, rb)


def test_mesh_deposit_coarse_far(case, ctxs):
    ctx = ctxs(case)
    load_state(ctx, case)
    LA, LB, boxB = boxes_from(gu.draws(case, "dep")[0])
    ctx.set_boxes(LA, LB, boxB)
    for b in ([0, 1] if case["name"] == "twobox" else [0]):
        ref = gu.mesh(case, b)
        mono_prev = np.zeros(4, np.float32)
        if b == 1:
            fields = json.loads(str(case["fields"]))
            m1 = float(case["mat"][case["pos"][:, 3] >= 3.5].sum())
            c = fields["boxes"][1]["cen"]
            mono_prev = np.array([c[0] * m1, c[1] * m1, c[2] * m1, m1], np.float32)
        ctx.set_mono(b, mono_prev)
        ctx.deposit(b)
        fine = ctx.read_mesh(b, "fine")
        # fixed-point sums against the reference's FP32 blend order
        np.testing.assert_allclose(fine[:512], ref["fine"][:512], rtol=1e-5, atol=1e-5)
        assert gu.rel_err(fine[512, 0], ref["fine"][512, 0]).max() < 1e-5
        ctx.coarse(b)
        np.testing.assert_allclose(ctx.read_mesh(b, "coarse"), ref["coarse"], rtol=1e-5, atol=1e-5)
        np.testing.assert_allclose(ctx.read_mesh(b, "mono"), ref["mono"], rtol=1e-5)
        # the far field from the reference's own coarse mesh
        ctx.write_mesh(b, "fine", ref["fine"])
        ctx.write_mesh(b, "coarse", ref["coarse"])
        ctx.cell(b)
        A = ctx.read_mesh(b, "farA")
        assert gu.rel_err_vec(A[..., :3], ref["farA"][..., :3]).max() < 1e-5
        assert gu.rel_err(A[..., 3], ref["farA"][..., 3]).max() < 1e-5
        # the tidal tensor as a whole: its six components (Txx Tyy Tzz Txy | Txz Tyz)
        T = np.concatenate([ctx.read_mesh(b, "farT0"), ctx.read_mesh(b, "farT1")[..., :2]], -1)
        Tr = np.concatenate([ref["farT0"], ref["farT1"][..., :2]], -1)
        assert gu.rel_err_vec(T, Tr).max() < 1e-5


def test_cell_vs_oracle(case, ctxs):
    """A7 on the GPU against the oracle's per-cell field from the same mesh
    (the force mesh that the step goldens were made with)."""
    ctx = ctxs(case)
    ref = gu.mesh(case, 0)
    for nm in ("fine", "coarse"):
        ctx.write_mesh(0, nm, ref[nm])
    ctx.cell(0)
    force = ctx.read_mesh(0, "force")
    want = gu.dense(case, "force0", (513, 512))
    occ = ref["fine"][:512, :, 3] > 0
    assert np.array_equal(force[:512][~occ], want[:512][~occ])
    assert gu.rel_err_vec(force[:512][occ][:, :3], want[:512][occ][:, :3]).max() < 1e-5
    assert gu.rel_err(force[:512][occ][:, 3], want[:512][occ][:, 3]).max() < 1e-5
    assert np.array_equal(force[512], want[512])


def test_pairwise(case, ctxs):
    ctx = ctxs(case)
    load_state(ctx, case)
    LA, LB, boxB = boxes_from(gu.draws(case, "ppgather")[0])
    ctx.set_boxes(LA, LB, boxB)
    ctx.set_pp_table(case["pptab"])
    ctx.set_pp(case["pp_index"])
    lst, box, force = ctx.read_pp()
    assert np.array_equal(lst, case["ppList"].reshape(-1, 4))
    assert np.array_equal(box, case["ppBox"][:, :, 0].ravel())
    ref = case["ppForce"].reshape(4, 8192, 4)
    assert gu.rel_err_vec(force[..., :3], ref[..., :3]).max() < 1e-5
    assert gu.rel_err(force[..., 3], ref[..., 3]).max() < 1e-5


def _prepare_step(ctx, fx):
    load_state(ctx, fx)
    u = gu.draws(fx, "sim")[0]
    LA, LB, boxB = boxes_from(u)
    ctx.set_boxes(LA, LB, boxB)
    ctx.slots()
    ctx.write_mesh(0, "force", gu.dense(fx, "force0", (513, 512)))
    ctx.write_mesh(0, "mono", gu.mesh(fx, 0)["mono"])
    if "force1_idx" in fx:
        ctx.write_mesh(1, "force", gu.dense(fx, "force1", (513, 512)))
        ctx.write_mesh(1, "mono", gu.mesh(fx, 1)["mono"])
    ctx.set_pp_table(fx["pptab"])
    ctx.set_pp(fx["pp_index"])


@pytest.mark.parametrize("variant", ["full", "settle", "kick"])
def test_step_pass(case, ctxs, variant):
    ctx = ctxs(case)
    _prepare_step(ctx, case)
    u = gu.draws(case, "sim")[["full", "settle", "kick"].index(variant)]
    args = ccamd.sim_args(settle=u["uSettle"][0] > 0.5, kick=u["uKick"][0] > 0.5, gdt=u["uGDt"][0],
                          bulk0=u["uBulk0"], bulk1=u["uBulk1"], cen0=u["uCen0"], cen1=u["uCen1"])
    ctx.sim(args)
    P, V, A, Q, _ = ctx.download()
    N = ctx.N
    rp = case["sim_%s_pos" % variant].reshape(N, 4)
    rv = case["sim_%s_vel" % variant].reshape(N, 4)
    ra = case["sim_%s_aux" % variant].reshape(N, 4)
    rq = case["sim_%s_q" % variant].reshape(N, 4)[:, 0]
    assert np.array_equal(A[:, 3], ra[:, 3]), "neighbour counts"
    assert np.array_equal(P[:, 3], rp[:, 3])
    for mine, ref, what in ((A[:, :3], ra[:, :3], "gravity"), (V[:, :3], rv[:, :3], "velocity"),
                            (P[:, :3], rp[:, :3], "position"), (Q, rq, "heat"), (V[:, 3], rv[:, 3], "temperature")):
        err = gu.rel_err(mine, ref)
        assert err.max() < 1e-5, (what, err.max(), int(err.argmax()))


def test_diag(case, ctxs):
    ctx = ctxs(case)
    _prepare_step(ctx, case)
    out = ctx.diag()
    ref = case["diag"].reshape(-1, 4)
    assert np.array_equal(out[:, 1:], ref[:, 1:])
    assert gu.rel_err(out[:, 0], ref[:, 0]).max() < 1e-5


def test_rigid(case, ctxs):
    ctx = ctxs(case)
    N = ctx.N
    load_state(ctx, case)
    u = gu.draws(case, "rigid")[0]
    ctx.rigid(u["uBulk0"], u["uBulk1"], u["uCen0"], u["uCen1"], u["uRot0"], u["uRot1"], u["uDt"][0])
    P, V, A, Q, _ = ctx.download()
    rp = case["rigid_pos"].reshape(N, 4)
    rv = case["rigid_vel"].reshape(N, 4)
    assert np.array_equal(P[:, 3], rp[:, 3])
    assert gu.rel_err(P[:, :3], rp[:, :3]).max() < 1e-5
    assert gu.rel_err(V, rv).max() < 1e-5


def test_morton_bit_exact(case, ctxs):
    ctx = ctxs(case)
    load_state(ctx, case)
    rep = json.loads(str(case["report"]))
    perm, into = ctx.morton(rep["core"])
    assert np.array_equal(perm.astype(np.float32), case["perm"])
    assert np.array_equal(into[perm], np.arange(ctx.N))
    P, V, A, Q, M = ctx.download()
    assert np.array_equal(P, case["pos"][perm])
    assert np.array_equal(M, case["mat"][perm])


def test_books_vs_oracle(case, ctxs):
    ctx = ctxs(case)
    _prepare_step(ctx, case)
    d = ctx.diag()
    b = ctx.books()
    s = json.loads(str(case["sim"]))
    ref = oracle.books(s["m"], case["pos"], case["vel"], case["mat"], d)
    for k in ("M", "px", "py", "pz", "Jx", "Jy", "Jz", "KE", "PE", "Q"):
        assert b[k] == pytest.approx(ref[k], rel=1e-12, abs=1e-14 * abs(ref["M"])), k
    assert b["unseated"] / ctx.N == pytest.approx(ref["unseated"], abs=0)
    assert b["full"] / ctx.N == pytest.approx(ref["full"], abs=0)
