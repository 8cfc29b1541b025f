"""The HIP library at the benchmark's full size (262,144 particles, contact
grid G = 128) against the CPU oracle on the same seeded state, plus
size-independent properties.  Needs the GPU."""
import math
import os
import sys

import numpy as np
import pytest

import golden_util as gu
import oracle

sys.path.insert(0, os.path.join(gu.REPO, "cosmic-collisions_amd"))
sys.path.insert(0, gu.REPO)
import synthetic as bench  # noqa: E402  (the twins scalars and the lattice state)
import ccamd  # noqa: E402

pytestmark = pytest.mark.gpu
FL, FH = 6.0, 12.0 / 64


@pytest.fixture(scope="module")
def run():
    sc = bench.twins_scalars()
    pos, vel, mat, q, c0, c1 = bench.synthetic_twins(sc, seed=0xC0FFEE)
    rng = np.random.default_rng(7)
    vel[:, :3] += rng.normal(0, 0.05, (sc["N"], 3)).astype(np.float32)   # relative motion at every contact
    vel[:, 3] = rng.uniform(0, 4000, sc["N"]).astype(np.float32)
    q = rng.uniform(0, 1e-5, sc["N"]).astype(np.float32)
    p = ccamd.params(N=sc["N"], W=sc["W"], G=sc["G"], m=sc["m"], a=sc["a"], cell=sc["cell"], dt=sc["dt"], k=sc["k"],
                     c=sc["c"], spin0=sc["spin0"], spin1=sc["spin1"])
    ctx = ccamd.Context(p)
    print("full-size state N=%d" % sc["N"], flush=True)
    yield sc, pos, vel, mat, q, c0, c1, ctx
    ctx.close()


def oracle_u(sc, LA, gdt=0.0, cen0=(0, 0, 0), cen1=(0, 0, 0)):
    u = oracle.SimU()
    u.W, u.H = sc["W"], sc["N"] // sc["W"]
    for nm, v in (("invCell", 1 / sc["cell"]), ("cell", sc["cell"]), ("dt", sc["dt"]), ("gdt", gdt),
                  ("touch", 2 * sc["a"]), ("link2", sc["cell"] ** 2), ("reach2", (sc["cell"] + 0.6 * sc["a"]) ** 2),
                  ("k", sc["k"]), ("c", sc["c"]), ("invM", 1 / sc["m"]), ("damp", math.exp(-sc["dt"] * 12)),
                  ("creep", 0.12), ("tK", 0.5 * 6.256e7 / 1200 / sc["m"]), ("kappa", 0.3), ("cpIron", 450 / 1200)):
        setattr(u, nm, float(np.float32(v)))
    u.cen0[:] = [float(np.float32(x)) for x in cen0]
    u.cen1[:] = [float(np.float32(x)) for x in cen1]
    u.spin0[:] = [float(np.float32(x)) for x in sc["spin0"]]
    u.spin1[:] = [float(np.float32(x)) for x in sc["spin1"]]
    u.grid = oracle.grid(sc["G"])
    u.box = oracle.box(float(np.float32(1 / FH)), [float(np.float32(x)) for x in LA], [FL] * 3, 0.0)
    u.gm = float(np.float32(sc["m"]))
    u.eps2 = float(np.float32(FH * FH))
    return u


def test_seats_bit_exact_full_size(run):
    sc, pos, vel, mat, q, c0, c1, ctx = run
    ctx.upload(pos, vel, mat, q)
    ctx.slots()
    a, b = ctx.read_seats()
    oa, ob = oracle.slots(pos, np.float32(1 / sc["cell"]), oracle.grid(sc["G"]))
    assert np.array_equal(a, oa) and np.array_equal(b, ob)
    # and the seats' own invariants: a particle is seated at most once, seats ascend within a cell
    s = np.concatenate([a, b], -1).reshape(-1, 2, 4).transpose(0, 2, 1).reshape(-1, 8)
    filled = s != 0xBF800000
    idx = (s & 0x3FFFF).astype(np.int64)
    seated = idx[filled]
    assert len(np.unique(seated)) == len(seated)
    d = np.diff(np.where(filled, idx, 1 << 30), axis=1)
    assert (d[filled[:, 1:]] > 0).all()


def test_mesh_conserves_mass(run):
    sc, pos, vel, mat, q, c0, c1, ctx = run
    ctx.upload(pos, vel, mat, q)
    LA = [FL, FL, FL]
    ctx.set_boxes(LA, [FL] * 3, False)
    ctx.mesh()
    fine = ctx.read_mesh(0, "fine").astype(np.float64)
    coarse = ctx.read_mesh(0, "coarse").astype(np.float64)
    mono = ctx.read_mesh(0, "mono").astype(np.float64)
    g = (pos[:, :3].astype(np.float32) + np.float32(FL)) * np.float32(1 / FH) - np.float32(0.5)
    inbox = ((g >= 0) & (g <= 63)).all(1)
    m = mat[inbox].astype(np.float64).sum()
    assert fine[:512, :, 3].sum() == pytest.approx(m, rel=1e-6)
    assert coarse[..., 3].sum() == pytest.approx(m, rel=1e-6)
    assert mono[3] == pytest.approx(m, rel=1e-6)
    com = (pos[inbox, :3].astype(np.float64) * mat[inbox, None]).sum(0)
    assert np.allclose(fine[:512, :, :3].reshape(-1, 3).sum(0), com, rtol=1e-5, atol=1e-3)


def test_step_pass_vs_oracle_full_size(run):
    """One step pass at 262,144 particles from the same state, seats and force
    mesh: the GPU against the oracle at the 1e-5 bar."""
    sc, pos, vel, mat, q, c0, c1, ctx = run
    ctx.upload(pos, vel, mat, q)
    LA = [FL, FL, FL]
    ctx.set_boxes(LA, [FL] * 3, False)
    ctx.slots()
    ctx.mesh()
    force = ctx.read_mesh(0, "force")
    mono = ctx.read_mesh(0, "mono")
    gdt = 8 * sc["dt"]
    ctx.sim(ccamd.sim_args(gdt=gdt, cen0=c0, cen1=c1))
    P, V, A, Q, _ = ctx.download()
    oa, ob = oracle.slots(pos, np.float32(1 / sc["cell"]), oracle.grid(sc["G"]))
    u = oracle_u(sc, LA, gdt, c0, c1)
    slot = np.full(sc["N"], -1, np.float32)
    oP, oV, oA, oQ = oracle.sim(u, pos, vel, mat, q, oa, ob, force, mono, np.zeros_like(force), np.zeros(4, np.float32),
                                slot, np.zeros((256, 128, 4), np.float32))
    assert np.array_equal(A[:, 3], oA[:, 3]), "neighbour counts"
    for mine, ref, what in ((A[:, :3], oA[:, :3], "gravity"), (V, oV, "velocity"), (P, oP, "position"), (Q, oQ, "heat")):
        err = gu.rel_err(mine, ref)
        assert err.max() < 1e-5, (what, err.max(), int(err.argmax()))


def test_steps_deterministic(run):
    """Two runs of 16 full steps from the same state are bitwise identical
    (fixed-point deposit, ordered sums everywhere)."""
    sc, pos, vel, mat, q, c0, c1, ctx = run
    outs = []
    for _ in range(2):
        ctx.upload(pos, vel, mat, q)
        ctx.step_no = 0
        ctx.set_boxes([FL] * 3, [FL] * 3, False)
        ctx.step_n(16, ccamd.sim_args(cen0=c0, cen1=c1))
        outs.append(ctx.download())
    for x, y in zip(outs[0], outs[1]):
        assert np.array_equal(x, y)
