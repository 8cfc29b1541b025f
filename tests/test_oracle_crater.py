"""The oracle against the reference's crater page (chicxulub.html) on
SwiftShader: its contact seats, the field-gravity step pass with the frozen
shell (simFS({gravity: 'field', frozen: true})) in settle and full mode, and
the coupling pass at touchdown (COUPLE_FS).  CPU only.

Every pass compares from the identical input state the reference's pass
read (the fixture holds the state dumped before and after each pass).
Bars: seats bit-exact; neighbour counts exact; FP32 values 1e-5 relative
(|x - ref| <= 1e-5 max(|ref|, 1e-3 rms(ref))).  The coupling's flow direction
comes from acos and tan, whose precision GLSL ES 3.00 leaves to the
implementation; SwiftShader's differ from correctly rounded ones by up to
3.2e-5 rad on this state, so the direction is held to 1e-4 rad (and pinned
exactly by the Z-model's closed form below), its speed, temperature and heat
to 1e-5.
"""
import numpy as np
import pytest

import golden_util as gu
import oracle


@pytest.fixture(scope="module")
def fx():
    return gu.fixture("c3_crater")


def _seats(fx, tag, g):
    return (gu.dense(fx, tag + "_seatA", (g.AH, g.AW), gu.EMPTY), gu.dense(fx, tag + "_seatB", (g.AH, g.AW), gu.EMPTY))


@pytest.mark.parametrize("src,tag", [("b0", "s1"), ("post", "f1")])
def test_crater_seats_bit_exact(fx, src, tag):
    u = gu.draws(fx, "sim")[0]
    g = oracle.grid(int(u["uGMask"][0]) + 1)
    pos, _, _ = gu.crater_state(fx, src)
    a, b = oracle.slots(pos, float(gu.f(u, "uInvCell")), g)
    ra, rb = _seats(fx, tag, g)
    assert np.array_equal(a, ra)
    assert np.array_equal(b, rb)


@pytest.mark.parametrize("k", range(4))
def test_crater_step_pass(fx, k):
    """A12 field variant: contacts, heat, constant gravity, the frozen shell
    held still; settle steps (b0 -> s1 -> s2) and full steps after the
    coupling (post -> f1 -> f2)."""
    src, dst, settle = gu.CRATER_STEPS[k]
    u = gu.draws(fx, "sim")[k]
    assert float(u["uSettle"][0]) == settle
    su = gu.sim_u(u, mode="field")
    assert su.frozenOn == 1 and su.field == 1
    g = su.grid
    seat_tag = "s1" if dst in ("s1", "s2") else "f1"
    sa, sb = _seats(fx, seat_tag, g)
    pos, vel, q = gu.crater_state(fx, src)
    N = pos.shape[0]
    z4, z1 = np.zeros((513, 512, 4), np.float32), np.zeros(4, np.float32)
    slot = np.full(N, -1.0, np.float32)
    ppf = np.zeros((256, 128, 4), np.float32)
    P, V, A, Q = oracle.sim(su, pos, vel, fx["mat"], q, sa, sb, z4, z1, z4, z1, slot, ppf)
    rp, rv, ra = (fx["%s_%s" % (dst, n)].reshape(N, 4) for n in ("pos", "vel", "aux"))
    rq = fx[dst + "_q"].reshape(N, 4)[:, 0]
    assert np.array_equal(A[:, 3], ra[:, 3]), "neighbour counts"
    assert np.array_equal(P[:, 3], rp[:, 3]), "materials"
    frozen = np.abs(pos[:, 3] - su.frozen) < 0.5
    assert frozen.any() and np.array_equal(P[frozen, :3], pos[frozen, :3]), "the frozen shell moved"
    for mine, ref, what in ((A[:, :3], ra[:, :3], "gravity"), (V[:, :3], rv[:, :3], "velocity"),
                            (P[:, :3], rp[:, :3], "position"), (Q, rq, "heat"), (V[:, 3], rv[:, 3], "temperature")):
        err = gu.rel_err(mine, ref)
        assert err.max() < 1e-5, (what, err.max(), int(err.argmax()))


def couple_args(u, melt):
    """The 15 floats of the coupling pass's uniforms, as the pass received them."""
    vals = list(u["uCentre"][:3]) + [u[k][0] for k in ("uVimp", "uAimp", "uC1", "uInvMu", "uPexp", "uStall", "uZ",
                                                       "uHeat", "uTk", "uFrozen", "uAmbient")] + [melt]
    return np.array(vals, np.float32)


def test_crater_couple(fx):
    """A20 COUPLE_FS at touchdown: Housen-Holsapple speed, Z-model direction,
    melt heat; the frozen shell stopped at ambient temperature."""
    u = gu.draws(fx, "couple")[0]
    melt = gu.crater_host(fx, "crater_build")[0]["consts"]["MAT_MELT"]
    N = fx["pre_pos"].size // 4
    pre = [fx["pre_" + n].reshape(N, 4) for n in ("pos", "vel", "aux")]
    P, V, A, Q = oracle.couple(couple_args(u, melt), pre[0], pre[1], pre[2], fx["mat"])
    rp, rv, ra = (fx["post_" + n].reshape(N, 4) for n in ("pos", "vel", "aux"))
    rq = fx["post_q"].reshape(N, 4)[:, 0]
    assert np.array_equal(A, ra)
    assert np.array_equal(P[:, :3], rp[:, :3])
    moved = (rp[:, 3] != pre[0][:, 3])
    assert moved.sum() > 0
    assert np.array_equal(P[:, 3], rp[:, 3]), "melt labels"
    sp = np.linalg.norm(V[:, :3].astype(np.float64), axis=1)
    rsp = np.linalg.norm(rv[:, :3].astype(np.float64), axis=1)
    moving = rsp > 1e-3 * np.sqrt(np.mean(rsp ** 2))
    assert moving.sum() > 1000
    turn = np.linalg.norm(V[moving, :3] / sp[moving, None] - rv[moving, :3] / rsp[moving, None], axis=1)
    assert turn.max() < 1e-4, turn.max()
    for mine, ref, what in ((sp, rsp, "speed"), (V[:, 3], rv[:, 3], "temperature"), (Q, rq, "heat")):
        err = gu.rel_err(mine, ref)
        assert err.max() < 1e-5, (what, err.max(), int(err.argmax()))


def test_crater_couple_closed_form(fx):
    """The coupling pinned by its published formulas at four points: a grain
    at the source's range 2a straight below (Housen-Holsapple speed, Z-model
    direction straight down), one level with the source (theta = 90 deg: the
    Z-model turns the radial flow up by (Z - 2) tan 45 deg), one past the stall
    radius (no flow), one frozen."""
    u = gu.draws(fx, "couple")[0]
    args = couple_args(u, 3.0)
    c = args[:3].astype(np.float64)
    vimp, aimp, c1, inv_mu, pexp, stall, z, heat = (float(x) for x in args[3:11])
    pos = np.array([[c[0], c[1] - 2 * aimp, c[2], 1.0], [c[0] + 3 * aimp, c[1], c[2], 2.0],
                    [c[0] + 1.01 * stall, c[1], c[2], 2.0], [c[0], c[1] + 1, c[2], 0.0]], np.float32)
    zeros = np.zeros((4, 4), np.float32)
    P, V, A, Q = oracle.couple(args, pos, zeros, zeros, np.ones(4, np.float32))

    def speed(r):
        return min(c1 * vimp * (r / aimp) ** (-inv_mu), vimp) * (1 - r / stall) ** pexp
    v = speed(2 * aimp)
    assert V[0, :3] == pytest.approx([0.0, -v, 0.0], rel=1e-5, abs=1e-6)
    assert V[0, 3] == pytest.approx(300 + heat * 0.5 * v * v, rel=1e-5)
    d = np.array([1.0, z - 2.0, 0.0]) / np.hypot(1.0, z - 2.0)
    assert V[1, :3] == pytest.approx(d * speed(3 * aimp), rel=1e-5, abs=1e-6)
    assert np.array_equal(V[2, :3], np.zeros(3, np.float32))
    assert np.array_equal(V[3], np.array([0, 0, 0, 300], np.float32)) and Q[3] == 0
