"""A7, the per-cell gravity pass (CELL_FS, src/glsl-sim.js:303-352), pinned
by what the physics fixes rather than by its own formula.  CPU only.

SwiftShader cannot run CELL_FS, so no reference output exists for it.  The
oracle's cco_cell (with cco_far feeding it) is checked here on constructed
meshes where the answer follows from the construction:

  * near field: every occupied cell inside one 3^3 cube of blocks, so that no
    block is beyond the near region and the far field is exactly zero; the
    field is then the softened direct sum over every other occupied cell
    (softening 0.5 FH^2, src/nbody.js:279), computed here by brute force over
    cell pairs with no block bookkeeping at all;
  * far field: isolated masses three and more blocks away from a block of
    target cells: the pass must give the first-order expansion about the
    target block's centre of mass of the unsoftened field of the far blocks
    as point masses at their own centres of mass -- the expansion's Jacobian
    taken here by central differences of that field, the potential to second
    order;
  * the mesh's own books, as the reference's meshCheck reads them
    (index.html:533-556): the net force and torque of the cells on each
    other, as fractions of the sums of their sizes, on the golden meshes.

Each check is shown to catch a planted mistake: the oracle built with one
of nine mistakes (a dropped, sign-flipped or mis-indexed tidal term, the
potential's second-order term dropped, a wrong softening, a 3^3 near region,
the own cell not skipped, near blocks counted as far, block centres in place
of centres of mass) fails at least one of the checks.
"""
import os
import subprocess

import numpy as np
import pytest

import golden_util as gu
import oracle

FL, FH = 6.0, 12.0 / 64
EPS2 = np.float32(0.5 * FH * FH)
GM = np.float32(0.37)
TOL = 1e-5


def fine_at(cx, cy, cz):
    return ((cz >> 3) * 64 + cy) * 512 + ((cz & 7) * 64 + cx)


def node(c):
    """A mesh node's position: the cell whose CIC weights peak there."""
    return (np.asarray(c, np.float64) + 0.5) * FH - FL


def mesh_of(cells):
    """A fine mesh (513 x 512 x 4) from [(cell index triple, mass, centre of mass)]."""
    fine = np.zeros((513 * 512, 4), np.float32)
    for c, m, x in cells:
        fine[fine_at(*c)] = [m * x[0], m * x[1], m * x[2], m]
    return fine.reshape(513, 512, 4)


def field(fine):
    """coarse, far field and per-cell field through the oracle (eps2 = 0.5 FH^2, G m = GM)."""
    co = oracle.coarse(fine)
    A, T0, T1 = oracle.far(co)
    return co, A, oracle.cell(fine, co, A, T0, T1, EPS2, GM)


def cells_in(rng, blocks, n):
    """n distinct occupied cells in the given blocks, masses 1..40, centres of mass jittered about the nodes."""
    out, seen = [], set()
    while len(out) < n:
        b = blocks[rng.integers(len(blocks))]
        c = tuple(int(4 * b[k] + rng.integers(4)) for k in range(3))
        if c in seen:
            continue
        seen.add(c)
        x = np.float32(node(c) + rng.uniform(-0.3, 0.3, 3) * FH).astype(np.float64)
        out.append((c, float(np.float32(rng.uniform(1, 40))), x))
    return out


def read(fine, cells):
    """(centre of mass as the mesh holds it, mass) of each listed cell, FP64."""
    f = fine.reshape(-1, 4).astype(np.float64)
    v = np.array([f[fine_at(*c)] for c, _, _ in cells])
    return v[:, :3] / v[:, 3:4], v[:, 3]


def near_sum(x, m, i, others):
    """The softened direct sum at cell i from the cells `others` (FP64)."""
    acc, pot = np.zeros(3), 0.0
    for j in others:
        if j == i:
            continue
        r = x[j] - x[i]
        inv = 1.0 / np.sqrt(r @ r + float(EPS2))
        acc += m[j] * r * inv ** 3
        pot -= m[j] * inv
    return acc, pot


def check_near():
    """Every occupied cell in blocks 7..9 per axis: the field is the softened
    sum over all other cells; the far field is exactly zero."""
    rng = np.random.default_rng(5)
    blocks = [(bx, by, bz) for bx in (7, 8, 9) for by in (7, 8, 9) for bz in (7, 8, 9)]
    cells = cells_in(rng, blocks, 60)
    fine = mesh_of(cells)
    co, A, F = field(fine)
    x, m = read(fine, cells)
    got = np.array([F.reshape(-1, 4)[fine_at(*c)] for c, _, _ in cells], np.float64)
    want = np.zeros_like(got)
    for i in range(len(cells)):
        a, p = near_sum(x, m, i, range(len(cells)))
        want[i, :3], want[i, 3] = a * float(GM), p * float(GM)
    return (not A.any(), gu.rel_err_vec(got[:, :3], want[:, :3]).max(), gu.rel_err(got[:, 3], want[:, 3]).max())


def mono_field(y, X, M):
    r = X - y
    d = np.sqrt((r ** 2).sum(1))
    return (M[:, None] * r / d[:, None] ** 3).sum(0), -(M / d).sum()


def check_far():
    """Target cells in block (8,8,8) with nothing else within two blocks; far
    blocks 3, 4, 5 and 7 blocks away, one of them holding three cells.  The
    pass must give near(x) + a(X_B) + J d for the acceleration and
    phi(X_B) - a(X_B).d - d.J d / 2 for the potential, where a, phi are the far
    blocks' field as point masses at their centres of mass, X_B the target
    block's centre of mass, d = x - X_B, J the Jacobian of a by central
    differences.  Returns the largest relative errors and the tidal term's
    size relative to the field (the margin by which a dropped tidal term fails)."""
    rng = np.random.default_rng(9)
    B = (8, 8, 8)
    target = cells_in(rng, [B], 7)
    far_blocks = [(11, 8, 8), (8, 4, 9), (3, 14, 8), (15, 15, 15)]
    far = [cells_in(rng, [b], 3 if k == 1 else 1) for k, b in enumerate(far_blocks)]
    fine = mesh_of(target + sum(far, []))
    co, A, F = field(fine)
    xt, mt = read(fine, target)
    XB = (mt[:, None] * xt).sum(0) / mt.sum()
    XK, MK = [], []
    for cl in far:
        xf, mf = read(fine, cl)
        XK.append((mf[:, None] * xf).sum(0) / mf.sum())
        MK.append(mf.sum())
    XK, MK = np.array(XK), np.array(MK)
    a0, p0 = mono_field(XB, XK, MK)
    h = 1e-4
    J = np.zeros((3, 3))
    for k in range(3):
        e = np.zeros(3)
        e[k] = h
        J[:, k] = (mono_field(XB + e, XK, MK)[0] - mono_field(XB - e, XK, MK)[0]) / (2 * h)
    got = np.array([F.reshape(-1, 4)[fine_at(*c)] for c, _, _ in target], np.float64)
    want = np.zeros_like(got)
    tidal = 0.0
    for i in range(len(target)):
        d = xt[i] - XB
        na, npot = near_sum(xt, mt, i, range(len(target)))
        want[i, :3] = (na + a0 + J @ d) * float(GM)
        want[i, 3] = (npot + p0 - a0 @ d - 0.5 * d @ J @ d) * float(GM)
        tidal = max(tidal, np.linalg.norm(J @ d) / np.linalg.norm(na + a0 + J @ d))
    return (gu.rel_err_vec(got[:, :3], want[:, :3]).max(), gu.rel_err(got[:, 3], want[:, 3]).max(), tidal)


def mesh_books(fine, F):
    """meshCheck's books (index.html:533-556): |net force| / sum |force| and
    |net torque| / sum |x| |force| over the occupied cells, forces m a."""
    f = fine[:512].reshape(-1, 4).astype(np.float64)
    a = F[:512].reshape(-1, 4).astype(np.float64)
    occ = f[:, 3] > 0
    m = f[occ, 3]
    x = f[occ, :3] / m[:, None]
    Fm = m[:, None] * a[occ, :3]
    tq = np.cross(x, Fm)
    mag = np.linalg.norm(Fm, axis=1)
    return (np.linalg.norm(Fm.sum(0)) / mag.sum(), np.linalg.norm(tq.sum(0)) / (mag * np.linalg.norm(x, axis=1)).sum())


def test_near_field_is_the_softened_direct_sum():
    far_zero, acc, pot = check_near()
    assert far_zero, "no block lies beyond the near region: the far field must vanish"
    assert acc < TOL and pot < TOL, (acc, pot)


def test_far_field_is_the_first_order_expansion_of_the_point_mass_field():
    acc, pot, tidal = check_far()
    assert acc < TOL and pot < TOL, (acc, pot)
    assert tidal > 100 * TOL, "the tidal term must matter at this tolerance"


@pytest.mark.parametrize("name", ["impact", "twobox"])
def test_mesh_books_net_force_and_torque(name):
    """The golden meshes' cells on each other: the near field is exactly
    antisymmetric and central; the far field's monopole-plus-tidal form leaves
    a residue of its truncation, which the bound below holds."""
    fx = gu.fixture("c1_" + name)
    ref = gu.mesh(fx, 0)
    _, _, F = field(ref["fine"])
    fr, tq = mesh_books(ref["fine"], F)
    assert fr < 2e-3 and tq < 2e-3, (fr, tq)


def test_each_planted_mistake_fails_a_check():
    subprocess.run(["make", "-s", "-C", os.path.join(gu.REPO, "oracle"), "mutants"], check=True)
    fx = gu.fixture("c1_impact")
    ref = gu.mesh(fx, 0)
    caught = {}
    for k in range(1, 10):
        with oracle.using(os.path.join(gu.REPO, "oracle", "_mut", "libcc_oracle_mut%d.so" % k)):
            far_zero, na, npot = check_near()
            fa, fp, _ = check_far()
            _, _, F = field(ref["fine"])
            fr, tq = mesh_books(ref["fine"], F)
        caught[k] = (not far_zero) or max(na, npot, fa, fp) >= TOL or fr >= 2e-3 or tq >= 2e-3
    assert all(caught.values()), caught
