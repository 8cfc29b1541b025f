"""Synthetic 262,144-particle twins states for the full-size GPU tests: the
sim scalars build() derives for the twins configuration and two jittered
lattice bodies touching at 30 degrees.  Test infrastructure."""
import math

import numpy as np


def twins_scalars(N=262144, q=1.0, dens=0.75, spin0=0.18, spin1=0.20, vfac=1.05, rest=0.3):
    """The sim scalars build() derives for the twins config (src/nbody.js:114-206)."""
    V1, V2 = 1.0, q / dens
    vp = (V1 + V2) / N
    s = ((16 * math.pi * vp / 3) ** (1.0 / 3.0))
    a = 0.96 * s / (2 * math.sqrt(2))
    n1 = round(N * V1 / (V1 + V2))
    M = 2.0                      # mass ratio 1: both bodies one Earth mass in the reference's units
    mref = 1.0 / (n1 * (0.325 / (2.5 - 0.325 * 1.5) * 2.5 + (1 - 0.325 / (2.5 - 0.325 * 1.5))))
    Rsum = 1.0 + ((V2) ** (1.0 / 3.0))
    vesc = math.sqrt(2 * M / Rsum)
    vc = vfac * vesc
    vref = max(1.5, vc)
    k = 4 * mref * vref * vref / (a * a)
    dt = 0.35 / math.sqrt(2 * k / (mref * min(1, dens)))
    lnE = math.log(rest)
    zeta = -lnE / math.sqrt(math.pi ** 2 + lnE ** 2)
    c = 2 * zeta * math.sqrt(k * mref / 2)
    return dict(N=N, W=512, G=128, s=s, a=a, n1=n1, m=mref, cell=2.6 * a, dt=dt, k=k, c=c, vc=vc,
                R1=1.0, R2=((V2) ** (1.0 / 3.0)), spin0=(0, spin0, 0), spin1=(0, spin1 * math.sqrt(dens), 0))


def lattice_ball(n, s, rng):
    """n points of a jittered FCC lattice of cube edge s, nearest to the centre."""
    r = (3 * n * s ** 3 / (16 * math.pi)) ** (1 / 3) * 1.1
    m = int(math.ceil(r / s)) + 1
    g = np.arange(-m, m + 1) * s
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    base = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    offs = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]]) * s
    pts = (base[:, None, :] + offs[None]).reshape(-1, 3)
    d = np.linalg.norm(pts, axis=1)
    pts = pts[np.argsort(d, kind="stable")[:n]]
    return pts + rng.uniform(-0.03, 0.03, pts.shape) * s


def synthetic_twins(sc, seed=0xC0FFEE):
    """Two lattice bodies of the twins' sizes, touching at 30 degrees, closing at
    the contact speed -- the configuration a step sees just after the kick."""
    rng = np.random.default_rng(seed)
    N, n1 = sc["N"], sc["n1"]
    b1 = lattice_ball(n1, sc["s"], rng)
    b2 = lattice_ball(N - n1, sc["s"], rng)
    ang = math.radians(30)
    dist = sc["R1"] + sc["R2"] - 0.05
    off = np.array([-dist * math.cos(ang), dist * math.sin(ang), 0.0])
    c0 = -off * 0.5
    c1 = off * 0.5
    pos = np.zeros((N, 4), np.float32)
    pos[:n1, :3] = b1 + c0
    pos[n1:, :3] = b2 + c1
    r1 = np.linalg.norm(b1, axis=1)
    r2 = np.linalg.norm(b2, axis=1)
    pos[:n1, 3] = np.where(r1 < 0.66 * sc["R1"], 0, np.where(r1 > 0.97 * sc["R1"], 2, 1))
    pos[n1:, 3] = 4 + np.where(r2 < 0.66 * sc["R2"], 0, np.where(r2 > 0.97 * sc["R2"], 2, 1))
    mat = np.where(pos[:, 3] % 4 < 0.5, 2.5, 1.0).astype(np.float32)
    mat[n1:] *= 0.75
    vel = np.zeros((N, 4), np.float32)
    vrel = np.array([sc["vc"], 0, 0]) * 0.5
    vel[:n1, :3] = vrel
    vel[n1:, :3] = -vrel
    w0, w1 = np.array(sc["spin0"]), np.array(sc["spin1"])
    vel[:n1, :3] += np.cross(w0, pos[:n1, :3] - c0)
    vel[n1:, :3] += np.cross(w1, pos[n1:, :3] - c1)
    return pos, vel, mat, np.zeros(N, np.float32), c0, c1
