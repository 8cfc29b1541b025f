"""The crater view's readout (cc_crater_stats, craterStats of
chicxulub.html:848-919) on states built so that every figure follows from
the construction: a settled surface with a floor, a rim and level ground
(the crater's diameter, rim, depth and the thrown mass from the 96 rings), a
frozen shell that must not count, a curtain moving at chosen speeds (airborne,
past orbit, past escape), melt and vapour by material and temperature, the
hot pool's centre, radius and light, and the books.  CPU only."""
import math
import os
import sys

import numpy as np
import pytest

import golden_util as gu

sys.path.insert(0, os.path.join(gu.REPO, "cosmic-collisions_amd"))
import ccamd  # noqa: E402

ARGS = dict(Rbox=1.0, lSI=1e4, vSI=1000.0, gSI=9.81, rhoSI=2500.0, vOrb=7.9, vEsc=11.2)
BIN = 1.0 / 96


def surface_state():
    """One settled particle per ring: floor -0.3 in rings 0..29, 0.12 at ring 30,
    0.1 in rings 31..34, level ground beyond but for a mound of 0.5 at ring 40;
    plus a frozen shell.  The readout walks the rings outwards and stops at the
    first ring past the crater that stands above -0.05, so the rim is ring 30
    and the higher mound beyond does not count as the rim."""
    P, V, R = [], [], []
    for b in range(96):
        r = (b + 0.5) * BIN
        y = -0.3 if b < 30 else 0.12 if b == 30 else 0.1 if b < 35 else 0.5 if b == 40 else 0.0
        P.append([r, y, 0.0, 1.0])
        V.append([0.0, 0.0, 0.0, 300.0])
        R.append(1.0)
    for k in range(5):   # frozen: far above, hot, fast -- and ignored
        P.append([0.01 * k, 5.0, 0.0, 0.0])
        V.append([20.0, 0.0, 0.0, 9000.0])
        R.append(100.0)
    return np.array(P, np.float32), np.array(V, np.float32), np.array(R, np.float32)


def test_surface_figures():
    P, V, R = surface_state()
    mref = 0.5
    d = ccamd.crater_stats(P, V, R, mref, **ARGS)
    assert d["rimH"] == float(np.float32(0.12))
    assert d["rimR"] == (30 + 0.5) * BIN
    assert d["depth"] == float(np.float32(0.3))
    assert d["craterD"] == 2 * 30 * BIN
    assert d["mass"] == 96 * mref                         # the frozen shell left out
    assert d["excavated"] == 6 * mref                     # rings 30..34 and the mound stand above the old ground
    assert d["thrown"] == 5 * mref                        # rings 31..34 and the mound, outside the rim's ring
    assert d["airborne"] == 0 and d["KE"] == 0 and d["drift"] == 0
    pe = sum(mref * ARGS["gSI"] * float(y) * ARGS["lSI"] for y in P[:96, 1])
    assert d["PE"] == pytest.approx(pe, rel=1e-12)
    assert d["Tmax"] == pytest.approx(300.0) and d["melt"] == 0 and d["vapour"] == 0
    assert d["R"] == 1e-3 and d["Tsurf"] == 0 and d["com"] == [0.0, 0.0, 0.0]


def test_curtain_melt_pool_and_books():
    mref, n = 2.0, 40
    P = np.zeros((n, 4), np.float32)
    V = np.zeros((n, 4), np.float32)
    R = np.ones(n, np.float32)
    P[:, 1] = 0.2                       # everything above the old ground
    P[:, 3] = 2.0                       # platform material
    speeds = np.array([0.1] * 10 + [0.5] * 10 + [9.0] * 10 + [12.0] * 10, np.float32)
    V[:, 0] = speeds
    V[:, 3] = 1000.0
    P[:5, 3] = 3.0                      # five melted by the shock
    V[5:8, 3] = 3000.0                  # three vaporised
    hot = list(range(5, 12))            # hot (T > 1200): seven, at known places and temperatures
    V[8:12, 3] = 1500.0
    P[hot, 0] = np.arange(len(hot), dtype=np.float32) * 0.1
    d = ccamd.crater_stats(P, V, R, mref, **ARGS)
    assert d["mass"] == n * mref and d["excavated"] == n * mref
    assert d["platform"] == (n - 5) * mref               # the melt (material 3) is not platform
    assert d["airborne"] == 30 * mref                    # faster than 0.15
    assert d["fastOrb"] == 20 * mref and d["fastEsc"] == 10 * mref
    assert d["melt"] == 5 * mref and d["vapour"] == 3 * mref and d["Tmax"] == 3000.0
    ke = sum(0.5 * mref * float(s) ** 2 * 1e6 for s in speeds)
    assert d["KE"] == pytest.approx(ke, rel=1e-12)
    assert d["drift"] == pytest.approx(float(np.sum(speeds.astype(np.float64))) * mref / (n * mref), rel=1e-12)
    hotM = len(hot) * mref
    assert d["com"][0] == pytest.approx(float(np.mean(P[hot, 0].astype(np.float64))), rel=1e-12)
    assert d["com"][1] == pytest.approx(0.2, rel=1e-7)
    Ts = np.array([3000.0] * 3 + [1500.0] * 4)
    assert d["Tsurf"] == pytest.approx(float(np.sum(Ts ** 3) / np.sum(Ts ** 2)), rel=1e-12)
    pool = (3 * (hotM / ARGS["rhoSI"]) / (4 * math.pi)) ** (1 / 3) / ARGS["lSI"]
    assert d["R"] == pytest.approx(max(pool, 1e-3), rel=1e-12)
    assert d["vel"] == [0.0, 0.0, 0.0]


def test_rejects_bad_arguments():
    P, V, R = surface_state()
    with pytest.raises(ccamd.CCError):
        ccamd.crater_stats(P, V, R, 1.0, **dict(ARGS, Rbox=0.0))
    with pytest.raises(ccamd.CCError):
        ccamd.crater_stats(P, V[:3], R, 1.0, **ARGS)


def test_far_and_nan_positions():
    """A grain thrown far past the box counts in the outermost ring (the page's
    Math.min(95, floor(r / binW)) with r / binW past any integer), and a grain
    whose place is NaN adds to the totals but to no ring."""
    P, V, R = surface_state()
    far = np.array([[1e30, 0.3, 0.0, 1.0], [np.nan, 0.3, 0.0, 1.0]], np.float32)
    still = np.array([[0.0, 0.0, 0.0, 300.0]] * 2, np.float32)
    base = ccamd.crater_stats(P, V, R, 0.5, **ARGS)
    d = ccamd.crater_stats(np.vstack([P, far]), np.vstack([V, still]), np.concatenate([R, [1.0, 1.0]]).astype(np.float32),
                           0.5, **ARGS)
    assert d["thrown"] == base["thrown"] + 0.5
    assert d["excavated"] == base["excavated"] + 1.0 and d["mass"] == base["mass"] + 1.0
    assert d["rimR"] == base["rimR"] and d["craterD"] == base["craterD"]


@pytest.mark.skipif(__import__("shutil").which("node") is None, reason="node is not installed")
def test_addon_matches_python_binding(tmp_path):
    """The N-API addon's craterStats (what the crater engine's analyze() calls)
    returns the same report as the ctypes binding on the same state."""
    import json
    import subprocess
    addon = os.path.join(gu.REPO, "cosmic-collisions_amd", "js", "cc_addon.node")
    if not os.path.exists(addon):
        pytest.skip("the N-API addon is not built")
    P, V, R = surface_state()
    P[40:50, 0] += 0.2          # some thrown, some moving
    V[50:60, 0] = 9.0
    V[60:65, 3] = 2600.0
    for a, f in ((P, "p"), (V, "v"), (R, "r")):
        a.tofile(tmp_path / f)
    script = tmp_path / "t.js"
    script.write_text("""
const addon = require('%s'); const fs = require('fs');
const rd = (f) => { const b = fs.readFileSync('%s/' + f); return new Float32Array(b.buffer.slice(b.byteOffset, b.byteOffset + b.length)); };
console.log(JSON.stringify(addon.craterStats(rd('p'), rd('v'), rd('r'), %d, 0.5, %s)));
""" % (addon, tmp_path, len(R), json.dumps(ARGS)))
    r = subprocess.run(["node", str(script)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    js = json.loads(r.stdout)
    d = ccamd.crater_stats(P, V, R, 0.5, **ARGS)
    assert set(js) == set(d)
    for k in d:
        assert js[k] == d[k], k
