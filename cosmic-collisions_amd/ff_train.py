"""Training the learned far field on the GPU from the library's own runs.

Samples are (coarse mesh -> exact far field) pairs recorded after the mesh
pass of every 8-step cycle of full-phase runs (runs.py); the network is
trained by the library's own BF16 matrix-core forward and backward kernels
with Adam (cc_ff_train).  The last quarter of each run's samples is held out.

    python ff_train.py --out weights.npy [--cycles 48] [--iters 3000]
"""
import argparse
import os
import sys
import tempfile
import time

import numpy as np

import ccamd
import runs


def collect(ff, runner, cycles):
    """cycles 8-step cycles of the runner with the exact far field, a sample
    per box that is up after each cycle's mesh pass; returns the indices."""
    first = ff.count
    for _ in range(cycles):
        runner.steps(8)
        ff.record(runner.ctx, 0)
        if runner.s["boxes"][1]["on"]:
            ff.record(runner.ctx, 1)
    return list(range(first, ff.count))


def train(ff, iters, batch=16, lr=2e-3, seed=1):
    """Adam with the step size stepped down by 4 for the last third."""
    a = (2 * iters) // 3
    l1 = ff.train(a, batch=batch, lr=lr, seed=seed)
    l2 = ff.train(iters - a, batch=batch, lr=lr / 4, seed=seed + 1)
    return np.concatenate([l1, l2])


def build(configs, state_root, cycles, iters, capacity=512, reuse=True, slope=0.25, log=print):
    """A FarField trained on the named configurations (runs.py): (ff, report)."""
    ff = ccamd.FarField(capacity=capacity)
    ff.set_slope(slope)
    t0 = time.perf_counter()
    train_set, held = [], []
    for name in configs:
        run, st = runs.prepare(getattr(runs, name), os.path.join(state_root, "cc_%s_state" % name.lower()), reuse=reuse)
        R = runs.Runner(run, st)
        idx = collect(ff, R, cycles)
        R.close()
        cut = len(idx) - len(idx) // 4
        train_set += [ff.sample(i) for i in idx[:cut]]
        held += [ff.sample(i) for i in idx[cut:]]
        log("ff samples: %s %d" % (name, len(idx)))
    t1 = time.perf_counter()
    ff.clear()   # the training samples first, the held-out after them once trained
    for x, t, m in train_set:
        ff.push(x, t, m)
    before = ff.eval()[0]
    losses = train(ff, iters)
    t2 = time.perf_counter()
    train_err = ff.eval()
    for x, t, m in held:
        ff.push(x, t, m)
    held_err = ff.eval(len(train_set), len(held)) if held else (float("nan"), float("nan"))
    rep = {"slope": slope, "samples": len(train_set), "held_out": len(held), "iters": iters, "collect_s": t1 - t0, "train_s": t2 - t1,
           "loss_first": float(losses[0]), "loss_last": float(np.mean(losses[-20:])),
           "train_rel_acc_err_init": before, "train_rel_acc_err": train_err[0], "train_rel_err_all": train_err[1],
           "held_rel_acc_err": held_err[0], "held_rel_err_all": held_err[1]}
    log("ff trained: %s" % rep)
    return ff, rep


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--configs", default="THEIA,TWINS,SHATTER")
    ap.add_argument("--cycles", type=int, default=48)
    ap.add_argument("--iters", type=int, default=3000)
    ap.add_argument("--slope", type=float, default=0.25, help="the leaky ReLU's negative slope (1: linear)")
    ap.add_argument("--state-root", default=tempfile.gettempdir())
    args = ap.parse_args()
    ff, rep = build(args.configs.split(","), args.state_root, args.cycles, args.iters, slope=args.slope,
                    log=lambda m: print(m, flush=True))
    np.save(args.out, ff.weights())
    ff.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
