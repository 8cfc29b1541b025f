"""Training the learned far field on the GPU from the library's own runs.

A run with the exact far field records a pass of each box that is up after
the mesh pass of every 8-step cycle (cc_ff_record: the coarse mesh's features
and the exact far field carried to the block centres).  A training sample
pairs two passes of one box as a mesh step would: the prediction from the
earlier pass (the base) to the later one, 8 steps on (a mesh step between
judges) or 16 (a judge step).  The runs are the named configurations (the
theia-like and twins impacts by default) and the instances of a what-if sweep
of their own (giant impacts of uneven size, mass ratio, angle and speed; the
sweep generator's next seed, so that the batched sweep a network is measured
on is none of them).
The network is trained by the library's own BF16 matrix-core forward and
backward kernels with Adam (cc_ff_train).  The last quarter of each run's
passes is held out.

The map from the coarse mesh's change to the far field's change is linear to
first order (the far field is a sum over blocks of mass times a fixed kernel of
their separation), and the network generalises best to an impact it has not
seen as a linear operator: the default negative slope of its activations is 1.

    python ff_train.py --out weights.npy [--configs THEIA,TWINS] [--cycles 250] [--sweep 64]
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
    """cycles 8-step cycles of the runner with the exact far field, a recorded
    pass per box that is up after each cycle's mesh pass; returns each box's
    passes in step order."""
    seqs = [[], []]
    for _ in range(cycles):
        runner.steps(8)
        seqs[0].append(ff.record(runner.ctx, 0))
        if runner.s["boxes"][1]["on"]:
            seqs[1].append(ff.record(runner.ctx, 1))
    return [q for q in seqs if len(q) > 2]


def pair_all(ff, seq):
    """Samples of one box's passes: each pass as the base of the next (8
    steps) and of the one after (16 steps); returns their indices."""
    out = []
    for i in range(len(seq) - 1):
        out.append(ff.pair(seq[i], seq[i + 1]))
        if i + 2 < len(seq):
            out.append(ff.pair(seq[i], seq[i + 2]))
    return out


def train(ff, iters, batch=16, lr=2e-3, seed=1):
    """Adam with the step size stepped down by 4 for the last third."""
    a = (2 * iters) // 3
    l1 = ff.train(a, batch=batch, lr=lr, seed=seed)
    l2 = ff.train(iters - a, batch=batch, lr=lr / 4, seed=seed + 1)
    return np.concatenate([l1, l2])


def capacity(configs, cycles, sweep, sweep_cycles):
    """Passes and samples the runs need (two boxes at most, two samples a pass)."""
    return 2 * 2 * (len(configs) * cycles + sweep * sweep_cycles) + 16


SWEEP_SEED = runs.BATCH_SEED + 1   # the training sweep: not the batched sweep's own seed


def collect_sweep(ff, batch_dir, count, cycles, seed=SWEEP_SEED, reuse=True, log=print):
    """A sweep's instances stepped together (runs.BatchRunner), a recorded
    pass per instance and box after each cycle's mesh pass."""
    B = runs.BatchRunner(runs.prepare_batch(batch_dir, count, seed=seed, reuse=reuse))
    seqs = [[[], []] for _ in B.rs]
    for _ in range(cycles):
        B.steps(8)
        for k, r in enumerate(B.rs):
            seqs[k][0].append(ff.record(r.ctx, 0))
            if r.s["boxes"][1]["on"]:
                seqs[k][1].append(ff.record(r.ctx, 1))
    B.close()
    out = [q for pair in seqs for q in pair if len(q) > 2]
    log("ff passes: sweep of %d, %d sequences" % (count, len(out)))
    return out


def build(configs, state_root, cycles, iters, reuse=True, slope=1.0, sweep=0, sweep_cycles=40, batch_dir=None,
          lr=2e-3, sweep_seed=SWEEP_SEED, log=print):
    """A FarField trained on the named configurations (runs.py) and, with
    sweep > 0, `sweep` instances of the sweep of seed `sweep_seed`: (ff, report).
    Training samples come first in the record, held-out ones after them."""
    ff = ccamd.FarField(capacity=capacity(configs, cycles, sweep, sweep_cycles))
    ff.set_slope(slope)
    t0 = time.perf_counter()
    seqs = []
    for name in configs:
        run, st = runs.prepare(getattr(runs, name), os.path.join(state_root, "cc_%s_state" % name.lower()), reuse=reuse)
        R = runs.Runner(run, st)
        got = collect(ff, R, cycles)
        R.close()
        log("ff passes: %s %s" % (name, [len(q) for q in got]))
        seqs += got
    if sweep > 0:
        seqs += collect_sweep(ff, batch_dir or os.path.join(state_root, "cc_train_sweep_state"), sweep, sweep_cycles,
                              seed=sweep_seed, reuse=reuse, log=log)
    # each box's passes: the first three quarters train, the last quarter is held out
    cut = [len(q) - len(q) // 4 for q in seqs]
    first = ff.count
    for q, c in zip(seqs, cut):
        pair_all(ff, q[:c])
    n_train = ff.count - first
    for q, c in zip(seqs, cut):
        pair_all(ff, q[c:])
    n_held = ff.count - first - n_train
    t1 = time.perf_counter()
    before = ff.eval(first, n_train)[0]
    losses = train(ff, iters, lr=lr)
    t2 = time.perf_counter()
    train_err = ff.eval(first, n_train)
    held_err = ff.eval(first + n_train, n_held) if n_held else (float("nan"), float("nan"))
    rep = {"configs": list(configs), "sweep_instances": sweep, "sweep_seed": sweep_seed, "slope": slope,
           "passes": ff.passes,
           "samples": n_train, "held_out": n_held, "iters": iters, "collect_s": t1 - t0, "train_s": t2 - t1,
           "loss_first": float(losses[0]), "loss_last": float(np.mean(losses[-20:])),
           "what": "relative L2 error of the predicted change of the blocks' far-field acceleration since the "
                   "base, each block weighted by its occupied fine texels (1.0: no prediction, the base kept)",
           "train_rel_change_err_init": before, "train_rel_change_err": train_err[0],
           "train_rel_change_err_all": train_err[1], "held_rel_change_err": held_err[0],
           "held_rel_change_err_all": held_err[1]}
    log("ff trained: %s" % rep)
    return ff, rep


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--configs", default="THEIA,TWINS")
    ap.add_argument("--cycles", type=int, default=250)
    ap.add_argument("--sweep", type=int, default=64, help="instances of the batched sweep to record (0: none)")
    ap.add_argument("--sweep-cycles", type=int, default=40)
    ap.add_argument("--iters", type=int, default=6000)
    ap.add_argument("--slope", type=float, default=1.0, help="the activations' negative slope (1: linear)")
    ap.add_argument("--state-root", default=tempfile.gettempdir())
    args = ap.parse_args()
    ff, rep = build(args.configs.split(","), args.state_root, args.cycles, args.iters, slope=args.slope,
                    sweep=args.sweep, sweep_cycles=args.sweep_cycles, log=lambda m: print(m, flush=True))
    np.save(args.out, ff.weights())
    ff.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
