"""Training driver of the learned far-field correction.

Everything numerical runs on the GPU through the C ABI of libnbg_hip.so:
  * nbg_learn_collect   -- direct-sum snapshot, Barnes-Hut walk listing its accepted
                           (leaf, cell) pairs, exact residual of each pair (FP64);
  * nbg_learn_standardize, nbg_learn_train -- BF16 matrix-core forward/backward and Adam.
This module only chooses the snapshots and the schedule and writes the model blob.

    python -m pynbg.learned --out opengl-physics_amd/learned/farfield_mlp.npy

The default run follows the training recipe of the learned configuration: Plummer and
Hernquist spheres of N = 131,072 (total mass 1, scale radius 1, G = 1), eps = 0.005,
theta = 0.8, pairs in the band size / (distance - group radius) > 0.6.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from . import (LEARN_ACT, LEARN_ACT_SLOT, LEARN_FEATURES, LEARN_HIDDEN, LEARN_NPARAM, LEARN_NTRAIN, Context,
               ic_hernquist, ic_plummer, make_params, to_f32)

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEFAULT_WEIGHTS = os.path.join(_PKG, "learned", "farfield_mlp.npy")


def init_blob(seed=0, act=LEARN_ACT["silu"]):
    """He-style initialisation of the 13-64-64-64-3 network (output layer scaled by 0.1);
    standardisation entries start as mean 0, 1/std 1; blob[LEARN_ACT_SLOT] = the
    activation code (0 SiLU, 1 ReLU)."""
    rng = np.random.default_rng(seed)
    blob = np.zeros(LEARN_NPARAM, np.float32)
    blob[16:32] = 1.0
    blob[LEARN_ACT_SLOT] = act
    sizes = [("W1", LEARN_FEATURES * LEARN_HIDDEN, LEARN_FEATURES), ("b1", LEARN_HIDDEN, 0),
             ("W2", LEARN_HIDDEN * LEARN_HIDDEN, LEARN_HIDDEN), ("b2", LEARN_HIDDEN, 0),
             ("W3", LEARN_HIDDEN * LEARN_HIDDEN, LEARN_HIDDEN), ("b3", LEARN_HIDDEN, 0),
             ("W4", LEARN_HIDDEN * 3, LEARN_HIDDEN), ("b4", 3, 0)]
    o = 32
    for name, size, fan_in in sizes:
        if fan_in:
            scale = np.sqrt(2.0 / fan_in) * (0.1 if name == "W4" else 1.0)
            blob[o:o + size] = rng.standard_normal(size) * scale
        o += size
    assert o == 32 + LEARN_NTRAIN
    return blob


def load_weights(path=DEFAULT_WEIGHTS):
    blob = np.load(path, allow_pickle=False).astype(np.float32)
    if blob.shape != (LEARN_NPARAM,):
        raise ValueError(f"{path}: expected {LEARN_NPARAM} floats, found {blob.shape}")
    return blob


def snapshot(kind, n, seed):
    ic = ic_plummer(n, seed) if kind == "plummer" else ic_hernquist(n, seed)
    return to_f32(ic)


def collect(ctx, snapshots, n, eps, theta, band, max_per_snapshot=0, log=print, amin=0.0):
    p = make_params("barnes_hut", "quadrupole", G=1.0, eps=eps, theta=theta, learn_band=band, learn_amin=amin)
    total = 0
    for kind, seed in snapshots:
        s = snapshot(kind, n, seed)
        ctx.upload(s["x"], s["y"], s["z"], s["m"])
        t0 = time.time()
        total = ctx.learn_collect(p, max_per_snapshot)
        log(f"collect {kind} seed {seed}: pool {total} ({time.time() - t0:.2f} s)")
    return total


def train(ctx, steps, batch, lr, seed=0, chunks=20, log=print):
    """Adam with a cosine learning-rate schedule applied per chunk of steps."""
    losses = []
    per = max(1, steps // chunks)
    done = 0
    while done < steps:
        k = min(per, steps - done)
        lr_k = 0.5 * lr * (1.0 + np.cos(np.pi * done / steps))
        losses.append(ctx.learn_train(k, batch, lr_k, seed + done))
        done += k
        log(f"train step {done}/{steps} lr {lr_k:.2e} loss {float(np.mean(losses[-1][-min(50, k):])):.5f}")
    return np.concatenate(losses)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    ap.add_argument("--n", type=int, default=131072)
    ap.add_argument("--plummer-seeds", type=int, nargs="*", default=[100, 101, 102, 103, 104, 105])
    ap.add_argument("--hernquist-seeds", type=int, nargs="*", default=[100, 101, 102, 103, 104, 105])
    ap.add_argument("--eps", type=float, default=0.005)
    ap.add_argument("--theta", type=float, default=0.8)
    ap.add_argument("--band", type=float, default=0.6)
    ap.add_argument("--amin", type=float, default=0.0,
                    help="pairs also need a correction scale G M s^3 / r^5 >= this (the evaluation's learn_amin)")
    ap.add_argument("--max-per-snapshot", type=int, default=0)
    ap.add_argument("--steps", type=int, default=100000)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--act", choices=sorted(LEARN_ACT), default="relu",
                    help="hidden-layer activation (ReLU: one packed instruction per two values in inference)")
    a = ap.parse_args(argv)
    log = lambda s: print(s, flush=True)
    snaps = [("plummer", s) for s in a.plummer_seeds] + [("hernquist", s) for s in a.hernquist_seeds]
    t0 = time.time()
    with Context(a.n) as ctx:
        ctx.learn_set_weights(init_blob(a.seed, LEARN_ACT[a.act]))
        pool = collect(ctx, snaps, a.n, a.eps, a.theta, a.band, a.max_per_snapshot, log, amin=a.amin)
        ctx.learn_standardize()
        t1 = time.time()
        loss = train(ctx, a.steps, a.batch, a.lr, a.seed, log=log)
        t2 = time.time()
        blob = ctx.learn_get_weights()
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    np.save(a.out, blob.astype(np.float32), allow_pickle=False)
    meta = dict(samples=int(pool), snapshots=[f"{k}:{s}" for k, s in snaps], n=a.n, eps=a.eps, theta=a.theta,
                band=a.band, amin=a.amin, activation=a.act, steps=a.steps, batch=a.batch, lr=a.lr, seed=a.seed,
                loss_first=float(np.mean(loss[:50])), loss_last=float(np.mean(loss[-200:])),
                collect_s=round(t1 - t0, 2), train_s=round(t2 - t1, 2))
    with open(os.path.splitext(a.out)[0] + ".json", "w") as f:
        json.dump(meta, f, indent=1)
    log(json.dumps(meta))
    return 0


if __name__ == "__main__":
    sys.exit(main())
