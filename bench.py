#!/usr/bin/env python3
"""Benchmark of the gravity path on one MI355X.

Workload (the configuration the headline metric is quoted on): Plummer sphere
N=1,048,576 (ic::Plummer<3>, seed 0, M=1, a=1, G=1), FP32 positions/velocities, Plummer
softening eps=0.005, Barnes-Hut theta=0.5 with quadrupoles, kick-drift-kick leapfrog
dt=5e-4.  A step = one KDK step: tree rebuilt from scratch (keys, sort, build, upsweep),
group walk, two kicks and a drift, replayed from a captured hipGraph.

Prints one JSON line.  `value` is seconds per step (lower is better); `roofline` is the
walk kernel's algorithmic FP32 rate (20 flop per P2P, 65 per quadrupole M2P, counted from
the actual interaction lists) over its HIP-event time inside the timed steps, against the
MI355X FP32 peak; `cpu_baseline` times the FP64 oracle (the reference CPU path restated,
OpenMP) on a bounded sample of the same workload on this host's cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "opengl-physics_amd"))

PEAK_FP32_TFLOPS = 157.3
FLOP_P2P = 20.0
FLOP_M2P_QUAD = 65.0
FLOP_M2P_MONO = 20.0


def walk_traffic(n, theta):
    """HBM bytes per walk launch from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE passes of
    this workload (scaled by a copy kernel of known size), or None for other workloads."""
    path = os.path.join(ROOT, "profiles", "walk_traffic_plummer_1m_theta0.5.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return d["bytes_per_launch"] if (d["n"] == n and d["theta"] == theta) else None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(state32, eps, theta, dt, multipole):
    """One KDK step of the same workload on the FP64 oracle (reference flags build)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    st = {k: np.asarray(v, np.float64) for k, v in state32.items()}
    quad = multipole == "quadrupole"
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    t0 = time.perf_counter()
    T = oracle.Tree(st["x"], st["y"], st["z"], st["m"], quad=quad, fast=True)
    a0 = T.accel_bh(1.0, eps * eps, theta)
    del T
    t1 = time.perf_counter()
    oracle.leapfrog(st, a0, 1.0, eps * eps, theta, dt, 1, quad=quad, fast=True)
    t2 = time.perf_counter()
    log(f"cpu baseline: prime {t1 - t0:.2f} s, step {t2 - t1:.2f} s on {threads} threads")
    return {"value": round(t2 - t1, 4), "unit": "s/step", "cores": threads, "kind": "port",
            "sample": f"one KDK leapfrog step of the full workload (N={len(st['x'])}), FP64 oracle "
                      f"(serial insertion tree + quadrupoles + OpenMP group walk), after one untimed force "
                      f"evaluation; built -O3 -march=x86-64-v3 -funroll-loops -fopenmp"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--theta", type=float, default=0.5)
    ap.add_argument("--eps", type=float, default=0.005)
    ap.add_argument("--dt", type=float, default=5e-4)
    ap.add_argument("--multipole", default="quadrupole", choices=["monopole", "quadrupole"])
    ap.add_argument("--no-kahan", action="store_true")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the multi-GPU decomposition is "
              f"design-only in this build", file=sys.stderr)
        sys.exit(2)

    import pynbg
    n = args.n
    log(f"generating Plummer N={n} seed 0")
    ic = pynbg.to_f32(pynbg.ic_plummer(n, 0))
    kahan = not args.no_kahan
    p = pynbg.make_params("barnes_hut", args.multipole, G=1.0, eps=args.eps, theta=args.theta, dt=args.dt,
                          kahan=kahan, stage_times=True)
    pc = pynbg.make_params("barnes_hut", args.multipole, G=1.0, eps=args.eps, theta=args.theta, dt=args.dt,
                           kahan=kahan, count=True)
    with pynbg.Context(n) as ctx:
        ctx.upload(ic["x"], ic["y"], ic["z"], ic["m"], ic["vx"], ic["vy"], ic["vz"])
        st0 = ctx.accel(pc)  # System::Prime, with interaction counting
        log(f"prime: nodes={st0.n_nodes} leaves={st0.n_leaves} depth={st0.max_depth} "
            f"p2p={st0.n_p2p} m2p={st0.n_m2p} walk={st0.ms_walk:.3f} ms")
        if args.warmup > 0:
            ctx.step(p, args.warmup)
        # timed region: the step call synchronises the device before it returns
        t0 = time.perf_counter()
        st = ctx.step(p, args.steps)
        t1 = time.perf_counter()
        state = ctx.download()
        st1 = ctx.accel(pc)  # interaction counts at the end state
        wall = t1 - t0
    ms_step = 1e3 * wall / args.steps
    n_p2p = 0.5 * (st0.n_p2p + st1.n_p2p)
    n_m2p = 0.5 * (st0.n_m2p + st1.n_m2p)
    fm2p = FLOP_M2P_QUAD if args.multipole == "quadrupole" else FLOP_M2P_MONO
    flops = FLOP_P2P * n_p2p + fm2p * n_m2p
    walk_ms = st.ms_walk / args.steps
    achieved = flops / (walk_ms * 1e-3) / 1e12
    stage = {k: round(getattr(st, "ms_" + k) / args.steps, 4)
             for k in ("keys", "sort", "build", "upsweep", "walk", "integrate")}
    log(f"steps={args.steps} wall={wall:.4f} s ({ms_step:.3f} ms/step) stages(ms/step)={stage}")
    out = {
        "metric": "s/step, 1M-body BH: plain θ=0.5 vs learned θ=0.8 at equal RMS force error",
        "value": round(wall / args.steps, 6),
        "unit": "s/step",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (ic::Plummer<3> seed 0, generated on the host)",
        "config": {"workload": "plummer_1m_barnes_hut_theta0.5" if args.theta == 0.5 else
                   f"plummer_bh_theta{args.theta}", "n": n, "theta": args.theta, "eps": args.eps, "dt": args.dt,
                   "multipole": args.multipole, "kahan": kahan, "variant": "plain"},
        "roofline": {"bound": "mfma", "pipe": "fp32-valu", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
                     "traffic": walk_traffic(n, args.theta),
                     "kernel": "k_walk_bins", "kernel_ms": round(walk_ms, 4),
                     "flop_per_launch": flops},
        "interactions_per_s": round((n_p2p + n_m2p) / (walk_ms * 1e-3), 1),
        "interactions_per_step": {"p2p": n_p2p, "m2p": n_m2p},
        "stage_ms_per_step": stage,
        "tree": {"nodes": st.n_nodes, "leaves": st.n_leaves, "max_depth": st.max_depth},
    }
    if not np.all(np.isfinite(state["ax"])):
        raise RuntimeError("non-finite accelerations")
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(ic, args.eps, args.theta, args.dt, args.multipole)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
