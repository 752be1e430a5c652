#!/usr/bin/env python3
"""Benchmark of the gravity path on one MI355X.

Workload (the configuration the headline metric is quoted on): Plummer sphere
N=1,048,576 (ic::Plummer<3>, seed 0, M=1, a=1, G=1), FP32 positions/velocities, Plummer
softening eps=0.005, quadrupoles, kick-drift-kick leapfrog dt=5e-4.  A step = one KDK
step: tree rebuilt from scratch (keys, sort, build, upsweep), group walk, two kicks and a
drift, replayed from a captured hipGraph.  Two variants are timed on the same state:
  plain    Barnes-Hut theta=0.5
  learned  Barnes-Hut theta=0.8 plus the learned far-field correction on the accepted
           (leaf, cell) pairs whose correction scale G M s^3 / r^5 is at least 1e-3
           (about 0.2% of a typical |a| of this sphere)
Each variant's force error is its relative RMS against the direct sum on 16,384 random
targets of the initial state.  Two more plain steps are timed for the comparison at equal
error: plain Barnes-Hut at the opening angle whose error matches the learned step's
(bisected), and plain Barnes-Hut at the learned step's own theta=0.8 (what the network
costs and what it buys).

Prints one JSON line.  `value` is seconds per step of the learned variant (lower is
better); `comparison` holds the variants' s/step and errors and the speedups, against
plain theta=0.5 and against plain Barnes-Hut at equal error.  `stage_hbm` gives the HBM
rate of the tree and integration stages against their algorithmic bytes.  `roofline`
is the dominant kernel's algorithmic rate over its HIP-event time inside the timed steps:
the walk against the FP32 peak (20 flop per P2P, 65 per quadrupole M2P, counted from the
actual interaction lists) or the MLP against the dense BF16 MFMA peak (18,432 flop per
pair).  `cpu_baseline` times the FP64 oracle (the reference CPU path restated, OpenMP,
plain theta=0.5), built with the reference's flags on this host, on a bounded sample of
the same workload on this host's physical cores (model, sockets, cores and threads
recorded).
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
PEAK_BF16_TFLOPS = 2500.0   # dense MFMA
FLOP_P2P = 20.0
FLOP_M2P_QUAD = 65.0
FLOP_M2P_MONO = 20.0
FLOP_MLP_PAIR = 2.0 * (13 * 64 + 64 * 64 + 64 * 64 + 64 * 3)   # 18,432 multiply-add flop per pair


def kernel_traffic(n, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    passes of the 1M Plummer workload (scaled by a copy kernel of known size), or None."""
    path = os.path.join(ROOT, "profiles", "traffic_plummer_1m.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return d["bytes_per_launch"].get(kernel) if d["n"] == n else None


PEAK_HBM_GBS = 8000.0


def sort_passes(levels):
    """LSD radix passes of 8-bit digits over a key of 3 bits per tree level (16 levels: 6)."""
    return (3 * levels + 7) // 8


def stage_hbm(st, steps, n, moments=False):
    """Per-stage HBM rate of the tree and integration stages from their HIP-event times
    inside the timed steps, against the algorithmic bytes of each (per particle unless
    noted; C = nodes): keys 28 (read x, y, z, m; write the 8-byte key and the 4-byte
    index), radix sort 24 per pass (a 64-bit key and a 32-bit value read and written) x the
    passes of the tree's key levels (5 for 13 levels, 6 for 16, 8 for 21), build 8 N + 40 C
    (keys in, node records out), upsweep 32 N + 56 C (particle records in and gathered into
    sorted order; centre of mass, mass, quadrupole, geometry and radius out), kick-drift-kick
    108.  counter_bytes: the same
    stage's HBM traffic per step from the committed FETCH_SIZE / WRITE_SIZE passes."""
    C = float(st.n_nodes)
    levels = st.key_levels
    algo = {"keys": 28.0 * n, "sort": 24.0 * sort_passes(levels) * n, "build": 8.0 * n + 40.0 * C,
            "upsweep": 32.0 * n + 56.0 * C, "integrate": 108.0 * n}
    counted = stage_counter_bytes(n, moments, levels)
    out = {}
    for k, b in algo.items():
        ms = getattr(st, "ms_" + k) / steps
        gbs = b / (ms * 1e-3) / 1e9 if ms > 0 else None
        out[k] = {"ms": round(ms, 4), "algorithmic_bytes": int(b),
                  "GB_s": round(gbs, 1) if gbs else None,
                  "frac_hbm": round(gbs / PEAK_HBM_GBS, 4) if gbs else None,
                  "counter_bytes": counted.get(k)}
    return out


# the kernels of each tree / integration stage and their launches per step (short names
# as profiles/traffic_plummer_1m.json keys them) for keys of L levels; the upsweep's
# internal nodes sit at depths 0..L-1: two depths per launch from the deepest down to depth
# 3 (depth 3 alone when left over), then depths 0..2 in one launch
def stage_kernels(levels, moments):
    P = sort_passes(levels)
    m = "true" if moments else "false"
    top = (levels if levels <= 21 else 40) - 1
    pairs, single = (top - 2) // 2, (top - 2) % 2
    return {
        "keys": [("void k_bbox<float>", 1), ("void k_root<float>", 1), (f"void k_keys<{levels}, float>", 1)],
        "sort": [("k_radix_hist", P), ("void k_scan_reduce<unsigned int>", P),
                 ("void k_scan_down_inline<unsigned int>", P), ("void k_radix_scatter<false>", P)],
        "build": [(f"void k_leaf_depth<{levels}>", 1), ("void k_scan_down_inline<unsigned long>", 1),
                  (f"void k_nodes<{levels}>", 1),
                  ("k_link", 1), (f"void k_node_finish<{levels}>", 1), ("void k_scan_reduce<unsigned int>", 1),
                  ("void k_scan_down_inline<unsigned int>", 1), ("k_depth_scatter", 1)],
        "upsweep": [(f"void k_leaves<true, {m}, false>", 1), (f"void k_up_pair<true, {m}, false>", pairs),
                    (f"void k_up_top<true, {m}, false>", 1)]
                   + ([(f"void k_up_internal<true, {m}, false>", 1)] if single else []),
        "integrate": [("k_kick_drift", 1), ("k_kick", 1)],
    }


def stage_counter_bytes(n, moments, levels):
    """HBM bytes per step of each stage from the committed counter passes, or {}."""
    path = os.path.join(ROOT, "profiles", "traffic_plummer_1m.json")
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        d = json.load(f)
    if d.get("n") != n:
        return {}
    t = d["bytes_per_launch"]
    out = {}
    for stage, ks in stage_kernels(levels, moments).items():
        if all(k in t for k, _ in ks):
            out[stage] = int(sum(t[k] * c for k, c in ks))
    return out


def _short_kernel_name(name):
    """'void nbg::k_walk_bins<true, ...>(nbg::BinArgs)' -> 'void k_walk_bins<true, ...>' (the
    template arguments kept, namespace and parameters dropped)."""
    name = name.replace("nbg::", "")
    depth, cut = 0, len(name)
    for i, ch in enumerate(name):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            cut = i
            break
    return name[:cut].strip()


def traffic_from_counters(fetch_csv, write_csv, out_path, n, workload, plain_csvs=None):
    """HBM bytes per launch of every kernel from two rocprofv3 --pmc passes over the same
    command (FETCH_SIZE, then WRITE_SIZE; counter_collection.csv of --output-format csv),
    each counter in KiB per dispatch.  The counters are calibrated on k_copy_f32, a plain
    copy of n floats (4 n bytes read, 4 n written): scale = algorithmic bytes / counter
    bytes of that copy, applied to every kernel.  plain_csvs: the same two passes over the
    plain theta=0.5 command, whose plain walk kernels replace those of the first pair (there
    the same kernel also runs the bisection probes and theta=0.8)."""
    import csv
    import collections

    def per_kernel(path, counter):
        per_disp = collections.defaultdict(float)
        names = {}
        with open(path) as f:
            for row in csv.DictReader(f):
                if row["Counter_Name"] != counter:
                    continue
                d = row["Dispatch_Id"]
                per_disp[d] += float(row["Counter_Value"])
                names[d] = _short_kernel_name(row["Kernel_Name"])
        acc = collections.defaultdict(list)
        for d, v in per_disp.items():
            acc[names[d]].append(v)
        return {k: sum(v) / len(v) for k, v in acc.items()}

    fetch = per_kernel(fetch_csv, "FETCH_SIZE")
    write = per_kernel(write_csv, "WRITE_SIZE")
    copy = "k_copy_f32"
    fs = 4.0 * n / (fetch[copy] * 1024.0)
    ws = 4.0 * n / (write[copy] * 1024.0)
    if plain_csvs:
        pf = per_kernel(plain_csvs[0], "FETCH_SIZE")
        pw = per_kernel(plain_csvs[1], "WRITE_SIZE")
        pfs = 4.0 * n / (pf[copy] * 1024.0)
        pws = 4.0 * n / (pw[copy] * 1024.0)
        for k in pf:
            if k.startswith("void k_walk_bins<") and ", false, 16, false, false>" in k:
                # rescaled to the first pair's calibration
                fetch[k] = pf[k] * pfs / fs
                write[k] = pw.get(k, 0.0) * pws / ws
    names = sorted(set(fetch) | set(write), key=lambda k: -(fetch.get(k, 0) * fs + write.get(k, 0) * ws))
    table = {k: int(1024.0 * (fetch.get(k, 0.0) * fs + write.get(k, 0.0) * ws)) for k in names}
    out = {"workload": workload, "n": n, "fetch_scale": round(fs, 4), "write_scale": round(ws, 4),
           "scale_meaning": "bytes = counter KiB * 1024 * scale; scale = algorithmic bytes of the copy / its counter bytes",
           "calibration": f"k_copy_f32 of {n:,} floats",
           "fetch_bytes_per_launch": {k: int(1024.0 * fetch.get(k, 0.0) * fs) for k in names},
           "write_bytes_per_launch": {k: int(1024.0 * write.get(k, 0.0) * ws) for k in names},
           "bytes_per_launch": table}
    with open(out_path, "w") as f:
        json.dump(out, f, indent=1)
    return out


def walk_kernel_name(quad, kahan, learn, near_tile, spline=False):
    b = lambda v: "true" if v else "false"
    return f"void k_walk_bins<{b(quad)}, {b(kahan)}, false, {b(learn)}, {near_tile}, {b(spline)}, false>"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


CPU_SHARE = 16  # host threads one GPU's job may use on the benchmark host


def host_cpu():
    """The host's CPU as lscpu reports it, restricted to the CPUs this process may run on:
    model, sockets, logical CPUs, physical cores, and the OpenMP threads the CPU baseline
    uses (one per physical core, at most the per-GPU share)."""
    import subprocess
    allowed = sorted(os.sched_getaffinity(0))
    info = {"model": None, "sockets": None, "logical_cpus": len(allowed), "physical_cores": None}
    try:
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, check=True).stdout
        for line in txt.splitlines():
            if line.startswith("Model name:"):
                info["model"] = line.split(":", 1)[1].strip()
        par = subprocess.run(["lscpu", "-p=CPU,CORE,SOCKET"], capture_output=True, text=True, check=True).stdout
        cores, sockets = set(), set()
        for line in par.splitlines():
            if line.startswith("#"):
                continue
            cpu, core, sock = (int(v) for v in line.split(",")[:3])
            if cpu in allowed:
                cores.add((sock, core))
                sockets.add(sock)
        info["physical_cores"] = len(cores)
        info["sockets"] = len(sockets)
    except (OSError, ValueError, subprocess.CalledProcessError):
        pass
    phys = info["physical_cores"] or len(allowed)
    info["threads_used"] = min(phys, CPU_SHARE)
    return info


def oracle_for_baseline():
    """The FP64 oracle built with the reference's flags (-O3 -march=native -funroll-loops
    -fopenmp, projects/nbody_core/CMakeLists.txt:36-38), compiled on this host so that
    -march=native means this CPU; the prebuilt x86-64-v3 build if no compiler is here.
    OpenMP placement is fixed before the library (and its OpenMP runtime) loads:
    OMP_PROC_BIND=spread, OMP_PLACES=cores, one thread per physical core
    (scripts/perlmutter_nbody_bench_cpu.sbatch:28-31)."""
    import subprocess
    cpu = host_cpu()
    os.environ["OMP_NUM_THREADS"] = str(cpu["threads_used"])
    os.environ["OMP_PROC_BIND"] = "spread"
    os.environ["OMP_PLACES"] = "cores"
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    flavour, flags = True, "-O3 -march=x86-64-v3 -funroll-loops -fopenmp (prebuilt)"
    try:
        subprocess.run(["make", "-s", "-B", "-C", os.path.join(ROOT, "oracle"), "native"], check=True,
                       capture_output=True, timeout=300)
        flavour, flags = "native", "-O3 -march=native -funroll-loops -fopenmp (built on this host)"
    except (OSError, subprocess.SubprocessError) as e:
        log(f"cpu baseline: no native build ({e}); using the x86-64-v3 build")
    oracle.lib(flavour)
    return oracle, flavour, flags, cpu


def cpu_baseline(state32, eps, theta, dt, multipole, reps=3):
    """KDK steps of the same workload on the FP64 oracle, the reference CPU path restated:
    one warm-up step, then the mean of `reps` timed steps (steady clock), as the reference's
    --bench harness times (projects/02_nbody_gravity/src/main.cpp:316-336).  The stages of
    one force evaluation are timed apart: the serial insertion tree with its quadrupoles and
    the OpenMP group walk; the kicks and the drift are a step minus its evaluation."""
    oracle, flavour, flags, cpu = oracle_for_baseline()
    st = {k: np.asarray(v, np.float64) for k, v in state32.items()}
    quad = multipole == "quadrupole"
    # an untimed build first: the first one also pays the first touch of the tree's memory
    T = oracle.Tree(st["x"], st["y"], st["z"], st["m"], quad=quad, fast=flavour)
    del T
    t0 = time.perf_counter()
    T = oracle.Tree(st["x"], st["y"], st["z"], st["m"], quad=quad, fast=flavour)
    t1 = time.perf_counter()
    a0 = T.accel_bh(1.0, eps * eps, theta)
    t2 = time.perf_counter()
    del T
    s1, a1 = oracle.leapfrog(st, a0, 1.0, eps * eps, theta, dt, 1, quad=quad, fast=flavour)  # warm-up step
    times = []
    for _ in range(reps):
        ts = time.perf_counter()
        s1, a1 = oracle.leapfrog(s1, a1, 1.0, eps * eps, theta, dt, 1, quad=quad, fast=flavour)
        times.append(time.perf_counter() - ts)
    step = float(np.mean(times))
    # one force evaluation timed by stage (the quadrupoles are built with the tree); the rest of
    # a step (kicks, drift) is the step mean minus both, which run-to-run noise can hide
    stages = {"tree_build_with_quadrupoles_s": round(t1 - t0, 4), "walk_s": round(t2 - t1, 4),
              "rest_of_step_s": round(max(step - (t2 - t0), 0.0), 4)}
    log(f"cpu baseline: tree (with quadrupoles) {t1 - t0:.2f} s, walk {t2 - t1:.2f} s; "
        f"step {step:.2f} s (mean of {reps}) on {cpu['threads_used']} threads")
    return {"value": round(step, 4), "unit": "s/step", "cores": cpu["threads_used"], "kind": "port",
            "host": cpu, "omp": {"OMP_NUM_THREADS": os.environ["OMP_NUM_THREADS"], "OMP_PROC_BIND": "spread",
                                 "OMP_PLACES": "cores"},
            "build": flags, "reps": reps, "warmup": 1, "step_times_s": [round(t, 4) for t in times],
            "stages_of_one_evaluation": stages,
            "sample": f"KDK leapfrog steps of the full workload (N={len(st['x'])}, theta={theta}), FP64 oracle "
                      f"(serial insertion tree + quadrupoles + OpenMP group walk, dynamic,32): 1 warm-up step, "
                      f"then the mean of {reps} timed steps; one force evaluation timed by stage"}


def run_variant(ctx, pynbg, ic, base, steps, warmup):
    """Prime (with interaction counting), warm up, then time `steps` graph-replayed KDK steps.
    Returns wall seconds, the stage stats of the timed steps and the interaction counts."""
    kw = dict(base)
    pc = pynbg.make_params("barnes_hut", count=True, **kw)
    p = pynbg.make_params("barnes_hut", stage_times=True, **kw)
    ctx.upload(ic["x"], ic["y"], ic["z"], ic["m"], ic["vx"], ic["vy"], ic["vz"])
    st0 = ctx.accel(pc)  # System::Prime, with interaction counting
    if warmup > 0:
        ctx.step(p, warmup)
    t0 = time.perf_counter()  # nbg_step synchronises the device before it returns
    st = ctx.step(p, steps)
    t1 = time.perf_counter()
    state = ctx.download()
    st1 = ctx.accel(pc)  # interaction counts at the end state
    if not np.all(np.isfinite(state["ax"])):
        raise RuntimeError("non-finite accelerations")
    return t1 - t0, st, st0, st1


def force_error(ctx, pynbg, ic, base, targets, ref):
    """Relative RMS force error of one evaluation at the initial state on the sampled targets."""
    ctx.upload(ic["x"], ic["y"], ic["z"], ic["m"])
    ctx.accel(pynbg.make_params("barnes_hut", **base))
    a = ctx.accelerations()[targets].astype(np.float64)
    return float(np.sqrt(np.sum((a - ref) ** 2) / np.sum(ref ** 2)))


def ensemble_main(args):
    """Batched ensemble: 512 Plummer/King(W0=6) clusters, N log-uniform in [1024, 65536]
    (mt19937 seed 0), total mass 1 each, eps=0.01, theta=0.6, quadrupole, dt=1e-3, one
    captured hipGraph per step over all clusters, tile sizes chosen by nbg_autotune."""
    import pynbg
    steps = args.steps if args.steps_given else 200
    log("generating the 512-cluster ensemble")
    off, st, kinds = pynbg.ensemble(512, 0, 1024, 65536, 6.0)
    n = int(off[-1])
    nsys = len(off) - 1
    kahan = not args.no_kahan
    base = dict(multipole="quadrupole", G=1.0, eps=0.01, theta=0.6, dt=1e-3, kahan=kahan)
    p = pynbg.make_params("barnes_hut", stage_times=True, **base)
    pc = pynbg.make_params("barnes_hut", count=True, **base)
    with pynbg.Context(n) as ctx:
        ctx.upload_batch(off, st["x"], st["y"], st["z"], st["m"], st["vx"], st["vy"], st["vz"])
        best, cand_ms = ctx.autotune(pynbg.make_params("barnes_hut", **base), reps=3)
        log(f"autotune: {best.as_dict()} candidates ms {np.round(cand_ms, 3).tolist()}")
        e0 = np.array([ctx.diagnostics_system(p, k)["energy"] for k in range(nsys)])
        c0 = ctx.accel(pc)  # System::Prime, with interaction counting
        if args.warmup > 0:
            ctx.step(p, args.warmup)
        t0 = time.perf_counter()
        stt = ctx.step(p, steps)
        t1 = time.perf_counter()
        e1 = np.array([ctx.diagnostics_system(p, k)["energy"] for k in range(nsys)])
        c1 = ctx.accel(pc)
    wall = t1 - t0
    n_p2p = 0.5 * (c0.n_p2p + c1.n_p2p)
    n_m2p = 0.5 * (c0.n_m2p + c1.n_m2p)
    flops = FLOP_P2P * n_p2p + FLOP_M2P_QUAD * n_m2p
    walk_ms = stt.ms_walk / steps
    ach = flops / (walk_ms * 1e-3) / 1e12
    drift = np.abs(e1 - e0) / np.abs(e0)
    stage = {k: round(getattr(stt, "ms_" + k) / steps, 4)
             for k in ("keys", "sort", "build", "upsweep", "walk", "integrate")}
    log(f"{steps} steps: {1e3 * wall / steps:.3f} ms/step, stages {stage}, max |dE/E| {drift.max():.2e}")
    out = {
        "metric": "s/step, batched ensemble of 512 clusters (one captured hipGraph per step)",
        "value": round(wall / steps, 6), "unit": "s/step", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * wall / steps, 4), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (512 clusters: ic::Plummer<3> / ic::King<3>(W0=6) alternating, cluster k seed k)",
        "config": {"workload": "ensemble_512_plummer_king", "clusters": nsys, "n_total": n,
                   "n_min": int(np.diff(off).min()), "n_max": int(np.diff(off).max()), "theta": 0.6, "eps": 0.01,
                   "dt": 1e-3, "multipole": "quadrupole", "kahan": kahan},
        "particle_steps_per_s": round(n * steps / wall, 1),
        "tuning": best.as_dict(),
        "autotune_candidates_ms": {"bin_near4": cand_ms[0], "bin_near8": cand_ms[1], "bin_near16": cand_ms[2],
                                   "leaf_groups": cand_ms[3]} if len(cand_ms) == 4 else cand_ms.tolist(),
        "energy_drift": {"max": float(drift.max()), "median": float(np.median(drift)),
                         "over_steps": steps + args.warmup},
        "roofline": {"bound": "valu", "pipe": "fp32-valu", "achieved": round(ach, 3), "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(ach / PEAK_FP32_TFLOPS, 4), "traffic": None,
                     "kernel": "k_walk_bins", "kernel_ms": round(walk_ms, 4), "flop_per_launch": flops},
        "interactions_per_step": {"p2p": n_p2p, "m2p": n_m2p},
        "interactions_per_s": round((n_p2p + n_m2p) / (walk_ms * 1e-3), 1),
        "stage_ms_per_step": stage,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = ensemble_cpu_baseline(off, st)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


def drift_main(args):
    """C2: 100 KDK steps of the 1M Plummer sphere (theta=0.5, quadrupole, eps=0.005,
    dt=5e-4, graph-replayed), FP64 direct-sum energy and momenta at steps 0 and 100."""
    import pynbg
    steps = args.steps if args.steps_given else 100
    n = args.n
    ic = pynbg.to_f32(pynbg.ic_plummer(n, 0))
    p = pynbg.make_params("barnes_hut", "quadrupole", G=1.0, eps=0.005, theta=0.5, dt=5e-4)
    with pynbg.Context(n) as ctx:
        ctx.upload(ic["x"], ic["y"], ic["z"], ic["m"], ic["vx"], ic["vy"], ic["vz"])
        t0 = time.perf_counter()
        d0 = ctx.diagnostics(p)
        t1 = time.perf_counter()
        ctx.accel(p)
        t2 = time.perf_counter()
        ctx.step(p, steps)
        t3 = time.perf_counter()
        d1 = ctx.diagnostics(p)
    out = {"metric": "energy and momentum drift over 100 KDK steps, FP64 direct-sum diagnostics",
           "value": (d1["energy"] - d0["energy"]) / abs(d0["energy"]), "unit": "dE/E", "n_gpus": 1,
           "steps": steps, "higher_is_better": False, "dtype": "fp32 (state, forces); fp64 (diagnostics)",
           "config": {"workload": f"plummer_{n}_theta0.5_drift", "n": n, "theta": 0.5, "eps": 0.005, "dt": 5e-4},
           "E0": d0["energy"], "E1": d1["energy"], "K0": d0["kinetic"], "W0": d0["potential"],
           "P0": d0["P"].tolist(), "P1": d1["P"].tolist(), "dP": float(np.linalg.norm(d1["P"] - d0["P"])),
           "L0": d0["L"].tolist(), "L1": d1["L"].tolist(),
           "dL_over_L": float(np.linalg.norm(d1["L"] - d0["L"]) / max(np.linalg.norm(d0["L"]), 1e-300)),
           "com_shift": float(np.linalg.norm(d1["com"] - d0["com"])),
           "s_per_step": round((t3 - t2) / steps, 6), "diagnostics_s": round(t1 - t0, 3),
           "energy_oracle": "FP64 O(N^2) softened potential + kinetic on the device (nbg_diagnostics); equal to the "
                            "oracle's total_energy to 1e-11 relative at N=8192 (tests/test_gpu_forces.py)"}
    log(f"dE/E {out['value']:.3e}  |dP| {out['dP']:.3e}  {1e3 * out['s_per_step']:.3f} ms/step")
    print(json.dumps(out), flush=True)


def modes_main(args):
    """The other deck-reachable force and step modes on the 1M Plummer state (seed 0,
    eps 0.005, quadrupoles): grouped walk with Plummer and spline softening, per-particle
    walk with the geometric and the relative MAC, one adaptive step and one block (rung)
    coarse step.  Force modes: mean HIP-event walk time of `steps` evaluations, interaction
    counts and the FP32 rate they imply, relative RMS against the direct sum on 16,384
    targets."""
    import pynbg
    n = args.n
    reps = args.steps if args.steps_given else 5
    eps, theta = args.eps, 0.5
    ic = pynbg.to_f32(pynbg.ic_plummer(n, 0))
    rng = np.random.default_rng(12345)
    targets = np.sort(rng.choice(n, min(n, args.error_sample), replace=False))
    base = dict(G=1.0, eps=eps, theta=theta, dt=args.dt)
    modes = [("grouped_plummer", dict()), ("grouped_spline", dict(softening="spline")),
             ("per_particle_geometric", dict(per_particle=True)),
             ("relative_alpha0.001", dict(mac="relative", alpha=0.001)),
             ("relative_alpha0.004", dict(mac="relative", alpha=0.004))]
    out_modes = {}
    with pynbg.Context(n) as ctx:
        ctx.upload(ic["x"], ic["y"], ic["z"], ic["m"], ic["vx"], ic["vy"], ic["vz"])
        ref_p = ctx.direct_targets(pynbg.make_params("direct", G=1.0, eps=eps), targets).astype(np.float64)
        ref_s = ctx.direct_targets(pynbg.make_params("direct", G=1.0, eps=eps, softening="spline"),
                                   targets).astype(np.float64)
        # history for the relative MAC: |a| of a grouped theta = 0.5 evaluation (as after System::Prime)
        ctx.accel(pynbg.make_params("barnes_hut", **base))
        a_hist = ctx.accelerations()
        for name, kw in modes:
            pc = pynbg.make_params("barnes_hut", count=True, **base, **kw)
            p = pynbg.make_params("barnes_hut", **base, **kw)
            ctx.upload_accel(a_hist[:, 0], a_hist[:, 1], a_hist[:, 2])
            stc = ctx.accel(pc)
            a = ctx.accelerations()[targets].astype(np.float64)
            ref = ref_s if kw.get("softening") == "spline" else ref_p
            err = float(np.sqrt(np.sum((a - ref) ** 2) / np.sum(ref ** 2)))
            ms = []
            for _ in range(reps + 1):
                ctx.upload_accel(a_hist[:, 0], a_hist[:, 1], a_hist[:, 2])
                ms.append(ctx.accel(p).ms_walk)
            walk_ms = float(np.mean(ms[1:]))
            flop = FLOP_P2P * stc.n_p2p + FLOP_M2P_QUAD * stc.n_m2p
            out_modes[name] = {"walk_ms": round(walk_ms, 4), "n_p2p": int(stc.n_p2p), "n_m2p": int(stc.n_m2p),
                               "tflops": round(flop / (walk_ms * 1e-3) / 1e12, 2),
                               "frac_fp32_peak": round(flop / (walk_ms * 1e-3) / 1e12 / PEAK_FP32_TFLOPS, 4),
                               "rel_rms_vs_direct": err}
            log(f"{name}: walk {walk_ms:.3f} ms, m2p {stc.n_m2p:.3e} p2p {stc.n_p2p:.3e}, "
                f"{out_modes[name]['tflops']} TFLOP/s, rel_rms {err:.3e}")
        # one adaptive step (graph) and one block coarse step from the primed state
        for name, kw in (("adaptive_step", dict(adaptive=True, eta=0.03)),
                         ("block_coarse_step", dict(block=True, eta=0.03, block_max_rung=8))):
            p = pynbg.make_params("barnes_hut", **base, **kw)
            ctx.upload(ic["x"], ic["y"], ic["z"], ic["m"], ic["vx"], ic["vy"], ic["vz"])
            ctx.accel(p)
            ctx.step(p, 1)  # warm-up (graph capture)
            t0 = time.perf_counter()
            st = ctx.step(p, reps)
            t1 = time.perf_counter()
            out_modes[name] = {"ms_per_step": round(1e3 * (t1 - t0) / reps, 3), "dt_last": st.dt_last,
                               "t_advanced": st.t_advanced, "rung_max": int(st.rung_max),
                               "particle_evals_per_step": int(st.particle_evals) // reps}
            log(f"{name}: {out_modes[name]}")
    main_mode = out_modes["per_particle_geometric"]
    out = {"metric": "force-evaluation walk time per mode, 1M-body Plummer", "value": main_mode["walk_ms"],
           "unit": "ms/evaluation (per-particle geometric walk)", "n_gpus": 1, "steps": reps,
           "higher_is_better": False, "dtype": "fp32", "data": "synthetic",
           "config": {"workload": f"plummer_{n}_modes", "n": n, "eps": eps, "theta": theta, "multipole": "quadrupole",
                      "dt": args.dt},
           "modes": out_modes,
           "roofline": {"bound": "valu", "pipe": "fp32-valu", "achieved": main_mode["tflops"], "peak": PEAK_FP32_TFLOPS,
                        "unit": "TFLOP/s", "frac": main_mode["frac_fp32_peak"], "traffic": None}}
    print(json.dumps(out), flush=True)


def ensemble_cpu_baseline(off, st, nclusters=None):
    """One KDK step of the first `nclusters` clusters (all of them by default) on the FP64
    oracle; each cluster's opening accelerations are computed untimed."""
    oracle, flavour, flags, cpu = oracle_for_baseline()
    tot = 0.0
    npart = 0
    nclusters = len(off) - 1 if nclusters is None else nclusters
    for k in range(nclusters):
        sl = slice(off[k], off[k + 1])
        s = {q: np.asarray(st[q][sl], np.float64) for q in st}
        T = oracle.Tree(s["x"], s["y"], s["z"], s["m"], quad=True, fast=flavour)
        a0 = T.accel_bh(1.0, 0.01 ** 2, 0.6)
        del T
        t0 = time.perf_counter()
        oracle.leapfrog(s, a0, 1.0, 0.01 ** 2, 0.6, 1e-3, 1, quad=True, fast=flavour)
        tot += time.perf_counter() - t0
        npart += off[k + 1] - off[k]
    return {"value": round(tot, 4), "unit": "s/step", "cores": cpu["threads_used"], "kind": "port", "host": cpu,
            "build": flags,
            "sample": f"one KDK step of {nclusters} of the {len(off) - 1} clusters ({npart} particles, "
                      f"{npart / off[-1]:.1%} of the ensemble) on the FP64 oracle, OpenMP, clusters one "
                      f"after another"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="plummer_1m", choices=["plummer_1m", "ensemble", "drift", "modes"],
                    help="plummer_1m: the headline learned/plain comparison; ensemble: 512 clusters; "
                         "drift: 100-step energy/momentum drift of the 1M Plummer sphere; modes: spline, "
                         "per-particle/relative-MAC walks, adaptive and block steps on the 1M sphere")
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--variant", default="learned", choices=["learned", "plain"],
                    help="learned: the learned theta=0.8 step timed next to plain theta=0.5 (the headline "
                         "metric); plain: plain Barnes-Hut at --theta only")
    ap.add_argument("--theta", type=float, default=0.5, help="plain Barnes-Hut opening angle")
    ap.add_argument("--learned-theta", type=float, default=0.8)
    ap.add_argument("--learn-band", type=float, default=0.0)
    ap.add_argument("--learn-amin", type=float, default=1e-3,
                    help="learned pairs also need a correction scale G M s^3 / r^5 >= this")
    ap.add_argument("--weights", default=None,
                    help="learned model blob (.npy); default: the committed opengl-physics_amd/learned/farfield_mlp.npy")
    ap.add_argument("--eps", type=float, default=0.005)
    ap.add_argument("--dt", type=float, default=5e-4)
    ap.add_argument("--multipole", default="quadrupole", choices=["monopole", "quadrupole"])
    ap.add_argument("--no-kahan", action="store_true")
    ap.add_argument("--error-sample", type=int, default=16384, help="targets of the direct-sum error check")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--traffic-from", nargs="+", metavar="CSV",
                    help="instead of running: write profiles/traffic_plummer_1m.json from the counter_collection.csv "
                         "files of two rocprofv3 --pmc passes (FETCH_SIZE, WRITE_SIZE) over `bench.py --no-cpu`, "
                         "optionally followed by the same two passes over `bench.py --no-cpu --variant plain` "
                         "(the plain theta=0.5 walk's own bytes)")
    args = ap.parse_args()
    if args.traffic_from:
        if len(args.traffic_from) not in (2, 4):
            print("--traffic-from takes FETCH WRITE [FETCH_PLAIN WRITE_PLAIN]", file=sys.stderr)
            sys.exit(2)
        out = traffic_from_counters(args.traffic_from[0], args.traffic_from[1],
                                    os.path.join(ROOT, "profiles", "traffic_plummer_1m.json"), args.n,
                                    "plummer_1m: bench.py --no-cpu (learned theta=0.8 and plain theta=0.5 steps, "
                                    "their error checks and the equal-error plain steps); the plain theta=0.5 walk "
                                    "from bench.py --no-cpu --variant plain" if len(args.traffic_from) == 4 else
                                    "plummer_1m: bench.py --no-cpu (learned theta=0.8 and plain theta=0.5 steps, "
                                    "their error checks and the equal-error plain steps)",
                                    plain_csvs=args.traffic_from[2:] or None)
        log(json.dumps({k: v for k, v in out.items() if not k.endswith("per_launch")}))
        return
    args.steps_given = args.steps is not None
    if args.steps is None:
        args.steps = 10

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the multi-GPU decomposition is "
              f"design-only in this build", file=sys.stderr)
        sys.exit(2)
    if args.workload == "ensemble":
        ensemble_main(args)
        return
    if args.workload == "drift":
        drift_main(args)
        return
    if args.workload == "modes":
        modes_main(args)
        return
    if args.variant == "learned" and args.multipole != "quadrupole":
        print("the learned variant needs quadrupoles", file=sys.stderr)
        sys.exit(2)

    import pynbg
    from pynbg import learned
    n = args.n
    log(f"generating Plummer N={n} seed 0")
    ic = pynbg.to_f32(pynbg.ic_plummer(n, 0))
    kahan = not args.no_kahan
    plain = dict(multipole=args.multipole, G=1.0, eps=args.eps, theta=args.theta, dt=args.dt, kahan=kahan)
    lrn = dict(multipole="quadrupole", G=1.0, eps=args.eps, theta=args.learned_theta, dt=args.dt, kahan=kahan,
               learned=True, learn_band=args.learn_band, learn_amin=args.learn_amin)
    rng = np.random.default_rng(12345)
    targets = np.sort(rng.choice(n, min(n, args.error_sample), replace=False))
    with pynbg.Context(n) as ctx:
        ctx.upload(ic["x"], ic["y"], ic["z"], ic["m"])
        tile = ctx.tuning().near_tile if kahan else 8
        ref = ctx.direct_targets(pynbg.make_params("direct", G=1.0, eps=args.eps), targets).astype(np.float64)
        err_plain = force_error(ctx, pynbg, ic, plain, targets, ref)
        wall_p, st_p, c0_p, c1_p = run_variant(ctx, pynbg, ic, plain, args.steps, args.warmup)
        log(f"plain theta={args.theta}: {1e3 * wall_p / args.steps:.3f} ms/step, rel_rms {err_plain:.3e}")
        if args.variant == "learned":
            blob = learned.load_weights(args.weights) if args.weights else learned.load_weights()
            act_name = {v: k for k, v in pynbg.LEARN_ACT.items()}[int(blob[pynbg.LEARN_ACT_SLOT])]
            ctx.learn_set_weights(blob)
            err_l = force_error(ctx, pynbg, ic, lrn, targets, ref)
            err_l_plain = force_error(ctx, pynbg, ic, dict(lrn, learned=False), targets, ref)
            wall_l, st_l, c0_l, c1_l = run_variant(ctx, pynbg, ic, lrn, args.steps, args.warmup)
            log(f"learned theta={args.learned_theta} band {args.learn_band}: {1e3 * wall_l / args.steps:.3f} "
                f"ms/step, rel_rms {err_l:.3e} (plain at that theta {err_l_plain:.3e}); "
                f"pairs {c0_l.n_learned}")
            # plain Barnes-Hut at the learned step's error: bisect theta between the plain
            # and the learned opening angles until the errors agree within 2%
            plain_at = dict(plain, multipole="quadrupole")
            lo, hi, e_lo, e_hi = args.theta, args.learned_theta, err_plain, err_l_plain
            probes = []
            theta_iso, err_iso = None, None
            if e_lo <= err_l <= e_hi:
                for _ in range(14):
                    mid = 0.5 * (lo + hi)
                    e = force_error(ctx, pynbg, ic, dict(plain_at, theta=mid), targets, ref)
                    probes.append([round(mid, 6), e])
                    theta_iso, err_iso = mid, e
                    if abs(e - err_l) <= 0.02 * err_l:
                        break
                    if e < err_l:
                        lo = mid
                    else:
                        hi = mid
                wall_iso, st_iso, _, _ = run_variant(ctx, pynbg, ic, dict(plain_at, theta=theta_iso), args.steps,
                                                     args.warmup)
                log(f"plain theta={theta_iso:.4f} (equal error {err_iso:.3e}): {1e3 * wall_iso / args.steps:.3f} ms/step")
            wall_08, st_08, _, _ = run_variant(ctx, pynbg, ic, dict(plain_at, theta=args.learned_theta), args.steps,
                                               args.warmup)
            log(f"plain theta={args.learned_theta}: {1e3 * wall_08 / args.steps:.3f} ms/step")

    def walk_roofline(st, c0, c1, multipole, theta, learn=False):
        n_p2p = 0.5 * (c0.n_p2p + c1.n_p2p)
        n_m2p = 0.5 * (c0.n_m2p + c1.n_m2p)
        fm2p = FLOP_M2P_QUAD if multipole == "quadrupole" else FLOP_M2P_MONO
        flops = FLOP_P2P * n_p2p + fm2p * n_m2p
        walk_ms = st.ms_walk / args.steps
        wtile = tile if (not learn or tile == 16) else 8
        ach = flops / (walk_ms * 1e-3) / 1e12
        return {"bound": "valu", "pipe": "fp32-valu", "achieved": round(ach, 3), "peak": PEAK_FP32_TFLOPS,
                "unit": "TFLOP/s", "frac": round(ach / PEAK_FP32_TFLOPS, 4),
                "traffic": kernel_traffic(n, walk_kernel_name(multipole == "quadrupole", kahan, learn, wtile)),
                "kernel": walk_kernel_name(multipole == "quadrupole", kahan, learn, wtile), "kernel_ms": round(walk_ms, 4),
                "flop_per_launch": flops}, n_p2p, n_m2p

    roof_p, p2p_p, m2p_p = walk_roofline(st_p, c0_p, c1_p, args.multipole, args.theta)
    stage_p = {k: round(getattr(st_p, "ms_" + k) / args.steps, 4)
               for k in ("keys", "sort", "build", "upsweep", "walk", "integrate")}
    plain_block = {"theta": args.theta, "s_per_step": round(wall_p / args.steps, 6), "rel_rms_force": err_plain,
                   "stage_ms_per_step": stage_p, "stage_hbm": stage_hbm(st_p, args.steps, n),
                   "interactions_per_step": {"p2p": p2p_p, "m2p": m2p_p},
                   "interactions_per_s": round((p2p_p + m2p_p) / (roof_p["kernel_ms"] * 1e-3), 1),
                   "roofline": roof_p}
    out = {
        "metric": "s/step, 1M-body BH: plain θ=0.5 vs learned θ=0.8 at equal RMS force error",
        "unit": "s/step",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (ic::Plummer<3> seed 0, generated on the host)",
    }
    if args.variant == "plain":
        out.update({"value": plain_block["s_per_step"], "ms_per_step": round(1e3 * wall_p / args.steps, 4),
                    "config": {"workload": f"plummer_{n}_barnes_hut_theta{args.theta}", "n": n,
                               "theta": args.theta, "eps": args.eps, "dt": args.dt, "multipole": args.multipole,
                               "kahan": kahan, "variant": "plain"},
                    "roofline": roof_p, "rel_rms_force": err_plain,
                    "interactions_per_s": plain_block["interactions_per_s"],
                    "interactions_per_step": plain_block["interactions_per_step"],
                    "stage_ms_per_step": stage_p, "stage_hbm": plain_block["stage_hbm"],
                    "tree": {"nodes": st_p.n_nodes, "leaves": st_p.n_leaves, "max_depth": st_p.max_depth}})
    else:
        roof_w, p2p_l, m2p_l = walk_roofline(st_l, c0_l, c1_l, "quadrupole", args.learned_theta, learn=True)
        pairs = 0.5 * (c0_l.n_learned + c1_l.n_learned)
        mlp_ms = st_l.ms_learned / args.steps
        mlp_ach = pairs * FLOP_MLP_PAIR / (mlp_ms * 1e-3) / 1e12
        roof_m = {"bound": "mfma", "pipe": "bf16-mfma", "achieved": round(mlp_ach, 3), "peak": PEAK_BF16_TFLOPS,
                  "unit": "TFLOP/s", "frac": round(mlp_ach / PEAK_BF16_TFLOPS, 5),
                  "traffic": kernel_traffic(n, f"void k_learn_infer<false, {pynbg.LEARN_ACT[act_name]}>"),
                  "kernel": "k_learn_infer (+ k_learn_zero, k_learn_apply)", "kernel_ms": round(mlp_ms, 4),
                  "flop_per_launch": pairs * FLOP_MLP_PAIR, "pairs_per_launch": pairs}
        stage_l = {k: round(getattr(st_l, "ms_" + k) / args.steps, 4)
                   for k in ("keys", "sort", "build", "upsweep", "walk", "learned", "integrate")}
        dominant, secondary = (roof_m, roof_w) if mlp_ms >= roof_w["kernel_ms"] else (roof_w, roof_m)
        s_l = wall_l / args.steps
        out.update({
            "value": round(s_l, 6),
            "ms_per_step": round(1e3 * s_l, 4),
            "config": {"workload": f"plummer_{n}_learned_theta{args.learned_theta}_vs_plain_theta{args.theta}",
                       "n": n, "eps": args.eps, "dt": args.dt, "multipole": "quadrupole", "kahan": kahan,
                       "variant": "learned", "theta": args.learned_theta, "learn_band": args.learn_band,
                       "learn_amin": args.learn_amin,
                       "model": f"13-64-64-64-3 {'ReLU' if act_name == 'relu' else 'SiLU'}, BF16 MFMA "
                                f"({os.path.relpath(args.weights, ROOT) if args.weights else 'opengl-physics_amd/learned/farfield_mlp.npy'})"},
            "rel_rms_force": err_l,
            "comparison": {
                "plain": plain_block,
                "learned": {"theta": args.learned_theta, "learn_band": args.learn_band, "learn_amin": args.learn_amin,
                            "s_per_step": round(s_l, 6),
                            "rel_rms_force": err_l, "rel_rms_force_without_model": err_l_plain,
                            "stage_ms_per_step": stage_l, "learned_pairs_per_step": pairs,
                            "interactions_per_step": {"p2p": p2p_l, "m2p": m2p_l}},
                "speedup_vs_plain": round(wall_p / wall_l, 4),
                "plain_iso_error": None if theta_iso is None else {
                    "theta": round(theta_iso, 6), "rel_rms_force": err_iso, "s_per_step": round(wall_iso / args.steps, 6),
                    "speedup_of_learned": round(wall_iso / wall_l, 4),
                    "stage_ms_per_step": {k: round(getattr(st_iso, "ms_" + k) / args.steps, 4)
                                          for k in ("keys", "sort", "build", "upsweep", "walk", "integrate")},
                    "bisection": probes,
                    "how": f"theta bisected in [{args.theta}, {args.learned_theta}] until the plain relative RMS "
                           f"error is within 2% of the learned step's, on the same targets"},
                "plain_at_learned_theta": {"theta": args.learned_theta, "rel_rms_force": err_l_plain,
                                           "s_per_step": round(wall_08 / args.steps, 6),
                                           "cost_of_the_network_s_per_step": round((wall_l - wall_08) / args.steps, 6)},
                "error_target": 1e-3,
                "error_target_met": bool(err_l <= 1e-3),
                "error_sample": f"{len(targets)} random targets (seed 12345) of the initial state, judged by the "
                                f"FP32 compensated direct sum over all {n} sources",
            },
            "roofline": dominant,
            "roofline_secondary": secondary,
            "interactions_per_s": round((p2p_l + m2p_l) / (roof_w["kernel_ms"] * 1e-3), 1),
            "stage_ms_per_step": stage_l,
            "stage_hbm": stage_hbm(st_l, args.steps, n, moments=True),
            "tree": {"nodes": st_l.n_nodes, "leaves": st_l.n_leaves, "max_depth": st_l.max_depth},
        })
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(ic, args.eps, args.theta, args.dt, args.multipole)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
