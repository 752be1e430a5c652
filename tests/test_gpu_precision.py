"""Precision study of the FP32 force path (C3) and energy/momentum drift against the FP64
oracle (C2), on the GPU.

C3: two equal-mass Hernquist spheres (ic::Hernquist<3>, projects/nbody_core/include/ngrav/ic/
Hernquist.hpp:131-181; N = 2,097,152 each, seeds 0 and 1, total mass 1 each, scale radius 1)
20 units apart on x, approaching head-on at the parabolic speed sqrt(2 G (M1+M2) / d) =
sqrt(0.2); eps = 0.01, theta = 0.5, quadrupole.  Per-particle force errors on 65,536
random particles (seed 7) against the FP64 direct sum (ComputeAccelDirectTargets,
projects/nbody_core/src/Solvers.cpp:108-134, restated in oracle/nbody_oracle.cpp) of
  * the FP64 state, and
  * the FP32-rounded state the device holds,
so that input rounding and arithmetic error are reported separately.

C2 (small N): 100 KDK steps (dt = 5e-4) of a Plummer sphere on the device against the
oracle's FP64 run of the same steps (Integrator.cpp:9-39 with System.hpp:43-78's
Barnes-Hut dispatch): energy and momentum drift of both, and the distance between the
final states.
"""
import json

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _per_particle(a, ref):
    e = np.linalg.norm(a - ref, axis=1) / np.linalg.norm(ref, axis=1)
    return dict(rel_rms=float(np.sqrt(np.sum((a - ref) ** 2) / np.sum(ref ** 2))),
                p50=float(np.percentile(e, 50)), p90=float(np.percentile(e, 90)),
                p99=float(np.percentile(e, 99)), max=float(e.max()))


def _collision(n_each):
    import pynbg
    a = pynbg.ic_hernquist(n_each, 0)
    b = pynbg.ic_hernquist(n_each, 1)
    v = np.sqrt(2.0 * 1.0 * 2.0 / 20.0)
    st = {}
    for k in ("x", "y", "z", "vx", "vy", "vz", "m"):
        st[k] = np.concatenate([a[k], b[k]])
    st["x"][:n_each] -= 10.0
    st["x"][n_each:] += 10.0
    st["vx"][:n_each] += 0.5 * v
    st["vx"][n_each:] -= 0.5 * v
    return st


def test_c3_hernquist_collision_precision(gpu_ctx_factory):
    import oracle
    import pynbg
    n_each = 2097152
    st64 = _collision(n_each)
    st32 = pynbg.to_f32(st64)
    n = 2 * n_each
    eps = 0.01
    rng = np.random.default_rng(7)
    tg = np.sort(rng.choice(n, 65536, replace=False))
    ctx = gpu_ctx_factory(n)
    ctx.upload(st32["x"], st32["y"], st32["z"], st32["m"])
    res = {}
    for name, kw in (("bh_kahan", dict(kahan=True)), ("bh_plain_sums", dict(kahan=False))):
        stt = ctx.accel(pynbg.make_params("barnes_hut", G=1.0, eps=eps, theta=0.5, **kw))
        res[name] = ctx.accelerations()[tg].astype(np.float64)
    res["direct_fp32"] = ctx.direct_targets(pynbg.make_params("direct", G=1.0, eps=eps), tg).astype(np.float64)
    ref64 = oracle.accel_direct_targets(st64["x"], st64["y"], st64["z"], st64["m"], 1.0, eps * eps, tg, fast=True)
    s32 = {k: st32[k].astype(np.float64) for k in ("x", "y", "z", "m")}
    ref32 = oracle.accel_direct_targets(s32["x"], s32["y"], s32["z"], s32["m"], 1.0, eps * eps, tg, fast=True)
    report = {"workload": "two Hernquist spheres 2 x 2,097,152, 20 apart, parabolic head-on, eps 0.01, theta 0.5",
              "targets": len(tg), "tree_depth": int(stt.max_depth), "nodes": int(stt.n_nodes),
              "input_rounding_fp32_vs_fp64_state": _per_particle(ref32, ref64)}
    for k, a in res.items():
        report[f"{k}_vs_fp64_state"] = _per_particle(a, ref64)
        report[f"{k}_vs_fp32_state"] = _per_particle(a, ref32)
    print("C3 " + json.dumps(report))
    # arithmetic of the direct sum against the same FP32 inputs: FP32 rounding only
    assert report["direct_fp32_vs_fp32_state"]["rel_rms"] < 1e-5
    # Barnes-Hut theta = 0.5 with quadrupoles: the expansion error dominates
    assert report["bh_kahan_vs_fp32_state"]["rel_rms"] < 1e-3
    assert report["bh_kahan_vs_fp64_state"]["rel_rms"] < 1e-3


def test_c2_drift_against_fp64_oracle(gpu_ctx_factory):
    import oracle
    import pynbg
    from conftest import plummer_f32
    n, eps, theta, dt, steps = 16384, 0.005, 0.5, 5e-4, 100
    s = plummer_f32(n, 0)
    st = {k: s[k].astype(np.float64) for k in s}
    ctx = gpu_ctx_factory(n)
    ctx.upload(s["x"], s["y"], s["z"], s["m"], s["vx"], s["vy"], s["vz"])
    p = pynbg.make_params("barnes_hut", G=1.0, eps=eps, theta=theta, dt=dt)
    d0 = ctx.diagnostics(p)
    ctx.accel(p)
    ctx.step(p, steps)
    d1 = ctx.diagnostics(p)
    g = ctx.download()
    # the oracle: FP64 state, FP64 Barnes-Hut with quadrupoles, same steps
    T = oracle.Tree(st["x"], st["y"], st["z"], st["m"], quad=True, fast=True)
    a0 = T.accel_bh(1.0, eps * eps, theta)
    del T
    o, _ = oracle.leapfrog(st, a0, 1.0, eps * eps, theta, dt, steps, quad=True, fast=True)
    e0 = oracle.total_energy(st, 1.0, eps * eps)
    e1 = oracle.total_energy(o, 1.0, eps * eps)
    _, _, P0 = oracle.moments(st)
    _, _, P1 = oracle.moments(o)
    pos_g = np.stack([g["x"], g["y"], g["z"]], 1).astype(np.float64)
    pos_o = np.stack([o["x"], o["y"], o["z"]], 1)
    report = {"workload": f"Plummer N={n} seed 0, eps {eps}, theta {theta}, quadrupole, {steps} KDK steps dt {dt}",
              "gpu_dE_over_E": (d1["energy"] - d0["energy"]) / abs(d0["energy"]),
              "oracle_dE_over_E": (e1 - e0) / abs(e0),
              "gpu_dP": float(np.linalg.norm(d1["P"] - d0["P"])),
              "oracle_dP": float(np.linalg.norm(P1 - P0)),
              "final_position_rel_rms_gpu_vs_oracle": float(np.sqrt(np.sum((pos_g - pos_o) ** 2) /
                                                                    np.sum(pos_o ** 2)))}
    print("C2 " + json.dumps(report))
    assert d0["energy"] == pytest.approx(e0, rel=1e-9)
    # same scheme, same tree decisions at step 0; FP32 state and forces on the device
    assert abs(report["gpu_dE_over_E"] - report["oracle_dE_over_E"]) < 1e-5
    assert abs(report["gpu_dE_over_E"]) < 1e-4
    assert report["gpu_dP"] < 1e-5
    assert report["final_position_rel_rms_gpu_vs_oracle"] < 1e-4
