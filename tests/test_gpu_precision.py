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
so that input rounding and arithmetic error are reported separately; and the FP64 state
(nbg_upload_f64: tree from FP64 positions, every force difference formed in FP32 from
offsets relative to leaf centres) against the oracle's FP64 grouped walk of that state,
next to the FP32 state against the same.

C2 (small N): 100 KDK steps (dt = 5e-4) of a Plummer sphere on the device against the
oracle's FP64 run of the same steps (Integrator.cpp:9-39 with System.hpp:43-78's
Barnes-Hut dispatch): energy and momentum drift of both, and the distance between the
final states.
"""
import json

import numpy as np
import pytest

from conftest import keys_at

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
    # the FP64 state: tree from the FP64 positions, cell-relative FP32 differences
    ctx.upload_f64(st64["x"], st64["y"], st64["z"], st32["m"])
    ctx.accel(pynbg.make_params("barnes_hut", G=1.0, eps=eps, theta=0.5, kahan=True))
    res["bh_f64state_rel"] = ctx.accelerations()[tg].astype(np.float64)
    ref64 = oracle.accel_direct_targets(st64["x"], st64["y"], st64["z"], st64["m"], 1.0, eps * eps, tg, fast=True)
    s32 = {k: st32[k].astype(np.float64) for k in ("x", "y", "z", "m")}
    ref32 = oracle.accel_direct_targets(s32["x"], s32["y"], s32["z"], s32["m"], 1.0, eps * eps, tg, fast=True)
    report = {"workload": "two Hernquist spheres 2 x 2,097,152, 20 apart, parabolic head-on, eps 0.01, theta 0.5",
              "targets": len(tg), "tree_depth": int(stt.max_depth), "nodes": int(stt.n_nodes),
              "input_rounding_fp32_vs_fp64_state": _per_particle(ref32, ref64)}
    for k, a in res.items():
        report[f"{k}_vs_fp64_state"] = _per_particle(a, ref64)
        report[f"{k}_vs_fp32_state"] = _per_particle(a, ref32)
    # the same method in FP64 on the FP64 state (the oracle's grouped walk, masses as the
    # device holds them): what is left is the device's own arithmetic and input rounding
    mq = st32["m"].astype(np.float64)
    bh64 = oracle.Tree(st64["x"], st64["y"], st64["z"], mq, quad=True).accel_bh(1.0, eps * eps, 0.5)[tg]
    report["bh_f64state_rel_vs_oracle_bh_fp64_state"] = _per_particle(res["bh_f64state_rel"], bh64)
    report["bh_kahan_fp32_state_vs_oracle_bh_fp64_state"] = _per_particle(res["bh_kahan"], bh64)
    print("C3 " + json.dumps(report))
    # arithmetic of the direct sum against the same FP32 inputs: FP32 rounding only
    assert report["direct_fp32_vs_fp32_state"]["rel_rms"] < 1e-5
    # Barnes-Hut theta = 0.5 with quadrupoles: the expansion error dominates
    assert report["bh_kahan_vs_fp32_state"]["rel_rms"] < 1e-3
    assert report["bh_kahan_vs_fp64_state"]["rel_rms"] < 1e-3
    # the FP64 state with cell-relative differences: no input-rounding outliers left
    # (measured: relative RMS 8.8e-8 and worst particle 3.4e-5, against 9.4e-6 and 1.5e-3
    # for the FP32 state, both against the oracle's FP64 walk of the FP64 state)
    rel = report["bh_f64state_rel_vs_oracle_bh_fp64_state"]
    assert rel["rel_rms"] < 1e-6 and rel["max"] < 1e-4
    assert rel["max"] < 0.1 * report["bh_kahan_fp32_state_vs_oracle_bh_fp64_state"]["max"]


def test_f64_state_dense_cluster_far_from_origin(gpu_ctx_factory):
    """A Plummer sphere of scale 1e-3 (N = 4,096, eps = 1e-5) centred at (20, 1, 0):
    rounding absolute positions to FP32 (ulp 1.9e-6 at 20, against separations of order
    1e-4) spoils the FP32 state's forces at the per-cent level, while the FP64 state's walk,
    differencing offsets from leaf centres, agrees with the oracle's FP64 grouped walk on the
    same positions to the FP32 rounding of its sums (measured: relative RMS 1.6e-2 against
    7.8e-7; worst particle 0.21 against 1.6e-5)."""
    import oracle
    import pynbg
    s = oracle.plummer(4096, 3)
    x, y, z = 20.0 + 1e-3 * s["x"], 1.0 + 1e-3 * s["y"], 1e-3 * s["z"]
    m32 = s["m"].astype(np.float32)
    eps = 1e-5
    T = oracle.Tree(x, y, z, m32.astype(np.float64), quad=True)
    ref = T.accel_bh(1.0, eps * eps, 0.5)
    ctx = gpu_ctx_factory(4096)
    p = pynbg.make_params("barnes_hut", G=1.0, eps=eps, theta=0.5)
    errs = {}
    for name in ("fp32", "fp64"):
        if name == "fp32":
            ctx.upload(x.astype(np.float32), y.astype(np.float32), z.astype(np.float32), m32)
        else:
            ctx.upload_f64(x, y, z, m32)
        ctx.accel(p)
        errs[name] = _per_particle(ctx.accelerations().astype(np.float64), ref)
    print("dense cluster at x = 20 vs the oracle's FP64 walk:", errs)
    assert errs["fp32"]["rel_rms"] > 1e-3
    assert errs["fp64"]["rel_rms"] < 2e-6 and errs["fp64"]["max"] < 1e-4


def test_f64_state_tree_lists_and_forces_match_oracle(gpu_ctx_factory):
    """An FP64 state (Plummer N = 65,536, seed 0, positions not rounded): keys, order,
    topology and FP64 cell geometry bit-exact against the oracle's insertion tree of the
    same FP64 positions, per-leaf lists exact, and accelerations within the FP32 rounding of
    the sums of the oracle's FP64 grouped walk (relative RMS <= 1e-6; the worst particles,
    where the pulls nearly cancel, <= 1e-4), below the FP32-rounded state's error against
    the same oracle (which carries the input rounding as well)."""
    import oracle
    import pynbg
    from test_gpu_tree import _compare_trees
    s = oracle.plummer(65536, 0)
    x, y, z = s["x"], s["y"], s["z"]
    m32 = s["m"].astype(np.float32)
    m = m32.astype(np.float64)
    eps2 = 0.01 ** 2
    ctx = gpu_ctx_factory(65536)
    ctx.upload_f64(x, y, z, m32)
    st = ctx.accel(pynbg.make_params("barnes_hut", "quadrupole", G=1.0, eps=0.01, theta=0.5, count=True))
    a = ctx.accelerations().astype(np.float64)
    g = ctx.tree_export()
    T = oracle.Tree(x, y, z, m, quad=True)
    o = T.export()
    keys_hi, _ = oracle.morton_keys(x, y, z, 21)
    np.testing.assert_array_equal(g["keys"], np.sort(keys_at(keys_hi, st.key_levels)))
    _compare_trees(g, o, 65536)
    ref, nfar, nnear = T.accel_bh(1.0, eps2, 0.5, counts=True)
    gfar, gnear = ctx.interaction_counts(g["n_nodes"])
    leaves = g["child_count"] == 0
    assert int(np.sum(gfar[leaves] != nfar[leaves])) == 0 and int(np.sum(gnear[leaves] != nnear[leaves])) == 0
    e64 = _per_particle(a, ref)
    ctx.upload(x.astype(np.float32), y.astype(np.float32), z.astype(np.float32), m32)
    ctx.accel(pynbg.make_params("barnes_hut", "quadrupole", G=1.0, eps=0.01, theta=0.5))
    e32 = _per_particle(ctx.accelerations().astype(np.float64), ref)
    print("FP64 state vs oracle BH:", e64, " FP32 state vs the same:", e32, "nodes", st.n_nodes)
    assert e64["rel_rms"] < 1e-6 and e64["max"] < 1e-4
    assert e64["rel_rms"] < e32["rel_rms"]


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
