"""GPU parity of the deck-reachable modes next to the grouped Plummer walk, against the
FP64 oracle on the same FP32-rounded inputs:
  * spline softening (AccumPairSpline, projects/nbody_core/src/Solvers.cpp:43-63) in the
    direct sum, the grouped walk and the FP64 energy (SplinePotential3D, Softening.hpp:49-56);
  * the per-particle walk (WalkBH, Solvers.cpp:181-227) with the geometric and the relative
    MAC (Mac.hpp:37-42), and onto target subsets (ComputeAccelBarnesHutTargets,
    Solvers.cpp:384-405);
  * adaptive global dt (ChooseAdaptiveDt, Integrator.cpp:41-59);
  * block / rung steps (System::BlockStep, System.hpp:239-294).
Force bar: relative RMS 1e-5.  A MAC decision that differed from the reference's would move
one particle's force by the truncation error of one cell (1e-4 .. 1e-3 relative), so the
per-particle walks are also held to a per-particle maximum of 2e-5.
"""
import numpy as np
import pytest

from conftest import f64, plummer_f32

pytestmark = pytest.mark.gpu

REL_RMS_TOL = 1e-5
PER_PARTICLE_TOL = 2e-5


def _per_particle_max(a, ref):
    return float(np.max(np.linalg.norm(a - ref, axis=1) / np.linalg.norm(ref, axis=1)))


def _amag64(a32):
    a = np.asarray(a32, np.float64)
    return np.sqrt((a[:, 0] * a[:, 0] + a[:, 1] * a[:, 1]) + a[:, 2] * a[:, 2])


# ------------------------------------------------------------------------- spline softening
def test_spline_direct_and_two_body(gpu_ctx_factory):
    import oracle
    import pynbg
    n, h = 8192, 0.05
    s = plummer_f32(n, 0)
    x, y, z, m = f64(s)
    ctx = gpu_ctx_factory(n)
    ctx.upload(s["x"], s["y"], s["z"], s["m"])
    ctx.accel(pynbg.make_params("direct", G=1.0, eps=h, softening="spline"))
    a = ctx.accelerations()
    ref = oracle.accel_direct(x, y, z, m, 1.0, h * h, spline=True)
    err = oracle.rel_rms(a, ref)
    print(f"spline direct N={n} h={h} rel_rms={err:.3e} max={_per_particle_max(a, ref):.3e}")
    assert err < REL_RMS_TOL
    # two bodies: Newtonian beyond 2h, chi(q)/r^2 inside, equal and opposite
    for r in (0.03, 0.07, 0.13, 0.5):
        ctx2 = gpu_ctx_factory(2)
        ctx2.upload(np.array([0.0, r], np.float32), np.zeros(2, np.float32), np.zeros(2, np.float32),
                    np.ones(2, np.float32))
        for solver in ("direct", "barnes_hut"):
            ctx2.accel(pynbg.make_params(solver, G=1.0, eps=h, softening="spline"))
            a2 = ctx2.accelerations().astype(np.float64)
            r32 = float(np.float32(r))
            ref2 = oracle.accel_direct(np.array([0.0, r32]), np.zeros(2), np.zeros(2), np.ones(2), 1.0, h * h,
                                       spline=True)
            np.testing.assert_allclose(a2, ref2, rtol=2e-6, atol=1e-12)


@pytest.mark.parametrize("multipole", ["monopole", "quadrupole"])
def test_spline_grouped_walk_forces_and_lists(gpu_ctx_factory, multipole):
    import oracle
    import pynbg
    n, h = 65536, 0.02
    s = plummer_f32(n, 0)
    x, y, z, m = f64(s)
    ctx = gpu_ctx_factory(n)
    ctx.upload(s["x"], s["y"], s["z"], s["m"])
    st = ctx.accel(pynbg.make_params("barnes_hut", multipole, G=1.0, eps=h, theta=0.5, softening="spline",
                                     count=True))
    a = ctx.accelerations()
    T = oracle.Tree(x, y, z, m, quad=(multipole == "quadrupole"))
    ref, nfar, nnear = T.accel_bh(1.0, h * h, 0.5, counts=True, spline=True)
    err = oracle.rel_rms(a, ref)
    g = ctx.tree_export()
    gfar, gnear = ctx.interaction_counts(g["n_nodes"])
    leaves = g["child_count"] == 0
    mism = int(np.sum(gfar[leaves] != nfar[leaves])) + int(np.sum(gnear[leaves] != nnear[leaves]))
    print(f"spline BH {multipole} rel_rms={err:.3e} list mismatches={mism} walk_ms={st.ms_walk:.3f}")
    assert err < REL_RMS_TOL and mism == 0
    # the tuned production kernel (no counting) gives the same forces
    ctx.accel(pynbg.make_params("barnes_hut", multipole, G=1.0, eps=h, theta=0.5, softening="spline"))
    assert oracle.rel_rms(ctx.accelerations(), ref) < REL_RMS_TOL


def test_spline_energy_and_steps(gpu_ctx_factory):
    import oracle
    import pynbg
    n, h, dt = 4096, 0.03, 1e-3
    s = plummer_f32(n, 1)
    st64 = {k: s[k].astype(np.float64) for k in s}
    ctx = gpu_ctx_factory(n)
    ctx.upload(s["x"], s["y"], s["z"], s["m"], s["vx"], s["vy"], s["vz"])
    p = pynbg.make_params("barnes_hut", G=1.0, eps=h, theta=0.5, dt=dt, softening="spline")
    d0 = ctx.diagnostics(p)
    e0 = oracle.total_energy(st64, 1.0, h * h, spline=True)
    print(f"spline energy gpu {d0['energy']:.15e} oracle {e0:.15e}")
    assert d0["energy"] == pytest.approx(e0, rel=1e-10)
    ctx.accel(p)
    ctx.step(p, 20)
    g = ctx.download()
    a0 = oracle.Tree(st64["x"], st64["y"], st64["z"], st64["m"], quad=True).accel_bh(1.0, h * h, 0.5, spline=True)
    o, _, _, t = oracle.steps(st64, a0, 1.0, h * h, 0.5, dt, 20, spline=True)
    pos_g = np.stack([g["x"], g["y"], g["z"]], 1).astype(np.float64)
    pos_o = np.stack([o["x"], o["y"], o["z"]], 1)
    assert oracle.rel_rms(pos_g, pos_o) < 1e-6
    d1 = ctx.diagnostics(p)
    assert abs(d1["energy"] - d0["energy"]) / abs(d0["energy"]) < 1e-4


# ------------------------------------------------------------------------- per-particle walks
@pytest.mark.parametrize("spline", [False, True])
def test_per_particle_geometric_walk(gpu_ctx_factory, spline):
    import oracle
    import pynbg
    n, eps = 32768, 0.01
    s = plummer_f32(n, 2)
    x, y, z, m = f64(s)
    soft = "spline" if spline else "plummer"
    ctx = gpu_ctx_factory(n)
    ctx.upload(s["x"], s["y"], s["z"], s["m"])
    st = ctx.accel(pynbg.make_params("barnes_hut", G=1.0, eps=eps, theta=0.5, per_particle=True, softening=soft,
                                     count=True))
    a = ctx.accelerations()
    T = oracle.Tree(x, y, z, m, quad=True)
    ref, n_m2p, n_p2p = T.accel_bh_per_particle(1.0, eps * eps, 0.5, spline=spline, counts=True)
    err, worst = oracle.rel_rms(a, ref), _per_particle_max(a, ref)
    print(f"per-particle walk ({soft}) rel_rms={err:.3e} max={worst:.3e} walk_ms={st.ms_walk:.3f} "
          f"m2p {st.n_m2p}/{n_m2p} p2p {st.n_p2p}/{n_p2p}")
    assert err < REL_RMS_TOL and worst < PER_PARTICLE_TOL
    # the interaction totals of every target's list are the reference's
    assert st.n_m2p == n_m2p and st.n_p2p == n_p2p
    # theta = 0: every leaf exact, the direct sum: n - 1 P2P sources per target, no M2P
    st0 = ctx.accel(pynbg.make_params("barnes_hut", G=1.0, eps=eps, theta=0.0, per_particle=True, softening=soft,
                                      count=True))
    assert st0.n_p2p == n * (n - 1) and st0.n_m2p == 0
    a0 = ctx.accelerations()
    ctx.accel(pynbg.make_params("direct", G=1.0, eps=eps, softening=soft))
    assert oracle.rel_rms(a0, ctx.accelerations()) < 2e-6


@pytest.mark.parametrize("alpha", [0.001, 0.004])
def test_relative_mac_walk(gpu_ctx_factory, alpha):
    import oracle
    import pynbg
    n, eps = 32768, 0.01
    s = plummer_f32(n, 3)
    x, y, z, m = f64(s)
    ctx = gpu_ctx_factory(n)
    ctx.upload(s["x"], s["y"], s["z"], s["m"])
    prm = pynbg.make_params("barnes_hut", G=1.0, eps=eps, theta=0.5, mac="relative", alpha=alpha)
    # no history: the geometric per-particle walk (System::SetParticles zeroes aOld)
    ctx.accel(prm)
    a_first = ctx.accelerations()
    ctx.upload(s["x"], s["y"], s["z"], s["m"])
    ctx.accel(pynbg.make_params("barnes_hut", G=1.0, eps=eps, theta=0.5, per_particle=True))
    assert oracle.rel_rms(a_first, ctx.accelerations()) < 1e-7
    # with history: aOld = |a| of the accelerations the context holds (a FP32 direct sum here)
    ctx.accel(pynbg.make_params("direct", G=1.0, eps=eps))
    a_prev = ctx.accelerations()
    st = ctx.accel(pynbg.make_params("barnes_hut", G=1.0, eps=eps, theta=0.5, mac="relative", alpha=alpha,
                                     count=True))
    a = ctx.accelerations()
    T = oracle.Tree(x, y, z, m, quad=True)
    ref, n_m2p, n_p2p = T.accel_bh_per_particle(1.0, eps * eps, 0.5, relative=True, alpha=alpha,
                                                a_old=_amag64(a_prev), counts=True)
    assert st.n_m2p == n_m2p and st.n_p2p == n_p2p
    # the production (non-counting) kernel on the same history
    ctx.upload_accel(a_prev[:, 0], a_prev[:, 1], a_prev[:, 2])
    st = ctx.accel(prm)
    a = ctx.accelerations()
    err, worst = oracle.rel_rms(a, ref), _per_particle_max(a, ref)
    exact = oracle.accel_direct(x, y, z, m, 1.0, eps * eps)
    print(f"relative MAC alpha={alpha} rel_rms vs oracle {err:.3e} max {worst:.3e}; "
          f"vs direct {oracle.rel_rms(a, exact):.3e}; walk_ms={st.ms_walk:.3f}")
    assert err < REL_RMS_TOL and worst < PER_PARTICLE_TOL


def test_bh_targets_walk(gpu_ctx_factory):
    import oracle
    import pynbg
    n, eps = 32768, 0.01
    s = plummer_f32(n, 4)
    x, y, z, m = f64(s)
    rng = np.random.default_rng(11)
    tg = rng.choice(n, 3000, replace=False)
    ctx = gpu_ctx_factory(n)
    ctx.upload(s["x"], s["y"], s["z"], s["m"])
    ctx.accel(pynbg.make_params("direct", G=1.0, eps=eps))
    before = ctx.accelerations()
    a = ctx.bh_targets(pynbg.make_params("barnes_hut", G=1.0, eps=eps, theta=0.5), tg)
    T = oracle.Tree(x, y, z, m, quad=True)
    ref = T.accel_bh_targets(1.0, eps * eps, 0.5, tg)
    err, worst = oracle.rel_rms(a, ref), _per_particle_max(a, ref)
    print(f"targets walk rel_rms={err:.3e} max={worst:.3e}")
    assert err < REL_RMS_TOL and worst < PER_PARTICLE_TOL
    # the context's accelerations are untouched
    np.testing.assert_array_equal(ctx.accelerations(), before)


# ------------------------------------------------------------------------- adaptive dt
def test_adaptive_dt_is_the_reference_rule_and_graph_equals_eager(gpu_ctx_factory):
    import oracle
    import pynbg
    n, eps, eta, dt0 = 4096, 0.02, 0.03, 1e-2
    s = plummer_f32(n, 5)
    runs = []
    for no_graph in (True, False):
        ctx = gpu_ctx_factory(n)
        ctx.upload(s["x"], s["y"], s["z"], s["m"], s["vx"], s["vy"], s["vz"])
        p = pynbg.make_params("barnes_hut", G=1.0, eps=eps, theta=0.5, dt=dt0, adaptive=True, eta=eta,
                              no_graph=no_graph)
        ctx.accel(p)
        a_t = ctx.accelerations()
        st = ctx.step(p, 1)
        want = oracle.adaptive_dt(a_t.astype(np.float64), eta, eps, dt0)
        print(f"adaptive dt gpu {st.dt_last!r} oracle rule {want!r} (graph={not no_graph})")
        assert st.dt_last == want and st.t_advanced == want and want < dt0
        st5 = ctx.step(p, 5)
        d = ctx.download()
        runs.append((d, st5.t_advanced))
    for k in ("x", "y", "z", "vx", "vy", "vz"):
        np.testing.assert_array_equal(runs[0][0][k], runs[1][0][k])
    assert runs[0][1] == runs[1][1]


def test_adaptive_steps_follow_the_oracle(gpu_ctx_factory):
    import oracle
    import pynbg
    n, eps, eta, dt0 = 2048, 0.02, 0.005, 5e-3
    s = plummer_f32(n, 6)
    st64 = {k: s[k].astype(np.float64) for k in s}
    ctx = gpu_ctx_factory(n)
    ctx.upload(s["x"], s["y"], s["z"], s["m"], s["vx"], s["vy"], s["vz"])
    p = pynbg.make_params("direct", G=1.0, eps=eps, dt=dt0, adaptive=True, eta=eta)
    ctx.accel(p)
    a0 = ctx.accelerations().astype(np.float64)
    st = ctx.step(p, 10)
    g = ctx.download()
    o, _, _, t = oracle.steps(st64, a0, 1.0, eps * eps, 0.5, dt0, 10, solver="direct", adaptive=True, eta=eta)
    print(f"adaptive 10 steps: t gpu {st.t_advanced:.9e} oracle {t:.9e} (nominal {10 * dt0:.3e})")
    assert st.t_advanced == pytest.approx(t, rel=1e-5) and t < 0.9 * 10 * dt0
    pos_g = np.stack([g["x"], g["y"], g["z"]], 1).astype(np.float64)
    pos_o = np.stack([o["x"], o["y"], o["z"]], 1)
    assert oracle.rel_rms(pos_g, pos_o) < 1e-6


def test_adaptive_eccentric_orbit_energy(gpu_ctx_factory):
    """SelfTest.cpp:391-455 on the device: e = 0.9 two-body orbit, direct sum,
    eps = 1e-4, eta = 0.015, three periods; energy drift within 1e-3."""
    import pynbg
    e, r0 = 0.9, 1.0
    a_sm = r0 / (1 + e)
    vrel = np.sqrt(2.0 * (2.0 / r0 - 1.0 / a_sm))
    period = 2 * np.pi * np.sqrt(a_sm ** 3 / 2.0)
    ctx = gpu_ctx_factory(2)
    f = np.float32
    ctx.upload(np.array([-0.5, 0.5], f), np.zeros(2, f), np.zeros(2, f), np.ones(2, f), np.zeros(2, f),
               np.array([-vrel / 2, vrel / 2], f), np.zeros(2, f))
    p = pynbg.make_params("direct", G=1.0, eps=1e-4, dt=period / 600.0, adaptive=True, eta=0.015)
    ctx.accel(p)
    e0 = ctx.diagnostics(p)["energy"]
    t, emin, emax, steps, dts = 0.0, e0, e0, 0, []
    while t < 3 * period:
        st = ctx.step(p, 32)
        t += st.t_advanced
        steps += 32
        dts.append(st.dt_last)
        en = ctx.diagnostics(p)["energy"]
        emin, emax = min(emin, en), max(emax, en)
    drift = (emax - emin) / abs(0.5 * (emax + emin))
    print(f"adaptive e=0.9: {steps} steps, dt {min(dts):.2e}..{max(dts):.2e}, energy drift {drift:.2e}")
    assert drift < 1e-3
    assert min(dts) < 0.1 * max(dts)


# ------------------------------------------------------------------------- block steps
def _cloud_and_binary_f32():
    """SelfTest.cpp:829-852: a 200-body Plummer cloud shifted to x = -8 and a tight fast
    binary at x = +8 (separation 0.08)."""
    import oracle
    c = oracle.plummer(200, 12)
    s = {k: np.zeros(202) for k in ("x", "y", "z", "vx", "vy", "vz", "m")}
    for k in s:
        s[k][:200] = c[k]
    s["x"][:200] -= 8.0
    s["x"][200], s["x"][201] = 7.96, 8.04
    vt = 0.5 * np.sqrt(2.0 / 0.08)
    s["vy"][200], s["vy"][201] = -vt, vt
    s["m"][200] = s["m"][201] = 1.0
    return {k: v.astype(np.float32) for k, v in s.items()}


@pytest.mark.parametrize("solver", ["direct", "barnes_hut"])
def test_block_step_rungs_evals_and_state(gpu_ctx_factory, solver):
    import oracle
    import pynbg
    s = _cloud_and_binary_f32()
    st64 = {k: s[k].astype(np.float64) for k in s}
    eps, base, eta = 0.01, 5e-3, 0.02
    ctx = gpu_ctx_factory(202)
    ctx.upload(s["x"], s["y"], s["z"], s["m"], s["vx"], s["vy"], s["vz"])
    p = pynbg.make_params(solver, G=1.0, eps=eps, theta=0.5, dt=base, block=True, eta=eta, block_max_rung=8)
    ctx.accel(p)
    a0 = ctx.accelerations().astype(np.float64)
    st = ctx.step(p, 1)
    r_gpu = ctx.rungs()
    g = ctx.download()
    o, _, r_o, rmax_o, ev_o = oracle.block_steps(st64, a0, 1.0, eps * eps, 0.5, base, 1, eta=eta, max_rung=8,
                                                 solver=solver)
    print(f"block {solver}: rung_max gpu {st.rung_max} oracle {rmax_o}; evals gpu {st.particle_evals} "
          f"oracle {ev_o}")
    np.testing.assert_array_equal(r_gpu, r_o)
    assert st.rung_max == rmax_o > 2 and st.particle_evals == ev_o
    assert st.t_advanced == base and st.dt_last == base
    pos_g = np.stack([g["x"], g["y"], g["z"]], 1).astype(np.float64)
    pos_o = np.stack([o["x"], o["y"], o["z"]], 1)
    vel_g = np.stack([g["vx"], g["vy"], g["vz"]], 1).astype(np.float64)
    vel_o = np.stack([o["vx"], o["vy"], o["vz"]], 1)
    assert oracle.rel_rms(pos_g, pos_o) < 1e-6
    assert oracle.rel_rms(vel_g, vel_o) < 1e-4


def test_block_scheme_energy_and_work(gpu_ctx_factory):
    """SelfTest.cpp:829-892 on the device: block steps to t = 2 keep |dE/E| < 1e-2 and use at
    least 2x fewer particle evaluations than global adaptive steps to the same time."""
    import pynbg
    s = _cloud_and_binary_f32()
    T = 2.0
    res = {}
    for mode in ("block", "adaptive"):
        ctx = gpu_ctx_factory(202)
        ctx.upload(s["x"], s["y"], s["z"], s["m"], s["vx"], s["vy"], s["vz"])
        p = pynbg.make_params("direct", G=1.0, eps=0.01, dt=5e-3, eta=0.02, block=(mode == "block"),
                              adaptive=(mode == "adaptive"), block_max_rung=8)
        ctx.accel(p)
        e0 = ctx.diagnostics(p)["energy"]
        t, evals = 0.0, 0
        while t < T:
            st = ctx.step(p, 10)
            t += st.t_advanced
            evals += st.particle_evals
        drift = abs((ctx.diagnostics(p)["energy"] - e0) / e0)
        res[mode] = (drift, evals, t)
    print(f"cloud + binary to t={T}: block drift {res['block'][0]:.2e} evals {res['block'][1]}; "
          f"adaptive drift {res['adaptive'][0]:.2e} evals {res['adaptive'][1]}")
    assert res["block"][0] < 1e-2
    assert res["adaptive"][1] >= 2 * res["block"][1]


def test_block_bh_cold_collapse_energy(gpu_ctx_factory):
    """SelfTest.cpp:895-923 on the device: cold Plummer sphere N=1200 seed 4, eps 0.02,
    theta 0.5, dt 3e-3, eta 0.03, max rung 5, 150 coarse steps: |dE/E| < 0.05."""
    import oracle
    import pynbg
    c = oracle.plummer(1200, 4)
    f = {k: c[k].astype(np.float32) for k in ("x", "y", "z", "m")}
    ctx = gpu_ctx_factory(1200)
    ctx.upload(f["x"], f["y"], f["z"], f["m"])
    p = pynbg.make_params("barnes_hut", G=1.0, eps=0.02, theta=0.5, dt=3e-3, eta=0.03, block=True,
                          block_max_rung=5)
    ctx.accel(p)
    e0 = ctx.diagnostics(p)["energy"]
    deepest = 0
    for _ in range(15):
        st = ctx.step(p, 10)
        deepest = max(deepest, st.rung_max)
    drift = abs((ctx.diagnostics(p)["energy"] - e0) / e0)
    print(f"BH block cold collapse: drift {drift:.2e}, deepest rung {deepest}")
    assert drift < 0.05


# ------------------------------------------------------------------------- edge cases
def test_relative_mac_steps_graph_equals_eager_and_follow_the_oracle(gpu_ctx_factory):
    """System::Step with the relative MAC: aOld is |a| captured after every step
    (System.hpp:302-310); the graph-replayed steps equal eager ones bit for bit and follow
    the oracle's FP64 steps."""
    import oracle
    import pynbg
    n, eps, dt = 4096, 0.02, 2e-3
    s = plummer_f32(n, 8)
    st64 = {k: s[k].astype(np.float64) for k in s}
    runs = []
    for no_graph in (True, False):
        ctx = gpu_ctx_factory(n)
        ctx.upload(s["x"], s["y"], s["z"], s["m"], s["vx"], s["vy"], s["vz"])
        p = pynbg.make_params("barnes_hut", G=1.0, eps=eps, theta=0.5, dt=dt, mac="relative", alpha=0.002,
                              no_graph=no_graph)
        ctx.accel(p)   # Prime: no history yet -> the geometric per-particle walk
        a0 = ctx.accelerations()
        ctx.step(p, 6)
        runs.append((ctx.download(), a0))
    for k in ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az"):
        np.testing.assert_array_equal(runs[0][0][k], runs[1][0][k])
    g, a0 = runs[1]
    T = oracle.Tree(st64["x"], st64["y"], st64["z"], st64["m"], quad=True)
    a0_ref = T.accel_bh_per_particle(1.0, eps * eps, 0.5)
    assert oracle.rel_rms(a0, a0_ref) < REL_RMS_TOL
    aold0 = np.sqrt((a0_ref[:, 0] ** 2 + a0_ref[:, 1] ** 2) + a0_ref[:, 2] ** 2)
    o, a_o, _, _ = oracle.steps(st64, a0_ref, 1.0, eps * eps, 0.5, dt, 6, relative=True, alpha=0.002, a_old=aold0)
    pos_g = np.stack([g["x"], g["y"], g["z"]], 1).astype(np.float64)
    pos_o = np.stack([o["x"], o["y"], o["z"]], 1)
    a_g = np.stack([g["ax"], g["ay"], g["az"]], 1).astype(np.float64)
    print(f"relative steps: pos rel {oracle.rel_rms(pos_g, pos_o):.2e}, a rel {oracle.rel_rms(a_g, a_o):.2e}")
    assert oracle.rel_rms(pos_g, pos_o) < 1e-6
    assert oracle.rel_rms(a_g, a_o) < 1e-4


@pytest.mark.parametrize("mode", ["per_particle", "relative"])
def test_per_particle_walks_on_a_batch(gpu_ctx_factory, mode):
    """Independent systems in one upload: every system's per-particle walk sees only its
    own tree (several systems can share a wave of targets)."""
    import oracle
    import pynbg
    off, st, _ = pynbg.ensemble(10, 2, 100, 3000)
    n = int(off[-1])
    eps = 0.01
    ctx = gpu_ctx_factory(n)
    ctx.upload_batch(off, st["x"], st["y"], st["z"], st["m"])
    kw = dict(per_particle=True) if mode == "per_particle" else dict(mac="relative", alpha=0.003)
    if mode == "relative":
        ctx.accel(pynbg.make_params("direct", G=1.0, eps=eps))
        a_prev = ctx.accelerations()
    ctx.accel(pynbg.make_params("barnes_hut", G=1.0, eps=eps, theta=0.6, **kw))
    a = ctx.accelerations().astype(np.float64)
    worst = 0.0
    for k in range(len(off) - 1):
        sl = slice(off[k], off[k + 1])
        x, y, z, m = (st[q][sl].astype(np.float64) for q in ("x", "y", "z", "m"))
        T = oracle.Tree(x, y, z, m, quad=True)
        if mode == "relative":
            ref = T.accel_bh_per_particle(1.0, eps * eps, 0.6, relative=True, alpha=0.003,
                                          a_old=_amag64(a_prev[sl]))
        else:
            ref = T.accel_bh_per_particle(1.0, eps * eps, 0.6)
        worst = max(worst, _per_particle_max(a[sl], ref))
    print(f"batch of {len(off) - 1} systems ({mode}): worst per-particle error {worst:.2e}")
    assert worst < PER_PARTICLE_TOL


def test_per_particle_walk_with_deep_keys(gpu_ctx_factory):
    """A cluster needing the 40-level key path (coincident points, a tight clump) under the
    per-particle walk and block steps."""
    import oracle
    import pynbg
    rng = np.random.default_rng(3)
    x = rng.uniform(-1, 1, 3000)
    y = rng.uniform(-1, 1, 3000)
    z = rng.uniform(-1, 1, 3000)
    x[:40] = 0.25 + 1e-9 * rng.standard_normal(40)
    y[:40] = -0.5
    z[:40] = 0.125
    m = np.full(3000, 1.0 / 3000)
    f = {k: v.astype(np.float32) for k, v in dict(x=x, y=y, z=z, m=m).items()}
    ctx = gpu_ctx_factory(3000)
    ctx.upload(f["x"], f["y"], f["z"], f["m"])
    st = ctx.accel(pynbg.make_params("barnes_hut", G=1.0, eps=0.02, theta=0.5, per_particle=True))
    assert st.deep_keys == 1
    a = ctx.accelerations()
    T = oracle.Tree(*(f[k].astype(np.float64) for k in ("x", "y", "z", "m")), quad=True)
    ref = T.accel_bh_per_particle(1.0, 0.02 ** 2, 0.5)
    assert _per_particle_max(a, ref) < PER_PARTICLE_TOL
    tg = np.arange(0, 3000, 7)
    at = ctx.bh_targets(pynbg.make_params("barnes_hut", G=1.0, eps=0.02, theta=0.5), tg)
    assert _per_particle_max(at, ref[tg]) < PER_PARTICLE_TOL


def test_mode_argument_errors(gpu_ctx_factory):
    import pynbg
    s = plummer_f32(256, 0)
    ctx = gpu_ctx_factory(256)
    ctx.upload(s["x"], s["y"], s["z"], s["m"])
    ctx.learn_set_weights(np.zeros(pynbg.LEARN_NPARAM, np.float32))
    bad = [
        (dict(softening="spline", eps=0.0), "spline"),
        (dict(mac="relative", alpha=-1.0), "alpha"),
        (dict(adaptive=True, eta=0.0), "eta"),
        (dict(block=True, block_max_rung=30), "block_max_rung"),
        (dict(learned=True, softening="spline"), "learned"),
        (dict(learned=True, mac="relative"), "learned"),
        (dict(learned=True, block=True), "learned"),
    ]
    for kw, word in bad:
        base = dict(G=1.0, eps=0.01, theta=0.5)
        base.update(kw)
        with pytest.raises(pynbg.NbgError, match=word):
            ctx.accel(pynbg.make_params("barnes_hut", **base))
    # a target index out of range is refused before any device work
    with pytest.raises(pynbg.NbgError, match="target"):
        ctx.bh_targets(pynbg.make_params("barnes_hut", G=1.0, eps=0.01), [0, 256])
    # the context still works
    ctx.accel(pynbg.make_params("barnes_hut", G=1.0, eps=0.01, theta=0.5, softening="spline"))
    assert np.all(np.isfinite(ctx.accelerations()))
