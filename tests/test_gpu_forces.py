"""GPU force parity against the FP64 oracle.

Bar: accelerations within a relative RMS of 1e-5 of the reference's FP64 CPU path on
identical FP32-rounded inputs; Barnes-Hut interaction lists
identical per leaf group (CollectGroupInteractions, projects/nbody_core/src/Solvers.cpp:248-277).
rel_rms = sqrt(sum |a - a_ref|^2 / sum |a_ref|^2).
"""
import numpy as np
import pytest

from conftest import f64, plummer_f32

pytestmark = pytest.mark.gpu

REL_RMS_TOL = 1e-5


def test_direct_fp32_vs_fp64_oracle_131072(gpu_ctx_factory):
    """Tiled FP32 direct sum on the 131,072-body Plummer sample, eps=0.005 (C2 check)."""
    import oracle
    import pynbg
    s = plummer_f32(131072, 0)
    x, y, z, m = f64(s)
    ctx = gpu_ctx_factory(131072)
    ctx.upload(s["x"], s["y"], s["z"], s["m"])
    ctx.accel(pynbg.make_params("direct", G=1.0, eps=0.005))
    a = ctx.accelerations()
    ref = oracle.accel_direct(x, y, z, m, 1.0, 0.005 ** 2)
    err = oracle.rel_rms(a, ref)
    print(f"direct N=131072 rel_rms={err:.3e}")
    assert err < REL_RMS_TOL


def test_direct_two_body_closed_form(gpu_ctx_factory):
    import pynbg
    # a = G m d / (d^2 + eps^2)^1.5 on each body, equal and opposite momenta
    x = np.array([-0.5, 0.5], np.float32)
    z0 = np.zeros(2, np.float32)
    m = np.array([1.0, 3.0], np.float32)
    ctx = gpu_ctx_factory(2)
    ctx.upload(x, z0, z0, m)
    for solver in ("direct", "barnes_hut"):
        ctx.accel(pynbg.make_params(solver, G=2.0, eps=0.1, theta=0.5))
        a = ctx.accelerations().astype(np.float64)
        f = 1.0 / (1.0 + 0.01) ** 1.5
        np.testing.assert_allclose(a[0], [2.0 * 3.0 * f, 0, 0], rtol=2e-6, atol=1e-12)
        np.testing.assert_allclose(a[1], [-2.0 * 1.0 * f, 0, 0], rtol=2e-6, atol=1e-12)


@pytest.mark.parametrize("multipole", ["monopole", "quadrupole"])
def test_bh_c1_plummer_65536(gpu_ctx_factory, multipole):
    """C1: Plummer N=65,536 seed 0, eps=0.01, theta=0.5 -- forces and interaction lists."""
    import oracle
    import pynbg
    s = plummer_f32(65536, 0)
    x, y, z, m = f64(s)
    ctx = gpu_ctx_factory(65536)
    ctx.upload(s["x"], s["y"], s["z"], s["m"])
    st = ctx.accel(pynbg.make_params("barnes_hut", multipole, G=1.0, eps=0.01, theta=0.5, count=True))
    a = ctx.accelerations()
    T = oracle.Tree(x, y, z, m, quad=(multipole == "quadrupole"))
    ref, nfar, nnear = T.accel_bh(1.0, 0.01 ** 2, 0.5, counts=True)
    err = oracle.rel_rms(a, ref)
    g = ctx.tree_export()
    gfar, gnear = ctx.interaction_counts(g["n_nodes"])
    leaves = g["child_count"] == 0
    far_mismatch = int(np.sum(gfar[leaves] != nfar[leaves]))
    near_mismatch = int(np.sum(gnear[leaves] != nnear[leaves]))
    print(f"BH {multipole} rel_rms={err:.3e} m2p={st.n_m2p} p2p={st.n_p2p} "
          f"list mismatches far={far_mismatch} near={near_mismatch} walk_ms={st.ms_walk:.3f}")
    assert err < REL_RMS_TOL
    assert far_mismatch == 0 and near_mismatch == 0
    counts = T.export()["count"]
    assert st.n_m2p == int(np.sum(nfar[leaves] * counts[leaves]))
    assert st.n_p2p == int(np.sum(nnear[leaves] * counts[leaves]))


def test_leapfrog_one_step_c1(gpu_ctx_factory):
    """C1 step: Prime, then one KDK step of dt=1e-3 (Integrator.cpp:9-39) vs the oracle."""
    import oracle
    import pynbg
    s = plummer_f32(65536, 0)
    p = pynbg.make_params("barnes_hut", "monopole", G=1.0, eps=0.01, theta=0.5, dt=1e-3)
    ctx = gpu_ctx_factory(65536)
    ctx.upload(s["x"], s["y"], s["z"], s["m"], s["vx"], s["vy"], s["vz"])
    ctx.accel(p)
    ctx.step(p, 1)
    d = ctx.download()
    st = {k: s[k].astype(np.float64) for k in s}
    T = oracle.Tree(st["x"], st["y"], st["z"], st["m"], quad=False)
    a0 = T.accel_bh(1.0, 1e-4, 0.5)
    s1, a1 = oracle.leapfrog(st, a0, 1.0, 1e-4, 0.5, 1e-3, 1, quad=False)
    for k in ("x", "y", "z"):
        np.testing.assert_allclose(d[k], s1[k], rtol=0, atol=2e-6 * (1 + np.abs(s1[k])).max())
    v_err = np.sqrt(sum(np.sum((d[k] - s1[k]) ** 2) for k in ("vx", "vy", "vz")) /
                    sum(np.sum(s1[k] ** 2) for k in ("vx", "vy", "vz")))
    assert v_err < 1e-6, v_err
    a_gpu = np.stack([d["ax"], d["ay"], d["az"]], 1)
    assert oracle.rel_rms(a_gpu, a1) < REL_RMS_TOL


def test_graph_replay_equals_eager_and_is_deterministic(gpu_ctx_factory):
    import pynbg
    s = plummer_f32(20000, 5)
    res = []
    for no_graph, timed in ((False, False), (True, False), (False, True)):
        p = pynbg.make_params("barnes_hut", "quadrupole", G=1.0, eps=0.01, theta=0.6, dt=1e-3, no_graph=no_graph,
                              stage_times=timed)
        ctx = gpu_ctx_factory(20000)
        ctx.upload(s["x"], s["y"], s["z"], s["m"], s["vx"], s["vy"], s["vz"])
        ctx.accel(p)
        st = ctx.step(p, 5)
        if timed:
            assert st.ms_walk > 0 and st.ms_sort > 0 and st.ms_integrate > 0
        d = ctx.download()
        res.append(np.stack([d[k] for k in ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az")]))
    np.testing.assert_array_equal(res[0], res[1])
    np.testing.assert_array_equal(res[0], res[2])


def test_fp64_diagnostics(gpu_ctx_factory):
    import oracle
    import pynbg
    s = plummer_f32(8192, 2)
    ctx = gpu_ctx_factory(8192)
    ctx.upload(s["x"], s["y"], s["z"], s["m"], s["vx"], s["vy"], s["vz"])
    p = pynbg.make_params("barnes_hut", G=1.0, eps=0.01)
    d = ctx.diagnostics(p)
    st = {k: s[k].astype(np.float64) for k in s}
    e = oracle.total_energy(st, 1.0, 1e-4)
    com, L, P = oracle.moments(st)
    np.testing.assert_allclose(d["energy"], e, rtol=1e-11)
    np.testing.assert_allclose(d["com"], com, rtol=0, atol=1e-14)
    np.testing.assert_allclose(d["L"], L, rtol=1e-10, atol=1e-15)
    np.testing.assert_allclose(d["P"], P, rtol=0, atol=1e-15)


def test_invalid_arguments_are_reported(gpu_ctx_factory):
    import pynbg
    ctx = gpu_ctx_factory(16)
    with pytest.raises(pynbg.NbgError, match="NBG_E_STATE"):
        ctx.accel(pynbg.make_params())
    x = np.zeros(32, np.float32)
    with pytest.raises(pynbg.NbgError, match="NBG_E_INVALID"):
        ctx.upload(x, x, x, x)
    ctx.upload(x[:4], x[:4], x[:4], x[:4] + 1)
    bad = pynbg.make_params(eps=0.1)
    bad.multipole = 7
    with pytest.raises(pynbg.NbgError, match="NBG_E_INVALID"):
        ctx.accel(bad)
