"""Parity at the configurations' full sizes.

* Plummer N = 1,048,576, seed 0, eps = 0.005, theta = 0.5, quadrupole (the 1M Barnes-Hut
  step): Morton keys, sort order, topology and FP64 cell geometry bit-exact against the
  oracle's serial-insertion tree (projects/nbody_core/src/Tree.cpp:33-254); per-leaf far
  and near interaction counts exact (CollectGroupInteractions, Solvers.cpp:248-277);
  accelerations within 1e-5 relative RMS of the grouped walk (Solvers.cpp:279-355); one
  KDK step of dt = 5e-4 (Integrator.cpp:9-39) against the oracle's.
* The same state with the learned far-field correction at theta = 0.8 (committed model,
  pairs with correction scale >= 1e-3): relative RMS <= 1e-3 against the oracle's FP64
  direct sum on 16,384 seeded random targets, and below plain theta = 0.8.
* The 512-cluster ensemble (sizes log-uniform in [1,024, 65,536], mt19937 seed 0,
  Plummer / King(W0 = 6) alternating, eps = 0.01, theta = 0.6): per-cluster forces of 16
  clusters spanning the size range, the largest included, against the oracle; FP64 energy
  of every cluster over 10 KDK steps of dt = 1e-3.
"""
import os

import numpy as np
import pytest

from conftest import f64, keys_at, plummer_f32

pytestmark = pytest.mark.gpu

N1M = 1 << 20
REL_RMS_TOL = 1e-5


@pytest.fixture(scope="module")
def plummer_1m():
    s = plummer_f32(N1M, 0)
    return s, f64(s)


def test_plummer_1m_tree_lists_forces_and_step(gpu_ctx_factory, plummer_1m):
    import oracle
    import pynbg
    from test_gpu_tree import _compare_trees
    s, (x, y, z, m) = plummer_1m
    eps2 = 0.005 ** 2
    ctx = gpu_ctx_factory(N1M)
    ctx.upload(s["x"], s["y"], s["z"], s["m"], s["vx"], s["vy"], s["vz"])
    st = ctx.accel(pynbg.make_params("barnes_hut", "quadrupole", G=1.0, eps=0.005, theta=0.5, count=True))
    a = ctx.accelerations().astype(np.float64)
    g = ctx.tree_export()
    gfar, gnear = ctx.interaction_counts(g["n_nodes"])
    print(f"device: nodes {g['n_nodes']} leaves {st.n_leaves} depth {st.max_depth} m2p {st.n_m2p} p2p {st.n_p2p}",
          flush=True)

    T = oracle.Tree(x, y, z, m, quad=True)
    o = T.export()
    keys_hi, _ = oracle.morton_keys(x, y, z, 21)
    assert st.key_levels == 13  # depth 12: 13-level keys, 5 sort passes
    np.testing.assert_array_equal(g["keys"], np.sort(keys_at(keys_hi, st.key_levels)))
    _compare_trees(g, o, N1M)
    assert st.deep_keys == 0 and st.max_depth == o["depth"].max()

    ref, nfar, nnear = T.accel_bh(1.0, eps2, 0.5, counts=True)
    leaves = g["child_count"] == 0
    far_mism = int(np.sum(gfar[leaves] != nfar[leaves]))
    near_mism = int(np.sum(gnear[leaves] != nnear[leaves]))
    err = oracle.rel_rms(a, ref)
    print(f"oracle: nodes {len(o['mass'])}; rel_rms {err:.3e}; list mismatches far {far_mism} near {near_mism}",
          flush=True)
    assert far_mism == 0 and near_mism == 0
    cnt = o["count"]
    assert st.n_m2p == int(np.sum(nfar[leaves] * cnt[leaves]))
    assert st.n_p2p == int(np.sum(nnear[leaves] * cnt[leaves]))
    assert err < REL_RMS_TOL

    # one KDK step from the primed state
    p = pynbg.make_params("barnes_hut", "quadrupole", G=1.0, eps=0.005, theta=0.5, dt=5e-4)
    ctx.step(p, 1)
    d = ctx.download()
    st64 = {k: s[k].astype(np.float64) for k in s}
    s1, a1 = oracle.leapfrog(st64, ref, 1.0, eps2, 0.5, 5e-4, 1, quad=True)
    for k in ("x", "y", "z"):
        np.testing.assert_allclose(d[k], s1[k], rtol=0, atol=2e-6 * (1 + np.abs(s1[k])).max())
    v_err = np.sqrt(sum(np.sum((d[k] - s1[k]) ** 2) for k in ("vx", "vy", "vz")) /
                    sum(np.sum(s1[k] ** 2) for k in ("vx", "vy", "vz")))
    a_err = oracle.rel_rms(np.stack([d["ax"], d["ay"], d["az"]], 1), a1)
    print(f"KDK step: velocity rel_rms {v_err:.2e}, a(t+dt) rel_rms {a_err:.2e}")
    assert v_err < 1e-6
    assert a_err < REL_RMS_TOL


def test_plummer_1m_learned_theta08_meets_target(gpu_ctx_factory, plummer_1m):
    import oracle
    import pynbg
    from pynbg import learned
    if not os.path.exists(learned.DEFAULT_WEIGHTS):
        pytest.skip("no trained model committed")
    s, (x, y, z, m) = plummer_1m
    eps2 = 0.005 ** 2
    targets = np.sort(np.random.default_rng(20240).choice(N1M, 16384, replace=False))
    ref = oracle.accel_direct_targets(x, y, z, m, 1.0, eps2, targets, fast=True)
    ctx = gpu_ctx_factory(N1M)
    ctx.upload(s["x"], s["y"], s["z"], s["m"])
    ctx.learn_set_weights(learned.load_weights())
    res, n_pairs = {}, 0
    for name, kw in (("plain 0.5", dict(theta=0.5)), ("plain 0.8", dict(theta=0.8)),
                     ("learned 0.8", dict(theta=0.8, learned=True, learn_band=0.0, learn_amin=1e-3))):
        st = ctx.accel(pynbg.make_params("barnes_hut", "quadrupole", G=1.0, eps=0.005, **kw))
        res[name] = oracle.rel_rms(ctx.accelerations()[targets].astype(np.float64), ref)
        if kw.get("learned"):
            n_pairs = st.n_learned
    print({k: f"{v:.3e}" for k, v in res.items()}, "corrected pairs", n_pairs)
    assert n_pairs > 0
    assert res["learned 0.8"] <= 1e-3
    assert res["learned 0.8"] < res["plain 0.8"]


def test_ensemble_512_forces_and_energy(gpu_ctx_factory):
    import oracle
    import pynbg
    off, st, kinds = pynbg.ensemble(512, 0, 1024, 65536, 6.0)
    sizes = np.diff(off)
    ntot = int(off[-1])
    ctx = gpu_ctx_factory(ntot)
    ctx.upload_batch(off, st["x"], st["y"], st["z"], st["m"], st["vx"], st["vy"], st["vz"])
    p = pynbg.make_params("barnes_hut", "quadrupole", G=1.0, eps=0.01, theta=0.6, dt=1e-3)
    ctx.accel(p)
    a = ctx.accelerations().astype(np.float64)
    by_size = np.argsort(sizes, kind="stable")
    pick = sorted(set(by_size[np.linspace(0, len(sizes) - 1, 16).round().astype(int)].tolist()))
    assert int(np.argmax(sizes)) in pick and int(np.argmin(sizes)) in pick
    worst = 0.0
    for k in pick:
        sl = slice(off[k], off[k + 1])
        xk, yk, zk, mk = (st[q][sl].astype(np.float64) for q in ("x", "y", "z", "m"))
        ref = oracle.Tree(xk, yk, zk, mk, quad=True).accel_bh(1.0, 1e-4, 0.6)
        e = oracle.rel_rms(a[sl], ref)
        worst = max(worst, e)
        assert e < REL_RMS_TOL, (k, int(sizes[k]), e)
    print(f"clusters {len(sizes)}, N {ntot} ({sizes.min()}..{sizes.max()}); "
          f"{len(pick)} checked, worst rel_rms {worst:.2e}", flush=True)

    nsys = len(sizes)
    e0 = np.array([ctx.diagnostics_system(p, k)["energy"] for k in range(nsys)])
    for k in (int(np.argmin(sizes)), 1):
        sl = slice(off[k], off[k + 1])
        stt = {q: st[q][sl].astype(np.float64) for q in ("x", "y", "z", "vx", "vy", "vz", "m")}
        assert e0[k] == pytest.approx(oracle.total_energy(stt, 1.0, 1e-4), rel=1e-9)
    ctx.step(p, 10)
    e1 = np.array([ctx.diagnostics_system(p, k)["energy"] for k in range(nsys)])
    drift = np.abs(e1 - e0) / np.abs(e0)
    print(f"|dE/E| over 10 steps: max {drift.max():.2e} median {np.median(drift):.2e}")
    assert np.all(np.isfinite(e1))
    assert drift.max() < 1e-4
