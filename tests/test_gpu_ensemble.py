"""Batched ensemble of independent clusters (no reference counterpart: the reference runs
one System at a time) and the autotuner of the force kernels' tile sizes.

Each cluster of a batch must get exactly what a one-system run of the reference's CPU path
gives it: per-cluster forces against the oracle's FP64 grouped walk
(projects/nbody_core/src/Solvers.cpp:279-355) within the 1e-5 relative RMS bar, and
per-cluster FP64 diagnostics (System.hpp:119-193) against the oracle's.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _small_ensemble(nclusters=12, seed=0, nmin=256, nmax=4096):
    import pynbg
    return pynbg.ensemble(nclusters, seed, nmin, nmax)


def _rel_rms(a, ref):
    return float(np.sqrt(np.sum((a - ref) ** 2) / np.sum(ref ** 2)))


def test_ensemble_forces_match_per_cluster_oracle(gpu_ctx_factory):
    import oracle
    import pynbg
    off, st, kinds = _small_ensemble()
    ctx = gpu_ctx_factory(int(off[-1]))
    ctx.upload_batch(off, st["x"], st["y"], st["z"], st["m"])
    stats = ctx.accel(pynbg.make_params("barnes_hut", G=1.0, eps=0.01, theta=0.6))
    a = ctx.accelerations().astype(np.float64)
    worst = 0.0
    for k in range(len(off) - 1):
        sl = slice(off[k], off[k + 1])
        x, y, z, m = (st[q][sl].astype(np.float64) for q in ("x", "y", "z", "m"))
        ref = oracle.Tree(x, y, z, m, quad=True).accel_bh(1.0, 0.01 ** 2, 0.6)
        worst = max(worst, _rel_rms(a[sl], ref))
    print(f"clusters {len(off) - 1} N {off[-1]} nodes {stats.n_nodes} worst rel_rms {worst:.2e}")
    assert worst < 1e-5


def test_autotune_picks_a_candidate_and_keeps_results(gpu_ctx_factory):
    import pynbg
    off, st, _ = _small_ensemble(16, 3, 1024, 16384)
    ctx = gpu_ctx_factory(int(off[-1]))
    ctx.upload_batch(off, st["x"], st["y"], st["z"], st["m"])
    p = pynbg.make_params("barnes_hut", G=1.0, eps=0.01, theta=0.6)
    ctx.set_tuning(0, 8, 4)
    ctx.accel(p)
    base = ctx.accelerations().astype(np.float64)
    results = {}
    for wk, nt in ((0, 4), (0, 16), (1, 8)):
        ctx.set_tuning(wk, nt, 4)
        ctx.accel(p)
        results[(wk, nt)] = _rel_rms(ctx.accelerations().astype(np.float64), base)
    best, ms = ctx.autotune(p, reps=2)
    print("candidate ms", ms, "best", best.as_dict(), "diffs", results)
    # same interaction lists, different summation order: rounding-level differences only
    assert all(v < 1e-6 for v in results.values()), results
    assert len(ms) == 4 and np.all(ms > 0)
    assert best.as_dict() == ctx.tuning().as_dict()
    assert (best.walk_kernel, best.near_tile) in ((0, 4), (0, 8), (0, 16), (1, 8))
    # the direct-sum tile
    pd = pynbg.make_params("direct", G=1.0, eps=0.01)
    bd, msd = ctx.autotune(pd, reps=1)
    assert bd.direct_tile in (2, 4, 8) and len(msd) == 3
    with pytest.raises(pynbg.NbgError, match="near_tile"):
        ctx.set_tuning(0, 5, 4)


def test_ensemble_steps_graph_equals_eager_and_conserve_energy(gpu_ctx_factory):
    import oracle
    import pynbg
    off, st, _ = _small_ensemble(8, 1, 512, 2048)
    n = int(off[-1])
    runs = []
    for no_graph in (True, False):
        ctx = gpu_ctx_factory(n)
        ctx.upload_batch(off, st["x"], st["y"], st["z"], st["m"], st["vx"], st["vy"], st["vz"])
        p = pynbg.make_params("barnes_hut", G=1.0, eps=0.01, theta=0.6, dt=1e-3, no_graph=no_graph)
        e0 = [ctx.diagnostics_system(p, k)["energy"] for k in range(len(off) - 1)]
        ctx.accel(p)
        ctx.step(p, 50)
        e1 = [ctx.diagnostics_system(p, k)["energy"] for k in range(len(off) - 1)]
        d = ctx.download()
        runs.append((np.stack([d["x"], d["y"], d["z"]], 1), np.array(e0), np.array(e1)))
    np.testing.assert_array_equal(runs[0][0], runs[1][0])
    drift = np.abs(runs[1][2] - runs[1][1]) / np.abs(runs[1][1])
    print("per-cluster |dE/E| after 50 steps:", np.array2string(drift, precision=2))
    assert drift.max() < 1e-3
    # per-system FP64 energy of the initial state against the oracle
    for k in (0, 3):
        sl = slice(off[k], off[k + 1])
        stt = {q: st[q][sl].astype(np.float64) for q in ("x", "y", "z", "vx", "vy", "vz", "m")}
        ke_pe = oracle.total_energy(stt, 1.0, 0.01 ** 2)
        assert runs[1][1][k] == pytest.approx(ke_pe, rel=1e-9)
