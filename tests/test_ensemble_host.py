"""CPU checks of the batched-ensemble layout (host C++ sampler, no GPU)."""
import numpy as np


def test_ensemble_sizes_follow_the_log_uniform_draw():
    """N_k = round(exp(U[ln 1024, ln 65536])), U from std::mt19937(0) through
    uniform_real_distribution<double> -- restated by the oracle's mt19937 sequence."""
    import ctypes
    import oracle
    import pynbg
    H = pynbg.load_host()
    n = 512
    sizes = np.empty(n, np.int64)
    kinds = np.empty(n, np.int32)
    total = H.ngh_ensemble_sizes(n, 0, 1024, 65536, sizes.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                 kinds.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    u = oracle.uniform_sequence(0, np.log(1024.0), np.log(65536.0), n)
    np.testing.assert_array_equal(sizes, np.clip(np.rint(np.exp(u)).astype(np.int64), 1024, 65536))
    np.testing.assert_array_equal(kinds, np.arange(n) % 2)
    assert total == sizes.sum()
    assert H.ngh_ensemble_sizes(0, 0, 1024, 65536, sizes.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), None) == -1


def test_ensemble_state_is_per_cluster_samplers():
    import pynbg
    off, st, kinds = pynbg.ensemble(6, 0, 64, 256)
    for k in range(6):
        ic = pynbg.ic_plummer(int(off[k + 1] - off[k]), k) if kinds[k] == 0 else \
            pynbg.ic_king(int(off[k + 1] - off[k]), k, w0=6.0)
        np.testing.assert_array_equal(st["x"][off[k]:off[k + 1]], ic["x"].astype(np.float32))
        assert abs(float(st["m"][off[k]:off[k + 1]].astype(np.float64).sum()) - 1.0) < 1e-5


def test_cluster_sampler_statistics():
    """Cluster<3> / ColdUniform (Scenarios.cpp:104-142, 201-214): a uniform unit sphere
    (r^3 uniform, isotropic), masses 1/n, Gaussian velocities of dispersion 0.3 per axis;
    the cold variant draws the same positions and has zero velocities."""
    import pynbg
    from scipy import stats
    n = 20000
    c = pynbg.ic_cluster(n, 1)
    cold = pynbg.ic_cluster(n, 1, cold=True)
    r = np.sqrt(c["x"] ** 2 + c["y"] ** 2 + c["z"] ** 2)
    assert r.max() <= 1.0
    assert stats.kstest(r ** 3, "uniform").pvalue > 1e-3
    assert stats.kstest((c["z"] / r + 1) / 2, "uniform").pvalue > 1e-3
    assert np.all(c["m"] == 1.0 / n)
    for k in ("vx", "vy", "vz"):
        assert abs(np.std(c[k]) - 0.3) < 0.01 and abs(np.mean(c[k])) < 0.01
        assert not np.any(cold[k])
    for k in ("x", "y", "z", "m"):
        np.testing.assert_array_equal(c[k], cold[k])
    # the stream is seeded: same seed, same sample; another seed, another sample
    np.testing.assert_array_equal(pynbg.ic_cluster(100, 7)["x"], pynbg.ic_cluster(100, 7)["x"])
    assert not np.array_equal(pynbg.ic_cluster(100, 7)["x"], pynbg.ic_cluster(100, 8)["x"])
