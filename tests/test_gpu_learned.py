"""GPU checks of the learned far-field correction against the oracle.

* the walk's (leaf, cell) pair list equals the oracle's per-leaf far lists
  (CollectGroupInteractions, projects/nbody_core/src/Solvers.cpp:248-277) -- bit-exact;
* sample generation (moments, features, FP64 residuals) against oracle/learned.py;
* forward/backward of the BF16 matrix-core kernels against the FP64 oracle MLP;
* training lowers the loss; the trained model lowers the force error of theta = 0.8;
* graph-replayed and eager learned steps agree.
"""
import os

import numpy as np
import pytest

from conftest import f64, plummer_f32

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
WEIGHTS = os.path.join(ROOT, "opengl-physics_amd", "learned", "farfield_mlp.npy")


def _ids(feats):
    return feats[:, 13].view(np.int32).copy(), feats[:, 14].view(np.int32).copy()


def _subtree_ranges(e):
    """Sorted-order particle range [first, first+count) of every node (pre-order ids)."""
    nn = len(e["parent"])
    leaf = e["child_count"] == 0
    first = np.zeros(nn, np.int64)
    cnt = np.zeros(nn, np.int64)
    cur = -1
    for v in range(nn - 1, -1, -1):
        if leaf[v]:
            cnt[v] += e["count"][v]
            cur = v
        first[v] = e["first"][cur]
        if e["parent"][v] >= 0:
            cnt[e["parent"][v]] += cnt[v]
    return first, cnt


def _blob_for_lists():
    from pynbg import learned
    return learned.init_blob(0)


def test_pair_list_equals_oracle_far_lists(gpu_ctx_factory):
    import oracle
    import pynbg
    s = plummer_f32(4096, 2)
    x, y, z, m = f64(s)
    ctx = gpu_ctx_factory(4096)
    ctx.upload(s["x"], s["y"], s["z"], s["m"])
    p = pynbg.make_params("barnes_hut", "quadrupole", G=1.0, eps=0.005, theta=0.8, learn_band=0.0)
    ctx.learn_clear()
    n = ctx.learn_collect(p)
    feats, _ = ctx.learn_pool(0, n)
    gl, gc = _ids(feats)
    T = oracle.Tree(x, y, z, m, quad=True)
    e = T.export()
    fp, fi, _, _ = T.group_lists(0.8)
    leaves = np.nonzero((e["child_count"] == 0) & (e["count"] > 0))[0]
    ol = np.repeat(leaves, np.diff(fp))
    print(f"pairs gpu {n} oracle {len(fi)}")
    got = np.sort(gl.astype(np.int64) << 32 | gc.astype(np.int64))
    want = np.sort(ol.astype(np.int64) << 32 | fi.astype(np.int64))
    np.testing.assert_array_equal(got, want)


def test_band_selects_the_marginal_pairs(gpu_ctx_factory):
    import oracle
    import pynbg
    s = plummer_f32(4096, 2)
    x, y, z, m = f64(s)
    ctx = gpu_ctx_factory(4096)
    ctx.upload(s["x"], s["y"], s["z"], s["m"])
    ctx.learn_clear()
    band = 0.6
    n = ctx.learn_collect(pynbg.make_params("barnes_hut", G=1.0, eps=0.005, theta=0.8, learn_band=band))
    feats, _ = ctx.learn_pool(0, n)
    gl, gc = _ids(feats)
    T = oracle.Tree(x, y, z, m, quad=True)
    e = T.export()
    fp, fi, _, _ = T.group_lists(0.8)
    leaves = np.nonzero((e["child_count"] == 0) & (e["count"] > 0))[0]
    ol = np.repeat(leaves, np.diff(fp))
    worst = np.linalg.norm(e["com"][fi] - e["com"][ol], axis=1) - e["max_radius"][ol]
    ratio = 2.0 * e["half"][fi] / worst
    sel = ratio > band
    want = set(zip(ol[sel].tolist(), fi[sel].tolist()))
    got = set(zip(gl.tolist(), gc.tolist()))
    # the device tests the band in FP32: only pairs within a rounding of the edge may differ
    diff = want ^ got
    edge = np.abs(ratio - band) < 1e-5 * band
    assert len(diff) <= int(edge.sum()), (len(diff), int(edge.sum()))
    assert len(got) == n and 0.1 < n / len(fi) < 0.5


def test_amin_selects_by_correction_scale(gpu_ctx_factory):
    """learn_amin: a listed pair also needs G M s^3 / r^5 >= learn_amin, r the distance of
    the cell's and the leaf group's centres of mass (the output scale of the model)."""
    import oracle
    import pynbg
    s = plummer_f32(4096, 2)
    x, y, z, m = f64(s)
    ctx = gpu_ctx_factory(4096)
    ctx.upload(s["x"], s["y"], s["z"], s["m"])
    T = oracle.Tree(x, y, z, m, quad=True)
    e = T.export()
    fp, fi, _, _ = T.group_lists(0.8)
    leaves = np.nonzero((e["child_count"] == 0) & (e["count"] > 0))[0]
    ol = np.repeat(leaves, np.diff(fp))
    r = np.linalg.norm(e["com"][fi] - e["com"][ol], axis=1)
    scale = e["mass"][fi] * (2.0 * e["half"][fi]) ** 3 / r ** 5
    for amin in (1e-4, 1e-3):
        ctx.learn_clear()
        n = ctx.learn_collect(pynbg.make_params("barnes_hut", G=1.0, eps=0.005, theta=0.8, learn_band=0.0,
                                                learn_amin=amin))
        feats, _ = ctx.learn_pool(0, n)
        gl, gc = _ids(feats)
        sel = scale >= amin
        want = set(zip(ol[sel].tolist(), fi[sel].tolist()))
        got = set(zip(gl.tolist(), gc.tolist()))
        edge = np.abs(scale / amin - 1.0) < 1e-5
        print(f"amin {amin}: listed {n} of {len(fi)}, expected {int(sel.sum())}, edge {int(edge.sum())}")
        assert len(want ^ got) <= int(edge.sum())
        assert 0 < n < len(fi)


def test_samples_match_oracle(gpu_ctx_factory):
    import learned as LO
    import oracle
    import pynbg
    s = plummer_f32(4096, 5)
    x, y, z, m = f64(s)
    eps = 0.005
    ctx = gpu_ctx_factory(4096)
    ctx.upload(s["x"], s["y"], s["z"], s["m"])
    ctx.learn_clear()
    n = ctx.learn_collect(pynbg.make_params("barnes_hut", G=1.0, eps=eps, theta=0.8, learn_band=0.6))
    feats, tg = ctx.learn_pool(0, n)
    g = ctx.tree_export()
    first, cnt = _subtree_ranges(g)
    P = np.stack([s["x"], s["y"], s["z"]], 1).astype(np.float64)
    w = s["m"].astype(np.float64)
    order = g["order"]
    # accelerations of the snapshot (direct, FP64) for the loss weights
    ad = oracle.accel_direct(x, y, z, m, 1.0, eps * eps)
    a2 = np.sum(ad * ad, axis=1)
    rng = np.random.default_rng(0)
    pick = rng.choice(n, 150, replace=False)
    gl, gc = _ids(feats)
    F_or, Y_or, W_or = [], [], []
    for k in pick:
        lf, cl = gl[k], gc[k]
        mem = order[first[lf]:first[lf] + cnt[lf]]
        src = order[first[cl]:first[cl] + cnt[cl]]
        com = g["com"][cl].astype(np.float64)
        M = float(np.float32(g["mass"][cl]))
        s2, s3 = LO.cell_moments(P[src], w[src], com)
        skip = [np.nonzero(src == i)[0] for i in mem]
        f, yv, scale = LO.pair_sample(P[mem], P[src], w[src], com, M, g["quad"][cl].astype(np.float64), s2, s3,
                                      2.0 * g["half"][cl], g["com"][lf].astype(np.float64),
                                      float(g["max_radius"][lf]), eps * eps,
                                      member_in_cell=[sk if len(sk) else None for sk in skip])
        F_or.append(f)
        Y_or.append(yv)
        W_or.append(scale * np.sqrt(np.sum(1.0 / a2[mem])))
    F_or, Y_or, W_or = np.array(F_or), np.array(Y_or), np.array(W_or)
    Fg, Yg, Wg = feats[pick, :13].astype(np.float64), tg[pick, :3].astype(np.float64), tg[pick, 3]
    # geometry and second-moment features: FP32 rounding
    np.testing.assert_allclose(Fg[:, [0, 1, 2, 3, 10, 11]], F_or[:, [0, 1, 2, 3, 10, 11]], rtol=1e-4, atol=1e-5)
    np.testing.assert_allclose(Fg[:, [7, 8, 9, 12]], F_or[:, [7, 8, 9, 12]], rtol=1e-3, atol=1e-4)
    # octupole features: FP32 parallel-axis sums with cancellation, relative to their spread
    err3 = np.abs(Fg[:, 4:7] - F_or[:, 4:7]).max()
    print(f"octupole feature max abs err {err3:.2e} (spread {np.abs(F_or[:, 4:7]).max():.2e})")
    assert err3 < 2e-3 * max(1.0, np.abs(F_or[:, 4:7]).max())
    # residual targets: FP64 sums on both sides over the same FP32 inputs
    errY = np.abs(Yg - Y_or).max() / np.abs(Y_or).max()
    print(f"target max rel err {errY:.2e}")
    assert errY < 1e-3
    np.testing.assert_allclose(Wg, W_or, rtol=2e-3)


@pytest.mark.parametrize("act", ["silu", "relu"])
def test_forward_backward_match_oracle(gpu_ctx_factory, act):
    import learned as LO
    import pynbg
    rng = np.random.default_rng(11)
    blob = LO.init_blob(11, pynbg.LEARN_ACT[act])
    blob[:13] = rng.normal(size=13) * 0.2
    blob[16:29] = 0.5 + rng.random(13)
    t = blob[32:]
    for k in ("b1", "b2", "b3", "b4"):
        o, sz = LO.OFF[k]
        t[o:o + sz] = rng.normal(size=sz) * 0.1
    n = 200  # not a multiple of 64: the last block is partly empty
    X = np.zeros((n, 16), np.float32)
    X[:, :13] = rng.normal(size=(n, 13)).astype(np.float32)
    T = np.concatenate([rng.normal(size=(n, 3)), 0.5 + rng.random((n, 1))], 1).astype(np.float32)
    ctx = gpu_ctx_factory(64)
    ctx.learn_set_weights(blob)
    y, loss, grad = ctx.learn_grad(X, T)
    g_or, y_or = LO.mlp_backward(blob.astype(np.float64), X, T, emulate_bf16=True)
    loss_or = float(np.sum(T[:, 3] * np.sum((y_or - T[:, :3]) ** 2, 1)) / n)
    ey = np.abs(y - y_or).max() / np.abs(y_or).max()
    eg = np.linalg.norm(grad - g_or) / np.linalg.norm(g_or)
    cos = float(grad @ g_or / (np.linalg.norm(grad) * np.linalg.norm(g_or)))
    print(f"{act}: forward max rel {ey:.2e}  loss {loss:.6f} vs {loss_or:.6f}  grad rel L2 {eg:.2e} cos {cos:.6f}")
    # BF16 operands: inputs, weights and activations round identically on both sides;
    # the device also rounds the backward deltas to BF16 (2^-8 relative each)
    assert ey < 5e-3
    assert abs(loss - loss_or) <= 1e-2 * loss_or
    assert eg < 3e-2 and cos > 0.999
    for k in ("W1", "b1", "W2", "b2", "W3", "b3", "W4", "b4"):
        o, sz = LO.OFF[k]
        e = np.linalg.norm(grad[o:o + sz] - g_or[o:o + sz]) / np.linalg.norm(g_or[o:o + sz])
        assert e < 5e-2, (k, e)


def test_training_lowers_the_loss(gpu_ctx_factory):
    import pynbg
    from pynbg import learned
    s = plummer_f32(16384, 7)
    ctx = gpu_ctx_factory(16384)
    ctx.upload(s["x"], s["y"], s["z"], s["m"])
    ctx.learn_set_weights(learned.init_blob(1))
    ctx.learn_clear()
    n = ctx.learn_collect(pynbg.make_params("barnes_hut", G=1.0, eps=0.005, theta=0.8, learn_band=0.6))
    ctx.learn_standardize()
    loss = ctx.learn_train(600, 4096, 2e-3, seed=3)
    first, last = float(np.mean(loss[:20])), float(np.mean(loss[-50:]))
    print(f"pool {n} loss {first:.4f} -> {last:.4f}")
    assert np.all(np.isfinite(loss))
    assert last < 0.6 * first


def _rel_rms(a, ref):
    return float(np.sqrt(np.sum((a - ref) ** 2) / np.sum(ref ** 2)))


@pytest.mark.parametrize("selection", ["band0.6", "shipped"])
def test_trained_model_gain_theta08(gpu_ctx_factory, selection):
    """The committed model on an unseen Plummer draw (N=65,536, seed 11, eps=0.005):
    learned theta=0.8 against plain theta=0.8, both judged by the FP32 direct sum.
    Selections: the band-0.6 pairs the model was trained on, and the shipped one (every
    accepted pair whose correction scale G M s^3 / r^5 is at least 1e-3).  The shipped
    selection halves the plain theta=0.8 error here (measured 1.07e-3 against 2.18e-3);
    the 1e-3 target itself is asserted at the 1M-body size it is stated for
    (test_gpu_fullsize.py), since at N=65,536 the per-leaf correction's floor (leaves of
    2.5x the 1M leaves' size) keeps the learned step just above it."""
    import pynbg
    from pynbg import learned
    if not os.path.exists(WEIGHTS):
        pytest.skip("no trained model committed")
    s = plummer_f32(65536, 11)
    ctx = gpu_ctx_factory(65536)
    ctx.upload(s["x"], s["y"], s["z"], s["m"])
    ctx.accel(pynbg.make_params("direct", G=1.0, eps=0.005))
    ref = ctx.accelerations().astype(np.float64)
    ctx.learn_set_weights(learned.load_weights(WEIGHTS))
    res = {}
    for name, kw in (("plain 0.8", dict(theta=0.8)), ("plain 0.5", dict(theta=0.5)),
                     ("learned 0.8", dict(theta=0.8, learned=True, learn_band=0.6) if selection == "band0.6" else
                      dict(theta=0.8, learned=True, learn_band=0.0, learn_amin=1e-3))):
        st = ctx.accel(pynbg.make_params("barnes_hut", G=1.0, eps=0.005, **kw))
        res[name] = _rel_rms(ctx.accelerations().astype(np.float64), ref)
        if "learned" in name:
            assert st.n_learned > 0
    print(selection, {k: f"{v:.3e}" for k, v in res.items()})
    assert res["learned 0.8"] < (0.8 if selection == "band0.6" else 0.6) * res["plain 0.8"]


def test_learned_steps_bitwise_reproducible(gpu_ctx_factory):
    """The learned evaluation sums each leaf's corrections exactly (64-bit fixed point),
    so the order in which the persistent walk lists the pairs does not reach the result:
    eager steps, graph-replayed steps and a second graph run are bitwise equal (the
    run-twice check that stands in for a race detector)."""
    import pynbg
    from pynbg import learned
    s = plummer_f32(65536, 4)
    blob = learned.load_weights(WEIGHTS) if os.path.exists(WEIGHTS) else learned.init_blob(2)
    out = []
    for no_graph in (True, False, False):
        ctx = gpu_ctx_factory(65536)
        ctx.upload(s["x"], s["y"], s["z"], s["m"], s["vx"], s["vy"], s["vz"])
        ctx.learn_set_weights(blob)
        p = pynbg.make_params("barnes_hut", G=1.0, eps=0.005, theta=0.8, dt=1e-3, learned=True, learn_band=0.0,
                              learn_amin=1e-3, no_graph=no_graph)
        st = ctx.accel(p)
        assert st.n_learned > 0
        ctx.step(p, 3)
        d = ctx.download()
        out.append(np.stack([d[k] for k in ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az")]))
    np.testing.assert_array_equal(out[0], out[1])
    np.testing.assert_array_equal(out[1], out[2])


def test_listing_walk_forces_equal_plain(gpu_ctx_factory):
    """The learned walk writes its accelerations by sorted position and the apply pass
    scatters them: with a model whose output is zero everywhere the learned evaluation,
    and the listing walk of sample collection (no correction), reproduce the plain walk
    at the same theta bit for bit."""
    import pynbg
    from pynbg import learned
    s = plummer_f32(8192, 6)
    ctx = gpu_ctx_factory(8192)
    ctx.upload(s["x"], s["y"], s["z"], s["m"])
    ctx.accel(pynbg.make_params("barnes_hut", G=1.0, eps=0.01, theta=0.8))
    plain = ctx.accelerations().copy()
    # W4 and b4 (the last 64*3 + 3 words of the blob) zero: the network outputs zero
    zero = learned.init_blob(3)
    zero[-(64 * 3 + 3):] = 0.0
    ctx.learn_set_weights(zero)
    st = ctx.accel(pynbg.make_params("barnes_hut", G=1.0, eps=0.01, theta=0.8, learned=True, learn_band=0.0))
    assert st.n_learned > 0
    np.testing.assert_array_equal(ctx.accelerations(), plain)
    ctx.learn_clear()
    n = ctx.learn_collect(pynbg.make_params("barnes_hut", G=1.0, eps=0.01, theta=0.8, learn_band=0.0))
    assert n > 0
    np.testing.assert_array_equal(ctx.accelerations(), plain)


def test_learned_walk_tiles_agree(gpu_ctx_factory):
    """The learned walk comes in two near-field tiles (16, the default, and 8 for tunings
    of 4 or 8): they list the same pairs and differ only in how the P2P sums are grouped,
    so with the trained model their forces agree to FP32 summation noise.  The tile-8
    plain walk and the tile-8 learned walk with a zero model are equal bit for bit."""
    import pynbg
    from pynbg import learned
    s = plummer_f32(16384, 8)
    ctx = gpu_ctx_factory(16384)
    ctx.upload(s["x"], s["y"], s["z"], s["m"])
    ctx.learn_set_weights(learned.load_weights())
    p = pynbg.make_params("barnes_hut", G=1.0, eps=0.01, theta=0.8, learned=True, learn_band=0.0, learn_amin=1e-3)
    out = {}
    for nt in (16, 8):
        ctx.set_tuning(0, nt, 4)
        st = ctx.accel(p)
        out[nt] = (ctx.accelerations().astype(np.float64), st.n_learned)
    assert out[16][1] == out[8][1] > 0
    a16, a8 = out[16][0], out[8][0]
    rel = np.sqrt(np.sum((a16 - a8) ** 2, axis=1)) / np.sqrt(np.sum(a16 ** 2, axis=1))
    assert rel.max() < 1e-5, rel.max()
    ctx.set_tuning(0, 8, 4)
    ctx.accel(pynbg.make_params("barnes_hut", G=1.0, eps=0.01, theta=0.8))
    plain8 = ctx.accelerations().copy()
    zero = learned.init_blob(3)
    zero[-(64 * 3 + 3):] = 0.0
    ctx.learn_set_weights(zero)
    ctx.accel(pynbg.make_params("barnes_hut", G=1.0, eps=0.01, theta=0.8, learned=True, learn_band=0.0))
    np.testing.assert_array_equal(ctx.accelerations(), plain8)


def test_learned_argument_errors(gpu_ctx_factory):
    import pynbg
    s = plummer_f32(512, 1)
    ctx = gpu_ctx_factory(512)
    ctx.upload(s["x"], s["y"], s["z"], s["m"])
    with pytest.raises(pynbg.NbgError, match="no learned model"):
        ctx.accel(pynbg.make_params("barnes_hut", theta=0.8, learned=True))
    with pytest.raises(pynbg.NbgError, match="NBG_LEARN_NPARAM"):
        ctx.learn_set_weights(np.zeros(10, np.float32))
    from pynbg import learned
    bad = learned.init_blob(0)
    bad[pynbg.LEARN_ACT_SLOT] = 2
    with pytest.raises(pynbg.NbgError, match="activation"):
        ctx.learn_set_weights(bad)
    ctx.learn_set_weights(learned.init_blob(0))
    with pytest.raises(pynbg.NbgError, match="quadrupoles"):
        ctx.accel(pynbg.make_params("barnes_hut", "monopole", theta=0.8, learned=True))
    with pytest.raises(pynbg.NbgError, match="multiple of 64"):
        ctx.learn_train(1, 100, 1e-3)


def _run_learned(ctx, s, **kw):
    import pynbg
    ctx.upload(s["x"], s["y"], s["z"], s["m"], s["vx"], s["vy"], s["vz"])
    st = ctx.accel(pynbg.make_params("barnes_hut", G=1.0, eps=0.005, learned=True, **kw))
    return ctx.accelerations().astype(np.float64), st


def test_pair_list_regrowth_eager_and_graph(gpu_ctx_factory):
    """More listed pairs than the initial list holds: the eager evaluation grows the list and
    evaluates again; a graph batch that overflows restores its state and reruns eagerly."""
    import pynbg
    from pynbg import learned
    s = plummer_f32(65536, 6)
    blob = learned.init_blob(3)
    # reference: a context whose list already holds every pair
    ref_ctx = gpu_ctx_factory(65536)
    ref_ctx.learn_set_weights(blob)
    a_ref, st_ref = _run_learned(ref_ctx, s, theta=0.8, learn_band=0.0)
    assert st_ref.n_learned > (1 << 20) * 2  # beyond the 2M initial capacity of this context size
    ctx = gpu_ctx_factory(65536)
    ctx.learn_set_weights(blob)
    a, st = _run_learned(ctx, s, theta=0.8, learn_band=0.0)
    assert st.n_learned == st_ref.n_learned
    np.testing.assert_array_equal(a, a_ref)
    # graph batch: captured while the list is small (band 0.9 lists few pairs), then run with
    # band 0 so that the replay overflows
    g_ctx = gpu_ctx_factory(65536)
    g_ctx.learn_set_weights(blob)
    g_ctx.upload(s["x"], s["y"], s["z"], s["m"], s["vx"], s["vy"], s["vz"])
    p_small = pynbg.make_params("barnes_hut", G=1.0, eps=0.005, theta=0.8, dt=1e-4, learned=True, learn_band=0.9)
    g_ctx.accel(p_small)
    p_all = pynbg.make_params("barnes_hut", G=1.0, eps=0.005, theta=0.8, dt=1e-4, learned=True, learn_band=0.0)
    e_ctx = gpu_ctx_factory(65536)
    e_ctx.learn_set_weights(blob)
    e_ctx.upload(s["x"], s["y"], s["z"], s["m"], s["vx"], s["vy"], s["vz"])
    e_ctx.accel(p_small)
    p_eager = pynbg.make_params("barnes_hut", G=1.0, eps=0.005, theta=0.8, dt=1e-4, learned=True, learn_band=0.0,
                                no_graph=True)
    g_ctx.step(p_all, 2)
    e_ctx.step(p_eager, 2)
    dg, de = g_ctx.download(), e_ctx.download()
    for k in ("x", "vx", "ax"):
        np.testing.assert_array_equal(dg[k], de[k])


def test_learned_batch_keeps_systems_apart(gpu_ctx_factory):
    """A batch of three systems: each gets the correction it gets alone."""
    from pynbg import learned
    blob = learned.load_weights(WEIGHTS) if os.path.exists(WEIGHTS) else learned.init_blob(4)
    parts = [plummer_f32(n, 20 + k) for k, n in enumerate((3000, 7000, 5000))]
    alone = []
    for sp in parts:
        ctx = gpu_ctx_factory(len(sp["x"]))
        ctx.learn_set_weights(blob)
        alone.append(_run_learned(ctx, sp, theta=0.8, learn_band=0.62)[0])
    off = np.concatenate([[0], np.cumsum([len(sp["x"]) for sp in parts])]).astype(np.int64)
    # systems of a batch are independent: they may overlap in space
    cat = {k: np.concatenate([sp[k] for sp in parts]) for k in parts[0]}
    ctx = gpu_ctx_factory(int(off[-1]))
    ctx.learn_set_weights(blob)
    import pynbg
    ctx.upload_batch(off, cat["x"], cat["y"], cat["z"], cat["m"])
    ctx.accel(pynbg.make_params("barnes_hut", G=1.0, eps=0.005, theta=0.8, learned=True, learn_band=0.62))
    a = ctx.accelerations().astype(np.float64)
    for i in range(3):
        sl = slice(off[i], off[i + 1])
        # same trees, bins and pairs; the walk keeps stack headroom for the deepest leaf of
        # the whole batch, so it pops nodes in other batches than alone and sums its far and
        # near sources in another order: FP32 rounding differences only
        err = np.sqrt(np.sum((a[sl] - alone[i]) ** 2) / np.sum(alone[i] ** 2))
        assert err < 1e-6, (i, err)


def test_learned_without_pairs_equals_plain_and_deep_keys(gpu_ctx_factory):
    import pynbg
    from pynbg import learned
    s = plummer_f32(8192, 9)
    ctx = gpu_ctx_factory(8192)
    ctx.learn_set_weights(learned.init_blob(2))
    ctx.upload(s["x"], s["y"], s["z"], s["m"])
    ctx.accel(pynbg.make_params("barnes_hut", G=1.0, eps=0.005, theta=0.3))
    plain = ctx.accelerations()
    # theta = 0.3 with band 0.95: no accepted cell is in the band
    st = ctx.accel(pynbg.make_params("barnes_hut", G=1.0, eps=0.005, theta=0.3, learned=True, learn_band=0.95))
    assert st.n_learned == 0
    np.testing.assert_array_equal(ctx.accelerations(), plain)
    # 40-level keys: 64 coincident particles force the deep path; the correction stays finite
    x = np.concatenate([s["x"][:4000], np.full(64, 0.25, np.float32)])
    y = np.concatenate([s["y"][:4000], np.full(64, -0.125, np.float32)])
    z = np.concatenate([s["z"][:4000], np.full(64, 0.5, np.float32)])
    m = np.full(len(x), 1.0 / len(x), np.float32)
    ctx.upload(x, y, z, m)
    st = ctx.accel(pynbg.make_params("barnes_hut", G=1.0, eps=0.005, theta=0.8, learned=True, learn_band=0.6))
    assert st.deep_keys == 1 and st.n_learned > 0
    assert np.all(np.isfinite(ctx.accelerations()))


def _two_scale_state(n=4096):
    """A Plummer sphere shrunk to 1e-5 of its size plus one light particle at distance 1:
    the core's corrections exceed the one-pass fixed-point range of the evaluation (set by
    G M / h^2 of the root cell) by orders of magnitude."""
    s = plummer_f32(n, 12)
    out = {k: s[k].astype(np.float64) for k in s}
    for k in ("x", "y", "z"):
        out[k] = out[k] * 1e-5
    for k in ("vx", "vy", "vz"):
        out[k] = out[k] * 0.0
    out["x"][0], out["y"][0], out["z"][0] = 1.0, 0.0, 0.0
    out["m"][0] = out["m"][1] * 1e-3
    return {k: v.astype(np.float32) for k, v in out.items()}


@pytest.mark.parametrize("act", ["silu", "relu"])
@pytest.mark.parametrize("case", ["plummer", "two_scale"])
def test_learned_correction_equals_oracle_mlp_sums(gpu_ctx_factory, case, act):
    """The learned evaluation minus the plain walk at the same theta, per particle, equals
    the sum over its leaf's listed pairs of scale * MLP(features) computed by the oracle
    (FP64 MLP with the kernels' BF16 roundings, scale = G M s^3 / r^5 from the exported
    tree).  The two-scale state drives the sums onto the per-leaf-unit fallback (its terms
    are checked to exceed the one-pass range), the Plummer state keeps the one-pass sums;
    both are also bitwise reproducible.  The model is the committed one when its activation
    is `act`, else a seeded initialisation with that activation."""
    import learned as learned_oracle
    import pynbg
    from pynbg import learned
    if case == "plummer":
        s, eps = plummer_f32(8192, 13), 0.005
    else:
        s, eps = _two_scale_state(), 5e-8
    n = len(s["x"])
    code = pynbg.LEARN_ACT[act]
    blob = learned.load_weights(WEIGHTS) if os.path.exists(WEIGHTS) else None
    if blob is None or int(blob[pynbg.LEARN_ACT_SLOT]) != code:
        blob = learned.init_blob(4, code)
    kw = dict(G=1.0, eps=eps, theta=0.8, learn_band=0.0)
    ctx = gpu_ctx_factory(n)
    ctx.learn_set_weights(blob)
    ctx.upload(s["x"], s["y"], s["z"], s["m"])
    ctx.accel(pynbg.make_params("barnes_hut", **kw))
    plain = ctx.accelerations().astype(np.float64)
    st = ctx.accel(pynbg.make_params("barnes_hut", learned=True, **kw))
    lrn = ctx.accelerations().astype(np.float64)
    again = ctx.accel(pynbg.make_params("barnes_hut", learned=True, **kw))
    np.testing.assert_array_equal(ctx.accelerations().astype(np.float64), lrn)
    assert again.n_learned == st.n_learned > 0
    e = ctx.tree_export()
    ctx.learn_clear()
    npairs = ctx.learn_collect(pynbg.make_params("barnes_hut", **kw))
    assert npairs == st.n_learned
    feats, _ = ctx.learn_pool(0, npairs)
    gl, gc = _ids(feats)
    y, _ = learned_oracle.mlp_forward(blob.astype(np.float64), feats.astype(np.float64), emulate_bf16=True)
    size = 2.0 * e["half"][gc]
    r = np.linalg.norm(e["com"][gc].astype(np.float64) - e["com"][gl].astype(np.float64), axis=1)
    scale = e["mass"][gc].astype(np.float64) * size ** 3 / r ** 5
    terms = scale[:, None] * y
    corr = np.zeros((e["n_nodes"], 3))
    np.add.at(corr, gl, terms)
    want = np.zeros((n, 3))
    leaves = np.nonzero(e["child_count"] == 0)[0]
    for lf in leaves:
        idx = e["order"][e["first"][lf]:e["first"][lf] + e["count"][lf]]
        want[idx] = corr[lf]
    got = lrn - plain
    err = np.sqrt(np.sum((got - want) ** 2) / np.sum(want ** 2))
    # the one-pass range: terms below 2^49 at 2^30 / (G M / h^2 of the root cell)
    root = int(np.argmin(e["depth"]))
    a_ref = float(e["mass"][root]) / float(e["half"][root]) ** 2
    over = np.abs(terms).max() * 2.0 ** (30 - np.floor(np.log2(a_ref))) >= 2.0 ** 49
    print(f"{case} {act}: pairs {npairs}, correction rel err vs oracle {err:.2e}, largest term / one-pass limit "
          f"{np.abs(terms).max() * 2.0 ** (30 - np.floor(np.log2(a_ref))) / 2.0 ** 49:.2e}")
    assert over == (case == "two_scale")
    assert err < 1e-2
