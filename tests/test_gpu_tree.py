"""GPU tree parity: Morton keys, sort permutation and octree topology bit-exact against
the oracle's serial-insertion AdaptiveTree<3> (projects/nbody_core/src/Tree.cpp:33-254)."""
import numpy as np
import pytest

from conftest import f64, keys_at, plummer_f32

pytestmark = pytest.mark.gpu


def _params(pynbg, multipole="quadrupole", theta=0.5, eps=0.01, count=False):
    return pynbg.make_params("barnes_hut", multipole, G=1.0, eps=eps, theta=theta, count=count)


def _compare_trees(g, o, n, moments=True):
    assert g["n_nodes"] == len(o["mass"]), (g["n_nodes"], len(o["mass"]))
    np.testing.assert_array_equal(g["parent"], o["parent"])
    np.testing.assert_array_equal(g["depth"], o["depth"])
    np.testing.assert_array_equal(g["child_count"], o["child_count"])
    np.testing.assert_array_equal(g["children"], o["children"])
    np.testing.assert_array_equal(g["first"], o["first"])
    np.testing.assert_array_equal(g["count"], o["count"])
    np.testing.assert_array_equal(g["order"], o["order"])
    # FP64 geometry: bit-exact
    np.testing.assert_array_equal(g["center"], o["center"])
    np.testing.assert_array_equal(g["half"], o["half"])
    if moments:
        live = o["mass"] > 0
        np.testing.assert_allclose(g["mass"][live], o["mass"][live], rtol=2e-6)
        scale = np.maximum(o["half"][live], 1e-30)[:, None]
        err = np.abs(g["com"][live] - o["com"][live]) / scale
        assert err.max() < 1e-4, err.max()


def _run(ctx_factory, pynbg, x, y, z, m, multipole="quadrupole"):
    ctx = ctx_factory(max(len(x), 1))
    ctx.upload(x, y, z, m)
    st = ctx.accel(_params(pynbg, multipole))
    return ctx, ctx.tree_export(), st


@pytest.mark.parametrize("n,seed", [(65536, 0), (4096, 3), (1000, 1)])
def test_plummer_tree_bit_exact(gpu_ctx_factory, n, seed):
    import oracle
    import pynbg
    s = plummer_f32(n, seed)
    x, y, z, m = f64(s)
    ctx, g, st = _run(gpu_ctx_factory, pynbg, s["x"], s["y"], s["z"], s["m"])
    T = oracle.Tree(x, y, z, m, quad=True)
    o = T.export()
    keys_hi, _ = oracle.morton_keys(x, y, z, 21)
    assert st.key_levels == 13  # no cell at depth 13 holds more than 8 particles here
    np.testing.assert_array_equal(g["keys"], np.sort(keys_at(keys_hi, st.key_levels)))
    _compare_trees(g, o, n)
    assert st.deep_keys == 0
    assert st.max_depth == o["depth"].max()


def test_depth13_overflow_takes_16_level_keys(gpu_ctx_factory):
    """A clump of depth 15 (Plummer x 3e-4 plus one particle at 1): the 13-level keys
    overflow and the build is redone with 16-level keys (6 sort passes instead of 5, the
    sorted data back in the other ping-pong buffers) -- keys and topology bit-exact against
    the oracle -- and the context keeps 16 levels for its next evaluation."""
    import oracle
    import pynbg
    s = plummer_f32(4096, 11)
    for k in ("x", "y", "z"):
        s[k] = (s[k] * np.float32(3e-4)).astype(np.float32)
    s["x"][0], s["y"][0], s["z"][0] = np.float32(1.0), np.float32(0.0), np.float32(0.0)
    x, y, z, m = f64(s)
    ctx, g, st = _run(gpu_ctx_factory, pynbg, s["x"], s["y"], s["z"], s["m"])
    o = oracle.Tree(x, y, z, m, quad=True).export()
    print(f"depth {st.max_depth} key levels {st.key_levels}", flush=True)
    assert 13 < o["depth"].max() <= 16
    assert st.key_levels == 16 and st.deep_keys == 0
    keys_hi, _ = oracle.morton_keys(x, y, z, 21)
    np.testing.assert_array_equal(g["keys"], np.sort(keys_at(keys_hi, 16)))
    _compare_trees(g, o, len(x))
    st2 = ctx.accel(_params(pynbg))
    assert st2.key_levels == 16
    np.testing.assert_array_equal(ctx.tree_export()["order"], g["order"])


def test_depth16_overflow_takes_21_level_keys(gpu_ctx_factory):
    """A dense clump far below a wide root box (Plummer x 2e-5 plus one particle at 1): its
    tree is deeper than 16 levels, so the 13- and then the 16-level keys overflow and the
    build is redone with 21-level keys -- topology bit-exact against the oracle -- and the
    context keeps 21 levels (graph-replayed steps then equal eager ones bitwise) until a new
    upload, which tries 13 again."""
    import oracle
    import pynbg
    s = plummer_f32(4096, 11)
    for k in ("x", "y", "z"):
        s[k] = (s[k] * np.float32(2e-5)).astype(np.float32)
    s["x"][0], s["y"][0], s["z"][0] = np.float32(1.0), np.float32(0.0), np.float32(0.0)
    x, y, z, m = f64(s)
    ctx, g, st = _run(gpu_ctx_factory, pynbg, s["x"], s["y"], s["z"], s["m"])
    o = oracle.Tree(x, y, z, m, quad=True).export()
    print(f"depth {st.max_depth} key levels {st.key_levels}", flush=True)
    assert 16 < o["depth"].max() <= 21
    assert st.key_levels == 21 and st.deep_keys == 0
    keys_hi, _ = oracle.morton_keys(x, y, z, 21)
    np.testing.assert_array_equal(g["keys"], np.sort(keys_hi))
    _compare_trees(g, o, len(x))
    # steps: the captured graph uses the 21-level keys too
    p = pynbg.make_params("barnes_hut", G=1.0, eps=1e-5, theta=0.5, dt=1e-6)
    out = []
    for flags in ({}, {"no_graph": True}):
        ctx.upload(s["x"], s["y"], s["z"], s["m"])
        ctx.accel(p)
        q = pynbg.make_params("barnes_hut", G=1.0, eps=1e-5, theta=0.5, dt=1e-6, **flags)
        ctx.step(q, 3)
        out.append(ctx.download())
    for k in out[0]:
        np.testing.assert_array_equal(out[0][k], out[1][k])
    # a shallow state in the same context starts from 13 levels again
    t = plummer_f32(4096, 12)
    ctx.upload(t["x"], t["y"], t["z"], t["m"])
    assert ctx.accel(_params(pynbg)).key_levels == 13


def test_tiny_and_degenerate_inputs(gpu_ctx_factory):
    """n=1, n=8 (root leaf), n=9 (first split), points exactly on centre planes."""
    import oracle
    import pynbg
    rng = np.random.default_rng(7)
    cases = []
    cases.append(np.zeros((1, 3)))
    cases.append(rng.uniform(-1, 1, (8, 3)))
    cases.append(rng.uniform(-1, 1, (9, 3)))
    # a lattice: many coordinates fall exactly on cell centres (>= sends them up)
    g1 = np.linspace(-1.0, 1.0, 9)
    lat = np.stack(np.meshgrid(g1, g1, g1, indexing="ij"), -1).reshape(-1, 3)
    cases.append(lat)
    cases.append(rng.integers(-4, 5, (300, 3)).astype(np.float64) * 0.25)
    for pts in cases:
        p32 = pts.astype(np.float32)
        m = np.full(len(p32), 1.0 / len(p32), np.float32)
        ctx, g, _ = _run(gpu_ctx_factory, pynbg, p32[:, 0], p32[:, 1], p32[:, 2], m)
        x, y, z = (p32[:, k].astype(np.float64) for k in range(3))
        o = oracle.Tree(x, y, z, m.astype(np.float64), quad=True).export()
        _compare_trees(g, o, len(p32))


def test_coincident_points_depth_cap(gpu_ctx_factory):
    """Coincident points reach the depth-40 cap (two-word keys) and stay in one leaf."""
    import oracle
    import pynbg
    rng = np.random.default_rng(11)
    pts = rng.uniform(-1, 1, (500, 3)).astype(np.float32)
    pts[100:130] = pts[5]          # 31 identical points
    pts[200:210] = pts[7] + np.float32(1e-7)  # near-coincident cluster
    m = np.full(len(pts), 1.0 / len(pts), np.float32)
    ctx, g, st = _run(gpu_ctx_factory, pynbg, pts[:, 0], pts[:, 1], pts[:, 2], m)
    assert st.deep_keys == 1
    x, y, z = (pts[:, k].astype(np.float64) for k in range(3))
    o = oracle.Tree(x, y, z, m.astype(np.float64), quad=True).export()
    _compare_trees(g, o, len(pts), moments=False)
    assert o["depth"].max() == 40


def test_batch_of_systems_matches_separate_trees(gpu_ctx_factory):
    import oracle
    import pynbg
    sizes = [1000, 1, 0, 3000, 9]
    xs, ys, zs, ms = [], [], [], []
    for k, n in enumerate(sizes):
        if n == 0:
            continue
        s = plummer_f32(max(n, 2), 100 + k)
        xs.append(s["x"][:n] + np.float32(3 * k))
        ys.append(s["y"][:n])
        zs.append(s["z"][:n])
        ms.append(s["m"][:n])
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    x, y, z, m = (np.concatenate(a) for a in (xs, ys, zs, ms))
    ctx = gpu_ctx_factory(len(x))
    ctx.upload_batch(off, x, y, z, m)
    ctx.accel(_params(pynbg))
    g = ctx.tree_export()
    # nodes of each system are contiguous in pre-order; compare system by system
    node_base = 0
    for k, n in enumerate(sizes):
        if n == 0:
            continue
        sl = slice(off[k], off[k + 1])
        o = oracle.Tree(*(a[sl].astype(np.float64) for a in (x, y, z, m)), quad=True).export()
        nn = len(o["mass"])
        sub = {key: g[key][node_base:node_base + nn] for key in
               ("parent", "depth", "child_count", "children", "first", "count", "center", "half", "com", "mass")}
        sub["n_nodes"] = nn
        sub["parent"] = np.where(sub["parent"] >= 0, sub["parent"] - node_base, -1)
        sub["children"] = np.where(sub["children"] >= 0, sub["children"] - node_base, -1)
        sub["first"] = np.where(sub["first"] >= 0, sub["first"] - off[k], -1)
        sub["order"] = g["order"][sl] - off[k]
        _compare_trees(sub, o, n)
        node_base += nn
    assert node_base == g["n_nodes"]
