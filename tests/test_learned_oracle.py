"""CPU checks of the learned-correction oracle (oracle/learned.py) and the committed model.

The reference has no learned component; these pin the oracle by mathematics:
finite differences for the backward pass, the far-field limit for the octupole term,
brute-force tensors for the moments.
"""
import os

import numpy as np
import pytest

import learned as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
WEIGHTS = os.path.join(ROOT, "opengl-physics_amd", "learned", "farfield_mlp.npy")


def _random_case(seed, n=9, act=L.ACT_SILU):
    rng = np.random.default_rng(seed)
    blob = L.init_blob(seed, act).astype(np.float64)
    blob[:13] = rng.normal(size=13) * 0.1
    blob[16:29] = 0.5 + rng.random(13)
    t = blob[32:]
    for k in ("b1", "b2", "b3", "b4"):
        o, s = L.OFF[k]
        t[o:o + s] = rng.normal(size=s) * 0.1
    X = rng.normal(size=(n, 16))
    T = np.concatenate([rng.normal(size=(n, 3)), 0.5 + rng.random((n, 1))], axis=1)
    return blob, X, T


def test_mlp_backward_matches_finite_differences():
    blob, X, T = _random_case(3)
    g, _ = L.mlp_backward(blob, X, T)
    rng = np.random.default_rng(0)
    # every block of the blob, including the first-layer weights of every input
    idx = np.concatenate([rng.choice(L.NTRAIN, 60, replace=False), [0, 12 * 64 + 5, L.OFF["b4"][0] + 2]])
    h = 1e-6
    for i in idx:
        b1 = blob.copy()
        b1[32 + i] += h
        b2 = blob.copy()
        b2[32 + i] -= h
        fd = (L.mlp_loss(b1, X, T) - L.mlp_loss(b2, X, T)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])) + 1e-9, (i, fd, g[i])


def test_relu_mlp_backward_matches_finite_differences():
    """The ReLU model (blob[ACT_SLOT] = 1): the same finite-difference check on a case
    whose pre-activations all stay at least 1e-3 away from the kink, so a +-1e-6 step
    never crosses it (seeds are tried in order until one qualifies)."""
    for seed in range(3, 40):
        blob, X, T = _random_case(seed, act=L.ACT_RELU)
        _, c = L.mlp_forward(blob, X)
        if min(np.min(np.abs(c[k])) for k in ("z1", "z2", "z3")) > 1e-3:
            break
    else:
        pytest.fail("no case with pre-activations away from zero")
    g, _ = L.mlp_backward(blob, X, T)
    assert np.all(c["h1"] >= 0.0) and np.any(c["h1"] == 0.0)      # the kink is exercised
    rng = np.random.default_rng(1)
    idx = np.concatenate([rng.choice(L.NTRAIN, 60, replace=False), [0, 12 * 64 + 5, L.OFF["b4"][0] + 2]])
    h = 1e-6
    for i in idx:
        b1 = blob.copy()
        b1[32 + i] += h
        b2 = blob.copy()
        b2[32 + i] -= h
        fd = (L.mlp_loss(b1, X, T) - L.mlp_loss(b2, X, T)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])) + 1e-9, (i, fd, g[i])


def test_activation_code_is_checked():
    blob, X, _ = _random_case(5)
    blob[L.ACT_SLOT] = 2
    with pytest.raises(ValueError):
        L.mlp_forward(blob, X)


def test_mlp_backward_planted_mistake_is_caught():
    """A dropped SiLU derivative in the backward pass must fail the finite-difference check."""
    blob, X, T = _random_case(4)
    y, c = L.mlp_forward(blob, X)
    n = len(T)
    dy = 2.0 * T[:, 3:4] * (y - T[:, :3]) / n
    wrong_dz3 = dy @ c["W"]["W4"].T                     # missing * dsilu(z3)
    wrong_gW3 = c["h2"].T @ wrong_dz3
    g, _ = L.mlp_backward(blob, X, T)
    o, s = L.OFF["W3"]
    assert np.max(np.abs(wrong_gW3.ravel() - g[o:o + s])) > 1e-3 * np.max(np.abs(g[o:o + s]))


def test_octupole_is_the_far_field_limit():
    rng = np.random.default_rng(5)
    P = rng.normal(size=(6, 3))
    w = rng.random(6) + 0.1
    com = (w[:, None] * P).sum(0) / w.sum()
    s2, s3 = L.cell_moments(P, w, com)
    d = P - com
    Qt = np.einsum("k,ki,kj->ij", w, 3 * d, d) - np.eye(3) * np.sum(w * (d * d).sum(1))
    q5 = [Qt[0, 0], Qt[1, 1], Qt[0, 1], Qt[0, 2], Qt[1, 2]]
    direction = np.array([0.3, -0.8, 0.52])
    errs = []
    for dist in (10.0, 100.0, 1000.0):
        p = com + dist * direction
        res = L.exact_force(p, P, w, 0.0) - L.multipole_force(p, com, w.sum(), q5, 0.0)
        o = L.octupole_force(p - com, s3)
        errs.append(np.linalg.norm(res - o) / np.linalg.norm(o))
    # the remainder is the hexadecapole: relative size ~ (size / r)
    assert errs[1] < 0.15 * errs[0] and errs[2] < 0.15 * errs[1], errs
    assert errs[2] < 5e-3


def test_cell_moments_match_brute_force_tensors():
    rng = np.random.default_rng(6)
    P = rng.normal(size=(11, 3))
    w = rng.random(11)
    c = rng.normal(size=3)
    s2, s3 = L.cell_moments(P, w, c)
    d = P - c
    T2 = np.einsum("k,ki,kj->ij", w, d, d)
    T3 = np.einsum("k,ki,kj,kl->ijl", w, d, d, d)
    np.testing.assert_allclose(s2, [T2[0, 0], T2[1, 1], T2[2, 2], T2[0, 1], T2[0, 2], T2[1, 2]], rtol=1e-12)
    np.testing.assert_allclose(L._s3_tensor(s3), T3, rtol=1e-12, atol=1e-14)


def test_features_geometry():
    s2 = np.array([1.0, 2.0, 3.0, 0.1, 0.2, 0.3])
    s3 = np.zeros(10)
    f, scale = L.features([3.0, 4.0, 0.0], 0.5, [0.0, 0.0, 0.0], 2.0, 1.0, s2, s3, 0.05)
    np.testing.assert_allclose(f[:3], [0.6, 0.8, 0.0])
    assert f[3] == pytest.approx(0.2)
    np.testing.assert_allclose(f[4:7], 0.0)
    assert f[10] == pytest.approx(0.1) and f[11] == pytest.approx(0.01)
    assert f[12] == pytest.approx(3.0)
    assert scale == pytest.approx(2.0 / 5.0 ** 5)


def test_product_initialisation_matches_oracle_recipe():
    from pynbg import learned as product
    np.testing.assert_array_equal(product.init_blob(7), L.init_blob(7))


def test_committed_model_blob():
    if not os.path.exists(WEIGHTS):
        pytest.skip("no trained model committed")
    blob = np.load(WEIGHTS, allow_pickle=False)
    assert blob.dtype == np.float32 and blob.shape == (L.NPARAM,)
    assert np.all(np.isfinite(blob))
    assert np.all(blob[16:16 + L.NF] > 0)
