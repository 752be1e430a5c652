"""CPU oracle of the learned far-field correction -- TEST INFRASTRUCTURE ONLY.

Imported by tests/ (and nothing shipped): it checks the HIP kernels of
opengl-physics_amd/csrc/nbg_learn.hip, never replaces them.

The reference has no learned component, so nothing here follows a reference file; the
parts that touch the reference's method restate it:
  * the multipole force an accepted cell contributes is the monopole (Plummer-softened)
    plus the unsoftened traceless quadrupole term of ComputeAccelBarnesHutGroup's far
    interaction (projects/nbody_core/src/Solvers.cpp:295-330, Green.hpp:23-33);
  * the exact force of a cell is the softened pair sum of PlummerAccel
    (projects/nbody_core/include/ngrav/Green.hpp:23-33) over the cell's particles.

Everything is FP64.  How each function is pinned:
  * mlp_backward -- central finite differences of mlp_loss (test_learned_oracle.py);
  * octupole_force -- the far-field limit: exact minus monopole+quadrupole of a point
    cluster converges to it as (size/r) -> 0 (test_learned_oracle.py);
  * cell_moments -- the parallel-axis identities, checked against moments summed from
    particles about a shifted origin (test_learned_oracle.py);
  * features -- u, s/r, R/r, eps/r, trace are definitions; the octupole entry is the
    pinned octupole_force divided by the scale.
"""
import numpy as np

NF = 13
NH = 64
OFF = {}
_o = 0
for _name, _size in (("W1", NF * NH), ("b1", NH), ("W2", NH * NH), ("b2", NH), ("W3", NH * NH), ("b3", NH),
                     ("W4", NH * 3), ("b4", 3)):
    OFF[_name] = (_o, _size)
    _o += _size
NTRAIN = _o            # 9411
NPARAM = 32 + NTRAIN   # standardisation (mean[16], 1/std[16]) first
ACT_SLOT = 15          # the unused mean[15] holds the activation code
ACT_SILU, ACT_RELU = 0, 1


# ---------------------------------------------------------------------------- model blob
def unpack(blob):
    """Model blob -> dict of FP64 arrays (W as [in][out])."""
    b = np.asarray(blob, np.float64)
    t = b[32:]
    d = {"mu": b[:16], "isd": b[16:32]}
    shapes = {"W1": (NF, NH), "b1": (NH,), "W2": (NH, NH), "b2": (NH,), "W3": (NH, NH), "b3": (NH,), "W4": (NH, 3),
              "b4": (3,)}
    for k, (o, s) in OFF.items():
        d[k] = t[o:o + s].reshape(shapes[k])
    return d


def init_blob(seed=0, act=ACT_SILU):
    """Seeded He-style initialisation (the same recipe the training driver uses)."""
    rng = np.random.default_rng(seed)
    blob = np.zeros(NPARAM, np.float32)
    blob[16:32] = 1.0
    blob[ACT_SLOT] = act
    t = blob[32:]
    for k, fan_in in (("W1", NF), ("W2", NH), ("W3", NH), ("W4", NH)):
        o, s = OFF[k]
        scale = np.sqrt(2.0 / fan_in) * (0.1 if k == "W4" else 1.0)
        t[o:o + s] = rng.standard_normal(s) * scale
    return blob


def bf16(x):
    """Round-to-nearest-even to bfloat16, returned as float64 (the kernels' rounding)."""
    a = np.ascontiguousarray(np.asarray(x, np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def relu(z):
    return np.maximum(z, 0.0)


def drelu(z):
    return (np.asarray(z) > 0.0).astype(np.float64)


def activation(blob):
    """(act, act') of the model: blob[ACT_SLOT] = 0 SiLU, 1 ReLU."""
    code = int(np.asarray(blob)[ACT_SLOT])
    if code == ACT_SILU:
        return silu, dsilu
    if code == ACT_RELU:
        return relu, drelu
    raise ValueError(f"unknown activation code {code}")


def mlp_forward(blob, feats, emulate_bf16=False):
    """feats (n, >=13) raw features -> (y (n,3), cache).  With emulate_bf16 the inputs,
    weights and hidden activations are rounded to bfloat16 as the matrix-core kernels do,
    and the activation takes the BF16-rounded pre-activation (accumulation stays FP64
    here, FP32 on the device; the backward uses the activation's derivative at the
    unrounded pre-activation).  The activation is the model's (blob[ACT_SLOT])."""
    P = unpack(blob)
    r = bf16 if emulate_bf16 else (lambda v: np.asarray(v, np.float64))
    x = r((np.asarray(feats, np.float64)[:, :NF] - P["mu"][:NF]) * P["isd"][:NF])
    W = {k: r(P[k]) for k in ("W1", "W2", "W3", "W4")}
    # the device activation is that of the BF16-rounded pre-activation, rounded to BF16
    f, _ = activation(blob)
    act = (lambda z: bf16(f(bf16(z)))) if emulate_bf16 else f
    z1 = x @ W["W1"] + P["b1"]
    h1 = act(z1)
    z2 = h1 @ W["W2"] + P["b2"]
    h2 = act(z2)
    z3 = h2 @ W["W3"] + P["b3"]
    h3 = act(z3)
    y = h3 @ W["W4"] + P["b4"]
    return y, dict(x=x, z1=z1, h1=h1, z2=z2, h2=h2, z3=z3, h3=h3, W=W)


def mlp_loss(blob, feats, targets):
    """sum_i w_i |y_i - t_i|^2 / n  (targets: (n, 4) = residual xyz, weight)."""
    y, _ = mlp_forward(blob, feats)
    t = np.asarray(targets, np.float64)
    return float(np.sum(t[:, 3] * np.sum((y - t[:, :3]) ** 2, axis=1)) / len(t))


def mlp_backward(blob, feats, targets, emulate_bf16=False):
    """Gradient of mlp_loss with respect to the trainable parameters, in blob order
    (NTRAIN entries), and the predictions."""
    y, c = mlp_forward(blob, feats, emulate_bf16)
    _, dact = activation(blob)
    t = np.asarray(targets, np.float64)
    n = len(t)
    dy = 2.0 * t[:, 3:4] * (y - t[:, :3]) / n
    W = c["W"]
    g = {}
    g["W4"] = c["h3"].T @ dy
    g["b4"] = dy.sum(0)
    dz3 = (dy @ W["W4"].T) * dact(c["z3"])
    g["W3"] = c["h2"].T @ dz3
    g["b3"] = dz3.sum(0)
    dz2 = (dz3 @ W["W3"].T) * dact(c["z2"])
    g["W2"] = c["h1"].T @ dz2
    g["b2"] = dz2.sum(0)
    dz1 = (dz2 @ W["W2"].T) * dact(c["z1"])
    g["W1"] = c["x"].T @ dz1
    g["b1"] = dz1.sum(0)
    out = np.zeros(NTRAIN)
    for k, (o, s) in OFF.items():
        out[o:o + s] = g[k].ravel()
    return out, y


# ---------------------------------------------------------------------------- physics
def cell_moments(pos, w, com):
    """Second (xx yy zz xy xz yz) and third (xxx yyy zzz xxy xxz xyy yyz xzz yzz xyz)
    moments of weights w at pos about com."""
    d = np.asarray(pos, np.float64) - np.asarray(com, np.float64)
    w = np.asarray(w, np.float64)
    x, y, z = d[:, 0], d[:, 1], d[:, 2]
    s2 = np.array([np.sum(w * x * x), np.sum(w * y * y), np.sum(w * z * z), np.sum(w * x * y), np.sum(w * x * z),
                   np.sum(w * y * z)])
    s3 = np.array([np.sum(w * x ** 3), np.sum(w * y ** 3), np.sum(w * z ** 3), np.sum(w * x * x * y),
                   np.sum(w * x * x * z), np.sum(w * x * y * y), np.sum(w * y * y * z), np.sum(w * x * z * z),
                   np.sum(w * y * z * z), np.sum(w * x * y * z)])
    return s2, s3


def _s3_tensor(s3):
    xxx, yyy, zzz, xxy, xxz, xyy, yyz, xzz, yzz, xyz = s3
    T = np.empty((3, 3, 3))
    vals = {(0, 0, 0): xxx, (1, 1, 1): yyy, (2, 2, 2): zzz, (0, 0, 1): xxy, (0, 0, 2): xxz, (0, 1, 1): xyy,
            (1, 1, 2): yyz, (0, 2, 2): xzz, (1, 2, 2): yzz, (0, 1, 2): xyz}
    for (i, j, k), v in vals.items():
        for a, b, cc in {(i, j, k), (i, k, j), (j, i, k), (j, k, i), (k, i, j), (k, j, i)}:
            T[a, b, cc] = v
    return T


def octupole_force(R, s3):
    """Acceleration at offset R (target minus the cell's com) from the third moments s3
    (weights include G): -grad of the l=3 term of the softening-free expansion,
    7.5 S.RR/r^7 - 17.5 (S:RRR) R/r^9 - 1.5 t/r^5 + 7.5 (t.R) R/r^7, t_i = S_ijj."""
    R = np.asarray(R, np.float64)
    S = _s3_tensor(s3)
    r = np.linalg.norm(R)
    SRR = np.einsum("ijk,j,k->i", S, R, R)
    t = np.einsum("ijj->i", S)
    return 7.5 * SRR / r ** 7 - 17.5 * (SRR @ R) * R / r ** 9 - 1.5 * t / r ** 5 + 7.5 * (t @ R) * R / r ** 7


def multipole_force(p, com, M, quad5, eps2):
    """The walk's far-cell force on a point p: softened monopole + quadrupole (traceless
    Q = sum w (3 d d - d^2 I), stored as Qxx Qyy Qxy Qxz Qyz), weights with G."""
    d = np.asarray(com, np.float64) - np.asarray(p, np.float64)
    d2 = d @ d
    a = M * d / (d2 + eps2) ** 1.5
    qxx, qyy, qxy, qxz, qyz = quad5
    Q = np.array([[qxx, qxy, qxz], [qxy, qyy, qyz], [qxz, qyz, -(qxx + qyy)]])
    qd = Q @ d
    r = np.sqrt(d2)
    return a + 2.5 * (d @ qd) * d / r ** 7 - qd / r ** 5


def exact_force(p, src_pos, src_w, eps2, skip=None):
    d = np.asarray(src_pos, np.float64) - np.asarray(p, np.float64)
    r2 = np.sum(d * d, axis=1) + eps2
    f = np.asarray(src_w, np.float64) / (r2 * np.sqrt(r2))
    if skip is not None:
        f[skip] = 0.0
    return np.sum(f[:, None] * d, axis=0)


def features(tgt_com, tgt_R, cell_com, cell_M, size, s2, s3, eps):
    """The 13 features of a (group, cell) pair and the output scale M s^3 / r^5."""
    R = np.asarray(tgt_com, np.float64) - np.asarray(cell_com, np.float64)
    r = np.linalg.norm(R)
    u = R / r
    sh2 = np.asarray(s2, np.float64) / (cell_M * size * size)
    Q = np.array([[sh2[0], sh2[3], sh2[4]], [sh2[3], sh2[1], sh2[5]], [sh2[4], sh2[5], sh2[2]]])
    scale = cell_M * size ** 3 / r ** 5
    a3 = octupole_force(R, s3) / scale
    f = np.concatenate([u, [size / r], a3, Q @ u, [tgt_R / r, eps / r, np.trace(Q)]])
    return f, scale


def pair_sample(members_pos, cell_pos, cell_w, cell_com, cell_M, quad5, s2, s3, size, tgt_com, tgt_R, eps2,
                member_in_cell=None):
    """(features, normalised residual) of one (group, cell) pair: the exact field of the
    cell minus its multipole field, averaged over the group's members, over the scale."""
    res = np.zeros(3)
    for k, p in enumerate(members_pos):
        skip = None if member_in_cell is None else member_in_cell[k]
        res += exact_force(p, cell_pos, cell_w, eps2, skip) - multipole_force(p, cell_com, cell_M, quad5, eps2)
    res /= len(members_pos)
    f, scale = features(tgt_com, tgt_R, cell_com, cell_M, size, s2, s3, np.sqrt(eps2))
    return f, res / scale, scale
