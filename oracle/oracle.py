"""ctypes binding of the FP64 oracle (oracle/nbody_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as the
checker, and by bench.py's CPU-baseline leg.  The product never imports it.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}

_d = ctypes.POINTER(ctypes.c_double)
_i = ctypes.POINTER(ctypes.c_int)
_i64 = ctypes.POINTER(ctypes.c_int64)
_u64 = ctypes.POINTER(ctypes.c_uint64)


def _ptr(a, t):
    return a.ctypes.data_as(t) if a is not None else None


def lib(fast=False):
    """Load liboracle.so (checker build) or, for fast=True, liboracle_fast.so (reference
    flags with -march=x86-64-v3); fast="native" loads liboracle_native.so (reference flags
    with -march=native, built on the host that runs it by `make -C oracle native`)."""
    name = {False: "liboracle.so", True: "liboracle_fast.so", "native": "liboracle_native.so"}[fast]
    if name in _LIBS:
        return _LIBS[name]
    path = os.path.join(_HERE, name)
    if not os.path.exists(path):
        raise RuntimeError(f"oracle library {path} is not built (run make -C oracle)")
    L = ctypes.CDLL(path)
    L.orc_ic_plummer.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_double, ctypes.c_double,
                                 ctypes.c_double] + [_d] * 7
    L.orc_mt19937_10000th.restype = ctypes.c_uint32
    L.orc_uniform_sequence.argtypes = [ctypes.c_uint32, ctypes.c_double, ctypes.c_double, ctypes.c_int, _d]
    L.orc_tree_build.restype = ctypes.c_void_p
    L.orc_tree_build.argtypes = [ctypes.c_int] + [_d] * 4 + [ctypes.c_int] * 3
    L.orc_tree_free.argtypes = [ctypes.c_void_p]
    L.orc_tree_num_nodes.argtypes = [ctypes.c_void_p]
    L.orc_tree_export.argtypes = [ctypes.c_void_p, _d, _d, _d, _i, _i, _i]
    L.orc_accel_bh.argtypes = [ctypes.c_void_p, ctypes.c_int] + [_d] * 4 + [ctypes.c_double] * 3 + \
        [ctypes.c_int] + [_d] * 3 + [_i64] * 2
    L.orc_accel_bh_pp.argtypes = [ctypes.c_void_p, ctypes.c_int] + [_d] * 4 + [ctypes.c_double] * 3 + \
        [ctypes.c_int, ctypes.c_int, ctypes.c_double, _d] + [_d] * 3 + [_i64]
    L.orc_accel_bh_targets.argtypes = [ctypes.c_void_p, ctypes.c_int] + [_d] * 4 + [ctypes.c_double] * 3 + \
        [ctypes.c_int, ctypes.c_int, _i] + [_d] * 3
    L.orc_group_lists.argtypes = [ctypes.c_void_p, ctypes.c_double, _i64, _i, _i64, _i]
    L.orc_accel_direct.argtypes = [ctypes.c_int] + [_d] * 4 + [ctypes.c_double] * 2 + [ctypes.c_int] + [_d] * 3
    L.orc_accel_direct_targets.argtypes = [ctypes.c_int] + [_d] * 4 + [ctypes.c_double] * 2 + \
        [ctypes.c_int, ctypes.c_int, _i] + [_d] * 3
    L.orc_steps.restype = ctypes.c_double
    L.orc_steps.argtypes = [ctypes.c_int] + [_d] * 11 + [ctypes.c_double] * 2 + [ctypes.c_int, ctypes.c_double] + \
        [ctypes.c_int] * 3 + [ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int]
    L.orc_adaptive_dt.restype = ctypes.c_double
    L.orc_adaptive_dt.argtypes = [ctypes.c_int] + [_d] * 3 + [ctypes.c_double] * 3
    L.orc_rungs.restype = ctypes.c_int
    L.orc_rungs.argtypes = [ctypes.c_int] + [_d] * 3 + [ctypes.c_double] * 3 + [ctypes.c_int, _i]
    L.orc_block_steps.restype = ctypes.c_int
    L.orc_block_steps.argtypes = [ctypes.c_int] + [_d] * 10 + [_i] + [ctypes.c_double] * 2 + \
        [ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
         ctypes.c_int, ctypes.c_int, _i64]
    L.orc_leapfrog.argtypes = [ctypes.c_int] + [_d] * 10 + [ctypes.c_double] * 3 + [ctypes.c_int] * 2 + \
        [ctypes.c_double, ctypes.c_int]
    L.orc_total_energy.restype = ctypes.c_double
    L.orc_total_energy.argtypes = [ctypes.c_int] + [_d] * 7 + [ctypes.c_double] * 2 + [ctypes.c_int]
    L.orc_moments.argtypes = [ctypes.c_int] + [_d] * 7 + [_d] * 3
    L.orc_root_box.argtypes = [ctypes.c_int] + [_d] * 3 + [_d, _d]
    L.orc_morton_keys.argtypes = [ctypes.c_int] + [_d] * 3 + [_d, ctypes.c_double, ctypes.c_int, _u64, _u64]
    L.orc_ic_hernquist.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_double] + [_d] * 7
    L.orc_ic_king.argtypes = [ctypes.c_int, ctypes.c_uint32] + [ctypes.c_double] * 4 + [_d] * 7
    L.orc_king_tidal_radius.restype = ctypes.c_double
    L.orc_king_tidal_radius.argtypes = [ctypes.c_double]
    _LIBS[name] = L
    return L


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def plummer(n, seed, total_mass=1.0, scale_radius=1.0, G=1.0, fast=False):
    """ic::Plummer<3> in FP64 -> dict of x,y,z,vx,vy,vz,m."""
    out = {k: np.empty(n, np.float64) for k in ("x", "y", "z", "vx", "vy", "vz", "m")}
    lib(fast).orc_ic_plummer(n, seed, total_mass, scale_radius, G,
                             *[_ptr(out[k], _d) for k in ("x", "y", "z", "vx", "vy", "vz", "m")])
    return out


def hernquist(n, seed, total_mass=1.0, scale_radius=1.0, G=1.0):
    """ic::Hernquist<3> in FP64 -> dict of x,y,z,vx,vy,vz,m."""
    out = {k: np.empty(n, np.float64) for k in ("x", "y", "z", "vx", "vy", "vz", "m")}
    lib().orc_ic_hernquist(n, seed, total_mass, scale_radius, G,
                           *[_ptr(out[k], _d) for k in ("x", "y", "z", "vx", "vy", "vz", "m")])
    return out


def king(n, seed, w0=6.0, total_mass=1.0, core_radius=1.0, G=1.0):
    """ic::King<3> in FP64 -> dict of x,y,z,vx,vy,vz,m."""
    out = {k: np.empty(n, np.float64) for k in ("x", "y", "z", "vx", "vy", "vz", "m")}
    lib().orc_ic_king(n, seed, total_mass, core_radius, w0, G,
                      *[_ptr(out[k], _d) for k in ("x", "y", "z", "vx", "vy", "vz", "m")])
    return out


def king_tidal_radius(w0):
    """Dimensionless tidal radius x_t where the King(W0) potential W reaches 0."""
    return float(lib().orc_king_tidal_radius(w0))


def uniform_sequence(seed, lo, hi, n):
    """n draws of std::uniform_real_distribution<double>(lo, hi) from std::mt19937(seed)."""
    out = np.empty(n, np.float64)
    lib().orc_uniform_sequence(seed, lo, hi, n, _ptr(out, _d))
    return out


def mt19937_10000th():
    return int(lib().orc_mt19937_10000th())


class Tree:
    """AdaptiveTree<3> built by serial insertion (Tree.cpp:33-203)."""

    def __init__(self, x, y, z, m, ncrit=8, max_depth=40, quad=True, fast=False):
        self._L = lib(fast)
        self.x, self.y, self.z, self.m = (_f64(a) for a in (x, y, z, m))
        self.n = len(self.x)
        self._h = self._L.orc_tree_build(self.n, _ptr(self.x, _d), _ptr(self.y, _d), _ptr(self.z, _d),
                                         _ptr(self.m, _d), ncrit, max_depth, 1 if quad else 0)
        self.num_nodes = self._L.orc_tree_num_nodes(self._h)

    def __del__(self):
        if getattr(self, "_h", None):
            self._L.orc_tree_free(self._h)
            self._h = None

    def export(self):
        """Canonical (DFS pre-order) node arrays."""
        nn = self.num_nodes
        geom = np.empty((nn, 4), np.float64)
        mom = np.empty((nn, 5), np.float64)
        quad = np.empty((nn, 5), np.float64)
        topo = np.empty((nn, 5), np.int32)
        ch = np.empty((nn, 8), np.int32)
        order = np.empty(self.n, np.int32)
        self._L.orc_tree_export(self._h, _ptr(geom, _d), _ptr(mom, _d), _ptr(quad, _d), _ptr(topo, _i),
                                _ptr(ch, _i), _ptr(order, _i))
        return dict(center=geom[:, :3], half=geom[:, 3], com=mom[:, :3], mass=mom[:, 3], max_radius=mom[:, 4],
                    quad=quad, parent=topo[:, 0], first=topo[:, 1], count=topo[:, 2], depth=topo[:, 3],
                    child_count=topo[:, 4], children=ch, order=order)

    def accel_bh(self, G, eps2, theta, counts=False, spline=False):
        n = self.n
        ax, ay, az = (np.empty(n, np.float64) for _ in range(3))
        nfar = np.empty(self.num_nodes, np.int64) if counts else None
        nnear = np.empty(self.num_nodes, np.int64) if counts else None
        self._L.orc_accel_bh(self._h, n, _ptr(self.x, _d), _ptr(self.y, _d), _ptr(self.z, _d), _ptr(self.m, _d),
                             G, eps2, theta, 1 if spline else 0, _ptr(ax, _d), _ptr(ay, _d), _ptr(az, _d),
                             _ptr(nfar, _i64),
                             _ptr(nnear, _i64))
        a = np.stack([ax, ay, az], axis=1)
        return (a, nfar, nnear) if counts else a

    def group_lists(self, theta):
        """Per-leaf CSR interaction lists (canonical ids, leaves in pre-order):
        (far_ptr, far_idx, near_ptr, near_idx)."""
        e = self.export()
        nl = int(np.sum((e["child_count"] == 0) & (e["count"] > 0)))
        fp = np.zeros(nl + 1, np.int64)
        qp = np.zeros(nl + 1, np.int64)
        self._L.orc_group_lists(self._h, theta, _ptr(fp, _i64), None, _ptr(qp, _i64), None)
        fi = np.empty(int(fp[-1]), np.int32)
        qi = np.empty(int(qp[-1]), np.int32)
        self._L.orc_group_lists(self._h, theta, _ptr(fp, _i64), _ptr(fi, _i), _ptr(qp, _i64), _ptr(qi, _i))
        return fp, fi, qp, qi

    def accel_bh_per_particle(self, G, eps2, theta, relative=False, alpha=0.001, a_old=None, spline=False,
                              counts=False):
        """ComputeAccelBarnesHutPerParticle (Solvers.cpp:359-382); with counts, also the
        totals (M2P, P2P) of the interactions."""
        n = self.n
        ax, ay, az = (np.empty(n, np.float64) for _ in range(3))
        ao = _f64(a_old) if a_old is not None else None
        cnt = np.zeros(2, np.int64)
        self._L.orc_accel_bh_pp(self._h, n, _ptr(self.x, _d), _ptr(self.y, _d), _ptr(self.z, _d),
                                _ptr(self.m, _d), G, eps2, theta, 1 if spline else 0, 1 if relative else 0, alpha,
                                _ptr(ao, _d), _ptr(ax, _d), _ptr(ay, _d), _ptr(az, _d), _ptr(cnt, _i64))
        a = np.stack([ax, ay, az], axis=1)
        return (a, int(cnt[0]), int(cnt[1])) if counts else a

    def accel_bh_targets(self, G, eps2, theta, targets, spline=False):
        """ComputeAccelBarnesHutTargets (Solvers.cpp:384-405), results by target slot."""
        t = np.ascontiguousarray(targets, dtype=np.int32)
        nt = len(t)
        ax, ay, az = (np.empty(nt, np.float64) for _ in range(3))
        self._L.orc_accel_bh_targets(self._h, self.n, _ptr(self.x, _d), _ptr(self.y, _d), _ptr(self.z, _d),
                                     _ptr(self.m, _d), G, eps2, theta, 1 if spline else 0, nt, _ptr(t, _i),
                                     _ptr(ax, _d), _ptr(ay, _d), _ptr(az, _d))
        return np.stack([ax, ay, az], axis=1)


def accel_direct(x, y, z, m, G, eps2, fast=False, spline=False):
    x, y, z, m = (_f64(a) for a in (x, y, z, m))
    n = len(x)
    ax, ay, az = (np.empty(n, np.float64) for _ in range(3))
    lib(fast).orc_accel_direct(n, _ptr(x, _d), _ptr(y, _d), _ptr(z, _d), _ptr(m, _d), G, eps2, 1 if spline else 0,
                               _ptr(ax, _d), _ptr(ay, _d), _ptr(az, _d))
    return np.stack([ax, ay, az], axis=1)


def accel_direct_targets(x, y, z, m, G, eps2, targets, fast=False, spline=False):
    x, y, z, m = (_f64(a) for a in (x, y, z, m))
    t = np.ascontiguousarray(targets, dtype=np.int32)
    nt = len(t)
    ax, ay, az = (np.empty(nt, np.float64) for _ in range(3))
    lib(fast).orc_accel_direct_targets(len(x), _ptr(x, _d), _ptr(y, _d), _ptr(z, _d), _ptr(m, _d), G, eps2,
                                       1 if spline else 0, nt,
                                       _ptr(t, _i), _ptr(ax, _d), _ptr(ay, _d), _ptr(az, _d))
    return np.stack([ax, ay, az], axis=1)


def leapfrog(state, a, G, eps2, theta, dt, nsteps, solver="barnes_hut", quad=True, fast=False):
    """KDK steps in place on a dict state (x..vz,m float64) with a (n,3) = a(t)."""
    s = {k: _f64(state[k]).copy() for k in ("x", "y", "z", "vx", "vy", "vz", "m")}
    ax, ay, az = (np.ascontiguousarray(a[:, k], dtype=np.float64).copy() for k in range(3))
    lib(fast).orc_leapfrog(len(s["x"]), *[_ptr(s[k], _d) for k in ("x", "y", "z", "vx", "vy", "vz", "m")],
                           _ptr(ax, _d), _ptr(ay, _d), _ptr(az, _d), G, eps2, theta,
                           0 if solver == "direct" else 1, 1 if quad else 0, dt, nsteps)
    return s, np.stack([ax, ay, az], axis=1)


def steps(state, a, G, eps2, theta, dt, nsteps, solver="barnes_hut", quad=True, spline=False, relative=False,
          alpha=0.001, a_old=None, adaptive=False, eta=0.03, fast=False):
    """System::Step x nsteps (global scheme) on copies: returns (state, a, a_old, time advanced).
    a_old (previous |a|, zeros if None) is read by the relative MAC and captured after every step."""
    s = {k: _f64(state[k]).copy() for k in ("x", "y", "z", "vx", "vy", "vz", "m")}
    ax, ay, az = (np.ascontiguousarray(a[:, k], dtype=np.float64).copy() for k in range(3))
    n = len(s["x"])
    ao = _f64(a_old).copy() if a_old is not None else np.zeros(n, np.float64)
    t = lib(fast).orc_steps(n, *[_ptr(s[k], _d) for k in ("x", "y", "z", "vx", "vy", "vz", "m")],
                            _ptr(ax, _d), _ptr(ay, _d), _ptr(az, _d), _ptr(ao, _d), G, eps2, 1 if spline else 0,
                            theta, 0 if solver == "direct" else 1, 1 if quad else 0, 1 if relative else 0, alpha,
                            1 if adaptive else 0, eta, dt, nsteps)
    return s, np.stack([ax, ay, az], axis=1), ao, t


def adaptive_dt(a, eta, eps, dt):
    """ChooseAdaptiveDt (Integrator.cpp:41-59) of accelerations a (n,3)."""
    ax, ay, az = (np.ascontiguousarray(a[:, k], dtype=np.float64) for k in range(3))
    return lib().orc_adaptive_dt(len(ax), _ptr(ax, _d), _ptr(ay, _d), _ptr(az, _d), eta, eps, dt)


def rungs(a, eta, eps, base, max_rung=8):
    """Block-step rungs (System.hpp:249-265) of accelerations a (n,3): (rungs, deepest)."""
    ax, ay, az = (np.ascontiguousarray(a[:, k], dtype=np.float64) for k in range(3))
    r = np.empty(len(ax), np.int32)
    rmax = lib().orc_rungs(len(ax), _ptr(ax, _d), _ptr(ay, _d), _ptr(az, _d), eta, eps, base, max_rung, _ptr(r, _i))
    return r, int(rmax)


def block_steps(state, a, G, eps2, theta, base, nsteps, eta=0.03, max_rung=8, solver="barnes_hut", quad=True,
                spline=False, fast=False):
    """System::BlockStep x nsteps on copies: returns (state, a, rungs, deepest rung, particle evaluations)."""
    s = {k: _f64(state[k]).copy() for k in ("x", "y", "z", "vx", "vy", "vz", "m")}
    ax, ay, az = (np.ascontiguousarray(a[:, k], dtype=np.float64).copy() for k in range(3))
    n = len(s["x"])
    r = np.zeros(n, np.int32)
    ev = np.zeros(1, np.int64)
    rmax = lib(fast).orc_block_steps(n, *[_ptr(s[k], _d) for k in ("x", "y", "z", "vx", "vy", "vz", "m")],
                                     _ptr(ax, _d), _ptr(ay, _d), _ptr(az, _d), _ptr(r, _i), G, eps2,
                                     1 if spline else 0, theta, 0 if solver == "direct" else 1, 1 if quad else 0,
                                     eta, base, max_rung, nsteps, _ptr(ev, _i64))
    return s, np.stack([ax, ay, az], axis=1), r, int(rmax), int(ev[0])


def total_energy(s, G, eps2, fast=False, spline=False):
    arrs = [_f64(s[k]) for k in ("x", "y", "z", "vx", "vy", "vz", "m")]
    return lib(fast).orc_total_energy(len(arrs[0]), *[_ptr(a, _d) for a in arrs], G, eps2, 1 if spline else 0)


def moments(s):
    arrs = [_f64(s[k]) for k in ("x", "y", "z", "vx", "vy", "vz", "m")]
    com, L, P = (np.empty(3, np.float64) for _ in range(3))
    lib().orc_moments(len(arrs[0]), *[_ptr(a, _d) for a in arrs], _ptr(com, _d), _ptr(L, _d), _ptr(P, _d))
    return com, L, P


def root_box(x, y, z):
    x, y, z = (_f64(a) for a in (x, y, z))
    c = np.empty(3, np.float64)
    h = np.empty(1, np.float64)
    lib().orc_root_box(len(x), _ptr(x, _d), _ptr(y, _d), _ptr(z, _d), _ptr(c, _d), _ptr(h, _d))
    return c, float(h[0])


def morton_keys(x, y, z, levels=21):
    """Derived key contract: octant digits per level from the FP64 centre recurrence."""
    x, y, z = (_f64(a) for a in (x, y, z))
    c, h = root_box(x, y, z)
    n = len(x)
    hi = np.empty(n, np.uint64)
    lo = np.empty(n, np.uint64)
    lib().orc_morton_keys(n, _ptr(x, _d), _ptr(y, _d), _ptr(z, _d), _ptr(c, _d), h, levels, _ptr(hi, _u64),
                          _ptr(lo, _u64))
    return hi, lo


def rel_rms(a, ref):
    """sqrt(sum |a-ref|^2 / sum |ref|^2)."""
    a = np.asarray(a, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(np.sqrt(np.sum((a - ref) ** 2) / np.sum(ref ** 2)))
