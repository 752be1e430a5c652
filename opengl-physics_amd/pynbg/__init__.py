"""Python binding of libnbg_hip.so (the C ABI declared in include/nbg.h).

Thin ctypes layer used by the tests, the benchmark and the smoke check: numpy arrays in,
numpy arrays out.  There is no CPU fallback -- if the HIP library is missing this module
raises on import of the library.
"""
import ctypes
import os

import numpy as np

_PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG_DIR, "libnbg_hip.so")
HOST_LIB_PATH = os.path.join(_PKG_DIR, "libngrav_host.so")

SOLVER_DIRECT = 0
SOLVER_BARNES_HUT = 1
MONOPOLE = 0
QUADRUPOLE = 1
FLAG_KAHAN = 1
FLAG_COUNT = 2
FLAG_LEARNED = 4
FLAG_NO_GRAPH = 8
FLAG_STAGE_TIMES = 16
FLAG_PER_PARTICLE = 32
SOFT_PLUMMER = 0
SOFT_SPLINE = 1
MAC_GEOMETRIC = 0
MAC_RELATIVE = 1

ERRORS = {0: "NBG_OK", -1: "NBG_E_INVALID", -2: "NBG_E_OOM", -3: "NBG_E_CAPACITY", -4: "NBG_E_HIP",
          -5: "NBG_E_STATE"}

EXPORTED = ["nbg_create", "nbg_destroy", "nbg_last_error", "nbg_upload", "nbg_upload_batch", "nbg_upload_accel",
            "nbg_accel", "nbg_step", "nbg_download", "nbg_tree_export", "nbg_interaction_counts", "nbg_diagnostics",
            "nbg_accel_direct_targets", "nbg_learn_set_weights", "nbg_learn_get_weights", "nbg_learn_collect",
            "nbg_learn_clear", "nbg_learn_pool", "nbg_learn_standardize", "nbg_learn_train", "nbg_learn_grad",
            "nbg_tuning_get", "nbg_tuning_set", "nbg_autotune", "nbg_diagnostics_system", "nbg_accel_bh_targets",
            "nbg_rungs", "nbg_accel_direct_noalloc", "nbg_upload_f64"]

LEARN_FEATURES = 13
LEARN_HIDDEN = 64
LEARN_NTRAIN = 9411
LEARN_NPARAM = 32 + LEARN_NTRAIN
LEARN_ACT_SLOT = 15          # blob[15] (an unused feature mean): the model's activation
LEARN_ACT = {"silu": 0, "relu": 1}


class Params(ctypes.Structure):
    _fields_ = [("solver", ctypes.c_int), ("multipole", ctypes.c_int), ("G", ctypes.c_double),
                ("eps2", ctypes.c_double), ("theta", ctypes.c_double), ("dt", ctypes.c_double),
                ("flags", ctypes.c_uint32), ("learn_band", ctypes.c_double), ("softening", ctypes.c_int32),
                ("mac", ctypes.c_int32), ("alpha", ctypes.c_double), ("adaptive", ctypes.c_int32),
                ("block", ctypes.c_int32), ("eta", ctypes.c_double), ("block_max_rung", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("learn_amin", ctypes.c_double)]


class Tuning(ctypes.Structure):
    _fields_ = [("walk_kernel", ctypes.c_int32), ("near_tile", ctypes.c_int32), ("direct_tile", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class Stats(ctypes.Structure):
    _fields_ = [("n_nodes", ctypes.c_int64), ("n_leaves", ctypes.c_int64), ("max_depth", ctypes.c_int32),
                ("deep_keys", ctypes.c_int32), ("n_p2p", ctypes.c_int64), ("n_m2p", ctypes.c_int64),
                ("ms_keys", ctypes.c_double), ("ms_sort", ctypes.c_double), ("ms_build", ctypes.c_double),
                ("ms_upsweep", ctypes.c_double), ("ms_walk", ctypes.c_double), ("ms_integrate", ctypes.c_double),
                ("ms_total", ctypes.c_double), ("n_learned", ctypes.c_int64), ("ms_learned", ctypes.c_double),
                ("t_advanced", ctypes.c_double), ("dt_last", ctypes.c_double), ("particle_evals", ctypes.c_int64),
                ("rung_max", ctypes.c_int32), ("key_levels", ctypes.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None
_f = ctypes.POINTER(ctypes.c_float)
_d = ctypes.POINTER(ctypes.c_double)
_i32 = ctypes.POINTER(ctypes.c_int32)
_i64 = ctypes.POINTER(ctypes.c_int64)
_u64 = ctypes.POINTER(ctypes.c_uint64)


def load():
    """Load libnbg_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built: run __graft_entry__.build() or make -C opengl-physics_amd")
    L = ctypes.CDLL(LIB_PATH)
    L.nbg_create.restype = ctypes.c_void_p
    L.nbg_create.argtypes = [ctypes.c_int, ctypes.c_int64]
    L.nbg_destroy.argtypes = [ctypes.c_void_p]
    L.nbg_last_error.restype = ctypes.c_char_p
    L.nbg_last_error.argtypes = [ctypes.c_void_p]
    L.nbg_upload.argtypes = [ctypes.c_void_p, ctypes.c_int64] + [_f] * 7
    L.nbg_upload_batch.argtypes = [ctypes.c_void_p, ctypes.c_int32, _i64] + [_f] * 7
    L.nbg_upload_accel.argtypes = [ctypes.c_void_p] + [_f] * 3
    L.nbg_upload_f64.argtypes = [ctypes.c_void_p, ctypes.c_int64] + [_d] * 3 + [_f] * 4
    L.nbg_accel.argtypes = [ctypes.c_void_p, ctypes.POINTER(Params), ctypes.POINTER(Stats)]
    L.nbg_step.argtypes = [ctypes.c_void_p, ctypes.POINTER(Params), ctypes.c_int32, ctypes.POINTER(Stats)]
    L.nbg_download.argtypes = [ctypes.c_void_p] + [_f] * 9
    L.nbg_tree_export.argtypes = [ctypes.c_void_p, _i64, _u64, _i32, _d, _i32, _i32, _f]
    L.nbg_interaction_counts.argtypes = [ctypes.c_void_p, ctypes.c_int64, _i64, _i64]
    L.nbg_diagnostics.argtypes = [ctypes.c_void_p, ctypes.POINTER(Params), ctypes.c_int32, _d]
    L.nbg_accel_direct_targets.argtypes = [ctypes.c_void_p, ctypes.POINTER(Params), ctypes.c_int64, _i64] + [_f] * 3
    L.nbg_learn_set_weights.argtypes = [ctypes.c_void_p, _f, ctypes.c_int64]
    L.nbg_learn_get_weights.argtypes = [ctypes.c_void_p, _f, ctypes.c_int64]
    L.nbg_learn_collect.argtypes = [ctypes.c_void_p, ctypes.POINTER(Params), ctypes.c_int64, _i64]
    L.nbg_learn_clear.argtypes = [ctypes.c_void_p]
    L.nbg_learn_pool.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, _f, _f]
    L.nbg_learn_standardize.argtypes = [ctypes.c_void_p, ctypes.c_int64]
    L.nbg_learn_train.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_float, ctypes.c_uint64, _f]
    L.nbg_learn_grad.argtypes = [ctypes.c_void_p, ctypes.c_int64, _f, _f, _f, _f, _f]
    L.nbg_tuning_get.argtypes = [ctypes.c_void_p, ctypes.POINTER(Tuning)]
    L.nbg_tuning_set.argtypes = [ctypes.c_void_p, ctypes.POINTER(Tuning)]
    L.nbg_autotune.argtypes = [ctypes.c_void_p, ctypes.POINTER(Params), ctypes.c_int32, ctypes.POINTER(Tuning), _d,
                               ctypes.c_int32, _i32]
    L.nbg_diagnostics_system.argtypes = [ctypes.c_void_p, ctypes.POINTER(Params), ctypes.c_int32, ctypes.c_int32, _d]
    L.nbg_accel_bh_targets.argtypes = [ctypes.c_void_p, ctypes.POINTER(Params), ctypes.c_int64, _i64] + [_f] * 3
    L.nbg_rungs.argtypes = [ctypes.c_void_p, _i32]
    L.nbg_destroy.restype = None
    for name in EXPORTED:
        if name not in ("nbg_create", "nbg_destroy", "nbg_last_error"):
            getattr(L, name).restype = ctypes.c_int
    _lib = L
    return L


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None else None


def _f32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float32)


class NbgError(RuntimeError):
    pass


def make_params(solver="barnes_hut", multipole="quadrupole", G=1.0, eps=None, eps2=None, theta=0.5, dt=1e-3,
                kahan=True, count=False, no_graph=False, learned=False, stage_times=False, learn_band=0.0,
                softening="plummer", mac="geometric", alpha=0.001, per_particle=False, adaptive=False,
                block=False, eta=0.03, block_max_rung=8, learn_amin=0.0):
    """nbg_params with the reference's StepParams defaults (Solvers.hpp:22-40): eps is the
    Plummer softening length or the spline length h (eps2 = eps^2 either way)."""
    if eps2 is None:
        eps2 = (eps if eps is not None else 0.05) ** 2
    if softening not in ("plummer", "spline") or mac not in ("geometric", "relative"):
        raise ValueError("softening must be plummer|spline and mac geometric|relative")
    flags = (FLAG_KAHAN if kahan else 0) | (FLAG_COUNT if count else 0) | (FLAG_NO_GRAPH if no_graph else 0) | \
        (FLAG_LEARNED if learned else 0) | (FLAG_STAGE_TIMES if stage_times else 0) | \
        (FLAG_PER_PARTICLE if per_particle else 0)
    return Params(SOLVER_DIRECT if solver == "direct" else SOLVER_BARNES_HUT,
                  QUADRUPOLE if multipole == "quadrupole" else MONOPOLE, float(G), float(eps2), float(theta),
                  float(dt), flags, float(learn_band), SOFT_SPLINE if softening == "spline" else SOFT_PLUMMER,
                  MAC_RELATIVE if mac == "relative" else MAC_GEOMETRIC, float(alpha), 1 if adaptive else 0,
                  1 if block else 0, float(eta), int(block_max_rung), 0, float(learn_amin))


class Context:
    """One device context (nbg_ctx*) owning the particle state and every device buffer."""

    def __init__(self, max_n, device=0):
        self._L = load()
        self._h = self._L.nbg_create(device, int(max_n))
        if not self._h:
            raise NbgError(f"nbg_create(device={device}, max_n={max_n}) failed")
        self.n = 0

    def close(self):
        if getattr(self, "_h", None):
            self._L.nbg_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc, what):
        if rc != 0:
            msg = self._L.nbg_last_error(self._h).decode()
            raise NbgError(f"{what} -> {ERRORS.get(rc, rc)}: {msg}")

    def upload(self, x, y, z, m, vx=None, vy=None, vz=None):
        arrs = [_f32(a) for a in (x, y, z, m, vx, vy, vz)]
        self.n = len(arrs[0])
        self._keep = arrs
        self._check(self._L.nbg_upload(self._h, self.n, *[_p(a, _f) for a in arrs]), "nbg_upload")

    def upload_f64(self, x, y, z, m, vx=None, vy=None, vz=None):
        """FP64 positions (nbg_upload_f64): the tree from them, the grouped walk on
        leaf-relative differences; masses and velocities are rounded to FP32."""
        pos = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z)]
        rest = [_f32(a) for a in (m, vx, vy, vz)]
        self.n = len(pos[0])
        self._keep = pos + rest
        self._check(self._L.nbg_upload_f64(self._h, self.n, *[_p(a, _d) for a in pos], *[_p(a, _f) for a in rest]),
                    "nbg_upload_f64")

    def upload_batch(self, offsets, x, y, z, m, vx=None, vy=None, vz=None):
        off = np.ascontiguousarray(offsets, dtype=np.int64)
        arrs = [_f32(a) for a in (x, y, z, m, vx, vy, vz)]
        self.n = int(off[-1])
        self._check(self._L.nbg_upload_batch(self._h, len(off) - 1, _p(off, _i64), *[_p(a, _f) for a in arrs]),
                    "nbg_upload_batch")

    def upload_accel(self, ax=None, ay=None, az=None):
        arrs = [_f32(a) for a in (ax, ay, az)]
        self._check(self._L.nbg_upload_accel(self._h, *[_p(a, _f) for a in arrs]), "nbg_upload_accel")

    def accel(self, params, stats=True):
        st = Stats()
        self._check(self._L.nbg_accel(self._h, ctypes.byref(params), ctypes.byref(st) if stats else None),
                    "nbg_accel")
        return st

    def step(self, params, nsteps, stats=True):
        st = Stats()
        self._check(self._L.nbg_step(self._h, ctypes.byref(params), int(nsteps),
                                     ctypes.byref(st) if stats else None), "nbg_step")
        return st

    def download(self):
        out = {k: np.empty(self.n, np.float32) for k in ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az")}
        self._check(self._L.nbg_download(self._h, *[_p(out[k], _f) for k in
                                                    ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az")]),
                    "nbg_download")
        return out

    def accelerations(self):
        d = self.download()
        return np.stack([d["ax"], d["ay"], d["az"]], axis=1)

    def tree_export(self):
        nn = ctypes.c_int64(0)
        self._check(self._L.nbg_tree_export(self._h, ctypes.byref(nn), None, None, None, None, None, None),
                    "nbg_tree_export(count)")
        n_nodes = nn.value
        keys = np.empty(self.n, np.uint64)
        order = np.empty(self.n, np.int32)
        center = np.empty((n_nodes, 4), np.float64)
        topo = np.empty((n_nodes, 5), np.int32)
        children = np.empty((n_nodes, 8), np.int32)
        moments = np.empty((n_nodes, 10), np.float32)
        nn = ctypes.c_int64(n_nodes)
        self._check(self._L.nbg_tree_export(self._h, ctypes.byref(nn), _p(keys, _u64), _p(order, _i32),
                                            _p(center, _d), _p(topo, _i32), _p(children, _i32), _p(moments, _f)),
                    "nbg_tree_export")
        return dict(keys=keys, order=order, center=center[:, :3], half=center[:, 3], parent=topo[:, 0],
                    first=topo[:, 1], count=topo[:, 2], depth=topo[:, 3], child_count=topo[:, 4], children=children,
                    com=moments[:, :3], mass=moments[:, 3], max_radius=moments[:, 4], quad=moments[:, 5:10],
                    n_nodes=n_nodes)

    def interaction_counts(self, n_nodes):
        f = np.empty(n_nodes, np.int64)
        q = np.empty(n_nodes, np.int64)
        self._check(self._L.nbg_interaction_counts(self._h, n_nodes, _p(f, _i64), _p(q, _i64)),
                    "nbg_interaction_counts")
        return f, q

    def diagnostics(self, params, with_potential=True):
        out = np.empty(11, np.float64)
        self._check(self._L.nbg_diagnostics(self._h, ctypes.byref(params), 1 if with_potential else 0,
                                            _p(out, _d)), "nbg_diagnostics")
        return dict(kinetic=out[0], potential=out[1], energy=out[0] + out[1], com=out[2:5].copy(),
                    L=out[5:8].copy(), P=out[8:11].copy())

    def direct_targets(self, params, targets):
        t = np.ascontiguousarray(targets, dtype=np.int64)
        ax, ay, az = (np.empty(len(t), np.float32) for _ in range(3))
        self._check(self._L.nbg_accel_direct_targets(self._h, ctypes.byref(params), len(t), _p(t, _i64),
                                                     _p(ax, _f), _p(ay, _f), _p(az, _f)), "nbg_accel_direct_targets")
        return np.stack([ax, ay, az], axis=1)

    # ------------------------------------------------------------------ tuning
    def bh_targets(self, params, targets):
        """Per-particle geometric Barnes-Hut walk onto targets (ComputeAccelBarnesHutTargets)."""
        t = np.ascontiguousarray(targets, dtype=np.int64)
        out = [np.empty(len(t), np.float32) for _ in range(3)]
        self._check(self._L.nbg_accel_bh_targets(self._h, ctypes.byref(params), len(t), _p(t, _i64),
                                                 *[_p(o, _f) for o in out]), "nbg_accel_bh_targets")
        return np.stack(out, axis=1)

    def rungs(self):
        r = np.empty(self.n, np.int32)
        self._check(self._L.nbg_rungs(self._h, _p(r, _i32)), "nbg_rungs")
        return r

    def tuning(self):
        t = Tuning()
        self._check(self._L.nbg_tuning_get(self._h, ctypes.byref(t)), "nbg_tuning_get")
        return t

    def set_tuning(self, walk_kernel=0, near_tile=16, direct_tile=4):
        t = Tuning(int(walk_kernel), int(near_tile), int(direct_tile), 0)
        self._check(self._L.nbg_tuning_set(self._h, ctypes.byref(t)), "nbg_tuning_set")

    def autotune(self, params, reps=3):
        best = Tuning()
        ms = np.zeros(16, np.float64)
        n = ctypes.c_int32(0)
        self._check(self._L.nbg_autotune(self._h, ctypes.byref(params), int(reps), ctypes.byref(best), _p(ms, _d), 16,
                                         ctypes.byref(n)), "nbg_autotune")
        return best, ms[:n.value].copy()

    def diagnostics_system(self, params, sys, with_potential=True):
        out = np.empty(11, np.float64)
        self._check(self._L.nbg_diagnostics_system(self._h, ctypes.byref(params), int(sys), 1 if with_potential else 0,
                                                   _p(out, _d)), "nbg_diagnostics_system")
        return dict(kinetic=out[0], potential=out[1], energy=out[0] + out[1], com=out[2:5].copy(),
                    L=out[5:8].copy(), P=out[8:11].copy())

    # ------------------------------------------------------------------ learned correction
    def learn_set_weights(self, blob):
        b = np.ascontiguousarray(blob, dtype=np.float32)
        self._check(self._L.nbg_learn_set_weights(self._h, _p(b, _f), len(b)), "nbg_learn_set_weights")

    def learn_get_weights(self):
        b = np.empty(LEARN_NPARAM, np.float32)
        self._check(self._L.nbg_learn_get_weights(self._h, _p(b, _f), len(b)), "nbg_learn_get_weights")
        return b

    def learn_collect(self, params, max_new=0):
        n = ctypes.c_int64(0)
        self._check(self._L.nbg_learn_collect(self._h, ctypes.byref(params), int(max_new), ctypes.byref(n)),
                    "nbg_learn_collect")
        return n.value

    def learn_clear(self):
        self._check(self._L.nbg_learn_clear(self._h), "nbg_learn_clear")

    def learn_pool(self, offset, count):
        f = np.empty((count, 16), np.float32)
        t = np.empty((count, 4), np.float32)
        self._check(self._L.nbg_learn_pool(self._h, int(offset), int(count), _p(f, _f), _p(t, _f)),
                    "nbg_learn_pool")
        return f, t

    def learn_standardize(self, count=0):
        self._check(self._L.nbg_learn_standardize(self._h, int(count)), "nbg_learn_standardize")

    def learn_train(self, steps, batch, lr, seed=0):
        loss = np.empty(max(1, steps), np.float32)
        self._check(self._L.nbg_learn_train(self._h, int(steps), int(batch), float(lr), int(seed), _p(loss, _f)),
                    "nbg_learn_train")
        return loss[:steps]

    def learn_grad(self, feats, targets):
        f = np.ascontiguousarray(feats, dtype=np.float32)
        t = np.ascontiguousarray(targets, dtype=np.float32)
        n = len(f)
        out = np.empty((n, 3), np.float32)
        grad = np.empty(LEARN_NTRAIN, np.float32)
        loss = ctypes.c_float(0)
        self._check(self._L.nbg_learn_grad(self._h, n, _p(f, _f), _p(t, _f), _p(out, _f), ctypes.byref(loss),
                                           _p(grad, _f)), "nbg_learn_grad")
        return out, float(loss.value), grad


# ---------------------------------------------------------------------------- host library
_host = None


def load_host():
    """libngrav_host.so: host C++ above the C ABI (initial-condition samplers)."""
    global _host
    if _host is not None:
        return _host
    if not os.path.exists(HOST_LIB_PATH):
        raise RuntimeError(f"{HOST_LIB_PATH} is not built")
    H = ctypes.CDLL(HOST_LIB_PATH)
    H.ngh_ic_plummer.argtypes = [ctypes.c_int, ctypes.c_uint32] + [ctypes.c_double] * 3 + [_d] * 7
    H.ngh_ic_hernquist.argtypes = [ctypes.c_int, ctypes.c_uint32] + [ctypes.c_double] * 3 + [_d] * 7
    H.ngh_ic_king.argtypes = [ctypes.c_int, ctypes.c_uint32] + [ctypes.c_double] * 4 + [_d] * 7
    H.ngh_ic_cluster.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_int] + [_d] * 7
    H.ngh_ic_cluster.restype = ctypes.c_int
    H.ngh_ensemble_sizes.argtypes = [ctypes.c_int32, ctypes.c_uint32, ctypes.c_int64, ctypes.c_int64, _i64, _i32]
    H.ngh_ensemble_sizes.restype = ctypes.c_int64
    for f in (H.ngh_ic_plummer, H.ngh_ic_hernquist, H.ngh_ic_king):
        f.restype = ctypes.c_int
    _host = H
    return H


def _ic(fn, n, seed, *scalars):
    out = {k: np.empty(n, np.float64) for k in ("x", "y", "z", "vx", "vy", "vz", "m")}
    rc = fn(int(n), int(seed), *[float(v) for v in scalars],
            *[_p(out[k], _d) for k in ("x", "y", "z", "vx", "vy", "vz", "m")])
    if rc != 0:
        raise NbgError(f"initial-condition sampler failed (rc={rc})")
    return out


def ic_plummer(n, seed, total_mass=1.0, scale_radius=1.0, G=1.0):
    """ic::Plummer<3>, FP64 arrays x,y,z,vx,vy,vz,m."""
    return _ic(load_host().ngh_ic_plummer, n, seed, total_mass, scale_radius, G)


def ic_hernquist(n, seed, total_mass=1.0, scale_radius=1.0, G=1.0):
    """ic::Hernquist<3>, FP64 arrays x,y,z,vx,vy,vz,m."""
    return _ic(load_host().ngh_ic_hernquist, n, seed, total_mass, scale_radius, G)


def ic_king(n, seed, w0=6.0, total_mass=1.0, core_radius=1.0, G=1.0):
    """ic::King<3>(W0), FP64 arrays x,y,z,vx,vy,vz,m."""
    return _ic(load_host().ngh_ic_king, n, seed, total_mass, core_radius, w0, G)


def ic_cluster(n, seed, cold=False):
    """Uniform-sphere cluster (Cluster<3>; cold=True: ColdUniform, the cold_collapse
    scenario), FP64 arrays x,y,z,vx,vy,vz,m."""
    out = {k: np.empty(n, np.float64) for k in ("x", "y", "z", "vx", "vy", "vz", "m")}
    rc = load_host().ngh_ic_cluster(int(n), int(seed), 1 if cold else 0,
                                    *[_p(out[k], _d) for k in ("x", "y", "z", "vx", "vy", "vz", "m")])
    if rc != 0:
        raise NbgError(f"initial-condition sampler failed (rc={rc})")
    return out


def ensemble(nclusters=512, seed=0, nmin=1024, nmax=65536, w0=6.0):
    """Batched ensemble: cluster sizes N_k = round(exp(U[ln nmin, ln nmax])) (mt19937(seed)),
    kinds alternating Plummer (even k) / King(w0) (odd k), cluster k drawn with seed k,
    total mass 1 each.  Returns (offsets int64[n+1], FP32 state dict, kinds)."""
    H = load_host()
    sizes = np.empty(nclusters, np.int64)
    kinds = np.empty(nclusters, np.int32)
    total = H.ngh_ensemble_sizes(int(nclusters), int(seed), int(nmin), int(nmax), _p(sizes, _i64), _p(kinds, _i32))
    if total < 0:
        raise ValueError("bad ensemble arguments")
    off = np.zeros(nclusters + 1, np.int64)
    off[1:] = np.cumsum(sizes)
    st = {k: np.empty(total, np.float32) for k in ("x", "y", "z", "vx", "vy", "vz", "m")}
    for k in range(nclusters):
        ic = ic_plummer(int(sizes[k]), k) if kinds[k] == 0 else ic_king(int(sizes[k]), k, w0=w0)
        for key in st:
            st[key][off[k]:off[k + 1]] = ic[key]
    return off, st, kinds


def to_f32(state):
    """Round an FP64 state dict to the FP32 arrays the device path takes."""
    return {k: np.asarray(v, np.float32) for k, v in state.items()}
