"""Pinning the oracle's softening kinds, per-particle walks and time-step schemes (CPU).

Each check is chosen so that a plausible mistake in the oracle fails it, without retyping
the oracle's own formula:
  * spline force (AccumPairSpline, Solvers.cpp:43-63 / SplineChi3D, Softening.hpp:39-45)
    against the enclosed mass of the published M4 cubic-spline density (Monaghan &
    Lattanzio 1985) integrated numerically, and against the reference self-test's
    "Newtonian for r >= 2h to 1e-12" (SelfTest.cpp:156-180);
  * spline potential (SplinePotential3D, Softening.hpp:49-56) against the force it must be
    the gradient of (finite differences of the oracle's energy), and the published central
    value phi(0) = -1.4 G m / h;
  * theta = 0 walks (grouped, per-particle, targets) against the direct sum (SelfTest.cpp:
    43-74 pins the same for Plummer);
  * the relative MAC (Mac.hpp:37-42) with no history equals the geometric per-particle walk,
    with alpha = 0 equals the direct sum, and its error shrinks with alpha;
  * adaptive dt (Integrator.cpp:41-59) and rungs (System.hpp:249-265) against hand values;
  * block steps (System.hpp:239-294): one rung reproduces the global KDK step bit for bit,
    the evaluation count is sum_i 2^rung_i, and the reference self-tests' energy and
    work claims hold (SelfTest.cpp:829-923).
"""
import numpy as np
import pytest
from scipy import integrate

import oracle


def _two_body(r, m_src=1.0):
    return (np.array([0.0, r]), np.zeros(2), np.zeros(2), np.array([1.0, m_src]))


def _m4_enclosed(q):
    """Enclosed-mass fraction of the normalised M4 spline density, W = 1/(pi h^3) w(q)."""
    def w(s):
        if s < 1.0:
            return 1.0 - 1.5 * s * s + 0.75 * s ** 3
        if s < 2.0:
            return 0.25 * (2.0 - s) ** 3
        return 0.0
    val, _ = integrate.quad(lambda s: 4.0 * np.pi * s * s * w(s) / np.pi, 0.0, min(q, 2.0), limit=200,
                            points=[1.0] if q > 1.0 else None)
    return val


@pytest.mark.parametrize("q", [0.05, 0.3, 0.7, 0.99, 1.0, 1.3, 1.7, 1.99, 2.0, 2.5])
def test_spline_force_is_the_m4_enclosed_mass(q):
    h = 0.1
    r = q * h
    x, y, z, m = _two_body(r)
    a = oracle.accel_direct(x, y, z, m, 1.0, h * h, spline=True)
    # G = m = 1: particle 0 is pulled toward +x by chi(q) / r^2
    chi = a[0, 0] * r * r
    assert chi == pytest.approx(_m4_enclosed(q), rel=1e-10, abs=1e-14)
    assert a[0, 1] == 0.0 and a[0, 2] == 0.0
    assert a[1, 0] == -a[0, 0]


def test_spline_is_newtonian_beyond_2h():
    h = 0.1
    rng = np.random.default_rng(777)
    worst = 0.0
    for r in rng.uniform(2.0 * h, 5.0 * h, 200):
        x, y, z, m = _two_body(r)
        a = oracle.accel_direct(x, y, z, m, 1.0, h * h, spline=True)
        worst = max(worst, abs(a[0, 0] - 1.0 / (r * r)) * r * r)
    assert worst < 1e-12


def test_spline_potential_is_the_force_integral():
    h = 0.1
    st = dict(vx=np.zeros(2), vy=np.zeros(2), vz=np.zeros(2))

    def energy(r):
        x, y, z, m = _two_body(r, 1.0)
        return oracle.total_energy(dict(st, x=x, y=y, z=z, m=m), 1.0, h * h, spline=True)

    # central value phi(0) = -1.4 G m / h (published), via two coincident particles
    x, y, z, m = _two_body(0.0)
    assert oracle.total_energy(dict(st, x=x, y=y, z=z, m=m), 1.0, h * h, spline=True) == pytest.approx(-14.0)
    for q in (0.2, 0.6, 0.95, 1.05, 1.5, 1.9, 2.5):
        r = q * h
        d = 1e-6 * h
        dEdr = (energy(r + d) - energy(r - d)) / (2 * d)
        x, y, z, m = _two_body(r)
        a = oracle.accel_direct(x, y, z, m, 1.0, h * h, spline=True)
        # dE/dr = G m1 m2 chi / r^2 = |a| of particle 0 (m2 = 1)
        assert dEdr == pytest.approx(a[0, 0], rel=1e-7)
    # continuity of the potential at the branch points q = 1 and q = 2
    for q in (1.0, 2.0):
        assert energy(q * h * (1 - 1e-12)) == pytest.approx(energy(q * h * (1 + 1e-12)), rel=1e-9)


def _cloud(n, seed=3):
    s = oracle.plummer(n, seed)
    return s, s["x"], s["y"], s["z"], s["m"]


@pytest.mark.parametrize("spline", [False, True])
def test_theta_zero_walks_equal_direct(spline):
    s, x, y, z, m = _cloud(600)
    eps2 = 0.05 ** 2
    ref = oracle.accel_direct(x, y, z, m, 1.0, eps2, spline=spline)
    T = oracle.Tree(x, y, z, m, quad=True)
    for a in (T.accel_bh(1.0, eps2, 0.0, spline=spline),
              T.accel_bh_per_particle(1.0, eps2, 0.0, spline=spline)):
        assert oracle.rel_rms(a, ref) < 1e-12
    tg = np.array([0, 5, 17, 599, 300])
    at = T.accel_bh_targets(1.0, eps2, 0.0, tg, spline=spline)
    assert oracle.rel_rms(at, ref[tg]) < 1e-12


def test_spline_bh_error_is_the_multipole_error():
    # theta = 0.5 with spline softening: the same expansion error as with Plummer softening
    # of the same length (the far field is Newtonian either way)
    s, x, y, z, m = _cloud(3000)
    eps2 = 0.02 ** 2
    T = oracle.Tree(x, y, z, m, quad=True)
    for spline in (False, True):
        ref = oracle.accel_direct(x, y, z, m, 1.0, eps2, spline=spline)
        err = oracle.rel_rms(T.accel_bh(1.0, eps2, 0.5, spline=spline), ref)
        assert 1e-6 < err < 2e-3


def test_relative_mac_limits_and_error_scaling():
    s, x, y, z, m = _cloud(2000)
    eps2 = 0.02 ** 2
    T = oracle.Tree(x, y, z, m, quad=True)
    ref = oracle.accel_direct(x, y, z, m, 1.0, eps2)
    geo = T.accel_bh_per_particle(1.0, eps2, 0.5)
    # no history (aOld = 0): the geometric per-particle walk
    np.testing.assert_array_equal(T.accel_bh_per_particle(1.0, eps2, 0.5, relative=True, alpha=0.01), geo)
    aold = np.linalg.norm(ref, axis=1)
    # alpha = 0: no cell passes, every leaf is summed exactly
    a0 = T.accel_bh_per_particle(1.0, eps2, 0.5, relative=True, alpha=0.0, a_old=aold)
    assert oracle.rel_rms(a0, ref) < 1e-12
    errs = []
    for alpha in (0.02, 0.005, 0.001):
        a = T.accel_bh_per_particle(1.0, eps2, 0.5, relative=True, alpha=alpha, a_old=aold)
        errs.append(oracle.rel_rms(a, ref))
    assert errs[0] > errs[1] > errs[2]
    # the criterion targets a fractional error of order alpha (Springel 2005)
    assert errs[2] < 1e-3 and errs[0] < 2e-2


def test_adaptive_dt_hand_values():
    a = np.array([[3.0, 4.0, 0.0], [0.0, 0.0, 0.0], [1.0, 0.0, 0.0]])
    # max |a| = 5: dt = eta sqrt(eps / 5)
    assert oracle.adaptive_dt(a, 0.1, 0.05, 1.0) == pytest.approx(0.1 * np.sqrt(0.01))
    # clamped to the nominal dt and to dt / 4096
    assert oracle.adaptive_dt(a, 100.0, 0.05, 1e-3) == 1e-3
    assert oracle.adaptive_dt(a, 1e-9, 0.05, 1.0) == 1.0 / 4096.0
    # no acceleration or no softening length: the nominal dt
    assert oracle.adaptive_dt(np.zeros((4, 3)), 0.1, 0.05, 0.3) == 0.3
    assert oracle.adaptive_dt(a, 0.1, 0.0, 0.3) == 0.3


def test_rungs_hand_values():
    eps, eta, base = 0.01, 0.1, 0.01
    # tau = eta sqrt(eps/|a|) = 0.01 / sqrt(|a|), rung = ceil(log2(base / tau)) = ceil(log2 sqrt|a|):
    # |a| = 0.5 -> tau > base (rung 0), 3 -> 1.73 (1), 10 -> 3.16 (2), 50 -> 7.07 (3), 1e12 -> cap
    amag = np.array([0.0, 0.5, 3.0, 10.0, 50.0, 1e12])
    a = np.stack([amag, np.zeros(6), np.zeros(6)], axis=1)
    r, rmax = oracle.rungs(a, eta, eps, base, max_rung=6)
    assert r.tolist() == [0, 0, 1, 2, 3, 6] and rmax == 6
    r, rmax = oracle.rungs(a, eta, eps, base, max_rung=-3)
    assert r.tolist() == [0] * 6 and rmax == 0


def _cloud_and_binary():
    """The reference's block-step test system (SelfTest.cpp:829-852)."""
    c = oracle.plummer(200, 12)
    s = {k: np.zeros(202) for k in ("x", "y", "z", "vx", "vy", "vz", "m")}
    for k in s:
        s[k][:200] = c[k]
    s["x"][:200] -= 8.0
    s["x"][200], s["x"][201] = 7.96, 8.04
    vt = 0.5 * np.sqrt(2.0 / 0.08)
    s["vy"][200], s["vy"][201] = -vt, vt
    s["m"][200] = s["m"][201] = 1.0
    return s


def test_block_step_single_rung_is_the_global_step():
    s, x, y, z, m = _cloud(300)
    eps2 = 0.05 ** 2
    a = oracle.accel_direct(x, y, z, m, 1.0, eps2)
    # eta large: every tau exceeds the coarse step -> rung 0 for all
    b, ab, r, rmax, ev = oracle.block_steps(s, a, 1.0, eps2, 0.5, 1e-3, 3, eta=1e6, solver="direct")
    g, ag, _, _ = oracle.steps(s, a, 1.0, eps2, 0.5, 1e-3, 3, solver="direct")
    assert rmax == 0 and ev == 3 * 300
    for k in ("x", "y", "z", "vx", "vy", "vz"):
        np.testing.assert_array_equal(b[k], g[k])
    np.testing.assert_array_equal(ab, ag)


def test_block_scheme_energy_and_work_on_cloud_plus_binary():
    s = _cloud_and_binary()
    eps2 = 0.01 ** 2
    a = oracle.accel_direct(s["x"], s["y"], s["z"], s["m"], 1.0, eps2)
    e0 = oracle.total_energy(s, 1.0, eps2)
    # 40 coarse steps of 5e-3 (t = 0.2) with the reference's eta and rung cap
    b, ab, r, rmax, ev = oracle.block_steps(s, a, 1.0, eps2, 0.5, 5e-3, 40, eta=0.02, max_rung=8, solver="direct")
    # the tight pair takes the deepest rung, the bulk of the cloud stays coarse
    assert rmax > 2 and r[200] == rmax and r[201] == rmax and np.mean(r[:200] <= 1) > 0.8
    # evaluations of the last coarse step = sum_i 2^rung_i (each particle once per own step)
    b1, _, r1, _, ev1 = oracle.block_steps(b, ab, 1.0, eps2, 0.5, 5e-3, 1, eta=0.02, max_rung=8, solver="direct")
    assert ev1 == int(np.sum(2 ** r1.astype(np.int64)))
    drift = abs(oracle.total_energy(b, 1.0, eps2) - e0) / abs(e0)
    assert drift < 1e-2
    # the same time with global adaptive steps takes at least twice the particle evaluations
    g, ag, t, nsteps = s, a, 0.0, 0
    while t < 40 * 5e-3:
        g, ag, _, dt = oracle.steps(g, ag, 1.0, eps2, 0.5, 5e-3, 16, solver="direct", adaptive=True, eta=0.02)
        t += dt
        nsteps += 16
    assert nsteps * 202 >= 2 * ev


def test_adaptive_eccentric_orbit_energy():
    """SelfTest.cpp:391-455: e = 0.9 Kepler orbit, adaptive dt (eta 0.015), 3 periods."""
    e, r0 = 0.9, 1.0
    a_sm = r0 / (1 + e)
    vrel = np.sqrt(2.0 * (2.0 / r0 - 1.0 / a_sm))
    period = 2 * np.pi * np.sqrt(a_sm ** 3 / 2.0)
    s = dict(x=np.array([-0.5, 0.5]), y=np.zeros(2), z=np.zeros(2), vx=np.zeros(2),
             vy=np.array([-vrel / 2, vrel / 2]), vz=np.zeros(2), m=np.ones(2))
    eps2 = 1e-4 ** 2
    a = oracle.accel_direct(s["x"], s["y"], s["z"], s["m"], 1.0, eps2)
    e0 = oracle.total_energy(s, 1.0, eps2)
    t, emin, emax, steps = 0.0, e0, e0, 0
    while t < 3 * period:
        s, a, _, dt = oracle.steps(s, a, 1.0, eps2, 0.5, period / 600.0, 32, solver="direct", adaptive=True,
                                   eta=0.015)
        t += dt
        steps += 32
        en = oracle.total_energy(s, 1.0, eps2)
        emin, emax = min(emin, en), max(emax, en)
    assert (emax - emin) / abs(0.5 * (emax + emin)) < 1e-3
