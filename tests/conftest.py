"""Shared test set-up: import paths, the `gpu` marker, seeded inputs."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "opengl-physics_amd"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def plummer_f32(n, seed, total_mass=1.0):
    """ic::Plummer<3> sample (FP64 oracle sampler) rounded to FP32, as the configs use."""
    import oracle
    s = oracle.plummer(n, seed, total_mass=total_mass)
    out = {k: s[k].astype(np.float32) for k in ("x", "y", "z", "vx", "vy", "vz", "m")}
    return out


def keys_at(keys21, levels):
    """The oracle's 21-level keys as a tree build with `levels`-level keys holds them
    (levels 1..levels at their places, the deeper levels zero; 40 keeps the short word)."""
    if levels >= 21:
        return keys21
    return keys21 & ~np.uint64((1 << (3 * (21 - levels))) - 1)


def f64(d, keys=("x", "y", "z", "m")):
    return [np.asarray(d[k], np.float64) for k in keys]


@pytest.fixture(scope="session")
def gpu_ctx_factory():
    import pynbg
    made = []

    def make(max_n):
        c = pynbg.Context(max_n)
        made.append(c)
        return c
    yield make
    for c in made:
        c.close()
