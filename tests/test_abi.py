"""CPU checks of the drop-in boundary: libnbg_hip.so loads and exports every entry point
that include/nbg.h declares (no compute calls -- no GPU is needed)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "nbg.h")
LIB = os.path.join(ROOT, "opengl-physics_amd", "libnbg_hip.so")


def declared_functions(path):
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nbg_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_api():
    names = declared_functions(HEADER)
    for must in ("nbg_create", "nbg_destroy", "nbg_upload", "nbg_accel", "nbg_step", "nbg_download",
                 "nbg_tree_export", "nbg_diagnostics"):
        assert must in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.fail(f"{LIB} not built")
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared_functions(HEADER) if not hasattr(lib, n)]
    assert not missing, missing


def test_exported_symbols_are_c_linkage():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (nbg_[a-z_0-9]+)$", out, flags=re.M))
    assert set(declared_functions(HEADER)) <= exported


def test_create_rejects_sizes_outside_its_range():
    # The range check comes before any HIP call, so it runs without a GPU.  The force
    # kernels address sorted particles with 32-bit byte offsets, hence the 2^28 - 1024 cap.
    lib = ctypes.CDLL(LIB)
    lib.nbg_create.restype = ctypes.c_void_p
    lib.nbg_create.argtypes = [ctypes.c_int, ctypes.c_int64]
    for bad in (0, -5, (1 << 28) - 1023, 1 << 30):
        assert lib.nbg_create(0, bad) is None, bad


def test_binding_covers_the_header():
    import sys
    sys.path.insert(0, os.path.join(ROOT, "opengl-physics_amd"))
    import pynbg
    assert set(declared_functions(HEADER)) <= set(pynbg.EXPORTED) | {"nbg_upload_batch"}


HOST_HEADER = os.path.join(ROOT, "include", "ngrav_host.h")
HOST_LIB = os.path.join(ROOT, "opengl-physics_amd", "libngrav_host.so")


def test_host_library_exports_its_header():
    names = [n for n in re.findall(r"\b(ngh_[a-z_0-9]+)\s*\(", re.sub(r"/\*.*?\*/", "", open(HOST_HEADER).read(),
                                                                    flags=re.S))]
    lib = ctypes.CDLL(HOST_LIB)
    assert names and all(hasattr(lib, n) for n in names)


def test_deck_runner_rejects_unsupported_features(tmp_path):
    """Error path on the CPU: the deck is validated before any device work."""
    exe = os.path.join(ROOT, "opengl-physics_amd", "nbg_deck")
    r = subprocess.run([exe, "--deck", os.path.join(ROOT, "tests", "decks", "fmm_unsupported.toml"), "--out",
                        str(tmp_path), "--quiet"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "fmm" in r.stderr


def test_dump_ic_writes_the_bench_sphere(tmp_path):
    """--dump-ic (main.cpp:249-280): the seed-1 Plummer sphere as raw float32 records
    x,y,z,m,vx,vy,vz -- no GPU involved."""
    import sys
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "opengl-physics_amd"))
    import pynbg
    exe = os.path.join(ROOT, "opengl-physics_amd", "nbg_deck")
    out = tmp_path / "ic.bin"
    r = subprocess.run([exe, "--dump-ic", str(out), "--bench-n", "1500"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert "N=1500" in r.stderr and "seed=1" in r.stderr
    rec = np.fromfile(out, dtype="<f4").reshape(-1, 7)
    ic = pynbg.ic_plummer(1500, 1)
    want = np.stack([ic[k] for k in ("x", "y", "z", "m", "vx", "vy", "vz")], axis=1).astype(np.float32)
    np.testing.assert_array_equal(rec, want)


def test_cli_rejects_unknown_bench_solver():
    exe = os.path.join(ROOT, "opengl-physics_amd", "nbg_deck")
    r = subprocess.run([exe, "--bench", "spherical_fmm", "--bench-n", "100"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 2 and "unknown solver" in r.stderr
