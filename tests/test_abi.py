"""Host-side checks of the C ABI: the library loads, exports every function
include/gcsim.h declares, and reports the no-device condition as an error
(no silent CPU fallback).  Runs without a GPU."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

import gcsim
from gcsim import _lib as L

ROOT = Path(__file__).resolve().parent.parent


def _declared():
    text = (ROOT / "include" / "gcsim.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gc_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = gcsim.load()
    names = _declared()
    assert len(names) >= 25
    for name in names:
        assert hasattr(lib, name), name
        assert name in L.SIGNATURES, f"{name} has no ctypes signature"
    out = subprocess.run(["nm", "-D", "--defined-only", str(L.LIB_PATH)], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (gc_\w+)", out))
    assert set(names) <= exported


def test_defaults_match_reference():
    p = L.Params()
    gcsim.load().gc_default_params(C.byref(p))
    # PhysicsParams (simulation_params.hpp:5-8), FmmOptions (fmm_solver.hpp:13-18)
    assert (p.gravitational_constant, p.softening) == (1.0, 0.02)
    assert (p.theta, p.leaf_capacity, p.max_depth, p.expansion_order) == (0.6, 16, 32, 4)


def test_status_strings():
    lib = gcsim.load()
    for code in range(8):
        assert lib.gc_status_string(code)
    assert b"unknown" in lib.gc_status_string(99)


@pytest.mark.skipif(gcsim.device_count() > 0, reason="a GPU is visible")
def test_no_device_is_an_error():
    assert not gcsim.solver_available()
    with pytest.raises(gcsim.GcError) as e:
        gcsim.Engine("hip-direct", "fp64")
    assert e.value.status == L.GC_ERR_NO_DEVICE


def test_bad_arguments_rejected_on_host():
    lib = gcsim.load()
    assert lib.gc_create(None, 0, None) == L.GC_ERR_INVALID_ARGUMENT
    ctx = C.c_void_p()
    p = gcsim.make_params()
    p.softening = -1.0
    assert lib.gc_create(C.byref(ctx), 0, C.byref(p)) == L.GC_ERR_INVALID_ARGUMENT
    assert b"softening" in lib.gc_last_error(None)
    for bad in (dict(solver="tree"), dict(precision="fp16")):
        with pytest.raises(ValueError):
            gcsim.make_params(**bad)
    assert lib.gc_step(None, 0.1, 1) == L.GC_ERR_INVALID_ARGUMENT


def test_cpp_interface_error_path():
    exe = ROOT / "tests" / "cpp" / "hip_solver_test"
    if gcsim.device_count() > 0:
        pytest.skip("GPU visible: covered by the GPU test")
    assert exe.exists(), "build() compiles tests/cpp/hip_solver_test"
    res = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "passed (no device)" in res.stdout


# ---------------------------------------------------------------- generators
def test_two_disk_config0_layout():
    pos, vel, mass = gcsim.generators.config0()
    assert pos.shape == (1000, 3) and mass.shape == (1000,)
    assert mass[0] == 1.0 and mass[500] == 1.0
    assert np.all(mass[1:500] == 2e-4) and np.all(mass[501:] == 2e-4)
    assert np.allclose(pos[0], [-2.5, -0.75, 0]) and np.allclose(pos[500], [2.5, 0.75, 0])
    assert np.allclose(vel[0], [0.35, 0, 0]) and np.allclose(vel[500], [-0.35, 0, 0])
    r = np.linalg.norm(pos[1:500, :2] - pos[0, :2], axis=1)
    assert 0.05 <= r.min() and r.max() <= 1.0 + 1e-12
    # exponential profile, scale 0.3, truncated to [0.05, 1]: mean radius
    h, a, b = 0.3, 0.05, 1.0
    mean = (a * np.exp(-a / h) - b * np.exp(-b / h)) / (np.exp(-a / h) - np.exp(-b / h)) + h
    assert abs(r.mean() - mean) < 0.03
    # galaxy B is inclined 45 degrees about x: its disk normal is (0, -sin, cos)
    d = pos[501:] - pos[500]
    normal = np.linalg.svd(d - d.mean(0))[2][-1]
    assert abs(abs(normal @ np.array([0, -np.sin(np.pi / 4), np.cos(np.pi / 4)])) - 1) < 1e-2


def test_generators_deterministic_and_seeded():
    a = gcsim.generators.config1(n=4096)
    b = gcsim.generators.config1(n=4096)
    c = gcsim.generators.config1(n=4096, seed=7)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    assert not np.array_equal(a[0], c[0])
    assert abs(a[2][1:2048].sum() - 0.1) < 1e-12


def test_plummer_statistics():
    pos, vel, mass = gcsim.generators.plummer(20000, seed=1)
    r = np.linalg.norm(pos, axis=1)
    # half-mass radius of a Plummer sphere: a / sqrt(2^(2/3) - 1)
    assert abs(np.median(r) - 1 / np.sqrt(2 ** (2 / 3) - 1)) < 0.03
    ke = 0.5 * np.sum(mass[:, None] * vel ** 2)
    # virial equilibrium: KE = -PE/2 = 3 pi / 64 for G = M = a = 1
    assert abs(ke - 3 * np.pi / 64) < 0.01
    assert np.all(np.linalg.norm(vel, axis=1) < np.sqrt(2) * (1 + r * r) ** -0.25 + 1e-12)


def test_collision_sampler():
    runs = gcsim.generators.collisions(256, seed=2026)
    ns = np.array([r["n"] for r in runs])
    assert len(runs) == 256 and np.all(ns % 2 == 0)
    assert ns.min() >= 512 and ns.max() <= 8192
    assert abs(np.log(ns).mean() - 0.5 * (np.log(512) + np.log(8192))) < 0.15
    assert all(0 <= r["inclination"] <= np.pi / 2 and 0.5 <= r["impact_parameter"] <= 2.0
               and 0.2 <= r["speed"] <= 0.5 for r in runs)
