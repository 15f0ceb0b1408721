"""The drop-in C++ interface (host/hip_solver.{hpp,cpp}) built the way the
reference would build it: with -DGCSIM_REFERENCE_TYPES against the
reference's own headers (cpp/core/particle.hpp, simulation_params.hpp,
diagnostics.hpp, config.hpp, cpp/fmm/fmm_solver.hpp), linked to libgcsim.
On a host without a device every solver call throws std::runtime_error
(the reference's error convention, cpp/cuda/cuda_solver.cu:55-61), and the
writer takes the reference's SimulationConfig.  Needs the reference sources
(build container only)."""
import subprocess
from pathlib import Path

import pytest

import gcsim

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "galaxy-collision-simulation_amd"
REF_CPP = Path("/root/reference/cpp")


@pytest.mark.skipif(not (REF_CPP / "core" / "particle.hpp").exists(),
                    reason="reference sources not present")
def test_hip_solver_builds_and_runs_with_reference_types(tmp_path):
    exe = tmp_path / "reference_types"
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-DGCSIM_REFERENCE_TYPES",
           f"-I{REF_CPP}", f"-I{ROOT / 'include'}", f"-I{PKG / 'host'}",
           str(ROOT / "tests" / "cpp" / "reference_types_main.cpp"), str(PKG / "host" / "hip_solver.cpp"),
           "-o", str(exe), f"-L{PKG}", "-lgcsim", f"-Wl,-rpath,{PKG}"]
    res = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stderr
    out = tmp_path / "run"
    res = subprocess.run([str(exe), str(out)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    if gcsim.device_count() == 0:
        assert "available=0 threw=3" in res.stdout
    meta = (out / "metadata.json").read_text()
    assert '"solver": "hip-direct"' in meta and '"particle_count": 8' in meta
    assert (out / "snapshot_000000.csv").exists()
