"""Host-side checks of the hip-direct building blocks (no GPU)."""
import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def test_symmetric_plan_covers_every_pair_once():
    # csrc/sym_plan.hpp: the symmetric kernel's work list over the
    # tile-pair triangle (single runs of 1 to 1,024 tiles and uneven batches,
    # default and set item lengths) covers every diagonal tile and every pair
    # I < J of a run exactly once, numbers row-partial slots in (row, J)
    # order, keeps pair-partial slots in range and launches longest first.
    exe = ROOT / "tests" / "cpp" / "sym_plan_check"
    res = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "sym_plan_check passed" in res.stdout
