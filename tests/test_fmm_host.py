"""Host-side checks of the hip-fmm building blocks (no GPU)."""
import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def test_device_m2p_matches_reference_m2p():
    # csrc/far_m2p.hpp (per-cell coefficient tables, radial-kernel
    # evaluation) against the oracle's restatement of multipole.cpp:263-296
    # for p = 0, 2, 4, targets well separated and at about the opening-angle
    # limit: the same formula in FP64, FP32 rounding only in single precision.
    exe = ROOT / "tests" / "cpp" / "m2p_check"
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True,
                         timeout=60).stdout
    rows = {int(l.split()[0][2:]): (float(l.split()[2]), float(l.split()[4]))
            for l in out.splitlines() if l.startswith("p=")}
    assert set(rows) == {0, 2, 4}
    for p, (d, f) in rows.items():
        assert d < 1e-13, (p, d)
        assert f < 5e-6, (p, f)
