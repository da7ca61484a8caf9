"""The C++ interface (include/cmag/systems.hpp) through a port of the
reference's Poiseuille report harness, run on the GPU."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from cmag_amd import scenarios as S

pytestmark = pytest.mark.gpu

EXE = Path(__file__).resolve().parent.parent / "cmag_amd" / "examples" / "poiseuille_report"


def test_cpp_report_harness_matches_series(device, tmp_path):
    assert EXE.exists(), "build the examples first (python -m cmag_amd.build)"
    proc = subprocess.run([str(EXE), str(tmp_path)], capture_output=True, text=True, timeout=240)
    assert proc.returncode == 0, proc.stderr
    lines = [l for l in proc.stdout.splitlines() if l.startswith("t=")]
    assert len(lines) == 5
    for line in lines:
        fields = dict(kv.split("=") for kv in line.split())
        t = float(fields["t"])
        elapsed = float(fields["elapsed"])
        data = np.loadtxt(tmp_path / f"XVelocityYPosition{fields['t']}.dat")
        assert data[0].tolist() == [0.0, 0.0] and data[-1].tolist() == [0.0, 0.001]
        u, y = data[1:-1, 0], data[1:-1, 1]
        assert u.size == 58
        want = S.poiseuille_series(y, elapsed, 1e-4, 1e-3, 1e-6)
        err = np.linalg.norm(u - want) / np.linalg.norm(want)
        assert err < 0.01, (t, err)
