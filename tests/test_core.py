"""Host-side tests of the ktt core through the compiled test program (no device needed):
configuration spaces of the reference's CoulombSum3d example, the results loader on the
reference's RTX 2080 Ti full-search dump, simulated tuning with every searcher, stop
conditions and the JSON round trip."""

from __future__ import annotations

import subprocess
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parents[1]
PROGRAM = REPO / "ktt_amd" / "bin" / "ktt_core_tests"
FIXTURE = REPO / "tests" / "golden" / "coulomb3d_2080_space.json"


def test_core_program_passes():
    if not PROGRAM.exists():
        pytest.fail("ktt_core_tests is not built: run python -m ktt_amd.build")
    result = subprocess.run([str(PROGRAM), str(FIXTURE)], capture_output=True, text=True, timeout=300)
    print(result.stdout, result.stderr)
    assert result.returncode == 0, result.stderr
    assert "0 failures" in result.stdout
