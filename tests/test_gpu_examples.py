"""The example programs: ktt::Tuner on ComputeApi::HIP tuning the gfx950 kernels, every
configuration validated against the reference kernel (SideBySideComparison at 0.01 for the
Coulomb examples, 0.001 for N-body, as the reference's examples use)."""

from __future__ import annotations

import json
import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

BIN = Path(__file__).resolve().parents[1] / "ktt_amd" / "bin"


def run_example(name, arguments, tmp_path, timeout=240):
    program = BIN / name
    if not program.exists():
        pytest.fail(f"{program} is not built: run python -m ktt_amd.build")
    output = tmp_path / name
    command = [str(program), *map(str, arguments), str(output)]
    print("running", " ".join(command), flush=True)
    result = subprocess.run(command, capture_output=True, text=True, timeout=timeout)
    print(result.stdout[-2000:], result.stderr[-2000:])
    assert result.returncode == 0, result.stderr[-2000:]
    return json.loads(output.with_suffix(".json").read_text()), result.stdout


@pytest.mark.parametrize("name,arguments,count", [
    ("CoulombSum3d", [0, "random", 5, 64, 700], 5),
    ("CoulombSum2d", [0, "mcmc", 6, 128, 900], 6),
    ("CoulombSum3dIterative", [0, "random", 4, 32, 300], 4),
    ("Nbody", ["random", 6, 8192], 6),
])
def test_example_tunes_and_validates(name, arguments, count, tmp_path):
    results, stdout = run_example(name, arguments, tmp_path)
    assert results["Metadata"]["ComputeApi"] == "HIP"
    assert len(results["Results"]) == count
    statuses = [r["Status"] for r in results["Results"]]
    # Every configuration of these spaces compiles and passes the reference validation.
    assert statuses == ["Ok"] * count, statuses
    assert "best" in stdout
