"""Symbol arguments (ktt::Tuner::AddArgumentSymbol, Source/Tuner.h:501-510) on the HIP engine:
the atoms, given as one symbol argument, fill a __constant__ array of the kernel
(tests/kernels/coulomb_symbol.hip); the potential it computes is checked against the oracle.
The program also checks that a symbol needs a name and that names are unique."""

from __future__ import annotations

import subprocess
from pathlib import Path

import numpy as np
import pytest

from ktt_amd import workloads
from oracle import coulomb_oracle as oracle

pytestmark = pytest.mark.gpu

REPO = Path(__file__).resolve().parents[1]
PROGRAM = REPO / "ktt_amd" / "bin" / "ktt_symbol_test"
KERNEL = REPO / "tests" / "kernels" / "coulomb_symbol.hip"


def test_atoms_in_a_symbol_match_the_oracle(capi, tmp_path):
    atoms = workloads.uniform_atoms(91, 200, box=16.0)
    grid, spacing = 32, 0.5
    atoms_file = tmp_path / "atoms.bin"
    atoms_file.write_bytes(np.ascontiguousarray(atoms, dtype=np.float32).tobytes())
    output = tmp_path / "grid.bin"
    print("symbol test: 200 atoms, 32^3 grid", flush=True)
    result = subprocess.run([str(PROGRAM), str(KERNEL), str(atoms_file), str(grid), str(spacing), str(output)],
                            capture_output=True, text=True, timeout=120)
    print(result.stdout, result.stderr, flush=True)
    assert result.returncode == 0
    values = np.fromfile(output, dtype=np.float32).reshape(grid, grid, grid)
    z, y, x = np.meshgrid(np.arange(grid), np.arange(grid), np.arange(grid), indexing="ij")
    points = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1).astype(np.int32)
    reference, norms = oracle.points_f64(atoms, spacing, points)
    worst, _ = oracle.normalised_error(values.ravel(), reference, norms)
    assert worst <= 1e-5
    assert np.max(np.abs(values - oracle.coulomb3d_f32(atoms, spacing, grid))) <= 0.01
