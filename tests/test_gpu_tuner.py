"""ktt::Tuner on ComputeApi::HIP end to end (hipRTC compilation per configuration, event timing,
reference-kernel validation, searchers, JSON results), through ktt_coulomb3d_tune."""

from __future__ import annotations

import json

import numpy as np
import pytest

from ktt_amd import workloads
from oracle import coulomb_oracle as oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("searcher", [0, 1, 2])
def test_tune_coulomb3d_small(capi, tmp_path, searcher):
    atoms = workloads.uniform_atoms(41, 96, box=16.0)
    grid, spacing = 32, 0.5
    results = tmp_path / f"results_{searcher}"
    print(f"tuning searcher={searcher}", flush=True)
    outcome = capi.coulomb3d_tune(atoms, spacing, grid, searcher=searcher, max_configurations=4, tolerance=0.01,
                                  results_file=str(results), want_grid=True)
    assert outcome["tested"] == 4
    assert outcome["valid"] >= 1
    assert outcome["best_ms"] > 0.0
    document = json.loads((tmp_path / f"results_{searcher}.json").read_text())
    assert document["Metadata"]["ComputeApi"] == "HIP"
    assert len(document["Results"]) == 4
    statuses = {entry["Status"] for entry in document["Results"]}
    assert statuses <= {"Ok", "CompilationFailed", "DeviceLimitsExceeded", "ComputationFailed", "ValidationFailed"}
    cpu = oracle.coulomb3d_f32(atoms, spacing, grid)
    assert np.max(np.abs(outcome["grid"] - cpu)) <= 0.01
