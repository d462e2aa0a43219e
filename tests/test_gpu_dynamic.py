"""Dynamic tuning run: ktt TuneIteration steps interleaved with production steps while the
atoms move under a seeded Gaussian jitter on the device."""

from __future__ import annotations

import json

import numpy as np
import pytest

from ktt_amd import workloads
from oracle import coulomb_oracle as oracle

pytestmark = pytest.mark.gpu

COMPENSATED_BAR = 1e-6


def run_small(capi, tmp_path, searcher, budget, steps=10, grid=64, atoms_count=3000, seed=77, sigma=0.05):
    atoms = workloads.uniform_atoms(seed, atoms_count, box=grid * 0.5, stream=5)
    prefix = tmp_path / f"dynamic_{searcher}"
    out = capi.coulomb3d_dynamic(atoms, 0.5, grid, steps, sigma, seed, searcher=searcher, tuning_budget=budget,
                                 results_file=str(prefix), want_final=True)
    return atoms, out, prefix


def test_dynamic_run_moves_atoms_like_the_oracle_and_keeps_the_map_exact(capi, tmp_path):
    steps, budget, seed, sigma = 10, 6, 77, 0.05
    atoms, out, prefix = run_small(capi, tmp_path, capi.SEARCHER_MCMC, budget, steps=steps, seed=seed, sigma=sigma)
    print("dynamic space", out["space"], "best", out["best_config"], flush=True)
    assert out["tested"] == budget
    np.testing.assert_array_equal(out["tuned"], np.arange(steps) < budget)
    assert np.all(out["kernel_ms"][budget:] > 0.0)

    expected = atoms
    for step in range(steps):
        expected = oracle.jitter(expected, seed, step, sigma)
    # Device FP64 log/sin/cos may differ from the host's in the last bit before rounding to
    # float32; one float32 ulp at |x| < 64 is below 1e-5 and accumulates over the steps.
    np.testing.assert_allclose(out["atoms"], expected, rtol=0, atol=1e-5 * steps)
    np.testing.assert_array_equal(out["atoms"][:, 3], atoms[:, 3])

    points = workloads.sample_points(78, 600, 64)
    values, norms = oracle.points_f64(out["atoms"], 0.5, points)
    error, rms = oracle.normalised_error(out["grid"][points[:, 2], points[:, 1], points[:, 0]], values, norms)
    print("dynamic final-map error", error, rms, flush=True)
    assert error <= COMPENSATED_BAR

    results = json.loads(prefix.with_suffix(".json").read_text())
    assert results["Metadata"]["ComputeApi"] == "HIP"
    assert len(results["Results"]) == budget
    assert results["UserData"]["workload"] == "dynamic"


def test_dynamic_full_size_short_run(capi, tmp_path):
    """The dynamic workload at its real size (192^3 grid, 131,072 atoms, sigma 0.05 A) for 20
    steps, 8 of them MCMC tuning steps: the device jitter follows the oracle's restatement and
    the final map meets the compensated bar against FP64 at 4,096 points."""
    steps, budget, sigma = 20, 8, workloads.DYNAMIC_SIGMA
    work = workloads.dynamic_initial()
    out = capi.coulomb3d_dynamic(work.atoms, work.spacing, work.grid, steps, sigma, workloads.SEED_DYNAMIC,
                                 searcher=capi.SEARCHER_MCMC, tuning_budget=budget, results_file=str(tmp_path / "dynamic_full"),
                                 want_final=True)
    print("dynamic full size: best", out["best_config"], "kernel ms", np.round(out["kernel_ms"], 2).tolist(), flush=True)
    assert out["tested"] == budget
    assert np.all(out["kernel_ms"] > 0.0)
    expected = work.atoms
    for step in range(steps):
        expected = oracle.jitter(expected, workloads.SEED_DYNAMIC, step, sigma)
    np.testing.assert_allclose(out["atoms"], expected, rtol=0, atol=1e-5 * steps)
    np.testing.assert_array_equal(out["atoms"][:, 3], work.atoms[:, 3])
    points = workloads.sample_points(79, 4096, work.grid)
    values, norms = oracle.points_f64(out["atoms"], work.spacing, points)
    error, rms = oracle.normalised_error(out["grid"][points[:, 2], points[:, 1], points[:, 0]], values, norms)
    print(f"dynamic full size final-map error max {error:.2e} rms {rms:.2e}", flush=True)
    assert error <= COMPENSATED_BAR


def test_dynamic_production_phase_runs_the_fastest_tested_configuration(capi, tmp_path):
    budget = 8
    _, out, prefix = run_small(capi, tmp_path, capi.SEARCHER_DETERMINISTIC, budget, steps=12, atoms_count=2000, seed=79)
    assert out["space"] > budget
    assert out["tested"] == budget
    results = json.loads(prefix.with_suffix(".json").read_text())["Results"]
    assert len(results) == budget
    fastest = min((r for r in results if r["Status"] == "Ok"), key=lambda r: r["TotalDuration"])
    chosen = ",".join(f"{p['Name']}={p['Value']}" for p in fastest["Configuration"])
    assert chosen == out["best_config"]
    assert np.all(out["kernel_ms"][budget:] > 0.0)


def test_dynamic_run_rejects_bad_input(capi):
    atoms = workloads.uniform_atoms(1, 10, box=8.0)
    with pytest.raises(capi.KttHipError):
        capi.coulomb3d_dynamic(atoms, 0.5, 64, 0, 0.05, 1)
    with pytest.raises(capi.KttHipError):
        capi.coulomb3d_dynamic(atoms, 0.5, 64, 2, -1.0, 1)
    with pytest.raises(capi.KttHipError):
        # No span of the space wastes at most an eighth of a 40-point row.
        capi.coulomb3d_dynamic(atoms, 0.5, 40, 2, 0.05, 1)
