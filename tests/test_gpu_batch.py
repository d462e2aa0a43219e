"""Batched form: many independent systems in one persistent, graph-captured launch."""

from __future__ import annotations

import json

import numpy as np
import pytest

from ktt_amd import workloads
from oracle import coulomb_oracle as oracle

pytestmark = pytest.mark.gpu

COMPENSATED_BAR = 1e-6
PLAIN_BAR = 1e-5


def small_batch(count=10, seed=51):
    sides = (32, 48, 64)
    systems = []
    for k in range(count):
        side = sides[k % len(sides)]
        atoms = workloads.uniform_atoms(seed * 100 + k, 40 + 37 * k, box=side * 0.5, stream=3)
        systems.append(workloads.BatchSystem(atoms, side, 0.5))
    return systems


def check_against_oracle(systems, grids, bar, samples=300):
    worst = 0.0
    for k, (system, grid) in enumerate(zip(systems, grids)):
        points = workloads.sample_points(60 + k, samples, system.grid)
        values, norms = oracle.points_f64(system.atoms, system.spacing, points)
        error, _ = oracle.normalised_error(grid[points[:, 2], points[:, 1], points[:, 0]], values, norms)
        worst = max(worst, error)
    assert worst <= bar, worst
    return worst


UNROLLED = ("BUCKET0_UNROLL=8,BUCKET1_UNROLL=8,BUCKET2_UNROLL=8,BUCKET1_TILE=32,BUCKET2_WGX=64,BUCKET2_PX=1,BUCKET2_ROWS=4,"
            "BUCKET2_TILE=128")


@pytest.mark.parametrize("config", [None, UNROLLED])
def test_batch_matches_oracle_and_graph_replays_match_eager(capi, config):
    systems = small_batch()
    print("batch config", config, flush=True)
    batch = capi.Batch(systems, config=config)
    info = batch.info()
    assert info["interactions"] == sum(s.interactions for s in systems)
    batch.zero()
    batch.run(use_graph=False)
    eager = [g.copy() for g in batch.grids()]
    print("batch items", info["items"], "error", check_against_oracle(systems, eager, COMPENSATED_BAR), flush=True)
    for replay in range(3):
        batch.zero()
        capi.synchronize()
        batch.run(use_graph=True)
        capi.synchronize()
        for k, grid in enumerate(batch.grids()):
            np.testing.assert_array_equal(grid, eager[k], err_msg=f"replay {replay} system {k}")


def test_batch_accumulates_across_runs(capi):
    systems = small_batch(count=4, seed=52)
    batch = capi.Batch(systems)
    batch.zero()
    batch.run(use_graph=True)
    once = [g.copy() for g in batch.grids()]
    batch.run(use_graph=True)
    for first, second in zip(once, batch.grids()):
        np.testing.assert_allclose(second, 2.0 * first, rtol=1e-6, atol=1e-7)


def test_per_system_baseline_matches_oracle(capi):
    systems = small_batch(count=6, seed=53)
    batch = capi.Batch(systems)
    batch.zero()
    batch.run_per_system(form=capi.PLAIN)
    check_against_oracle(systems, batch.grids(), PLAIN_BAR)
    batch.zero()
    batch.run_per_system(form=capi.COMPENSATED)
    check_against_oracle(systems, batch.grids(), COMPENSATED_BAR)


def test_bucket_tuning_keeps_results_correct(capi, tmp_path):
    systems = small_batch(count=6, seed=54)
    batch = capi.Batch(systems)
    prefix = tmp_path / "batched"
    best = batch.tune(searcher=capi.SEARCHER_RANDOM, max_configurations=3, results_file=str(prefix))
    assert "BUCKET0_WGX" in best
    winners = json.loads((tmp_path / "batched_winners.json").read_text())
    assert len(winners["buckets"]) == 3
    for bucket in range(3):
        results = json.loads((tmp_path / f"batched_bucket{bucket}.json").read_text())
        assert results["Metadata"]["ComputeApi"] == "HIP"
        assert 1 <= len(results["Results"]) <= 3
    batch.zero()
    batch.run(use_graph=True)
    check_against_oracle(systems, batch.grids(), COMPENSATED_BAR)


def test_full_batch_every_system_against_oracle(capi):
    """The 1,024-system batch of the benchmark (atom counts log-uniform in [200, 20000], sides
    32/48/64/96/128, so every bucket's tile shape runs) through the graph-launched persistent
    kernel: every system at 64 sampled points against FP64, compensated bar; a graph replay
    equals the first run bit for bit."""
    systems = workloads.batch()
    assert sorted({s.grid for s in systems}) == list(workloads.BATCH_SIDES)
    batch = capi.Batch(systems)
    info = batch.info()
    print("full batch:", len(systems), "systems,", info["items"], "items,", info["shapes"], flush=True)
    assert info["interactions"] == sum(s.interactions for s in systems)
    batch.zero()
    ms = batch.run(use_graph=True)
    first = [g.copy() for g in batch.grids()]
    worst_by_side = {}
    rms_all = []
    for k, (system, grid) in enumerate(zip(systems, first)):
        points = workloads.sample_points(4000 + k, 64, system.grid)
        values, norms = oracle.points_f64(system.atoms, system.spacing, points)
        worst, rms = oracle.normalised_error(grid[points[:, 2], points[:, 1], points[:, 0]], values, norms)
        worst_by_side[system.grid] = max(worst_by_side.get(system.grid, 0.0), worst)
        rms_all.append(rms)
    print(f"full batch {ms:.1f} ms; worst normalised error by side {worst_by_side}; mean rms {np.mean(rms_all):.2e}", flush=True)
    assert max(worst_by_side.values()) <= COMPENSATED_BAR
    batch.zero()
    batch.run(use_graph=True)
    for k, grid in enumerate(batch.grids()):
        np.testing.assert_array_equal(grid, first[k], err_msg=f"replay, system {k}")
    batch.close()


def test_batch_rejects_bad_input(capi):
    with pytest.raises(capi.KttHipError):
        capi.Batch([workloads.BatchSystem(np.zeros((3, 4), np.float32), side, 0.5) for side in (8, 16, 24, 32, 40, 56)])
    with pytest.raises(capi.KttHipError):
        capi.Batch([workloads.BatchSystem(np.zeros((3, 4), np.float32), 0, 0.5)])


def test_batch_system_equals_single_system_run(capi):
    """Each system of a batch against the compensated single-system kernel on its own atoms and
    grid: the two summation orders agree to twice the compensated bar at every point."""
    systems = small_batch(count=3, seed=53)
    batch = capi.Batch(systems)
    batch.zero()
    batch.run(use_graph=False)
    grids = batch.grids()
    for k, system in enumerate(systems):
        atoms = capi.DeviceBuffer.from_array(system.atoms)
        energy = capi.DeviceBuffer(4 * system.grid ** 3)
        energy.zero()
        capi.coulomb3d(atoms, system.atoms.shape[0], system.spacing, system.grid, energy, form=capi.COMPENSATED)
        single = energy.download((system.grid, system.grid, system.grid))
        z, y, x = np.meshgrid(*(np.arange(system.grid),) * 3, indexing="ij")
        points = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1).astype(np.int32)
        _, norms = oracle.points_f64(system.atoms, system.spacing, points)
        difference = np.abs(grids[k].astype(np.float64) - single).ravel() / norms
        print(f"system {k}: side {system.grid}, {system.atoms.shape[0]} atoms, max {difference.max():.2e}", flush=True)
        assert difference.max() <= 2 * COMPENSATED_BAR
