"""Parity of the gfx950 Coulomb kernels with the CPU oracle, through the C ABI.

Bars: the plain FP32 form within a normalised error of 1e-5 of the FP64 oracle, the
compensated form within 1e-6 (normalised by sum_i |q_i| / r_i, as the benchmark reports it),
and agreement with the reference-semantics kernels at the reference's own validation
tolerance (SideBySideComparison, absolute 0.01: Examples/CoulombSum3d/CoulombSum3d.cpp:226).
"""

from __future__ import annotations

import ctypes
import json
from pathlib import Path

import numpy as np
import pytest

from ktt_amd import workloads
from oracle import coulomb_oracle as oracle

pytestmark = pytest.mark.gpu

PLAIN_BAR = 1e-5
COMPENSATED_BAR = 1e-6
SIDE_BY_SIDE = 0.01
GOLDEN = Path(__file__).resolve().parent / "golden"


def run_volume(capi, atoms, spacing, grid, form=0, config=None, initial=None):
    atoms_buffer = capi.DeviceBuffer.from_array(atoms)
    energy = capi.DeviceBuffer(4 * grid ** 3)
    if initial is None:
        energy.zero()
    else:
        energy.upload(initial)
    capi.coulomb3d(atoms_buffer, atoms.shape[0], spacing, grid, energy, form=form, config=config)
    return energy.download((grid, grid, grid))


def all_points(grid):
    z, y, x = np.meshgrid(np.arange(grid), np.arange(grid), np.arange(grid), indexing="ij")
    return np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1).astype(np.int32)


def normalised(values_grid, atoms, spacing, points):
    reference, norms = oracle.points_f64(atoms, spacing, points)
    sampled = values_grid[points[:, 2], points[:, 1], points[:, 0]]
    return oracle.normalised_error(sampled, reference, norms)


def test_device_is_gfx950(capi):
    info = capi.device_info()
    print("device:", info, flush=True)
    assert info["arch"] == "gfx950"
    assert info["compute_units"] >= 1
    assert info["name"], "the device name falls back to hipDeviceGetName, then to the architecture"


CONFIGS_3D = [
    None,
    "WORK_GROUP_SIZE_X=32,WORK_GROUP_SIZE_Y=2,POINTS_X=1,ROWS_PER_THREAD=4,ATOM_TILE=32",
    "WORK_GROUP_SIZE_X=16,WORK_GROUP_SIZE_Y=8,POINTS_X=8,ROWS_PER_THREAD=8,ATOM_TILE=128,INNER_UNROLL_FACTOR=4",
    "WORK_GROUP_SIZE_X=64,WORK_GROUP_SIZE_Y=1,POINTS_X=2,ROWS_PER_THREAD=16,ATOM_TILE=64,ATOM_BLOCKS=3",
    # The default shape for rows that are a multiple of 128 points, here over a 40-point row.
    "WORK_GROUP_SIZE_X=128,WORK_GROUP_SIZE_Y=4,POINTS_X=1,ROWS_PER_THREAD=8,ATOM_TILE=128,INNER_UNROLL_FACTOR=8",
]


@pytest.mark.parametrize("config", CONFIGS_3D)
def test_plain_3d_matches_fp64_oracle(capi, config):
    atoms = workloads.uniform_atoms(21, 257, box=20.0)
    grid, spacing = 40, 0.5
    print(f"plain 3d config={config}", flush=True)
    result = run_volume(capi, atoms, spacing, grid, config=config)
    worst, rms = normalised(result, atoms, spacing, all_points(grid))
    assert worst <= PLAIN_BAR, (worst, rms)
    cpu = oracle.coulomb3d_f32(atoms, spacing, grid)
    assert np.max(np.abs(result - cpu)) <= SIDE_BY_SIDE


@pytest.mark.parametrize("config", [None, "ATOM_BLOCKS=4", "POINTS_X=2,ROWS_PER_THREAD=8,ATOM_TILE=32,ATOM_BLOCKS=7",
                                    # the default shape for rows that are a multiple of 256 points
                                    "WORK_GROUP_SIZE_X=128,WORK_GROUP_SIZE_Y=2,POINTS_X=2,ROWS_PER_THREAD=8,ATOM_TILE=256,"
                                    "INNER_UNROLL_FACTOR=8",
                                    "WORK_GROUP_SIZE_X=128,WORK_GROUP_SIZE_Y=2,POINTS_X=2,ROWS_PER_THREAD=8,ATOM_TILE=256,"
                                    "INNER_UNROLL_FACTOR=8,ATOM_BLOCKS=3"])
def test_compensated_3d_matches_fp64_oracle(capi, config):
    atoms = workloads.uniform_atoms(22, 1000, box=20.0)
    grid, spacing = 40, 0.5
    print(f"compensated 3d config={config}", flush=True)
    result = run_volume(capi, atoms, spacing, grid, form=capi.COMPENSATED, config=config)
    worst, rms = normalised(result, atoms, spacing, all_points(grid))
    assert worst <= COMPENSATED_BAR, (worst, rms)


PHASE_SHAPES = [
    None,
    # scalar bracket reads (ROWS_PER_THREAD not a multiple of 4), 3 points per thread, padded tiles
    "WORK_GROUP_SIZE_X=32,WORK_GROUP_SIZE_Y=2,POINTS_X=3,ROWS_PER_THREAD=2,ATOM_TILE=32,INNER_UNROLL_FACTOR=4",
]


@pytest.mark.parametrize("form", [0, 1])
@pytest.mark.parametrize("shape", PHASE_SHAPES)
def test_issue_phases_are_bit_identical(capi, form, shape):
    """ISSUE_PHASES changes the order in which a wave issues its instructions, not the arithmetic:
    every point adds its atoms in the same order with the same operations, so the interleaved
    loop (0), the phased loop (1) and the phased loop with a one-atom LDS read-ahead (2) give
    bit-identical grids (one atom block: no atomics)."""
    atoms = workloads.uniform_atoms(31, 777, box=20.0)
    grid, spacing = 40, 0.5
    grids = []
    for phases in (0, 1, 2):
        config = f"ISSUE_PHASES={phases}" if shape is None else f"{shape},ISSUE_PHASES={phases}"
        print(f"form={form} config={config}", flush=True)
        grids.append(run_volume(capi, atoms, spacing, grid, form=form, config=config))
    for other in grids[1:]:
        np.testing.assert_array_equal(other.view(np.uint32), grids[0].view(np.uint32))
    worst, rms = normalised(grids[1], atoms, spacing, all_points(grid))
    assert worst <= (COMPENSATED_BAR if form == 1 else PLAIN_BAR), (worst, rms)


@pytest.mark.parametrize("shape", [None, "WORK_GROUP_SIZE_X=32,POINTS_X=2,ROWS_PER_THREAD=2,ATOM_TILE=32"])
def test_slice_issue_phases_are_bit_identical(capi, shape):
    """The slice kernel with one atom block (plain stores, no atomics): ISSUE_PHASES 0, 1 and 2
    agree bit for bit on the 256^2 x 4,000 slice."""
    work = workloads.slice2d()
    atoms = capi.DeviceBuffer.from_array(work.slice_atoms)
    energy = capi.DeviceBuffer(4 * work.grid_x * work.grid_y)
    grids = []
    for phases in (0, 1, 2):
        config = f"ATOM_BLOCKS=1,ISSUE_PHASES={phases}" if shape is None else f"{shape},ATOM_BLOCKS=1,ISSUE_PHASES={phases}"
        print(f"slice config={config}", flush=True)
        energy.zero()
        capi.coulomb2d(atoms, work.slice_atoms.shape[0], work.spacing, work.grid_x, work.grid_y, energy, config=config)
        grids.append(energy.download((work.grid_y, work.grid_x)))
    for other in grids[1:]:
        np.testing.assert_array_equal(other.view(np.uint32), grids[0].view(np.uint32))
    cpu = oracle.coulomb2d_f32(work.slice_atoms, work.spacing, work.grid_x, work.grid_y)
    assert np.max(np.abs(grids[1] - cpu)) <= SIDE_BY_SIDE


def test_compensated_repeated_launch_reuses_counters(capi):
    """The per-tile arrival counters must be left at zero by every launch."""
    atoms = workloads.uniform_atoms(23, 700, box=16.0)
    grid, spacing = 32, 0.5
    first = run_volume(capi, atoms, spacing, grid, form=capi.COMPENSATED, config="ATOM_BLOCKS=5")
    second = run_volume(capi, atoms, spacing, grid, form=capi.COMPENSATED, config="ATOM_BLOCKS=5")
    np.testing.assert_array_equal(first, second)


def test_reference_kernel_matches_oracle(capi):
    atoms = workloads.uniform_atoms(24, 300, box=16.0)
    grid, spacing = 32, 0.5
    atoms_buffer = capi.DeviceBuffer.from_array(atoms)
    energy = capi.DeviceBuffer(4 * grid ** 3)
    energy.zero()
    capi.coulomb3d_reference(atoms_buffer, atoms.shape[0], spacing, grid, energy)
    result = energy.download((grid, grid, grid))
    worst, _ = normalised(result, atoms, spacing, all_points(grid))
    assert worst <= PLAIN_BAR


def test_slice_2d_reference_kernel_matches_oracle(capi):
    """CoulombSum2dReference (one point per thread, "=" store) on the 256^2 x 4,000 slice
    against the oracle's FP32 restatement of Examples/CoulombSum2d/CoulombSum2dReference.cl:8-20
    (SideBySide 0.01) and against FP64 at every point; the "=" store overwrites what the
    output held."""
    work = workloads.slice2d()
    atoms = capi.DeviceBuffer.from_array(work.slice_atoms)
    energy = capi.DeviceBuffer(4 * work.grid_x * work.grid_y)
    energy.upload(np.full((work.grid_y, work.grid_x), 7.0, dtype=np.float32))
    capi.coulomb2d_reference(atoms, work.slice_atoms.shape[0], work.spacing, work.grid_x, work.grid_y, energy)
    result = energy.download((work.grid_y, work.grid_x))
    cpu = oracle.coulomb2d_f32(work.slice_atoms, work.spacing, work.grid_x, work.grid_y)
    assert np.max(np.abs(result - cpu)) <= SIDE_BY_SIDE
    y, x = np.meshgrid(np.arange(work.grid_y), np.arange(work.grid_x), indexing="ij")
    points = np.stack([x.ravel(), y.ravel()], axis=1).astype(np.int32)
    values, norms = oracle.points2d_f64(work.slice_atoms, work.spacing, points)
    worst, rms = oracle.normalised_error(result[points[:, 1], points[:, 0]], values, norms)
    print(f"2d reference kernel: error max {worst:.2e} rms {rms:.2e}", flush=True)
    assert worst <= PLAIN_BAR


def test_iterative_reference_kernel_matches_oracle(capi):
    """CoulombSum3dIterativeReference (one point per thread over the whole volume with the
    atoms' real z, "+=" store) against the FP64 oracle at every point and the FP32 restatement
    of the reference generator (SideBySide 0.01), added onto a non-zero grid."""
    atoms = workloads.uniform_atoms(34, 300, box=16.0)
    grid, spacing = 32, 0.5
    atoms_buffer = capi.DeviceBuffer.from_array(atoms)
    energy = capi.DeviceBuffer(4 * grid ** 3)
    start = np.full((grid, grid, grid), -1.25, dtype=np.float32)
    energy.upload(start)
    capi.coulomb3d_iterative_reference(atoms_buffer, atoms.shape[0], spacing, grid, energy)
    result = energy.download((grid, grid, grid))
    cpu = oracle.coulomb3d_f32(atoms, spacing, grid, out=start.copy())
    assert np.max(np.abs(result - cpu)) <= SIDE_BY_SIDE
    worst, rms = normalised(result - start, atoms, spacing, all_points(grid))
    print(f"iterative reference kernel: error max {worst:.2e} rms {rms:.2e}", flush=True)
    assert worst <= PLAIN_BAR
    with pytest.raises(capi.KttHipError, match="multiple of 16"):
        capi.coulomb3d_iterative_reference(atoms_buffer, atoms.shape[0], spacing, 40, energy)


@pytest.mark.parametrize("config", [None, "ATOM_BLOCKS=1", "WORK_GROUP_SIZE_X=32,POINTS_X=2,ROWS_PER_THREAD=8,ATOM_TILE=32,ATOM_BLOCKS=5",
                                    "WORK_GROUP_SIZE_X=16,WORK_GROUP_SIZE_Y=8,POINTS_X=4,ROWS_PER_THREAD=2,ATOM_TILE=64,"
                                    "INNER_UNROLL_FACTOR=4,ATOM_BLOCKS=3"])
def test_fused_slice_kernel_matches_oracle(capi, config):
    """directCoulombSumSlice (dz^2 formed on the device, one launch per slice) against the FP64
    oracle at every point of a 64^3 grid, with one and several atom blocks (FP32 atomics)."""
    atoms = workloads.uniform_atoms(35, 1100, box=32.0)
    grid, spacing = 64, 0.5
    atoms_buffer = capi.DeviceBuffer.from_array(atoms)
    energy = capi.DeviceBuffer(4 * grid ** 3)
    energy.zero()
    print(f"fused slices config={config}", flush=True)
    capi.coulomb3d_slices(atoms_buffer, atoms.shape[0], spacing, grid, energy, config=config)
    result = energy.download((grid, grid, grid))
    worst, rms = normalised(result, atoms, spacing, all_points(grid))
    print(f"  error max {worst:.2e} rms {rms:.2e}", flush=True)
    assert worst <= PLAIN_BAR
    cpu = oracle.coulomb3d_f32(atoms, spacing, grid)
    assert np.max(np.abs(result - cpu)) <= SIDE_BY_SIDE


def test_accumulates_into_grid_and_empty_atoms_leave_it(capi):
    atoms = workloads.uniform_atoms(25, 100, box=8.0)
    grid, spacing = 16, 0.5
    once = run_volume(capi, atoms, spacing, grid)
    twice = run_volume(capi, atoms, spacing, grid, initial=once)
    np.testing.assert_allclose(twice, 2.0 * once, rtol=1e-6, atol=1e-6)
    start = np.full((grid, grid, grid), 3.5, dtype=np.float32)
    empty = run_volume(capi, np.zeros((0, 4), dtype=np.float32), spacing, grid, initial=start)
    np.testing.assert_array_equal(empty, start)


@pytest.mark.parametrize("form,config", [(0, None), (0, "ATOM_BLOCKS=3"), (1, None), (1, "ATOM_BLOCKS=3"),
                                         (1, "POINTS_X=1,ROWS_PER_THREAD=4,ATOM_TILE=32,ATOM_BLOCKS=2")])
def test_atom_on_a_grid_point(capi, form, config):
    """An atom exactly on a grid point gives r = 0 there: q / 0 is +inf or -inf, and 0 / 0 is NaN,
    in the reference generator's arithmetic (Examples/JSON/generate-coulomb-output.c:28) and in
    both forms. The compensated form's two-sum turns an infinite term into a NaN error term,
    which must not reach the grid; every other point stays finite and within the usual bars."""
    grid, spacing = 24, 0.5
    atoms = workloads.uniform_atoms(29, 150, box=12.0)
    on_points = {(5, 7, 9): 0.5, (17, 3, 20): -0.25, (11, 11, 2): 0.0}
    extra = np.array([[x * spacing, y * spacing, z * spacing, q] for (x, y, z), q in on_points.items()], dtype=np.float32)
    atoms = np.concatenate([atoms[:70], extra, atoms[70:]])
    print(f"atom on a grid point form={form} config={config}", flush=True)
    result = run_volume(capi, atoms, spacing, grid, form=form, config=config)
    cpu = oracle.coulomb3d_f32(atoms, spacing, grid)
    special = np.zeros(result.shape, dtype=bool)
    for (x, y, z), q in on_points.items():
        special[z, y, x] = True
        if q > 0:
            assert np.isposinf(result[z, y, x]) and np.isposinf(cpu[z, y, x])
        elif q < 0:
            assert np.isneginf(result[z, y, x]) and np.isneginf(cpu[z, y, x])
        else:
            assert np.isnan(result[z, y, x]) and np.isnan(cpu[z, y, x])
    assert np.isfinite(result[~special]).all()
    assert np.max(np.abs(result[~special] - cpu[~special])) <= SIDE_BY_SIDE
    points = all_points(grid)
    points = points[~special[points[:, 2], points[:, 1], points[:, 0]]]
    worst, rms = normalised(result, atoms, spacing, points)
    assert worst <= (COMPENSATED_BAR if form == 1 else PLAIN_BAR), (worst, rms)


def test_known_answer_grid_of_reference_generator(capi):
    """Examples/JSON/generate-coulomb-output.c grid (64 atoms, 256^3, h = 0.49) at its sampled points."""
    fixture = json.loads((GOLDEN / "coulomb3d_kat.json").read_text())
    atoms = workloads.kat_atoms(fixture["atom_count"])
    grid = fixture["grid"]
    result = run_volume(capi, atoms, workloads.KAT_SPACING, grid)
    points = np.array(fixture["points"], dtype=np.int32)
    expected = np.array(fixture["values"], dtype=np.float32)
    got = result[points[:, 2], points[:, 1], points[:, 0]]
    assert np.max(np.abs(got - expected)) <= SIDE_BY_SIDE
    _, norms = oracle.points_f64(atoms, workloads.KAT_SPACING, points)
    assert np.max(np.abs(got.astype(np.float64) - expected) / norms) <= PLAIN_BAR
    compensated = run_volume(capi, atoms, workloads.KAT_SPACING, grid, form=capi.COMPENSATED)
    values, norms = oracle.points_f64(atoms, workloads.KAT_SPACING, points)
    sampled = compensated[points[:, 2], points[:, 1], points[:, 0]]
    assert oracle.normalised_error(sampled, values, norms)[0] <= COMPENSATED_BAR


def test_slice_2d_full_size(capi):
    work = workloads.slice2d()
    atoms = capi.DeviceBuffer.from_array(work.slice_atoms)
    energy = capi.DeviceBuffer(4 * work.grid_x * work.grid_y)
    energy.zero()
    capi.coulomb2d(atoms, work.slice_atoms.shape[0], work.spacing, work.grid_x, work.grid_y, energy)
    result = energy.download((work.grid_y, work.grid_x))
    cpu = oracle.coulomb2d_f32(work.slice_atoms, work.spacing, work.grid_x, work.grid_y)
    assert np.max(np.abs(result - cpu)) <= SIDE_BY_SIDE
    points = workloads.sample_points(31, 2048, work.grid_x)[:, :2]
    values, norms = oracle.points2d_f64(work.slice_atoms, work.spacing, points)
    assert oracle.normalised_error(result[points[:, 1], points[:, 0]], values, norms)[0] <= PLAIN_BAR
    reference = capi.DeviceBuffer(4 * work.grid_x * work.grid_y)
    capi.coulomb2d_reference(atoms, work.slice_atoms.shape[0], work.spacing, work.grid_x, work.grid_y, reference)
    assert np.max(np.abs(reference.download((work.grid_y, work.grid_x)) - result)) <= SIDE_BY_SIDE


@pytest.mark.parametrize("config", [None, "WORK_GROUP_SIZE_X=32,POINTS_X=1,ROWS_PER_THREAD=4,ATOM_BLOCKS=1"])
def test_iterative_and_fused_slices_match_volume(capi, config):
    atoms = workloads.uniform_atoms(26, 333, box=64.0)
    grid, spacing = 128, 0.5
    volume = run_volume(capi, atoms, spacing, grid)
    energy = capi.DeviceBuffer(4 * grid ** 3)
    energy.zero()
    kernel_ms, wall_ms = capi.coulomb3d_iterative(atoms, spacing, grid, energy, config=config)
    iterative = energy.download((grid, grid, grid))
    assert kernel_ms > 0.0 and wall_ms >= kernel_ms
    assert np.max(np.abs(iterative - volume)) <= SIDE_BY_SIDE
    worst, _ = normalised(iterative, atoms, spacing, workloads.sample_points(32, 1024, grid))
    assert worst <= PLAIN_BAR
    atoms_buffer = capi.DeviceBuffer.from_array(atoms)
    energy.zero()
    capi.coulomb3d_slices(atoms_buffer, atoms.shape[0], spacing, grid, energy, config=config)
    fused = energy.download((grid, grid, grid))
    assert np.max(np.abs(fused - volume)) <= SIDE_BY_SIDE


def test_full_size_iterative_forms(capi):
    """CoulombSum3dIterative at 256^3 x 65,536 atoms, both launchers with their default slice
    shape (256 atom blocks per slice adding with FP32 atomics): the reference's host dz^2
    scheme (Examples/CoulombSum3dIterative/CoulombSum3dIterative.cpp:146-184) and the fused
    device dz^2, each against FP64 at 4,096 sampled points and against the volume kernel
    (SideBySide 0.01) everywhere."""
    work = workloads.volume3d()
    points = workloads.sample_points(36, 4096, work.grid)
    values, norms = oracle.points_f64(work.atoms, work.spacing, points)
    atoms_buffer = capi.DeviceBuffer.from_array(work.atoms)
    energy = capi.DeviceBuffer(4 * work.grid ** 3)
    energy.zero()
    capi.coulomb3d(atoms_buffer, work.atoms.shape[0], work.spacing, work.grid, energy)
    volume = energy.download((work.grid, work.grid, work.grid))
    energy.zero()
    kernel_ms, wall_ms = capi.coulomb3d_iterative(work.atoms, work.spacing, work.grid, energy)
    host = energy.download((work.grid, work.grid, work.grid))
    energy.zero()
    slices_ms = capi.coulomb3d_slices(atoms_buffer, work.atoms.shape[0], work.spacing, work.grid, energy)
    fused = energy.download((work.grid, work.grid, work.grid))
    for name, result, ms in (("host dz2", host, kernel_ms), ("device dz2", fused, slices_ms)):
        worst, rms = oracle.normalised_error(result[points[:, 2], points[:, 1], points[:, 0]], values, norms)
        difference = float(np.max(np.abs(result - volume)))
        print(f"iterative {name}: {ms:.1f} ms, error max {worst:.2e} rms {rms:.2e}, |iterative - volume| {difference:.2e}",
              flush=True)
        assert worst <= PLAIN_BAR
        assert difference <= SIDE_BY_SIDE
    assert wall_ms >= kernel_ms > 0.0


def test_full_size_volume_properties(capi):
    """256^3 x 65,536 atoms: FP64 error at 4,096 sampled points, linearity and sign flip."""
    work = workloads.volume3d()
    points = workloads.sample_points(33, 4096, work.grid)
    atoms_buffer = capi.DeviceBuffer.from_array(work.atoms)
    energy = capi.DeviceBuffer(4 * work.grid ** 3)
    for form, bar in ((capi.PLAIN, PLAIN_BAR), (capi.COMPENSATED, COMPENSATED_BAR)):
        energy.zero()
        ms = capi.coulomb3d(atoms_buffer, work.atoms.shape[0], work.spacing, work.grid, energy, form=form)
        result = energy.download((work.grid, work.grid, work.grid))
        worst, rms = normalised(result, work.atoms, work.spacing, points)
        print(f"form {form}: {ms:.2f} ms, {work.interactions / ms / 1e9:.3e} interactions/s, error max {worst:.2e} rms {rms:.2e}",
              flush=True)
        assert worst <= bar
    flipped = work.atoms.copy()
    flipped[:, 3] *= -2.0
    energy.zero()
    capi.coulomb3d(capi.DeviceBuffer.from_array(flipped), flipped.shape[0], work.spacing, work.grid, energy)
    scaled = energy.download((work.grid, work.grid, work.grid))
    energy.zero()
    capi.coulomb3d(atoms_buffer, work.atoms.shape[0], work.spacing, work.grid, energy)
    base = energy.download((work.grid, work.grid, work.grid))
    np.testing.assert_array_equal(scaled, np.float32(-2.0) * base)


@pytest.mark.parametrize("pairs,grid,centres", [(32768, 96, 16), (524288, 256, 64)])
def test_clustered_cancellation_workload(capi, pairs, grid, centres):
    """+-1 pairs 1 A apart in Gaussian clusters: the compensated form stays within 1e-6 of FP64
    where the potential cancels strongly; the plain form is reported beside it."""
    work = workloads.clustered(pair_count=pairs, centres=centres, grid=grid)
    points = workloads.sample_points(3, 4096 if grid == 256 else 1024, work.grid)
    atoms_buffer = capi.DeviceBuffer.from_array(work.atoms)
    energy = capi.DeviceBuffer(4 * work.grid ** 3)
    errors = {}
    for form in (capi.PLAIN, capi.COMPENSATED):
        energy.zero()
        print(f"clustered {pairs} pairs grid {grid} form {form}", flush=True)
        ms = capi.coulomb3d(atoms_buffer, work.atoms.shape[0], work.spacing, work.grid, energy, form=form)
        result = energy.download((work.grid, work.grid, work.grid))
        errors[form] = normalised(result, work.atoms, work.spacing, points)
        print(f"  {ms:.1f} ms, error max {errors[form][0]:.2e} rms {errors[form][1]:.2e}", flush=True)
    assert errors[capi.COMPENSATED][0] <= COMPENSATED_BAR
    assert errors[capi.PLAIN][0] <= PLAIN_BAR
    assert errors[capi.COMPENSATED][1] <= errors[capi.PLAIN][1]


@pytest.mark.parametrize("form", [0, 1])
def test_single_point_grid_and_single_atom(capi, form):
    """G = 1 (one grid point, at the origin) and N = 1 (V = q / r at every point), both forms."""
    bar = PLAIN_BAR if form == 0 else COMPENSATED_BAR
    atoms = workloads.uniform_atoms(27, 333, box=12.0)
    one = run_volume(capi, atoms, 0.5, 1, form=form)
    worst, _ = normalised(one, atoms, 0.5, np.zeros((1, 3), dtype=np.int32))
    assert worst <= bar
    charge = np.array([[3.3, 4.1, 5.7, 0.8]], dtype=np.float32)
    grid = 24
    result = run_volume(capi, charge, 0.5, grid, form=form)
    points = all_points(grid)
    d = points.astype(np.float64) * 0.5 - charge[0, :3].astype(np.float64)
    exact = float(charge[0, 3]) / np.sqrt((d * d).sum(axis=1))
    # One term: the FP32 displacement, r^2 and v_rsq_f32 are each within a few ulp.
    np.testing.assert_allclose(result[points[:, 2], points[:, 1], points[:, 0]], exact, rtol=2e-6)


@pytest.mark.parametrize("count", [1, 3, 127, 129, 255, 257])
def test_atom_counts_around_the_tile(capi, count):
    """Atom counts below, at and just past the default atom tiles (128 plain, 256 compensated):
    the last tile is padded with zero charges. The reference's vector kernels drop such a tail
    (Examples/CoulombSum3d/CoulombSum3d.cu:21); these kernels must not."""
    atoms = workloads.uniform_atoms(280 + count, count, box=10.0)
    points = all_points(20)
    for form, bar in ((0, PLAIN_BAR), (1, COMPENSATED_BAR)):
        print(f"count {count} form {form}", flush=True)
        result = run_volume(capi, atoms, 0.5, 20, form=form)
        worst, _ = normalised(result, atoms, 0.5, points)
        assert worst <= bar, (count, form, worst)


def test_translation_and_axis_swap(capi):
    """Moving every atom by whole grid steps along x moves the map with it, and swapping the
    atoms' x and y swaps the map's x and y axes, to the error bar of each form (per point,
    normalised by sum_i |q_i| / r_i)."""
    grid, spacing, shift = 32, 0.5, 6
    atoms = workloads.uniform_atoms(29, 400, box=12.0)
    # Coordinates on multiples of 2^-8 A: the shift and every displacement are exact in FP32, so
    # the comparison sees the kernels' rounding only, not a rounding of the moved inputs.
    atoms[:, :3] = np.round(atoms[:, :3] * 256.0) / 256.0
    on_grid = np.all(np.mod(atoms[:, :3], spacing) == 0.0, axis=1)
    atoms[on_grid, 2] += np.float32(1.0 / 256.0)
    _, norms = oracle.points_f64(atoms, spacing, all_points(grid))
    norms = norms.reshape(grid, grid, grid)
    moved = atoms.copy()
    moved[:, 0] += np.float32(shift * spacing)
    swapped = np.ascontiguousarray(atoms[:, [1, 0, 2, 3]])
    for form, bar in ((0, PLAIN_BAR), (1, COMPENSATED_BAR)):
        base = run_volume(capi, atoms, spacing, grid, form=form)
        shifted = run_volume(capi, moved, spacing, grid, form=form)
        difference = np.abs(shifted[:, :, shift:].astype(np.float64) - base[:, :, :-shift]) / norms[:, :, :-shift]
        assert difference.max() <= 2 * bar, (form, difference.max())
        transposed = run_volume(capi, swapped, spacing, grid, form=form)
        difference = np.abs(transposed.astype(np.float64) - base.transpose(0, 2, 1)) / norms.transpose(0, 2, 1)
        assert difference.max() <= 2 * bar, (form, difference.max())


def test_compensated_tail_split(capi):
    """The compensated default splits only the tiles of the last, partial round of workgroups over
    atom blocks; those meet in FP64 partials, the others store directly. The grid side is the first
    of 160..256 whose tile count leaves this device (at the kernel's occupancy) a partial round to
    split, as the launch plan reports. Points of the split tiles and sampled points everywhere meet
    the bar, a second launch is bitwise equal (the merge order is fixed), and TAIL_SPLIT=0 (no
    split) agrees on the split tiles to twice the bar."""
    atoms = workloads.uniform_atoms(30, 600, box=80.0)
    spacing = 0.5
    for grid in (160, 176, 192, 208, 224, 240, 256):
        plan = capi.coulomb3d_launch_plan(atoms.shape[0], grid, form=capi.COMPENSATED)
        if plan["grid"][2] > 1:
            break
    print(f"tail split plan at {grid}^3:", plan, flush=True)
    # The split path must run: more than one atom block, and some (not all) tiles split.
    assert plan["grid"][2] > 1 and 0 < plan["split_from"] < plan["groups"] and plan["scratch_bytes"] > 0
    assert capi.coulomb3d_launch_plan(atoms.shape[0], grid, form=capi.COMPENSATED, config="TAIL_SPLIT=0")["grid"][2] == 1
    first = run_volume(capi, atoms, spacing, grid, form=capi.COMPENSATED)
    second = run_volume(capi, atoms, spacing, grid, form=capi.COMPENSATED)
    np.testing.assert_array_equal(first, second)
    # Rows (z * G + y) of the split tiles: from the first split tile's row group to the end.
    rows_per_group = (grid * grid) // plan["grid"][1]
    first_row = (plan["split_from"] // plan["grid"][0]) * rows_per_group
    rows = first_row + (workloads.splitmix64(32, 4096, stream=7) % np.uint64(grid * grid - first_row)).astype(np.int64)
    xs = (workloads.splitmix64(33, 4096, stream=7) % np.uint64(grid)).astype(np.int64)
    split_points = np.stack([xs, rows % grid, rows // grid], axis=1).astype(np.int32)
    for points in (split_points, workloads.sample_points(31, 4096, grid)):
        worst, _ = normalised(first, atoms, spacing, points)
        assert worst <= COMPENSATED_BAR, worst
    whole = run_volume(capi, atoms, spacing, grid, form=capi.COMPENSATED, config="TAIL_SPLIT=0")
    _, norms = oracle.points_f64(atoms, spacing, split_points)
    index = (split_points[:, 2], split_points[:, 1], split_points[:, 0])
    assert (np.abs(first[index].astype(np.float64) - whole[index]) / norms).max() <= 2 * COMPENSATED_BAR


def read_points(capi, energy, grid, points):
    """Values of a device grid at points (M, 3) of x, y, z, read one float at a time (the grid may
    be larger than host memory should hold twice)."""
    library = capi.library()
    out = np.empty(points.shape[0], dtype=np.float32)
    base = energy.ptr.value
    for i, (x, y, z) in enumerate(points.astype(np.int64)):
        offset = 4 * ((int(z) * grid + int(y)) * grid + int(x))
        status = library.ktt_hip_memcpy_dtoh(out[i:].ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(base + offset), 4)
        assert status == 0
    return out


def all_positive(capi, energy, grid, chunk=1 << 26):
    """Whether every point of a device grid holds a positive value, read in chunks of floats."""
    library = capi.library()
    total = grid ** 3
    buffer = np.empty(chunk, dtype=np.float32)
    for start in range(0, total, chunk):
        count = min(chunk, total - start)
        status = library.ktt_hip_memcpy_dtoh(buffer.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(energy.ptr.value + 4 * start),
                                             4 * count)
        assert status == 0
        if not (buffer[:count] > 0).all():
            return False
    return True


@pytest.mark.parametrize("form,grids,config", [
    # the compensated defaults (64 x 2 threads, 2 x 8 points): 16 rows per group, 75,625 row groups
    # at 1,100, fold 2; the first side from 1,100 whose last round of workgroups the tail split
    # divides over atom blocks on this device, if one is found below 1,300
    (1, range(1100, 1300, 4), None),
    # plain, 4 rows per group: 90,000 row groups at 600, fold 2
    (0, [600], "WORK_GROUP_SIZE_X=64,WORK_GROUP_SIZE_Y=1,POINTS_X=4,ROWS_PER_THREAD=4,ATOM_TILE=64"),
])
def test_large_grid_folds_row_groups(capi, form, grids, config):
    """A grid with more than 65,535 row groups folds them into the launch's x dimension; the
    points of the last row groups (which only a folded launch reaches) and sampled points
    everywhere meet the bar, and every point is written once (all positive for positive charges;
    a point computed twice would be off by its own value). Above 1,024^3 points the point index
    also passes 2^30 floats (2^32 bytes)."""
    atoms = workloads.uniform_atoms(35, 600, box=grids[0] * 0.5)
    atoms[:, 3] = np.abs(atoms[:, 3]) + np.float32(0.01)
    spacing = 0.5
    for grid in grids:
        plan = capi.coulomb3d_launch_plan(atoms.shape[0], grid, form=form, config=config)
        if plan["grid"][2] > 1:
            break
    else:
        grid = grids[0]
        plan = capi.coulomb3d_launch_plan(atoms.shape[0], grid, form=form, config=config)
    print(f"large grid form={form} grid={grid} plan={plan}", flush=True)
    span = plan["block"][0] * (4 if config else 2)
    assert plan["grid"][1] <= 65535 and plan["grid"][0] >= 2 * -(-grid // span)
    atoms_buffer = capi.DeviceBuffer.from_array(atoms)
    energy = capi.DeviceBuffer(4 * grid ** 3)
    try:
        energy.zero()
        capi.coulomb3d(atoms_buffer, atoms.shape[0], spacing, grid, energy, form=form, config=config)
        capi.synchronize()
        last_rows = np.array([[x, grid - 1 - k, grid - 1] for k in range(3) for x in range(grid)], dtype=np.int32)
        for points in (last_rows, workloads.sample_points(36, 2048, grid)):
            reference, norms = oracle.points_f64(atoms, spacing, points)
            values = read_points(capi, energy, grid, points)
            worst, rms = oracle.normalised_error(values, reference, norms)
            assert worst <= (COMPENSATED_BAR if form == 1 else PLAIN_BAR), (worst, rms)
        assert all_positive(capi, energy, grid)
    finally:
        energy.close()
        atoms_buffer.close()
