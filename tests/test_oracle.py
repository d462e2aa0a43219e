"""Pins the CPU oracle before it is trusted as the checker.

* FP32 restatement: bit-exact against the reference's own CPU generator
  (Examples/JSON/generate-coulomb-output.c) through the committed known-answer fixture, and
  against the compiled reference program itself when it is present (oracle/_ref).
* FP64 evaluation: closed forms (single charge, symmetric pair), linearity, sign flip,
  translation by whole grid steps, numpy brute force on tiny grids.
* 2D slice restatement: equal to the 3D restatement on the slice (same operation order), and
  close to FP64.
A plausible slip (dropped dz^2 term, swapped x/y, wrong sign, missing charge factor) breaks at
least one of these.
"""

from __future__ import annotations

import hashlib
import json
import subprocess
import tempfile
from pathlib import Path

import numpy as np
import pytest

from ktt_amd import workloads
from oracle import coulomb_oracle as oracle

GOLDEN = Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def kat():
    return json.loads((GOLDEN / "coulomb3d_kat.json").read_text())


def test_fp32_restatement_reproduces_reference_generator_bit_exactly(kat):
    atoms = workloads.kat_atoms(kat["atom_count"])
    grid = oracle.coulomb3d_f32(atoms, workloads.KAT_SPACING, kat["grid"])
    assert hashlib.sha256(grid.tobytes()).hexdigest() == kat["sha256"]
    points = np.array(kat["points"])
    got = grid[points[:, 2], points[:, 1], points[:, 0]].view(np.uint32)
    assert [format(int(v), "08x") for v in got] == kat["values_hex"]


def test_reference_program_output_matches_fixture(kat):
    binary = oracle.REFERENCE_BINARY
    if not binary.exists():
        pytest.skip("oracle/_ref/generate-coulomb-output is built only where the reference sources are mounted")
    with tempfile.TemporaryDirectory() as directory:
        subprocess.run([str(binary)], cwd=directory, check=True)
        data = (Path(directory) / "Coulomb3d.bin").read_bytes()
    assert hashlib.sha256(data).hexdigest() == kat["sha256"]


def test_fixture_statistics_are_consistent(kat):
    values = np.array(kat["values"])
    assert kat["stats"]["min"] <= values.min() and values.max() <= kat["stats"]["max"]
    assert np.all(values > 0.0)   # all KAT charges are non-negative


def test_fp64_single_charge_is_q_over_r():
    atom = np.array([[1.3, 2.7, 0.9, -0.75]], dtype=np.float32)
    grid, spacing = 6, 0.5
    z, y, x = np.meshgrid(np.arange(grid), np.arange(grid), np.arange(grid), indexing="ij")
    points = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1).astype(np.int32)
    values, norms = oracle.points_f64(atom, spacing, points)
    distance = np.sqrt(((points * spacing - atom[0, :3].astype(np.float64)) ** 2).sum(axis=1))
    np.testing.assert_allclose(values, -0.75 / distance, rtol=1e-15)
    np.testing.assert_allclose(norms, 0.75 / distance, rtol=1e-15)


def test_fp64_symmetric_pair_vanishes_on_bisector_and_matches_axis_closed_form():
    # +q at (2, 2, 1.5), -q at (2, 2, 2.5): the plane z = 2 is equipotential 0.
    atoms = np.array([[2.0, 2.0, 1.5, 1.0], [2.0, 2.0, 2.5, -1.0]], dtype=np.float32)
    spacing = 0.5
    plane = np.array([[x, y, 4] for x in range(9) for y in range(9)], dtype=np.int32)
    values, norms = oracle.points_f64(atoms, spacing, plane)
    assert np.max(np.abs(values) / norms) < 1e-15
    axis = np.array([[4, 4, z] for z in range(6, 12)], dtype=np.int32)
    values, _ = oracle.points_f64(atoms, spacing, axis)
    zc = axis[:, 2] * spacing
    np.testing.assert_allclose(values, 1.0 / (zc - 1.5) - 1.0 / (zc - 2.5), rtol=1e-14)


def test_fp64_matches_numpy_brute_force():
    atoms = workloads.uniform_atoms(5, 7, box=3.0)
    shape = (3, 4, 5)
    brute = oracle.brute_force_f64(atoms, 0.5, shape)
    z, y, x = np.meshgrid(np.arange(shape[0]), np.arange(shape[1]), np.arange(shape[2]), indexing="ij")
    points = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1).astype(np.int32)
    values, _ = oracle.points_f64(atoms, 0.5, points)
    np.testing.assert_allclose(values.reshape(shape), brute, rtol=1e-13)


def test_fp64_linearity_sign_and_translation():
    atoms = workloads.uniform_atoms(6, 50, box=10.0)
    # positions on a 2^-10 lattice so that the translation below is exact in float32
    atoms[:, :3] = np.round(atoms[:, :3] * 1024.0) / 1024.0
    points = workloads.sample_points(3, 64, 20)
    base, _ = oracle.points_f64(atoms, 0.5, points)
    doubled = atoms.copy()
    doubled[:, 3] *= -2.0
    flipped, _ = oracle.points_f64(doubled, 0.5, points)
    np.testing.assert_allclose(flipped, -2.0 * base, rtol=1e-13)
    moved = atoms.copy()
    moved[:, :3] += np.float32(1.5)   # three grid steps of 0.5
    shifted, _ = oracle.points_f64(moved, 0.5, points + 3)
    np.testing.assert_allclose(shifted, base, rtol=1e-13)
    swapped = atoms[:, [1, 0, 2, 3]].copy()
    swapped_values, _ = oracle.points_f64(swapped, 0.5, points[:, [1, 0, 2]])
    np.testing.assert_allclose(swapped_values, base, rtol=1e-13)


def test_fp32_restatement_close_to_fp64():
    atoms = workloads.uniform_atoms(8, 400, box=16.0)
    grid = oracle.coulomb3d_f32(atoms, 0.5, 32)
    points = workloads.sample_points(9, 500, 32)
    values, norms = oracle.points_f64(atoms, 0.5, points)
    worst, _ = oracle.normalised_error(grid[points[:, 2], points[:, 1], points[:, 0]], values, norms)
    assert worst < 1e-6


def test_fp32_slab_decomposition_is_exact():
    atoms = workloads.uniform_atoms(10, 100, box=8.0)
    whole = oracle.coulomb3d_f32(atoms, 0.5, 16)
    pieces = np.zeros_like(whole)
    for z0 in range(0, 16, 5):
        oracle.coulomb3d_f32(atoms, 0.5, 16, z0, min(16, z0 + 5), out=pieces)
    np.testing.assert_array_equal(whole, pieces)


def test_slice_restatement_equals_volume_slice():
    atoms = workloads.uniform_atoms(12, 300, box=16.0)
    grid, spacing, z = 32, 0.5, 13
    volume = oracle.coulomb3d_f32(atoms, spacing, grid)
    slice_atoms = oracle.slice_dz2(atoms, spacing, z)
    np.testing.assert_array_equal(slice_atoms, workloads.slice_dz2(atoms, spacing * z))
    flat = oracle.coulomb2d_f32(slice_atoms, spacing, grid, grid)
    np.testing.assert_array_equal(flat, volume[z])
    points = workloads.sample_points(13, 200, grid)[:, :2]
    values, norms = oracle.points2d_f64(slice_atoms, spacing, points)
    worst, _ = oracle.normalised_error(flat[points[:, 1], points[:, 0]], values, norms)
    assert worst < 1e-6


def test_planted_mistakes_are_caught():
    """Each plausible slip in a restatement changes the answer far beyond the test bars."""
    atoms = workloads.uniform_atoms(14, 40, box=8.0)
    points = workloads.sample_points(15, 100, 16)
    values, norms = oracle.points_f64(atoms, 0.5, points)
    p = points * 0.5
    d = p[:, None, :] - atoms[None, :, :3].astype(np.float64)
    q = atoms[:, 3].astype(np.float64)
    mistakes = {
        "dropped z term": (q / np.sqrt(d[..., 0] ** 2 + d[..., 1] ** 2)).sum(axis=1),
        "x and y swapped": (q / np.sqrt((p[:, None, 1] - atoms[None, :, 0]) ** 2 + (p[:, None, 0] - atoms[None, :, 1]) ** 2
                                        + d[..., 2] ** 2)).sum(axis=1),
        "missing charge": (1.0 / np.sqrt((d ** 2).sum(axis=2))).sum(axis=1),
        "r instead of 1/r": (q * np.sqrt((d ** 2).sum(axis=2))).sum(axis=1),
    }
    for name, wrong in mistakes.items():
        assert np.max(np.abs(wrong - values) / norms) > 1e-3, name


def test_jitter_hash_matches_published_splitmix64_outputs():
    # SplitMix64 seeded with 0 returns 0xE220A8397B1DCDAF, then 0x6E789E6AA1B965F4.
    state = np.array([0, 0x9E3779B97F4A7C15], dtype=np.uint64)
    outputs = oracle._mix64(state)
    assert int(outputs[0]) == 0xE220A8397B1DCDAF
    assert int(outputs[1]) == 0x6E789E6AA1B965F4


def test_jitter_is_standard_normal_per_axis_and_independent_across_steps():
    atoms = np.zeros((200_000, 4), np.float32)
    atoms[:, 3] = 0.25
    first = oracle.jitter(atoms, seed=5, step=3, sigma=1.0)
    second = oracle.jitter(atoms, seed=5, step=4, sigma=1.0)
    np.testing.assert_array_equal(first[:, 3], atoms[:, 3])
    displacement = first[:, :3].astype(np.float64)
    # Sample moments of 2e5 draws: mean within ~5 standard errors, variance and kurtosis close.
    assert np.all(np.abs(displacement.mean(axis=0)) < 0.012)
    assert np.all(np.abs(displacement.var(axis=0) - 1.0) < 0.02)
    assert np.all(np.abs((displacement ** 4).mean(axis=0) - 3.0) < 0.1)
    assert abs(np.corrcoef(displacement[:, 0], displacement[:, 1])[0, 1]) < 0.012
    assert abs(np.corrcoef(displacement[:, 0], second[:, 0])[0, 1]) < 0.012
    # sigma scales the displacement; the same key reproduces it.
    np.testing.assert_allclose(oracle.jitter(atoms, 5, 3, 0.05)[:, :3], 0.05 * displacement, rtol=1e-6, atol=1e-9)
    np.testing.assert_array_equal(oracle.jitter(atoms, 5, 3, 1.0), first)
