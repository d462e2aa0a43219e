"""The entry point that allocates nothing (ktt_coulomb3d_compensated_noalloc) against the entry
point it is a variant of (ktt_coulomb3d, compensated form, default configuration).

Two calls in a row share the library's static scratch: the first at 256^3 x 65,536 atoms (the
tail split's partials fill about 40 MiB of it), the second smaller (160^3 x 600 atoms, a
different split). Each must agree with ktt_coulomb3d to the compensated form's bar (normalised
by sum_i |q_i| / r_i at every sampled point) and meet that bar against the FP64 oracle.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

from ktt_amd import workloads
from oracle import coulomb_oracle as oracle

pytestmark = pytest.mark.gpu

COMPENSATED_BAR = 1e-6


def run_both(capi, atoms, spacing, grid):
    atoms_buffer = capi.DeviceBuffer.from_array(atoms)
    energy = capi.DeviceBuffer(4 * grid ** 3)
    energy.zero()
    capi.coulomb3d(atoms_buffer, atoms.shape[0], spacing, grid, energy, form=capi.COMPENSATED)
    existing = energy.download((grid, grid, grid))
    energy.zero()
    capi.coulomb3d_compensated_noalloc(atoms_buffer, atoms.shape[0], spacing, grid, energy)
    capi.synchronize()
    no_alloc = energy.download((grid, grid, grid))
    atoms_buffer.close()
    energy.close()
    return existing, no_alloc


def test_no_alloc_matches_existing_entry_point_on_two_calls(capi):
    cases = [(workloads.volume3d().atoms, 0.5, 256, 41), (workloads.uniform_atoms(30, 600, box=80.0), 0.5, 160, 42)]
    for atoms, spacing, grid, seed in cases:
        print(f"no-alloc call: grid {grid}, {atoms.shape[0]} atoms", flush=True)
        existing, no_alloc = run_both(capi, atoms, spacing, grid)
        points = workloads.sample_points(seed, 4096, grid)
        values, norms = oracle.points_f64(atoms, spacing, points)
        index = (points[:, 2], points[:, 1], points[:, 0])
        difference = np.abs(no_alloc[index].astype(np.float64) - existing[index]) / norms
        worst, rms = oracle.normalised_error(no_alloc[index], values, norms)
        print(f"  |no-alloc - ktt_coulomb3d| max {difference.max():.2e}; bitwise equal: {np.array_equal(no_alloc, existing)}; "
              f"vs FP64 max {worst:.2e} rms {rms:.2e}", flush=True)
        assert difference.max() <= COMPENSATED_BAR
        assert worst <= COMPENSATED_BAR


def test_no_alloc_rejects_bad_input(capi):
    atoms = capi.DeviceBuffer.from_array(workloads.uniform_atoms(1, 10, box=8.0))
    energy = capi.DeviceBuffer(4 * 16 ** 3)
    with pytest.raises(capi.KttHipError, match="grid size"):
        capi.coulomb3d_compensated_noalloc(atoms, 10, 0.5, 0, energy)
    with pytest.raises(capi.KttHipError, match="grid size"):
        capi.coulomb3d_compensated_noalloc(atoms, 10, 0.5, 46341, energy)
    with pytest.raises(capi.KttHipError, match="negative"):
        capi.coulomb3d_compensated_noalloc(atoms, -1, 0.5, 16, energy)


def test_no_alloc_large_grid_folds_row_groups(capi):
    """A grid with more row groups than a launch's y dimension holds (1,108^3: 76,729 groups of 16
    rows, folded into x): the entry point that allocates nothing agrees with ktt_coulomb3d to the
    bar at sampled points and at the last rows, which only the folded launch reaches."""
    grid, spacing = 1108, 0.5
    atoms = workloads.uniform_atoms(37, 600, box=grid * 0.5)
    atoms_buffer = capi.DeviceBuffer.from_array(atoms)
    energy = capi.DeviceBuffer(4 * grid ** 3)
    points = np.concatenate([workloads.sample_points(38, 2048, grid),
                             np.array([[x, grid - 1, grid - 1] for x in range(grid)], dtype=np.int32)])
    offsets = 4 * ((points[:, 2].astype(np.int64) * grid + points[:, 1]) * grid + points[:, 0])
    library = capi.library()

    def read():
        out = np.empty(points.shape[0], dtype=np.float32)
        for i, offset in enumerate(offsets):
            assert library.ktt_hip_memcpy_dtoh(out[i:].ctypes.data_as(ctypes.c_void_p),
                                               ctypes.c_void_p(energy.ptr.value + int(offset)), 4) == 0
        return out

    try:
        print(f"no-alloc call: grid {grid}, {atoms.shape[0]} atoms", flush=True)
        energy.zero()
        capi.coulomb3d(atoms_buffer, atoms.shape[0], spacing, grid, energy, form=capi.COMPENSATED)
        capi.synchronize()
        existing = read()
        energy.zero()
        capi.coulomb3d_compensated_noalloc(atoms_buffer, atoms.shape[0], spacing, grid, energy)
        capi.synchronize()
        no_alloc = read()
    finally:
        energy.close()
        atoms_buffer.close()
    values, norms = oracle.points_f64(atoms, spacing, points)
    assert (np.abs(no_alloc.astype(np.float64) - existing) / norms).max() <= COMPENSATED_BAR
    worst, _ = oracle.normalised_error(no_alloc, values, norms)
    assert worst <= COMPENSATED_BAR
