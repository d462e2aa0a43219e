"""Seeded synthetic inputs of the Coulomb workloads.

Every generator is a fixed algorithm (SplitMix64 bit streams, uniform floats formed from the
top 24 bits so that they are exact in float32), so the GPU harness, the oracle and the
benchmark see identical atoms for a given seed. Atoms are float32 records (x, y, z, q) in an
(N, 4) array, the layout of Examples/CoulombSum3d/CoulombSum3d.cpp:95-98.

Workloads (sizes of the direct-summation benchmark configurations):
  slice2d     one 256x256 slice at z = 31.25 A, h = 0.5 A, 4000 atoms U[0,128)^3, q U[-1,1];
              the slice kernel reads dz^2 in .z (Examples/CoulombSum2d/CoulombSum2d.cl:126).
  volume3d    256^3 grid, h = 0.5 A, 65,536 atoms U[0,128)^3, q U[-1,1].
  clustered   256^3 grid, 1,048,576 atoms: 524,288 +/-1 pairs 1 A apart along random unit
              vectors, pair midpoints N(centre_k, 4^2 I) around 64 centres U[0,128)^3.
  batch       1024 systems, atom counts log-uniform in [200, 20000], cubic grids with side
              in {32, 48, 64, 96, 128}, h = 0.5 A, atoms uniform in each box, q U[-1,1].
  dynamic     192^3 grid, 131,072 atoms U[0,96)^3, jittered by N(0, 0.05^2 I) per step.
  reference_example  the reference CoulombSum3d example's own workload: 256^3 grid, 64 atoms
              U[0,20)^3, q in [0, 0.5) (the only workload the reference publishes a time for).
The reference's own examples draw from random_device (CoulombSum3d.cpp:86-98) and are not
reproducible; these generators replace them on both sides of every comparison.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
MIX1 = np.uint64(0xBF58476D1CE4E5B9)
MIX2 = np.uint64(0x94D049BB133111EB)

SEED_SLICE2D = 1
SEED_VOLUME3D = 2
SEED_CLUSTERED = 3
SEED_BATCH = 4
SEED_DYNAMIC = 5


def splitmix64(seed: int, count: int, stream: int = 0) -> np.ndarray:
    """count SplitMix64 outputs of the sequence (seed, stream)."""
    base = np.uint64((seed * 0x2545F4914F6CDD1D + stream * 0x632BE59BD9B4E019) % (1 << 64))
    with np.errstate(over="ignore"):
        z = base + np.arange(1, count + 1, dtype=np.uint64) * GOLDEN
        z = (z ^ (z >> np.uint64(30))) * MIX1
        z = (z ^ (z >> np.uint64(27))) * MIX2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, count: int, stream: int = 0) -> np.ndarray:
    """Uniform float32 in [0, 1) with 24 random bits (exactly representable)."""
    bits = splitmix64(seed, count, stream) >> np.uint64(40)
    return bits.astype(np.float32) * np.float32(2.0 ** -24)


def normal(seed: int, count: int, stream: int = 0) -> np.ndarray:
    """Standard normal float64 values by Box-Muller on two uniform streams."""
    u1 = (splitmix64(seed, count, 2 * stream + 1) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    u2 = (splitmix64(seed, count, 2 * stream + 2) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    u1 = np.maximum(u1, 2.0 ** -60)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)


def uniform_atoms(seed: int, count: int, box: float, charge: float = 1.0, stream: int = 0) -> np.ndarray:
    """count atoms with x, y, z ~ U[0, box) and q ~ U[-charge, charge), float32 (N, 4)."""
    values = uniform01(seed, 4 * count, stream).reshape(count, 4)
    atoms = np.empty((count, 4), dtype=np.float32)
    atoms[:, :3] = values[:, :3] * np.float32(box)
    atoms[:, 3] = (values[:, 3] * np.float32(2.0) - np.float32(1.0)) * np.float32(charge)
    return atoms


def slice_dz2(atoms: np.ndarray, z_coordinate: float) -> np.ndarray:
    """Copy of atoms whose .z holds (z - Z)^2 in float32, the host precompute of
    Examples/CoulombSum3dIterative/CoulombSum3dIterative.cpp:160-168."""
    out = atoms.copy()
    d = np.float32(z_coordinate) - atoms[:, 2]
    out[:, 2] = d * d
    return out


@dataclass
class Slice2d:
    atoms: np.ndarray          # (N, 4) with real z
    slice_atoms: np.ndarray    # (N, 4) with dz^2 in .z
    z: float
    spacing: float
    grid_x: int
    grid_y: int

    @property
    def interactions(self) -> int:
        return self.grid_x * self.grid_y * self.atoms.shape[0]


def slice2d(atom_count: int = 4000, grid: int = 256, spacing: float = 0.5, z: float = 31.25, box: float = 128.0,
            seed: int = SEED_SLICE2D) -> Slice2d:
    atoms = uniform_atoms(seed, atom_count, box)
    return Slice2d(atoms, slice_dz2(atoms, z), z, spacing, grid, grid)


@dataclass
class Volume:
    atoms: np.ndarray
    spacing: float
    grid: int

    @property
    def interactions(self) -> int:
        return self.grid ** 3 * self.atoms.shape[0]


def volume3d(atom_count: int = 65536, grid: int = 256, spacing: float = 0.5, box: float | None = None,
             seed: int = SEED_VOLUME3D) -> Volume:
    box = grid * spacing if box is None else box
    return Volume(uniform_atoms(seed, atom_count, box), spacing, grid)


def clustered(pair_count: int = 524288, centres: int = 64, grid: int = 256, spacing: float = 0.5, sigma: float = 4.0,
              separation: float = 1.0, seed: int = SEED_CLUSTERED) -> Volume:
    box = grid * spacing
    centre_positions = uniform01(seed, 3 * centres, stream=1).reshape(centres, 3).astype(np.float64) * box
    owner = (splitmix64(seed, pair_count, stream=2) % np.uint64(centres)).astype(np.int64)
    offsets = normal(seed, 3 * pair_count, stream=3).reshape(pair_count, 3) * sigma
    midpoints = centre_positions[owner] + offsets
    direction = normal(seed, 3 * pair_count, stream=4).reshape(pair_count, 3)
    direction /= np.linalg.norm(direction, axis=1, keepdims=True)
    half = 0.5 * separation * direction
    atoms = np.empty((2 * pair_count, 4), dtype=np.float32)
    atoms[0::2, :3] = (midpoints + half).astype(np.float32)
    atoms[1::2, :3] = (midpoints - half).astype(np.float32)
    atoms[0::2, 3] = 1.0
    atoms[1::2, 3] = -1.0
    return Volume(atoms, spacing, grid)


def sample_points(seed: int, count: int, grid: int) -> np.ndarray:
    """count grid indices (x, y, z) drawn uniformly, int32 (count, 3)."""
    values = splitmix64(seed, 3 * count, stream=7) % np.uint64(grid)
    return values.astype(np.int32).reshape(count, 3)


@dataclass
class BatchSystem:
    atoms: np.ndarray
    grid: int
    spacing: float

    @property
    def interactions(self) -> int:
        return self.grid ** 3 * self.atoms.shape[0]


BATCH_SIDES = (32, 48, 64, 96, 128)


def batch(systems: int = 1024, min_atoms: int = 200, max_atoms: int = 20000, spacing: float = 0.5,
          sides: tuple = BATCH_SIDES, seed: int = SEED_BATCH) -> list[BatchSystem]:
    u = uniform01(seed, systems, stream=1).astype(np.float64)
    counts = np.rint(np.exp(math.log(min_atoms) + u * (math.log(max_atoms) - math.log(min_atoms)))).astype(np.int64)
    side_index = (splitmix64(seed, systems, stream=2) % np.uint64(len(sides))).astype(np.int64)
    result = []
    for k in range(systems):
        side = sides[side_index[k]]
        atoms = uniform_atoms(seed * 1000003 + k, int(counts[k]), side * spacing, stream=3)
        result.append(BatchSystem(atoms, side, spacing))
    return result


def dynamic_initial(atom_count: int = 131072, grid: int = 192, spacing: float = 0.5, seed: int = SEED_DYNAMIC) -> Volume:
    """Starting atoms of the dynamic run; each step's displacement is drawn on the device by
    kttJitterAtoms from (seed, step, atom index)."""
    return Volume(uniform_atoms(seed, atom_count, grid * spacing), spacing, grid)


DYNAMIC_STEPS = 500
DYNAMIC_SIGMA = 0.05


SEED_REFERENCE_EXAMPLE = 6


def reference_example(atom_count: int = 64, grid: int = 256, spacing: float = 0.5,
                      seed: int = SEED_REFERENCE_EXAMPLE) -> Volume:
    """The reference CoulombSum3d example's workload (Examples/CoulombSum3d/CoulombSum3d.cpp:67-98):
    a 256^3 grid at h = 0.5 A and 64 atoms with x, y, z ~ U[0, 20) A and q = U[0, 20) / 40, i.e.
    q in [0, 0.5), drawn from this module's seeded stream instead of random_device."""
    values = uniform01(seed, 4 * atom_count).reshape(atom_count, 4) * np.float32(20.0)
    atoms = values.astype(np.float32)
    atoms[:, 3] = values[:, 3] / np.float32(40.0)
    return Volume(atoms, spacing, grid)


def kat_atoms(count: int = 64) -> np.ndarray:
    """Deterministic atoms of Examples/JSON/generate-coulomb-output.c:15-20 in float32 arithmetic:
    x = 1 + 0.1 i, y = 1 + 0.15 i, z = 1 + 0.2 i, q = (i % 10) * 0.1."""
    i = np.arange(count, dtype=np.float32)
    atoms = np.empty((count, 4), dtype=np.float32)
    atoms[:, 0] = np.float32(1.0) + i * np.float32(0.1)
    atoms[:, 1] = np.float32(1.0) + i * np.float32(0.15)
    atoms[:, 2] = np.float32(1.0) + i * np.float32(0.2)
    atoms[:, 3] = (np.arange(count) % 10).astype(np.float32) * np.float32(0.1)
    return atoms


KAT_GRID = 256
KAT_SPACING = np.float32(0.49)
