"""Python front end of the CPU oracle (coulomb_oracle.c) plus numpy brute force.

Test infrastructure only. Functions:
  coulomb3d_f32(atoms, spacing, grid, z0, z1)   FP32 restatement of
      Examples/JSON/generate-coulomb-output.c:22-31 over z-slabs [z0, z1) (C, OpenMP).
  coulomb2d_f32(slice_atoms, spacing, gx, gy)   FP32 restatement of
      Examples/CoulombSum2d/CoulombSum2dReference.cl:8-20, .z holding dz^2 (C, OpenMP).
  points_f64(atoms, spacing, points)            FP64 V and sum |q|/r at grid points (C, OpenMP).
  points2d_f64(slice_atoms, spacing, points)    the same on a dz^2 slice.
  brute_force_f64(atoms, spacing, shape)        numpy FP64 over a whole small grid; the
      independent cross-check of the C oracle for tiny inputs.
  normalised_error(values, reference, normalisers)  max and RMS of |V - V64| / sum_i |q_i|/r_i.
  jitter(atoms, seed, step, sigma)              one Gaussian displacement step of the dynamic
      workload (no reference counterpart: the workload is defined by its seed and sigma);
      pinned by the published SplitMix64 outputs and by the normal distribution's moments.
How each is pinned is listed in coulomb_oracle.c and exercised by tests/test_oracle.py.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIBRARY = HERE / "liboracle.so"
REFERENCE_BINARY = HERE / "_ref" / "generate-coulomb-output"

_library = None


def build() -> Path:
    """Compiles liboracle.so (and the reference generator when the reference is mounted)."""
    subprocess.run(["make", "-s", "-C", str(HERE), "liboracle.so"], check=True)
    if Path(os.environ.get("KTT_REFERENCE", "/root/reference"), "Examples/JSON/generate-coulomb-output.c").exists():
        subprocess.run(["make", "-s", "-C", str(HERE), "ref"], check=True)
    return LIBRARY


def _lib():
    global _library
    if _library is None:
        if not LIBRARY.exists():
            build()
        library = ctypes.CDLL(str(LIBRARY))
        f32p = ctypes.POINTER(ctypes.c_float)
        f64p = ctypes.POINTER(ctypes.c_double)
        i32p = ctypes.POINTER(ctypes.c_int32)
        library.oracle_coulomb3d_f32.argtypes = [f32p, ctypes.c_int, ctypes.c_float, ctypes.c_int, ctypes.c_int, ctypes.c_int, f32p]
        library.oracle_coulomb2d_f32.argtypes = [f32p, ctypes.c_int, ctypes.c_float, ctypes.c_int, ctypes.c_int, f32p]
        library.oracle_slice_dz2.argtypes = [f32p, ctypes.c_int, ctypes.c_float, ctypes.c_int, f32p]
        library.oracle_coulomb_points_f64.argtypes = [f32p, ctypes.c_int, ctypes.c_float, i32p, ctypes.c_int, f64p, f64p]
        library.oracle_coulomb2d_points_f64.argtypes = [f32p, ctypes.c_int, ctypes.c_float, i32p, ctypes.c_int, f64p, f64p]
        library.oracle_max_threads.restype = ctypes.c_int
        _library = library
    return _library


def _f32(array: np.ndarray):
    return array.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _atoms(atoms: np.ndarray) -> np.ndarray:
    atoms = np.ascontiguousarray(atoms, dtype=np.float32)
    if atoms.ndim != 2 or atoms.shape[1] != 4:
        raise ValueError("atoms must be an (N, 4) float32 array")
    return atoms


def threads() -> int:
    return int(_lib().oracle_max_threads())


def coulomb3d_f32(atoms: np.ndarray, spacing: float, grid: int, z0: int = 0, z1: int | None = None,
                  out: np.ndarray | None = None) -> np.ndarray:
    """grid[z0:z1] += V in FP32 with the reference generator's operation order."""
    atoms = _atoms(atoms)
    z1 = grid if z1 is None else z1
    if out is None:
        out = np.zeros((grid, grid, grid), dtype=np.float32)
    assert out.dtype == np.float32 and out.flags.c_contiguous and out.size == grid ** 3
    _lib().oracle_coulomb3d_f32(_f32(atoms), atoms.shape[0], ctypes.c_float(spacing), grid, z0, z1, _f32(out))
    return out


def coulomb2d_f32(slice_atoms: np.ndarray, spacing: float, grid_x: int, grid_y: int) -> np.ndarray:
    atoms = _atoms(slice_atoms)
    out = np.zeros((grid_y, grid_x), dtype=np.float32)
    _lib().oracle_coulomb2d_f32(_f32(atoms), atoms.shape[0], ctypes.c_float(spacing), grid_x, grid_y, _f32(out))
    return out


def slice_dz2(atoms: np.ndarray, spacing: float, z_index: int) -> np.ndarray:
    atoms = _atoms(atoms)
    out = np.empty_like(atoms)
    _lib().oracle_slice_dz2(_f32(atoms), atoms.shape[0], ctypes.c_float(spacing), z_index, _f32(out))
    return out


def points_f64(atoms: np.ndarray, spacing: float, points: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """FP64 potential and normaliser at grid indices points (M, 3)."""
    atoms = _atoms(atoms)
    points = np.ascontiguousarray(points, dtype=np.int32)
    values = np.empty(points.shape[0], dtype=np.float64)
    norms = np.empty(points.shape[0], dtype=np.float64)
    _lib().oracle_coulomb_points_f64(_f32(atoms), atoms.shape[0], ctypes.c_float(spacing),
                                     points.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), points.shape[0],
                                     values.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                     norms.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return values, norms


def points2d_f64(slice_atoms: np.ndarray, spacing: float, points: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """FP64 potential and normaliser at slice indices points (M, 2) = (x, y)."""
    atoms = _atoms(slice_atoms)
    points = np.ascontiguousarray(points, dtype=np.int32)
    values = np.empty(points.shape[0], dtype=np.float64)
    norms = np.empty(points.shape[0], dtype=np.float64)
    _lib().oracle_coulomb2d_points_f64(_f32(atoms), atoms.shape[0], ctypes.c_float(spacing),
                                       points.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), points.shape[0],
                                       values.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                       norms.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return values, norms


def brute_force_f64(atoms: np.ndarray, spacing: float, shape: tuple[int, int, int]) -> np.ndarray:
    """numpy FP64 V over a whole (G_z, G_y, G_x) grid; for tiny inputs only."""
    atoms = np.asarray(atoms, dtype=np.float64)
    h = float(np.float32(spacing))
    gz, gy, gx = shape
    z, y, x = np.meshgrid(np.arange(gz) * h, np.arange(gy) * h, np.arange(gx) * h, indexing="ij")
    points = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
    out = np.zeros(points.shape[0], dtype=np.float64)
    for record in atoms:
        distance = np.sqrt(((points - record[:3]) ** 2).sum(axis=1))
        out += record[3] / distance
    return out.reshape(shape)


def normalised_error(values: np.ndarray, reference: np.ndarray, normalisers: np.ndarray) -> tuple[float, float]:
    """(max, RMS) of |values - reference| / normalisers."""
    error = np.abs(np.asarray(values, dtype=np.float64) - reference) / normalisers
    return float(error.max()), float(np.sqrt(np.mean(error ** 2)))


_GAMMA = np.uint64(0x9E3779B97F4A7C15)


def _mix64(z: np.ndarray) -> np.ndarray:
    """SplitMix64 finaliser on uint64 arrays (wrapping arithmetic)."""
    z = z + _GAMMA
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def jitter(atoms: np.ndarray, seed: int, step: int, sigma: float) -> np.ndarray:
    """One step of the dynamic workload's displacement: every atom moves by sigma times a
    standard normal triple from Box-Muller on counter-based SplitMix64 draws keyed by
    (seed, step, atom index); charges are unchanged. The checker of the device kernel
    kttJitterAtoms, computed in FP64 and rounded once to float32."""
    atoms = np.asarray(atoms, dtype=np.float32)
    with np.errstate(over="ignore"):
        index = np.arange(atoms.shape[0], dtype=np.uint64)
        step_key = _mix64(np.array([step], dtype=np.uint64) * np.uint64(0x632BE59BD9B4E019))
        key = _mix64(np.uint64(seed) ^ step_key) ^ (index * np.uint64(0xD1B54A32D192ED03))
        normals = []
        for k in range(2):
            u1 = (_mix64(key + np.uint64(2 * k + 1)) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
            u2 = (_mix64(key + np.uint64(2 * k + 2)) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
            radius = np.sqrt(-2.0 * np.log(np.maximum(u1, 1.0e-300)))
            normals += [radius * np.cos(6.283185307179586 * u2), radius * np.sin(6.283185307179586 * u2)]
    out = atoms.copy()
    for axis in range(3):
        out[:, axis] = (atoms[:, axis].astype(np.float64) + float(np.float32(sigma)) * normals[axis]).astype(np.float32)
    return out
