"""Python front end of the N-body oracle (nbody_oracle.c) plus a numpy brute force.

Test infrastructure only. Functions:
  step_f32(pos, vel, dt, damping, soft2)   FP32 restatement of Examples/Nbody/NbodyReference.cu:40-75.
  step_f64(pos, vel, dt, damping, soft2)   FP64 with compensated sums; returns (pos', vel', acc).
  acceleration_brute_f64(pos, soft2)       numpy FP64 accelerations; the independent check of the
      C oracle on small inputs.
pos and vel are (n, 4) float32 arrays: (x, y, z, mass) and (vx, vy, vz, 0).
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import coulomb_oracle


def _lib():
    library = coulomb_oracle._lib()
    if not getattr(library, "_nbody_bound", False):
        f32p = ctypes.POINTER(ctypes.c_float)
        f64p = ctypes.POINTER(ctypes.c_double)
        library.oracle_nbody_f32.argtypes = [f32p, f32p, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_float, f32p, f32p]
        library.oracle_nbody_f64.argtypes = [f32p, f32p, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_float, f64p, f64p, f64p]
        library._nbody_bound = True
    return library


def _f32(array: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(array, dtype=np.float32)


def step_f32(pos: np.ndarray, vel: np.ndarray, dt: float, damping: float, soft2: float) -> tuple[np.ndarray, np.ndarray]:
    pos, vel = _f32(pos), _f32(vel)
    n = pos.shape[0]
    out_pos = np.empty_like(pos)
    out_vel = np.empty_like(vel)
    f32p = ctypes.POINTER(ctypes.c_float)
    _lib().oracle_nbody_f32(pos.ctypes.data_as(f32p), vel.ctypes.data_as(f32p), n, dt, damping, soft2,
                            out_pos.ctypes.data_as(f32p), out_vel.ctypes.data_as(f32p))
    return out_pos, out_vel


def step_f64(pos: np.ndarray, vel: np.ndarray, dt: float, damping: float, soft2: float) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    pos, vel = _f32(pos), _f32(vel)
    n = pos.shape[0]
    out_pos = np.empty((n, 4), dtype=np.float64)
    out_vel = np.empty((n, 4), dtype=np.float64)
    acc = np.empty((n, 3), dtype=np.float64)
    f32p = ctypes.POINTER(ctypes.c_float)
    f64p = ctypes.POINTER(ctypes.c_double)
    _lib().oracle_nbody_f64(pos.ctypes.data_as(f32p), vel.ctypes.data_as(f32p), n, dt, damping, soft2,
                            out_pos.ctypes.data_as(f64p), out_vel.ctypes.data_as(f64p), acc.ctypes.data_as(f64p))
    return out_pos, out_vel, acc


def acceleration_brute_f64(pos: np.ndarray, soft2: float) -> np.ndarray:
    p = np.asarray(pos, dtype=np.float64)
    d = p[None, :, :3] - p[:, None, :3]
    r2 = (d * d).sum(axis=2) + soft2
    f = p[None, :, 3] / (r2 * np.sqrt(r2))
    return (f[:, :, None] * d).sum(axis=1)
