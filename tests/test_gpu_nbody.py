"""The gfx950 N-body kernel (ktt_amd/kernels/Nbody.hip) and its reference kernel against the
FP64 oracle, over the configurations' code paths, and the N-body example's tuning run
(Examples/Nbody/Nbody.cpp) through ktt::Tuner.

Tolerance: the kernels sum n terms in FP32 with an approximate reciprocal square root, so a
velocity may differ from FP64 by 2e-5 of |v| + dt * sum_j |m_j d_ij| / (r^2 + eps^2)^(3/2) (the
size of the terms summed), and a position by 2e-6 of |p| plus the same share of its update."""

from __future__ import annotations

import numpy as np
import pytest

from ktt_amd import capi
from oracle import nbody_oracle as nb

pytestmark = pytest.mark.gpu

DT, DAMPING, SOFT2 = 0.001, 0.5, 0.01  # Nbody.cpp:71-73


def bodies(n: int, seed: int = 5) -> tuple[np.ndarray, np.ndarray]:
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0.0, 20.0, (n, 4)).astype(np.float32)
    vel = np.zeros((n, 4), dtype=np.float32)
    vel[:, :3] = rng.uniform(0.0, 20.0, (n, 3))
    return pos, vel


def magnitude(pos: np.ndarray) -> np.ndarray:
    p = pos.astype(np.float64)
    total = np.zeros(p.shape[0])
    for start in range(0, p.shape[0], 512):
        d = p[None, :, :3] - p[start:start + 512, None, :3]
        r2 = (d * d).sum(axis=2) + SOFT2
        total[start:start + 512] = (p[None, :, 3] * np.sqrt((d * d).sum(axis=2)) / (r2 * np.sqrt(r2))).sum(axis=1)
    return total


def check(pos, vel, new_pos, new_vel):
    p64, v64, _ = nb.step_f64(pos, vel, DT, DAMPING, SOFT2)
    scale = DT * magnitude(pos)[:, None]
    dv = np.abs(new_vel[:, :3] - v64[:, :3])
    dp = np.abs(new_pos[:, :3] - p64[:, :3])
    assert (dv <= 2e-5 * (np.abs(vel[:, :3]) + scale)).all(), dv.max()
    assert (dp <= 2e-6 * np.abs(pos[:, :3]) + 2e-5 * DAMPING * DT * scale + 1e-6).all(), dp.max()
    assert np.array_equal(new_pos[:, 3], pos[:, 3])
    assert np.array_equal(new_vel[:, 3], np.zeros(pos.shape[0], dtype=np.float32))


def run(pos, vel, config=None, reference=False):
    n = pos.shape[0]
    d_pos, d_vel = capi.DeviceBuffer.from_array(pos), capi.DeviceBuffer.from_array(vel)
    d_new_pos, d_new_vel = capi.DeviceBuffer(16 * n), capi.DeviceBuffer(16 * n)
    if reference:
        capi.nbody_step_reference(d_pos, d_vel, n, DT, DAMPING, SOFT2, d_new_pos, d_new_vel)
    else:
        capi.nbody_step(d_pos, d_vel, n, DT, DAMPING, SOFT2, d_new_pos, d_new_vel, config=config)
    capi.synchronize()
    return d_new_pos.download((n, 4)), d_new_vel.download((n, 4))


CONFIGS = [
    None,
    "WORK_GROUP_SIZE_X=64,WORK_GROUP_SIZE_Y=1,BODIES_PER_THREAD=1,ATOM_TILE=64,INNER_UNROLL_FACTOR=1",
    "WORK_GROUP_SIZE_X=128,WORK_GROUP_SIZE_Y=2,BODIES_PER_THREAD=4,ATOM_TILE=128,INNER_UNROLL_FACTOR=8",
    "WORK_GROUP_SIZE_X=64,WORK_GROUP_SIZE_Y=4,BODIES_PER_THREAD=8,ATOM_TILE=512,INNER_UNROLL_FACTOR=2",
    "WORK_GROUP_SIZE_X=256,WORK_GROUP_SIZE_Y=4,BODIES_PER_THREAD=2,ATOM_TILE=256,INNER_UNROLL_FACTOR=4",
]


@pytest.mark.parametrize("config", CONFIGS)
def test_nbody_kernel_against_fp64(config):
    pos, vel = bodies(4096)
    print("nbody", config, flush=True)
    check(pos, vel, *run(pos, vel, config))


@pytest.mark.parametrize("shape", [None, "WORK_GROUP_SIZE_X=64,WORK_GROUP_SIZE_Y=2,BODIES_PER_THREAD=3,ATOM_TILE=64,"
                                          "INNER_UNROLL_FACTOR=4"])
def test_nbody_issue_phases_are_bit_identical(shape):
    """ISSUE_PHASES=1 (a phase of plain VALU work at wave priority 0, then the v_rsq_f32 at
    priority 3, per source) reorders the issue, not the arithmetic: every target adds its
    sources in the same order with the same operations as the interleaved loop."""
    n = 3072 if shape is not None else 4096
    pos, vel = bodies(n, seed=13)
    outputs = []
    for phases in (0, 1):
        config = f"ISSUE_PHASES={phases}" if shape is None else f"{shape},ISSUE_PHASES={phases}"
        print("nbody", config, flush=True)
        outputs.append(run(pos, vel, config))
    for interleaved, phased in zip(*outputs):
        np.testing.assert_array_equal(phased.view(np.uint32), interleaved.view(np.uint32))
    check(pos, vel, *outputs[1])


def test_nbody_reference_kernel_against_fp64():
    pos, vel = bodies(2048)
    check(pos, vel, *run(pos, vel, reference=True))


def test_nbody_momentum_at_the_example_size():
    """128K bodies (the example's size): sum_i m_i a_i of the antisymmetric pair forces is 0; the
    FP32 kernel keeps it to 1e-5 of sum_i m_i |a_i|."""
    n = 128 * 1024
    pos, vel = bodies(n, seed=7)
    new_pos, new_vel = run(pos, vel)
    a = (new_vel[:, :3].astype(np.float64) - vel[:, :3]) / DT
    weighted = pos[:, 3:4].astype(np.float64) * a
    assert np.abs(weighted.sum(axis=0)).max() <= 1e-5 * np.abs(weighted).sum()


def test_nbody_tuning_run(tmp_path):
    pos, vel = bodies(4096, seed=11)
    out = capi.nbody_tune(pos, vel, DT, DAMPING, SOFT2, searcher=capi.SEARCHER_RANDOM, max_configurations=12, tolerance=0.001,
                          results_file=str(tmp_path / "nbody"))
    print(out, flush=True)
    assert out["tested"] == 12 and out["valid"] == 12
    assert out["best_ms"] > 0
    assert (tmp_path / "nbody.json").exists()
