"""The N-body oracle (oracle/nbody_oracle.c) against what the mathematics fixes: the two-body
closed form, zero acceleration at the centre of a symmetric cube, momentum conservation of the
antisymmetric pair forces, the update rule's special cases (dt = 0, damping = 0), a numpy brute
force on small inputs, and the FP32 restatement within rounding of the FP64 one. Each check
must fail for a planted mistake (a sign, the softening, the power of r, the mass)."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import nbody_oracle as nb

SOFT2 = 2.0 ** -6  # exactly representable in FP32, so the C oracle and the closed form agree


def bodies(n: int, seed: int = 3) -> tuple[np.ndarray, np.ndarray]:
    """The example's data: positions, masses and velocities uniform in [0, 20) (Nbody.cpp:78-100)."""
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0.0, 20.0, (n, 4)).astype(np.float32)
    vel = np.zeros((n, 4), dtype=np.float32)
    vel[:, :3] = rng.uniform(0.0, 20.0, (n, 3))
    return pos, vel


def two_body_ok(acceleration) -> bool:
    d, m0, m1 = 3.0, 2.0, 5.0
    pos = np.array([[0, 0, 0, m0], [d, 0, 0, m1]], dtype=np.float32)
    a = acceleration(pos, SOFT2)
    r3 = (d * d + SOFT2) ** 1.5
    expected = np.array([[m1 * d / r3, 0, 0], [-m0 * d / r3, 0, 0]])
    return np.allclose(a, expected, rtol=1e-12, atol=1e-15)


def cube_centre_ok(acceleration) -> bool:
    corners = np.array([[x, y, z, 7.0] for x in (0, 2) for y in (0, 2) for z in (0, 2)], dtype=np.float32)
    pos = np.vstack([corners, [[1.0, 1.0, 1.0, 3.0]]]).astype(np.float32)
    a = acceleration(pos, SOFT2)
    return bool(np.abs(a[-1]).max() < 1e-12)


def momentum_ok(acceleration) -> bool:
    pos, _ = bodies(128)
    a = acceleration(pos, SOFT2)
    weighted = pos[:, 3:4].astype(np.float64) * a
    return bool(np.abs(weighted.sum(axis=0)).max() < 1e-12 * np.abs(weighted).sum())


def oracle_acceleration(pos, soft2):
    vel = np.zeros_like(pos)
    return nb.step_f64(pos, vel, 0.001, 0.5, soft2)[2]


def mutant(kind: str):
    def acceleration(pos, soft2):
        p = np.asarray(pos, dtype=np.float64)
        d = p[None, :, :3] - p[:, None, :3]
        if kind == "sign":
            d = -d
        r2 = (d * d).sum(axis=2) + (0.0 if kind == "softening" else soft2)
        with np.errstate(divide="ignore", invalid="ignore"):
            power = r2 if kind == "power" else r2 * np.sqrt(r2)
            mass = p[:, None, 3] if kind == "mass" else p[None, :, 3]
            f = np.where(r2 > 0, mass / power, 0.0)
        return (f[:, :, None] * d).sum(axis=1)
    return acceleration


CHECKS = [two_body_ok, cube_centre_ok, momentum_ok]


def test_oracle_passes_every_check():
    for check in CHECKS:
        assert check(oracle_acceleration), check.__name__
        assert check(nb.acceleration_brute_f64), check.__name__


@pytest.mark.parametrize("kind", ["sign", "softening", "power", "mass"])
def test_planted_mistakes_fail(kind):
    assert not all(check(mutant(kind)) for check in CHECKS), kind


def test_oracle_matches_brute_force():
    pos, vel = bodies(300)
    _, _, a = nb.step_f64(pos, vel, 0.001, 0.5, SOFT2)
    brute = nb.acceleration_brute_f64(pos, SOFT2)
    assert np.abs(a - brute).max() <= 1e-12 * np.abs(brute).max()


def test_update_rule_special_cases():
    pos, vel = bodies(64)
    for step in (nb.step_f64, nb.step_f32):
        p, v = step(pos, vel, 0.0, 0.5, SOFT2)[:2]
        assert np.array_equal(p, pos.astype(p.dtype)) and np.array_equal(v, vel.astype(v.dtype))
    # Without damping the positions move by dt v exactly as in FP64.
    p, v, a = nb.step_f64(pos, vel, 0.25, 0.0, SOFT2)
    assert np.allclose(p[:, :3], pos[:, :3].astype(np.float64) + 0.25 * vel[:, :3], rtol=0, atol=1e-12)
    assert np.allclose(v[:, :3], vel[:, :3] + 0.25 * a, rtol=1e-15, atol=1e-12)
    # Mass and the velocity's w pass through.
    p32, v32 = nb.step_f32(pos, vel, 0.001, 0.5, SOFT2)
    assert np.array_equal(p32[:, 3], pos[:, 3]) and np.array_equal(v32[:, 3], vel[:, 3])


def test_fp32_restatement_within_rounding():
    pos, vel = bodies(512)
    p32, v32 = nb.step_f32(pos, vel, 0.001, 0.5, SOFT2)
    p64, v64, a = nb.step_f64(pos, vel, 0.001, 0.5, SOFT2)
    scale = 0.001 * nb_magnitude(pos)
    assert (np.abs(v32[:, :3] - v64[:, :3]) <= 1e-5 * (np.abs(vel[:, :3]) + scale[:, None])).all()
    assert (np.abs(p32[:, :3] - p64[:, :3]) <= 1e-6 * np.abs(pos[:, :3]) + 1e-5 * 0.0005 * scale[:, None] + 1e-6).all()


def nb_magnitude(pos: np.ndarray) -> np.ndarray:
    """sum_j |m_j d_ij| / (r^2 + eps^2)^(3/2): the size of the terms an acceleration sums."""
    p = np.asarray(pos, dtype=np.float64)
    d = p[None, :, :3] - p[:, None, :3]
    r2 = (d * d).sum(axis=2) + SOFT2
    return (p[None, :, 3] * np.sqrt((d * d).sum(axis=2)) / (r2 * np.sqrt(r2))).sum(axis=1)
