"""Argument checks of the N-body entry points, which fail before any device work (no GPU
needed): null buffers, a body count that the configuration does not tile, USE_SOA on the
float4 entry point, configurations beyond the LDS or thread limits."""

from __future__ import annotations

import ctypes

import pytest

from ktt_amd import capi

FAKE = ctypes.c_void_p(4096)  # never dereferenced: every call below fails its checks first


def step(n, config=None, bodies=FAKE):
    status = capi.library().ktt_nbody_step(bodies, FAKE, n, 0.001, 0.5, 0.01, FAKE, FAKE, config, None, None)
    return status, capi.library().ktt_hip_last_error().decode()


@pytest.mark.parametrize("n,config,bodies,message", [
    (1024, None, None, "must not be null"),
    (0, None, FAKE, "must be positive"),
    (1000, None, FAKE, "must be a multiple"),
    (1024, b"WORK_GROUP_SIZE_X=64,WORK_GROUP_SIZE_Y=4,ATOM_TILE=512,BODIES_PER_THREAD=8", FAKE, "must be a multiple"),
    (1024, b"USE_SOA=1", FAKE, "USE_SOA=1"),
    (4096, b"WORK_GROUP_SIZE_X=512,WORK_GROUP_SIZE_Y=4", FAKE, "invalid N-body configuration"),
    (65536, b"WORK_GROUP_SIZE_X=256,WORK_GROUP_SIZE_Y=4,BODIES_PER_THREAD=8,ATOM_TILE=1024", FAKE, "64 KiB of LDS"),
])
def test_nbody_step_rejects(n, config, bodies, message):
    status, error = step(n, config, bodies)
    assert status == -1 and message in error, error


def test_nbody_reference_needs_whole_workgroups():
    status = capi.library().ktt_nbody_step_reference(FAKE, FAKE, 100, 0.001, 0.5, 0.01, FAKE, FAKE, None, None)
    assert status == -1 and b"multiple of 64" in capi.library().ktt_hip_last_error()


class _Unallocated:
    """Stands in for a DeviceBuffer; the calls below fail their checks before using it."""

    ptr = FAKE


def test_python_wrappers_marshal_their_arguments():
    for wrapper, message in ((capi.nbody_step, "must be a multiple"), (capi.nbody_step_reference, "multiple of 64")):
        with pytest.raises(capi.KttHipError, match=message):
            wrapper(_Unallocated, _Unallocated, 1000, 0.001, 0.5, 0.01, _Unallocated, _Unallocated, timed=True)
    import numpy as np
    with pytest.raises(capi.KttHipError, match="multiple of 64"):
        capi.nbody_tune(np.zeros((100, 4), np.float32), np.zeros((100, 4), np.float32), 0.001, 0.5, 0.01)
