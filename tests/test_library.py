"""CPU checks of the native library: it loads, exports every function include/ktt_hip.h declares,
compiles every kernel for gfx950 through hipRTC (no GPU needed), and rejects bad input."""

from __future__ import annotations

import ctypes
import re
from pathlib import Path

import pytest

from ktt_amd import build as native_build
from ktt_amd import capi

REPO = Path(__file__).resolve().parents[1]
HEADER = REPO / "include" / "ktt_hip.h"


def declared_functions() -> list[str]:
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ktt_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    library = capi.library()
    names = declared_functions()
    assert len(names) >= 20
    missing = [name for name in names if not hasattr(library, name)]
    assert not missing, missing
    assert set(names) == set(capi.SIGNATURES), set(names) ^ set(capi.SIGNATURES)
    assert library.ktt_hip_abi_version() == 100


@pytest.mark.parametrize("source, kernel, config", native_build.KERNEL_CHECKS)
def test_kernel_compiles_for_gfx950(source, kernel, config):
    capi.compile_check(str(native_build.KERNELS / source), kernel, config)


def test_invalid_configuration_fails_at_compile_time():
    # 3 rows per workgroup and 64 threads do not divide one another: rejected by a static_assert.
    with pytest.raises(capi.KttHipError, match="divide"):
        capi.compile_check(str(native_build.KERNELS / "CoulombSum3d.hip"), "directCoulombSum",
                           "WORK_GROUP_SIZE_X=64,WORK_GROUP_SIZE_Y=1,ROWS_PER_THREAD=3")


def test_malformed_configuration_string_is_rejected():
    with pytest.raises(capi.KttHipError, match="NAME=VALUE"):
        capi.compile_check(str(native_build.KERNELS / "CoulombSum3d.hip"), "directCoulombSum", "POINTS_X")


def test_launch_shape_follows_configuration():
    grid, block = capi.launch_shape(3, 256, None)
    # Plain-form defaults: 128 x 4 threads, 2 x 8 points per thread.
    assert block == (128, 4, 1)
    assert grid == (256 // 256, 256 * 256 // 32, 1)
    # Rows that are not a multiple of 256 points: 64 x 8 threads (a 128-point span).
    grid, block = capi.launch_shape(3, 96, None)
    assert block == (64, 8, 1)
    assert grid == (1, 96 * 96 // 64, 1)
    grid, block = capi.launch_shape(3, 100, "WORK_GROUP_SIZE_X=32,POINTS_X=2,WORK_GROUP_SIZE_Y=2,ROWS_PER_THREAD=8,ATOM_BLOCKS=3")
    assert block == (32, 2, 1)
    assert grid == (2, (100 * 100 + 15) // 16, 3)
    with pytest.raises(capi.KttHipError):
        capi.launch_shape(3, 256, "WORK_GROUP_SIZE_X=64,WORK_GROUP_SIZE_Y=32")


def test_null_buffers_are_rejected_before_any_launch():
    library = capi.library()
    status = library.ktt_coulomb3d(None, 10, 0.5, 16, None, 0, None, None, None)
    assert status != 0
    message = library.ktt_hip_last_error().decode()
    assert "null" in message


def test_unknown_form_is_rejected():
    library = capi.library()
    dummy = ctypes.c_void_p(16)
    status = library.ktt_coulomb3d(dummy, 0, 0.5, 16, dummy, 7, None, None, None)
    assert status != 0
    assert "form" in library.ktt_hip_last_error().decode()


def test_no_alloc_library_loads_and_checks_input_before_the_device():
    """ktt_coulomb3d_compensated_noalloc loads libktt_hip_noalloc.so (built by hipcc) on its first
    call; input errors are reported before any device work, and without a device the set-up
    reports that instead of launching."""
    assert native_build.NOALLOC_LIBRARY.exists()
    exported = ctypes.CDLL(str(native_build.NOALLOC_LIBRARY))
    assert hasattr(exported, "ktt_noalloc_coulomb3d_compensated")
    library = capi.library()
    dummy = ctypes.c_void_p(16)
    assert library.ktt_coulomb3d_compensated_noalloc(dummy, 10, 0.5, 1024, dummy, None) != 0
    assert "grid size" in library.ktt_hip_last_error().decode()
    assert library.ktt_coulomb3d_compensated_noalloc(None, 10, 0.5, 16, dummy, None) != 0
    assert "null" in library.ktt_hip_last_error().decode()
    assert library.ktt_coulomb3d_compensated_noalloc(ctypes.c_void_p(20), 10, 0.5, 16, dummy, None) != 0
    assert "aligned" in library.ktt_hip_last_error().decode()
