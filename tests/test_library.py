"""CPU checks of the native library: it loads, exports every function include/ktt_hip.h declares,
compiles every kernel for gfx950 through hipRTC (no GPU needed), and rejects bad input."""

from __future__ import annotations

import ctypes
import re
from pathlib import Path

import numpy as np
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


def fold_cover(grid_side, grid, block, span, rows_per_group):
    """The tiles (row group * x groups + x group) of the workgroups of a volume launch that do
    not exit, as the kernels derive them (kc::volumeGroup in ktt_amd/kernels/coulomb_core.h),
    restated over all workgroups; with the x and row group counts."""
    groups_x = -(-grid_side // span)
    groups_y = -(-grid_side * grid_side // rows_per_group)
    fold = max(1, grid[0] // groups_x)
    bx = np.arange(grid[0])[None, :]
    by = np.arange(grid[1])[:, None]
    row = by * fold + bx // groups_x
    valid = (bx < fold * groups_x) & (row < groups_y)
    tiles = (row * groups_x + bx % groups_x)[valid]
    return groups_x, groups_y, tiles


def test_volume_launch_folds_row_groups_beyond_the_y_limit():
    """More than 65,535 row groups fold into x; every (x group, row group) pair is then covered
    by exactly one workgroup, and a launch within the limit is not folded."""
    # Plain defaults at 1,500 (not a multiple of 256): 64 x 8 threads, 2 x 8 points: a 128-point
    # span and 64 rows per group, 35,157 row groups, no fold.
    grid, block = capi.launch_shape(3, 1500, None)
    assert block == (64, 8, 1) and grid == (12, -(-1500 * 1500 // 64), 1)
    # The compensated defaults' 16 rows per group at 1,100: 75,625 row groups, fold 2.
    config = "WORK_GROUP_SIZE_X=128,WORK_GROUP_SIZE_Y=2,POINTS_X=2,ROWS_PER_THREAD=8"
    grid, block = capi.launch_shape(3, 1100, config)
    groups_x, groups_y, tiles = fold_cover(1100, grid, block, 256, 16)
    assert (groups_x, groups_y) == (5, 75625)
    assert grid == (2 * groups_x, -(-groups_y // 2), 1)
    assert np.array_equal(np.sort(tiles), np.arange(groups_x * groups_y))
    # A tall grid of narrow groups: 2,000 with 4 rows per group is 1,000,000 row groups, fold 16.
    config = "WORK_GROUP_SIZE_X=64,WORK_GROUP_SIZE_Y=1,POINTS_X=4,ROWS_PER_THREAD=4"
    grid, block = capi.launch_shape(3, 2000, config)
    groups_x, groups_y, tiles = fold_cover(2000, grid, block, 256, 4)
    fold = -(-groups_y // 65535)
    assert fold == 16 and grid == (fold * groups_x, -(-groups_y // fold), 1) and grid[1] <= 65535
    assert np.array_equal(np.sort(tiles), np.arange(groups_x * groups_y))
    # The slice kernels take their row length from the launch and never fold.
    grid, _ = capi.launch_shape(2, 1100, config)
    assert grid[0] == -(-1100 // 256)


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
    assert library.ktt_coulomb3d_compensated_noalloc(dummy, 10, 0.5, 46341, dummy, None) != 0
    assert "grid size" in library.ktt_hip_last_error().decode()
    assert library.ktt_coulomb3d_compensated_noalloc(None, 10, 0.5, 16, dummy, None) != 0
    assert "null" in library.ktt_hip_last_error().decode()
    assert library.ktt_coulomb3d_compensated_noalloc(ctypes.c_void_p(20), 10, 0.5, 16, dummy, None) != 0
    assert "aligned" in library.ktt_hip_last_error().decode()
