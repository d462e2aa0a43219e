"""The reference's own CoulombSum3d tuning space on ComputeApi::HIP (ktt_amd/reference_space.py).

A KTT JSON tuning script holds the parameters, conditions and launch sizes of the reference's
example (Examples/CoulombSum3d/CoulombSum3d.cpp:139-209, CUDA variant, 210 configurations)
with ktt_amd/kernels/CoulombSum3dZ.hip in place of CoulombSum3d.cu. Validation is the
reference's (SideBySide against a reference grid, here the oracle's FP32 restatement of the
reference generator, Examples/JSON/generate-coulomb-output.c:22-31) at 1e-3 instead of its
0.01: grid values reach 52 near the atoms, where FP32 with an approximate rsqrt leaves about
1e-5 relative.
"""

from __future__ import annotations

import json

import pytest

from ktt_amd import capi as capi_module
from ktt_amd import reference_space, workloads
from oracle import coulomb_oracle as oracle

THRESHOLD = 1e-3


def test_reference_space_has_the_reference_size(tmp_path):
    """CPU: the script's space is the reference's 210 configurations (the count of the
    reference's CUDA full-search dump, coulomb_2080_full_search_space.json)."""
    path = reference_space.write_script(tmp_path, workloads.kat_atoms(64), 64, 0.49, count=1)
    assert capi_module.tuning_script_space(str(path)) == 210


@pytest.mark.gpu
def test_reference_space_sample_validates(capi, tmp_path):
    """24 configurations drawn at random from the reference's space, at 64^3 with the reference
    generator's 64 atoms, each validated against the oracle."""
    atoms = workloads.kat_atoms(64)
    reference = oracle.coulomb3d_f32(atoms, 0.49, 64)
    path = reference_space.write_script(tmp_path, atoms, 64, 0.49, searcher="Random", count=24, reference=reference,
                                        threshold=THRESHOLD)
    print("script", path, flush=True)
    summary = capi.run_tuning_script(str(path))
    print(summary, flush=True)
    assert summary["space"] == 210
    assert summary["tested"] == 24 and summary["valid"] == 24
    results = json.loads((tmp_path / "ReferenceSpaceOutput.json").read_text())
    assert [r["Status"] for r in results["Results"]] == ["Ok"] * 24


@pytest.mark.gpu
def test_reference_space_edge_configurations(capi, tmp_path):
    """Every code path of the kernel on a grid that the blocks do not tile (40 points, partial
    workgroups in x and y, a partial z run) with 300 atoms (a partial second LDS tile):
    constant-memory reads, SoA with 4-atom staging, partial z unrolls."""
    atoms = workloads.uniform_atoms(5, 300, box=19.0)
    reference = oracle.coulomb3d_f32(atoms, 0.5, 40)
    space = [("WORK_GROUP_SIZE_X", [32]), ("WORK_GROUP_SIZE_Y", [2]), ("WORK_GROUP_SIZE_Z", [1]),
             ("Z_ITERATIONS", [3, 16]), ("INNER_UNROLL_FACTOR", [0, 2]), ("USE_CONSTANT_MEMORY", [0, 1]),
             ("USE_SOA", [0, 1]), ("VECTOR_SIZE", [1, 4])]
    conditions = [(["USE_SOA", "VECTOR_SIZE"], "USE_SOA == 1 or VECTOR_SIZE == 1")]
    path = reference_space.write_script(tmp_path, atoms, 40, 0.5, space, conditions, count=24, reference=reference,
                                        threshold=THRESHOLD)
    summary = capi.run_tuning_script(str(path))
    print(summary, flush=True)
    assert summary["space"] == 24
    assert summary["tested"] == 24 and summary["valid"] == 24
