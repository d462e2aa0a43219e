"""A KTT JSON tuning script (Language "HIP") for this library's CoulombSum3d kernel, run end to
end: parameters, conditions and launch sizes as script expressions, arguments from generators
and constants, validation against a BinaryRaw reference grid written by the oracle."""

from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

from ktt_amd import workloads
from oracle import coulomb_oracle as oracle

pytestmark = pytest.mark.gpu

KERNELS = Path(__file__).resolve().parents[1] / "ktt_amd" / "kernels"
G = 64
N = 64
SPACING = 0.49

# The atoms of the reference generator (x = 1 + 0.1k, y = 1 + 0.15k, z = 1 + 0.2k, q = (k % 10) / 10).
AOS = ("1 + (i//4)*0.1 if i%4 == 0 else (1 + (i//4)*0.15 if i%4 == 1 else "
       "(1 + (i//4)*0.2 if i%4 == 2 else ((i//4)%10)*0.1))")


def write_script(directory: Path, searcher: str, count: int) -> Path:
    reference = oracle.coulomb3d_f32(workloads.kat_atoms(N), SPACING, G)
    (directory / "reference.bin").write_bytes(np.ascontiguousarray(reference, dtype=np.float32).tobytes())
    span = "(WORK_GROUP_SIZE_X * POINTS_X)"
    script = {
        "ConfigurationSpace": {
            "TuningParameters": [
                {"Name": "WORK_GROUP_SIZE_X", "Type": "uint", "Values": "[32, 64]"},
                {"Name": "WORK_GROUP_SIZE_Y", "Type": "uint", "Values": "[2, 4]"},
                {"Name": "POINTS_X", "Type": "uint", "Values": "[1, 2]"},
                {"Name": "ROWS_PER_THREAD", "Type": "uint", "Values": "[4, 8]"},
                {"Name": "ATOM_TILE", "Type": "uint", "Values": "[32, 64]"},
                {"Name": "USE_SOA", "Type": "uint", "Values": "[0, 1]"},
            ],
            "Conditions": [
                {"Parameters": ["WORK_GROUP_SIZE_X", "WORK_GROUP_SIZE_Y"],
                 "Expression": "64 <= WORK_GROUP_SIZE_X * WORK_GROUP_SIZE_Y <= 256"},
            ],
        },
        "Search": {"Name": searcher},
        "Budget": [{"Type": "ConfigurationCount", "BudgetValue": count}],
        "General": {"LoggingLevel": "Info", "TimeUnit": "Microseconds", "OutputFile": "CoulombScriptOutput", "OutputFormat": "JSON"},
        "KernelSpecification": {
            "Language": "HIP",
            "KernelName": "directCoulombSum",
            "KernelFile": str(KERNELS / "CoulombSum3d.hip"),
            "CompilerOptions": [],
            "Profiling": False,
            "GlobalSizeType": "CUDA",
            "GlobalSize": {"X": f"({G} + {span} - 1) // {span}", "Y": f"{G * G} // (WORK_GROUP_SIZE_Y * ROWS_PER_THREAD)", "Z": "1"},
            "LocalSize": {"X": "WORK_GROUP_SIZE_X", "Y": "WORK_GROUP_SIZE_Y", "Z": "1"},
            "Arguments": [
                {"Name": "atomInfo", "Type": "float", "Size": 4 * N, "FillType": "Generator", "DataSource": AOS,
                 "AccessType": "ReadOnly", "MemoryType": "Vector"},
                {"Name": "atomInfoX", "Type": "float", "Size": N, "FillType": "Generator", "DataSource": "1+i*0.1",
                 "AccessType": "ReadOnly", "MemoryType": "Vector"},
                {"Name": "atomInfoY", "Type": "float", "Size": N, "FillType": "Generator", "DataSource": "1+i*0.15",
                 "AccessType": "ReadOnly", "MemoryType": "Vector"},
                {"Name": "atomInfoZ", "Type": "float", "Size": N, "FillType": "Generator", "DataSource": "1+i*0.2",
                 "AccessType": "ReadOnly", "MemoryType": "Vector"},
                {"Name": "atomInfoW", "Type": "float", "Size": N, "FillType": "Generator", "DataSource": "(i%10)*0.1",
                 "AccessType": "ReadOnly", "MemoryType": "Vector"},
                {"Name": "atoms", "Type": "int32", "FillType": "Constant", "FillValue": N, "MemoryType": "Scalar"},
                {"Name": "gridSpacing", "Type": "float", "FillType": "Constant", "FillValue": SPACING, "MemoryType": "Scalar"},
                {"Name": "gridDimensions", "Type": "int32", "FillType": "Constant", "FillValue": G, "MemoryType": "Scalar"},
                {"Name": "grid", "Type": "float", "Size": G ** 3, "FillType": "Constant", "FillValue": 0.0,
                 "AccessType": "ReadWrite", "MemoryType": "Vector"},
            ],
            "ReferenceArguments": [
                {"Name": "reference", "FillType": "BinaryRaw", "DataSource": "reference.bin", "TargetName": "grid",
                 "ValidationMethod": "SideBySideComparison", "ValidationThreshold": 0.01},
            ],
        },
    }
    path = directory / "coulomb3d_hip.json"
    path.write_text(json.dumps(script, indent=1))
    return path


@pytest.mark.parametrize("searcher,count", [("Random", 5), ("MCMC", 4), ("Deterministic", 3)])
def test_tuning_script_runs_and_validates(capi, tmp_path, searcher, count):
    path = write_script(tmp_path, searcher, count)
    print("script", path, flush=True)
    summary = capi.run_tuning_script(str(path))
    print(summary, flush=True)
    assert summary["space"] == capi.tuning_script_space(str(path))
    assert summary["tested"] == count
    assert summary["valid"] == count
    results = json.loads((tmp_path / "CoulombScriptOutput.json").read_text())
    assert results["Metadata"]["ComputeApi"] == "HIP"
    assert [r["Status"] for r in results["Results"]] == ["Ok"] * count


def test_tuning_script_detects_a_wrong_reference(capi, tmp_path):
    path = write_script(tmp_path, "Deterministic", 2)
    reference = np.fromfile(tmp_path / "reference.bin", dtype=np.float32)
    reference[12345] += 1.0
    reference.tofile(tmp_path / "reference.bin")
    summary = capi.run_tuning_script(str(path))
    assert summary["tested"] == 2 and summary["valid"] == 0
    statuses = [r["Status"] for r in json.loads((tmp_path / "CoulombScriptOutput.json").read_text())["Results"]]
    assert statuses == ["ValidationFailed"] * 2
