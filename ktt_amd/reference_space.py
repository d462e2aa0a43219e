"""The reference's own CoulombSum3d tuning space as a KTT JSON tuning script for ComputeApi::HIP.

The reference example (Examples/CoulombSum3d/CoulombSum3d.cpp:139-209) tunes its kernel
(CoulombSum3d.cu, identical to Examples/JSON/CoulombSum3d.cu) over WORK_GROUP_SIZE_X/Y/Z,
Z_ITERATIONS, INNER_UNROLL_FACTOR, USE_CONSTANT_MEMORY, USE_SOA and VECTOR_SIZE, with the
launch grid (ceil(G / WORK_GROUP_SIZE_X), ceil(G / WORK_GROUP_SIZE_Y), ceil(G / Z_ITERATIONS)).
ktt_amd/kernels/CoulombSum3dZ.hip has that kernel's name, arguments, parameters and launch, so
the same space runs on this library with the kernel file and the language switched. The CUDA
variant of the space (below) holds the 210 configurations of the reference's published RTX
2080 Ti full search (coulomb_2080_full_search_space.json).
"""

from __future__ import annotations

import json
from pathlib import Path

import numpy as np

KERNEL = Path(__file__).resolve().parent / "kernels" / "CoulombSum3dZ.hip"

# CoulombSum3d.cpp:139-209, the CUDA branch (USE_CONSTANT_MEMORY and VECTOR_SIZE fixed there).
REFERENCE_SPACE = [
    ("WORK_GROUP_SIZE_X", [16, 32]),
    ("WORK_GROUP_SIZE_Y", [1, 2, 4, 8]),
    ("WORK_GROUP_SIZE_Z", [1]),
    ("Z_ITERATIONS", [1, 2, 4, 8, 16, 32]),
    ("INNER_UNROLL_FACTOR", [0, 1, 2, 4, 8, 16, 32]),
    ("USE_CONSTANT_MEMORY", [0]),
    ("USE_SOA", [0, 1]),
    ("VECTOR_SIZE", [1]),
]
REFERENCE_CONDITIONS = [
    (["INNER_UNROLL_FACTOR", "Z_ITERATIONS"], "INNER_UNROLL_FACTOR < Z_ITERATIONS"),
    (["WORK_GROUP_SIZE_X", "WORK_GROUP_SIZE_Y"], "WORK_GROUP_SIZE_X * WORK_GROUP_SIZE_Y >= 64"),
]


def write_script(directory: Path, atoms: np.ndarray, grid: int, spacing: float, space=REFERENCE_SPACE,
                 conditions=REFERENCE_CONDITIONS, searcher: str = "Deterministic", count: int = 210,
                 reference: np.ndarray | None = None, threshold: float = 0.01, logging: str = "Warning",
                 output: str = "ReferenceSpaceOutput") -> Path:
    """Writes `directory`/reference_space.json: a tuning script over `space` for CoulombSum3dZ.hip
    with the given atoms (BinaryRaw inputs in both layouts), a zeroed grid and, when `reference`
    is given, SideBySide validation against it at `threshold` (the reference validates at 0.01).
    The reference's option -use_fast_math is kept: hipRTC accepts it and it changes nothing."""
    directory = Path(directory)
    atoms = np.ascontiguousarray(atoms, dtype=np.float32)
    n = atoms.shape[0]
    atoms.tofile(directory / "atoms.bin")
    for k, name in enumerate("XYZW"):
        np.ascontiguousarray(atoms[:, k]).tofile(directory / f"atoms{name}.bin")
    arguments = [{"Name": "atomInfo", "Type": "float", "Size": 4 * n, "FillType": "BinaryRaw", "DataSource": "atoms.bin",
                  "AccessType": "ReadOnly", "MemoryType": "Vector"}]
    arguments += [{"Name": f"atomInfo{name}", "Type": "float", "Size": n, "FillType": "BinaryRaw",
                   "DataSource": f"atoms{name}.bin", "AccessType": "ReadOnly", "MemoryType": "Vector"} for name in "XYZW"]
    arguments += [
        {"Name": "atoms", "Type": "int32", "FillType": "Constant", "FillValue": n, "MemoryType": "Scalar"},
        {"Name": "gridSpacing", "Type": "float", "FillType": "Constant", "FillValue": spacing, "MemoryType": "Scalar"},
        {"Name": "gridDimensions", "Type": "int32", "FillType": "Constant", "FillValue": grid, "MemoryType": "Scalar"},
        {"Name": "grid", "Type": "float", "Size": grid ** 3, "FillType": "Constant", "FillValue": 0.0,
         "AccessType": "WriteOnly", "MemoryType": "Vector"},
    ]
    specification = {
        "Language": "HIP",
        "KernelName": "directCoulombSum",
        "KernelFile": str(KERNEL),
        "CompilerOptions": ["-use_fast_math"],
        "Profiling": False,
        "GlobalSizeType": "CUDA",
        "GlobalSize": {"X": f"({grid} + WORK_GROUP_SIZE_X - 1) // WORK_GROUP_SIZE_X",
                       "Y": f"({grid} + WORK_GROUP_SIZE_Y - 1) // WORK_GROUP_SIZE_Y",
                       "Z": f"({grid} + Z_ITERATIONS - 1) // Z_ITERATIONS"},
        "LocalSize": {"X": "WORK_GROUP_SIZE_X", "Y": "WORK_GROUP_SIZE_Y", "Z": "1"},
        "Arguments": arguments,
    }
    if reference is not None:
        np.ascontiguousarray(reference, dtype=np.float32).tofile(directory / "reference.bin")
        specification["ReferenceArguments"] = [
            {"Name": "reference", "FillType": "BinaryRaw", "DataSource": "reference.bin", "TargetName": "grid",
             "ValidationMethod": "SideBySideComparison", "ValidationThreshold": threshold}]
    script = {
        "ConfigurationSpace": {
            "TuningParameters": [{"Name": name, "Type": "uint", "Values": json.dumps(values)} for name, values in space],
            "Conditions": [{"Parameters": parameters, "Expression": expression} for parameters, expression in conditions],
        },
        "Search": {"Name": searcher},
        "Budget": [{"Type": "ConfigurationCount", "BudgetValue": count}],
        "General": {"LoggingLevel": logging, "TimeUnit": "Microseconds", "OutputFile": output, "OutputFormat": "JSON"},
        "KernelSpecification": specification,
    }
    path = directory / "reference_space.json"
    path.write_text(json.dumps(script, indent=1))
    return path
