"""Writes tests/golden/coulomb3d_kat.json, the known-answer fixture of the 3D Coulomb grid.

The reference's only Coulomb known answer is the grid written by its CPU generator
Examples/JSON/generate-coulomb-output.c (64 deterministic atoms, 256^3 grid, spacing 0.49,
FP32; consumed by Examples/JSON/CoulombSum3d.json:172-180). This script compiles that program
from the reference sources (oracle/Makefile, target ref), runs it in a temporary directory and
records: the SHA-256 of the whole 64 MiB grid, 4096 seeded sample points with their values
(float32 bit patterns), and summary statistics. The grid itself is not kept.

Usage: python tests/golden/make_coulomb3d_kat.py   (needs /root/reference; run in the container)
"""

from __future__ import annotations

import hashlib
import json
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(REPO))

from ktt_amd import workloads  # noqa: E402

OUTPUT = Path(__file__).resolve().parent / "coulomb3d_kat.json"
SAMPLES = 2048
SAMPLE_SEED = 11


def main() -> None:
    subprocess.run(["make", "-s", "-C", str(REPO / "oracle"), "ref"], check=True)
    binary = REPO / "oracle" / "_ref" / "generate-coulomb-output"
    with tempfile.TemporaryDirectory() as directory:
        subprocess.run([str(binary)], cwd=directory, check=True)
        data = (Path(directory) / "Coulomb3d.bin").read_bytes()
    grid = workloads.KAT_GRID
    if len(data) != 4 * grid ** 3:
        raise SystemExit(f"unexpected output size {len(data)}")
    values = np.frombuffer(data, dtype=np.float32)
    points = workloads.sample_points(SAMPLE_SEED, SAMPLES, grid)
    index = (points[:, 2].astype(np.int64) * grid + points[:, 1]) * grid + points[:, 0]
    sampled = values[index]
    fixture = {
        "source": "Examples/JSON/generate-coulomb-output.c (compiled by oracle/Makefile target ref)",
        "atoms": "workloads.kat_atoms(64): x=1+0.1i, y=1+0.15i, z=1+0.2i, q=(i%10)*0.1 in float32",
        "grid": grid,
        "spacing": float(workloads.KAT_SPACING),
        "atom_count": 64,
        "sha256": hashlib.sha256(data).hexdigest(),
        "sample_seed": SAMPLE_SEED,
        "points": points.tolist(),
        "values_hex": [format(int(v), "08x") for v in sampled.view(np.uint32)],
        "values": [float(v) for v in sampled],
        "stats": {"min": float(values.min()), "max": float(values.max()), "mean": float(values.astype(np.float64).mean())},
    }
    OUTPUT.write_text(json.dumps(fixture, separators=(",", ":")) + "\n")
    print(f"wrote {OUTPUT} (sha256 {fixture['sha256']})")


if __name__ == "__main__":
    main()
