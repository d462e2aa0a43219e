"""Writes coulomb3d_2080_space.json: the configurations, statuses and durations of the
reference's full CoulombSum3d search on an RTX 2080 Ti
(Examples/CoulombSum3d/coulomb_2080_full_search_space.json), in KTT's JSON results format,
without the per-run profiling counters and compilation data.

The fixture pins the tuner's configuration space (the CUDA space of
Examples/CoulombSum3d/CoulombSum3d.cpp:139-209 has exactly these 210 configurations), the
results loader, and simulated tuning (the best configuration is the one with the smallest
TotalDuration, 769.696 us).

Usage: python tests/golden/make_coulomb3d_2080_space.py [reference root]
"""

from __future__ import annotations

import json
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent


def main() -> int:
    root = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    source = json.loads((root / "Examples/CoulombSum3d/coulomb_2080_full_search_space.json").read_text())
    results = []
    for entry in source["Results"]:
        computations = []
        for computation in entry["ComputationResults"]:
            computations.append({key: computation[key] for key in ("KernelFunction", "Duration", "Overhead", "GlobalSize",
                                                                    "LocalSize")})
        kept = {key: entry[key] for key in ("KernelName", "Status", "TotalDuration", "TotalOverhead", "ExtraDuration",
                                            "DataMovementOverhead", "ValidationOverhead", "SearcherOverhead", "Configuration")}
        kept["ComputationResults"] = computations
        results.append(kept)
    # One result per line keeps the file small and diffable.
    lines = [json.dumps(result, sort_keys=True, separators=(",", ":")) for result in results]
    text = '{"Metadata":' + json.dumps(source["Metadata"], sort_keys=True) + ',\n"Results":[\n' + ",\n".join(lines) + "\n]}\n"
    json.loads(text)
    (HERE / "coulomb3d_2080_space.json").write_text(text)
    print(len(results), "results")
    return 0


if __name__ == "__main__":
    sys.exit(main())
