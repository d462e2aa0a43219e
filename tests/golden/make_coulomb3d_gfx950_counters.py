"""Writes coulomb3d_gfx950_<grid>x<atoms>[_<space>]_counters.json: a profiled run of every
configuration of the reduced gfx950 CoulombSum3d space (288 shapes: at most 64 threads in x, atom
tiles of at most 128, no unroll, float4 atoms, the interleaved inner loop; with the space
"phases", the same shapes with ISSUE_PHASES 0 and 1) on one MI355X, in the compact table of
coulomb3d_2080_counters.json. Each configuration holds the rocprofiler-sdk counters of the HIP
backend's default set and the kernel's static data (VGPRs, LDS, scratch, occupancy).

The input is the KTT JSON results file that the profiling example writes on the GPU:

    ktt_amd/bin/ProfiledCoulombSum3d profile-all 0 <grid> <atoms> <results> "" reduced
    ktt_amd/bin/ProfiledCoulombSum3d profile-all 0 <grid> <atoms> <results> "" reduced-phases

Usage: python tests/golden/make_coulomb3d_gfx950_counters.py <results.json> <grid> <atoms> [phases]
"""

from __future__ import annotations

import json
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent


def compact(source: dict) -> dict:
    results = source["Results"]
    parameters = [pair["Name"] for pair in results[0]["Configuration"]]
    types = [pair["ValueType"] for pair in results[0]["Configuration"]]
    counters = [counter["Name"] for counter in results[0]["ComputationResults"][0]["ProfilingData"]["Counters"]]
    rows = []
    for result in results:
        computation = result["ComputationResults"][0]
        assert [pair["Name"] for pair in result["Configuration"]] == parameters
        values = {counter["Name"]: counter["Value"] for counter in computation["ProfilingData"]["Counters"]}
        row = [pair["Value"] for pair in result["Configuration"]]
        row += [result["Status"], result["TotalDuration"], computation["Duration"]]
        row += [computation["GlobalSize"][axis] for axis in "XYZ"] + [computation["LocalSize"][axis] for axis in "XYZ"]
        row += [values[name] for name in counters]
        rows.append(row)
    return {
        "Metadata": source["Metadata"],
        "KernelName": results[0]["KernelName"],
        "KernelFunction": results[0]["ComputationResults"][0]["KernelFunction"],
        "Parameters": parameters,
        "ValueTypes": types,
        "Counters": counters,
        "Columns": parameters + ["Status", "TotalDuration", "Duration", "GlobalX", "GlobalY", "GlobalZ", "LocalX", "LocalY",
                                 "LocalZ"] + counters,
        "Rows": rows,
    }


def main() -> int:
    if len(sys.argv) not in (4, 5):
        print(__doc__)
        return 2
    table = compact(json.loads(Path(sys.argv[1]).read_text()))
    rows = table.pop("Rows")
    text = json.dumps(table, indent=1)[:-2] + ',\n "Rows": [\n' + ",\n".join(
        json.dumps(row, separators=(",", ":")) for row in rows) + "\n ]\n}\n"
    json.loads(text)
    space = f"_{sys.argv[4]}" if len(sys.argv) == 5 else ""
    target = HERE / f"coulomb3d_gfx950_{sys.argv[2]}x{sys.argv[3]}{space}_counters.json"
    target.write_text(text)
    print(target.name, len(rows), "rows,", len(table["Counters"]), "counters")
    return 0


if __name__ == "__main__":
    sys.exit(main())
