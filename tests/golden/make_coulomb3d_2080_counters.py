"""Writes coulomb3d_2080_counters.json: the reference's full CoulombSum3d search on an RTX 2080 Ti
(Examples/CoulombSum3d/coulomb_2080_full_search_space.json, 210 configurations, every one
profiled with the 33 CUPTI metrics of the reference's default set) in a compact table: one row
per result with the parameter values, TotalDuration, the kernel's Duration, global and local
size and the counter values. tests/test_profile_searcher.py expands it back into KTT JSON
results for simulated tuning with the profile-based searcher.

Usage: python tests/golden/make_coulomb3d_2080_counters.py [reference root]
"""

from __future__ import annotations

import json
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent


def main() -> int:
    root = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    source = json.loads((root / "Examples/CoulombSum3d/coulomb_2080_full_search_space.json").read_text())
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
    table = {
        "Metadata": source["Metadata"],
        "KernelName": results[0]["KernelName"],
        "KernelFunction": results[0]["ComputationResults"][0]["KernelFunction"],
        "Parameters": parameters,
        "ValueTypes": types,
        "Counters": counters,
        "Columns": parameters + ["Status", "TotalDuration", "Duration", "GlobalX", "GlobalY", "GlobalZ", "LocalX", "LocalY",
                                 "LocalZ"] + counters,
    }
    text = json.dumps(table, indent=1)[:-2] + ',\n "Rows": [\n' + ",\n".join(
        json.dumps(row, separators=(",", ":")) for row in rows) + "\n ]\n}\n"
    json.loads(text)
    (HERE / "coulomb3d_2080_counters.json").write_text(text)
    print(len(rows), "rows,", len(counters), "counters")
    return 0


if __name__ == "__main__":
    sys.exit(main())
