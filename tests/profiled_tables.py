"""Expands the compact counter tables under tests/golden (one row per profiled configuration)
back into KTT JSON results, the input of simulated tuning."""

from __future__ import annotations

import json
from pathlib import Path

GOLDEN = Path(__file__).resolve().parent / "golden"


def expand(table: dict) -> dict:
    """KTT JSON results of a compact counter table."""
    count = len(table["Parameters"])
    results = []
    for row in table["Rows"]:
        configuration = [{"Name": n, "Value": v, "ValueType": t} for n, v, t in zip(table["Parameters"], row[:count], table["ValueTypes"])]
        status, total, duration = row[count:count + 3]
        global_size, local_size, values = row[count + 3:count + 6], row[count + 6:count + 9], row[count + 9:]
        computation = {
            "KernelFunction": table["KernelFunction"], "Duration": duration, "Overhead": 0.0,
            "GlobalSize": dict(zip("XYZ", global_size)), "LocalSize": dict(zip("XYZ", local_size)),
            "ProfilingData": {"RemainingProfilingRuns": 0,
                              "Counters": [{"Name": n, "Type": "Double", "Value": v} for n, v in zip(table["Counters"], values)]},
        }
        results.append({"KernelName": table["KernelName"], "Status": status, "TotalDuration": total, "TotalOverhead": 0.0,
                        "ExtraDuration": max(0.0, total - duration), "DataMovementOverhead": 0.0, "ValidationOverhead": 0.0,
                        "SearcherOverhead": 0.0, "Configuration": configuration, "ComputationResults": [computation]})
    return {"Metadata": table["Metadata"], "Results": results}


def write_expanded(name: str, directory: Path) -> tuple[Path, float]:
    """Writes the expanded table tests/golden/<name> into directory; returns its path and the
    fastest TotalDuration in milliseconds."""
    table = json.loads((GOLDEN / name).read_text())
    path = directory / name.replace("_counters.json", "_profiled.json")
    path.write_text(json.dumps(expand(table)))
    best_ms = min(row[len(table["Parameters"]) + 1] for row in table["Rows"] if row[len(table["Parameters"])] == "Ok") / 1000.0
    return path, best_ms
