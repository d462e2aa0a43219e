"""Conversion between KTT's JSON results format and the T4 common autotuning results format
(schema_version 1.0.0 of the TuningSchema T4 results schema), in both directions.

KTT -> T4 follows the reference's converter (Scripts/ConvertJsonOutput.py:13-60):

* one T4 result per KTT result, in order; ``timestamp`` is the KTT file's Metadata.Timestamp
  (the KTT format has no per-result time);
* ``configuration`` maps each parameter name to its value, in the order KTT lists them;
* ``objectives`` is ["TotalDuration"]; ``times.runtimes`` is [TotalDuration];
  ``times.compilation_time`` is CompilationOverhead; ``times.framework`` is
  DataMovementOverhead + ProfilingOverhead; ``times.search_algorithm`` is SearcherOverhead;
  ``times.validation`` is ValidationOverhead;
* ``invalidity`` maps the KTT status (Ok -> correct, ComputationFailed -> runtime,
  ValidationFailed -> correctness, CompilationFailed -> compile, DeviceLimitsExceeded -> runtime);
  ``correctness`` is 0 for ValidationFailed and 1 otherwise;
* ``measurements`` holds TotalDuration (in the file's time unit) and then the profiling counters
  of the result's first kernel computation, unit "".

KTT files written by KTT 2.1.0 (the reference's saved searches) carry neither
CompilationOverhead nor ProfilingOverhead. Their overhead is then split as
compilation = ((TotalOverhead - DataMovementOverhead) - SearcherOverhead) - ValidationOverhead,
profiling = 0, which reproduces the reference's converted files
(Tutorials/03KernelTuning/SimulationInput_converted.json,
Examples/CoulombSum3d/coulomb_2080_full_search_space_converted.json) bit for bit.

T4 -> KTT builds a results file that ``Tuner::LoadResults`` reads and ``Tuner::SimulateTuning``
replays: one kernel computation per result with Duration = runtimes[0] and Overhead =
compilation_time, the measurements after TotalDuration as counters of that computation. T4 keeps
no launch sizes and no value types, so sizes are written as 0 and value types follow the JSON
value (bool -> Bool, non-negative int -> UnsignedInt, negative int -> Int, float -> Double,
string -> String) unless ``value_types`` names them. Invalidity maps back with
correct -> Ok, runtime and timeout -> ComputationFailed, correctness -> ValidationFailed,
compile and constraints -> CompilationFailed (KTT reports time-outs as failed computations and
constraint violations as failed compilations).

Usage:
    python -m ktt_amd.t4 <ktt results.json> <t4 output.json>
    python -m ktt_amd.t4 --to-ktt <t4 results.json> <ktt output.json> [--kernel NAME]
"""

from __future__ import annotations

import argparse
import json
import sys
from pathlib import Path
from typing import Any

SCHEMA_VERSION = "1.0.0"

_INVALIDITY = {
    "Ok": "correct",
    "ComputationFailed": "runtime",
    "ValidationFailed": "correctness",
    "CompilationFailed": "compile",
    "DeviceLimitsExceeded": "runtime",
}

_STATUS = {
    "correct": "Ok",
    "runtime": "ComputationFailed",
    "timeout": "ComputationFailed",
    "correctness": "ValidationFailed",
    "compile": "CompilationFailed",
    "constraints": "CompilationFailed",
}


class ConversionError(ValueError):
    """A results file that does not have the fields its format requires."""


def _field(entry: dict, name: str, where: str) -> Any:
    if name not in entry:
        raise ConversionError(f"{where}: missing field {name!r}")
    return entry[name]


def _overheads(result: dict, where: str) -> tuple[float, float]:
    """(compilation, profiling) overhead of one KTT result."""
    profiling = result.get("ProfilingOverhead", 0.0)
    if "CompilationOverhead" in result:
        return result["CompilationOverhead"], profiling
    total = _field(result, "TotalOverhead", where)
    compilation = total - _field(result, "DataMovementOverhead", where)
    compilation -= _field(result, "SearcherOverhead", where)
    compilation -= _field(result, "ValidationOverhead", where)
    if "ProfilingOverhead" in result:
        compilation -= profiling
    return compilation, profiling


def ktt_to_t4(document: dict) -> dict:
    """The T4 results document of a KTT JSON results document."""
    metadata = _field(document, "Metadata", "KTT results")
    timestamp = _field(metadata, "Timestamp", "Metadata")
    unit = _field(metadata, "TimeUnit", "Metadata")
    converted = []
    for index, result in enumerate(_field(document, "Results", "KTT results")):
        where = f"Results[{index}]"
        status = _field(result, "Status", where)
        if status not in _INVALIDITY:
            raise ConversionError(f"{where}: unknown status {status!r}")
        total = _field(result, "TotalDuration", where)
        compilation, profiling = _overheads(result, where)
        measurements = [{"name": "TotalDuration", "value": total, "unit": unit}]
        computations = _field(result, "ComputationResults", where)
        if computations and "ProfilingData" in computations[0]:
            for counter in computations[0]["ProfilingData"].get("Counters", []):
                measurements.append({"name": counter["Name"], "value": counter["Value"], "unit": ""})
        converted.append({
            "timestamp": timestamp,
            "configuration": {pair["Name"]: pair["Value"] for pair in _field(result, "Configuration", where)},
            "objectives": ["TotalDuration"],
            "times": {
                "compilation_time": compilation,
                "runtimes": [total],
                "framework": _field(result, "DataMovementOverhead", where) + profiling,
                "search_algorithm": _field(result, "SearcherOverhead", where),
                "validation": _field(result, "ValidationOverhead", where),
            },
            "invalidity": _INVALIDITY[status],
            "correctness": 0 if status == "ValidationFailed" else 1,
            "measurements": measurements,
        })
    return {"schema_version": SCHEMA_VERSION, "results": converted}


def _value_type(value: Any) -> str:
    if isinstance(value, bool):
        return "Bool"
    if isinstance(value, int):
        return "UnsignedInt" if value >= 0 else "Int"
    if isinstance(value, float):
        return "Double"
    if isinstance(value, str):
        return "String"
    raise ConversionError(f"configuration value {value!r} has no KTT parameter type")


def _counter_type(value: Any) -> str:
    return "Int" if isinstance(value, int) and not isinstance(value, bool) and value < 0 else (
        "UnsignedInt" if isinstance(value, int) and not isinstance(value, bool) else "Double")


def t4_to_ktt(document: dict, kernel: str = "", value_types: dict[str, str] | None = None,
              metadata: dict[str, Any] | None = None) -> dict:
    """A KTT JSON results document of a T4 results document.

    ``kernel`` names the kernel (and its one computation) in every result; ``value_types`` maps
    parameter names to KTT value types where the JSON value does not fix them; ``metadata``
    overrides entries of the generated Metadata."""
    if document.get("schema_version") != SCHEMA_VERSION:
        raise ConversionError(f"T4 schema_version {document.get('schema_version')!r} is not {SCHEMA_VERSION}")
    value_types = value_types or {}
    results = []
    unit = None
    timestamp = None
    for index, entry in enumerate(_field(document, "results", "T4 results")):
        where = f"results[{index}]"
        invalidity = _field(entry, "invalidity", where)
        if invalidity not in _STATUS:
            raise ConversionError(f"{where}: unknown invalidity {invalidity!r}")
        times = _field(entry, "times", where)
        runtimes = _field(times, "runtimes", where)
        measurements = entry.get("measurements", [])
        total = runtimes[0] if runtimes else 0.0
        counters = []
        for measurement in measurements:
            if measurement["name"] == "TotalDuration":
                unit = unit or measurement.get("unit")
                continue
            counters.append({"Name": measurement["name"], "Type": _counter_type(measurement["value"]),
                             "Value": measurement["value"]})
        timestamp = timestamp or entry.get("timestamp")
        compilation = times.get("compilation_time", 0.0)
        framework = times.get("framework", 0.0)
        searcher = times.get("search_algorithm", 0.0)
        validation = times.get("validation", 0.0)
        computation = {
            "KernelFunction": kernel,
            "Duration": total,
            "Overhead": compilation,
            "CompilationOverhead": compilation,
            "GlobalSize": {"X": 0, "Y": 0, "Z": 0},
            "LocalSize": {"X": 0, "Y": 0, "Z": 0},
        }
        if counters:
            computation["ProfilingData"] = {"Counters": counters, "RemainingProfilingRuns": 0}
        configuration = [{"Name": name, "Value": value, "ValueType": value_types.get(name, _value_type(value))}
                         for name, value in _field(entry, "configuration", where).items()]
        results.append({
            "KernelName": kernel,
            "Status": _STATUS[invalidity],
            "TotalDuration": total,
            "TotalOverhead": compilation + framework + searcher + validation,
            "ExtraDuration": 0.0,
            "DataMovementOverhead": framework,
            "ValidationOverhead": validation,
            "SearcherOverhead": searcher,
            "CompilationOverhead": compilation,
            "ProfilingOverhead": 0.0,
            "Configuration": configuration,
            "ComputationResults": [computation],
        })
    generated = {
        "ComputeApi": "Unknown",
        "Device": "Unknown",
        "GlobalSizeType": "CUDA",
        "KttVersion": "2.1.0",
        "Platform": "Unknown",
        "TimeUnit": unit or "Microseconds",
        "Timestamp": timestamp or "",
    }
    generated.update(metadata or {})
    return {"Metadata": generated, "Results": results}


def main(argv: list[str] | None = None) -> int:
    parser = argparse.ArgumentParser(description="Convert between KTT JSON results and T4 results.")
    parser.add_argument("input")
    parser.add_argument("output")
    parser.add_argument("--to-ktt", action="store_true", help="convert a T4 file to KTT JSON results")
    parser.add_argument("--kernel", default="", help="kernel name written into KTT results (with --to-ktt)")
    args = parser.parse_args(argv)
    document = json.loads(Path(args.input).read_text())
    try:
        converted = t4_to_ktt(document, kernel=args.kernel) if args.to_ktt else ktt_to_t4(document)
    except ConversionError as error:
        print(f"error: {error}", file=sys.stderr)
        return 1
    Path(args.output).write_text(json.dumps(converted, indent=4))
    return 0


if __name__ == "__main__":
    sys.exit(main())
