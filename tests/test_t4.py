"""KTT JSON <-> T4 results conversion against the reference's own converted files, and simulated
tuning over results that went through T4 and back.

Fixtures (tests/golden/make_t4_fixtures.py): the reference's vectorAdd simulation input and its
T4 conversion (Tutorials/03KernelTuning), and the first 12 results of the RTX 2080 Ti
CoulombSum3d full search with their conversion (Examples/CoulombSum3d)."""

from __future__ import annotations

import json
import subprocess
from pathlib import Path

import pytest

from ktt_amd import t4

REPO = Path(__file__).resolve().parents[1]
GOLDEN = REPO / "tests" / "golden"
CORE_TESTS = REPO / "ktt_amd" / "bin" / "ktt_core_tests"


def load(name: str) -> dict:
    return json.loads((GOLDEN / name).read_text())


def same(a, b) -> bool:
    """Deep equality that also compares the key order of every object and the JSON types."""
    if isinstance(a, dict):
        return isinstance(b, dict) and list(a) == list(b) and all(same(a[k], b[k]) for k in a)
    if isinstance(a, list):
        return isinstance(b, list) and len(a) == len(b) and all(same(x, y) for x, y in zip(a, b))
    return type(a) is type(b) and a == b


@pytest.mark.parametrize("fixture", ["t4_vector_add.json", "t4_coulomb2080_head.json"])
def test_ktt_to_t4_matches_the_reference_conversion(fixture):
    pair = load(fixture)
    converted = t4.ktt_to_t4(pair["ktt"])
    for ours, theirs in zip(converted["results"], pair["t4"]["results"]):
        assert same(ours, theirs), (ours, theirs)
    assert same(converted, pair["t4"])


@pytest.mark.parametrize("fixture", ["t4_vector_add.json", "t4_coulomb2080_head.json"])
def test_t4_round_trip_is_idempotent(fixture):
    reference = load(fixture)["t4"]
    back = t4.t4_to_ktt(reference, kernel="k")
    assert same(t4.ktt_to_t4(back), reference)
    assert back["Metadata"]["TimeUnit"] == "Microseconds"
    assert [r["Status"] for r in back["Results"]] == ["Ok"] * len(reference["results"])


def test_statuses_and_explicit_overheads():
    document = load("t4_vector_add.json")["ktt"]
    statuses = ["Ok", "ComputationFailed", "ValidationFailed", "CompilationFailed"]
    for result, status in zip(document["Results"], statuses):
        result["Status"] = status
    document["Results"][0]["CompilationOverhead"] = 123.5
    document["Results"][0]["ProfilingOverhead"] = 7.25
    converted = t4.ktt_to_t4(document)["results"]
    assert [r["invalidity"] for r in converted] == ["correct", "runtime", "correctness", "compile"]
    assert [r["correctness"] for r in converted] == [1, 1, 0, 1]
    assert converted[0]["times"]["compilation_time"] == 123.5
    assert converted[0]["times"]["framework"] == document["Results"][0]["DataMovementOverhead"] + 7.25
    back = t4.t4_to_ktt(t4.ktt_to_t4(document))
    assert [r["Status"] for r in back["Results"]] == statuses


def test_value_types_of_the_reverse_conversion():
    document = {"schema_version": "1.0.0", "results": [{
        "timestamp": "t", "configuration": {"A": 3, "B": -2, "C": 0.5, "D": True, "E": "x"},
        "objectives": ["TotalDuration"], "times": {"runtimes": [1.0]}, "invalidity": "timeout",
        "correctness": 1, "measurements": []}]}
    back = t4.t4_to_ktt(document, value_types={"A": "Int"})
    assert [p["ValueType"] for p in back["Results"][0]["Configuration"]] == ["Int", "Int", "Double", "Bool", "String"]
    assert back["Results"][0]["Status"] == "ComputationFailed"
    with pytest.raises(t4.ConversionError):
        t4.t4_to_ktt({"schema_version": "2.0", "results": []})
    with pytest.raises(t4.ConversionError):
        t4.ktt_to_t4({"Metadata": {"Timestamp": "t", "TimeUnit": "Microseconds"}, "Results": [{"Status": "Odd"}]})


def test_simulated_tuning_over_t4_results(tmp_path):
    """The 210-result CoulombSum3d search, converted to T4 and back, still drives simulated tuning
    with every searcher to the published best configuration (769.696 us)."""
    if not CORE_TESTS.exists():
        pytest.fail("ktt_core_tests is not built: run python -m ktt_amd.build")
    original = load("coulomb3d_2080_space.json")
    via_t4 = t4.t4_to_ktt(t4.ktt_to_t4(original), kernel=original["Results"][0]["KernelName"],
                          metadata={"ComputeApi": original["Metadata"]["ComputeApi"]})
    path = tmp_path / "coulomb3d_2080_via_t4.json"
    path.write_text(json.dumps(via_t4))
    result = subprocess.run([str(CORE_TESTS), str(path)], capture_output=True, text=True, timeout=300)
    print(result.stdout, result.stderr)
    assert result.returncode == 0, result.stderr
    assert "0 failures" in result.stdout


def test_command_line(tmp_path):
    source = tmp_path / "ktt.json"
    source.write_text(json.dumps(load("t4_vector_add.json")["ktt"]))
    assert t4.main([str(source), str(tmp_path / "t4.json")]) == 0
    assert same(json.loads((tmp_path / "t4.json").read_text()), load("t4_vector_add.json")["t4"])
    assert t4.main(["--to-ktt", "--kernel", "vectorAddition", str(tmp_path / "t4.json"), str(tmp_path / "back.json")]) == 0
    assert json.loads((tmp_path / "back.json").read_text())["Results"][0]["KernelName"] == "vectorAddition"
