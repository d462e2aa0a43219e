"""KTT JSON tuning scripts without a device: the expression evaluator against CPython's own
evaluation of the same expressions, and configuration-space counting.

The scripts are evaluated without Python in the library; Python's semantics are the
definition, so every expression is checked against eval() here, bit for bit."""

from __future__ import annotations

import json

import numpy as np
import pytest

from ktt_amd import capi

GENERATORS = [
    "1+i*0.1",
    "1+i*0.15",
    "(i%10)*0.1",
    "1 + (i//4)*0.1 if i%4 == 0 else (1 + (i//4)*0.15 if i%4 == 1 else (1 + (i//4)*0.2 if i%4 == 2 else ((i//4)%10)*0.1))",
    "(i - 50) // 7 + (i - 50) % 7 * 0.5",
    "-(i - 33) // -4",
    "(i * 0.37 - 9.0) // 1.3",
    "(i * 0.37 - 9.0) % -1.3",
    "(i - 20) % 6",
    "2 ** (i % 5) - 2 ** -(i % 3)",
    "i / 3 if 2 < i % 7 <= 5 else -i / 7",
    "1 if (i % 3 == 0 or i % 5 == 0) and not i % 15 == 0 else 0",
    "max(i % 4, 2) + min([i % 3, 1]) + abs(10 - i) + int(i / 4) + float(i % 2)",
    "len(range(i % 6)) + len(range(10, i % 13, -3))",
    "1e-3 * i + .5",
]


@pytest.mark.parametrize("expression", GENERATORS)
def test_generator_matches_python(expression):
    count = 128
    ours = capi.script_generate(expression, count)
    expected = np.array([float(eval(expression, {"__builtins__": {}},  # noqa: S307 - fixed test expressions
                                    {"i": i, "max": max, "min": min, "abs": abs, "int": int, "float": float, "len": len,
                                     "range": lambda *a: list(range(*a))}))
                         for i in range(count)])
    np.testing.assert_array_equal(ours, expected)


def test_generator_errors_are_reported():
    for bad in ("1 +", "i // 0", "j + 1", "(1, 2", "1 if i else"):
        with pytest.raises(capi.KttHipError):
            capi.script_generate(bad, 4)


def test_integer_results_beyond_int64_are_refused_not_wrapped():
    """Python's integers are unbounded; the evaluator keeps int64 and reports a result outside it
    instead of wrapping. Powers are square-and-multiply, so a huge exponent ends at once."""
    np.testing.assert_array_equal(capi.script_generate("2 ** 62 + (2 ** 62 - 1)", 1), [float(2 ** 62 + (2 ** 62 - 1))])
    np.testing.assert_array_equal(capi.script_generate("(-2) ** 63 // 1", 1), [float((-2) ** 63)])
    np.testing.assert_array_equal(capi.script_generate("3 ** 39 * (i + 1)", 2), [float(3 ** 39), float(2 * 3 ** 39)])
    for bad in ("2 ** 64", "2 ** 10 ** 12", "2 ** 62 * 2", "-(2 ** 62) - 2 ** 62 - 1", "(-2) ** 63 // -1",
                "-((-2) ** 63)", "4611686018427387904 + 4611686018427387904"):
        with pytest.raises(capi.KttHipError, match="overflow"):
            capi.script_generate(bad, 1)


def reference_space_script(path, open_cl=False, conditions=True):
    parameters = [
        ("WORK_GROUP_SIZE_X", "[16, 32]"), ("WORK_GROUP_SIZE_Y", "[1, 2, 4, 8]"), ("WORK_GROUP_SIZE_Z", "[1]"),
        ("Z_ITERATIONS", "[1, 2, 4, 8, 16, 32]"), ("INNER_UNROLL_FACTOR", "[0, 1, 2, 4, 8, 16, 32]"),
        ("USE_CONSTANT_MEMORY", "[0, 1]" if open_cl else "[0]"), ("USE_SOA", "[0, 1]"),
        ("VECTOR_SIZE", "[1, 2, 4, 8, 16]" if open_cl else "[1]"),
    ]
    rules = [
        (["INNER_UNROLL_FACTOR", "Z_ITERATIONS"], "INNER_UNROLL_FACTOR < Z_ITERATIONS"),
        (["WORK_GROUP_SIZE_X", "WORK_GROUP_SIZE_Y"], "WORK_GROUP_SIZE_X * WORK_GROUP_SIZE_Y >= 64"),
    ]
    if open_cl:
        rules.append((["USE_SOA", "VECTOR_SIZE"], "USE_SOA or VECTOR_SIZE == 1"))
    script = {
        "ConfigurationSpace": {
            "TuningParameters": [{"Name": n, "Type": "uint", "Values": v} for n, v in parameters],
            "Conditions": [{"Parameters": p, "Expression": e} for p, e in rules] if conditions else [],
        },
        "General": {"TimeUnit": "Microseconds", "OutputFile": "out"},
        "KernelSpecification": {"Language": "HIP", "KernelName": "k", "KernelFile": "k.hip",
                                "GlobalSize": {"X": "1"}, "LocalSize": {"X": "1"}, "Arguments": []},
    }
    path.write_text(json.dumps(script))
    return path


def test_space_of_the_coulomb_example_scripts(tmp_path):
    # The CUDA variant of Examples/CoulombSum3d has 210 configurations (the reference's own
    # full-search dump holds exactly 210 results); the OpenCL variant 1,260.
    assert capi.tuning_script_space(reference_space_script(tmp_path / "cuda.json")) == 210
    assert capi.tuning_script_space(reference_space_script(tmp_path / "opencl.json", open_cl=True)) == 1260
    unconstrained = 2 * 4 * 1 * 6 * 7 * 1 * 2 * 1
    assert capi.tuning_script_space(reference_space_script(tmp_path / "free.json", conditions=False)) == unconstrained


def test_space_errors(tmp_path):
    path = reference_space_script(tmp_path / "bad.json")
    script = json.loads(path.read_text())
    script["ConfigurationSpace"]["Conditions"].append({"Parameters": ["USE_SOA"], "Expression": "USE_SOA < Z_ITERATIONS"})
    path.write_text(json.dumps(script))
    with pytest.raises(capi.KttHipError):
        capi.tuning_script_space(path)
