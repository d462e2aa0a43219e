"""Hardware counters of profiled runs on the HIP backend (rocprofiler-sdk dispatch counting,
replacing the reference's CUPTI metric interface) and the profile-based searcher on a real
device. Each case runs ktt_amd/bin/ProfiledCoulombSum3d in a fresh process, because the counter
service must be registered before the HIP runtime initialises.

The launch profiled is the CoulombSum3d kernel with 64 x 4 threads per group, 2 points and 16
rows per thread on a 128^3 grid: 1 x 256 groups of 4 waves, so SQ_WAVES is exactly 1024."""

from __future__ import annotations

import json
import math
import os
import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

PROGRAM = Path(__file__).resolve().parents[1] / "ktt_amd" / "bin" / "ProfiledCoulombSum3d"
GRID = 128
ATOMS = 1024


def run(*arguments: str, timeout: int = 240) -> dict:
    if not PROGRAM.exists():
        pytest.fail("ProfiledCoulombSum3d is not built: run python -m ktt_amd.build")
    print("running", PROGRAM.name, *arguments, flush=True)
    completed = subprocess.run([str(PROGRAM), *arguments], capture_output=True, text=True, timeout=timeout,
                               env=dict(os.environ))
    print(completed.stdout[-4000:], completed.stderr[-4000:], flush=True)
    assert completed.returncode == 0, completed.stderr
    return json.loads(completed.stdout.strip().splitlines()[-1])


def test_named_counters_in_passes():
    out = run("counters", str(GRID), str(ATOMS), "SQ_WAVES,SQ_INSTS_VALU,GRBM_GUI_ACTIVE,FETCH_SIZE,WRITE_SIZE")
    first, second = out["collections"]
    # FETCH_SIZE needs 3 TCC counters and WRITE_SIZE 2 of the 4 one pass holds: two counter
    # passes after the timing run, counted down in the runs' profiling data.
    assert first["runs"] == 3 and first["remaining"] == [2, 1, 0]
    waves = first["groups"] * math.ceil(first["group_threads"] / 64)
    assert waves == 1024
    for collection in (first, second):
        counters = collection["counters"]
        assert counters["SQ_WAVES"] == waves
        assert counters["GRBM_GUI_ACTIVE"] > 0
        # Each wave runs 32 points x 1024 atoms = 32768 wave-wide interactions at 2 VALU
        # instructions each plus loop and tile overhead.
        per_wave = counters["SQ_INSTS_VALU"] / waves
        assert 2.0 * 32768 <= per_wave <= 8.0 * 32768, per_wave
        # The kernel accumulates into the 8 MiB grid: it is written once.
        grid_kib = GRID ** 3 * 4 / 1024
        assert 0.5 * grid_kib <= counters["WRITE_SIZE"] <= 2.0 * grid_kib, counters["WRITE_SIZE"]
        assert counters["FETCH_SIZE"] > 0
        assert counters["kernel_duration_ns"] > 0
    assert first["counters"]["SQ_INSTS_VALU"] == second["counters"]["SQ_INSTS_VALU"]
    assert first["counters"]["SQ_WAVES"] == second["counters"]["SQ_WAVES"]


def test_default_counters():
    out = run("counters", str(GRID), str(ATOMS))
    collection = out["collections"][0]
    counters = collection["counters"]
    for name in ("FETCH_SIZE", "WRITE_SIZE", "VALUBusy", "OccupancyPercent", "SQ_WAVES", "SQ_INSTS_VALU", "SQ_INSTS_VALU_TRANS_F32",
                 "SQ_INSTS_LDS", "TCC_HIT_sum", "SIMD_UTILIZATION"):
        assert name in counters, name
        assert math.isfinite(counters[name]) and counters[name] >= 0, (name, counters[name])
    assert counters["SQ_WAVES"] == 1024
    assert 0 < counters["VALUBusy"] <= 100
    # One reciprocal square root per interaction: 32768 per wave.
    assert counters["SQ_INSTS_VALU_TRANS_F32"] == pytest.approx(1024 * 32768, rel=0.05)
    assert collection["runs"] >= 3
    # Dual issue: a v_rsq_f32 holds two issue quad-cycles and any other VALU instruction one, and
    # at most half of the others can issue in a pair with another wave's.
    valu, trans = counters["SQ_INSTS_VALU"], counters["SQ_INSTS_VALU_TRANS_F32"]
    assert counters["SQ_ACTIVE_INST_VALU"] == pytest.approx(valu + trans, rel=1e-3)
    assert 0 <= counters["SQ_ACTIVE_INST_VALU2"] <= (valu - trans) / 2


def test_profile_based_search(tmp_path):
    results = tmp_path / "profiled"
    out = run("search", "25", "64", "512", str(results))
    assert out["tested"] == 25
    assert out["valid"] >= 20
    # Batches of 5, each followed by the profiled run of its fastest configuration.
    assert out["profiled"] >= 3
    saved = json.loads((tmp_path / "profiled.json").read_text())
    with_counters = [r for r in saved["Results"] if any(c["Name"] == "VALUBusy" for c in
                                                        r["ComputationResults"][0].get("ProfilingData", {}).get("Counters", []))]
    assert len(with_counters) == out["profiled"]


def test_profile_based_search_with_a_smaller_grid_knowledge_base(tmp_path):
    """A knowledge base profiled on a 32^3 grid drives a search on 64^3. The searcher scales the
    base's instruction counts to each candidate's launch on 64^3, which it computes from the
    tuning space's thread modifiers."""
    knowledge = tmp_path / "knowledge"
    base = run("profile-all", "30", "32", "256", str(knowledge), "", "reduced")
    assert base["tested"] == 30 and base["profiled"] == 30
    results = tmp_path / "steered"
    out = run("search", "15", "64", "512", str(results), str(knowledge) + ".json", "reduced")
    assert out["tested"] == 15
    assert out["valid"] == 15
    assert out["profiled"] >= 2
