"""Profile-based searcher (Source/Api/Searcher/ProfileBasedSearcher.h:728-858) on CDNA counters:
simulated tuning over profiled runs of the 288-shape reduced gfx950 CoulombSum3d space on one
MI355X (tests/golden/coulomb3d_gfx950_*_counters.json, written by
make_coulomb3d_gfx950_counters.py from ProfiledCoulombSum3d profile-all). The target is the
128^3 grid with 8,192 atoms; the knowledge bases are the same space profiled at other sizes.

Mean steps to the fastest configuration over 100 seeds, as measured: Random 133.7, MCMC 67.8,
profile-based learning from its own profiled runs 121.8, with the target's own counters as
knowledge base 103.0, with 128^3 x 1,024 atoms as knowledge base 105.0, with 64^3 x 1,024 atoms
76.6. On the 64^3 grid a shape's workgroups overhang the grid, so its instruction counts grow
with its launched threads, not with the target's, and a shape of few workgroups leaves compute
units idle; the searcher scales a knowledge base's counts by the ratio of the target's launched
threads to the base's, and its device-wide utilisations by the ratio of the compute units the
two launches occupy (identity on the same grid). Unscaled, the 64^3 knowledge base took 224.1
steps, more than random search."""

from __future__ import annotations

import json
import math

import numpy as np
import pytest

from ktt_amd import capi
from profiled_tables import GOLDEN, expand, write_expanded

SEEDS = range(1, 101)
CU_COUNT = 256  # MI355X


@pytest.fixture(scope="module")
def tables(tmp_path_factory):
    directory = tmp_path_factory.mktemp("gfx950")
    target, best_ms = write_expanded("coulomb3d_gfx950_128x8192_counters.json", directory)
    same_grid, _ = write_expanded("coulomb3d_gfx950_128x1024_counters.json", directory)
    small_grid, _ = write_expanded("coulomb3d_gfx950_64x1024_counters.json", directory)
    return target, best_ms, same_grid, small_grid


def steps_to_best(durations: np.ndarray, best_ms: float) -> int:
    running = np.minimum.accumulate(np.where(durations < 0, np.inf, durations))
    hit = np.nonzero(running <= best_ms * (1 + 1e-12))[0]
    return int(hit[0]) + 1 if len(hit) else len(durations) + 1


def mean_steps(path, best_ms, searcher, knowledge=None) -> float:
    return float(np.mean([steps_to_best(capi.simulate_search(path, searcher, knowledge, seed=seed, compute_units=CU_COUNT), best_ms)
                          for seed in SEEDS]))


def test_tables_hold_the_cdna_counters(tables):
    """Every shape ran, validated and was profiled with the default counter set, and the counters
    obey what the kernel does: one v_rsq_f32 per wave-interaction of the launch (x span and rows
    rounded up to whole workgroups, atoms to whole tiles), no LDS bank conflicts, no MFMA."""
    for name, grid, atoms in (("coulomb3d_gfx950_128x8192_counters.json", 128, 8192),
                              ("coulomb3d_gfx950_128x1024_counters.json", 128, 1024),
                              ("coulomb3d_gfx950_64x1024_counters.json", 64, 1024)):
        table = json.loads((GOLDEN / name).read_text())
        assert len(table["Rows"]) == 288
        column = {name: i for i, name in enumerate(table["Columns"])}
        for row in table["Rows"]:
            assert row[column["Status"]] == "Ok"
            wgx, wgy, px, rows, tile = (row[column[key]] for key in
                                        ("WORK_GROUP_SIZE_X", "WORK_GROUP_SIZE_Y", "POINTS_X", "ROWS_PER_THREAD", "ATOM_TILE"))
            span, group_rows = wgx * px, wgy * rows
            points = math.ceil(grid / span) * span * math.ceil(grid * grid / group_rows) * group_rows
            wave_interactions = points * math.ceil(atoms / tile) * tile / 64
            assert row[column["SQ_INSTS_VALU_TRANS_F32"]] == pytest.approx(wave_interactions, rel=1e-3)
            assert row[column["LDSBankConflict"]] == 0
            assert row[column["SQ_INSTS_MFMA"]] == 0


def test_same_grid_knowledge_base_beats_random(tables):
    target, best_ms, same_grid, _ = tables
    random = mean_steps(target, best_ms, "Random")
    with_same_grid = mean_steps(target, best_ms, "ProfileBased", same_grid)
    with_itself = mean_steps(target, best_ms, "ProfileBased", target)
    print(f"random {random:.1f}, knowledge 128^3 x 1024 {with_same_grid:.1f}, knowledge = target {with_itself:.1f}")
    assert with_same_grid < 0.85 * random
    assert with_itself < 0.85 * random


def test_small_grid_knowledge_base_scales_to_the_target_launch(tables):
    """A knowledge base profiled on a grid 8 times smaller, whose launches overhang it by shape,
    steers the search once its counts are scaled to the target's launched threads."""
    target, best_ms, _, small_grid = tables
    seeds = range(1, 51)
    random = float(np.mean([steps_to_best(capi.simulate_search(target, "Random", None, seed=seed, compute_units=CU_COUNT), best_ms)
                            for seed in seeds]))
    with_small_grid = float(np.mean([steps_to_best(capi.simulate_search(target, "ProfileBased", small_grid, seed=seed,
                                                                        compute_units=CU_COUNT), best_ms) for seed in seeds]))
    print(f"random {random:.1f}, knowledge 64^3 x 1024 {with_small_grid:.1f}")
    assert with_small_grid < 0.85 * random


def test_every_searcher_finds_the_best_of_the_space(tables):
    target, best_ms, same_grid, small_grid = tables
    for searcher, knowledge in (("Random", None), ("MCMC", None), ("ProfileBased", None), ("ProfileBased", same_grid),
                                ("ProfileBased", small_grid)):
        durations = capi.simulate_search(target, searcher, knowledge, seed=5, compute_units=CU_COUNT)
        assert durations.min() == pytest.approx(best_ms), (searcher, knowledge)


# The same 288 shapes, each with the interleaved (ISSUE_PHASES=0) and the phased (1) inner loop,
# profiled with the default counter set that includes the dual-issue counters SQ_ACTIVE_INST_VALU
# and SQ_ACTIVE_INST_VALU2 (coulomb3d_gfx950_<grid>x<atoms>_phases_counters.json, from
# ProfiledCoulombSum3d profile-all ... reduced-phases). The 50 fastest of the 576 run the phased
# loop. Mean steps to the fastest over 100 seeds, as measured: Random 276.9, MCMC 139.0,
# profile-based learning from its own runs 120.5, with 128^3 x 1,024 atoms as knowledge base
# 30.5; without the two dual-issue counters 195.4 and 115.3.
DUAL_ISSUE = ("SQ_ACTIVE_INST_VALU", "SQ_ACTIVE_INST_VALU2")
PHASES_SEEDS = range(1, 51)


def write_table(name: str, directory, drop=()):
    """The expanded table, without the counters in drop; its path and fastest duration (ms)."""
    table = json.loads((GOLDEN / name).read_text())
    if drop:
        keep = [i for i, counter in enumerate(table["Counters"]) if counter not in drop]
        fixed = len(table["Parameters"]) + 9
        table["Rows"] = [row[:fixed] + [row[fixed + i] for i in keep] for row in table["Rows"]]
        table["Counters"] = [table["Counters"][i] for i in keep]
    path = directory / name.replace("_counters.json", "_dropped.json" if drop else "_profiled.json")
    path.write_text(json.dumps(expand(table)))
    count = len(table["Parameters"])
    return path, min(row[count + 1] for row in table["Rows"] if row[count] == "Ok") / 1000.0


@pytest.fixture(scope="module")
def phase_tables(tmp_path_factory):
    directory = tmp_path_factory.mktemp("gfx950_phases")
    target, best_ms = write_table("coulomb3d_gfx950_128x8192_phases_counters.json", directory)
    knowledge, _ = write_table("coulomb3d_gfx950_128x1024_phases_counters.json", directory)
    target_dropped, _ = write_table("coulomb3d_gfx950_128x8192_phases_counters.json", directory, DUAL_ISSUE)
    knowledge_dropped, _ = write_table("coulomb3d_gfx950_128x1024_phases_counters.json", directory, DUAL_ISSUE)
    return target, best_ms, knowledge, target_dropped, knowledge_dropped


def test_phase_tables_hold_the_dual_issue_counters():
    """Every configuration ran, validated and was profiled; a v_rsq_f32 counts two issue
    quad-cycles and any other VALU instruction one (SQ_ACTIVE_INST_VALU = VALU + TRANS), at most
    half the other instructions issue in a pair (none when a launch has fewer waves than the
    GPU's 1,024 SIMDs can pair), and the phased loop pairs more than the interleaved loop of the
    same shape."""
    for name in ("coulomb3d_gfx950_128x8192_phases_counters.json", "coulomb3d_gfx950_128x1024_phases_counters.json",
                 "coulomb3d_gfx950_64x1024_phases_counters.json"):
        table = json.loads((GOLDEN / name).read_text())
        assert len(table["Rows"]) == 576
        column = {name: i for i, name in enumerate(table["Columns"])}
        shapes = {}
        for row in table["Rows"]:
            assert row[column["Status"]] == "Ok"
            valu, trans = row[column["SQ_INSTS_VALU"]], row[column["SQ_INSTS_VALU_TRANS_F32"]]
            active, paired = row[column["SQ_ACTIVE_INST_VALU"]], row[column["SQ_ACTIVE_INST_VALU2"]]
            assert active == pytest.approx(valu + trans, rel=1e-3)
            assert 0 <= paired <= (valu - trans) / 2
            if row[column["SQ_WAVES"]] >= 4096:
                assert paired > 0
            shape = tuple(row[column[key]] for key in ("WORK_GROUP_SIZE_X", "WORK_GROUP_SIZE_Y", "POINTS_X", "ROWS_PER_THREAD",
                                                       "ATOM_TILE"))
            shapes.setdefault(shape, {})[row[column["ISSUE_PHASES"]]] = paired / (valu - trans)
        assert len(shapes) == 288
        assert np.median([pair[1] / pair[0] for pair in shapes.values() if pair[0] > 0]) > 1.2


def test_dual_issue_counters_steer_the_profile_searcher(phase_tables):
    target, best_ms, knowledge, target_dropped, knowledge_dropped = phase_tables

    def mean(path, searcher, base=None):
        return float(np.mean([steps_to_best(capi.simulate_search(path, searcher, base, seed=seed, compute_units=CU_COUNT), best_ms)
                              for seed in PHASES_SEEDS]))

    random = mean(target, "Random")
    with_pairing = mean(target, "ProfileBased", knowledge)
    without_pairing = mean(target_dropped, "ProfileBased", knowledge_dropped)
    print(f"random {random:.1f}, profile-based with dual-issue counters {with_pairing:.1f}, without {without_pairing:.1f}")
    assert with_pairing < 0.5 * without_pairing
    assert with_pairing < 0.25 * random


def test_small_grid_knowledge_base_on_the_phases_space(tmp_path):
    """The 576 configurations with both inner loops, steered by a knowledge base profiled on a
    64^3 grid with 1,024 atoms (coulomb3d_gfx950_64x1024_phases_counters.json). Measured over 100
    seeds: 135.0 steps to the fastest with the base scaled to the target's launches, 426.1 with
    the same base given the target's launch sizes (so that nothing is scaled), 276.9 at random."""
    target, best_ms = write_table("coulomb3d_gfx950_128x8192_phases_counters.json", tmp_path)
    small_grid, _ = write_table("coulomb3d_gfx950_64x1024_phases_counters.json", tmp_path)

    def mean(searcher, knowledge):
        return float(np.mean([steps_to_best(capi.simulate_search(target, searcher, knowledge, seed=seed, compute_units=CU_COUNT),
                                            best_ms) for seed in PHASES_SEEDS]))

    random = mean("Random", None)
    steered = mean("ProfileBased", small_grid)
    print(f"random {random:.1f}, knowledge 64^3 x 1024 {steered:.1f}")
    assert steered < 0.7 * random
