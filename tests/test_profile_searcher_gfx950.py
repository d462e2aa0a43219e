"""Profile-based searcher (Source/Api/Searcher/ProfileBasedSearcher.h:728-858) on CDNA counters:
simulated tuning over profiled runs of the 288-shape reduced gfx950 CoulombSum3d space on one
MI355X (tests/golden/coulomb3d_gfx950_*_counters.json, written by
make_coulomb3d_gfx950_counters.py from ProfiledCoulombSum3d profile-all). The target is the
128^3 grid with 8,192 atoms; the knowledge bases are the same space profiled at other sizes.

Mean steps to the fastest configuration over 100 seeds, as measured: Random 133.7, MCMC 67.8,
profile-based learning from its own profiled runs 131.5, with the target's own counters as
knowledge base 100.8, with 128^3 x 1,024 atoms as knowledge base 104.2, with 64^3 x 1,024 atoms
213.5. A 64^3 launch has too few workgroups to fill 256 CUs, so its occupancy and busy counters
rank the shapes for another bottleneck than the target's: a knowledge base must come from a
launch of the target's grid; fewer atoms are enough."""

from __future__ import annotations

import json
import math

import numpy as np
import pytest

from ktt_amd import capi
from profiled_tables import GOLDEN, write_expanded

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


def test_every_searcher_finds_the_best_of_the_space(tables):
    target, best_ms, same_grid, small_grid = tables
    for searcher, knowledge in (("Random", None), ("MCMC", None), ("ProfileBased", None), ("ProfileBased", same_grid),
                                ("ProfileBased", small_grid)):
        durations = capi.simulate_search(target, searcher, knowledge, seed=5, compute_units=CU_COUNT)
        assert durations.min() == pytest.approx(best_ms), (searcher, knowledge)
