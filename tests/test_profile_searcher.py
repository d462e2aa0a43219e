"""Profile-based searcher (Source/Api/Searcher/ProfileBasedSearcher.h) in simulated tuning over the
reference's profiled RTX 2080 Ti CoulombSum3d search: 210 configurations, each with the 33 CUPTI
metrics of the reference's default set (tests/golden/coulomb3d_2080_counters.json, made by
make_coulomb3d_2080_counters.py from Examples/CoulombSum3d/coulomb_2080_full_search_space.json).

The searcher is replayed against the saved results with fixed seeds; every profiled run is
answered with the saved counters of its configuration. Steps count every proposal, the
profiling re-runs included, as a real search pays for them."""

from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

from ktt_amd import capi
from profiled_tables import expand

GOLDEN = Path(__file__).resolve().parent / "golden"
SEEDS = range(1, 101)
SM_COUNT = 68  # RTX 2080 Ti


@pytest.fixture(scope="module")
def profiled(tmp_path_factory):
    table = json.loads((GOLDEN / "coulomb3d_2080_counters.json").read_text())
    path = tmp_path_factory.mktemp("profiled") / "coulomb3d_2080_profiled.json"
    path.write_text(json.dumps(expand(table)))
    best_ms = min(row[len(table["Parameters"]) + 1] for row in table["Rows"]) / 1000.0
    return path, best_ms


def steps_to_best(durations: np.ndarray, best_ms: float) -> int:
    running = np.minimum.accumulate(np.where(durations < 0, np.inf, durations))
    hit = np.nonzero(running <= best_ms * (1 + 1e-12))[0]
    return int(hit[0]) + 1 if len(hit) else len(durations) + 1


def mean_steps(path, best_ms, searcher, knowledge=None) -> float:
    return float(np.mean([steps_to_best(capi.simulate_search(path, searcher, knowledge, seed=seed, compute_units=SM_COUNT), best_ms)
                          for seed in SEEDS]))


def test_profile_searcher_covers_the_space(profiled):
    path, best_ms = profiled
    for knowledge in (None, path):
        durations = capi.simulate_search(path, "ProfileBased", knowledge, seed=3, compute_units=SM_COUNT)
        # Every configuration once, plus one profiled re-run per batch of 5.
        assert 210 <= len(durations) <= 210 + 210 // 5 + 1
        assert durations.min() == pytest.approx(best_ms)
        assert (durations > 0).all()


def test_profile_searcher_reaches_the_best_sooner_than_random(profiled):
    """Mean steps to the fastest configuration over 100 seeded searches (the published best,
    769.696 us). Random search needs about 67; the profile-based searcher with the profiled
    search as knowledge base (the reference's exact-mapping variant) about 41, and learning
    only from its own profiled runs about 52."""
    path, best_ms = profiled
    random = mean_steps(path, best_ms, "Random")
    with_knowledge = mean_steps(path, best_ms, "ProfileBased", path)
    online = mean_steps(path, best_ms, "ProfileBased")
    print(f"random {random:.1f}, profile-based with knowledge {with_knowledge:.1f}, online {online:.1f}")
    assert with_knowledge < 0.75 * random
    assert online < 0.9 * random


def test_profile_searcher_without_counters_is_a_batched_random_search(profiled, tmp_path):
    path, best_ms = profiled
    document = json.loads(Path(path).read_text())
    for result in document["Results"]:
        del result["ComputationResults"][0]["ProfilingData"]
    bare = tmp_path / "bare.json"
    bare.write_text(json.dumps(document))
    durations = capi.simulate_search(bare, "ProfileBased", None, seed=11, compute_units=SM_COUNT)
    assert durations.min() == pytest.approx(best_ms)
    with pytest.raises(capi.KttHipError):
        capi.simulate_search(bare, "ProfileBased", bare, seed=11)  # a knowledge base without counters


def test_simulated_searchers_match_the_saved_space(profiled):
    path, best_ms = profiled
    for searcher in ("Deterministic", "Random", "MCMC"):
        durations = capi.simulate_search(path, searcher, None, seed=5)
        assert len(durations) == 210
        assert durations.min() == pytest.approx(best_ms)
    limited = capi.simulate_search(path, "Random", None, seed=5, max_steps=17)
    assert len(limited) == 17
