"""Autotuner host logic (no GPU): candidate set, selection, cache."""
import json

from geopvi_amd import autotune, synth


def test_candidates_cover_all_knobs():
    c = autotune.candidates()
    assert len(c) == 6 * 3 * 2
    assert {v for v, _, _ in c} == set(range(6))
    assert {t for _, _, t in c} == {1, 2}
    assert (0, 0, 1) in c


def test_select_picks_fastest_first_on_ties():
    t = {(0, 0, 1): 5.0, (1, 0, 1): 3.0, (2, 0, 2): 3.0, (3, 8, 1): 4.0}
    assert autotune.select(t) == (1, 0, 1)


def test_autotune_with_timer_and_cache(tmp_path):
    p = synth.workload_params("small")
    path = str(tmp_path / "tune.json")
    cost = {(v, g, t): 10.0 + v + (g > 0) - 0.5 * (t == 2) for v, g, t in autotune.candidates()}
    calls = []

    def timer(c):
        calls.append(c)
        return cost[c]

    r1 = autotune.autotune(p, 4, cache_path=path, timer=timer)
    assert (r1["variant"], r1["grid"], r1["tb"]) == (0, 0, 2)
    assert not r1["cached"] and len(calls) == len(cost)
    data = json.load(open(path))
    key = autotune.shape_key(p, 4, "test") + "|evaluation"
    assert data[key]["variant"] == 0 and data[key]["tb"] == 2
    # second call: cache hit, nothing measured
    r2 = autotune.autotune(p, 4, cache_path=path, timer=timer)
    assert r2["cached"] and len(calls) == len(cost)
    # another shape is a miss
    r3 = autotune.autotune(p, 8, cache_path=path, timer=timer)
    assert not r3["cached"] and len(calls) == 2 * len(cost)
    # force re-measures
    autotune.autotune(p, 4, cache_path=path, timer=timer, force=True)
    assert len(calls) == 3 * len(cost)


def test_corrupt_cache_is_ignored(tmp_path):
    path = tmp_path / "tune.json"
    path.write_text("{not json")
    assert autotune.load_cache(str(path)) == {}
