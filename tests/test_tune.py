"""The autotuner's search on the CPU: a simulated workload whose cost depends on the applied
parameters, timed by a simulated clock, so that the expected choice is known exactly."""

from __future__ import annotations

from scatrs_amd import tune


def test_coordinate_search_picks_each_minimum_and_keeps_it_applied():
    state = dict(tune.DEFAULTS)
    now = [0.0]
    cost = {"sphere_waves": {4: 3.0, 6: 1.0, 8: 2.0, 12: 2.5, 16: 4.0, 20: 4.5},
            "chunk_share": {2: 0.5, 3: 0.7, 4: 0.9, 6: 1.1}}

    def run():
        now[0] += sum(cost[k][state[k]] for k in cost)

    def set_param(name, value):
        state[name] = value

    res = tune.tune(run, set_param, ["sphere_waves", "chunk_share"], repeats=3,
                    clock=lambda: now[0])
    assert res["best"] == {"sphere_waves": 6, "chunk_share": 2}
    assert state["sphere_waves"] == 6 and state["chunk_share"] == 2
    # while sphere_waves was tuned, chunk_share sat at its default (4)
    assert abs(res["timings"]["sphere_waves"][6] - (1.0 + 0.9)) < 1e-12
    assert abs(res["timings"]["chunk_share"][2] - (1.0 + 0.5)) < 1e-12


def test_ties_go_to_the_first_candidate_and_cache_round_trips(tmp_path):
    assert tune.best_of({8: 1.0, 4: 1.0, 6: 2.0}) == 8
    c = tune.TuningCache(tmp_path / "t.json")
    assert c.get("gfx950", "spheres:1048576") is None
    c.put("gfx950", "spheres:1048576", {"sphere_waves": 6})
    c2 = tune.TuningCache(tmp_path / "t.json")
    assert c2.get("gfx950", "spheres:1048576") == {"sphere_waves": 6}
